// gfx950 kernels either side of the dedispersion launch: the input corner
// turn (time-major spectra -> channel-major rows) and the per-segment
// detrend/normalise of dedispersed rows.  Both are HBM-bound.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mdsm_internal.h"

namespace mdsm {
namespace {

// ---------------------------------------------------------------------------
// Corner turn: out[c][col0 + t] = in[t][invert ? C-1-c : c]
// One workgroup moves a 64-sample x 64-channel tile through LDS, so both the
// read (along channels) and the write (along time) are contiguous per wave.
// ---------------------------------------------------------------------------
constexpr int kCT = 64;

template <typename TI, typename TO>
__global__ void __launch_bounds__(256)
    corner_turn_kernel(const TI *__restrict__ in, int64_t in_stride, uint32_t nsamp,
                       uint32_t nchans, int invert, TO *__restrict__ out, int64_t out_stride,
                       int64_t col0)
{
    __shared__ TO tile[kCT][kCT + 1];
    const uint32_t t0 = blockIdx.x * kCT, c0 = blockIdx.y * kCT;
    const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
#pragma unroll 4
    for (int r = ly; r < kCT; r += 4) {  // r: sample within the tile
        const uint32_t t = t0 + r, c = c0 + lx;
        if (t < nsamp && c < nchans) {
            const uint32_t src = invert ? nchans - 1 - c : c;
            tile[lx][r] = (TO)in[(int64_t)t * in_stride + src];
        }
    }
    __syncthreads();
#pragma unroll 4
    for (int r = ly; r < kCT; r += 4) {  // r: channel within the tile
        const uint32_t c = c0 + r, t = t0 + lx;
        if (t < nsamp && c < nchans) out[(int64_t)c * out_stride + col0 + t] = tile[r][lx];
    }
}

// uint8 corner turn without LDS.  A lane owns 4 channels x 16 consecutive
// spectra: it loads one dword (its 4 channels) of each of the 16 spectra,
// transposes the four 4x4 byte blocks with v_perm_b32 and stores 16
// consecutive samples of each channel with one 16-byte store.  A wave is CG
// channel groups x (64/CG) spectrum groups, so a load instruction reads 4*CG
// contiguous bytes of 64/CG spectra and a store writes 16*64/CG contiguous
// bytes of CG channel rows; the workgroup's four waves take consecutive runs
// of spectra.  Needs dword-aligned input rows, 16-byte-aligned output rows
// and nchans % 4 == 0 (the host checks).
__device__ __forceinline__ void transpose4x4(uint32_t a, uint32_t b, uint32_t c, uint32_t d,
                                             uint32_t &p0, uint32_t &p1, uint32_t &p2, uint32_t &p3)
{
    // rows a..d hold channels 0..3 of samples 0..3; p_j = channel j, samples 0..3
    const uint32_t t0 = __builtin_amdgcn_perm(b, a, 0x05010400u);  // a0 b0 a1 b1
    const uint32_t t1 = __builtin_amdgcn_perm(b, a, 0x07030602u);  // a2 b2 a3 b3
    const uint32_t t2 = __builtin_amdgcn_perm(d, c, 0x05010400u);  // c0 d0 c1 d1
    const uint32_t t3 = __builtin_amdgcn_perm(d, c, 0x07030602u);  // c2 d2 c3 d3
    p0 = __builtin_amdgcn_perm(t2, t0, 0x05040100u);  // a0 b0 c0 d0
    p1 = __builtin_amdgcn_perm(t2, t0, 0x07060302u);  // a1 b1 c1 d1
    p2 = __builtin_amdgcn_perm(t3, t1, 0x05040100u);
    p3 = __builtin_amdgcn_perm(t3, t1, 0x07060302u);
}

#ifndef MDSM_CT_CG
#define MDSM_CT_CG 16
#endif
#ifndef MDSM_CT_NT
#define MDSM_CT_NT 16
#endif
constexpr int kCtCG = MDSM_CT_CG;     // channel groups (of 4) per wave
constexpr int kCtNT = MDSM_CT_NT;     // spectra per lane (multiple of 16)
constexpr int kCtTQ = 64 / kCtCG;     // spectrum groups per wave
constexpr int kCtWaveT = kCtNT * kCtTQ;  // spectra per wave
constexpr int kCtBlockT = 4 * kCtWaveT;
constexpr int kCtBlockC = 4 * kCtCG;

__global__ void __launch_bounds__(256)
    corner_turn_u8_kernel(const uint8_t *__restrict__ in, int64_t in_stride, uint32_t nsamp,
                          uint32_t nchans, int invert, uint8_t *__restrict__ out,
                          int64_t out_stride, int64_t col0)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int cg = lane % kCtCG, tq = lane / kCtCG;
    const uint32_t cbase = blockIdx.y * (uint32_t)kCtBlockC + 4u * (uint32_t)cg;
    const uint32_t tl = blockIdx.x * (uint32_t)kCtBlockT + (uint32_t)wave * kCtWaveT + kCtNT * tq;
    if (cbase >= nchans || tl >= nsamp) return;  // nchans % 4 == 0: all 4 channels valid
    // input column of the lane's channel group (read reversed when inverting)
    const uint32_t icol = invert ? nchans - 4u - cbase : cbase;
#pragma unroll
    for (int h = 0; h < kCtNT / 16; ++h) {
    const uint32_t t0 = tl + 16u * h;
    if (t0 >= nsamp) break;
    uint32_t v[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const uint32_t t = t0 + i;
        v[i] = t < nsamp ? *(const uint32_t *)(in + (int64_t)t * in_stride + icol) : 0u;
        if (invert) v[i] = __builtin_amdgcn_perm(v[i], v[i], 0x00010203u);  // reverse bytes
    }
    uint32_t q[4][4];  // q[j][k]: channel j, samples 4k..4k+3
#pragma unroll
    for (int k = 0; k < 4; ++k)
        transpose4x4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3], q[0][k], q[1][k],
                     q[2][k], q[3][k]);
    const bool full = t0 + 16u <= nsamp;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        uint8_t *o = out + (int64_t)(cbase + j) * out_stride + col0 + t0;
        if (full) {
            *(uint4 *)o = make_uint4(q[j][0], q[j][1], q[j][2], q[j][3]);
        } else {
            for (uint32_t i = 0; i < 16u && t0 + i < nsamp; ++i)
                o[i] = (uint8_t)(q[j][i >> 2] >> (8 * (i & 3)));
        }
    }
    }
}

// Wide variant: a lane owns 16 channels x 16 spectra (16-byte stores, a
// 16x16 byte transpose as sixteen 4x4 blocks).  A wave is 8 channel groups x
// 8 spectrum groups: a store writes 128 contiguous bytes of 8 channel rows,
// a load reads 16*NB contiguous bytes of 8 spectra.  NB = 8 is the uint8
// corner turn; NB = 1, 2, 4 unpack packed SIGPROC spectra on the way.  Needs
// input rows aligned to 2*NB bytes, 16-byte-aligned output rows and
// nchans % 16 == 0.
constexpr int kCw = 8;                  // 16-channel groups per wave
constexpr int kCwT = 16 * (64 / kCw);   // spectra per wave
constexpr int kCwBlockT = 4 * kCwT, kCwBlockC = 16 * kCw;

// 16 consecutive samples of one spectrum as 16 bytes (channel order),
// from NB-bit packed samples (least significant bits first within a byte):
// 2*NB bytes at p, aligned to 2*NB.
template <int NB>
__device__ __forceinline__ uint4 load16(const uint8_t *p)
{
    if constexpr (NB == 8) {
        return *(const uint4 *)p;
    } else if constexpr (NB == 4) {
        const uint2 u = *(const uint2 *)p;
        auto spread = [](uint32_t x) {  // 2 bytes of nibbles -> 4 bytes
            const uint32_t y = (x & 0xffu) | ((x & 0xff00u) << 8);
            return (y | (y << 4)) & 0x0f0f0f0fu;
        };
        return make_uint4(spread(u.x & 0xffffu), spread(u.x >> 16), spread(u.y & 0xffffu),
                          spread(u.y >> 16));
    } else if constexpr (NB == 2) {
        const uint32_t d = *(const uint32_t *)p;
        auto spread = [](uint32_t b) {  // 1 byte of 2-bit samples -> 4 bytes
            return (b | (b << 6) | (b << 12) | (b << 18)) & 0x03030303u;
        };
        return make_uint4(spread(d & 0xffu), spread((d >> 8) & 0xffu), spread((d >> 16) & 0xffu),
                          spread(d >> 24));
    } else {
        static_assert(NB == 1, "1, 2, 4 or 8 bits");
        const uint32_t h = *(const uint16_t *)p;
        auto spread = [](uint32_t n) {  // 4 bits -> 4 bytes
            return (n | (n << 7) | (n << 14) | (n << 21)) & 0x01010101u;
        };
        return make_uint4(spread(h & 0xfu), spread((h >> 4) & 0xfu), spread((h >> 8) & 0xfu),
                          spread(h >> 12));
    }
}

template <int NB>
__global__ void __launch_bounds__(256)
    corner_turn_u8x16_kernel(const uint8_t *__restrict__ in, int64_t in_stride, uint32_t nsamp,
                             uint32_t nchans, int invert, uint8_t *__restrict__ out,
                             int64_t out_stride, int64_t col0)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int cg = lane % kCw, tq = lane / kCw;
    const uint32_t cbase = blockIdx.y * (uint32_t)kCwBlockC + 16u * (uint32_t)cg;
    const uint32_t t0 = blockIdx.x * (uint32_t)kCwBlockT + (uint32_t)wave * kCwT + 16u * tq;
    if (cbase >= nchans || t0 >= nsamp) return;  // nchans % 16 == 0: all 16 channels valid
    const uint32_t icol = invert ? nchans - 16u - cbase : cbase;
    uint32_t v[16][4];  // v[i][g]: spectrum t0+i, channels 4g..4g+3 of the lane's 16
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const uint32_t t = t0 + i;
        uint4 w = make_uint4(0u, 0u, 0u, 0u);
        if (t < nsamp) w = load16<NB>(in + (int64_t)t * in_stride + icol * NB / 8);
        if (invert) {  // reverse the 16 bytes
            v[i][0] = __builtin_amdgcn_perm(w.w, w.w, 0x00010203u);
            v[i][1] = __builtin_amdgcn_perm(w.z, w.z, 0x00010203u);
            v[i][2] = __builtin_amdgcn_perm(w.y, w.y, 0x00010203u);
            v[i][3] = __builtin_amdgcn_perm(w.x, w.x, 0x00010203u);
        } else {
            v[i][0] = w.x;
            v[i][1] = w.y;
            v[i][2] = w.z;
            v[i][3] = w.w;
        }
    }
    const bool full = t0 + 16u <= nsamp;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        uint32_t q[4][4];  // q[j][k]: channel 4g+j, samples 4k..4k+3
#pragma unroll
        for (int k = 0; k < 4; ++k)
            transpose4x4(v[4 * k][g], v[4 * k + 1][g], v[4 * k + 2][g], v[4 * k + 3][g], q[0][k],
                         q[1][k], q[2][k], q[3][k]);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            uint8_t *o = out + (int64_t)(cbase + 4 * g + j) * out_stride + col0 + t0;
            if (full) {
                *(uint4 *)o = make_uint4(q[j][0], q[j][1], q[j][2], q[j][3]);
            } else {
                for (uint32_t i = 0; i < 16u && t0 + i < nsamp; ++i)
                    o[i] = (uint8_t)(q[j][i >> 2] >> (8 * (i & 3)));
            }
        }
    }
}

// Packed SIGPROC spectra: sample c of a spectrum starts at bit c*NB, least
// significant bits first within a byte (NB <= 8); 16-bit unsigned and 32-bit
// float samples are plain arrays.
template <int NB>
__device__ __forceinline__ float packed_sample(const uint8_t *spec, uint32_t c)
{
    if constexpr (NB == 32) {
        return ((const float *)spec)[c];
    } else if constexpr (NB == 16) {
        return (float)((const uint16_t *)spec)[c];
    } else {
        const uint32_t bit = c * NB;
        return (float)((spec[bit >> 3] >> (bit & 7)) & ((1u << NB) - 1u));
    }
}

template <int NB, typename TO>
__global__ void __launch_bounds__(256)
    unpack_kernel(const uint8_t *__restrict__ raw, int64_t in_stride, uint32_t nsamp,
                  uint32_t nchans, int invert, TO *__restrict__ out, int64_t out_stride,
                  int64_t col0)
{
    __shared__ TO tile[kCT][kCT + 1];
    const uint32_t t0 = blockIdx.x * kCT, c0 = blockIdx.y * kCT;
    const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
#pragma unroll 4
    for (int r = ly; r < kCT; r += 4) {
        const uint32_t t = t0 + r, c = c0 + lx;
        if (t < nsamp && c < nchans) {
            const uint32_t src = invert ? nchans - 1 - c : c;
            tile[lx][r] = (TO)packed_sample<NB>(raw + (int64_t)t * in_stride, src);
        }
    }
    __syncthreads();
#pragma unroll 4
    for (int r = ly; r < kCT; r += 4) {
        const uint32_t c = c0 + r, t = t0 + lx;
        if (t < nsamp && c < nchans) out[(int64_t)c * out_stride + col0 + t] = tile[r][lx];
    }
}

// ---------------------------------------------------------------------------
// Detrend and normalise: per row and per segment of L samples, subtract the
// least-squares line a + b x (x centred on the segment) and divide by the
// RMS of the residual.  Sums are formed in FP64.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum_d(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ double block_sum_d(double v, double *red)
{
    v = wave_sum_d(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();  // red[] may still be read from the previous call
    if (l == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) s += red[i];
    return s;
}

__global__ void __launch_bounds__(256)
    detrend_kernel(float *x, int64_t stride, uint32_t n, uint32_t seg_len)
{
    __shared__ double red[4];
    const uint32_t row = blockIdx.y;
    const uint32_t s0 = blockIdx.x * seg_len;
    const uint32_t L = min(seg_len, n - s0);
    float *p = x + (int64_t)row * stride + s0;
    const double xc = -(double)L / 2.0 + 0.5;  // abscissa of sample 0
    double sy = 0.0, sxy = 0.0;
    for (uint32_t i = threadIdx.x; i < L; i += 256) {
        const double y = p[i], xi = xc + (double)i;
        sy += y;
        sxy += xi * y;
    }
    sy = block_sum_d(sy, red);
    sxy = block_sum_d(sxy, red);
    // sum of x^2 over the centred abscissae: L (L^2 - 1) / 12
    const double sxx = (double)L * ((double)L * (double)L - 1.0) / 12.0;
    const double a = sy / (double)L;
    const double b = sxx > 0.0 ? sxy / sxx : 0.0;
    double ss = 0.0;
    for (uint32_t i = threadIdx.x; i < L; i += 256) {
        const double r = (double)p[i] - (a + b * (xc + (double)i));
        ss += r * r;
    }
    ss = block_sum_d(ss, red);
    const double rms = sqrt(ss / (double)L);
    for (uint32_t i = threadIdx.x; i < L; i += 256) {
        const double r = (double)p[i] - (a + b * (xc + (double)i));
        p[i] = rms > 0.0 ? (float)(r / rms) : 0.0f;
    }
}

// Segments of at most 1024 samples: one wave per segment, the samples held
// in registers (16 per lane), FP64 sums by wave reductions: the segment is
// read once and written once.  Whole 1024-sample segments of 16-byte aligned
// rows take 16 consecutive samples per lane (four 16-byte loads and stores);
// other segments are lane-strided (256 contiguous bytes per wave and
// instruction).  The normalisation multiplies by the segment's 1/rms.
constexpr int kDtK = 16;  // samples per lane

template <bool kVec>
__device__ __forceinline__ void detrend_wave(float *p, uint32_t L, int lane)
{
    const double xc = -(double)L / 2.0 + 0.5;  // abscissa of sample 0
    auto idx = [&](int k) -> uint32_t {
        return kVec ? (uint32_t)(kDtK * lane + k) : (uint32_t)lane + 64u * (uint32_t)k;
    };
    float y[kDtK];
    if constexpr (kVec) {
#pragma unroll
        for (int q = 0; q < kDtK / 4; ++q) {
            const float4 v = ((const float4 *)(p + kDtK * lane))[q];
            y[4 * q] = v.x;
            y[4 * q + 1] = v.y;
            y[4 * q + 2] = v.z;
            y[4 * q + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kDtK; ++k) y[k] = idx(k) < L ? p[idx(k)] : 0.0f;
    }
    double sy = 0.0, sxy = 0.0;
#pragma unroll
    for (int k = 0; k < kDtK; ++k) {
        const double v = y[k];
        sy += v;
        sxy += (xc + (double)idx(k)) * v;  // zero where the sample is past L
    }
    sy = wave_sum_d(sy);
    sxy = wave_sum_d(sxy);
    const double sxx = (double)L * ((double)L * (double)L - 1.0) / 12.0;
    const double a = sy / (double)L;
    const double b = sxx > 0.0 ? sxy / sxx : 0.0;
    double ss = 0.0;
#pragma unroll
    for (int k = 0; k < kDtK; ++k) {
        const double r = (double)y[k] - (a + b * (xc + (double)idx(k)));
        if (idx(k) < L) ss += r * r;
    }
    ss = wave_sum_d(ss);
    const double rms = sqrt(ss / (double)L);
    const double inv = rms > 0.0 ? 1.0 / rms : 0.0;
    float o[kDtK];
#pragma unroll
    for (int k = 0; k < kDtK; ++k)
        o[k] = (float)(((double)y[k] - (a + b * (xc + (double)idx(k)))) * inv);
    if constexpr (kVec) {
#pragma unroll
        for (int q = 0; q < kDtK / 4; ++q)
            ((float4 *)(p + kDtK * lane))[q] =
                make_float4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
    } else {
#pragma unroll
        for (int k = 0; k < kDtK; ++k)
            if (idx(k) < L) p[idx(k)] = o[k];
    }
}

__global__ void __launch_bounds__(256)
    detrend_wave_kernel(float *x, int64_t stride, uint32_t n, uint32_t seg_len, uint32_t nseg)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t seg = blockIdx.x * 4u + (uint32_t)wave;
    if (seg >= nseg) return;
    const uint32_t row = blockIdx.y;
    const uint32_t s0 = seg * seg_len;
    const uint32_t L = min(seg_len, n - s0);
    float *p = x + (int64_t)row * stride + s0;
    if (L == 64u * kDtK && ((uintptr_t)p & 15u) == 0)
        detrend_wave<true>(p, L, lane);
    else
        detrend_wave<false>(p, L, lane);
}

}  // namespace
}  // namespace mdsm

using namespace mdsm;

extern "C" int mdsm_corner_turn(const void *d_in, int32_t dtype_in, uint32_t nsamp,
                                uint32_t nchans, int64_t in_stride, int32_t invert, void *d_out,
                                int32_t dtype_out, int64_t out_stride, int64_t out_col0,
                                void *stream)
{
    if (!d_in || !d_out) return fail(MDSM_ERR_INVALID, "NULL argument");
    if (dtype_in < MDSM_U8 || dtype_in > MDSM_F32) return fail(MDSM_ERR_INVALID, "bad dtype_in");
    if (dtype_out != dtype_in && dtype_out != MDSM_F32)
        return fail(MDSM_ERR_INVALID, "dtype_out must equal dtype_in or be MDSM_F32");
    if (nsamp == 0 || nchans == 0) return MDSM_OK;
    if (in_stride < (int64_t)nchans) return fail(MDSM_ERR_INVALID, "in_stride < nchans");
    if (out_col0 < 0 || out_stride < out_col0 + (int64_t)nsamp)
        return fail(MDSM_ERR_INVALID, "output rows too short for out_col0 + nsamp");
    if (nchans > 65535u * kCT) return fail(MDSM_ERR_INVALID, "nchans too large");
    hipStream_t s = (hipStream_t)stream;
    TimedLaunch tl("corner_turn", s);
    if (dtype_in == MDSM_U8 && dtype_out == MDSM_U8 && nchans % 16 == 0 &&
        (uintptr_t)d_in % 16 == 0 && in_stride % 16 == 0 &&
        ((uintptr_t)d_out + (uint64_t)out_col0) % 16 == 0 && out_stride % 16 == 0 &&
        (nchans + kCwBlockC - 1) / kCwBlockC <= 65535u) {
        dim3 g16((nsamp + kCwBlockT - 1) / kCwBlockT, (nchans + kCwBlockC - 1) / kCwBlockC);
        hipLaunchKernelGGL(corner_turn_u8x16_kernel<8>, g16, dim3(256), 0, s, (const uint8_t *)d_in,
                           in_stride, nsamp, nchans, invert ? 1 : 0, (uint8_t *)d_out, out_stride,
                           out_col0);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "corner turn launch");
        return MDSM_OK;
    }
    if (dtype_in == MDSM_U8 && dtype_out == MDSM_U8 && nchans % 4 == 0 &&
        (uintptr_t)d_in % 4 == 0 && in_stride % 4 == 0 &&
        ((uintptr_t)d_out + (uint64_t)out_col0) % 16 == 0 && out_stride % 16 == 0 &&
        (nchans + kCtBlockC - 1) / kCtBlockC <= 65535u) {
        dim3 g8((nsamp + kCtBlockT - 1) / kCtBlockT, (nchans + kCtBlockC - 1) / kCtBlockC);
        hipLaunchKernelGGL(corner_turn_u8_kernel, g8, dim3(256), 0, s, (const uint8_t *)d_in,
                           in_stride, nsamp, nchans, invert ? 1 : 0, (uint8_t *)d_out, out_stride,
                           out_col0);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "corner turn launch");
        return MDSM_OK;
    }
    dim3 grid((nsamp + kCT - 1) / kCT, (nchans + kCT - 1) / kCT);
    const int inv = invert ? 1 : 0;
#define MDSM_CT(TI, TO)                                                                        \
    hipLaunchKernelGGL((corner_turn_kernel<TI, TO>), grid, dim3(256), 0, s, (const TI *)d_in,   \
                       in_stride, nsamp, nchans, inv, (TO *)d_out, out_stride, out_col0)
    if (dtype_in == MDSM_U8 && dtype_out == MDSM_U8)
        MDSM_CT(uint8_t, uint8_t);
    else if (dtype_in == MDSM_U8)
        MDSM_CT(uint8_t, float);
    else if (dtype_in == MDSM_F16 && dtype_out == MDSM_F16)
        MDSM_CT(_Float16, _Float16);
    else if (dtype_in == MDSM_F16)
        MDSM_CT(_Float16, float);
    else
        MDSM_CT(float, float);
#undef MDSM_CT
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "corner turn launch");
    return MDSM_OK;
}

extern "C" int mdsm_detrend(float *d_x, uint32_t nrows, uint32_t n, int64_t stride,
                            uint32_t seg_len, void *stream)
{
    if (!d_x) return fail(MDSM_ERR_INVALID, "NULL argument");
    if (seg_len < 2) return fail(MDSM_ERR_INVALID, "seg_len must be >= 2");
    if (stride < (int64_t)n) return fail(MDSM_ERR_INVALID, "stride < n");
    if (nrows == 0 || n == 0) return MDSM_OK;
    if (nrows > 65535) return fail(MDSM_ERR_INVALID, "nrows > 65535");
    hipStream_t s = (hipStream_t)stream;
    TimedLaunch tl("detrend", s);
    const uint32_t nseg = (uint32_t)(((uint64_t)n + seg_len - 1) / seg_len);
    if (seg_len <= 64u * kDtK) {
        dim3 g((nseg + 3) / 4, nrows);
        hipLaunchKernelGGL(detrend_wave_kernel, g, dim3(256), 0, s, d_x, stride, n, seg_len, nseg);
    } else {
        dim3 grid(nseg, nrows);
        hipLaunchKernelGGL(detrend_kernel, grid, dim3(256), 0, s, d_x, stride, n, seg_len);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "detrend launch");
    return MDSM_OK;
}

extern "C" int mdsm_unpack_spectra(const void *d_raw, int32_t nbits, uint32_t nsamp,
                                   uint32_t nchans, int64_t in_stride_bytes, int32_t invert,
                                   void *d_out, int32_t dtype_out, int64_t out_stride,
                                   int64_t out_col0, void *stream)
{
    if (!d_raw || !d_out) return fail(MDSM_ERR_INVALID, "NULL argument");
    if (nbits != 1 && nbits != 2 && nbits != 4 && nbits != 8 && nbits != 16 && nbits != 32)
        return fail(MDSM_ERR_INVALID, "nbits must be 1, 2, 4, 8, 16 or 32");
    if (dtype_out != MDSM_F32 && !(dtype_out == MDSM_U8 && nbits <= 8))
        return fail(MDSM_ERR_INVALID, "dtype_out must be MDSM_F32, or MDSM_U8 for nbits <= 8");
    if (nsamp == 0 || nchans == 0) return MDSM_OK;
    if (in_stride_bytes * 8 < (int64_t)nchans * nbits)
        return fail(MDSM_ERR_INVALID, "in_stride_bytes shorter than a spectrum");
    if (out_col0 < 0 || out_stride < out_col0 + (int64_t)nsamp)
        return fail(MDSM_ERR_INVALID, "output rows too short for out_col0 + nsamp");
    if (nchans > 65535u * kCT) return fail(MDSM_ERR_INVALID, "nchans too large");
    hipStream_t s = (hipStream_t)stream;
    TimedLaunch tl("unpack", s);
    if (dtype_out == MDSM_U8 && nbits <= 8 && nchans % 16 == 0 &&
        (uintptr_t)d_raw % (2 * nbits) == 0 && in_stride_bytes % (2 * nbits) == 0 &&
        ((uintptr_t)d_out + (uint64_t)out_col0) % 16 == 0 && out_stride % 16 == 0 &&
        (nchans + kCwBlockC - 1) / kCwBlockC <= 65535u) {
        dim3 g16((nsamp + kCwBlockT - 1) / kCwBlockT, (nchans + kCwBlockC - 1) / kCwBlockC);
        const uint8_t *r = (const uint8_t *)d_raw;
        uint8_t *o = (uint8_t *)d_out;
        const int inv = invert ? 1 : 0;
        switch (nbits) {
        case 1: hipLaunchKernelGGL(corner_turn_u8x16_kernel<1>, g16, dim3(256), 0, s, r, in_stride_bytes, nsamp, nchans, inv, o, out_stride, out_col0); break;
        case 2: hipLaunchKernelGGL(corner_turn_u8x16_kernel<2>, g16, dim3(256), 0, s, r, in_stride_bytes, nsamp, nchans, inv, o, out_stride, out_col0); break;
        case 4: hipLaunchKernelGGL(corner_turn_u8x16_kernel<4>, g16, dim3(256), 0, s, r, in_stride_bytes, nsamp, nchans, inv, o, out_stride, out_col0); break;
        default: hipLaunchKernelGGL(corner_turn_u8x16_kernel<8>, g16, dim3(256), 0, s, r, in_stride_bytes, nsamp, nchans, inv, o, out_stride, out_col0); break;
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "unpack launch");
        return MDSM_OK;
    }
    dim3 grid((nsamp + kCT - 1) / kCT, (nchans + kCT - 1) / kCT);
    const int inv = invert ? 1 : 0;
    const uint8_t *raw = (const uint8_t *)d_raw;
#define MDSM_UNPACK(NB, TO)                                                                    \
    hipLaunchKernelGGL((unpack_kernel<NB, TO>), grid, dim3(256), 0, s, raw, in_stride_bytes,    \
                       nsamp, nchans, inv, (TO *)d_out, out_stride, out_col0)
    const bool u8 = dtype_out == MDSM_U8;
    switch (nbits) {
    case 1: if (u8) MDSM_UNPACK(1, uint8_t); else MDSM_UNPACK(1, float); break;
    case 2: if (u8) MDSM_UNPACK(2, uint8_t); else MDSM_UNPACK(2, float); break;
    case 4: if (u8) MDSM_UNPACK(4, uint8_t); else MDSM_UNPACK(4, float); break;
    case 8: if (u8) MDSM_UNPACK(8, uint8_t); else MDSM_UNPACK(8, float); break;
    case 16: MDSM_UNPACK(16, float); break;
    default: MDSM_UNPACK(32, float); break;
    }
#undef MDSM_UNPACK
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "unpack launch");
    return MDSM_OK;
}
