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
    dim3 grid((n + seg_len - 1) / seg_len, nrows);
    hipLaunchKernelGGL(detrend_kernel, grid, dim3(256), 0, s, d_x, stride, n, seg_len);
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
