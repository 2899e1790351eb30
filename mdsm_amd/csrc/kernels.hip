// gfx950 kernels of the dedispersion path.
//
// dedisperse_kernel: one workgroup (8 waves) per work item = (block of
//   kDB=64 DM trials sharing one set of input rows, tile of 256 output
//   samples).  Each wave holds 8 trials x 4 samples of accumulators in
//   registers and reads its per-(channel, trial) offsets as wave-uniform
//   scalars.  Integer data (uint8, and the uint16 sums of decimated ranges)
//   is staged a chunk of channels at a time into double-buffered LDS in a
//   four-phase uint16 layout, so that any offset is one aligned ds_read_b64
//   per lane; float data is staged as float32 rows.  Work items are ordered
//   tile-major and remapped so that each XCD (8 per MI355X, round-robin
//   dispatch) receives a contiguous run: blocks sharing an input tile meet in
//   the same L2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <type_traits>

#include "mdsm_internal.h"

namespace mdsm {

// Small fill kernel: used instead of hipMemsetAsync so that the library's
// launch sequences can be captured into HIP graphs.
__global__ void fill_u32_kernel(uint32_t *p, uint32_t n, uint32_t v)
{
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}

hipError_t launch_fill_u32(uint32_t *p, uint32_t n, uint32_t v, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((n + 255) / 256 > 1024 ? 1024 : (n + 255) / 256),
                       dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

// 2D device copy (rows of `width` bytes): a kernel rather than
// hipMemcpy2DAsync, whose device-to-device path runs on a copy engine at a
// fraction of HBM bandwidth.  16-byte words when everything is 16-aligned.
__global__ void copy2d_u4_kernel(uint4 *dst, int64_t dpitch, const uint4 *src, int64_t spitch,
                                 uint32_t words, uint32_t height)
{
    const uint32_t w = blockIdx.x * 256u + threadIdx.x;
    if (w >= words) return;
    for (int64_t r = blockIdx.y; r < height; r += gridDim.y) dst[r * dpitch + w] = src[r * spitch + w];
}

__global__ void copy2d_u8_kernel(uint8_t *dst, int64_t dpitch, const uint8_t *src, int64_t spitch,
                                 uint32_t width, uint32_t height)
{
    for (int64_t r = blockIdx.y; r < height; r += gridDim.y)
        for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < width; i += gridDim.x * 256u)
            dst[r * dpitch + i] = src[r * spitch + i];
}

hipError_t launch_copy2d(void *dst, int64_t dpitch, const void *src, int64_t spitch,
                         int64_t width, uint32_t height, hipStream_t s)
{
    if (width <= 0 || height == 0) return hipSuccess;
    if (width > 0xffffffffLL) return hipErrorInvalidValue;
    const uint32_t gy = std::min<uint32_t>(height, 65535u);
    const bool al = (((uintptr_t)dst | (uintptr_t)src | (uint64_t)dpitch | (uint64_t)spitch |
                      (uint64_t)width) & 15) == 0;
    if (al) {
        const uint32_t words = (uint32_t)(width / 16);
        hipLaunchKernelGGL(copy2d_u4_kernel, dim3((words + 255) / 256, gy), dim3(256), 0, s,
                           (uint4 *)dst, dpitch / 16, (const uint4 *)src, spitch / 16, words,
                           height);
    } else {
        const uint32_t bx = (uint32_t)std::min<int64_t>((width + 255) / 256, 1024);
        hipLaunchKernelGGL(copy2d_u8_kernel, dim3(bx, gy), dim3(256), 0, s, (uint8_t *)dst,
                           dpitch, (const uint8_t *)src, spitch, (uint32_t)width, height);
    }
    return hipGetLastError();
}

// 2D zero fill (rows of `width` bytes), a kernel so that it can be captured.
__global__ void zero2d_kernel(uint8_t *dst, int64_t dpitch, uint32_t width, uint32_t height)
{
    for (int64_t r = blockIdx.y; r < height; r += gridDim.y)
        for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < width; i += gridDim.x * 256u)
            dst[r * dpitch + i] = 0;
}

hipError_t launch_zero2d(void *dst, int64_t dpitch, int64_t width, uint32_t height, hipStream_t s)
{
    if (width <= 0 || height == 0) return hipSuccess;
    if (width > 0xffffffffLL) return hipErrorInvalidValue;
    const uint32_t bx = (uint32_t)std::min<int64_t>((width + 255) / 256, 1024);
    hipLaunchKernelGGL(zero2d_kernel, dim3(bx, std::min<uint32_t>(height, 65535u)), dim3(256), 0,
                       s, (uint8_t *)dst, dpitch, (uint32_t)width, height);
    return hipGetLastError();
}

// Plain single-beam float32 dedispersion in the reference GPU kernel's shape
// (src/mdsm/dedispersion_kernel.cu:51-77, launched at
// src/mdsm/dedispersion_thread.cu:229-231): one thread per (DM, output
// sample), a channel loop of cached global-memory gathers.  It is the
// baseline the staged kernel is compared with, not a product path.
__global__ void __launch_bounds__(256)
    plain_kernel(const float *__restrict__ in, int64_t in_stride, uint32_t nchans,
                 const int32_t *__restrict__ shift, uint32_t nout, float *__restrict__ out)
{
    extern __shared__ int32_t sh[];
    const uint32_t d = blockIdx.y;
    for (uint32_t c = threadIdx.x; c < nchans; c += 256) sh[c] = shift[(int64_t)d * nchans + c];
    __syncthreads();
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= nout) return;
    float acc = 0.0f;
    for (uint32_t c = 0; c < nchans; ++c) acc += in[(int64_t)c * in_stride + sh[c] + t];
    out[(int64_t)d * nout + t] = acc;
}

hipError_t launch_plain(const float *in, int64_t in_stride, uint32_t nchans,
                        const int32_t *shift, uint32_t ndms, uint32_t nout, float *out,
                        hipStream_t s)
{
    if (ndms == 0 || nout == 0) return hipSuccess;
    hipLaunchKernelGGL(plain_kernel, dim3((nout + 255) / 256, ndms), dim3(256),
                       (size_t)nchans * sizeof(int32_t), s, in, in_stride, nchans, shift, nout,
                       out);
    return hipGetLastError();
}

// Loads through the constant address space are scalar (s_load) when the
// address is wave-uniform: the per-(row, DM) offsets are read this way.
typedef __attribute__((address_space(4))) const int32_t cint32_t;
__device__ __forceinline__ const cint32_t *as_const(const int32_t *p)
{
    return (const cint32_t *)(p);
}

// Constant-address-space view of a read-only table: uniform loads through it
// are scalar (s_load); loads through generic pointers never are.
#if defined(__HIP_DEVICE_COMPILE__)
#define MDSM_CONST_AS __attribute__((address_space(4)))
#else
#define MDSM_CONST_AS  // the host pass never runs device code
#endif
template <typename T>
__device__ __forceinline__ const MDSM_CONST_AS T *cptr(const T *p)
{
    return (const MDSM_CONST_AS T *)(p);
}

// A pointer every lane holds the same value of, moved to scalar registers so
// that loads through it compile to s_load.
template <typename T>
__device__ __forceinline__ T *uniform_ptr(T *p)
{
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (T *)(((uint64_t)hi << 32) | lo);
}

struct DedispArgs {
    const SegDesc *segs;
    int64_t items;
    int32_t seg_begin;
    int32_t nseg;
    uint32_t *counter;  // persistent mode when non-NULL
    float one;          // 1.0f (see dedisperse_lds_t)
    int32_t group;      // tiles per item group (>= 1)
    JobBufs jobs[kMaxJobs];
};

// What a work item resolves to: its block, the tables and buffers it uses.
struct ItemCtx {
    const int32_t *rel;
    const int32_t *abase;
    const void *in;
    float *out;
    char *work;
    int64_t in_stride;
};

// Butterfly halving step of a wave reduction of 2H values per lane: lanes
// with bit m set keep the upper H values, the others the lower H, and each
// adds its partner's (lane ^ m) copy of the half it keeps.
template <int H>
__device__ __forceinline__ void bfly_half(double *v, int lane, int m)
{
    const bool up = (lane & m) != 0;
#pragma unroll
    for (int i = 0; i < H; ++i) {
        const double keep = up ? v[H + i] : v[i];
        const double send = up ? v[i] : v[H + i];
        v[i] = keep + __shfl_xor(send, m, 64);
    }
}

// ---- LDS layouts ------------------------------------------------------------
// Every path stages a chunk of channels into LDS in phase-shifted copies, so
// that any per-(channel, DM) offset o is one read per lane of aligned
// addresses that returns the lane's 4 samples:
//  * uint8 / uint16: four phases of 8-byte entries; entry k of phase s packs
//    samples u = 4k+s+{0..3} as 16-bit lanes, .x = x[u] | x[u+2] << 16,
//    .y = x[u+1] | x[u+3] << 16 (one ds_read_b64 at entry (o>>2) + L of
//    phase o&3: lane L holds samples 4L..4L+3 of the tile);
//  * float16: two planes of 16-bit pairs, dword j of plane p =
//    x[2j+p] | x[2j+p+1] << 16; offset o reads dwords (o>>1) + L and
//    (o>>1) + L + 64 of plane o&1 (one ds_read2st64_b32): lane L holds
//    samples 2L, 2L+1, 128+2L, 129+2L.  Half the staged bytes of four phases;
//    the reads run slower (ds_read2st64_b32 75 TB/s against 94 for
//    ds_read_b64 at 4 waves per SIMD), which the integer rows, bound by their
//    reads, do not repay (the three-range uint8 plan of 1024 channels x 2^20
//    samples 46.9 -> 56.0 ms) and float16 rows, bound by their
//    v_fma_mix_f32, do.  (An 8-byte ds_read_b64 at a 4-byte-aligned address
//    runs 19x slower, ds_read2_b32 of adjacent dwords at 40% of the rate.)
//  * float32: two phases of float2 entries, entry j of phase p =
//    (x[2j+p], x[2j+p+1]), read like the float16 planes (ds_read2st64_b64).
//    (Four phases of float4 entries read by ds_read_b128 read faster, but
//    their 16 bytes per staged sample leave wide windows 2 rows per chunk.)
// Integer samples are summed two channels at a time with one v_add3_u32 per
// half into 16-bit lanes, flushed to float32 before they can overflow;
// float16 samples are converted and added in one v_fma_mix_f32 each (x*1+acc,
// a single rounding: the same value as converting and adding); float32
// samples are added directly.  Channels are summed in ascending order.
__device__ __forceinline__ uint32_t perm(uint32_t hi, uint32_t lo, uint32_t sel)
{
    return __builtin_amdgcn_perm(hi, lo, sel);
}

// Global-address-space views.  Loads through generic pointers compile to
// FLAT instructions, which count against LGKM_CNT as well as VM_CNT: every
// s_waitcnt lgkmcnt(0) for a scalar (offset) load would then also wait for
// the next chunk's prefetch.  Through these views they are global_load_*.
#if defined(__HIP_DEVICE_COMPILE__)
#define MDSM_GLOBAL_AS __attribute__((address_space(1)))
#else
#define MDSM_GLOBAL_AS
#endif
template <typename T>
__device__ __forceinline__ const MDSM_GLOBAL_AS T *gptr(const void *p)
{
    return (const MDSM_GLOBAL_AS T *)(p);
}
template <typename T>
__device__ __forceinline__ MDSM_GLOBAL_AS T *gptr_w(void *p)
{
    return (MDSM_GLOBAL_AS T *)(p);
}

template <int DT>
__device__ __forceinline__ uint32_t load_u(const void *base, int64_t i)
{
    if constexpr (DT == DT_U8)
        return gptr<uint8_t>(base)[i];
    else if constexpr (DT == DT_F32)
        return gptr<uint32_t>(base)[i];
    else
        return gptr<uint16_t>(base)[i];
}


template <int DT>
struct Layout {
    static constexpr int es = DT == DT_U8 ? 1 : (DT == DT_F32 ? 4 : 2);  // bytes per sample
    static constexpr int raw = DT == DT_U8 ? 2 : (DT == DT_F32 ? 8 : 4);  // dwords per 8 samples
    static constexpr bool planes = two_planes16(DT);  // two planes of 16-bit pairs
    // bytes between neighbouring lanes' reads (the entry or dword size)
    static constexpr int eb = DT == DT_F32 ? 8 : (planes ? 4 : 8);
    // lane L holds samples 2L, 2L+1, 128+2L, 129+2L of the tile (else 4L..4L+3)
    static constexpr bool split = DT == DT_F32 || planes;
};

// Parts of a channel pair's reads in the pipelined trial loop (halves;
// quarters measured 0.5-0.8% slower on the integer launches).
constexpr int kPipe = 2;

// Tile sample of a lane's accumulator q (0..3).
template <int DT>
__device__ __forceinline__ int lane_sample(int lane, int q)
{
    if constexpr (Layout<DT>::split)
        return 128 * (q >> 1) + 2 * lane + (q & 1);
    else
        return 4 * lane + q;
}

// Sample-by-sample fetch of 8 samples at a row end or an odd row start (zero
// past in_len): out of line, it is rarely taken and would otherwise be
// inlined into every staging loop.
template <int DT>
struct Raw8 {
    uint32_t v[Layout<DT>::raw];
};
template <int DT>
__device__ __attribute__((noinline)) Raw8<DT> fetch8_slow(const void *src, int64_t rowbase, int col,
                                                          int in_len)
{
    Raw8<DT> r;
#pragma unroll
    for (int i = 0; i < Layout<DT>::raw; ++i) r.v[i] = 0u;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const uint32_t v = (col + j < in_len) ? load_u<DT>(src, rowbase + col + j) : 0u;
        if constexpr (DT == DT_U8)
            r.v[j >> 2] |= v << (8 * (j & 3));
        else if constexpr (DT == DT_F32)
            r.v[j] = v;
        else
            r.v[j >> 1] |= v << (16 * (j & 1));
    }
    return r;
}

// Global fetch of 8 consecutive samples (one LDS entry column) into raw
// registers.
template <int DT>
__device__ __forceinline__ void fetch8(const void *src, int64_t rowbase, int col, int in_len,
                                       bool aligned, uint32_t *raw)
{
    constexpr int n = Layout<DT>::raw;
    constexpr int es = Layout<DT>::es;
    if (aligned && col + 8 <= in_len) {
        const MDSM_GLOBAL_AS uint32_t *p = gptr<uint32_t>((const char *)src + (rowbase + col) * es);
#pragma unroll
        for (int i = 0; i < n; ++i) raw[i] = p[i];
    } else if constexpr (DT == DT_F16 || DT == DT_F32) {
        const Raw8<DT> r = fetch8_slow<DT>(src, rowbase, col, in_len);
#pragma unroll
        for (int i = 0; i < n; ++i) raw[i] = r.v[i];
    } else {  // integer data: inline (the out-of-line call costs this path more than it saves)
#pragma unroll
        for (int i = 0; i < n; ++i) raw[i] = 0u;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uint32_t v = (col + j < in_len) ? load_u<DT>(src, rowbase + col + j) : 0u;
            if constexpr (DT == DT_U8)
                raw[j >> 2] |= v << (8 * (j & 3));
            else
                raw[j >> 1] |= v << (16 * (j & 1));
        }
    }
}

// The four 16-bit phase entries of one entry column from its 8 raw samples:
// phase ph is (d0[ph], d1[ph]).  Phase ph+1's first dword is phase ph's
// second (d0[ph+1] == d1[ph]).
template <int DT>
__device__ __forceinline__ void phases16(const uint32_t *raw, uint32_t (&d0)[4], uint32_t (&d1)[4])
{
    if constexpr (DT == DT_U8) {
        const uint32_t a = raw[0], b = raw[1];  // samples 0-3, 4-7
#pragma unroll
        for (int ph = 0; ph < 4; ++ph) {
            d0[ph] = perm(b, a, 0x0c000c00u | (uint32_t)ph | ((uint32_t)(ph + 2) << 16));
            d1[ph] = perm(b, a, 0x0c000c00u | (uint32_t)(ph + 1) | ((uint32_t)(ph + 3) << 16));
        }
    } else {
        const uint32_t h0 = raw[0], h1 = raw[1], h2 = raw[2], h3 = raw[3];
        const uint32_t P0 = perm(h1, h0, 0x05040100u), Q0 = perm(h1, h0, 0x07060302u);
        const uint32_t P1 = perm(h2, h1, 0x05040100u), Q1 = perm(h2, h1, 0x07060302u);
        const uint32_t P2 = perm(h3, h2, 0x05040100u);
        d0[0] = P0; d1[0] = Q0;
        d0[1] = Q0; d1[1] = P1;
        d0[2] = P1; d1[2] = Q1;
        d0[3] = Q1; d1[3] = P2;
    }
}

// The two 16-bit planes of entry column k (samples 4k..4k+4 of its raw
// registers, 16-bit samples): dwords 2k and 2k+1 of plane p are w[p][0],
// w[p][1].
__device__ __forceinline__ void planes16(const uint32_t *raw, uint32_t (&w)[2][2])
{
    const uint32_t h0 = raw[0], h1 = raw[1], h2 = raw[2];  // samples 0-1, 2-3, 4-5
    w[0][0] = h0;
    w[0][1] = h1;
    w[1][0] = perm(h1, h0, 0x05040302u);
    w[1][1] = perm(h2, h1, 0x05040302u);
}

// A lane's read of 4 staged 16-bit samples at p: two dwords 256 bytes apart
// (one ds_read2st64_b32) in the two-plane layout, one ds_read_b64 of a
// four-phase entry otherwise.
template <int DT>
__device__ __forceinline__ uint2 ld16(const char *p)
{
    if constexpr (Layout<DT>::planes) {
        const uint32_t *q = (const uint32_t *)p;
        return make_uint2(q[0], q[64]);
    } else {
        return *(const uint2 *)p;
    }
}

// Write entry column k of a float32 row (samples 4k..4k+4 of its raw
// registers): entries 2k and 2k+1 of both phases (2K entries each).
__device__ __forceinline__ void store_f32_phases(char *row, int K, int k, const uint32_t *raw)
{
    const int K2 = 2 * K;
    uint2 *r2 = (uint2 *)row;
    r2[2 * k] = make_uint2(raw[0], raw[1]);
    r2[2 * k + 1] = make_uint2(raw[2], raw[3]);
    r2[K2 + 2 * k] = make_uint2(raw[1], raw[2]);
    r2[K2 + 2 * k + 1] = make_uint2(raw[3], raw[4]);
}

// A lane's read of 4 staged float32 samples at p: two float2 entries 512
// bytes apart (one ds_read2st64_b64), samples 2L, 2L+1, 128+2L, 129+2L.
__device__ __forceinline__ float4 ld_f32(const char *p)
{
    const float2 *q = (const float2 *)p;
    const float2 lo = q[0], hi = q[64];
    return make_float4(lo.x, lo.y, hi.x, hi.y);
}

// Register prefetch of one chunk: wave w owns rows w, w+kWaves, ... of the
// chunk (<= 8 rows), each up to pf_iters entry columns per lane: K <= 256
// (W <= 1024), or K <= 384 (W <= 1536) for float16.
constexpr int kPfRows = (8 + kWaves - 1) / kWaves;
template <int DT>
constexpr int pf_iters() { return max_window(DT) / 256; }

template <int DT>
using PreRegs = uint32_t[kPfRows][pf_iters<DT>()][Layout<DT>::raw];

// The staged window of one row: its first sample column, and its info --
// bits 0-9 the window's entries per phase (a quarter of its samples), bits
// 10- the row's byte place in its chunk buffer.
struct RowWin {
    int base, info;
};

// Windows of the rows a wave stages in a chunk: a vector load (a scalar one
// would tie the prefetch to the LDS waits).  The integer paths load them a
// chunk ahead of their use, through a global-address-space view; the float
// paths load them just before the prefetch through a generic pointer, whose
// flat load waits for both counters, the scalar offsets of the chunk's first
// row (pair) included (measured faster for them).
template <bool kAhead>
__device__ __forceinline__ void load_colbase(RowWin (&cb)[kPfRows], const int32_t *ab, int cc, int wave)
{
#pragma unroll
    for (int rr = 0; rr < kPfRows; ++rr) {
        const int i = wave + kWaves * rr;
        int2 v = make_int2(0, 0);
        if (i < cc) {
            if constexpr (kAhead)
                v = gptr<int2>(ab)[i];
            else
                v = ((const int2 *)ab)[i];
        }
        cb[rr].base = v.x;
        cb[rr].info = v.y;
    }
}

template <int DT>
__device__ __forceinline__ void fetch_chunk(PreRegs<DT> &pre, const void *src, int64_t in_off,
                                            int64_t row0, int64_t rs, const RowWin (&cb)[kPfRows],
                                            int t0, int cc, int in_len, int wave, int lane)
{
    constexpr int es = Layout<DT>::es;
#pragma unroll
    for (int rr = 0; rr < kPfRows; ++rr) {
        const int i = wave + kWaves * rr;
        if (i >= cc) break;
        const int64_t rowbase = in_off + (row0 + i) * rs;
        const int col0 = cb[rr].base + t0;
        const int K = cb[rr].info & 1023;
        const bool aligned = ((rowbase * es) & 3) == 0;
        // Integer rows whose whole window is inside the row (every item but
        // a row's last) take one address per row, entry column lane + 64*it
        // at a constant offset from it, and no per-column bounds tests.  The
        // test is wave-uniform, but is made per column on a value the
        // compiler must treat as per-lane (the empty asm), so that both
        // branches are exec-masked and load into the same registers: as a
        // scalar branch the compiler merged the two paths by copying the
        // loaded registers, waiting for the loads first.
        constexpr bool kRowFast = DT == DT_U8 || DT == DT_U16;
        int row_fast = kRowFast && aligned && col0 + 4 * K + 4 <= in_len;
        if constexpr (kRowFast) asm volatile("" : "+v"(row_fast));
        const MDSM_GLOBAL_AS uint32_t *q =
            gptr<uint32_t>((const char *)src + (rowbase + col0) * es) + lane * es;
#pragma unroll
        for (int it = 0; it < pf_iters<DT>(); ++it) {
            const int k = lane + 64 * it;
            if (64 * it >= K) break;
            if (k < K) {
                if (row_fast) {
#pragma unroll
                    for (int j = 0; j < Layout<DT>::raw; ++j) pre[rr][it][j] = q[64 * es * it + j];
                } else {
                    fetch8<DT>(src, rowbase, col0 + 4 * k, in_len, aligned, pre[rr][it]);
                }
            }
        }
    }
}

template <int DT>
__device__ __forceinline__ void store_chunk(const PreRegs<DT> &pre, char *buf,
                                            const RowWin (&cb)[kPfRows], int cc, int wave, int lane)
{
#pragma unroll
    for (int rr = 0; rr < kPfRows; ++rr) {
        const int i = wave + kWaves * rr;
        if (i >= cc) break;
        char *row = buf + (cb[rr].info >> 10);
        const int K = cb[rr].info & 1023;
        if constexpr (DT == DT_F32) {
#pragma unroll
            for (int it = 0; it < pf_iters<DT>(); ++it) {
                const int k = lane + 64 * it;
                if (64 * it >= K) break;
                if (k < K) store_f32_phases(row, K, k, pre[rr][it]);
            }
        } else if constexpr (Layout<DT>::planes) {
            // the lane's column in both planes (2W = 8K bytes apart),
            // computed once per row; entry column lane + 64*it sits 512*it
            // bytes further.  Contiguous 8-byte stores: no bank conflicts.
            uint2 *q0 = (uint2 *)(row + 8 * lane);
            uint2 *q1 = (uint2 *)(row + 8 * K + 8 * lane);
#pragma unroll
            for (int it = 0; it < pf_iters<DT>(); ++it) {
                const int k = lane + 64 * it;
                if (64 * it >= K) break;
                if (k < K) {
                    uint32_t w[2][2];
                    planes16(pre[rr][it], w);
                    q0[64 * it] = make_uint2(w[0][0], w[0][1]);
                    q1[64 * it] = make_uint2(w[1][0], w[1][1]);
                }
            }
        } else {
            // the lane's entry in each of the four phase planes (8*K bytes
            // apart), computed once per row; entry column lane + 64*it sits
            // 512*it bytes further.  Each entry is written as two dwords, so
            // the shared dword of neighbouring phases needs no register copy.
            uint32_t *q[4];
            q[0] = (uint32_t *)(row + 8 * lane);
#pragma unroll
            for (int ph = 1; ph < 4; ++ph) q[ph] = (uint32_t *)((char *)q[ph - 1] + 8 * K);
#pragma unroll
            for (int it = 0; it < pf_iters<DT>(); ++it) {
                const int k = lane + 64 * it;
                if (64 * it >= K) break;
                if (k < K) {
                    uint32_t d0[4], d1[4];
                    phases16<DT>(pre[rr][it], d0, d1);
#pragma unroll
                    for (int ph = 0; ph < 4; ++ph) {
                        q[ph][128 * it] = d0[ph];
                        q[ph][128 * it + 1] = d1[ph];
                    }
                }
            }
        }
    }
}

__device__ __forceinline__ float half_lo(uint32_t v)
{
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(v & 0xffffu));
}
__device__ __forceinline__ float half_hi(uint32_t v)
{
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(v >> 16));
}

// Lane k of each row of 16 lanes, to every lane of that row (DPP
// row_newbcast; folded into the consuming VALU instruction by the compiler).
template <int K>
__device__ __forceinline__ int bcast16_c(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, 0x150 + K, 0xf, 0xf, false);
}
__device__ __forceinline__ int bcast16(int v, int k)
{
    switch (k) {
    case 0: return bcast16_c<0>(v);
    case 1: return bcast16_c<1>(v);
    case 2: return bcast16_c<2>(v);
    case 3: return bcast16_c<3>(v);
    case 4: return bcast16_c<4>(v);
    case 5: return bcast16_c<5>(v);
    case 6: return bcast16_c<6>(v);
    case 7: return bcast16_c<7>(v);
    case 8: return bcast16_c<8>(v);
    case 9: return bcast16_c<9>(v);
    case 10: return bcast16_c<10>(v);
    case 11: return bcast16_c<11>(v);
    case 12: return bcast16_c<12>(v);
    case 13: return bcast16_c<13>(v);
    case 14: return bcast16_c<14>(v);
    default: return bcast16_c<15>(v);
    }
}

// Chunks of ccr <= 8 channels are double-buffered in LDS; the next chunk's
// global loads are in flight (in registers) while the current chunk is
// summed.  Each wave takes its kRPW trials' offsets of a channel (pair) from
// per-lane offset registers (integer rows) or scalar loads (float rows) and
// issues all of that channel (pair)'s LDS reads before the adds.  `one` is
// 1.0f, passed at run time so that x*one+acc stays a fused convert-and-add.
template <int DT, int R, bool kStats, bool kFloatCls>
__device__ __forceinline__ void dedisperse_lds_t(const ItemCtx &a, const BlockDesc &bd,
                                                 const BlockDesc *bdp, int t0, int wave, int lane,
                                                 char *lds, float one)
{
    constexpr bool kInt = DT == DT_U8 || DT == DT_U16;
    constexpr int kEB = Layout<DT>::eb;
    const int64_t rs = bd.src == BUF_IN ? a.in_stride : bd.row_stride;
    const void *src = bd.src == BUF_IN ? a.in : (const void *)a.work;
    const int32_t *relb = uniform_ptr(a.rel + bd.tab_off);  // [nrows][kDB] (+8 zero rows)
    const int32_t *ab = uniform_ptr(a.abase + bd.base_off);  // [nrows] (start, info) pairs
    const int ccr = bd.ccr;
    const size_t bufsz = (size_t)bd.bufsz;      // bytes per chunk buffer
    const int nch = (bd.nrows + ccr - 1) / ccr;

    float f[R][4];
    uint32_t ae[R], ao[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        ae[r] = ao[r] = 0u;
#pragma unroll
        for (int j = 0; j < 4; ++j) f[r][j] = 0.0f;
    }
    // accumulators of the 16-bit halves of a read's second (.x hi) and
    // third (.y lo) samples: 2L+1 and 128+2L (two planes), 4L+2 and 4L+1
    // (four phases)
    constexpr int qxh = Layout<DT>::planes ? 1 : 2, qyl = Layout<DT>::planes ? 2 : 1;
    auto flush = [&]() {
        if constexpr (kInt) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                f[r][0] += (float)(ae[r] & 0xffffu);
                f[r][qxh] += (float)(ae[r] >> 16);
                f[r][qyl] += (float)(ao[r] & 0xffffu);
                f[r][3] += (float)(ao[r] >> 16);
                ae[r] = ao[r] = 0u;
            }
        }
    };
    auto add_entry = [&](int r, uint2 x) {  // one 16-bit read into trial r
        if constexpr (DT == DT_F16) {
            f[r][0] = __builtin_fmaf(half_lo(x.x), one, f[r][0]);
            f[r][qxh] = __builtin_fmaf(half_hi(x.x), one, f[r][qxh]);
            f[r][qyl] = __builtin_fmaf(half_lo(x.y), one, f[r][qyl]);
            f[r][3] = __builtin_fmaf(half_hi(x.y), one, f[r][3]);
        } else {
            ae[r] += x.x;
            ao[r] += x.y;
        }
    };
    // sum rows [0, cc) of the staged chunk into the accumulators.  The next
    // row's (pair's) offsets are loaded while this row's LDS reads are in
    // flight, so the scalar-load latency hides under the LDS latency instead
    // of adding to it (the last row's load fetches the next chunk's first
    // rows); the scheduling barriers keep that order.  Offsets past the block
    // read the next block's rows or the zero rows after the segment's tables,
    // never beyond them.
    // oi/oj hold the offsets of the chunk's first row (pair) on entry and of
    // the next chunk's on exit.
    int oi[R], oj[R];
    auto load_soffsets = [&](const int32_t *g) {
        const cint32_t *r0 = as_const(g);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            oi[r] = r0[r];
            if constexpr (DT != DT_F32) oj[r] = r0[kDB + r];
        }
    };
    auto sum_chunk_s = [&](const char *lane_base, const int32_t *grow, int cc) {
        if constexpr (DT == DT_F32) {
            for (int i = 0; i < cc; ++i) {
                float4 x[R];
#pragma unroll
                for (int r = 0; r < R; ++r) x[r] = ld_f32(lane_base + oi[r]);
                __builtin_amdgcn_sched_barrier(0);
                {
                    const cint32_t *rn = as_const(grow + (i + 1) * kDB);
#pragma unroll
                    for (int r = 0; r < R; ++r) oi[r] = rn[r];
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    f[r][0] += x[r].x;
                    f[r][1] += x[r].y;
                    f[r][2] += x[r].z;
                    f[r][3] += x[r].w;
                }
            }
        } else {
            int i = 0;
            for (; i + 1 < cc; i += 2) {
                uint2 x[R], y[R];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    x[r] = ld16<DT>(lane_base + oi[r]);
                    y[r] = ld16<DT>(lane_base + oj[r]);
                }
                __builtin_amdgcn_sched_barrier(0);
                {
                    const cint32_t *rn = as_const(grow + (i + 2) * kDB);
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        oi[r] = rn[r];
                        oj[r] = rn[kDB + r];
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if constexpr (kInt) {
                        ae[r] = ae[r] + x[r].x + y[r].x;
                        ao[r] = ao[r] + x[r].y + y[r].y;
                    } else {
                        add_entry(r, x[r]);
                        add_entry(r, y[r]);
                    }
                }
            }
            if (i < cc) {
                uint2 x[R];
#pragma unroll
                for (int r = 0; r < R; ++r) x[r] = ld16<DT>(lane_base + oi[r]);
                __builtin_amdgcn_sched_barrier(0);
                load_soffsets(grow + cc * kDB);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int r = 0; r < R; ++r) add_entry(r, x[r]);
            }
        }
    };


    // Offsets: a wave's trials of a channel pair sit in one VGPR, lane 16q + m
    // holding (row 2p + m/8, trial m%8) for every q, and a row_newbcast DPP
    // folded into the address add hands lane m's value to every lane (no
    // extra instruction).  A chunk's offset registers (up to 4 pairs) are
    // vector loads issued when the previous chunk's LDS reads are done, in
    // flight across its staging stores and the barrier; they read rows up to
    // c0 + 7: the block's table, the next block's or the zero rows after the
    // segment's tables.
    static_assert(kRPW == 8, "the offset registers hold 2 channels x 8 trials per 16 lanes");
    int vo[4];
    const int orow = (lane >> 3) & 1, ocol = lane & 7;
    auto load_offsets = [&](int (&v)[4], int c0) {
        const MDSM_GLOBAL_AS int32_t *g =
            gptr<int32_t>(relb + (size_t)(c0 + orow) * kDB + wave * kRPW + ocol);
#pragma unroll
        for (int p = 0; p < 4; ++p) v[p] = g[2 * p * kDB];
    };
    // sum rows [0, cc) of the staged chunk into the accumulators
    auto sum_chunk = [&](const char *lane_base, int cc) {
        {
            // A rolled loop over the chunk's full pairs (integer rows; float32
            // rows take the scalar offsets, sum_chunk_s, and float16 rows on
            // per-lane offsets sum_chunk_staged below): the accumulators are
            // loop-carried and updated in place.  (Unrolled with an exit per
            // pair count, the exits' merges cost 16 register copies per pair.)
            // The pair's offset register is vo[0]; the others move down one
            // place per pair.
            int v0 = vo[0], v1 = vo[1], v2 = vo[2], v3 = vo[3];
            const int np = cc >> 1;
            // Pipelined by parts of a pair (kPipe parts of R/kPipe trials):
            // the next pair's reads of a part go into the registers the
            // current pair's adds of that part have just freed, so the
            // LDS queue does not drain between pairs; the data registers
            // stay 2R.
            constexpr int G = kPipe, H = R / kPipe;
            static_assert(R % kPipe == 0, "parts of whole trials");
            uint2 x[R], y[R];
            auto issue = [&](int v, int r0) {
#pragma unroll
                for (int r = r0; r < r0 + H; ++r) {
                    x[r] = ld16<DT>(lane_base + bcast16(v, r));
                    y[r] = ld16<DT>(lane_base + bcast16(v, 8 + r));
                }
            };
            auto add = [&](int r0) {
#pragma unroll
                for (int r = r0; r < r0 + H; ++r) {
                    if constexpr (kInt) {
                        ae[r] = ae[r] + x[r].x + y[r].x;
                        ao[r] = ao[r] + x[r].y + y[r].y;
                    } else {
                        add_entry(r, x[r]);
                        add_entry(r, y[r]);
                    }
                }
            };
            if (np > 0) {
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int g = 0; g < G; ++g) issue(v0, g * H);
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll 1
            for (int pr = 0; pr + 1 < np; ++pr) {
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    add(g * H);
                    __builtin_amdgcn_sched_barrier(0);
                    issue(v1, g * H);
                    __builtin_amdgcn_sched_barrier(0);
                }
                v0 = v1;
                v1 = v2;
                v2 = v3;
            }
            if (np > 0) {
#pragma unroll
                for (int g = 0; g < G; ++g) add(g * H);
                v0 = v1;
            }
            if (cc & 1) {  // the chunk's last row on its own
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int r = 0; r < R; ++r)
                    add_entry(r, ld16<DT>(lane_base + bcast16(v0, r)));
            }
        }
    };

    // The same loop for float16 rows on per-lane offsets, with the next
    // chunk's staging (the other buffer) run between the last pair's first-
    // and second-half adds, its stores queued behind that pair's reads (the
    // two-stage subband plan of 4096 float16 channels 18.0 -> 17.3 ms; for
    // integer rows, bound by their reads, it cost 1.5%).  A lambda of its
    // own: a template parameter on sum_chunk changes the register allocation
    // of the integer instances.
    auto sum_chunk_staged = [&](const char *lane_base, int cc, auto &&stage) {
        {
            // A rolled loop over the chunk's full pairs (integer rows; float32
            // rows take the scalar offsets, sum_chunk_s, and float16 rows on
            // per-lane offsets sum_chunk_staged below): the accumulators are
            // loop-carried and updated in place.  (Unrolled with an exit per
            // pair count, the exits' merges cost 16 register copies per pair.)
            // The pair's offset register is vo[0]; the others move down one
            // place per pair.
            int v0 = vo[0], v1 = vo[1], v2 = vo[2], v3 = vo[3];
            const int np = cc >> 1;
            // Pipelined by parts of a pair (kPipe parts of R/kPipe trials):
            // the next pair's reads of a part go into the registers the
            // current pair's adds of that part have just freed, so the
            // LDS queue does not drain between pairs; the data registers
            // stay 2R.
            constexpr int G = kPipe, H = R / kPipe;
            static_assert(R % kPipe == 0, "parts of whole trials");
            uint2 x[R], y[R];
            auto issue = [&](int v, int r0) {
#pragma unroll
                for (int r = r0; r < r0 + H; ++r) {
                    x[r] = ld16<DT>(lane_base + bcast16(v, r));
                    y[r] = ld16<DT>(lane_base + bcast16(v, 8 + r));
                }
            };
            auto add = [&](int r0) {
#pragma unroll
                for (int r = r0; r < r0 + H; ++r) {
                    if constexpr (kInt) {
                        ae[r] = ae[r] + x[r].x + y[r].x;
                        ao[r] = ao[r] + x[r].y + y[r].y;
                    } else {
                        add_entry(r, x[r]);
                        add_entry(r, y[r]);
                    }
                }
            };
            if (np > 0) {
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int g = 0; g < G; ++g) issue(v0, g * H);
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll 1
            for (int pr = 0; pr + 1 < np; ++pr) {
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    add(g * H);
                    __builtin_amdgcn_sched_barrier(0);
                    issue(v1, g * H);
                    __builtin_amdgcn_sched_barrier(0);
                }
                v0 = v1;
                v1 = v2;
                v2 = v3;
            }
            if (np > 0) add(0);
            __builtin_amdgcn_sched_barrier(0);
            stage();
            __builtin_amdgcn_sched_barrier(0);
            if (np > 0) {
#pragma unroll
                for (int g = 1; g < G; ++g) add(g * H);
                v0 = v1;
            }
            if (cc & 1) {  // the chunk's last row on its own
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int r = 0; r < R; ++r)
                    add_entry(r, ld16<DT>(lane_base + bcast16(v0, r)));
            }
        }
    };

    // float32 rows on per-lane offsets (the kernel built for float data
    // alone): one row of R float4 reads at a time, pipelined by half rows --
    // the next row's reads of trials 0..R/2-1 go into the registers the
    // current row's adds of those trials have just freed; row i is half i&1
    // of offset register i>>1.  The next chunk's staging runs after the
    // chunk (between the last row's halves, as for float16 rows, it measured
    // 1.5-2.5% slower on the 256-channel float32 plan of 512 DMs).
    auto sum_chunk_f32v = [&](const char *lane_base, int cc, auto &&stage) {
        constexpr int H = R / 2;
        int v0 = vo[0], v1 = vo[1], v2 = vo[2], v3 = vo[3];
        const int np = cc >> 1;
        float4 x[R];
        auto issue = [&](int v, auto half, int r0) {
#pragma unroll
            for (int r = r0; r < r0 + H; ++r)
                x[r] = ld_f32(lane_base + bcast16(v, 8 * decltype(half)::value + r));
        };
        auto add = [&](int r0) {
#pragma unroll
            for (int r = r0; r < r0 + H; ++r) {
                f[r][0] += x[r].x;
                f[r][1] += x[r].y;
                f[r][2] += x[r].z;
                f[r][3] += x[r].w;
            }
        };
        using Lo = std::integral_constant<int, 0>;
        using Hi = std::integral_constant<int, 1>;
        // add the row in flight while the next one (register v, half) is issued
        auto step = [&](int v, auto half) {
            add(0);
            __builtin_amdgcn_sched_barrier(0);
            issue(v, half, 0);
            __builtin_amdgcn_sched_barrier(0);
            add(H);
            __builtin_amdgcn_sched_barrier(0);
            issue(v, half, H);
            __builtin_amdgcn_sched_barrier(0);
        };
        if (cc > 0) {
            __builtin_amdgcn_sched_barrier(0);
            issue(v0, Lo{}, 0);
            issue(v0, Lo{}, H);
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll 1
        for (int pr = 0; pr + 1 < np; ++pr) {
            step(v0, Hi{});  // row 2pr added, row 2pr+1 issued
            step(v1, Lo{});  // row 2pr+1 added, row 2pr+2 issued
            v0 = v1;
            v1 = v2;
            v2 = v3;
        }
        if (np > 0) {
            step(v0, Hi{});
            if (cc & 1) step(v1, Lo{});  // the odd last row
        }
        if (cc > 0) {
            add(0);
            add(H);
        }
        __builtin_amdgcn_sched_barrier(0);
        stage();
    };

    // integer and float16 rows take the per-lane offset registers, and so do
    // float32 rows in the kernel built for float data alone; float32 rows in
    // the kernel of all dtypes (mixed batches), the scalar offsets.  (uint16
    // rows on scalar offsets: the three-range uint8 plan of 1024 channels x
    // 2^20 samples 48.7 ms against 46.9 with the registers.  Float16 rows on
    // them in the kernel of all dtypes, with their own pipelined loop: the
    // served 1024-channel uint8 block 4.67 -> 4.66 ms, the 64-beam batch
    // 1.642 -> 1.631 s per step; before the pipelined loops they had cost the
    // served block 0.7%.)
    constexpr bool kVOff = DT == DT_U8 || DT == DT_U16 || DT == DT_F16 || (DT == DT_F32 && kFloatCls);
    int acc_rows = 0;
    PreRegs<DT> pre;
    if constexpr (kVOff)
        load_offsets(vo, 0);
    else
        load_soffsets(relb + wave * kRPW);
    RowWin cb[kPfRows];  // windows of the rows the next prefetch stages
    __syncthreads();  // the previous item (persistent mode) is done with the buffers
    {
        const int cc0 = min(ccr, bd.nrows);
        load_colbase<kVOff>(cb, ab, cc0, wave);
        fetch_chunk<DT>(pre, src, bd.in_off, bd.row0, rs, cb, t0, cc0, bd.in_len, wave, lane);
        store_chunk<DT>(pre, lds, cb, cc0, wave, lane);
        if constexpr (kVOff)
            if (nch > 1) load_colbase<true>(cb, ab + 2 * ccr, min(ccr, bd.nrows - ccr), wave);
    }
    __syncthreads();
    for (int ch = 0; ch < nch; ++ch) {
        const int c0 = ch * ccr;
        const int cc = min(ccr, bd.nrows - c0);
        const bool more = ch + 1 < nch;
        const int c1 = c0 + ccr;
        const int cc1 = more ? min(ccr, bd.nrows - c1) : 0;
        // integer rows: every vector load issued before this point (this chunk's
        // offset registers, the window starts) has landed by now; waiting
        // here, on every path, keeps the trial loop's reads of the offset
        // registers from waiting for the prefetch below (the counter is in
        // order).
        if constexpr (kVOff) __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0)
        // the windows of the rows the prefetch below stages, kept (wave-
        // uniform, in scalar registers) for their staging stores: the integer
        // paths reload cb for the chunk after before those stores
        RowWin cs[kPfRows];
        if (more) {  // next chunk's global loads overlap this chunk's compute
            if constexpr (!kVOff) load_colbase<false>(cb, ab + 2 * c1, cc1, wave);
#pragma unroll
            for (int rr = 0; rr < kPfRows; ++rr) {
                cs[rr].base = __builtin_amdgcn_readfirstlane(cb[rr].base);
                cs[rr].info = __builtin_amdgcn_readfirstlane(cb[rr].info);
            }
            fetch_chunk<DT>(pre, src, bd.in_off, bd.row0 + c1, rs, cs, t0, cc1, bd.in_len,
                            wave, lane);
        }
        if constexpr (kInt) {
            if (acc_rows + cc > bd.flush) {
                flush();
                acc_rows = 0;
            }
        }
        if constexpr (kVOff && (DT == DT_F16 || DT == DT_F32)) {
            auto stage = [&] {
                if (more) {
                    if (ch + 2 < nch)
                        load_colbase<true>(cb, ab + 2 * (c1 + ccr), min(ccr, bd.nrows - c1 - ccr),
                                           wave);
                    load_offsets(vo, c1);
                    store_chunk<DT>(pre, lds + ((ch + 1) & 1) * bufsz, cs, cc1, wave, lane);
                }
            };
            if constexpr (DT == DT_F16)
                sum_chunk_staged(lds + (ch & 1) * bufsz + lane * kEB, cc, stage);
            else
                sum_chunk_f32v(lds + (ch & 1) * bufsz + lane * kEB, cc, stage);
            acc_rows += cc;
            __syncthreads();
            continue;
        }
        if constexpr (kVOff)
            sum_chunk(lds + (ch & 1) * bufsz + lane * kEB, cc);
        else
            sum_chunk_s(lds + (ch & 1) * bufsz + lane * kEB, relb + (size_t)c0 * kDB + wave * kRPW,
                        cc);
        acc_rows += cc;
        if (more) {
            // in flight across the staging stores and the barrier
            if constexpr (kVOff) {
                if (ch + 2 < nch)
                    load_colbase<true>(cb, ab + 2 * (c1 + ccr), min(ccr, bd.nrows - c1 - ccr),
                                       wave);
                load_offsets(vo, c1);
            }
            store_chunk<DT>(pre, lds + ((ch + 1) & 1) * bufsz, cs, cc1, wave, lane);
        }
        __syncthreads();
    }
    flush();

    // Row statistics of integer full-rate rows (launches with stats on): the
    // sum and sum of squares of each trial's samples t < stat_n in this tile.
    // The values are integers below 2^24, so both sums are exact in FP64 in
    // any order.  A butterfly over the wave leaves value v (trial v's sum for
    // v < 8, trial v-8's sum of squares for v >= 8) in lanes 4v..4v+3.
    // (The statistics fields are read here, through the descriptor's
    // pointer, so that they hold no scalar registers across the chunk loop.)
    if constexpr (kInt && kStats) {
        const int tile_i = t0 / kTile;
        const MDSM_CONST_AS BlockDesc *sq = cptr(bdp);
        const int stat_tiles = sq->stat_tiles;
        if (tile_i < stat_tiles) {
            const int stat_n = sq->stat_n;
            static_assert(R == 8, "the butterfly reduces 2 x 8 values per lane");
            double v[2 * R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                double s1 = 0.0, s2 = 0.0;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const double x =
                        (t0 + lane_sample<DT>(lane, q) < stat_n) ? (double)f[r][q] : 0.0;
                    s1 += x;
                    s2 += x * x;
                }
                v[r] = s1;
                v[R + r] = s2;
            }
            bfly_half<8>(v, lane, 32);
            bfly_half<4>(v, lane, 16);
            bfly_half<2>(v, lane, 8);
            bfly_half<1>(v, lane, 4);
            v[0] += __shfl_xor(v[0], 2, 64);
            v[0] += __shfl_xor(v[0], 1, 64);
            const int vi = lane >> 2;
            const int j = wave * R + (vi & 7);
            if ((lane & 3) == 0 && j < bd.ntrials) {
                MDSM_GLOBAL_AS double *st = gptr_w<double>(a.work + sq->stat_off);
                st[2 * ((int64_t)j * stat_tiles + tile_i) + (vi >> 3)] = v[0];
            }
        }
    }

    float *ob = (bd.dst == BUF_OUT ? a.out : (float *)a.work) + bd.out_off;
    const float kf = (float)bd.k;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = wave * R + r;
        if (j >= bd.ntrials) continue;
        float v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = bd.k > 1 ? __fdiv_rn(f[r][q], kf) : f[r][q];
        float *og = ob + (int64_t)j * bd.out_trial_stride;
        MDSM_GLOBAL_AS float *o = gptr_w<float>(og);
        if constexpr (Layout<DT>::split) {
            // samples t0 + 2L + {0, 1} and t0 + 128 + 2L + {0, 1}
            const bool al = (((uintptr_t)(og + t0)) & 7) == 0;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int t = t0 + 128 * h + 2 * lane;
                if (al && t + 1 < bd.nout) {
                    *(MDSM_GLOBAL_AS float2 *)(o + t) = make_float2(v[2 * h], v[2 * h + 1]);
                } else {
                    if (t < bd.nout) o[t] = v[2 * h];
                    if (t + 1 < bd.nout) o[t + 1] = v[2 * h + 1];
                }
            }
        } else {
            const int t = t0 + 4 * lane;
            const bool al = (((uintptr_t)(og + t0)) & 15) == 0;
            if (al && t + 3 < bd.nout) {
                *(MDSM_GLOBAL_AS float4 *)(o + t) = make_float4(v[0], v[1], v[2], v[3]);
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (t + q < bd.nout) o[t + q] = v[q];
            }
        }
    }
}

// Kernel classes: the dtypes a dedispersion kernel is built for.  Each class
// is its own kernel with its own register allocation, so the integer path's
// offset registers do not crowd the float paths: launches of float blocks
// only take the float kernel (the two-stage subband plan of 4096 float16
// channels x 2^19 samples, 4096 DMs: 20.1 -> 19.6 ms), every other launch the
// kernel of all dtypes (an integer-only kernel was about 1% slower than it on
// the three-range uint8 plan of 1024 channels x 2^20 samples).
constexpr int kClsInt = 1, kClsFloat = 2, kClsAll = 3;
// Mixed launches (integer and float rows) without row statistics: the kernel
// of all dtypes, with float32 rows on the per-lane offsets as in the float
// kernel.  (The 64-beam batch, whose decimated float16 ranges are float32
// rows: 1.630 -> 1.574 s per step.  The same change in the kernel that
// integer-only and served launches take cost them 0.5-1.9%.)
constexpr int kClsMixed = 7;

template <int kCls, bool kStats>
__device__ __forceinline__ void dedisperse_item(const DedispArgs &a, int64_t f, int wave,
                                                int lane, char *lds)
{
    // Segment lookup (wave-uniform binary search over item_start).
    int lo = a.seg_begin, hi = a.seg_begin + a.nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cptr(a.segs)[mid].item_start <= f)
            lo = mid;
        else
            hi = mid - 1;
    }
    lo = __builtin_amdgcn_readfirstlane(lo);
    const SegDesc sg = *cptr(uniform_ptr(a.segs + lo));
    // Items of a segment run in groups of G consecutive tiles; within a group
    // the G tiles of one DM block are consecutive items, so the block's offset
    // table is re-read from L2 by neighbouring items (G = 1: tile-major).
    const int64_t fl = f - sg.item_start;
    const int64_t gsz = (int64_t)a.group * sg.nblocks;
    const int64_t g = fl / gsz;
    const int64_t r = fl - g * gsz;
    const int gt = (int)min<int64_t>(a.group, sg.ntiles - g * a.group);  // tiles in this group
    const int tile_i = __builtin_amdgcn_readfirstlane((int)(g * a.group + r % gt));
    const int bi = __builtin_amdgcn_readfirstlane(sg.first_block + (int)(r / gt));
    const BlockDesc *bdp = uniform_ptr(sg.blocks + bi);
    const BlockDesc bd = *cptr(bdp);
    const JobBufs jb = a.jobs[sg.job];
    ItemCtx ctx{uniform_ptr(sg.rel), uniform_ptr(sg.abase), uniform_ptr(jb.in),
                uniform_ptr(jb.out), uniform_ptr(jb.work), jb.in_stride};
    const int t0 = tile_i * kTile;
    // float32 rows on per-lane offsets: the kernel built for float data alone,
    // and the one for mixed launches
    constexpr bool kFloat = kCls == kClsFloat || kCls == kClsMixed;
    // float blocks of at most kDB/2 trials run kRPW/2 trials per wave
    if constexpr ((kCls & kClsInt) != 0) {
        if (bd.dtype == DT_U8) {
            dedisperse_lds_t<DT_U8, kRPW, kStats, kFloat>(ctx, bd, bdp, t0, wave, lane, lds,
                                                          a.one);
            return;
        }
        if (bd.dtype == DT_U16) {
            dedisperse_lds_t<DT_U16, kRPW, kStats, kFloat>(ctx, bd, bdp, t0, wave, lane, lds,
                                                           a.one);
            return;
        }
    }
    if constexpr ((kCls & kClsFloat) != 0) {
        constexpr int H = kRPW / 2;
        if (bd.rpw == kRPW) {
            if (bd.dtype == DT_F16)
                dedisperse_lds_t<DT_F16, kRPW, kStats, kFloat>(ctx, bd, bdp, t0, wave, lane, lds,
                                                               a.one);
            else
                dedisperse_lds_t<DT_F32, kRPW, kStats, kFloat>(ctx, bd, bdp, t0, wave, lane, lds,
                                                               a.one);
        } else {  // integer blocks always run kRPW trials per wave
            if (bd.dtype == DT_F16)
                dedisperse_lds_t<DT_F16, H, kStats, kFloat>(ctx, bd, bdp, t0, wave, lane, lds,
                                                            a.one);
            else
                dedisperse_lds_t<DT_F32, H, kStats, kFloat>(ctx, bd, bdp, t0, wave, lane, lds,
                                                            a.one);
        }
    }
}

#ifndef MDSM_DEDISP_WAVES
#define MDSM_DEDISP_WAVES 4
#endif
template <int kCls, bool kStats>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(MDSM_DEDISP_WAVES)))
dedisperse_kernel(DedispArgs a)
{
    extern __shared__ __attribute__((aligned(16))) char lds_raw[];
    __shared__ int64_t sh_item;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (a.counter == nullptr) {
        // One workgroup per item; XCD-aware remap: XCD x = blockIdx % 8 gets a
        // contiguous run of items, so blocks sharing an input tile share an L2.
        const int64_t g = blockIdx.x;
        const int64_t q = a.items >> 3, rem = a.items & 7;
        const int64_t xcd = g & 7, loc = g >> 3;
        const int64_t f = xcd * q + (xcd < rem ? xcd : rem) + loc;
        dedisperse_item<kCls, kStats>(a, f, wave, lane, lds_raw);
        return;
    }
    // Persistent mode: workgroups pull items from a device counter until the
    // queue is empty; every workgroup reaches the exit test.
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) sh_item = (int64_t)atomicAdd(a.counter, 1u);
        __syncthreads();
        const int64_t fv = sh_item;
        // the item index is the same in every lane: make that explicit so the
        // descriptor and table loads that depend on it stay scalar
        const uint32_t flo = __builtin_amdgcn_readfirstlane((uint32_t)fv);
        const uint32_t fhi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)fv >> 32));
        const int64_t f = (int64_t)(((uint64_t)fhi << 32) | flo);
        if (f >= a.items) break;
        dedisperse_item<kCls, kStats>(a, f, wave, lane, lds_raw);
    }
}

hipError_t launch_dedisperse_spec(const LaunchSpec &L, hipStream_t s)
{
    if (L.items <= 0) return hipSuccess;
    if (L.items > 0x7fffffffLL || L.njobs > kMaxJobs) return hipErrorInvalidValue;
    DedispArgs a{};
    a.segs = L.segs;
    a.items = L.items;
    a.seg_begin = L.seg_begin;
    a.nseg = L.nseg;
    a.counter = L.counter;
    a.one = 1.0f;
    a.group = std::max(1, L.group);
    for (int i = 0; i < L.njobs; ++i) a.jobs[i] = L.jobs[i];
    const size_t lds = (size_t)L.lds_bytes;
    if (lds > (size_t)kLdsMax) return hipErrorInvalidValue;
    const bool ints = (L.dtypes & ((1 << DT_U8) | (1 << DT_U16))) != 0;
    const bool floats = (L.dtypes & ((1 << DT_F16) | (1 << DT_F32))) != 0;
    // launches with row statistics (the server's) take a kernel instance of
    // their own, so that the statistics epilogue costs the others nothing
    const int cls = (floats && !ints) ? kClsFloat
                    : L.stats         ? kClsAll + 1
                    : floats          ? kClsMixed
                                      : kClsAll;
    void (*kern)(DedispArgs) = cls == kClsFloat   ? dedisperse_kernel<kClsFloat, false>
                               : cls == kClsAll   ? dedisperse_kernel<kClsAll, false>
                               : cls == kClsMixed ? dedisperse_kernel<kClsMixed, false>
                                                  : dedisperse_kernel<kClsAll, true>;
    // the full gfx950 LDS to one workgroup, once per kernel (host threads may
    // launch concurrently; setting the attribute twice is harmless)
    static std::atomic<bool> lds_attr[8];
    if (!lds_attr[cls].load(std::memory_order_acquire)) {
        hipError_t e = hipFuncSetAttribute((const void *)kern,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, kLdsMax);
        if (e != hipSuccess) return e;
        lds_attr[cls].store(true, std::memory_order_release);
    }
    unsigned grid = (unsigned)L.items;
    // A persistent grid at least as large as the queue would give every
    // workgroup one item anyway: launch one workgroup per item instead, which
    // keeps the XCD-aware item order and needs no counter.
    const int64_t pgrid = L.grid > 0 ? L.grid : 2048;
    if (L.counter && L.items > pgrid) {
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, s, L.counter, 1u, 0u);
        grid = (unsigned)pgrid;
    } else {
        a.counter = nullptr;
        // every item is resident at once: tile-major order lets the blocks of
        // one tile, running side by side, share their input windows in L2
        if (L.counter) a.group = 1;
    }
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kThreads), lds, s, a);
    return hipGetLastError();
}

hipError_t dedisperse_occupancy(int32_t lds_bytes, int32_t dtypes, DedispOccupancy *o)
{
    const bool ints = (dtypes & ((1 << DT_U8) | (1 << DT_U16))) != 0;
    const bool floats = (dtypes & ((1 << DT_F16) | (1 << DT_F32))) != 0;
    const void *kern = (floats && !ints) ? (const void *)dedisperse_kernel<kClsFloat, false>
                                         : (const void *)dedisperse_kernel<kClsAll, false>;
    hipFuncAttributes at{};
    hipError_t e = hipFuncGetAttributes(&at, kern);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsMax);
    if (e != hipSuccess) return e;
    int nb = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, kThreads, (size_t)lds_bytes);
    if (e != hipSuccess) return e;
    o->workgroups_per_cu = nb;
    o->waves_per_simd = nb * kWaves / 4;
    o->vgprs = at.numRegs;
    o->static_lds = (int32_t)at.sharedSizeBytes;
    o->dynamic_lds = lds_bytes;
    o->scratch_bytes = (int32_t)at.localSizeBytes;
    return hipSuccess;
}

hipError_t launch_dedisperse(const mdsm_plan *p, const Stage &st, const void *in,
                             int64_t in_stride, float *out, void *work, hipStream_t s)
{
    return launch_dedisperse_ctr(p, st, in, in_stride, out, work,
                                 p->dev.counter + (&st - p->stages.data()), s);
}

hipError_t launch_dedisperse_ctr(const mdsm_plan *p, const Stage &st, const void *in,
                                 int64_t in_stride, float *out, void *work, uint32_t *counter,
                                 hipStream_t s)
{
    LaunchSpec L;
    L.segs = p->dev.segs;
    L.seg_begin = st.seg_begin;
    L.nseg = st.seg_end - st.seg_begin;
    L.items = st.items;
    L.lds_bytes = st.lds_bytes;
    L.dtypes = st.dtypes;
    L.group = (int32_t)p->tune.tile_group;
    if (p->tune.persistent) {
        L.counter = counter;
        L.grid = p->tune.grid ? (int32_t)p->tune.grid : 8 * std::max(1, p->cus);
    }
    L.njobs = 1;
    L.jobs[0] = JobBufs{in, out, (char *)work, in_stride};
    return launch_dedisperse_spec(L, s);
}

// ---------------------------------------------------------------------------
// Decimation: out[c][t] = sum_{j<k} in[c][k t + j]
// ---------------------------------------------------------------------------

template <int DT>
__device__ __forceinline__ float load_sample(const void *base, int64_t i)
{
    if constexpr (DT == DT_F16)
        return (float)((const _Float16 *)base)[i];
    else
        return ((const float *)base)[i];
}

template <int DTI, int DTO>
__global__ void __launch_bounds__(256)
    decimate_kernel(const void *in, int64_t in_stride, uint32_t nsamp_out, uint32_t k, void *out,
                    int64_t out_stride)
{
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    const int64_t c = blockIdx.y;
    if (t >= nsamp_out) return;
    const int64_t base = c * in_stride + (int64_t)t * k;
    if constexpr (DTI == DT_U8) {
        uint32_t acc = 0;
        for (uint32_t j = 0; j < k; ++j) acc += ((const uint8_t *)in)[base + j];
        if constexpr (DTO == DT_U16)
            ((uint16_t *)out)[c * out_stride + t] = (uint16_t)acc;
        else
            ((float *)out)[c * out_stride + t] = (float)acc;
    } else {
        float acc = 0.0f;
        for (uint32_t j = 0; j < k; ++j) acc += load_sample<DTI>(in, base + j);
        ((float *)out)[c * out_stride + t] = acc;
    }
}

// uint8 -> uint16 k-sums for every k in {2, 4, 8, 16} a plan needs, in one
// pass: a thread loads 16 consecutive samples (one 16-byte load) and forms the
// pairwise tree s2, s4, s8, s16.  Integer sums are exact in any order, so each
// level equals the sequential k-sum.  Output i of level k covers input samples
// [k i, k i + k); only i < floor(nsamp / k) is written.
struct PyrLevel {
    uint16_t *out;
    int64_t stride;   // elements per output row
    uint32_t nout;
    uint32_t lg;      // log2 k, 1..4
};
struct PyrArgs {
    const uint8_t *in;
    int64_t in_stride;
    uint32_t nsamp;
    int32_t nlev;
    int32_t vec_in;   // 16-byte aligned rows: vector loads
    PyrLevel lev[4];
};

__global__ void __launch_bounds__(256) decimate_u8_pyramid_kernel(PyrArgs a)
{
    const uint32_t g = blockIdx.x * 256u + threadIdx.x;
    const int64_t c = blockIdx.y;
    const uint32_t col = g * 16u;
    if (col >= a.nsamp) return;
    const uint8_t *row = a.in + c * a.in_stride;
    uint32_t w[4];
    if (a.vec_in && col + 16u <= a.nsamp) {
        const uint4 v = *(const uint4 *)(row + col);
        w[0] = v.x;
        w[1] = v.y;
        w[2] = v.z;
        w[3] = v.w;
    } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            w[q] = 0u;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const uint32_t t = col + 4u * q + b;
                if (t < a.nsamp) w[q] |= (uint32_t)row[t] << (8 * b);
            }
        }
    }
    // s2 as 16-bit pairs: p[q] = (b0+b1) | (b2+b3) << 16 of dword q
    uint32_t s2[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) s2[q] = (w[q] & 0x00ff00ffu) + ((w[q] >> 8) & 0x00ff00ffu);
    uint32_t s4[4];  // 4-sums of samples 4q..4q+3 as plain integers
#pragma unroll
    for (int q = 0; q < 4; ++q) s4[q] = (s2[q] & 0xffffu) + (s2[q] >> 16);
    const uint32_t s8[2] = {s4[0] + s4[1], s4[2] + s4[3]};
    const uint32_t s16 = s8[0] + s8[1];
    for (int l = 0; l < a.nlev; ++l) {
        const PyrLevel L = a.lev[l];
        uint16_t *o = L.out + c * L.stride;
        const uint32_t i0 = col >> L.lg;  // first output of this thread at level l
        const uint32_t n = 16u >> L.lg;   // outputs per thread
        if (i0 + n <= L.nout) {
            if (L.lg == 1)
                *(uint4 *)(o + i0) = make_uint4(s2[0], s2[1], s2[2], s2[3]);
            else if (L.lg == 2)
                *(uint2 *)(o + i0) = make_uint2(s4[0] | (s4[1] << 16), s4[2] | (s4[3] << 16));
            else if (L.lg == 3)
                *(uint32_t *)(o + i0) = s8[0] | (s8[1] << 16);
            else
                o[i0] = (uint16_t)s16;
        } else {
            for (uint32_t j = 0; j < n && i0 + j < L.nout; ++j) {
                uint32_t v;
                if (L.lg == 1)
                    v = (j & 1) ? (s2[j >> 1] >> 16) : (s2[j >> 1] & 0xffffu);
                else if (L.lg == 2)
                    v = s4[j];
                else if (L.lg == 3)
                    v = s8[j];
                else
                    v = s16;
                o[i0 + j] = (uint16_t)v;
            }
        }
    }
}

hipError_t launch_decimate_steps(const void *in, uint32_t nchans, int64_t in_stride,
                                 uint32_t nsamp, const DecimStep *steps, size_t nsteps, char *work,
                                 hipStream_t s)
{
    PyrArgs a{};
    a.in = (const uint8_t *)in;
    a.in_stride = in_stride;
    a.nsamp = nsamp;
    a.vec_in = ((uintptr_t)in % 16 == 0 && in_stride % 16 == 0) ? 1 : 0;
    for (size_t i = 0; i < nsteps; ++i) {
        const DecimStep &d = steps[i];
        const int k = d.k;
        const bool pyr = d.dtype_in == DT_U8 && d.dtype_out == DT_U16 &&
                         (k == 2 || k == 4 || k == 8 || k == 16) && a.nlev < 4 &&
                         ((uintptr_t)(work + d.work_off) % 16) == 0 && d.stride % 8 == 0;
        if (pyr) {
            PyrLevel &L = a.lev[a.nlev++];
            L.out = (uint16_t *)(work + d.work_off);
            L.stride = d.stride;
            L.nout = nsamp / (uint32_t)k;
            L.lg = k == 2 ? 1 : k == 4 ? 2 : k == 8 ? 3 : 4;
        } else {
            hipError_t e = launch_decimate(in, d.dtype_in, nchans, in_stride, nsamp, (uint32_t)k,
                                           work + d.work_off, d.dtype_out, d.stride, s);
            if (e != hipSuccess) return e;
        }
    }
    if (a.nlev == 0 || nchans == 0 || nsamp == 0) return hipSuccess;
    dim3 grid(((nsamp + 15) / 16 + 255) / 256, nchans);
    hipLaunchKernelGGL(decimate_u8_pyramid_kernel, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_decimate(const void *in, int dtype_in, uint32_t nchans, int64_t in_stride,
                           uint32_t nsamp, uint32_t k, void *out, int dtype_out,
                           int64_t out_stride, hipStream_t s)
{
    const uint32_t nout = nsamp / k;
    if (nout == 0 || nchans == 0) return hipSuccess;
    dim3 grid((nout + 255) / 256, nchans);
    if (dtype_in == DT_U8 && dtype_out == DT_U16)
        hipLaunchKernelGGL((decimate_kernel<DT_U8, DT_U16>), grid, dim3(256), 0, s, in, in_stride,
                           nout, k, out, out_stride);
    else if (dtype_in == DT_U8 && dtype_out == DT_F32)
        hipLaunchKernelGGL((decimate_kernel<DT_U8, DT_F32>), grid, dim3(256), 0, s, in, in_stride,
                           nout, k, out, out_stride);
    else if (dtype_in == DT_F16 && dtype_out == DT_F32)
        hipLaunchKernelGGL((decimate_kernel<DT_F16, DT_F32>), grid, dim3(256), 0, s, in,
                           in_stride, nout, k, out, out_stride);
    else if (dtype_in == DT_F32 && dtype_out == DT_F32)
        hipLaunchKernelGGL((decimate_kernel<DT_F32, DT_F32>), grid, dim3(256), 0, s, in,
                           in_stride, nout, k, out, out_stride);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Wave-level helpers (64 lanes)
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, int lane)
{
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

__device__ __forceinline__ double wave_sum_f64(double v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Rank query on a 256-bin LDS histogram: the smallest value v whose
// cumulative count exceeds `rank`.  Executed by wave 0 (4 bins per lane).
__device__ int hist_rank(const uint32_t *hist, uint32_t rank, int lane)
{
    const uint32_t h0 = hist[4 * lane], h1 = hist[4 * lane + 1], h2 = hist[4 * lane + 2],
                   h3 = hist[4 * lane + 3];
    const uint32_t mine = h0 + h1 + h2 + h3;
    const uint32_t incl = wave_incl_scan(mine, lane);
    const uint32_t excl = incl - mine;
    // lane holding the answer: excl <= rank < incl
    const bool here = excl <= rank && rank < incl;
    int v = 255;
    if (here) {
        uint32_t acc = excl + h0;
        v = 4 * lane;
        if (acc <= rank) {
            acc += h1;
            v++;
            if (acc <= rank) {
                acc += h2;
                v++;
                if (acc <= rank) v++;
            }
        }
    }
    // broadcast: the unique lane with `here`, found via ballot
    const uint64_t m = __ballot(here);
    const int src = m ? (int)__ffsll((unsigned long long)m) - 1 : 0;
    return __shfl(v, src, 64);
}

// ---------------------------------------------------------------------------
// RFI clipper (uint8): one workgroup per channel.
// ---------------------------------------------------------------------------

constexpr int kClipRegs = 16;  // 16-byte words per thread kept in registers (65536 samples a row)

__global__ void __launch_bounds__(256)
    clip_u8_kernel(const uint8_t *in, int64_t in_stride, int64_t t0, uint32_t n, float nsigma,
                   uint8_t *out, int64_t out_stride, int32_t *med_out, int32_t *mad_out,
                   int32_t *ncl_out)
{
    __shared__ uint32_t hist[256];
    __shared__ uint32_t whist[4][256];  // one histogram per wave: a quarter of the atomic contention
    __shared__ int sh_med, sh_mad;
    __shared__ uint32_t sh_cnt;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t c = blockIdx.x;
    const uint8_t *x = in + c * in_stride + t0;
    uint8_t *o = out + c * out_stride + t0;
    const uint32_t rank = (n - 1) / 2;

    // 16-byte loads and stores when both rows allow them (uniform per
    // workgroup); the remaining n % 16 samples one at a time.  The first
    // kClipRegs 16-byte words of each thread (rows up to 65536 samples: all
    // of them) stay in registers from the histogram pass to the clip pass,
    // so the row is read once.
    const bool vec = ((uintptr_t)x % 16 == 0) && ((uintptr_t)o % 16 == 0);
    const uint32_t nv = vec ? n / 16u : 0u;
    uint4 keep[kClipRegs];
#pragma unroll
    for (int w = 0; w < 4; ++w) whist[w][tid] = 0;
    if (tid == 0) sh_cnt = 0;
    __syncthreads();
    auto hist_word = [&](const uint4 &w) {
        const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int k = 0; k < 16; ++k) atomicAdd(&whist[wv][(ws[k >> 2] >> (8 * (k & 3))) & 0xffu], 1u);
    };
#pragma unroll
    for (int i = 0; i < kClipRegs; ++i) {
        const uint32_t q = tid + 256u * i;
        keep[i] = make_uint4(0u, 0u, 0u, 0u);
        if (q < nv) {
            keep[i] = ((const uint4 *)x)[q];
            hist_word(keep[i]);
        }
    }
    for (uint32_t q = tid + 256u * kClipRegs; q < nv; q += 256) hist_word(((const uint4 *)x)[q]);
    for (uint32_t t = 16u * nv + tid; t < n; t += 256) atomicAdd(&whist[wv][x[t]], 1u);
    __syncthreads();
    hist[tid] = (whist[0][tid] + whist[1][tid]) + (whist[2][tid] + whist[3][tid]);
    __syncthreads();
    if (tid < 64) {
        const int m = hist_rank(hist, rank, lane);
        if (lane == 0) sh_med = m;
    }
    __syncthreads();
    const int med = sh_med;
    // the histogram of |x - med| follows from the value histogram: no second
    // pass over the row
    {
        const int d = tid;
        const uint32_t up = med + d <= 255 ? hist[med + d] : 0u;
        const uint32_t dn = (d > 0 && med - d >= 0) ? hist[med - d] : 0u;
        whist[0][d] = up + dn;
    }
    __syncthreads();
    if (tid < 64) {
        const int m = hist_rank(whist[0], rank, lane);
        if (lane == 0) sh_mad = m;
    }
    __syncthreads();
    const int mad = sh_mad;
    const float thr = (nsigma * 1.4826f) * (float)mad;
    uint32_t mine = 0;
    auto clip_word = [&](uint4 w, uint32_t q) {
        uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int sh = 8 * (k & 3);
            const int v = (int)((ws[k >> 2] >> sh) & 0xffu);
            if ((float)abs(v - med) > thr) {
                ++mine;
                ws[k >> 2] = (ws[k >> 2] & ~(0xffu << sh)) | ((uint32_t)med << sh);
            }
        }
        ((uint4 *)o)[q] = make_uint4(ws[0], ws[1], ws[2], ws[3]);
    };
#pragma unroll
    for (int i = 0; i < kClipRegs; ++i) {
        const uint32_t q = tid + 256u * i;
        if (q < nv) clip_word(keep[i], q);
    }
    for (uint32_t q = tid + 256u * kClipRegs; q < nv; q += 256) clip_word(((const uint4 *)x)[q], q);
    for (uint32_t t = 16u * nv + tid; t < n; t += 256) {
        const int v = x[t];
        const bool clip = (float)abs(v - med) > thr;
        mine += clip;
        o[t] = clip ? (uint8_t)med : (uint8_t)v;
    }
    // wave reduction of the clip count, then one LDS atomic per wave
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) mine += __shfl_xor(mine, d, 64);
    if (lane == 0) atomicAdd(&sh_cnt, mine);
    __syncthreads();
    if (tid == 0) {
        if (med_out) med_out[c] = med;
        if (mad_out) mad_out[c] = mad;
        if (ncl_out) ncl_out[c] = (int32_t)sh_cnt;
    }
}

hipError_t launch_clip_u8(const uint8_t *in, int64_t in_stride, uint32_t nchans, int64_t t0,
                          uint32_t n, float nsigma, uint8_t *out, int64_t out_stride,
                          int32_t *med, int32_t *mad, int32_t *ncl, hipStream_t s)
{
    if (nchans == 0 || n == 0) return hipSuccess;
    hipLaunchKernelGGL(clip_u8_kernel, dim3(nchans), dim3(256), 0, s, in, in_stride, t0, n,
                       nsigma, out, out_stride, med, mad, ncl);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Row statistics (FP64): one workgroup per row.  Writes sums s1, s2.
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(256)
    row_sums_kernel(const float *x, uint32_t n, int64_t stride, double *s1_out, double *s2_out)
{
    __shared__ double p1[4], p2[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const float *row = x + (int64_t)blockIdx.x * stride;
    // 16-byte loads where the row allows, four of them in flight per thread,
    // into four independent accumulator pairs (FP64 sums of float32 values:
    // exact for the integer-valued rows of uint8 dedispersion in any order)
    double a1[4] = {0.0, 0.0, 0.0, 0.0}, a2[4] = {0.0, 0.0, 0.0, 0.0};
    uint32_t t = 0;
    if (((uintptr_t)row & 15) == 0) {
        const uint32_t n16 = n & ~1023u;  // whole rounds of 256 threads x 4 floats
        for (; t < n16; t += 4096) {
            float4 v[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t i = t + 1024u * j + 4u * tid;
                v[j] = make_float4(0.f, 0.f, 0.f, 0.f);
                if (i + 3 < n16) v[j] = *gptr<float4>(row + i);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const double e0 = v[j].x, e1 = v[j].y, e2 = v[j].z, e3 = v[j].w;
                a1[j] += (e0 + e1) + (e2 + e3);
                a2[j] += (e0 * e0 + e1 * e1) + (e2 * e2 + e3 * e3);
            }
        }
        t = n16;
    }
    for (uint32_t i = t + tid; i < n; i += 256) {
        const double v = (double)row[i];
        a1[0] += v;
        a2[0] += v * v;
    }
    double s1 = (a1[0] + a1[1]) + (a1[2] + a1[3]);
    double s2 = (a2[0] + a2[1]) + (a2[2] + a2[3]);
    s1 = wave_sum_f64(s1);
    s2 = wave_sum_f64(s2);
    if (lane == 0) {
        p1[wave] = s1;
        p2[wave] = s2;
    }
    __syncthreads();
    if (tid == 0) {
        s1_out[blockIdx.x] = (p1[0] + p1[1]) + (p1[2] + p1[3]);
        s2_out[blockIdx.x] = (p2[0] + p2[1]) + (p2[2] + p2[3]);
    }
}

// Row statistics from the per-tile (sum, sum of squares) partials the
// dedispersion launch writes (integer values: the sums are exact in FP64 in
// any order), finished as row_stats_finish does.  One wave per row.
__global__ void __launch_bounds__(256)
    row_partials_kernel(const double2 *part, uint32_t ndms, uint32_t tiles, uint32_t n,
                        double *mean, double *sigma)
{
    const uint32_t row = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    if (row >= ndms) return;  // whole waves; nothing below synchronises the workgroup
    const double2 *p = part + (size_t)row * tiles;
    double a1 = 0.0, a2 = 0.0;
    for (uint32_t i = lane; i < tiles; i += 64u) {
        const double2 v = p[i];
        a1 += v.x;
        a2 += v.y;
    }
    a1 = wave_sum_f64(a1);
    a2 = wave_sum_f64(a2);
    if (lane == 0) {
        const double m = a1 / (double)n;
        const double ex2 = a2 / (double)n;
        const double mm = m * m;
        double var = ex2 - mm;
        if (var < 0.0) var = 0.0;
        mean[row] = m;
        sigma[row] = sqrt(var);
    }
}

// mean = s1/n, sigma = sqrt(s2/n - mean*mean) (each a single IEEE operation;
// this file is compiled with -ffp-contract=off).
__global__ void row_stats_finish(uint32_t ndms, uint32_t n, double *mean, double *sigma)
{
    const uint32_t d = blockIdx.x * 256u + threadIdx.x;
    if (d >= ndms) return;
    const double m = mean[d] / (double)n;
    const double ex2 = sigma[d] / (double)n;
    const double mm = m * m;
    double var = ex2 - mm;
    if (var < 0.0) var = 0.0;
    mean[d] = m;
    sigma[d] = sqrt(var);
}

// ---------------------------------------------------------------------------
// Boxcar detector: grid (tiles of kBT samples, rows).
// ---------------------------------------------------------------------------

constexpr int kBT = 1024;           // samples per tile
constexpr int kMaxBoxW = 1024;      // widest boxcar
constexpr int kBoxL = kBT + kMaxBoxW + 1;

__global__ void __launch_bounds__(256)
    boxcar_kernel(const float *x, uint32_t n, int64_t stride, BoxStream bs, uint32_t wmask,
                  uint32_t maxw, double thr, SqrtW sqrtw, mdsm_cand *cands, uint32_t cap,
                  uint32_t *count, const double *mean, const double *sigma)
{
    __shared__ double buf[2][kBoxL];
    const int tid = threadIdx.x;
    const uint32_t d = blockIdx.y;
    const double sg = sigma[d];
    if (!(sg > 0.0)) return;
    const double mu = mean[d];
    const float *row = x + (int64_t)d * stride;
    const float *crow = bs.carry + (int64_t)d * bs.carry_stride;
    const int64_t H = bs.H, ne = (int64_t)n + H;        // extended series length
    const int64_t b = (int64_t)blockIdx.x * kBT - 1;     // sample at local index 0
    const int L = kBT + (int)maxw + 1;
    for (int u = tid; u < L; u += 256) {
        const int64_t t = b + u;
        buf[0][u] = (t >= 0 && t < ne) ? (double)(t < H ? crow[t] : row[t - H]) : 0.0;
    }
    __syncthreads();
    int cur = 0;
    for (uint32_t w = 1, wi = 0; w <= maxw; w <<= 1, ++wi) {
        if (wmask & w) {
            const double wm = (double)w * mu;
            const double den = sg * sqrtw.v[wi];
            // (num / den > thr) with den > 0 implies num > thr*den - |thr*den| 2^-40
            // (rounding errors are below 2^-52 relative): samples at or below
            // that bound skip the division without changing any decision
            const double td = thr * den;
            const double lo = td - fabs(td) * 0x1p-40;
            const int64_t last = ne - w;
            const int64_t tmin = H > (int64_t)w ? H - w : 0;
            for (int u = 1 + tid; u <= kBT; u += 256) {
                const int64_t t = b + u;
                if (t > last - bs.tail) break;
                if (t < tmin) continue;
                const double sv = buf[cur][u];
                const double num = sv - wm;
                if (!(num > lo)) continue;
                const double snr = num / den;
                if (!(snr > thr)) continue;
                if (t > 0 && !(sv >= buf[cur][u - 1])) continue;
                if (t < last && !(sv > buf[cur][u + 1])) continue;
                const uint32_t slot = atomicAdd(count, 1u);
                if (slot < cap) {
                    mdsm_cand cd;
                    cd.dm = d;
                    cd.t = (int32_t)(t - H);
                    cd.width = w;
                    cd.snr = (float)snr;
                    cands[slot] = cd;
                }
            }
        }
        if (2 * w > maxw) break;
        // next level: S_2w[u] = S_w[u] + S_w[u+w], u <= L - 2w
        const int lim = L - 2 * (int)w;
        for (int u = tid; u <= lim; u += 256) buf[cur ^ 1][u] = buf[cur][u] + buf[cur][u + w];
        __syncthreads();
        cur ^= 1;
    }
}

// Register form for widths up to 64: a wave holds kBW = 1024 consecutive
// samples of a row (16 per lane, starting 4 before its tile) and builds
// S_2w(t) = S_w(t) + S_w(t+w) in registers -- the same tree and FP64
// arithmetic as above.  Partners of levels w < 16 are mostly in the lane; the
// rest come from the next lane by a DPP wave shift (wave_shl:1, lane i reads
// lane i+1), level 16 takes the next lane's whole set the same way and only
// level 32 needs a two-lane shuffle.  Each wave reports the kBR tile
// positions whose widest sum and right neighbour lie inside its window.
constexpr int kBL = 16;                 // samples per lane
constexpr int kBW = 64 * kBL;           // samples per wave
constexpr int kBR = kBW - 4 - 64;       // positions reported per wave (maxw <= 64)

__device__ __forceinline__ double shfl_down_d(double v, int k) { return __shfl_down(v, k, 64); }
__device__ __forceinline__ double shfl_up_d(double v, int k) { return __shfl_up(v, k, 64); }
__device__ __forceinline__ float shfl_down_d(float v, int k) { return __shfl_down(v, k, 64); }
__device__ __forceinline__ float shfl_up_d(float v, int k) { return __shfl_up(v, k, 64); }
// v of lane i+1 (lane 63 keeps its own): DPP wave_shl:1
__device__ __forceinline__ float next_lane_d(float v)
{
    const int u = __builtin_bit_cast(int, v);
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(u, u, 0x130, 0xf, 0xf, false));
}
// v of lane i+1 (lane 63 keeps its own): DPP wave_shl:1 on both halves
__device__ __forceinline__ double next_lane_d(double v)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const int lo = (int)(uint32_t)u, hi = (int)(uint32_t)(u >> 32);
    const uint32_t nlo = (uint32_t)__builtin_amdgcn_update_dpp(lo, lo, 0x130, 0xf, 0xf, false);
    const uint32_t nhi = (uint32_t)__builtin_amdgcn_update_dpp(hi, hi, 0x130, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((uint64_t)nhi << 32) | nlo);
}

// T = float: for series of integers whose window sums stay below 2^24 (the
// caller guarantees it: uint8 data summed over C channels, 255 C maxw < 2^24),
// every sum of the tree is exact in float32 and equal to the FP64 one.  The
// pre-filter then runs in float32 with a margin that covers its two
// roundings (below 1 + 2^-23 |lo| at these magnitudes; 4 + 2^-20 |lo| is
// used), so every sample the FP64 pre-filter passes also passes it, and the
// passing samples are decided in FP64 exactly as in the double form: the
// candidates are the same.
template <typename T>
__global__ void __launch_bounds__(256)
    boxcar_reg_kernel(const float *x, uint32_t n, int64_t stride, BoxStream bs, uint32_t wmask,
                      uint32_t maxw, double thr, SqrtW sqrtw, mdsm_cand *cands, uint32_t cap,
                      uint32_t *count, const double *mean, const double *sigma)
{
    constexpr bool kF32 = sizeof(T) == 4;
    const uint32_t d = blockIdx.y;
    const double sg = sigma[d];
    if (!(sg > 0.0)) return;
    const double mu = mean[d];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int64_t H = bs.H, ne = (int64_t)n + H;  // extended series: carry, then the row
    const int64_t T0 = ((int64_t)blockIdx.x * 4 + wave) * kBR;
    if (T0 >= ne) return;  // the whole wave
    const float *row = x + (int64_t)d * stride;
    const float *crow = bs.carry + (int64_t)d * bs.carry_stride;
    const int64_t b = T0 - 4;       // sample held by lane 0, slot 0
    const int64_t tl = b + kBL * lane;
    T s[kBL];
    // 16-byte loads when every float4 of the lane's 16 samples is aligned and
    // lies in one source (H, tl and both row bases are multiples of 4 floats)
    const bool vec = tl >= 0 && tl + kBL <= ne && ((tl | H) & 3) == 0 &&
                     (((uintptr_t)row | (uintptr_t)crow | ((uint64_t)stride << 2) |
                       ((uint64_t)bs.carry_stride << 2)) & 15) == 0;
    if (vec) {
#pragma unroll
        for (int q = 0; q < kBL / 4; ++q) {
            const int64_t e = tl + 4 * q;
            const float4 v = e < H ? *(const float4 *)(crow + e) : *(const float4 *)(row + e - H);
            s[4 * q] = v.x;
            s[4 * q + 1] = v.y;
            s[4 * q + 2] = v.z;
            s[4 * q + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < kBL; ++j) {
            const int64_t t = tl + j;
            s[j] = (t >= 0 && t < ne) ? (T)(t < H ? crow[t] : row[t - H]) : (T)0;
        }
    }
    // slots j of this lane whose position u = kBL*lane + j is one the wave
    // reports, u in [4, 4 + kBR)
    uint32_t vm;
    {
        const int lo = max(0, 4 - kBL * lane), hi = min(kBL, 4 + kBR - kBL * lane);
        vm = hi > lo ? ((1u << hi) - 1u) & ~((1u << lo) - 1u) : 0u;
    }
#pragma unroll
    for (int lv = 0; lv < 7; ++lv) {
        const uint32_t w = 1u << lv;
        if (w > maxw) break;
        if (wmask & w) {
            const double wm = (double)w * mu;
            const double den = sg * sqrtw.v[lv];
            const double td = thr * den;
            const double lo = td - fabs(td) * 0x1p-40;  // exact pre-filter, as above
            const float wm32 = (float)wm;
            const float lo32 = (float)lo - (4.0f + fabsf((float)lo) * 0x1p-20f);
            const int64_t last = ne - w;
            const int64_t tmax = last - bs.tail;
            const int64_t tmin = H > (int64_t)w ? H - w : 0;
            auto pass = [&](T v) {
                if constexpr (kF32)
                    return (v - wm32) > lo32;
                else
                    return (v - wm) > lo;
            };
            // Quick rejection: the pre-filter is monotone in the sum (a
            // correctly rounded subtraction is), so no slot of the lane passes
            // when its largest sum does not.  Slots outside the reported
            // positions take part here and only cost the exact pass below.
            T m = s[0];
#pragma unroll
            for (int j = 1; j < kBL; ++j) m = s[j] > m ? s[j] : m;
            if (__any(pass(m))) {
                const double prevL = (double)shfl_up_d(s[kBL - 1], 1);
                const double nextF = (double)next_lane_d(s[0]);
#pragma unroll
                for (int j = 0; j < kBL; ++j) {
                    const int64_t t = b + kBL * lane + j;
                    if (!(((vm >> j) & 1u) != 0u && t <= tmax && t >= tmin && pass(s[j])))
                        continue;
                    const double sv = (double)s[j];
                    const double snr = (sv - wm) / den;
                    if (!(snr > thr)) continue;
                    const double pv = j > 0 ? (double)s[j > 0 ? j - 1 : 0] : prevL;
                    const double nv = j < kBL - 1 ? (double)s[j < kBL - 1 ? j + 1 : 0] : nextF;
                    if (t > 0 && !(sv >= pv)) continue;
                    if (t < last && !(sv > nv)) continue;
                    const uint32_t slot = atomicAdd(count, 1u);
                    if (slot < cap) {
                        mdsm_cand cd;
                        cd.dm = d;
                        cd.t = (int32_t)(t - H);
                        cd.width = w;
                        cd.snr = (float)snr;
                        cands[slot] = cd;
                    }
                }
            }
        }
        if (2 * w > maxw) break;
        // next level, S_2w(t) = S_w(t) + S_w(t + w), reading only old values
        if (w < (uint32_t)kBL) {
            T nx[kBL];
#pragma unroll
            for (int j = 0; j < kBL; ++j)
                if (j < (int)w) nx[j] = next_lane_d(s[j]);
#pragma unroll
            for (int j = 0; j < kBL; ++j)
                s[j] += (j + (int)w < kBL) ? s[(j + (int)w) % kBL] : nx[(j + (int)w) % kBL];
        } else if (w == (uint32_t)kBL) {
#pragma unroll
            for (int j = 0; j < kBL; ++j) s[j] += next_lane_d(s[j]);
        } else {
            const int k = (int)(w / kBL);
#pragma unroll
            for (int j = 0; j < kBL; ++j) s[j] += shfl_down_d(s[j], k);
        }
    }
}

uint32_t boxcar_maxw(uint32_t wmask)
{
    uint32_t maxw = 1;
    while ((maxw << 1) <= wmask && (maxw << 1) <= (uint32_t)kMaxBoxW) maxw <<= 1;
    return maxw;
}

uint32_t boxcar_carry_len(uint32_t wmask)
{
    return (boxcar_maxw(wmask) + 1 + 3) & ~3u;
}

hipError_t launch_boxcar_stream(const float *x, uint32_t ndms, uint32_t n, int64_t stride,
                                const BoxStream &bs, uint32_t wmask, double thr,
                                const SqrtW &sqrtw, mdsm_cand *cands, uint32_t cap,
                                uint32_t *count, double *mean, double *sigma, bool stats,
                                hipStream_t s, bool int_exact)
{
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, s, count, 1u, 0u);
    if (ndms == 0 || n == 0) return hipGetLastError();
    const uint32_t maxw = boxcar_maxw(wmask);
    if (stats) {
        hipLaunchKernelGGL(row_sums_kernel, dim3(ndms), dim3(256), 0, s, x, n, stride, mean,
                           sigma);
        hipLaunchKernelGGL(row_stats_finish, dim3((ndms + 255) / 256), dim3(256), 0, s, ndms, n,
                           mean, sigma);
    }
    const int64_t ne = (int64_t)n + bs.H;
    if (maxw <= 64) {
        const uint32_t tiles = (uint32_t)((ne + kBR - 1) / kBR);
        if (int_exact)
            hipLaunchKernelGGL(boxcar_reg_kernel<float>, dim3((tiles + 3) / 4, ndms), dim3(256), 0,
                               s, x, n, stride, bs, wmask, maxw, thr, sqrtw, cands, cap, count,
                               mean, sigma);
        else
            hipLaunchKernelGGL(boxcar_reg_kernel<double>, dim3((tiles + 3) / 4, ndms), dim3(256),
                               0, s, x, n, stride, bs, wmask, maxw, thr, sqrtw, cands, cap, count,
                               mean, sigma);
        return hipGetLastError();
    }
    dim3 grid((uint32_t)((ne + kBT - 1) / kBT), ndms);
    hipLaunchKernelGGL(boxcar_kernel, grid, dim3(256), 0, s, x, n, stride, bs, wmask, maxw, thr,
                       sqrtw, cands, cap, count, mean, sigma);
    return hipGetLastError();
}

hipError_t launch_row_partials(const void *part, uint32_t ndms, uint32_t tiles, uint32_t n,
                               double *mean, double *sigma, hipStream_t s)
{
    if (ndms == 0) return hipSuccess;
    hipLaunchKernelGGL(row_partials_kernel, dim3((ndms + 3) / 4), dim3(256), 0, s,
                       (const double2 *)part, ndms, tiles, n, mean, sigma);
    return hipGetLastError();
}

hipError_t launch_boxcar(const float *x, uint32_t ndms, uint32_t n, int64_t stride, uint32_t wmask,
                         double thr, const SqrtW &sqrtw, mdsm_cand *cands, uint32_t cap,
                         uint32_t *count, double *mean, double *sigma, hipStream_t s)
{
    BoxStream bs{};  // one block on its own: no carry, the last window reported
    bs.carry = x;
    return launch_boxcar_stream(x, ndms, n, stride, bs, wmask, thr, sqrtw, cands, cap, count,
                                mean, sigma, true, s);
}

// ---------------------------------------------------------------------------
// Median-of-five threshold detector (compat with the reference output thread)
// ---------------------------------------------------------------------------

__device__ __forceinline__ float med5(float a, float b, float c, float d, float e)
{
    // sorting network on 5 values; the middle element is the median
    float t;
#define CSWAP(p, q) \
    if (q < p) {    \
        t = p;      \
        p = q;      \
        q = t;      \
    }
    CSWAP(a, b);
    CSWAP(d, e);
    CSWAP(c, e);
    CSWAP(c, d);
    CSWAP(b, e);
    CSWAP(a, d);
    CSWAP(a, c);
    CSWAP(b, d);
    CSWAP(b, c);
#undef CSWAP
    return c;
}

// Whole-block statistics of the reference's brute-force output thread
// without noise arrays (src/mdsm/dedispersion_output.cpp:143-167).  The
// ndms x n values, rows read one after another as one series, are taken in
// chunks of mf = min(vals, 32768) values; chunk i's sum and sum of squares
// are formed in double in index order (one thread per chunk, so the
// rounding of every addition is the reference's); the loop stops before the
// chunk that would reach the end of the values, so iters = max(1,
// ceil(vals / mf) - 1) chunks are read.
uint32_t brute_stat_chunks(uint32_t ndms, uint32_t n, uint64_t *mf)
{
    const uint64_t vals = (uint64_t)ndms * n;
    *mf = vals < 32768 ? vals : 32768;
    if (*mf == 0) return 0;
    const uint64_t c = (vals + *mf - 1) / *mf;
    return (uint32_t)(c > 1 ? c - 1 : 1);
}

__global__ void __launch_bounds__(256)
    brute_chunk_sums_kernel(const float *x, uint32_t n, int64_t stride, uint64_t mf,
                            uint32_t iters, double *tot, double *tot2)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= iters) return;
    const uint64_t f = (uint64_t)i * mf;
    const uint64_t row = f / n;
    uint32_t col = (uint32_t)(f - row * n);
    const float *p = x + (int64_t)row * stride;
    double t = 0.0, t2 = 0.0;
    for (uint64_t j = 0; j < mf; ++j) {
        const double v = (double)p[col];
        t += v;
        t2 += v * v;  // the square of a float is exact in double, as pow(x, 2) is
        if (++col == n) {
            col = 0;
            p += stride;
        }
    }
    tot[i] = t;
    tot2[i] = t2;
}

// The reference's combination, in its arithmetic: the float running mean
// gains each chunk mean (the addition formed in double, then rounded to
// float), the double mean square each chunk's; mean = float(sum) / iters in
// float; rms = float(sqrt(msq / iters - float(mean * mean))).
__global__ void brute_stats_finish(const double *tot, const double *tot2, uint32_t iters,
                                   uint64_t mf, double *stats)
{
    if (threadIdx.x != 0) return;
    const double mfd = (double)mf;
    float lm = 0.0f;
    double m2 = 0.0;
    for (uint32_t i = 0; i < iters; ++i) {
        lm = (float)((double)lm + tot[i] / mfd);
        m2 += tot2[i] / mfd;
    }
    lm = lm / (float)iters;
    const float lm2 = lm * lm;
    stats[0] = (double)lm;
    stats[1] = (double)(float)sqrt(m2 / (double)iters - (double)lm2);
}

constexpr uint32_t kM5 = 16;  // median-of-five outputs per thread

// mode 0: statistics over the whole block (stats[0..1]);
// mode 1: per-row statistics from the row sums (stats[2..2+2*ndms)), as the
//         reference's subband output path (dedispersion_output.cpp:86-96);
// mode 2: caller-supplied per-chunk mean/rms, chunk = l / spc
//         (dedispersion_output.cpp:120-126,199-208).
__global__ void __launch_bounds__(256)
    median5_kernel(const float *x, uint32_t n, int64_t stride, const float *dmv, float thr,
                   int mode, const float *cmean, const float *crms, uint32_t spc,
                   mdsm_cand *cands, uint32_t cap, uint32_t *count, const double *stats,
                   uint32_t ndms)
{
    // a thread takes kM5 consecutive outputs l0.. from samples l0..l0+kM5+3
    const uint32_t d = blockIdx.y;
    const uint32_t l0 = kM5 * (blockIdx.x * 256u + threadIdx.x);
    if (n < 5 || l0 >= n - 4) return;
    if (!(dmv[d] > 1.0f)) return;
    float rm = 0.0f, rr = 0.0f;  // row statistics (modes 0 and 1)
    if (mode == 0) {
        rm = (float)stats[0];
        rr = (float)stats[1];
    } else if (mode == 1) {
        rm = (float)(stats[2 + d] / (double)n);
        const double v = stats[2 + ndms + d] / (double)n - (double)rm * (double)rm;
        rr = (float)sqrt(v);
    }
    const float *r = x + (int64_t)d * stride + l0;
    float v[kM5 + 4];
    if (l0 + kM5 + 4u <= n && ((uintptr_t)r % 16) == 0) {
#pragma unroll
        for (int q = 0; q < kM5 / 4 + 1; ++q) {
            const float4 a = ((const float4 *)r)[q];
            v[4 * q] = a.x;
            v[4 * q + 1] = a.y;
            v[4 * q + 2] = a.z;
            v[4 * q + 3] = a.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < kM5 + 4; ++i) v[i] = l0 + i < n ? r[i] : 0.0f;
    }
#pragma unroll
    for (int q = 0; q < kM5; ++q) {
        const uint32_t l = l0 + q;
        if (l >= n - 4) break;
        float lm = rm, lr = rr;
        if (mode == 2) {
            lm = cmean[l / spc];
            lr = crms[l / spc];
        }
        const float m = med5(v[q], v[q + 1], v[q + 2], v[q + 3], v[q + 4]);
        if (m - lm >= lr * thr) {
            const uint32_t slot = atomicAdd(count, 1u);
            if (slot < cap) {
                mdsm_cand cd;
                cd.dm = d;
                cd.t = l;
                cd.width = 5;
                cd.snr = mode == 2 ? (m - lm) / lr : m / lr;
                cands[slot] = cd;
            }
        }
    }
}

hipError_t launch_median5(const float *x, uint32_t ndms, uint32_t n, int64_t stride,
                          const float *dmv, double thr, int mode, const float *cmean,
                          const float *crms, uint32_t spc, mdsm_cand *cands, uint32_t cap,
                          uint32_t *count, double *stats, hipStream_t s)
{
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, s, count, 1u, 0u);
    if (ndms == 0 || n < 5) return hipGetLastError();
    // stats[2 .. 2+2*ndms) holds the per-row sums (mode 1), stats[2 ..
    // 2+2*chunks) the chunk sums (mode 0)
    if (mode == 1) {
        double *s1 = stats + 2, *s2 = stats + 2 + ndms;
        hipLaunchKernelGGL(row_sums_kernel, dim3(ndms), dim3(256), 0, s, x, n, stride, s1, s2);
    } else if (mode == 0) {
        uint64_t mf = 0;
        const uint32_t iters = brute_stat_chunks(ndms, n, &mf);
        double *t1 = stats + 2, *t2 = stats + 2 + iters;
        hipLaunchKernelGGL(brute_chunk_sums_kernel, dim3((iters + 255) / 256), dim3(256), 0, s, x, n,
                           stride, mf, iters, t1, t2);
        hipLaunchKernelGGL(brute_stats_finish, dim3(1), dim3(64), 0, s, t1, t2, iters, mf, stats);
    }
    dim3 grid(((n - 4 + kM5 - 1) / kM5 + 255) / 256, ndms);
    hipLaunchKernelGGL(median5_kernel, grid, dim3(256), 0, s, x, n, stride, dmv, (float)thr, mode,
                       cmean, crms, spc, cands, cap, count, stats, ndms);
    return hipGetLastError();
}

}  // namespace mdsm
