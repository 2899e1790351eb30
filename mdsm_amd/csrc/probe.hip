// LDS read- and write-bandwidth probes: the roofs the dedispersion kernel's
// trial reads and staging stores are measured against, and the read forms
// its layouts chose between.  Each wave reads 64 lanes x `W` contiguous bytes per instruction
// (ds_read_b32 / _b64 / _b128) at a wave-uniform offset that moves on every
// read, the access shape of the dedispersion kernel's trial reads; the offset
// is masked with a run-time mask so that the compiler cannot merge reads.
// Every address is a multiple of 16 bytes (of W for the lane part), so each
// read is naturally aligned and ends inside the masked region.
// The number of resident waves per SIMD is set through the workgroup's LDS
// allocation (one wave per SIMD per workgroup of 256 threads).
#include <hip/hip_runtime.h>

#include "mdsm_internal.h"

using namespace mdsm;

namespace {

constexpr int kProbeUnroll = 16;

template <int W>
__global__ void __launch_bounds__(256) lds_probe_kernel(int iters, uint32_t mask, uint32_t *out)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t i = threadIdx.x; i < (mask + 1) / 4; i += 256)
        ((uint32_t *)lds)[i] = i * 2654435761u;
    __syncthreads();
    uint32_t acc = 0;
    uint32_t off = wave * 1024u;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int u = 0; u < kProbeUnroll; ++u) {
            const uint32_t a = (off + (uint32_t)u * 528u + lane * W) & mask;
            if constexpr (W == 4) {
                acc += *(const uint32_t *)(lds + a);
            } else if constexpr (W == 8) {
                const uint2 v = *(const uint2 *)(lds + a);
                acc += v.x ^ v.y;
            } else {
                const uint4 v = *(const uint4 *)(lds + a);
                acc += (v.x ^ v.y) + (v.z ^ v.w);
            }
        }
        off += kProbeUnroll * 528u + 16u;
    }
    out[blockIdx.x * 256 + threadIdx.x] = acc;
}

// Write form: each wave writes 64 lanes x W contiguous bytes per
// instruction (ds_write_b32 / _b64 / _b128) at a wave-uniform offset that
// moves on every write, the access shape of the dedispersion kernel's
// staging stores.  Every address is a multiple of 64*W bytes (of W for the
// lane part) inside the masked region; the barrier and the read-back keep
// the stores live.
template <int W>
__global__ void __launch_bounds__(256) lds_write_probe_kernel(int iters, uint32_t mask, uint32_t *out)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t v = threadIdx.x * 2654435761u;
    uint32_t off = wave * 1024u;
    constexpr uint32_t kAlign = 64u * W;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int u = 0; u < kProbeUnroll; ++u) {
            const uint32_t a = ((off + (uint32_t)u * 528u) & mask & ~(kAlign - 1)) + lane * W;
            if constexpr (W == 4)
                *(uint32_t *)(lds + a) = v + u;
            else if constexpr (W == 8)
                *(uint2 *)(lds + a) = make_uint2(v + u, v ^ u);
            else
                *(uint4 *)(lds + a) = make_uint4(v + u, v ^ u, v, u);
        }
        off += kProbeUnroll * 528u + 16u;
        v += 1u;
    }
    __syncthreads();
    out[blockIdx.x * 256 + threadIdx.x] = ((const uint32_t *)lds)[threadIdx.x];
}

// Other read forms, each at a wave-uniform offset that moves on every read
// as above (F = the mdsm_lds_read_form value):
//  1 two dwords 256 bytes apart per lane, lanes contiguous (ds_read2st64_b32:
//    the float16 rows' reads), 8 bytes per lane;
//  2 two 8-byte entries 512 bytes apart (ds_read2st64_b64: the float32
//    rows'), 16 bytes per lane;
//  3 8 bytes at a 4-byte-aligned, not 8-byte-aligned, address (ds_read_b64;
//    the misalignment `skew` = 4 is a run-time value, so the compiler keeps
//    the 8-byte read the type promises);
//  4 the same 8 bytes as two adjacent dwords (ds_read2_b32).
// Reads end at most 1 KiB past the masked region; the allocation covers it.
template <int F>
__global__ void __launch_bounds__(256) lds_form_probe_kernel(int iters, uint32_t mask, uint32_t skew,
                                                             uint32_t *out)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t i = threadIdx.x; i < (mask + 1) / 4 + 256; i += 256)
        ((uint32_t *)lds)[i] = i * 2654435761u;
    __syncthreads();
    uint32_t acc = 0;
    uint32_t off = wave * 1024u;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int u = 0; u < kProbeUnroll; ++u) {
            const uint32_t a = (off + (uint32_t)u * 528u) & mask;
            if constexpr (F == 1) {
                const uint32_t *q = (const uint32_t *)(lds + a + lane * 4);
                acc += q[0] ^ q[64];
            } else if constexpr (F == 2) {
                const uint2 *q = (const uint2 *)(lds + a + lane * 8);
                const uint2 lo = q[0], hi = q[64];
                acc += (lo.x ^ lo.y) + (hi.x ^ hi.y);
            } else if constexpr (F == 3) {
                const uint2 v = *(const uint2 *)(lds + a + skew + lane * 8);
                acc += v.x ^ v.y;
            } else {
                const uint32_t *q = (const uint32_t *)(lds + a + skew + lane * 8);
                acc += q[0] ^ q[1];
            }
        }
        off += kProbeUnroll * 528u + 16u;
    }
    out[blockIdx.x * 256 + threadIdx.x] = acc;
}

// Shared driver of the probes: occupancy set through the workgroup's LDS
// allocation, one warm-up launch, one timed launch.  form 0: the plain
// read or write of bytes_per_lane; forms 1..4: the read forms above.
int probe_lds(bool write, uint32_t bytes_per_lane, uint32_t waves_per_simd, double *tbs,
              uint32_t form = 0)
{
    if (!tbs) return fail(MDSM_ERR_INVALID, "NULL argument");
    if (form > 4) return fail(MDSM_ERR_INVALID, "form must be 1..4");
    if (form) bytes_per_lane = form == 2 ? 16 : 8;
    if (bytes_per_lane != 4 && bytes_per_lane != 8 && bytes_per_lane != 16)
        return fail(MDSM_ERR_INVALID, "bytes_per_lane must be 4, 8 or 16");
    if (waves_per_simd < 1 || waves_per_simd > 8)
        return fail(MDSM_ERR_INVALID, "waves_per_simd must be 1..8");
    int dev = 0, cus = 256;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return hip_fail(e, "device query");
    // LDS per workgroup: the CU's 160 KiB shared by waves_per_simd workgroups,
    // rounded down to a power of two for the address mask (>= 16 KiB)
    uint32_t lds = 16384;
    while (lds * 2 * waves_per_simd <= 163840u - 4096u && lds < 65536) lds *= 2;
    const uint32_t need = 163840u / waves_per_simd - 1024u;  // occupancy set by allocation
    const uint32_t alloc = need > lds + 1024u ? need : lds + 1024u;  // + the read forms' overrun
    const uint32_t mask = lds - 1;
    const int grid = cus * (int)waves_per_simd * 8;
    const int iters = 512;
    uint32_t *out = nullptr;
    e = hipMalloc(&out, (size_t)grid * 256 * 4);
    if (e != hipSuccess) return hip_fail(e, "probe buffer");
    void (*k)(int, uint32_t, uint32_t *) = nullptr;
    void (*kf)(int, uint32_t, uint32_t, uint32_t *) = nullptr;
    if (form)
        kf = form == 1 ? lds_form_probe_kernel<1>
                       : (form == 2 ? lds_form_probe_kernel<2>
                                    : (form == 3 ? lds_form_probe_kernel<3> : lds_form_probe_kernel<4>));
    else if (write)
        k = bytes_per_lane == 4 ? lds_write_probe_kernel<4>
                                : (bytes_per_lane == 8 ? lds_write_probe_kernel<8>
                                                       : lds_write_probe_kernel<16>);
    else
        k = bytes_per_lane == 4 ? lds_probe_kernel<4>
                                : (bytes_per_lane == 8 ? lds_probe_kernel<8> : lds_probe_kernel<16>);
    e = hipFuncSetAttribute(form ? (const void *)kf : (const void *)k,
                            hipFuncAttributeMaxDynamicSharedMemorySize, 163840);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    float ms = 0.0f;
    if (e == hipSuccess) {
        auto run = [&]() {
            if (form)
                hipLaunchKernelGGL(kf, dim3(grid), dim3(256), alloc, 0, iters, mask, 4u, out);
            else
                hipLaunchKernelGGL(k, dim3(grid), dim3(256), alloc, 0, iters, mask, out);
        };
        run();  // warm-up
        (void)hipEventRecord(e0, 0);
        run();
        (void)hipEventRecord(e1, 0);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
    if (e != hipSuccess) return hip_fail(e, "LDS probe");
    const double bytes = (double)grid * 256.0 * iters * kProbeUnroll * bytes_per_lane;
    *tbs = bytes / (ms * 1e-3) / 1e12;
    return MDSM_OK;
}

}  // namespace

extern "C" int mdsm_probe_lds_read(uint32_t bytes_per_lane, uint32_t waves_per_simd, double *tbs)
{
    return probe_lds(false, bytes_per_lane, waves_per_simd, tbs);
}

extern "C" int mdsm_probe_lds_write(uint32_t bytes_per_lane, uint32_t waves_per_simd, double *tbs)
{
    return probe_lds(true, bytes_per_lane, waves_per_simd, tbs);
}

extern "C" int mdsm_probe_lds_read_form(uint32_t form, uint32_t waves_per_simd, double *tbs)
{
    if (form < 1) return fail(MDSM_ERR_INVALID, "form must be 1..4");
    return probe_lds(false, 0, waves_per_simd, tbs, form);
}
