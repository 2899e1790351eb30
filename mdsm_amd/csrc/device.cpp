// Device plumbing of the C ABI: memory, copies, streams, device queries.
// These replace the reference worker's direct CUDA runtime calls
// (src/mdsm/dedispersion_thread.cu:258-263 cudaSetDevice/cudaMalloc/
// cudaMemcpyToSymbol, :285,297 H2D copies, :324 cudaMemset, :353-355 D2H).
#include <cstring>

#include "mdsm_internal.h"

using namespace mdsm;

extern "C" int mdsm_device_count(int *count)
{
    if (!count) return fail(MDSM_ERR_INVALID, "NULL argument");
    *count = 0;
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) {
        *count = 0;
        return hip_fail(e, "hipGetDeviceCount");
    }
    return MDSM_OK;
}

extern "C" int mdsm_device_set(int device)
{
    hipError_t e = hipSetDevice(device);
    return e == hipSuccess ? MDSM_OK : hip_fail(e, "hipSetDevice");
}

extern "C" int mdsm_device_query(int device, mdsm_device_props *props)
{
    if (!props) return fail(MDSM_ERR_INVALID, "NULL argument");
    hipDeviceProp_t p;
    hipError_t e = hipGetDeviceProperties(&p, device);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
    std::memset(props, 0, sizeof(*props));
    std::strncpy(props->name, p.name, sizeof(props->name) - 1);
    std::strncpy(props->arch, p.gcnArchName, sizeof(props->arch) - 1);
    props->total_mem = p.totalGlobalMem;
    props->cus = p.multiProcessorCount;
    props->clock_khz = p.clockRate;
    props->mem_clock_khz = p.memoryClockRate;
    props->mem_bus_bits = p.memoryBusWidth;
    props->lds_per_block = (uint32_t)p.sharedMemPerBlock;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess) props->free_mem = fr;
    return MDSM_OK;
}

extern "C" int mdsm_device_malloc(void **ptr, uint64_t bytes)
{
    if (!ptr) return fail(MDSM_ERR_INVALID, "NULL argument");
    *ptr = nullptr;
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
    if (e != hipSuccess) {
        *ptr = nullptr;
        return fail(MDSM_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " +
                                        hipGetErrorString(e));
    }
    return MDSM_OK;
}

extern "C" int mdsm_device_free(void *ptr)
{
    if (!ptr) return MDSM_OK;
    hipError_t e = hipFree(ptr);
    return e == hipSuccess ? MDSM_OK : hip_fail(e, "hipFree");
}

extern "C" int mdsm_host_malloc_pinned(void **ptr, uint64_t bytes)
{
    if (!ptr) return fail(MDSM_ERR_INVALID, "NULL argument");
    hipError_t e = hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault);
    if (e != hipSuccess) {
        *ptr = nullptr;
        return fail(MDSM_ERR_NOMEM, "hipHostMalloc failed");
    }
    return MDSM_OK;
}

extern "C" int mdsm_host_free_pinned(void *ptr)
{
    if (!ptr) return MDSM_OK;
    hipError_t e = hipHostFree(ptr);
    return e == hipSuccess ? MDSM_OK : hip_fail(e, "hipHostFree");
}

extern "C" int mdsm_memcpy(void *dst, const void *src, uint64_t bytes, int kind, void *stream)
{
    if ((!dst || !src) && bytes) return fail(MDSM_ERR_INVALID, "NULL argument");
    hipMemcpyKind k;
    switch (kind) {
    case MDSM_H2D: k = hipMemcpyHostToDevice; break;
    case MDSM_D2H: k = hipMemcpyDeviceToHost; break;
    case MDSM_D2D: k = hipMemcpyDeviceToDevice; break;
    default: return fail(MDSM_ERR_INVALID, "bad copy kind");
    }
    if (!bytes) return MDSM_OK;
    hipError_t e = stream ? hipMemcpyAsync(dst, src, bytes, k, (hipStream_t)stream)
                          : hipMemcpy(dst, src, bytes, k);
    return e == hipSuccess ? MDSM_OK : hip_fail(e, "hipMemcpy");
}

extern "C" int mdsm_memcpy_2d(void *dst, uint64_t dpitch, const void *src, uint64_t spitch,
                              uint64_t width, uint64_t rows, int kind, void *stream)
{
    hipMemcpyKind k;
    switch (kind) {
    case MDSM_H2D: k = hipMemcpyHostToDevice; break;
    case MDSM_D2H: k = hipMemcpyDeviceToHost; break;
    case MDSM_D2D: k = hipMemcpyDeviceToDevice; break;
    default: return fail(MDSM_ERR_INVALID, "bad copy kind");
    }
    if (!width || !rows) return MDSM_OK;
    hipError_t e = hipMemcpy2DAsync(dst, dpitch, src, spitch, width, rows, k, (hipStream_t)stream);
    if (e == hipSuccess && !stream) e = hipStreamSynchronize(nullptr);
    return e == hipSuccess ? MDSM_OK : hip_fail(e, "hipMemcpy2D");
}

extern "C" int mdsm_fill_u32(void *dst, uint32_t value, uint64_t count, void *stream)
{
    if (!dst && count) return fail(MDSM_ERR_INVALID, "NULL argument");
    if (count > 0xffffffffull) return fail(MDSM_ERR_INVALID, "count too large");
    hipError_t e = launch_fill_u32((uint32_t *)dst, (uint32_t)count, value, (hipStream_t)stream);
    return e == hipSuccess ? MDSM_OK : hip_fail(e, "fill launch");
}

extern "C" int mdsm_stream_create(void **stream)
{
    if (!stream) return fail(MDSM_ERR_INVALID, "NULL argument");
    hipStream_t s;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
    *stream = (void *)s;
    return MDSM_OK;
}

extern "C" int mdsm_stream_destroy(void *stream)
{
    if (!stream) return MDSM_OK;
    hipError_t e = hipStreamDestroy((hipStream_t)stream);
    return e == hipSuccess ? MDSM_OK : hip_fail(e, "hipStreamDestroy");
}

extern "C" int mdsm_stream_synchronize(void *stream)
{
    hipError_t e = hipStreamSynchronize((hipStream_t)stream);
    return e == hipSuccess ? MDSM_OK : hip_fail(e, "hipStreamSynchronize");
}

extern "C" int mdsm_device_synchronize(void)
{
    hipError_t e = hipDeviceSynchronize();
    return e == hipSuccess ? MDSM_OK : hip_fail(e, "hipDeviceSynchronize");
}
