// Dedispersion without allocation: mdsm_dedisperse_noalloc.
//
// The workspace of a plan (decimated copies, subband stage-1 series) and the
// persistent work-queue counters live in one file-scope device array,
// `scratch`, instead of memory the plan or the caller allocates.  Each GPU
// the library runs on holds its own copy of the array (kScratchBytes) until
// the process ends.  The array's name is a host-side stand-in: its device
// address is looked up once per device with hipGetSymbolAddress and only
// that address is handed to kernels.  Every call on one GPU shares its array
// and finds the previous call's contents: the counters are cleared by each
// launch (a fill kernel), the workspace is written before it is read, and
// calls that overlap on one GPU (two streams, two threads) race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "mdsm_internal.h"

using namespace mdsm;

namespace {

constexpr uint64_t kScratchBytes = 256ull << 20;  // 256 MiB per GPU
constexpr uint64_t kCounterBytes = 256;           // work-queue counters, one per stage
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

// Device address of this device's scratch array (looked up on the first call
// on the device, with the device current).
int scratch_address(int dev, char **out)
{
    if (dev < 0 || dev >= kMaxDevices) return fail(MDSM_ERR_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!ptr) return fail(MDSM_ERR_DEVICE, "no device address for the scratch array");
        g_scratch[dev] = ptr;
    }
    *out = (char *)g_scratch[dev];
    return MDSM_OK;
}

}  // namespace

extern "C" uint64_t mdsm_noalloc_max_work(void)
{
    return kScratchBytes - kCounterBytes;
}

extern "C" int mdsm_dedisperse_noalloc(mdsm_plan *p, const void *d_in, int64_t in_stride,
                                       float *d_out, void *stream)
{
    if (!p || !d_in || !d_out) return fail(MDSM_ERR_INVALID, "NULL argument");
    if (in_stride < (int64_t)p->obs.nsamp_in)
        return fail(MDSM_ERR_INVALID, "in_stride < nsamp_in");
    if (p->work_bytes > kScratchBytes - kCounterBytes)
        return fail(MDSM_ERR_INVALID,
                    "plan needs " + std::to_string(p->work_bytes) +
                        " bytes of workspace; the no-allocation entry point holds at most " +
                        std::to_string(kScratchBytes - kCounterBytes));
    if (p->stages.size() > kCounterBytes / sizeof(uint32_t))
        return fail(MDSM_ERR_INVALID, "too many launch stages for the scratch counters");
    if (!p->uploaded)
        return fail(MDSM_ERR_STATE,
                    "plan tables are not on the device: call mdsm_plan_upload at set-up");
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (cur != p->device) return fail(MDSM_ERR_STATE, "plan belongs to another device");
    char *base = nullptr;
    int rc = scratch_address(cur, &base);
    if (rc) return rc;
    uint32_t *counters = (uint32_t *)base;
    char *work = base + kCounterBytes;
    hipStream_t s = (hipStream_t)stream;
    if (!p->decim.empty()) {
        e = launch_decimate_steps(d_in, p->obs.nchans, in_stride, p->obs.nsamp_in,
                                  p->decim.data(), p->decim.size(), work, s);
        if (e != hipSuccess) return hip_fail(e, "decimate launch");
    }
    for (size_t k = 0; k < p->stages.size(); ++k) {
        e = launch_dedisperse_ctr(p, p->stages[k], d_in, in_stride, d_out, work, counters + k, s);
        if (e != hipSuccess) return hip_fail(e, "dedisperse launch");
    }
    return MDSM_OK;
}
