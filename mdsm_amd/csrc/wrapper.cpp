// The reference's host/device ABI (include/mdsm_wrapper.h): device
// enumeration, and one GPU's worker of the reference manager's protocol
// (src/mdsm/dedispersion_thread.cu:246-401) running this library's plans.
//
// The protocol is the reference's, step for step: barriers, rw lock, the
// iterations/maxiters pipeline warm-up and the stopping clause, the host-side
// carry of the last maxshift samples of each channel to the front of the
// input rows.  The computation is a plan of this thread's share of the
// trials, dedispersed on a stream of its own.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/mdsm_wrapper.h"
#include "mdsm_internal.h"

using namespace mdsm;

extern "C" DEVICES *call_initialise_devices(SURVEY *survey)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        fail(MDSM_ERR_DEVICE, "no GPU found");
        std::fprintf(stderr, "No HIP device found\n");
        return nullptr;
    }
    auto *d = (DEVICES *)std::malloc(sizeof(DEVICES));
    if (d) d->devices = (DEVICE_INFO *)std::malloc((size_t)n * sizeof(DEVICE_INFO));
    if (!d || !d->devices) {
        std::free(d);
        fail(MDSM_ERR_NOMEM, "host allocation failed");
        return nullptr;
    }
    d->num_devices = 0;
    d->minTotalGlobalMem = 1024ul * 1024ul * 16ul;  // KiB, as the reference starts it
    for (int i = 0; i < n; ++i) {
        hipDeviceProp_t pr;
        if (hipGetDeviceProperties(&pr, i) != hipSuccess) continue;
        const unsigned long kib = (unsigned long)(pr.totalGlobalMem / 1024);
        if (!(kib > 1024ul * 1024ul * 5ul / 2ul)) continue;  // more than 2.5 GiB
        bool use = survey == nullptr || survey->gpu_ids == nullptr;
        if (!use)
            for (unsigned j = 0; j < survey->num_gpus; ++j) use = use || (int)survey->gpu_ids[j] == i;
        if (!use) continue;
        DEVICE_INFO &di = d->devices[d->num_devices++];
        di.multiprocessor_count = pr.multiProcessorCount;
        di.constant_memory = (int)std::min<size_t>(pr.totalConstMem, 0x7fffffff);
        di.shared_memory = (int)std::min<size_t>(pr.sharedMemPerBlock, 0x7fffffff);
        di.register_count = pr.regsPerBlock;
        di.thread_count = pr.maxThreadsPerBlock;
        di.clock_rate = pr.clockRate;
        di.device_id = i;
        d->minTotalGlobalMem = std::min(d->minTotalGlobalMem, kib);
    }
    if (d->num_devices == 0) {
        std::free(d->devices);
        std::free(d);
        fail(MDSM_ERR_DEVICE, "no GPU with more than 2.5 GiB in the survey's list");
        std::fprintf(stderr, "No HIP device found\n");
        return nullptr;
    }
    return d;
}

namespace {

void wait_barrier(pthread_barrier_t *b, const char *which)
{
    const int r = pthread_barrier_wait(b);
    if (!(r == 0 || r == PTHREAD_BARRIER_SERIAL_THREAD))
        std::fprintf(stderr, "Error during barrier synchronisation (%s) [thread]\n", which);
}

// This thread's plan: its share of the survey's trials over input rows of
// `stride` samples.
int make_plan(const THREAD_PARAMS *p, uint32_t stride, mdsm_plan **out)
{
    const SURVEY *s = p->survey;
    mdsm_obs obs{};
    obs.nchans = s->nchans;
    obs.fch1 = s->fch1;
    obs.foff = s->foff;
    obs.tsamp = s->tsamp;
    obs.nsamp_in = stride;
    obs.dtype = MDSM_F32;
    obs.shift_rule = MDSM_SHIFT_REF;
    const unsigned tid = p->thread_num;
    if (s->useBruteForce) {
        // the survey's thread count and the first trial as the reference
        // computes it, in float, left to right (src/mdsm/dedispersion_thread.cu:225)
        const unsigned nt = std::max<unsigned>(1, s->num_threads ? s->num_threads : p->num_threads);
        float startdm = s->dmstep * (float)s->tdms;
        startdm = startdm / (float)nt;
        startdm = startdm * (float)tid;
        startdm = s->lowdm + startdm;
        const mdsm_range r{startdm, s->dmstep, s->tdms / nt, 1};
        if (r.ndms < 1) return fail(MDSM_ERR_INVALID, "fewer trials than threads");
        return mdsm_plan_create(&obs, &r, 1, out);
    }
    const unsigned nt = std::max<unsigned>(1, p->num_threads);
    std::vector<mdsm_pass> passes;
    for (unsigned i = 0; i < s->num_passes; ++i) {
        const SUBBAND_PASSES &q = s->pass_parameters[i];
        const int ncalls = q.ncalls / (int)nt;
        // src/mdsm/dedispersion_thread.cu:161-163
        float startdm = q.sub_dmstep * (float)ncalls;
        startdm = startdm * (float)tid;
        startdm = q.lowdm + startdm;
        passes.push_back({startdm, q.highdm, q.dmstep, q.sub_dmstep, (uint32_t)q.binsize,
                          (uint32_t)(ncalls * q.calldms), (uint32_t)ncalls, (uint32_t)q.calldms});
    }
    return mdsm_plan_create_subband(&obs, s->nsubs, passes.data(), (uint32_t)passes.size(), out);
}

struct Worker {
    mdsm_plan *plan = nullptr;
    float *d_in = nullptr, *d_out = nullptr;
    void *d_work = nullptr;
    hipStream_t st = nullptr;
    bool ok = false;

    ~Worker()
    {
        if (st) (void)hipStreamDestroy(st);
        if (d_in) (void)hipFree(d_in);
        if (d_out) (void)hipFree(d_out);
        if (d_work) (void)hipFree(d_work);
        if (plan) mdsm_plan_destroy(plan);
    }
};

}  // namespace

extern "C" void *call_dedisperse(void *thread_params)
{
    auto *params = (THREAD_PARAMS *)thread_params;
    SURVEY *survey = params->survey;
    const size_t nchans = survey->nchans, maxshift = (size_t)std::max(0, params->maxshift);
    int nsamp = (int)survey->nsamp;
    const size_t stride = (size_t)nsamp + maxshift;  // the input buffer's row length
    int loop_counter = 0, iters = 0;
    std::vector<float> tail(maxshift * nchans);

    Worker w;
    {
        const char *why = nullptr;
        if (hipSetDevice(params->device_id) != hipSuccess)
            why = "hipSetDevice failed";
        else if (make_plan(params, (uint32_t)stride, &w.plan))
            why = mdsm_last_error();
        else if (hipStreamCreateWithFlags(&w.st, hipStreamNonBlocking) != hipSuccess ||
                 hipMalloc((void **)&w.d_in, stride * nchans * sizeof(float)) != hipSuccess ||
                 hipMalloc((void **)&w.d_out, w.plan->info.out_elems * sizeof(float)) !=
                     hipSuccess ||
                 hipMalloc(&w.d_work, std::max<uint64_t>(w.plan->info.work_bytes, 256)) !=
                     hipSuccess)
            why = "device allocation failed";
        else
            for (const RangeHost &rh : w.plan->ranges)
                if (rh.info.nout < (uint32_t)nsamp / rh.info.downsample)
                    why = "the manager's maxshift is shorter than the plan's delays";
        w.ok = why == nullptr;
        if (!w.ok)
            std::fprintf(stderr, "mdsm worker %u: %s; writing zeros\n", params->thread_num, why);
    }
    auto keep_tail = [&]() {  // last maxshift samples of this block, per channel
        for (size_t c = 0; c < nchans; ++c)
            std::memcpy(tail.data() + maxshift * c, params->input + c * stride + nsamp,
                        maxshift * sizeof(float));
    };

    for (;;) {
        if (loop_counter >= params->iterations && w.ok) {
            if (loop_counter != 1)  // the previous block's tail to the front of each row
                for (size_t c = 0; c < nchans; ++c)
                    std::memcpy(params->input + c * stride, tail.data() + maxshift * c,
                                maxshift * sizeof(float));
            w.ok = hipMemcpyAsync(w.d_in, params->input, stride * nchans * sizeof(float),
                                  hipMemcpyHostToDevice, w.st) == hipSuccess &&
                   hipStreamSynchronize(w.st) == hipSuccess;
            keep_tail();
        }

        wait_barrier(params->input_barrier, "input");

        if (loop_counter >= params->iterations && w.ok)
            w.ok = mdsm_dedisperse(w.plan, w.d_in, (int64_t)stride, w.d_out, w.d_work, w.st) ==
                       MDSM_OK &&
                   hipStreamSynchronize(w.st) == hipSuccess;

        wait_barrier(params->output_barrier, "output");

        if (loop_counter >= params->iterations) {
            // rows [trials][nsamp/binsize], ranges (passes) one after another
            size_t o = 0;
            for (size_t r = 0; w.ok && r < w.plan->ranges.size(); ++r) {
                const mdsm_range_info &ri = w.plan->ranges[r].info;
                const size_t n = (size_t)nsamp / ri.downsample;
                const size_t rows =
                    n ? std::min<size_t>(ri.ndms, (params->dedispersed_size - o) / n) : 0;
                if (rows)
                    w.ok = hipMemcpy2DAsync(params->output + o, n * sizeof(float),
                                            w.d_out + ri.out_offset, ri.nout * sizeof(float),
                                            n * sizeof(float), rows, hipMemcpyDeviceToHost,
                                            w.st) == hipSuccess &&
                           hipStreamSynchronize(w.st) == hipSuccess;
                o += rows * n;
            }
            if (!w.ok) std::memset(params->output, 0, params->dedispersed_size * sizeof(float));
        }

        if (pthread_rwlock_rdlock(params->rw_lock))
            std::fprintf(stderr, "Unable to acquire rw_lock [thread]\n");
        nsamp = (int)std::min<size_t>(survey->nsamp, stride - maxshift);
        if (params->stop) {
            if (iters >= params->iterations - 1) {
                pthread_rwlock_unlock(params->rw_lock);
                for (int i = 0; i < params->maxiters - params->iterations; ++i) {
                    wait_barrier(params->input_barrier, "input");
                    wait_barrier(params->output_barrier, "output");
                }
                break;
            }
            ++iters;
        }
        pthread_rwlock_unlock(params->rw_lock);
        ++loop_counter;
    }
    return thread_params;
}
