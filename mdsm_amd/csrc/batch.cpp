// Batched execution of many plans (beams) in one dedispersion launch per
// stage, optionally captured into a HIP graph for steady-state replay.
//
// The union of the plans' stage-k segments forms one work queue: segments are
// ordered longest item first (channels x trials per item), so a persistent
// launch drains the long items early and ends on short ones.
#include <algorithm>
#include <cstring>
#include <limits>

#include "mdsm_internal.h"

using namespace mdsm;

struct mdsm_batch {
    std::vector<mdsm_plan *> plans;
    int device = 0;
    int cus = 0;
    size_t nstages = 0;
    std::vector<SegDesc> segs_host;
    std::vector<int32_t> stage_seg_begin, stage_nseg;
    std::vector<int64_t> stage_items;
    SegDesc *segs_dev = nullptr;
    uint32_t *counters = nullptr;  // one persistent-queue counter per stage
    std::vector<void *> own_work;  // per-plan workspace when the caller passes none
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
};

namespace {

void drop_graph(mdsm_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

int build_queue(mdsm_batch *b)
{
    b->nstages = 0;
    for (mdsm_plan *p : b->plans) b->nstages = std::max(b->nstages, p->stages.size());
    b->segs_host.clear();
    b->stage_seg_begin.assign(b->nstages, 0);
    b->stage_nseg.assign(b->nstages, 0);
    b->stage_items.assign(b->nstages, 0);
    for (size_t k = 0; k < b->nstages; ++k) {
        struct Ent {
            double cost;
            SegDesc d;
        };
        std::vector<Ent> ents;
        for (size_t pi = 0; pi < b->plans.size(); ++pi) {
            const mdsm_plan *p = b->plans[pi];
            if (k >= p->stages.size()) continue;
            const Stage &st = p->stages[k];
            for (int i = st.seg_begin; i < st.seg_end; ++i) {
                SegDesc d = p->segs[i];
                const BlockDesc &b0 = p->blocks[d.first_block];
                d.job = (int32_t)pi;
                ents.push_back({(double)b0.nrows * b0.ntrials, d});
            }
        }
        std::stable_sort(ents.begin(), ents.end(),
                         [](const Ent &a, const Ent &c) { return a.cost > c.cost; });
        b->stage_seg_begin[k] = (int32_t)b->segs_host.size();
        int64_t items = 0;
        for (Ent &e : ents) {
            e.d.item_start = items;
            items += (int64_t)e.d.nblocks * e.d.ntiles;
            b->segs_host.push_back(e.d);
        }
        if (items > 0x7fffffffLL) return fail(MDSM_ERR_INVALID, "batch has more than 2^31 work items");
        b->stage_nseg[k] = (int32_t)ents.size();
        b->stage_items[k] = items;
    }
    if (b->segs_dev) (void)hipFree(b->segs_dev);
    b->segs_dev = nullptr;
    hipError_t e = hipMalloc((void **)&b->segs_dev,
                             std::max<size_t>(b->segs_host.size(), 1) * sizeof(SegDesc));
    if (e == hipSuccess && !b->segs_host.empty())
        e = hipMemcpy(b->segs_dev, b->segs_host.data(), b->segs_host.size() * sizeof(SegDesc),
                      hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "uploading the batch work queue");
    return MDSM_OK;
}

// Enqueue one step: every plan's decimation, then one launch per stage.
int enqueue(mdsm_batch *b, const void *const *ins, const int64_t *strides, float *const *outs,
            void *const *works, hipStream_t s)
{
    const size_t n = b->plans.size();
    std::vector<void *> wk(n);
    for (size_t i = 0; i < n; ++i) {
        const mdsm_plan *p = b->plans[i];
        wk[i] = (works && works[i]) ? works[i] : b->own_work[i];
        if (p->work_bytes && !wk[i]) return fail(MDSM_ERR_STATE, "plan workspace missing");
        if (!p->decim.empty()) {
            hipError_t e = launch_decimate_steps(ins[i], p->obs.nchans, strides[i], p->obs.nsamp_in,
                                                 p->decim.data(), p->decim.size(), (char *)wk[i], s);
            if (e != hipSuccess) return hip_fail(e, "batch decimate launch");
        }
    }
    for (size_t k = 0; k < b->nstages; ++k) {
        LaunchSpec L;
        L.segs = b->segs_dev;
        L.seg_begin = b->stage_seg_begin[k];
        L.nseg = b->stage_nseg[k];
        L.items = b->stage_items[k];
        bool persistent = false;
        uint32_t grid = 0, group = 1;
        for (size_t i = 0; i < n; ++i) {
            const mdsm_plan *p = b->plans[i];
            if (k < p->stages.size()) {
                L.lds_bytes = std::max(L.lds_bytes, p->stages[k].lds_bytes);
                L.dtypes |= p->stages[k].dtypes;
            }
            persistent = persistent || p->tune.persistent;
            grid = std::max(grid, p->tune.grid);
            group = std::max(group, p->tune.tile_group);
        }
        L.group = (int32_t)group;
        if (persistent) {
            L.counter = b->counters + k;
            L.grid = grid ? (int32_t)grid : 8 * std::max(1, b->cus);
        }
        L.njobs = (int32_t)n;
        for (size_t i = 0; i < n; ++i)
            L.jobs[i] = JobBufs{ins[i], outs[i], (char *)wk[i], strides[i]};
        TimedLaunch tl("dedisperse", s);
        hipError_t e = launch_dedisperse_spec(L, s);
        if (e != hipSuccess) return hip_fail(e, "batch dedisperse launch");
    }
    return MDSM_OK;
}

int check_args(mdsm_batch *b, const void *const *ins, const int64_t *strides, float *const *outs)
{
    if (!ins || !strides || !outs) return fail(MDSM_ERR_INVALID, "NULL argument");
    for (size_t i = 0; i < b->plans.size(); ++i) {
        if (!ins[i] || !outs[i]) return fail(MDSM_ERR_INVALID, "NULL buffer for a plan");
        if (strides[i] < (int64_t)b->plans[i]->obs.nsamp_in)
            return fail(MDSM_ERR_INVALID, "in_stride < nsamp_in for a plan");
    }
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (cur != b->device) return fail(MDSM_ERR_STATE, "batch belongs to another device");
    return MDSM_OK;
}

int ensure_work(mdsm_batch *b, void *const *works)
{
    for (size_t i = 0; i < b->plans.size(); ++i) {
        const mdsm_plan *p = b->plans[i];
        if (!p->work_bytes || (works && works[i]) || b->own_work[i]) continue;
        hipError_t e = hipMalloc(&b->own_work[i], p->work_bytes);
        if (e != hipSuccess) {
            b->own_work[i] = nullptr;
            return fail(MDSM_ERR_NOMEM, "batch workspace allocation failed");
        }
    }
    return MDSM_OK;
}

}  // namespace

extern "C" int mdsm_batch_create(mdsm_plan *const *plans, uint32_t nplans, mdsm_batch **out)
{
    if (!out || !plans) return fail(MDSM_ERR_INVALID, "NULL argument");
    *out = nullptr;
    if (nplans < 1 || nplans > (uint32_t)kMaxJobs)
        return fail(MDSM_ERR_INVALID, "a batch holds 1.." + std::to_string(kMaxJobs) + " plans");
    auto *b = new mdsm_batch();
    hipError_t e = hipGetDevice(&b->device);
    if (e != hipSuccess) {
        delete b;
        return hip_fail(e, "hipGetDevice");
    }
    for (uint32_t i = 0; i < nplans; ++i) {
        mdsm_plan *p = plans[i];
        if (!p) {
            delete b;
            return fail(MDSM_ERR_INVALID, "NULL plan in batch");
        }
        int rc = upload_plan(p);
        if (rc) {
            delete b;
            return rc;
        }
        if (p->device != b->device) {
            delete b;
            return fail(MDSM_ERR_STATE, "batch plans must live on the current device");
        }
        b->plans.push_back(p);
    }
    b->own_work.assign(nplans, nullptr);
    b->cus = b->plans[0]->cus;
    int rc = build_queue(b);
    if (rc == MDSM_OK) {
        e = hipMalloc((void **)&b->counters, std::max<size_t>(b->nstages, 1) * sizeof(uint32_t));
        if (e != hipSuccess) rc = hip_fail(e, "batch counters");
    }
    if (rc) {
        if (b->segs_dev) (void)hipFree(b->segs_dev);
        delete b;
        return rc;
    }
    *out = b;
    return MDSM_OK;
}

extern "C" void mdsm_batch_destroy(mdsm_batch *b)
{
    if (!b) return;
    drop_graph(b);
    if (b->segs_dev) (void)hipFree(b->segs_dev);
    if (b->counters) (void)hipFree(b->counters);
    for (void *w : b->own_work)
        if (w) (void)hipFree(w);
    delete b;
}

extern "C" int mdsm_batch_run(mdsm_batch *b, const void *const *d_ins, const int64_t *in_strides,
                              float *const *d_outs, void *const *d_works, void *stream)
{
    if (!b) return fail(MDSM_ERR_INVALID, "NULL batch");
    int rc = check_args(b, d_ins, in_strides, d_outs);
    if (rc) return rc;
    if ((rc = ensure_work(b, d_works))) return rc;
    return enqueue(b, d_ins, in_strides, d_outs, d_works, (hipStream_t)stream);
}

extern "C" int mdsm_batch_capture(mdsm_batch *b, const void *const *d_ins,
                                  const int64_t *in_strides, float *const *d_outs,
                                  void *const *d_works, void *stream)
{
    if (!b) return fail(MDSM_ERR_INVALID, "NULL batch");
    if (!stream) return fail(MDSM_ERR_INVALID, "graph capture needs a non-null stream");
    int rc = check_args(b, d_ins, in_strides, d_outs);
    if (rc) return rc;
    if ((rc = ensure_work(b, d_works))) return rc;
    drop_graph(b);
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
    // launch timing events are not recorded into the graph
    const bool was = timing_swap(false);
    rc = enqueue(b, d_ins, in_strides, d_outs, d_works, s);
    timing_swap(was);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(s, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    b->graph = g;
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        drop_graph(b);
        return hip_fail(e, "hipGraphInstantiate");
    }
    return MDSM_OK;
}

extern "C" int mdsm_batch_replay(mdsm_batch *b, void *stream)
{
    if (!b) return fail(MDSM_ERR_INVALID, "NULL batch");
    if (!b->exec) return fail(MDSM_ERR_STATE, "no captured graph (call mdsm_batch_capture)");
    TimedLaunch tl("batch_graph", (hipStream_t)stream);
    hipError_t e = hipGraphLaunch(b->exec, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
    return MDSM_OK;
}

extern "C" int mdsm_batch_autotune(mdsm_batch *b, const void *const *d_ins,
                                   const int64_t *in_strides, float *const *d_outs,
                                   void *const *d_works, void *stream, uint32_t reps,
                                   mdsm_tuning *best)
{
    if (!b) return fail(MDSM_ERR_INVALID, "NULL batch");
    if (reps < 1 || reps > 1000) return fail(MDSM_ERR_INVALID, "reps must be 1..1000");
    int rc = check_args(b, d_ins, in_strides, d_outs);
    if (rc) return rc;
    if ((rc = ensure_work(b, d_works))) return rc;
    drop_graph(b);  // a captured graph holds the old launch shapes
    std::vector<mdsm_tuning> saved;
    for (mdsm_plan *p : b->plans) saved.push_back(p->tune);
    hipStream_t s = (hipStream_t)stream;
    auto set_all = [&](const mdsm_tuning &t) -> int {
        for (mdsm_plan *p : b->plans) {
            int r = mdsm_plan_set_tuning(p, &t);
            if (r) return r;
        }
        return MDSM_OK;
    };
    mdsm_tuning found{};
    rc = tune_search(saved[0], set_all,
                     [&]() { return enqueue(b, d_ins, in_strides, d_outs, d_works, s); }, s, reps,
                     &found);
    if (rc) {
        for (size_t i = 0; i < b->plans.size(); ++i)
            (void)mdsm_plan_set_tuning(b->plans[i], &saved[i]);
        return rc;
    }
    if (best) *best = found;
    return MDSM_OK;
}
