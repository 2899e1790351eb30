// Streaming server: device-resident beams, HBM-budget admission and
// continuous batching of blocks from many beams into one dedispersion launch.
//
// Reference counterpart: the per-GPU worker loop and barrier protocol of
// src/mdsm/dedispersion_thread.cu:246-401 and src/mdsm/dedispersion_manager.cpp
// :324-501, which re-upload every block with a host-side overlap copy
// (dedispersion_thread.cu:266-303) and process one block at a time.
//
// Here a beam owns a device ring [C][M+N] whose first M samples per channel
// are the previous block's tail.  A block needs its beam's previous block to
// have finished (the overlap), so each beam has at most one block on the
// device; concurrency comes from many beams.  One scheduler thread admits
// blocks while transient buffers fit the HBM budget, launches the batch on
// one of two streams (H2D -> clip -> decimate -> one dedispersion launch for
// all blocks -> tail save -> boxcar detection -> D2H of candidates) and
// retires completed batches.  Every block's computation is a function of its
// own beam's data and tables only, so results do not depend on what it was
// batched with.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "mdsm_internal.h"

using namespace mdsm;

extern "C" double mdsm_now(void)
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

namespace {

constexpr int kStreams = 2;
inline uint64_t up256(uint64_t v) { return (v + 255) / 256 * 256; }

// Budget-capped caching device allocator.  Cached blocks count as used.
struct DevPool {
    uint64_t budget = 0, used = 0, peak = 0;
    std::multimap<uint64_t, void *> cache;

    // *held receives the size class the pointer carries (pass it to release)
    void *alloc(uint64_t bytes, uint64_t *held)
    {
        bytes = up256(bytes);
        auto it = cache.lower_bound(bytes);
        if (it != cache.end() && it->first <= bytes * 2) {
            void *p = it->second;
            *held = it->first;
            cache.erase(it);
            return p;
        }
        while (used + bytes > budget && !cache.empty()) {
            auto big = std::prev(cache.end());
            (void)hipFree(big->second);
            used -= big->first;
            cache.erase(big);
        }
        if (used + bytes > budget) return nullptr;
        void *p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
        used += bytes;
        peak = std::max(peak, used);
        *held = bytes;
        return p;
    }
    void release(void *p, uint64_t held)
    {
        if (p) cache.emplace(held, p);
    }
    void clear()
    {
        for (auto &kv : cache) {
            (void)hipFree(kv.second);
            used -= kv.first;
        }
        cache.clear();
    }
};

// Kinds of queued blocks.
enum : int32_t {
    PK_HOST = 0,    // samples in a pinned staging block
    PK_DEVICE = 1,  // samples in the caller's device buffer (mdsm_beam_push_device)
    PK_ZERO = 2,    // a missing block, processed as zeros
    PK_SKIP = 3     // a missing block whose history is all zeros: output known to be zero
};

struct Pending {
    uint64_t seq;
    double t_arrival;
    void *staging;        // PK_HOST
    const void *dsrc;     // PK_DEVICE
    int64_t dstride;      // PK_DEVICE, elements
    int32_t kind;
};

struct Beam;

struct Job {
    Beam *beam = nullptr;
    uint64_t seq = 0;
    double t_arrival = 0, t_admit = 0;
    int32_t kind = PK_HOST;
    const void *dsrc = nullptr;
    int64_t dstride = 0;
    bool carry_valid = false;  // the beam's carried samples continue this block's series
    bool carry_zero = false;   // ... and are zeros not yet written (after skipped blocks)
    void *staging = nullptr;
    void *buf = nullptr;     // transient device allocation
    uint64_t held = 0;       // its size class
    float *out = nullptr;
    char *work = nullptr;
    mdsm_cand *cands = nullptr;
    uint32_t *count = nullptr;
    double *mean = nullptr, *sigma = nullptr;
    int32_t *ncl = nullptr;
    char *dstage = nullptr;  // device copy of the new samples [C][N]
    int status = MDSM_OK;
    uint32_t batch = 0;
};

struct Beam {
    uint32_t id = 0;
    mdsm_obs obs{};
    uint32_t N = 0, M = 0;
    mdsm_plan *plan = nullptr;
    mdsm_detect_config det{};
    uint32_t wmask = 0;
    void *ring = nullptr, *tail = nullptr;
    float *carry = nullptr;    // detector carry [ndms_total][H]
    uint32_t H = 0;            // carried samples per DM row (0: no detection)
    bool carry_valid = false;  // set once a block has filled the carry
    bool carry_zero = false;   // carry is zeros not yet written (skipped blocks)
    uint32_t zero_run = 0;     // consecutive zero-filled blocks at the end of the queue
    uint64_t nfill = 0;        // zero-filled entries in the queue (not counted against its depth)
    uint32_t zneed = 1;        // zero blocks after which the history is all zeros
    int pushing = 0;           // pushes copying into staging without the lock
    bool draining = false;     // admitted while the server is held (close, export)
    uint64_t resident = 0;
    size_t esize = 1;
    std::deque<Pending> queue;
    std::vector<void *> free_staging;
    std::vector<void *> all_staging;
    std::vector<void *> acquired;   // handed out by mdsm_beam_acquire, not yet committed
    bool busy = false;
    bool closing = false;
    bool cancelling = false;
    uint64_t next_seq = 0;
    uint64_t trans_bytes = 0;   // transient bytes per block
    uint64_t off_out = 0, off_work = 0, off_cands = 0, off_count = 0, off_mean = 0,
             off_sigma = 0, off_ncl = 0, off_in = 0;
    // host result staging (pinned)
    mdsm_cand *h_cands = nullptr;
    uint32_t *h_count = nullptr;
    int32_t *h_ncl = nullptr;
    std::map<uint64_t, Job *> retained;
};

struct Batch {
    std::vector<Job *> jobs;
    hipEvent_t done = nullptr;
    int stream = 0;
    void *segs_dev = nullptr;
    uint64_t segs_held = 0;
    std::vector<SegDesc> segs_host;
};

struct Done {
    mdsm_result res;
    std::vector<mdsm_cand> cands;
};

Done host_result(Beam *b, const Pending &p, int status, uint32_t flags)
{
    Done d{};
    d.res.beam = b->id;
    d.res.seq = p.seq;
    d.res.status = status;
    d.res.flags = flags;
    d.res.t_arrival = p.t_arrival;
    d.res.t_admit = d.res.t_done = mdsm_now();
    return d;
}

}  // namespace

struct mdsm_server {
    mdsm_server_config cfg{};
    int device = 0;
    DevPool pool;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::map<uint32_t, Beam *> beams;
    uint32_t next_id = 1, rr = 0;
    std::deque<Batch *> inflight;
    uint32_t inflight_blocks = 0;
    std::deque<Done> done;
    std::thread sched;
    bool stop = false;
    bool held = false;  // admissions paused (mdsm_server_hold)
    hipStream_t streams[kStreams] = {nullptr, nullptr};
    uint32_t *counters = nullptr;  // persistent-mode work counters, one per stream
    // end of the latest dedispersion launch: the next one (on the other
    // stream) waits for it, so dedispersion launches never share the GPU and
    // each launch's duration is its own, while the small kernels and copies
    // of one stream still overlap the other's dedispersion
    hipEvent_t ded_done = nullptr;
    bool ded_recorded = false;
    int next_stream = 0;
    mdsm_server_stats st{};
    std::string sched_error;
};

namespace {

int beam_transient_layout(Beam *b, uint32_t cap)
{
    const mdsm_plan_info &pi = b->plan->info;
    const uint32_t nr = pi.nranges;
    uint64_t o = 0;
    b->off_out = o;
    o = up256(o + pi.out_elems * 4);
    b->off_work = o;  // the plan's workspace, then its row-statistics partials
    o = up256(up256(o + pi.work_bytes) + b->plan->stat_bytes);
    b->off_cands = o;
    o = up256(o + (uint64_t)nr * cap * sizeof(mdsm_cand));
    b->off_count = o;
    o = up256(o + nr * 4);
    b->off_mean = o;
    o = up256(o + (uint64_t)pi.ndms_total * 8);
    b->off_sigma = o;
    o = up256(o + (uint64_t)pi.ndms_total * 8);
    b->off_ncl = o;
    o = up256(o + (uint64_t)b->obs.nchans * 4);
    b->off_in = o;
    o = up256(o + (uint64_t)b->obs.nchans * b->N * b->esize);
    b->trans_bytes = o;
    return MDSM_OK;
}

void finish_job(mdsm_server *s, Job *j)
{
    Beam *b = j->beam;
    const uint32_t nr = b->plan->info.nranges;
    const uint32_t cap = b->det.cand_cap;
    Done d{};
    d.res.beam = b->id;
    d.res.seq = j->seq;
    d.res.status = j->status;
    d.res.batch = j->batch;
    d.res.t_arrival = j->t_arrival;
    d.res.t_admit = j->t_admit;
    if (j->kind == PK_ZERO) d.res.flags |= MDSM_RESULT_ZERO_FILLED;
    uint32_t total = 0;
    if (j->status == MDSM_OK && b->wmask) {
        for (uint32_t r = 0; r < nr; ++r) {
            const uint32_t n = b->h_count[r];
            total += n;
            if (n > cap) d.res.flags |= MDSM_RESULT_OVERFLOW;
            const uint32_t first = b->plan->ranges[r].info.first_dm;
            for (uint32_t i = 0; i < std::min(n, cap); ++i) {
                mdsm_cand c = b->h_cands[(size_t)r * cap + i];
                c.dm += first;
                d.cands.push_back(c);
            }
        }
        std::sort(d.cands.begin(), d.cands.end(), [](const mdsm_cand &a, const mdsm_cand &c) {
            if (a.dm != c.dm) return a.dm < c.dm;
            if (a.width != c.width) return a.width < c.width;
            return a.t < c.t;
        });
    }
    d.res.ncand = total;
    if (j->status == MDSM_OK && b->det.clip_nsigma > 0.0f && b->obs.dtype == MDSM_U8 &&
        j->kind != PK_ZERO) {
        uint64_t ncl = 0;
        for (uint32_t c = 0; c < b->obs.nchans; ++c) ncl += (uint32_t)b->h_ncl[c];
        d.res.nclipped = (uint32_t)std::min<uint64_t>(ncl, 0xffffffffu);
    }
    d.res.t_done = mdsm_now();
    if (j->staging) b->free_staging.push_back(j->staging);
    b->busy = false;
    if (b->det.retain_output && j->status == MDSM_OK) {
        b->retained[j->seq] = j;
    } else {
        s->pool.release(j->buf, j->held);
        delete j;
    }
    s->done.push_back(std::move(d));
    s->st.completed++;
}

// Enqueue one batch on a stream.  Called without the lock held; the jobs'
// beams are busy, so nothing else touches their rings.
int launch_batch(mdsm_server *s, Batch *bt)
{
    hipStream_t st = s->streams[bt->stream];
    // 1. copy-in, clip, decimate
    for (Job *j : bt->jobs) {
        Beam *b = j->beam;
        const size_t C = b->obs.nchans, M = b->M, N = b->N, es = b->esize;
        const size_t S = M + N;
        // one contiguous DMA of the block (or the caller's device block, or
        // zeros), then a kernel places it behind the overlap (the clipper
        // when it runs, a copy otherwise)
        hipError_t e = hipSuccess;
        const void *src = j->dstage;
        int64_t sstride = (int64_t)N;
        if (j->kind == PK_HOST) {
            e = hipMemcpyAsync(j->dstage, j->staging, C * N * es, hipMemcpyHostToDevice, st);
            if (e != hipSuccess) return hip_fail(e, "server H2D");
        } else if (j->kind == PK_DEVICE) {
            src = j->dsrc;
            sstride = j->dstride;
        }
        if (j->kind == PK_ZERO) {
            // a zero block is unchanged by the clipper (median 0, MAD 0)
            e = launch_zero2d((char *)b->ring + M * es, (int64_t)(S * es), (int64_t)(N * es),
                              (uint32_t)C, st);
            if (e != hipSuccess) return hip_fail(e, "server zero fill");
        } else if (b->det.clip_nsigma > 0.0f && b->obs.dtype == MDSM_U8) {
            TimedLaunch tl("clip", st);
            e = launch_clip_u8((const uint8_t *)src, sstride, (uint32_t)C, 0, (uint32_t)N,
                               b->det.clip_nsigma, (uint8_t *)b->ring + M, (int64_t)S, nullptr,
                               nullptr, j->ncl, st);
            if (e != hipSuccess) return hip_fail(e, "server clip");
        } else {
            e = launch_copy2d((char *)b->ring + M * es, (int64_t)(S * es), src,
                              (int64_t)(sstride * es), (int64_t)(N * es), (uint32_t)C, st);
            if (e != hipSuccess) return hip_fail(e, "server block placement");
        }
        if (!b->plan->decim.empty()) {
            TimedLaunch tl("decimate", st);
            e = launch_decimate_steps(b->ring, (uint32_t)C, (int64_t)S, (uint32_t)S,
                                      b->plan->decim.data(), b->plan->decim.size(), j->work, st);
            if (e != hipSuccess) return hip_fail(e, "server decimate");
        }
    }
    // 2. one dedispersion launch per stage for the whole batch
    size_t nstages = 0;
    for (Job *j : bt->jobs) nstages = std::max(nstages, j->beam->plan->stages.size());
    size_t seg_total = 0;
    for (Job *j : bt->jobs)
        for (const Stage &sg : j->beam->plan->stages) seg_total += sg.seg_end - sg.seg_begin;
    bt->segs_host.clear();
    bt->segs_host.reserve(seg_total);
    std::vector<LaunchSpec> specs(nstages);
    for (size_t k = 0; k < nstages; ++k) {
        LaunchSpec &L = specs[k];
        L.seg_begin = (int32_t)bt->segs_host.size();
        // longest items first, for the persistent work queue
        std::vector<std::pair<double, const SegDesc *>> order;
        std::vector<int> jobof;
        for (size_t ji = 0; ji < bt->jobs.size(); ++ji) {
            const mdsm_plan *p = bt->jobs[ji]->beam->plan;
            if (k >= p->stages.size()) continue;
            const Stage &sg = p->stages[k];
            for (int i = sg.seg_begin; i < sg.seg_end; ++i) {
                const SegDesc &d = p->segs[i];
                const BlockDesc &b0 = p->blocks[d.first_block];
                order.push_back({-(double)b0.nrows * b0.ntrials, &d});
                jobof.push_back((int)ji);
            }
            L.lds_bytes = std::max(L.lds_bytes, sg.lds_bytes);
            L.dtypes |= sg.dtypes;
        }
        std::vector<size_t> idx(order.size());
        for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
        std::stable_sort(idx.begin(), idx.end(),
                         [&](size_t a, size_t c) { return order[a].first < order[c].first; });
        int64_t items = 0;
        for (size_t i : idx) {
            SegDesc d = *order[i].second;
            d.job = jobof[i];
            d.item_start = items;
            items += (int64_t)d.nblocks * d.ntiles;
            bt->segs_host.push_back(d);
        }
        L.nseg = (int32_t)bt->segs_host.size() - L.seg_begin;
        L.items = items;
        L.njobs = (int32_t)bt->jobs.size();
        for (Job *j : bt->jobs)
            L.group = std::max<int32_t>(L.group, (int32_t)j->beam->plan->tune.tile_group);
        L.stats = 1;  // detecting beams' plans carry row statistics (plan_row_stats)
        for (size_t ji = 0; ji < bt->jobs.size(); ++ji) {
            Job *j = bt->jobs[ji];
            L.jobs[ji] = JobBufs{j->beam->ring, j->out, j->work, (int64_t)(j->beam->M + j->beam->N)};
        }
        if (s->cfg.persistent) {
            L.counter = s->counters + bt->stream;
            int dev = 0, cus = 256;
            if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                hipSuccess)
                cus = 256;
            L.grid = s->cfg.persistent > 1 ? (int32_t)s->cfg.persistent : cus * 4;
        }
    }
    hipError_t e = hipMemcpyAsync(bt->segs_dev, bt->segs_host.data(),
                                  bt->segs_host.size() * sizeof(SegDesc), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return hip_fail(e, "server segment upload");
    if (s->ded_recorded) {
        e = hipStreamWaitEvent(st, s->ded_done, 0);
        if (e != hipSuccess) return hip_fail(e, "server launch order");
    }
    for (LaunchSpec &L : specs) {
        L.segs = (const SegDesc *)bt->segs_dev;
        TimedLaunch tl("dedisperse", st);
        e = launch_dedisperse_spec(L, st);
        if (e != hipSuccess) return hip_fail(e, "server dedisperse");
        s->st.launches++;
    }
    e = hipEventRecord(s->ded_done, st);
    if (e != hipSuccess) return hip_fail(e, "server launch order");
    s->ded_recorded = true;
    // 3. overlap for the next block, detection, results back
    for (Job *j : bt->jobs) {
        Beam *b = j->beam;
        const size_t C = b->obs.nchans, M = b->M, N = b->N, es = b->esize, S = M + N;
        if (M) {
            if (N >= M) {
                e = launch_copy2d(b->ring, (int64_t)(S * es), (char *)b->ring + N * es,
                                  (int64_t)(S * es), (int64_t)(M * es), (uint32_t)C, st);
            } else {
                e = launch_copy2d(b->tail, (int64_t)(M * es), (char *)b->ring + N * es,
                                  (int64_t)(S * es), (int64_t)(M * es), (uint32_t)C, st);
                if (e == hipSuccess)
                    e = launch_copy2d(b->ring, (int64_t)(S * es), b->tail, (int64_t)(M * es),
                                      (int64_t)(M * es), (uint32_t)C, st);
            }
            if (e != hipSuccess) return hip_fail(e, "server overlap save");
        }
        const uint32_t nr = b->plan->info.nranges;
        const uint32_t cap = b->det.cand_cap;
        if (b->wmask) {
            SqrtW sw;
            for (int i = 0; i < 11; ++i) sw.v[i] = std::sqrt((double)(1u << i));
            const uint32_t H = b->H;
            if (j->carry_zero) {  // skipped blocks left a zero history
                e = launch_zero2d(b->carry, (int64_t)H * 4, (int64_t)H * 4,
                                  b->plan->info.ndms_total, st);
                if (e != hipSuccess) return hip_fail(e, "server carry reset");
            }
            for (uint32_t r = 0; r < nr; ++r) {
                const mdsm_range_info &ri = b->plan->ranges[r].info;
                // the block's share of the range's series (the rest of each row
                // repeats the next block's first samples)
                const uint32_t n = b->N / ri.downsample;
                float *carry = b->carry + (size_t)ri.first_dm * H;
                const BoxStream bs{carry, (int64_t)H, j->carry_valid ? (int32_t)H : 0, 1};
                // uint8 ranges at full rate: the series are integer sums of at
                // most 255 per channel (the clipper writes integer values), so
                // window sums below 2^24 are exact in float32
                const bool int_exact =
                    b->obs.dtype == MDSM_U8 && ri.downsample == 1 &&
                    255.0 * (double)C * (double)boxcar_maxw(b->wmask) < 16777216.0;
                // row statistics: from the dedispersion launch's partials
                // where the plan has them (uint8 at full rate), else a pass
                // over the output
                const RangeHost &rh = b->plan->ranges[r];
                const bool fused = rh.stat_tiles > 0 && (uint32_t)rh.stat_n == n;
                {
                    TimedLaunch tl("detect", st);
                    if (fused)
                        e = launch_row_partials(j->work + rh.stat_off, ri.ndms,
                                                (uint32_t)rh.stat_tiles, n, j->mean + ri.first_dm,
                                                j->sigma + ri.first_dm, st);
                    if (e == hipSuccess)
                        e = launch_boxcar_stream(j->out + ri.out_offset, ri.ndms, n, ri.nout, bs,
                                                 b->wmask, b->det.snr_thr, sw,
                                                 j->cands + (size_t)r * cap, cap, j->count + r,
                                                 j->mean + ri.first_dm, j->sigma + ri.first_dm,
                                                 !fused, st, int_exact);
                }
                if (e != hipSuccess) return hip_fail(e, "server detect");
                // the last H samples of each row carry over to the next block
                e = launch_copy2d(carry, (int64_t)H * 4, j->out + ri.out_offset + (n - H),
                                  (int64_t)ri.nout * 4, (int64_t)H * 4, ri.ndms, st);
                if (e != hipSuccess) return hip_fail(e, "server carry save");
            }
            e = hipMemcpyAsync(b->h_count, j->count, nr * 4, hipMemcpyDeviceToHost, st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(b->h_cands, j->cands, (size_t)nr * cap * sizeof(mdsm_cand),
                                   hipMemcpyDeviceToHost, st);
            if (e != hipSuccess) return hip_fail(e, "server D2H");
        }
        if (b->det.clip_nsigma > 0.0f && b->obs.dtype == MDSM_U8 && j->kind != PK_ZERO) {
            e = hipMemcpyAsync(b->h_ncl, j->ncl, C * 4, hipMemcpyDeviceToHost, st);
            if (e != hipSuccess) return hip_fail(e, "server D2H");
        }
    }
    e = hipEventRecord(bt->done, st);
    if (e != hipSuccess) return hip_fail(e, "server event");
    return MDSM_OK;
}

void scheduler(mdsm_server *s)
{
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (!s->stop) {
        bool progress = false;
        // retire completed batches (in order per stream, any order overall)
        for (auto it = s->inflight.begin(); it != s->inflight.end();) {
            Batch *bt = *it;
            hipError_t q = hipEventQuery(bt->done);
            if (q == hipErrorNotReady) {
                ++it;
                continue;
            }
            if (q != hipSuccess)
                for (Job *j : bt->jobs) j->status = MDSM_ERR_DEVICE;
            for (Job *j : bt->jobs) finish_job(s, j);
            s->inflight_blocks -= (uint32_t)bt->jobs.size();
            s->pool.release(bt->segs_dev, bt->segs_held);
            (void)hipEventDestroy(bt->done);
            delete bt;
            it = s->inflight.erase(it);
            progress = true;
        }
        if (progress) s->cv_done.notify_all();
        // blocks whose output is known to be zero complete on the host, in
        // their beam's order
        for (auto &kv : s->beams) {
            Beam *b = kv.second;
            while (!b->busy && !b->queue.empty() && b->queue.front().kind == PK_SKIP) {
                const Pending pd = b->queue.front();
                b->queue.pop_front();
                b->nfill--;
                s->done.push_back(
                    host_result(b, pd, MDSM_OK, MDSM_RESULT_ZERO_FILLED | MDSM_RESULT_SKIPPED));
                s->st.completed++;
                if (b->wmask) {
                    b->carry_zero = true;
                    b->carry_valid = true;
                }
                progress = true;
            }
        }
        if (progress) s->cv_done.notify_all();
        // admit a batch
        Batch *bt = nullptr;
        if (s->inflight.size() < (size_t)kStreams * 2) {
            std::vector<Beam *> order;
            for (auto &kv : s->beams) order.push_back(kv.second);
            const size_t nb = order.size();
            for (size_t i = 0; i < nb; ++i) {
                Beam *b = order[(s->rr + i) % nb];
                if (b->busy || b->queue.empty() || b->queue.front().kind == PK_SKIP) continue;
                if (s->held && !b->draining) continue;
                if (!bt) {
                    bt = new Batch();
                    bt->stream = s->next_stream;
                }
                if (bt->jobs.size() >= s->cfg.max_batch) break;
                if (s->inflight_blocks + bt->jobs.size() >= s->cfg.max_inflight) {
                    s->st.deferred_inflight++;
                    break;
                }
                uint64_t held = 0;
                void *buf = s->pool.alloc(b->trans_bytes, &held);
                if (!buf) {
                    s->st.deferred_budget++;
                    break;
                }
                Pending pd = b->queue.front();
                b->queue.pop_front();
                if (pd.kind == PK_ZERO) b->nfill--;
                Job *j = new Job();
                j->beam = b;
                j->seq = pd.seq;
                j->t_arrival = pd.t_arrival;
                j->t_admit = mdsm_now();
                j->staging = pd.staging;
                j->kind = pd.kind;
                j->dsrc = pd.dsrc;
                j->dstride = pd.dstride;
                j->carry_valid = b->carry_valid;
                j->carry_zero = b->carry_zero;
                b->carry_valid = b->wmask != 0;
                b->carry_zero = false;
                j->buf = buf;
                j->held = held;
                char *base = (char *)buf;
                j->out = (float *)(base + b->off_out);
                j->work = base + b->off_work;
                j->cands = (mdsm_cand *)(base + b->off_cands);
                j->count = (uint32_t *)(base + b->off_count);
                j->mean = (double *)(base + b->off_mean);
                j->sigma = (double *)(base + b->off_sigma);
                j->ncl = (int32_t *)(base + b->off_ncl);
                j->dstage = base + b->off_in;
                b->busy = true;
                bt->jobs.push_back(j);
                s->st.admitted++;
            }
            s->rr = nb ? (s->rr + 1) % nb : 0;
            if (bt && bt->jobs.empty()) {
                delete bt;
                bt = nullptr;
            }
        }
        if (bt) {
            size_t nseg = 0;
            for (Job *j : bt->jobs) nseg += j->beam->plan->segs.size();
            bt->segs_dev = s->pool.alloc(nseg * sizeof(SegDesc), &bt->segs_held);
            for (Job *j : bt->jobs) j->batch = (uint32_t)bt->jobs.size();
            int rc = bt->segs_dev ? MDSM_OK : MDSM_ERR_BUDGET;
            if (rc == MDSM_OK && hipEventCreateWithFlags(&bt->done, hipEventDisableTiming) != hipSuccess)
                rc = MDSM_ERR_DEVICE;
            if (rc == MDSM_OK) {
                s->next_stream = (s->next_stream + 1) % kStreams;
                s->inflight_blocks += (uint32_t)bt->jobs.size();
                lk.unlock();
                rc = launch_batch(s, bt);
                lk.lock();
                if (rc == MDSM_OK) {
                    s->inflight.push_back(bt);
                    progress = true;
                } else {
                    s->inflight_blocks -= (uint32_t)bt->jobs.size();
                }
            }
            if (rc != MDSM_OK) {
                // fail the blocks of this batch; the server stays usable
                if (bt->done) {
                    (void)hipStreamSynchronize(s->streams[bt->stream]);
                    (void)hipEventDestroy(bt->done);
                }
                for (Job *j : bt->jobs) {
                    j->status = rc;
                    finish_job(s, j);
                }
                s->pool.release(bt->segs_dev, bt->segs_held);
                delete bt;
                s->cv_done.notify_all();
            }
        }
        if (!progress) {
            // nothing to do right now: wait briefly for pushes or the GPU
            s->cv_work.wait_for(lk, std::chrono::microseconds(s->inflight.empty() ? 2000 : 100));
        }
    }
}

int upload_and_alloc_beam(mdsm_server *s, Beam *b)
{
    int rc = upload_plan(b->plan);
    if (rc) return rc;
    const size_t C = b->obs.nchans, S = (size_t)b->M + b->N;
    uint64_t ring_bytes = up256(C * S * b->esize);
    uint64_t tail_bytes = (b->N < b->M) ? up256(C * b->M * b->esize) : 0;
    uint64_t tables = b->plan->blocks.size() * sizeof(BlockDesc) +
                      b->plan->rel.size() * 4 + b->plan->abase.size() * 4 +
                      b->plan->segs.size() * sizeof(SegDesc);
    const uint64_t carry_bytes = up256((uint64_t)b->plan->info.ndms_total * b->H * 4);
    b->resident = ring_bytes + tail_bytes + carry_bytes + up256(tables);
    if (s->pool.used + b->resident + up256(b->trans_bytes) > s->pool.budget) {
        s->pool.clear();
        if (s->pool.used + b->resident + up256(b->trans_bytes) > s->pool.budget)
            return fail(MDSM_ERR_BUDGET, "beam needs " + std::to_string(b->resident) +
                                             " resident + " + std::to_string(b->trans_bytes) +
                                             " per-block bytes; budget " +
                                             std::to_string(s->pool.budget) + ", in use " +
                                             std::to_string(s->pool.used));
    }
    hipError_t e = hipMalloc(&b->ring, ring_bytes);
    if (e != hipSuccess) return hip_fail(e, "beam ring");
    e = hipMemset(b->ring, 0, ring_bytes);  // block 0 sees a zero history
    if (e != hipSuccess) return hip_fail(e, "beam ring clear");
    if (tail_bytes) {
        e = hipMalloc(&b->tail, tail_bytes);
        if (e != hipSuccess) return hip_fail(e, "beam tail");
    }
    if (b->H) {
        e = hipMalloc((void **)&b->carry, carry_bytes);
        if (e != hipSuccess) return hip_fail(e, "beam detector carry");
    }
    s->pool.used += b->resident;
    s->pool.peak = std::max(s->pool.peak, s->pool.used);
    const uint32_t nr = b->plan->info.nranges;
    if (hipHostMalloc((void **)&b->h_cands, (size_t)nr * b->det.cand_cap * sizeof(mdsm_cand)) !=
            hipSuccess ||
        hipHostMalloc((void **)&b->h_count, nr * 4) != hipSuccess ||
        hipHostMalloc((void **)&b->h_ncl, C * 4) != hipSuccess)
        return fail(MDSM_ERR_NOMEM, "beam host staging");
    const size_t block_bytes = C * b->N * b->esize;
    for (uint32_t i = 0; i < s->cfg.queue_depth + 1; ++i) {
        void *p = nullptr;
        if (hipHostMalloc(&p, block_bytes) != hipSuccess)
            return fail(MDSM_ERR_NOMEM, "beam block staging");
        b->all_staging.push_back(p);
        b->free_staging.push_back(p);
    }
    return MDSM_OK;
}

void free_beam(mdsm_server *s, Beam *b)
{
    for (auto &kv : b->retained) {
        s->pool.release(kv.second->buf, kv.second->held);
        delete kv.second;
    }
    b->retained.clear();
    if (b->ring) {
        (void)hipFree(b->ring);
        s->pool.used -= b->resident;
    }
    if (b->tail) (void)hipFree(b->tail);
    if (b->carry) (void)hipFree(b->carry);
    for (void *p : b->all_staging) (void)hipHostFree(p);
    if (b->h_cands) (void)hipHostFree(b->h_cands);
    if (b->h_count) (void)hipHostFree(b->h_count);
    if (b->h_ncl) (void)hipHostFree(b->h_ncl);
    if (b->plan) mdsm_plan_destroy(b->plan);
    delete b;
}

}  // namespace

extern "C" int mdsm_server_create(const mdsm_server_config *cfg, mdsm_server **out)
{
    if (!out) return fail(MDSM_ERR_INVALID, "out is NULL");
    *out = nullptr;
    auto *s = new mdsm_server();
    if (cfg) s->cfg = *cfg;
    if (!s->cfg.max_inflight) s->cfg.max_inflight = 16;
    if (!s->cfg.max_batch) s->cfg.max_batch = 16;
    s->cfg.max_batch = std::min<uint32_t>(s->cfg.max_batch, kMaxJobs);
    if (!s->cfg.queue_depth) s->cfg.queue_depth = 4;
    hipError_t e = hipGetDevice(&s->device);
    if (e != hipSuccess) {
        delete s;
        return hip_fail(e, "hipGetDevice");
    }
    size_t fr = 0, tot = 0;
    e = hipMemGetInfo(&fr, &tot);
    if (e != hipSuccess) {
        delete s;
        return hip_fail(e, "hipMemGetInfo");
    }
    s->pool.budget = s->cfg.hbm_budget ? s->cfg.hbm_budget : (uint64_t)(fr * 0.9);
    s->st.budget = s->pool.budget;
    for (auto &st : s->streams) {
        e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e != hipSuccess) {
            mdsm_server_destroy(s);
            return hip_fail(e, "server stream");
        }
    }
    e = hipMalloc((void **)&s->counters, kStreams * sizeof(uint32_t));
    if (e != hipSuccess) {
        mdsm_server_destroy(s);
        return hip_fail(e, "server counters");
    }
    e = hipEventCreateWithFlags(&s->ded_done, hipEventDisableTiming);
    if (e != hipSuccess) {
        mdsm_server_destroy(s);
        return hip_fail(e, "server event");
    }
    s->sched = std::thread(scheduler, s);
    *out = s;
    return MDSM_OK;
}

extern "C" void mdsm_server_destroy(mdsm_server *s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->sched.joinable()) s->sched.join();
    for (auto &st : s->streams)
        if (st) (void)hipStreamSynchronize(st);
    for (Batch *bt : s->inflight) {
        for (Job *j : bt->jobs) {
            s->pool.release(j->buf, j->held);
            delete j;
        }
        s->pool.release(bt->segs_dev, bt->segs_held);
        if (bt->done) (void)hipEventDestroy(bt->done);
        delete bt;
    }
    for (auto &kv : s->beams) free_beam(s, kv.second);
    s->pool.clear();
    if (s->counters) (void)hipFree(s->counters);
    if (s->ded_done) (void)hipEventDestroy(s->ded_done);
    for (auto &st : s->streams)
        if (st) (void)hipStreamDestroy(st);
    delete s;
}

extern "C" int mdsm_beam_open(mdsm_server *s, const mdsm_obs *obs, const mdsm_range *ranges,
                              uint32_t nranges, const mdsm_detect_config *det, uint32_t *beam_id)
{
    if (!s || !obs || !beam_id) return fail(MDSM_ERR_INVALID, "NULL argument");
    if (obs->nsamp_in < 1) return fail(MDSM_ERR_INVALID, "block length N must be >= 1");
    auto *b = new Beam();
    b->obs = *obs;
    b->N = obs->nsamp_in;
    b->esize = obs->dtype == MDSM_U8 ? 1 : (obs->dtype == MDSM_F16 ? 2 : 4);
    if (det) b->det = *det;
    if (!b->det.cand_cap) b->det.cand_cap = 65536;
    b->wmask = b->det.widths_mask;
    if (b->wmask & ~0x7ffu) {
        delete b;
        return fail(MDSM_ERR_INVALID, "boxcar widths must be powers of two <= 1024");
    }
    if (b->det.clip_nsigma > 0.0f && obs->dtype != MDSM_U8) {
        delete b;
        return fail(MDSM_ERR_INVALID, "the clipper runs on uint8 beams");
    }
    // overlap = the plan's maximum delay, rounded to whole decimated samples
    mdsm_obs o = *obs;
    o.nsamp_in = 1u << 30;
    mdsm_plan *probe = nullptr;
    int rc = mdsm_plan_create(&o, ranges, nranges, &probe);
    if (rc) {
        delete b;
        return rc;
    }
    uint32_t kmax = 1;
    for (auto &rh : probe->ranges) kmax = std::max(kmax, rh.info.downsample);
    uint32_t M = probe->info.maxshift;
    mdsm_plan_destroy(probe);
    M = (M + kmax - 1) / kmax * kmax;
    if (b->N % kmax) {
        delete b;
        return fail(MDSM_ERR_INVALID, "block length N must be a multiple of every downsample");
    }
    b->M = M;
    b->zneed = M ? (M + b->N - 1) / b->N : 1;
    o.nsamp_in = M + b->N;
    rc = mdsm_plan_create(&o, ranges, nranges, &b->plan);
    if (rc) {
        delete b;
        return rc;
    }
    if (b->wmask) {  // the detector's row statistics come with the dedispersion
        rc = plan_row_stats(b->plan, b->N);
        if (rc) {
            mdsm_plan_destroy(b->plan);
            delete b;
            return rc;
        }
    }
    if (b->wmask) {
        b->H = boxcar_carry_len(b->wmask);
        for (auto &rh : b->plan->ranges)
            if (b->N / rh.info.downsample < b->H) {
                mdsm_plan_destroy(b->plan);
                delete b;
                return fail(MDSM_ERR_INVALID,
                            "block too short for the detector: N/downsample must be >= " +
                                std::to_string(boxcar_carry_len(b->wmask)));
            }
    }
    beam_transient_layout(b, b->det.cand_cap);
    std::lock_guard<std::mutex> g(s->mu);
    (void)hipSetDevice(s->device);
    rc = upload_and_alloc_beam(s, b);
    if (rc) {
        free_beam(s, b);
        return rc;
    }
    b->id = s->next_id++;
    s->beams[b->id] = b;
    s->st.beams = (uint32_t)s->beams.size();
    *beam_id = b->id;
    return MDSM_OK;
}

extern "C" int mdsm_beam_info(mdsm_server *s, uint32_t id, uint32_t *overlap, mdsm_plan_info *info)
{
    if (!s) return fail(MDSM_ERR_INVALID, "NULL server");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->beams.find(id);
    if (it == s->beams.end()) return fail(MDSM_ERR_INVALID, "no such beam");
    if (overlap) *overlap = it->second->M;
    if (info) *info = it->second->plan->info;
    return MDSM_OK;
}

extern "C" int mdsm_beam_push(mdsm_server *s, uint32_t id, const void *host, int64_t stride,
                              double arrival_s, uint64_t *seq)
{
    if (!s || !host) return fail(MDSM_ERR_INVALID, "NULL argument");
    void *stg = nullptr;
    Beam *b = nullptr;
    {
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->beams.find(id);
        if (it == s->beams.end()) return fail(MDSM_ERR_INVALID, "no such beam");
        b = it->second;
        if (b->closing) return fail(MDSM_ERR_STATE, "beam is closing");
        if (stride < (int64_t)b->N) return fail(MDSM_ERR_INVALID, "stride < block length");
        if (b->queue.size() - b->nfill >= s->cfg.queue_depth || b->free_staging.empty())
            return fail(MDSM_ERR_AGAIN, "beam queue full");
        stg = b->free_staging.back();
        b->free_staging.pop_back();
        b->pushing++;  // close and cancel wait for the copy below
    }
    const size_t C = b->obs.nchans, N = b->N, es = b->esize;
    auto copy_rows = [&](size_t c0, size_t c1) {
        for (size_t c = c0; c < c1; ++c)
            std::memcpy((char *)stg + c * N * es, (const char *)host + c * (size_t)stride * es,
                        N * es);
    };
    const size_t bytes = C * N * es;
    const size_t nt = std::min<size_t>({8, C, bytes / (4u << 20) + 1});
    if (nt <= 1) {
        copy_rows(0, C);
    } else {  // a large block: copy channel ranges on several threads
        std::vector<std::thread> th;
        for (size_t i = 1; i < nt; ++i) th.emplace_back(copy_rows, C * i / nt, C * (i + 1) / nt);
        copy_rows(0, C / nt);
        for (auto &t : th) t.join();
    }
    {
        std::lock_guard<std::mutex> g(s->mu);
        b->pushing--;
        if (b->cancelling) {
            b->free_staging.push_back(stg);
            s->cv_done.notify_all();
            return fail(MDSM_ERR_STATE, "beam was cancelled");
        }
        Pending p{b->next_seq++, arrival_s >= 0 ? arrival_s : mdsm_now(), stg, nullptr, 0,
                  PK_HOST};
        b->queue.push_back(p);
        b->zero_run = 0;
        s->st.pushed++;
        if (seq) *seq = p.seq;
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
    return MDSM_OK;
}

extern "C" int mdsm_beam_push_seq(mdsm_server *s, uint32_t id, const void *host, int64_t stride,
                                  uint64_t seq, double arrival_s)
{
    if (!s || !host) return fail(MDSM_ERR_INVALID, "NULL argument");
    {
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->beams.find(id);
        if (it == s->beams.end()) return fail(MDSM_ERR_INVALID, "no such beam");
        Beam *b = it->second;
        if (b->closing) return fail(MDSM_ERR_STATE, "beam is closing");
        if (seq < b->next_seq)
            return fail(MDSM_ERR_STATE, "block " + std::to_string(seq) +
                                            " is a duplicate or late (next expected " +
                                            std::to_string(b->next_seq) + ")");
        const uint64_t gap = seq - b->next_seq;
        if (gap > (1u << 20)) return fail(MDSM_ERR_INVALID, "sequence gap longer than 2^20 blocks");
        if (stride < (int64_t)b->N) return fail(MDSM_ERR_INVALID, "stride < block length");
        if (b->queue.size() - b->nfill >= s->cfg.queue_depth || b->free_staging.empty())
            return fail(MDSM_ERR_AGAIN, "beam queue full");
        // the missing blocks go first, as zeros; once M samples of zeros are
        // in the history a zero block's output is exactly zero
        const double now = arrival_s >= 0 ? arrival_s : mdsm_now();
        for (uint64_t i = 0; i < gap; ++i) {
            const int32_t kind = b->zero_run >= b->zneed ? PK_SKIP : PK_ZERO;
            b->queue.push_back(Pending{b->next_seq++, now, nullptr, nullptr, 0, kind});
            b->zero_run++;
            b->nfill++;
            s->st.pushed++;
        }
    }
    uint64_t got = 0;
    int rc = mdsm_beam_push(s, id, host, stride, arrival_s, &got);
    if (rc == MDSM_OK && got != seq) return fail(MDSM_ERR_STATE, "concurrent pushes to one beam");
    return rc;
}

extern "C" int mdsm_beam_push_device(mdsm_server *s, uint32_t id, const void *d_block,
                                     int64_t stride, double arrival_s, uint64_t *seq)
{
    if (!s || !d_block) return fail(MDSM_ERR_INVALID, "NULL argument");
    {
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->beams.find(id);
        if (it == s->beams.end()) return fail(MDSM_ERR_INVALID, "no such beam");
        Beam *b = it->second;
        if (b->closing) return fail(MDSM_ERR_STATE, "beam is closing");
        if (stride < (int64_t)b->N) return fail(MDSM_ERR_INVALID, "stride < block length");
        if (b->queue.size() - b->nfill >= s->cfg.queue_depth)
            return fail(MDSM_ERR_AGAIN, "beam queue full");
        Pending p{b->next_seq++, arrival_s >= 0 ? arrival_s : mdsm_now(), nullptr, d_block,
                  stride, PK_DEVICE};
        b->queue.push_back(p);
        b->zero_run = 0;
        s->st.pushed++;
        if (seq) *seq = p.seq;
    }
    s->cv_work.notify_all();
    return MDSM_OK;
}

extern "C" int mdsm_server_poll(mdsm_server *s, mdsm_result *res, mdsm_cand *cands, uint32_t cap,
                                int32_t wait_ms)
{
    if (!s || !res) return fail(MDSM_ERR_INVALID, "NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto ready = [&] { return !s->done.empty(); };
    if (!ready()) {
        if (wait_ms < 0)
            s->cv_done.wait(lk, ready);
        else if (wait_ms > 0)
            s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), ready);
    }
    if (!ready()) return MDSM_ERR_AGAIN;
    Done d = std::move(s->done.front());
    s->done.pop_front();
    *res = d.res;
    const size_t nret = (cands && cap) ? std::min<size_t>(cap, d.cands.size()) : 0;
    if (nret) std::memcpy(cands, d.cands.data(), nret * sizeof(mdsm_cand));
    res->nreturned = (uint32_t)nret;
    return MDSM_OK;
}

extern "C" int mdsm_server_output(mdsm_server *s, uint32_t id, uint64_t seq, const float **d_out)
{
    if (!s || !d_out) return fail(MDSM_ERR_INVALID, "NULL argument");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->beams.find(id);
    if (it == s->beams.end()) return fail(MDSM_ERR_INVALID, "no such beam");
    auto r = it->second->retained.find(seq);
    if (r == it->second->retained.end()) return fail(MDSM_ERR_STATE, "block output not retained");
    *d_out = r->second->out;
    return MDSM_OK;
}

extern "C" int mdsm_server_release(mdsm_server *s, uint32_t id, uint64_t seq)
{
    if (!s) return fail(MDSM_ERR_INVALID, "NULL argument");
    {
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->beams.find(id);
        if (it == s->beams.end()) return fail(MDSM_ERR_INVALID, "no such beam");
        auto r = it->second->retained.find(seq);
        if (r == it->second->retained.end()) return fail(MDSM_ERR_STATE, "block output not retained");
        s->pool.release(r->second->buf, r->second->held);
        delete r->second;
        it->second->retained.erase(r);
    }
    s->cv_work.notify_all();
    return MDSM_OK;
}

extern "C" int mdsm_server_hold(mdsm_server *s, int hold)
{
    if (!s) return fail(MDSM_ERR_INVALID, "NULL argument");
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->held = hold != 0;
    }
    s->cv_work.notify_all();
    return MDSM_OK;
}

extern "C" int mdsm_server_drain(mdsm_server *s)
{
    if (!s) return fail(MDSM_ERR_INVALID, "NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    s->held = false;  // a held server would never drain
    for (;;) {
        bool idle = s->inflight.empty();
        for (auto &kv : s->beams) idle = idle && kv.second->queue.empty() && !kv.second->busy;
        if (idle) return MDSM_OK;
        s->cv_work.notify_all();
        s->cv_done.wait_for(lk, std::chrono::milliseconds(5));
    }
}

namespace {

// Wait (lock held) until the beam has nothing queued, on the device or being
// copied in.  Its blocks are admitted even while the server is held.
void wait_beam_idle(mdsm_server *s, Beam *b, std::unique_lock<std::mutex> &lk)
{
    b->draining = true;
    while (!b->queue.empty() || b->busy || b->pushing > 0) {
        s->cv_work.notify_all();
        s->cv_done.wait_for(lk, std::chrono::milliseconds(5));
    }
    b->draining = false;
}

int close_beam(mdsm_server *s, uint32_t id, bool cancel)
{
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->beams.find(id);
    if (it == s->beams.end()) return fail(MDSM_ERR_INVALID, "no such beam");
    Beam *b = it->second;
    if (b->closing) return fail(MDSM_ERR_STATE, "beam is already closing");
    b->closing = true;
    if (cancel) {
        b->cancelling = true;
        for (;;) {
            // drop what is queued (pushes still copying drop their own block)
            bool dropped = false;
            while (!b->queue.empty()) {
                const Pending pd = b->queue.front();
                b->queue.pop_front();
                if (pd.kind == PK_ZERO || pd.kind == PK_SKIP) b->nfill--;
                if (pd.staging) b->free_staging.push_back(pd.staging);
                s->done.push_back(host_result(b, pd, MDSM_ERR_CANCELLED, 0));
                dropped = true;
            }
            if (dropped) s->cv_done.notify_all();
            if (!b->busy && b->pushing == 0) break;
            s->cv_done.wait_for(lk, std::chrono::milliseconds(5));
        }
    } else {
        wait_beam_idle(s, b, lk);
    }
    s->beams.erase(it);
    s->st.beams = (uint32_t)s->beams.size();
    free_beam(s, b);
    return MDSM_OK;
}

}  // namespace

extern "C" int mdsm_beam_close(mdsm_server *s, uint32_t id)
{
    if (!s) return fail(MDSM_ERR_INVALID, "NULL argument");
    return close_beam(s, id, false);
}

extern "C" int mdsm_beam_cancel(mdsm_server *s, uint32_t id)
{
    if (!s) return fail(MDSM_ERR_INVALID, "NULL argument");
    return close_beam(s, id, true);
}

namespace {

// Serialised beam state (mdsm_beam_export): this header, the overlap
// [nchans][M] samples of the beam's dtype, then the carry [ndms][H] floats.
struct BeamStateHeader {
    char magic[8];
    uint32_t nchans, N, M, esize, ndms, H, carry_valid, zero_run;
    uint64_t next_seq;
    uint64_t shape;  // hash of the plan's ranges, widths and clip setting
};
constexpr char kStateMagic[8] = {'M', 'D', 'S', 'M', 'B', 'S', 'T', '1'};

uint64_t beam_shape_hash(const Beam *b)
{
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void *p, size_t n) {
        for (size_t i = 0; i < n; ++i) {
            h ^= ((const unsigned char *)p)[i];
            h *= 1099511628211ull;
        }
    };
    mix(&b->obs, sizeof(b->obs));
    for (const auto &rh : b->plan->ranges) {
        mix(&rh.info.lowdm, sizeof(float));
        mix(&rh.info.dmstep, sizeof(float));
        mix(&rh.info.ndms, sizeof(uint32_t));
        mix(&rh.info.downsample, sizeof(uint32_t));
    }
    mix(&b->wmask, sizeof(b->wmask));
    mix(&b->det.clip_nsigma, sizeof(float));
    return h;
}

uint64_t beam_state_bytes(const Beam *b)
{
    return sizeof(BeamStateHeader) + (uint64_t)b->obs.nchans * b->M * b->esize +
           (uint64_t)b->plan->info.ndms_total * b->H * 4;
}

}  // namespace

extern "C" int mdsm_beam_export(mdsm_server *s, uint32_t id, void *host, uint64_t cap,
                                uint64_t *bytes)
{
    if (!s || !bytes) return fail(MDSM_ERR_INVALID, "NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->beams.find(id);
    if (it == s->beams.end()) return fail(MDSM_ERR_INVALID, "no such beam");
    Beam *b = it->second;
    *bytes = beam_state_bytes(b);
    if (!host || cap < *bytes) return fail(MDSM_ERR_CAPACITY, "state buffer too small");
    if (b->closing) return fail(MDSM_ERR_STATE, "beam is closing");
    b->closing = true;  // no pushes while the state is read
    wait_beam_idle(s, b, lk);
    BeamStateHeader h{};
    std::memcpy(h.magic, kStateMagic, 8);
    h.nchans = b->obs.nchans;
    h.N = b->N;
    h.M = b->M;
    h.esize = (uint32_t)b->esize;
    h.ndms = b->plan->info.ndms_total;
    h.H = b->H;
    h.carry_valid = b->carry_valid ? 1u : 0u;
    h.zero_run = b->zero_run;
    h.next_seq = b->next_seq;
    h.shape = beam_shape_hash(b);
    char *p = (char *)host;
    std::memcpy(p, &h, sizeof(h));
    p += sizeof(h);
    (void)hipSetDevice(s->device);
    const size_t C = b->obs.nchans, S = (size_t)b->M + b->N, es = b->esize;
    hipError_t e = hipSuccess;
    if (b->M)
        e = hipMemcpy2D(p, b->M * es, b->ring, S * es, b->M * es, C, hipMemcpyDeviceToHost);
    p += C * b->M * es;
    const size_t cbytes = (size_t)h.ndms * b->H * 4;
    if (e == hipSuccess && cbytes) {
        if (b->carry_zero)
            std::memset(p, 0, cbytes);
        else
            e = hipMemcpy(p, b->carry, cbytes, hipMemcpyDeviceToHost);
    }
    b->closing = false;
    if (e != hipSuccess) return hip_fail(e, "beam export");
    return MDSM_OK;
}

extern "C" int mdsm_beam_import(mdsm_server *s, uint32_t id, const void *host, uint64_t bytes)
{
    if (!s || !host) return fail(MDSM_ERR_INVALID, "NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->beams.find(id);
    if (it == s->beams.end()) return fail(MDSM_ERR_INVALID, "no such beam");
    Beam *b = it->second;
    if (b->next_seq != 0 || b->busy || !b->queue.empty() || b->pushing || b->closing)
        return fail(MDSM_ERR_STATE, "import needs a beam no block was pushed to");
    if (bytes != beam_state_bytes(b)) return fail(MDSM_ERR_INVALID, "state size does not match");
    BeamStateHeader h;
    std::memcpy(&h, host, sizeof(h));
    if (std::memcmp(h.magic, kStateMagic, 8) != 0) return fail(MDSM_ERR_INVALID, "not a beam state");
    if (h.nchans != b->obs.nchans || h.N != b->N || h.M != b->M || h.esize != b->esize ||
        h.ndms != b->plan->info.ndms_total || h.H != b->H || h.shape != beam_shape_hash(b))
        return fail(MDSM_ERR_INVALID, "beam state is of another observation, plan or detector");
    const char *p = (const char *)host + sizeof(h);
    (void)hipSetDevice(s->device);
    const size_t C = b->obs.nchans, S = (size_t)b->M + b->N, es = b->esize;
    hipError_t e = hipSuccess;
    if (b->M) e = hipMemcpy2D(b->ring, S * es, p, b->M * es, b->M * es, C, hipMemcpyHostToDevice);
    p += C * b->M * es;
    const size_t cbytes = (size_t)h.ndms * b->H * 4;
    if (e == hipSuccess && cbytes) e = hipMemcpy(b->carry, p, cbytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "beam import");
    b->next_seq = h.next_seq;
    b->carry_valid = h.carry_valid != 0;
    b->carry_zero = false;
    b->zero_run = h.zero_run;
    return MDSM_OK;
}

extern "C" int mdsm_server_get_stats(mdsm_server *s, mdsm_server_stats *st)
{
    if (!s || !st) return fail(MDSM_ERR_INVALID, "NULL argument");
    std::lock_guard<std::mutex> g(s->mu);
    *st = s->st;
    st->bytes_in_use = s->pool.used;
    st->bytes_peak = s->pool.peak;
    st->inflight = s->inflight_blocks;
    return MDSM_OK;
}

extern "C" int mdsm_beam_acquire(mdsm_server *s, uint32_t id, void **host_block)
{
    if (!s || !host_block) return fail(MDSM_ERR_INVALID, "NULL argument");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->beams.find(id);
    if (it == s->beams.end()) return fail(MDSM_ERR_INVALID, "no such beam");
    Beam *b = it->second;
    if (b->closing) return fail(MDSM_ERR_STATE, "beam is closing");
    if (b->queue.size() - b->nfill + b->acquired.size() >= s->cfg.queue_depth || b->free_staging.empty())
        return fail(MDSM_ERR_AGAIN, "beam queue full");
    void *stg = b->free_staging.back();
    b->free_staging.pop_back();
    b->acquired.push_back(stg);
    *host_block = stg;
    return MDSM_OK;
}

extern "C" int mdsm_beam_commit(mdsm_server *s, uint32_t id, void *host_block, double arrival_s,
                                uint64_t *seq)
{
    if (!s || !host_block) return fail(MDSM_ERR_INVALID, "NULL argument");
    {
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->beams.find(id);
        if (it == s->beams.end()) return fail(MDSM_ERR_INVALID, "no such beam");
        Beam *b = it->second;
        auto a = std::find(b->acquired.begin(), b->acquired.end(), host_block);
        if (a == b->acquired.end())
            return fail(MDSM_ERR_INVALID, "buffer was not acquired from this beam");
        b->acquired.erase(a);
        if (b->closing) {
            b->free_staging.push_back(host_block);
            return fail(MDSM_ERR_STATE, "beam is closing");
        }
        Pending p{b->next_seq++, arrival_s >= 0 ? arrival_s : mdsm_now(), host_block, nullptr, 0,
                  PK_HOST};
        b->queue.push_back(p);
        b->zero_run = 0;
        s->st.pushed++;
        if (seq) *seq = p.seq;
    }
    s->cv_work.notify_all();
    return MDSM_OK;
}
