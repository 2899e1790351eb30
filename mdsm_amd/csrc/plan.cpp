// Host-side plan builder: shift tables, output layout, device work tables.
//
// Shift tables are integer tables built once per plan on the host, so the
// rounding of the delay arithmetic is decided here, in one place, and every
// kernel reads plain int32 offsets.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdio>
#include <functional>
#include <limits>
#include <map>
#include <mutex>

#include "mdsm_internal.h"

namespace mdsm {
namespace {

inline int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

// Delay model of one observation under one shift rule.
struct ShiftModel {
    int rule = MDSM_SHIFT_REF;
    std::vector<float> delay_f;   // REF: seconds per unit DM, float32
    std::vector<double> delay_d;  // NS: f^-2 - f_top^-2 (MHz^-2)

    ShiftModel(const mdsm_obs &o)
    {
        rule = o.shift_rule;
        delay_f.resize(o.nchans);
        delay_d.resize(o.nchans);
        for (uint32_t c = 0; c < o.nchans; ++c) {
            // REF: channel frequency formed in float, delay evaluated in double
            // and narrowed to float (src/mdsm/dedispersion_manager.cpp:262-265,
            // 287-289; src/prototypes/cpu_dedisp.c:52-55).
            float fc = o.foff * (float)c;
            fc = o.fch1 + fc;
            double a = 1.0 / (double)fc;
            a = a / (double)fc;
            double b = 1.0 / (double)o.fch1;
            b = b / (double)o.fch1;
            delay_f[c] = (float)(4148.741601 * (a - b));
            // NS: FP64 physical delay, constant 4.148808e3 s MHz^2.
            double fd = (double)o.fch1 + (double)o.foff * (double)c;
            delay_d[c] = 1.0 / (fd * fd) - 1.0 / ((double)o.fch1 * (double)o.fch1);
        }
    }

    // REF: (int)(delay * dm / tsamp) in float32, left to right, truncated
    // (src/prototypes/cpu_dedisp.c:72).  NS: nearest sample in FP64.
    int32_t shift(uint32_t c, float dm_f, double dm_d, float tsamp_f, double tsamp_d) const
    {
        if (rule == MDSM_SHIFT_REF) {
            float v = delay_f[c] * dm_f;
            v = v / tsamp_f;
            return (int32_t)v;
        }
        double v = 4.148808e3 * dm_d * delay_d[c];
        return (int32_t)std::llround(v / tsamp_d);
    }
};

int validate_obs(const mdsm_obs *o)
{
    if (!o) return fail(MDSM_ERR_INVALID, "obs is NULL");
    if (o->nchans < 1 || o->nchans > (1u << 20))
        return fail(MDSM_ERR_INVALID, "nchans out of range");
    if (o->nsamp_in < 1) return fail(MDSM_ERR_INVALID, "nsamp_in must be >= 1");
    if (!(o->tsamp > 0.0f) || !std::isfinite(o->tsamp))
        return fail(MDSM_ERR_INVALID, "tsamp must be > 0");
    if (!(o->fch1 > 0.0f) || !std::isfinite(o->fch1))
        return fail(MDSM_ERR_INVALID, "fch1 must be > 0");
    if (!(o->foff < 0.0f))
        return fail(MDSM_ERR_INVALID, "foff must be negative (channel 0 is the top of the band)");
    float flow = o->fch1 + o->foff * (float)(o->nchans - 1);
    if (!(flow > 0.0f)) return fail(MDSM_ERR_INVALID, "lowest channel frequency must be > 0");
    if (o->dtype < MDSM_U8 || o->dtype > MDSM_F32) return fail(MDSM_ERR_INVALID, "bad dtype");
    if (o->shift_rule != MDSM_SHIFT_REF && o->shift_rule != MDSM_SHIFT_NS)
        return fail(MDSM_ERR_INVALID, "bad shift rule");
    return MDSM_OK;
}

// Appends one segment: trials [0, ntrials) over rows [0, nrows), split into
// blocks of at most kDB trials whose LDS window fits max_window(dtype).
struct SegSpec {
    int ntrials, nrows;
    std::function<int32_t(int, int)> sh;  // (trial, row) -> shift >= 0
    int src;
    int64_t row0, row_stride, in_off;
    int in_len, dtype;
    int dst;
    int64_t out_off, out_trial_stride;
    int nout, k;
    int maxval;  // largest sample value of integer data (uint8: 255; k-sums: 255k)
};

int add_segment(mdsm_plan *p, Stage &st, const SegSpec &s)
{
    SegDesc seg{};
    seg.first_block = (int32_t)p->blocks.size();
    seg.ntiles = (int32_t)((s.nout + kTile - 1) / kTile);
    std::vector<int32_t> base(s.nrows), rspread(s.nrows);
    int j0 = 0;
    while (j0 < s.ntrials) {
        int nt = std::min(kDB, s.ntrials - j0);
        int64_t relmax = 0;
        for (;;) {
            relmax = 0;
            for (int i = 0; i < s.nrows; ++i) {
                int32_t lo = std::numeric_limits<int32_t>::max(), hi = 0;
                for (int j = 0; j < nt; ++j) {
                    int32_t v = s.sh(j0 + j, i);
                    lo = std::min(lo, v);
                    hi = std::max(hi, v);
                }
                if (lo < 0) return fail(MDSM_ERR_INVALID, "negative shift in plan");
                base[i] = lo & ~15;
                rspread[i] = hi - base[i];
                relmax = std::max<int64_t>(relmax, hi - base[i]);
            }
            if (round_up(kTile + relmax, 64) <= max_window(s.dtype) || nt == 1) break;
            nt = (nt + 1) / 2;
        }
        BlockDesc b{};
        b.row0 = s.row0;
        b.row_stride = s.row_stride;
        b.in_off = s.in_off;
        b.out_off = s.out_off + (int64_t)j0 * s.out_trial_stride;
        b.out_trial_stride = s.out_trial_stride;
        b.in_len = s.in_len;
        b.nrows = s.nrows;
        b.ntrials = nt;
        b.W = (int32_t)round_up(kTile + relmax, 64);
        b.nout = s.nout;
        b.k = s.k;
        b.src = s.src;
        b.dst = s.dst;
        b.dtype = s.dtype;
        b.flush = int_path(s.dtype) ? std::max(1, 65535 / std::max(1, s.maxval)) : 0;
        b.tab_off = (int32_t)p->rel0.size();
        b.base_off = (int32_t)p->abase.size();
        if (p->rel0.size() + (size_t)s.nrows * kDB > (size_t)std::numeric_limits<int32_t>::max())
            return fail(MDSM_ERR_INVALID, "plan shift tables exceed 2^31 entries");
        // Offsets are stored as byte offsets within the block's four-phase LDS
        // row (phase o&3, entry o>>2); apply_tuning adds the row's place in
        // its chunk, so the kernel adds them to a lane address without
        // further arithmetic.
        // Trial t of a block sits in the table column its wave reads: column
        // t with kRPW trials per wave; with kRPW/2 (float blocks of <= kDB/2
        // trials, which then keep every wave busy) column (t / R) * kRPW + t % R.
        b.rpw = (nt <= kDB / 2 && !int_path(s.dtype)) ? kRPW / 2 : kRPW;
        std::vector<int> col_trial(kDB, -1);
        for (int t = 0; t < nt; ++t) col_trial[(t / b.rpw) * kRPW + t % b.rpw] = t;
        // Each row is staged over its own window: the tile plus its own
        // spread of shifts over the block's trials (rows of the upper band
        // spread far less than the block's widest), rounded up to 64
        // samples.  The base table holds (window start, row info) pairs; the
        // row info carries the window's entries per phase (bits 0-9) and,
        // from apply_tuning, the row's byte place in its chunk buffer.
        for (int i = 0; i < s.nrows; ++i) {
            const int Wi = (int)round_up(kTile + rspread[i], 64);
            p->abase.push_back(base[i]);
            p->abase.push_back(Wi / 4);
            for (int j = 0; j < kDB; ++j) {
                const int t = col_trial[j];
                const int32_t o = t >= 0 ? s.sh(j0 + t, i) - base[i] : 0;
                p->rel0.push_back(lds_offset(s.dtype, Wi, o));
            }
        }
        st.wmax = std::max(st.wmax, b.W);
        st.dtypes |= 1 << b.dtype;
        p->blocks.push_back(b);
        j0 += nt;
    }
    // 8 zero rows after every block's table: the integer path stages 8 rows of
    // offsets per chunk without a bounds test
    for (int i = 0; i < 8 * kDB; ++i) p->rel0.push_back(0);
    seg.nblocks = (int32_t)p->blocks.size() - seg.first_block;
    seg.item_start = st.items;
    st.items += (int64_t)seg.nblocks * seg.ntiles;
    p->segs.push_back(seg);
    st.seg_end = (int32_t)p->segs.size();
    return MDSM_OK;
}

int add_decimation(mdsm_plan *p, int k, uint64_t &work)
{
    for (size_t i = 0; i < p->decim.size(); ++i)
        if (p->decim[i].k == k) return (int)i;
    DecimStep d;
    d.k = k;
    d.dtype_in = p->obs.dtype;
    d.dtype_out = (p->obs.dtype == MDSM_U8 && k <= 16) ? DT_U16 : DT_F32;
    d.nsamp_out = (int32_t)(p->obs.nsamp_in / (uint32_t)k);
    d.stride = round_up(d.nsamp_out, 64);
    d.work_off = (int64_t)round_up((int64_t)work, 256);
    work = d.work_off + (uint64_t)p->obs.nchans * d.stride * dtype_size(d.dtype_out);
    p->decim.push_back(d);
    return (int)p->decim.size() - 1;
}

// Channels per LDS chunk of every block and each launch's dynamic LDS, from
// the plan's tuning: the requested rows (2..8, 8 by default), or fewer where
// a block's window is too wide for the LDS budget.  Chunks are
// double-buffered.
void apply_tuning(mdsm_plan *p)
{
    const int budget = p->tune.lds_budget ? (int)p->tune.lds_budget : kLdsBudget;
    p->rel = p->rel0;
    for (Stage &st : p->stages) {
        st.lds_bytes = 0;
        for (int si = st.seg_begin; si < st.seg_end; ++si) {
            const SegDesc &sg = p->segs[si];
            for (int bi = sg.first_block; bi < sg.first_block + sg.nblocks; ++bi) {
                BlockDesc &b = p->blocks[bi];
                const int rb = lds_row_bytes(b.dtype, b.W);  // the block's widest row
                const int want = p->tune.chunk_rows ? (int)p->tune.chunk_rows : kChunkRows;
                b.ccr = std::max(2, std::min({8, want, budget / (2 * rb)}));
                // rows of a chunk sit one after another in its buffer, each
                // at its own width; the buffer is the widest chunk's rows
                int place = 0, widest = 0;
                for (int i = 0; i < b.nrows; ++i) {
                    if (i % b.ccr == 0) place = 0;
                    int32_t &info = p->abase[(size_t)b.base_off + 2 * (size_t)i + 1];
                    const int Ki = info & 1023;
                    info = Ki | (place << 10);
                    int32_t *row = &p->rel[(size_t)b.tab_off + (size_t)i * kDB];
                    for (int j = 0; j < kDB; ++j) row[j] += place;
                    place += lds_row_bytes(b.dtype, 4 * Ki);
                    widest = std::max(widest, place);
                }
                b.bufsz = widest;
                st.lds_bytes = std::max(st.lds_bytes, 2 * b.bufsz);
            }
        }
    }
}

int upload(mdsm_plan *p)
{
    hipError_t e = hipGetDevice(&p->device);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    p->uploaded = true;
    auto up = [&](void **dst, const void *src, size_t bytes) -> hipError_t {
        hipError_t r = hipMalloc(dst, bytes ? bytes : 16);
        if (r != hipSuccess) return r;
        if (src && bytes) return hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
        return hipSuccess;
    };
    if ((e = up((void **)&p->dev.blocks, p->blocks.data(), p->blocks.size() * sizeof(BlockDesc))) !=
            hipSuccess ||
        (e = up((void **)&p->dev.rel, p->rel.data(), p->rel.size() * sizeof(int32_t))) !=
            hipSuccess ||
        (e = up((void **)&p->dev.abase, p->abase.data(), p->abase.size() * sizeof(int32_t))) !=
            hipSuccess)
        return hip_fail(e, "uploading plan tables");
    for (SegDesc &sg : p->segs) {
        sg.blocks = p->dev.blocks;
        sg.rel = p->dev.rel;
        sg.abase = p->dev.abase;
        sg.job = 0;
    }
    if ((e = up((void **)&p->dev.segs, p->segs.data(), p->segs.size() * sizeof(SegDesc))) !=
            hipSuccess ||
        (e = up((void **)&p->dev.counter, nullptr, p->stages.size() * sizeof(uint32_t))) !=
            hipSuccess)
        return hip_fail(e, "uploading plan tables");
    if ((e = hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, p->device)) !=
        hipSuccess)
        return hip_fail(e, "hipDeviceGetAttribute");
    return MDSM_OK;
}

// Launch shapes chosen by the autotuner, per plan shape, for this process.
std::mutex g_tune_mu;
std::map<std::string, mdsm_tuning> g_tune_cache;

std::string shape_key_of(const mdsm_obs &o, const std::string &plan_part)
{
    char buf[256];
    std::snprintf(buf, sizeof buf, "%u|%.9g|%.9g|%.9g|%u|%d|%d|", o.nchans, (double)o.fch1,
                  (double)o.foff, (double)o.tsamp, o.nsamp_in, o.dtype, o.shift_rule);
    return std::string(buf) + plan_part;
}

void adopt_cached_tuning(mdsm_plan *p)
{
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tune_cache.find(p->shape_key);
    if (it != g_tune_cache.end()) p->tune = it->second;
}

void finish_info(mdsm_plan *p, double adds, double bytes, double dcs)
{
    mdsm_plan_info &in = p->info;
    in.nranges = (uint32_t)p->ranges.size();
    in.nchans = p->obs.nchans;
    in.work_bytes = p->work_bytes;
    in.adds = adds;
    in.bytes = bytes;
    in.dm_chan_samples = dcs;
    in.is_subband = p->subband ? 1u : 0u;
    in.nblocks = (uint32_t)p->blocks.size();
}

}  // namespace

int upload_plan(mdsm_plan *p) { return p->uploaded ? MDSM_OK : upload(p); }

// Row statistics written by the dedispersion launch (launches with stats
// on, the server's): uint8 ranges at full rate, over each DM row's first n
// samples.  The values are integers (float32 sums below 2^24), so the sums
// are exact in FP64 in any order while n (255 C)^2 < 2^53; ranges beyond
// that keep none.  The partials sit after the plan's workspace (stat_bytes
// more).
int plan_row_stats(mdsm_plan *p, uint32_t n)
{
    if (p->subband || n < 1) return MDSM_ERR_INVALID;
    const uint32_t C = p->obs.nchans;
    const double vmax = 255.0 * (double)C;
    const int64_t base = round_up((int64_t)p->work_bytes, 256);
    int64_t so = base;
    for (size_t r = 0; r < p->ranges.size(); ++r) {
        RangeHost &rh = p->ranges[r];
        rh.stat_tiles = rh.stat_n = 0;
        rh.stat_off = 0;
        if (p->obs.dtype != MDSM_U8 || rh.decim >= 0 || n > rh.info.nout ||
            vmax >= 16777216.0 || (double)n * vmax * vmax >= 9007199254740992.0)
            continue;
        rh.stat_n = (int32_t)n;
        rh.stat_tiles = (int32_t)((n + kTile - 1) / kTile);
        rh.stat_off = so;
        so = round_up(so + (int64_t)rh.info.ndms * rh.stat_tiles * 16, 256);
    }
    p->stat_bytes = (uint64_t)(so - base);
    // a brute-force plan has one segment per range, in range order
    for (size_t r = 0; r < p->ranges.size() && r < p->segs.size(); ++r) {
        const RangeHost &rh = p->ranges[r];
        const SegDesc &sg = p->segs[r];
        for (int bi = sg.first_block; bi < sg.first_block + sg.nblocks; ++bi) {
            BlockDesc &b = p->blocks[bi];
            const int64_t j0 = (b.out_off - (int64_t)rh.info.out_offset) / (int64_t)rh.info.nout;
            b.stat_tiles = rh.stat_tiles;
            b.stat_n = rh.stat_n;
            b.stat_off = rh.stat_off + j0 * rh.stat_tiles * 16;
        }
    }
    if (p->uploaded) {
        hipError_t e = hipMemcpy(p->dev.blocks, p->blocks.data(),
                                 p->blocks.size() * sizeof(BlockDesc), hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e, "uploading plan blocks");
    }
    return MDSM_OK;
}

mdsm_tuning default_tuning()
{
    mdsm_tuning t{};
    t.chunk_rows = 0;   // kChunkRows
    t.persistent = 1;   // persistent workgroups over the work queue
    t.tile_group = 64;  // 64 tiles of a DM block per item group
    return t;
}

int tune_search(const mdsm_tuning &start, const std::function<int(const mdsm_tuning &)> &apply,
                const std::function<int()> &step, hipStream_t s, uint32_t reps,
                mdsm_tuning *best_out)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return hip_fail(e, "hipEventCreate");
    }
    auto time_one = [&](const mdsm_tuning &t, float *ms) -> int {
        int rc = apply(t);
        if (!rc) rc = step();  // warm-up
        if (rc) return rc;
        hipError_t r = hipEventRecord(e0, s);
        if (r != hipSuccess) return hip_fail(r, "hipEventRecord");
        for (uint32_t i = 0; i < reps && !rc; ++i) rc = step();
        if (rc) return rc;
        if ((r = hipEventRecord(e1, s)) != hipSuccess || (r = hipEventSynchronize(e1)) != hipSuccess ||
            (r = hipEventElapsedTime(ms, e0, e1)) != hipSuccess)
            return hip_fail(r, "timing a candidate");
        return MDSM_OK;
    };
    // coordinate search: channels per chunk first, then work order
    mdsm_tuning best = start;
    float best_ms = std::numeric_limits<float>::max();
    int rc = MDSM_OK;
    for (uint32_t r : {2u, 3u, 4u, 6u, 8u}) {
        mdsm_tuning t = start;
        t.chunk_rows = r;
        float ms = 0.0f;
        if ((rc = time_one(t, &ms))) break;
        if (ms < best_ms) {
            best_ms = ms;
            best = t;
        }
    }
    const uint32_t orders[4][2] = {{0, 1}, {1, 16}, {1, 64}, {1, 128}};
    const mdsm_tuning base = best;
    for (int i = 0; i < 4 && !rc; ++i) {
        mdsm_tuning t = base;
        t.persistent = orders[i][0];
        t.tile_group = orders[i][1];
        float ms = 0.0f;
        if ((rc = time_one(t, &ms))) break;
        if (ms < best_ms) {
            best_ms = ms;
            best = t;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) {
        (void)apply(start);
        return rc;
    }
    if ((rc = apply(best))) return rc;
    if (best_out) *best_out = best;
    return MDSM_OK;
}

}  // namespace mdsm

using namespace mdsm;

extern "C" int mdsm_plan_create(const mdsm_obs *obs, const mdsm_range *ranges, uint32_t nranges,
                                mdsm_plan **out)
{
    if (!out) return fail(MDSM_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int rc = validate_obs(obs);
    if (rc) return rc;
    if (!ranges || nranges < 1 || nranges > 64)
        return fail(MDSM_ERR_INVALID, "need 1..64 DM ranges");
    auto *p = new mdsm_plan();
    p->tune = default_tuning();
    p->obs = *obs;
    ShiftModel model(*obs);
    const uint32_t C = obs->nchans, S = obs->nsamp_in;
    uint64_t work = 0, out_elems = 0;
    uint32_t dm_row = 0, maxshift = 0;
    double adds = 0, bytes = (double)C * S * dtype_size(obs->dtype), dcs = 0;
    for (uint32_t r = 0; r < nranges; ++r) {
        const mdsm_range &R = ranges[r];
        if (R.ndms < 1 || R.downsample < 1 || R.downsample > 256 || !(R.lowdm >= 0.0f) ||
            !(R.dmstep >= 0.0f) || !std::isfinite(R.lowdm) || !std::isfinite(R.dmstep)) {
            delete p;
            return fail(MDSM_ERR_INVALID, "range " + std::to_string(r) + " is malformed");
        }
        RangeHost rh;
        const int k = (int)R.downsample;
        const uint32_t Sk = S / (uint32_t)k;
        const float tsf = obs->tsamp * (float)k;
        const double tsd = (double)obs->tsamp * (double)k;
        rh.shift.resize((size_t)R.ndms * C);
        rh.dm_values.resize(R.ndms);
        int32_t M = 0;
        for (uint32_t n = 0; n < R.ndms; ++n) {
            float dmf = (float)n * R.dmstep;
            dmf = R.lowdm + dmf;
            double dmd = (double)R.lowdm + (double)n * (double)R.dmstep;
            rh.dm_values[n] = dmf;
            for (uint32_t c = 0; c < C; ++c) {
                int32_t v = model.shift(c, dmf, dmd, tsf, tsd);
                rh.shift[(size_t)n * C + c] = v;
                M = std::max(M, v);
            }
        }
        if ((int64_t)Sk - M < 1) {
            delete p;
            return fail(MDSM_ERR_INVALID, "range " + std::to_string(r) +
                                              ": block too short for its maximum delay (" +
                                              std::to_string(M) + " samples at tsamp*k)");
        }
        mdsm_range_info &ri = rh.info;
        ri.lowdm = R.lowdm;
        ri.dmstep = R.dmstep;
        ri.ndms = R.ndms;
        ri.downsample = R.downsample;
        ri.maxshift = (uint32_t)M;
        ri.nout = Sk - (uint32_t)M;
        ri.out_offset = out_elems;
        ri.first_dm = dm_row;
        ri.ncalls = R.ndms;
        out_elems += (uint64_t)R.ndms * ri.nout;
        dm_row += R.ndms;
        maxshift = std::max<uint32_t>(maxshift, (uint32_t)k * (uint32_t)M);
        if (k > 1) rh.decim = add_decimation(p, k, work);
        adds += (double)R.ndms * C * ri.nout;
        dcs += (double)R.ndms * C * ri.nout;
        bytes += (double)R.ndms * ri.nout * 4.0;
        p->ranges.push_back(std::move(rh));
    }
    for (auto &d : p->decim) {
        adds += (double)C * d.nsamp_out * (d.k - 1);
        bytes += (double)C * d.nsamp_out * dtype_size(d.dtype_out) * 2.0;  // written, read back
    }
    p->work_bytes = work;
    // One launch covers every range.
    Stage st;
    st.seg_begin = 0;
    for (uint32_t r = 0; r < nranges; ++r) {
        RangeHost &rh = p->ranges[r];
        const int32_t *sh = rh.shift.data();
        SegSpec s;
        s.ntrials = (int)rh.info.ndms;
        s.nrows = (int)C;
        s.sh = [sh, C](int j, int i) { return sh[(size_t)j * C + i]; };
        if (rh.decim < 0) {
            s.src = BUF_IN;
            s.row_stride = 0;  // caller's in_stride
            s.in_off = 0;
            s.in_len = (int)S;
            s.dtype = obs->dtype;
            s.maxval = 255;
        } else {
            const DecimStep &d = p->decim[rh.decim];
            s.src = BUF_WORK;
            s.row_stride = d.stride;
            s.in_off = d.work_off / (int64_t)dtype_size(d.dtype_out);
            s.in_len = d.nsamp_out;
            s.dtype = d.dtype_out;
            s.maxval = 255 * d.k;
        }
        s.row0 = 0;
        s.dst = BUF_OUT;
        s.out_off = (int64_t)rh.info.out_offset;
        s.out_trial_stride = rh.info.nout;
        s.nout = (int)rh.info.nout;
        s.k = (int)rh.info.downsample;
        int rc2 = add_segment(p, st, s);
        if (rc2) {
            delete p;
            return rc2;
        }
    }
    p->stages.push_back(st);
    {
        std::string part = "B";
        for (uint32_t r = 0; r < nranges; ++r) {
            char buf[128];
            std::snprintf(buf, sizeof buf, "|%.9g,%.9g,%u,%u", (double)ranges[r].lowdm,
                          (double)ranges[r].dmstep, ranges[r].ndms, ranges[r].downsample);
            part += buf;
        }
        p->shape_key = shape_key_of(*obs, part);
    }
    adopt_cached_tuning(p);
    apply_tuning(p);
    p->info.ndms_total = dm_row;
    p->info.maxshift = maxshift;
    p->info.out_elems = out_elems;
    bytes += (double)p->rel.size() * 4.0;
    finish_info(p, adds, bytes, dcs);
    *out = p;
    return MDSM_OK;
}

extern "C" int mdsm_plan_create_subband(const mdsm_obs *obs, uint32_t nsubs, const mdsm_pass *passes,
                                        uint32_t npasses, mdsm_plan **out)
{
    if (!out) return fail(MDSM_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int rc = validate_obs(obs);
    if (rc) return rc;
    const uint32_t C = obs->nchans, S = obs->nsamp_in;
    if (nsubs < 1 || C % nsubs != 0)
        return fail(MDSM_ERR_INVALID, "nsubs must divide nchans");
    if (!passes || npasses < 1 || npasses > 64)
        return fail(MDSM_ERR_INVALID, "need 1..64 subband passes");
    const uint32_t cps = C / nsubs;
    auto *p = new mdsm_plan();
    p->tune = default_tuning();
    p->obs = *obs;
    p->subband = true;
    p->nsubs = nsubs;
    ShiftModel model(*obs);
    uint64_t work = 0, out_elems = 0;
    uint32_t dm_row = 0, maxshift = 0;
    double adds = 0, bytes = (double)C * S * dtype_size(obs->dtype), dcs = 0;
    for (uint32_t q = 0; q < npasses; ++q) {
        const mdsm_pass &P = passes[q];
        if (P.binsize < 1 || P.binsize > 256 || P.ncalls < 1 || P.calldms < 1 ||
            P.ndms != P.ncalls * P.calldms || !(P.lowdm >= 0.0f) || !(P.dmstep >= 0.0f) ||
            !(P.sub_dmstep >= 0.0f)) {
            delete p;
            return fail(MDSM_ERR_INVALID, "pass " + std::to_string(q) +
                                              " is malformed (ndms must equal ncalls*calldms)");
        }
        RangeHost rh;
        rh.pass = P;
        const int k = (int)P.binsize;
        const uint32_t Sk = S / (uint32_t)k;
        const float tsf = obs->tsamp * (float)k;
        const double tsd = (double)obs->tsamp * (double)k;
        rh.shift.resize((size_t)P.ncalls * C);
        rh.fine.resize((size_t)P.ndms * nsubs);
        rh.dm_values.resize(P.ndms);
        int32_t M1 = 0, M2 = 0;
        for (uint32_t j = 0; j < P.ncalls; ++j) {
            // Coarse DM of call j (src/mdsm/dedispersion_thread.cu:161-168).
            float cdm = (float)j * P.sub_dmstep;
            cdm = P.lowdm + cdm;
            double cdmd = (double)P.lowdm + (double)j * (double)P.sub_dmstep;
            int32_t *co = &rh.shift[(size_t)j * C];
            for (uint32_t c = 0; c < C; ++c) co[c] = model.shift(c, cdm, cdmd, tsf, tsd);
            for (uint32_t c = 0; c < C; ++c) {
                int32_t rel = co[c] - co[(c / cps) * cps];
                if (rel < 0) {
                    delete p;
                    return fail(MDSM_ERR_INVALID, "negative subband-relative shift");
                }
                M1 = std::max(M1, rel);
            }
            // Fine DMs of call j centred on the coarse DM
            // (src/mdsm/dedispersion_thread.cu:195-205).
            float half = (float)P.calldms * P.dmstep;
            half = half / 2.0f;
            float start = cdm - half;
            if (!(start > 0.0f)) start = 0.0f;
            double startd = std::max(cdmd - (double)P.calldms * (double)P.dmstep / 2.0, 0.0);
            for (uint32_t i = 0; i < P.calldms; ++i) {
                float fdm = (float)i * P.dmstep;
                fdm = start + fdm;
                double fdmd = startd + (double)i * (double)P.dmstep;
                size_t d = (size_t)j * P.calldms + i;
                rh.dm_values[d] = fdm;
                for (uint32_t b = 0; b < nsubs; ++b) {
                    int32_t v = model.shift(b * cps, fdm, fdmd, tsf, tsd);
                    rh.fine[d * nsubs + b] = v;
                    M2 = std::max(M2, v);
                }
            }
        }
        int64_t N = (int64_t)Sk - M1 - M2;
        if (N < 1) {
            delete p;
            return fail(MDSM_ERR_INVALID, "pass " + std::to_string(q) +
                                              ": block too short for its maximum delay");
        }
        rh.len1 = (int32_t)(N + M2);
        mdsm_range_info &ri = rh.info;
        ri.lowdm = P.lowdm;
        ri.dmstep = P.dmstep;
        ri.ndms = P.ndms;
        ri.downsample = P.binsize;
        ri.maxshift = (uint32_t)(M1 + M2);
        ri.nout = (uint32_t)N;
        ri.out_offset = out_elems;
        ri.first_dm = dm_row;
        ri.ncalls = P.ncalls;
        out_elems += (uint64_t)P.ndms * ri.nout;
        dm_row += P.ndms;
        maxshift = std::max<uint32_t>(maxshift, (uint32_t)k * (uint32_t)(M1 + M2));
        if (k > 1) rh.decim = add_decimation(p, k, work);
        work = (uint64_t)round_up((int64_t)work, 256);
        rh.sb_off = (int64_t)work;
        work += (uint64_t)P.ncalls * nsubs * (uint64_t)rh.len1 * 4u;
        adds += (double)P.ncalls * C * rh.len1 + (double)P.ndms * nsubs * N;
        dcs += (double)P.ndms * C * N;
        bytes += (double)P.ncalls * nsubs * rh.len1 * 4.0 * 2.0 + (double)P.ndms * N * 4.0;
        p->ranges.push_back(std::move(rh));
    }
    for (auto &d : p->decim) {
        adds += (double)C * d.nsamp_out * (d.k - 1);
        bytes += (double)C * d.nsamp_out * dtype_size(d.dtype_out) * 2.0;
    }
    p->work_bytes = work;
    // Stage 1: per pass and subband, coarse DMs over the subband's channels.
    Stage s1;
    s1.seg_begin = 0;
    for (auto &rh : p->ranges) {
        const int32_t *co = rh.shift.data();
        for (uint32_t b = 0; b < nsubs; ++b) {
            SegSpec s;
            s.ntrials = (int)rh.pass.ncalls;
            s.nrows = (int)cps;
            const uint32_t top = b * cps;
            s.sh = [co, C, top](int j, int i) {
                const int32_t *row = co + (size_t)j * C;
                return row[top + i] - row[top];
            };
            if (rh.decim < 0) {
                s.src = BUF_IN;
                s.row_stride = 0;
                s.in_off = 0;
                s.in_len = (int)S;
                s.dtype = obs->dtype;
                s.maxval = 255;
            } else {
                const DecimStep &d = p->decim[rh.decim];
                s.src = BUF_WORK;
                s.row_stride = d.stride;
                s.in_off = d.work_off / (int64_t)dtype_size(d.dtype_out);
                s.in_len = d.nsamp_out;
                s.dtype = d.dtype_out;
                s.maxval = 255 * d.k;
            }
            s.row0 = top;
            s.dst = BUF_WORK;
            s.out_off = rh.sb_off / 4 + (int64_t)b * rh.len1;
            s.out_trial_stride = (int64_t)nsubs * rh.len1;
            s.nout = rh.len1;
            s.k = 1;
            int rc2 = add_segment(p, s1, s);
            if (rc2) {
                delete p;
                return rc2;
            }
        }
    }
    // Stage 2: per pass and call, fine DMs over the call's subband series.
    Stage s2;
    s2.seg_begin = (int32_t)p->segs.size();
    for (auto &rh : p->ranges) {
        const int32_t *fi = rh.fine.data();
        for (uint32_t j = 0; j < rh.pass.ncalls; ++j) {
            SegSpec s;
            s.ntrials = (int)rh.pass.calldms;
            s.nrows = (int)nsubs;
            const size_t d0 = (size_t)j * rh.pass.calldms;
            const uint32_t ns = nsubs;
            s.sh = [fi, d0, ns](int t, int b) { return fi[(d0 + t) * ns + b]; };
            s.src = BUF_WORK;
            s.row0 = 0;
            s.row_stride = rh.len1;
            s.in_off = rh.sb_off / 4 + (int64_t)j * nsubs * rh.len1;
            s.in_len = rh.len1;
            s.dtype = DT_F32;
            s.maxval = 0;
            s.dst = BUF_OUT;
            s.out_off = (int64_t)rh.info.out_offset + (int64_t)d0 * rh.info.nout;
            s.out_trial_stride = rh.info.nout;
            s.nout = (int)rh.info.nout;
            s.k = (int)rh.pass.binsize;
            int rc2 = add_segment(p, s2, s);
            if (rc2) {
                delete p;
                return rc2;
            }
        }
    }
    p->stages.push_back(s1);
    p->stages.push_back(s2);
    {
        std::string part = "S" + std::to_string(nsubs);
        for (uint32_t q = 0; q < npasses; ++q) {
            const mdsm_pass &P = passes[q];
            char buf[192];
            std::snprintf(buf, sizeof buf, "|%.9g,%.9g,%.9g,%.9g,%u,%u,%u,%u", (double)P.lowdm,
                          (double)P.highdm, (double)P.dmstep, (double)P.sub_dmstep, P.binsize,
                          P.ndms, P.ncalls, P.calldms);
            part += buf;
        }
        p->shape_key = shape_key_of(*obs, part);
    }
    adopt_cached_tuning(p);
    apply_tuning(p);
    p->info.ndms_total = dm_row;
    p->info.maxshift = maxshift;
    p->info.out_elems = out_elems;
    bytes += (double)p->rel.size() * 4.0;
    finish_info(p, adds, bytes, dcs);
    *out = p;
    return MDSM_OK;
}

extern "C" void mdsm_plan_destroy(mdsm_plan *p)
{
    if (!p) return;
    if (!p->uploaded) {
        delete p;
        return;
    }
    int cur = 0;
    bool sw = hipGetDevice(&cur) == hipSuccess && cur != p->device;
    if (sw) (void)hipSetDevice(p->device);
    if (p->dev.blocks) (void)hipFree(p->dev.blocks);
    if (p->dev.segs) (void)hipFree(p->dev.segs);
    if (p->dev.rel) (void)hipFree(p->dev.rel);
    if (p->dev.abase) (void)hipFree(p->dev.abase);
    if (p->dev.counter) (void)hipFree(p->dev.counter);
    if (p->dev.plain_shift) (void)hipFree(p->dev.plain_shift);
    if (p->own_work) (void)hipFree(p->own_work);
    if (sw) (void)hipSetDevice(cur);
    delete p;
}

extern "C" int mdsm_plan_get_info(const mdsm_plan *p, mdsm_plan_info *info)
{
    if (!p || !info) return fail(MDSM_ERR_INVALID, "NULL argument");
    *info = p->info;
    return MDSM_OK;
}

extern "C" int mdsm_plan_get_range(const mdsm_plan *p, uint32_t r, mdsm_range_info *info)
{
    if (!p || !info) return fail(MDSM_ERR_INVALID, "NULL argument");
    if (r >= p->ranges.size()) return fail(MDSM_ERR_INVALID, "range index out of bounds");
    *info = p->ranges[r].info;
    return MDSM_OK;
}

extern "C" int mdsm_plan_shift_table(const mdsm_plan *p, uint32_t r, int32_t *host_out)
{
    if (!p || !host_out) return fail(MDSM_ERR_INVALID, "NULL argument");
    if (r >= p->ranges.size()) return fail(MDSM_ERR_INVALID, "range index out of bounds");
    const auto &v = p->ranges[r].shift;
    std::memcpy(host_out, v.data(), v.size() * sizeof(int32_t));
    return MDSM_OK;
}

extern "C" int mdsm_plan_fine_table(const mdsm_plan *p, uint32_t r, int32_t *host_out)
{
    if (!p || !host_out) return fail(MDSM_ERR_INVALID, "NULL argument");
    if (!p->subband) return fail(MDSM_ERR_INVALID, "not a subband plan");
    if (r >= p->ranges.size()) return fail(MDSM_ERR_INVALID, "pass index out of bounds");
    const auto &v = p->ranges[r].fine;
    std::memcpy(host_out, v.data(), v.size() * sizeof(int32_t));
    return MDSM_OK;
}

extern "C" int mdsm_plan_dm_values(const mdsm_plan *p, float *host_out)
{
    if (!p || !host_out) return fail(MDSM_ERR_INVALID, "NULL argument");
    size_t o = 0;
    for (const auto &rh : p->ranges) {
        std::memcpy(host_out + o, rh.dm_values.data(), rh.dm_values.size() * sizeof(float));
        o += rh.dm_values.size();
    }
    return MDSM_OK;
}

extern "C" int mdsm_dedisperse(mdsm_plan *p, const void *d_in, int64_t in_stride, float *d_out,
                               void *d_work, void *stream)
{
    if (!p || !d_in || !d_out) return fail(MDSM_ERR_INVALID, "NULL argument");
    if (in_stride < (int64_t)p->obs.nsamp_in)
        return fail(MDSM_ERR_INVALID, "in_stride < nsamp_in");
    int cur = 0;
    hipError_t e;
    if (!p->uploaded) {
        int rc = upload(p);
        if (rc) return rc;
    }
    e = hipGetDevice(&cur);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (cur != p->device) return fail(MDSM_ERR_STATE, "plan belongs to another device");
    hipStream_t s = (hipStream_t)stream;
    void *work = d_work;
    if (!work && p->work_bytes) {
        if (!p->own_work) {
            e = hipMalloc(&p->own_work, p->work_bytes);
            if (e != hipSuccess) {
                p->own_work = nullptr;
                return fail(MDSM_ERR_NOMEM, "workspace allocation of " +
                                                std::to_string(p->work_bytes) + " bytes failed");
            }
        }
        work = p->own_work;
    }
    if (!p->decim.empty()) {
        TimedLaunch tl("decimate", s);
        e = launch_decimate_steps(d_in, p->obs.nchans, in_stride, p->obs.nsamp_in, p->decim.data(),
                                  p->decim.size(), (char *)work, s);
        if (e != hipSuccess) return hip_fail(e, "decimate launch");
    }
    for (const Stage &st : p->stages) {
        TimedLaunch tl("dedisperse", s);
        e = launch_dedisperse(p, st, d_in, in_stride, d_out, work, s);
        if (e != hipSuccess) return hip_fail(e, "dedisperse launch");
    }
    return MDSM_OK;
}

extern "C" int mdsm_plan_upload(mdsm_plan *p)
{
    if (!p) return fail(MDSM_ERR_INVALID, "NULL argument");
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (cur != p->device) return fail(MDSM_ERR_STATE, "plan belongs to another device");
    return upload_plan(p);
}

extern "C" int mdsm_plan_occupancy(mdsm_plan *p, uint32_t stage, mdsm_occupancy *out)
{
    if (!p || !out) return fail(MDSM_ERR_INVALID, "NULL argument");
    if (stage >= p->stages.size()) return fail(MDSM_ERR_INVALID, "no such stage");
    const Stage &st = p->stages[stage];
    DedispOccupancy o{};
    hipError_t e = dedisperse_occupancy(st.lds_bytes, st.dtypes, &o);
    if (e != hipSuccess) return hip_fail(e, "occupancy query");
    out->workgroups_per_cu = o.workgroups_per_cu;
    out->waves_per_simd = o.waves_per_simd;
    out->vgprs = o.vgprs;
    out->static_lds = o.static_lds;
    out->dynamic_lds = o.dynamic_lds;
    out->scratch_bytes = o.scratch_bytes;
    return MDSM_OK;
}

extern "C" int mdsm_plan_lds_bytes(const mdsm_plan *p, uint32_t stage, uint64_t *reads,
                                   uint64_t *writes)
{
    if (!p || !reads || !writes) return fail(MDSM_ERR_INVALID, "NULL argument");
    if (stage >= p->stages.size()) return fail(MDSM_ERR_INVALID, "no such stage");
    const Stage &st = p->stages[stage];
    uint64_t r = 0, w = 0;
    for (int si = st.seg_begin; si < st.seg_end; ++si) {
        const SegDesc &sg = p->segs[si];
        for (int bi = sg.first_block; bi < sg.first_block + sg.nblocks; ++bi) {
            const BlockDesc &b = p->blocks[bi];
            const uint64_t rows = (uint64_t)b.nrows * (uint64_t)sg.ntiles;
            // every (row, trial slot) of a tile: 8 bytes of LDS per lane for
            // 16-bit entries, 16 for float32 (padding slots read too)
            const uint64_t per_sample = b.dtype == DT_F32 ? 4 : 2;
            r += rows * (uint64_t)(kWaves * b.rpw) * 256u * per_sample;
            for (int i = 0; i < b.nrows; ++i)
                w += (uint64_t)sg.ntiles *
                     (uint64_t)lds_row_bytes(b.dtype, 4 * (p->abase[(size_t)b.base_off + 2 * i + 1] & 1023));
        }
    }
    *reads = r;
    *writes = w;
    return MDSM_OK;
}

extern "C" int mdsm_plan_set_tuning(mdsm_plan *p, const mdsm_tuning *t)
{
    if (!p || !t) return fail(MDSM_ERR_INVALID, "NULL argument");
    if (t->chunk_rows == 1 || t->chunk_rows > 8)
        return fail(MDSM_ERR_INVALID, "chunk_rows must be 0 or 2..8");
    if (t->lds_budget > (uint32_t)kLdsMax)
        return fail(MDSM_ERR_INVALID, "lds_budget must be <= 162816");
    if (t->lds_budget && t->lds_budget < 4096)
        return fail(MDSM_ERR_INVALID, "lds_budget must be 0 or >= 4096");
    if (t->persistent > 1) return fail(MDSM_ERR_INVALID, "persistent must be 0 or 1");
    if (t->grid > (1u << 20)) return fail(MDSM_ERR_INVALID, "grid must be <= 2^20");
    if (t->tile_group > 4096) return fail(MDSM_ERR_INVALID, "tile_group must be <= 4096");
    p->tune = *t;
    apply_tuning(p);
    if (p->uploaded) {
        // only chunk sizes changed: refresh the block, offset and base tables in place
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
        if (cur != p->device) return fail(MDSM_ERR_STATE, "plan belongs to another device");
        e = hipMemcpy(p->dev.blocks, p->blocks.data(), p->blocks.size() * sizeof(BlockDesc),
                      hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(p->dev.rel, p->rel.data(), p->rel.size() * sizeof(int32_t),
                          hipMemcpyHostToDevice);
        // the rows' places in their chunk buffers (row info) move with the chunk size
        if (e == hipSuccess)
            e = hipMemcpy(p->dev.abase, p->abase.data(), p->abase.size() * sizeof(int32_t),
                          hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e, "updating the block and offset tables");
    }
    return MDSM_OK;
}

extern "C" int mdsm_plan_get_tuning(const mdsm_plan *p, mdsm_tuning *t)
{
    if (!p || !t) return fail(MDSM_ERR_INVALID, "NULL argument");
    *t = p->tune;
    return MDSM_OK;
}

extern "C" int mdsm_plan_autotune(mdsm_plan *p, const void *d_in, int64_t in_stride, float *d_out,
                                  void *d_work, void *stream, uint32_t reps, mdsm_tuning *best)
{
    if (!p || !d_in || !d_out) return fail(MDSM_ERR_INVALID, "NULL argument");
    if (reps < 1 || reps > 1000) return fail(MDSM_ERR_INVALID, "reps must be 1..1000");
    mdsm_tuning found{};
    int rc = tune_search(
        p->tune, [&](const mdsm_tuning &t) { return mdsm_plan_set_tuning(p, &t); },
        [&]() { return mdsm_dedisperse(p, d_in, in_stride, d_out, d_work, stream); },
        (hipStream_t)stream, reps, &found);
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        g_tune_cache[p->shape_key] = found;
    }
    if (best) *best = found;
    return MDSM_OK;
}

extern "C" int mdsm_dedisperse_plain(mdsm_plan *p, const float *d_in, int64_t in_stride,
                                     float *d_out, void *stream)
{
    if (!p || !d_in || !d_out) return fail(MDSM_ERR_INVALID, "NULL argument");
    if (p->subband || p->obs.dtype != MDSM_F32)
        return fail(MDSM_ERR_INVALID, "the plain kernel runs float32 brute-force plans only");
    for (const RangeHost &rh : p->ranges)
        if (rh.info.downsample != 1)
            return fail(MDSM_ERR_INVALID, "the plain kernel runs undecimated ranges only");
    if (in_stride < (int64_t)p->obs.nsamp_in)
        return fail(MDSM_ERR_INVALID, "in_stride < nsamp_in");
    if (p->obs.nchans > 16384) return fail(MDSM_ERR_INVALID, "plain kernel: nchans > 16384");
    int rc = upload_plan(p);
    if (rc) return rc;
    hipError_t e;
    if (!p->dev.plain_shift) {
        size_t n = 0;
        for (const RangeHost &rh : p->ranges) n += rh.shift.size();
        e = hipMalloc((void **)&p->dev.plain_shift, n * sizeof(int32_t));
        if (e != hipSuccess) {
            p->dev.plain_shift = nullptr;
            return hip_fail(e, "plain shift table");
        }
        size_t o = 0;
        for (const RangeHost &rh : p->ranges) {
            e = hipMemcpy(p->dev.plain_shift + o, rh.shift.data(), rh.shift.size() * sizeof(int32_t),
                          hipMemcpyHostToDevice);
            if (e != hipSuccess) return hip_fail(e, "plain shift table");
            o += rh.shift.size();
        }
    }
    hipStream_t s = (hipStream_t)stream;
    size_t o = 0;
    for (const RangeHost &rh : p->ranges) {
        TimedLaunch tl("plain", s);
        e = launch_plain(d_in, in_stride, p->obs.nchans, p->dev.plain_shift + o, rh.info.ndms,
                         rh.info.nout, d_out + rh.info.out_offset, s);
        if (e != hipSuccess) return hip_fail(e, "plain launch");
        o += rh.shift.size();
    }
    return MDSM_OK;
}
