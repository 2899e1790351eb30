// C-ABI entry points that are not plan methods: errors, timing, decimation,
// clipper, detectors and the seeded synthetic filterbank generator.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "mdsm_internal.h"

namespace mdsm {

namespace {
thread_local std::string g_last_error;

struct TimingState {
    std::mutex mu;
    bool on = false;
    struct Pending {
        std::string name;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::map<std::string, std::pair<double, uint64_t>> done;

    void drain_locked()
    {
        for (auto &p : pending) {
            float ms = 0.0f;
            if (hipEventSynchronize(p.b) == hipSuccess &&
                hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
                auto &d = done[p.name];
                d.first += ms;
                d.second += 1;
            }
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        pending.clear();
    }
};

TimingState &timing()
{
    static TimingState t;
    return t;
}
}  // namespace

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? MDSM_ERR_NOMEM : MDSM_ERR_DEVICE;
}

TimedLaunch::TimedLaunch(const char *n, hipStream_t s) : name(n), stream(s)
{
    TimingState &t = timing();
    if (!t.on) return;
    if (hipEventCreate(&ev0) != hipSuccess || hipEventCreate(&ev1) != hipSuccess) {
        ev0 = ev1 = nullptr;
        return;
    }
    (void)hipEventRecord(ev0, stream);
}

TimedLaunch::~TimedLaunch()
{
    if (!ev0) return;
    (void)hipEventRecord(ev1, stream);
    TimingState &t = timing();
    std::lock_guard<std::mutex> g(t.mu);
    t.pending.push_back({name, ev0, ev1});
    if (t.pending.size() > 4096) t.drain_locked();
}

bool timing_swap(bool on)
{
    TimingState &t = timing();
    std::lock_guard<std::mutex> g(t.mu);
    const bool was = t.on;
    t.on = on;
    return was;
}

}  // namespace mdsm

using namespace mdsm;

extern "C" const char *mdsm_last_error(void) { return g_last_error.c_str(); }

extern "C" const char *mdsm_version(void) { return "mdsm-mi355x 0.1 (gfx950)"; }

extern "C" int mdsm_timing_enable(int on)
{
    TimingState &t = timing();
    std::lock_guard<std::mutex> g(t.mu);
    t.on = on != 0;
    return MDSM_OK;
}

extern "C" int mdsm_timing_reset(void)
{
    TimingState &t = timing();
    std::lock_guard<std::mutex> g(t.mu);
    t.drain_locked();
    t.done.clear();
    return MDSM_OK;
}

extern "C" int mdsm_timing_get(const char *kernel, double *total_ms, uint64_t *launches)
{
    if (!kernel || !total_ms || !launches) return fail(MDSM_ERR_INVALID, "NULL argument");
    TimingState &t = timing();
    std::lock_guard<std::mutex> g(t.mu);
    t.drain_locked();
    auto it = t.done.find(kernel);
    *total_ms = it == t.done.end() ? 0.0 : it->second.first;
    *launches = it == t.done.end() ? 0 : it->second.second;
    return MDSM_OK;
}

extern "C" int mdsm_decimate(const void *d_in, int32_t dtype, uint32_t nchans, int64_t in_stride,
                             uint32_t nsamp, uint32_t k, void *d_out, int64_t out_stride,
                             void *stream)
{
    if (!d_in || !d_out) return fail(MDSM_ERR_INVALID, "NULL argument");
    if (dtype < MDSM_U8 || dtype > MDSM_F32) return fail(MDSM_ERR_INVALID, "bad dtype");
    if (k < 1 || k > 256) return fail(MDSM_ERR_INVALID, "k must be in 1..256");
    if (in_stride < (int64_t)nsamp || out_stride < (int64_t)(nsamp / k))
        return fail(MDSM_ERR_INVALID, "stride smaller than row length");
    TimedLaunch tl("decimate", (hipStream_t)stream);
    hipError_t e = launch_decimate(d_in, dtype, nchans, in_stride, nsamp, k, d_out,
                                   dtype == MDSM_U8 ? DT_U16 : DT_F32, out_stride,
                                   (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "decimate launch");
    return MDSM_OK;
}

extern "C" int mdsm_clip_u8(const uint8_t *d_in, int64_t in_stride, uint32_t nchans, int64_t t0,
                            uint32_t n, float nsigma, uint8_t *d_out, int64_t out_stride,
                            int32_t *d_med, int32_t *d_mad, int32_t *d_nclipped, void *stream)
{
    if (!d_in || !d_out) return fail(MDSM_ERR_INVALID, "NULL argument");
    if (n < 1 || t0 < 0 || t0 + (int64_t)n > in_stride || t0 + (int64_t)n > out_stride)
        return fail(MDSM_ERR_INVALID, "clip window outside the rows");
    if (!(nsigma > 0.0f)) return fail(MDSM_ERR_INVALID, "nsigma must be > 0");
    TimedLaunch tl("clip", (hipStream_t)stream);
    hipError_t e = launch_clip_u8(d_in, in_stride, nchans, t0, n, nsigma, d_out, out_stride, d_med,
                                  d_mad, d_nclipped, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "clip launch");
    return MDSM_OK;
}

extern "C" int mdsm_detect_boxcar(const float *d_x, uint32_t ndms, uint32_t n, int64_t stride,
                                  const uint32_t *widths, uint32_t nwidths, double thr,
                                  mdsm_cand *d_cands, uint32_t cap, uint32_t *d_count,
                                  double *d_mean, double *d_sigma, void *stream)
{
    if (!d_x || !widths || !d_cands || !d_count || !d_mean || !d_sigma)
        return fail(MDSM_ERR_INVALID, "NULL argument");
    if (stride < (int64_t)n || n < 1) return fail(MDSM_ERR_INVALID, "bad row length/stride");
    if (ndms > 65535) return fail(MDSM_ERR_INVALID, "ndms > 65535");
    uint32_t mask = 0;
    for (uint32_t i = 0; i < nwidths; ++i) {
        uint32_t w = widths[i];
        if (w == 0 || (w & (w - 1)) || w > 1024)
            return fail(MDSM_ERR_INVALID, "boxcar widths must be powers of two <= 1024");
        mask |= w;
    }
    if (!mask) return fail(MDSM_ERR_INVALID, "no widths");
    SqrtW sw;
    for (int i = 0; i < 11; ++i) sw.v[i] = std::sqrt((double)(1u << i));
    TimedLaunch tl("detect", (hipStream_t)stream);
    hipError_t e = launch_boxcar(d_x, ndms, n, stride, mask, thr, sw, d_cands, cap, d_count, d_mean,
                                 d_sigma, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "detect launch");
    return MDSM_OK;
}

extern "C" uint32_t mdsm_boxcar_carry_len(uint32_t widths_mask)
{
    return boxcar_carry_len(widths_mask);
}

extern "C" int mdsm_detect_boxcar_stream(const float *d_x, uint32_t ndms, uint32_t n,
                                         int64_t stride, float *d_carry, int64_t carry_stride,
                                         int32_t carry_valid, const uint32_t *widths,
                                         uint32_t nwidths, double thr, mdsm_cand *d_cands,
                                         uint32_t cap, uint32_t *d_count, double *d_mean,
                                         double *d_sigma, void *stream)
{
    if (!d_x || !d_carry || !widths || !d_cands || !d_count || !d_mean || !d_sigma)
        return fail(MDSM_ERR_INVALID, "NULL argument");
    if (stride < (int64_t)n || n < 1) return fail(MDSM_ERR_INVALID, "bad row length/stride");
    if (ndms > 65535) return fail(MDSM_ERR_INVALID, "ndms > 65535");
    uint32_t mask = 0;
    for (uint32_t i = 0; i < nwidths; ++i) {
        uint32_t w = widths[i];
        if (w == 0 || (w & (w - 1)) || w > 1024)
            return fail(MDSM_ERR_INVALID, "boxcar widths must be powers of two <= 1024");
        mask |= w;
    }
    if (!mask) return fail(MDSM_ERR_INVALID, "no widths");
    const uint32_t H = boxcar_carry_len(mask);
    if (carry_stride < (int64_t)H) return fail(MDSM_ERR_INVALID, "carry_stride < carry length");
    if (n < H) return fail(MDSM_ERR_INVALID, "block shorter than the carried samples");
    SqrtW sw;
    for (int i = 0; i < 11; ++i) sw.v[i] = std::sqrt((double)(1u << i));
    BoxStream bs{d_carry, carry_stride, carry_valid ? (int32_t)H : 0, 1};
    hipStream_t st = (hipStream_t)stream;
    {
        TimedLaunch tl("detect", st);
        hipError_t e = launch_boxcar_stream(d_x, ndms, n, stride, bs, mask, thr, sw, d_cands, cap,
                                            d_count, d_mean, d_sigma, true, st);
        if (e != hipSuccess) return hip_fail(e, "detect launch");
    }
    // the block's last H samples are the next block's carry
    hipError_t e = launch_copy2d(d_carry, carry_stride * 4, d_x + (n - H), stride * 4,
                                 (int64_t)H * 4, ndms, st);
    if (e != hipSuccess) return hip_fail(e, "carry save");
    return MDSM_OK;
}

extern "C" uint64_t mdsm_median5_stats_len(uint32_t ndms, uint32_t n)
{
    uint64_t mf = 0;
    const uint64_t chunks = brute_stat_chunks(ndms, n, &mf);
    return 2 + 2 * std::max<uint64_t>(ndms, chunks);
}

extern "C" int mdsm_detect_median5(const float *d_x, uint32_t ndms, uint32_t n, int64_t stride,
                                   const float *d_dm_values, double thr, int32_t per_row_stats,
                                   const float *d_chunk_mean, const float *d_chunk_rms,
                                   uint32_t samples_per_chunk, mdsm_cand *d_cands, uint32_t cap,
                                   uint32_t *d_count, double *d_stats, void *stream)
{
    if (!d_x || !d_dm_values || !d_cands || !d_count || !d_stats)
        return fail(MDSM_ERR_INVALID, "NULL argument");
    if (stride < (int64_t)n) return fail(MDSM_ERR_INVALID, "bad row length/stride");
    if (ndms > 65535) return fail(MDSM_ERR_INVALID, "ndms > 65535");
    int mode = per_row_stats ? 1 : 0;
    if (d_chunk_mean || d_chunk_rms) {
        if (!d_chunk_mean || !d_chunk_rms || samples_per_chunk == 0)
            return fail(MDSM_ERR_INVALID, "chunk statistics need mean, rms and samples_per_chunk");
        mode = 2;
    }
    TimedLaunch tl("median5", (hipStream_t)stream);
    hipError_t e = launch_median5(d_x, ndms, n, stride, d_dm_values, thr, mode, d_chunk_mean,
                                  d_chunk_rms, samples_per_chunk, d_cands, cap, d_count, d_stats,
                                  (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "median5 launch");
    return MDSM_OK;
}

// ---------------------------------------------------------------------------
// Seeded synthetic filterbank
// ---------------------------------------------------------------------------

namespace {

inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Standard normal from a counter: Box-Muller on two 53-bit uniforms.
inline double normal_at(uint64_t seed, uint64_t idx)
{
    const uint64_t h1 = mix64(seed ^ mix64(idx));
    const uint64_t h2 = mix64(h1 ^ 0xD1B54A32D192ED03ull);
    const double u1 = ((double)(h1 >> 11) + 1.0) * (1.0 / 9007199254740992.0);  // (0,1]
    const double u2 = (double)(h2 >> 11) * (1.0 / 9007199254740992.0);          // [0,1)
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

inline double uniform_at(uint64_t seed, uint64_t idx)
{
    return (double)(mix64(seed ^ mix64(idx ^ 0xA5A5A5A5DEADBEEFull)) >> 11) *
           (1.0 / 9007199254740992.0);
}

}  // namespace

extern "C" int mdsm_synth_filterbank(const mdsm_synth_spec *sp, void *host_out, int64_t out_stride)
{
    if (!sp || !host_out) return fail(MDSM_ERR_INVALID, "NULL argument");
    if (sp->dtype < MDSM_U8 || sp->dtype > MDSM_F32) return fail(MDSM_ERR_INVALID, "bad dtype");
    if (out_stride < (int64_t)sp->nsamp) return fail(MDSM_ERR_INVALID, "out_stride < nsamp");
    if (sp->npulses && !sp->pulses) return fail(MDSM_ERR_INVALID, "pulses is NULL");
    if (!(sp->tsamp > 0.0f)) return fail(MDSM_ERR_INVALID, "tsamp must be > 0");
    const uint32_t C = sp->nchans, S = sp->nsamp;
    const uint64_t seed = mix64(sp->seed);
    const double itop = 1.0 / ((double)sp->fch1 * (double)sp->fch1);
    unsigned nth = sp->nthreads ? sp->nthreads : std::max(1u, std::thread::hardware_concurrency());
    nth = std::min<unsigned>(nth, std::max<uint32_t>(C, 1));
    auto work = [&](uint32_t c_begin, uint32_t c_end) {
        std::vector<double> row(S);
        for (uint32_t c = c_begin; c < c_end; ++c) {
            const uint64_t rowid = (uint64_t)c * S;
            for (uint32_t t = 0; t < S; ++t) row[t] = normal_at(seed, rowid + t);
            const double f = (double)sp->fch1 + (double)sp->foff * (double)c;
            const double dinv = 1.0 / (f * f) - itop;
            for (uint32_t k = 0; k < sp->npulses; ++k) {
                const mdsm_pulse &P = sp->pulses[k];
                // MDSM_SHIFT_NS rule: nearest sample of the FP64 delay.
                const double v = 4.148808e3 * P.dm * dinv;
                const int64_t sh = std::llround(v / (double)sp->tsamp);
                const int64_t a = (int64_t)P.t0 + sh;
                for (int64_t t = std::max<int64_t>(a, 0);
                     t < std::min<int64_t>(a + (int64_t)P.width, (int64_t)S); ++t)
                    row[t] += (double)P.amp;
            }
            const double g = (double)sp->gain_lo +
                             ((double)sp->gain_hi - (double)sp->gain_lo) * uniform_at(seed, c);
            const double m = sp->mean, s = sp->sigma;
            if (sp->dtype == MDSM_U8) {
                uint8_t *o = (uint8_t *)host_out + (size_t)c * out_stride;
                for (uint32_t t = 0; t < S; ++t) {
                    double v = std::floor(m + s * g * row[t] + 0.5);
                    o[t] = (uint8_t)std::min(255.0, std::max(0.0, v));
                }
            } else if (sp->dtype == MDSM_F16) {
                _Float16 *o = (_Float16 *)host_out + (size_t)c * out_stride;
                for (uint32_t t = 0; t < S; ++t) o[t] = (_Float16)(float)(m + s * g * row[t]);
            } else {
                float *o = (float *)host_out + (size_t)c * out_stride;
                for (uint32_t t = 0; t < S; ++t) o[t] = (float)(m + s * g * row[t]);
            }
        }
    };
    std::vector<std::thread> th;
    const uint32_t per = (C + nth - 1) / nth;
    for (unsigned i = 0; i < nth; ++i) {
        uint32_t a = i * per, b = std::min(C, a + per);
        if (a >= b) break;
        th.emplace_back(work, a, b);
    }
    for (auto &t : th) t.join();
    return MDSM_OK;
}
