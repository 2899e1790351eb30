// Drop-in replacement of MDSM's dedispersion manager (include/mdsm_compat.h).
//
// Reference: src/mdsm/dedispersion_manager.cpp (processSurveyParameters
// :40-199, initialiseMDSM :269-382, tearDownMDSM :385-413, next_chunk
// :416-484, start_processing :487-501) and the output thread of
// src/mdsm/dedispersion_output.cpp:56-356.
//
// MI355X design: no worker pthreads or barriers.  start_processing enqueues
// block k on one HIP stream -- overlap restore (device to device), H2D of the
// new samples, dedispersion, median-of-five detection, D2H of the output and
// the candidates into host slot k%3 -- and returns once the H2D has
// completed, so the caller may refill the input buffer.  next_chunk waits
// for block k-2 (the reference's pipeline latency) and writes its candidates
// in the reference's text format.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/mdsm_compat.h"
#include "mdsm_internal.h"

using namespace mdsm;

// ---------------------------------------------------------------------------
// Minimal XML reader: elements, attributes, text; comments, declarations and
// CDATA-free documents (the observation schema needs nothing more).
// ---------------------------------------------------------------------------
namespace {

struct XNode {
    std::string tag;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::string text;
    std::vector<XNode> kids;

    const std::string *attr(const char *name) const
    {
        for (auto &a : attrs)
            if (a.first == name) return &a.second;
        return nullptr;
    }
};

std::string lower(std::string s)
{
    for (auto &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

std::string decode_entities(const std::string &s)
{
    std::string o;
    for (size_t i = 0; i < s.size(); ++i) {
        if (s[i] == '&') {
            size_t e = s.find(';', i);
            if (e != std::string::npos) {
                std::string ent = s.substr(i + 1, e - i - 1);
                if (ent == "lt") o += '<';
                else if (ent == "gt") o += '>';
                else if (ent == "amp") o += '&';
                else if (ent == "quot") o += '"';
                else if (ent == "apos") o += '\'';
                else o += s.substr(i, e - i + 1);
                i = e;
                continue;
            }
        }
        o += s[i];
    }
    return o;
}

struct XParser {
    const char *p, *end;
    std::string err;

    void skip_ws()
    {
        while (p < end && std::isspace((unsigned char)*p)) ++p;
    }
    bool starts(const char *s) const
    {
        size_t n = std::strlen(s);
        return (size_t)(end - p) >= n && std::memcmp(p, s, n) == 0;
    }
    bool skip_misc()
    {
        for (;;) {
            skip_ws();
            if (starts("<?")) {
                const char *q = std::strstr(p, "?>");
                if (!q || q >= end) return fail_at("unterminated declaration");
                p = q + 2;
            } else if (starts("<!--")) {
                const char *q = std::strstr(p, "-->");
                if (!q || q >= end) return fail_at("unterminated comment");
                p = q + 3;
            } else if (starts("<!")) {
                const char *q = (const char *)std::memchr(p, '>', end - p);
                if (!q) return fail_at("unterminated DOCTYPE");
                p = q + 1;
            } else {
                return true;
            }
        }
    }
    bool fail_at(const std::string &m)
    {
        if (err.empty()) err = m;
        return false;
    }
    bool name(std::string &out)
    {
        const char *s = p;
        while (p < end && (std::isalnum((unsigned char)*p) || *p == '_' || *p == '-' || *p == ':' ||
                           *p == '.'))
            ++p;
        out.assign(s, p);
        return !out.empty();
    }
    bool element(XNode &n)
    {
        if (p >= end || *p != '<') return fail_at("expected '<'");
        ++p;
        if (!name(n.tag)) return fail_at("expected element name");
        for (;;) {
            skip_ws();
            if (p >= end) return fail_at("unterminated tag <" + n.tag + ">");
            if (*p == '/') {
                if (p + 1 >= end || p[1] != '>') return fail_at("bad empty tag");
                p += 2;
                return true;
            }
            if (*p == '>') {
                ++p;
                break;
            }
            std::string an;
            if (!name(an)) return fail_at("bad attribute in <" + n.tag + ">");
            skip_ws();
            if (p >= end || *p != '=') return fail_at("expected '=' after " + an);
            ++p;
            skip_ws();
            if (p >= end || (*p != '"' && *p != '\'')) return fail_at("unquoted attribute " + an);
            char q = *p++;
            const char *s = p;
            while (p < end && *p != q) ++p;
            if (p >= end) return fail_at("unterminated attribute " + an);
            n.attrs.emplace_back(an, decode_entities(std::string(s, p)));
            ++p;
        }
        for (;;) {
            const char *s = p;
            while (p < end && *p != '<') ++p;
            n.text += decode_entities(std::string(s, p));
            if (p >= end) return fail_at("unterminated element <" + n.tag + ">");
            if (starts("<!--")) {
                const char *q = std::strstr(p, "-->");
                if (!q || q >= end) return fail_at("unterminated comment");
                p = q + 3;
                continue;
            }
            if (starts("</")) {
                p += 2;
                std::string cn;
                name(cn);
                skip_ws();
                if (p >= end || *p != '>' || cn != n.tag)
                    return fail_at("mismatched </" + cn + "> for <" + n.tag + ">");
                ++p;
                return true;
            }
            n.kids.emplace_back();
            if (!element(n.kids.back())) return false;
        }
    }
};

// Qt's QString::toFloat/toUInt: the whole (trimmed) string must parse, else 0.
float to_float(const std::string *s, float def = 0.0f)
{
    if (!s) return def;
    std::string t = *s;
    while (!t.empty() && std::isspace((unsigned char)t.back())) t.pop_back();
    size_t i = 0;
    while (i < t.size() && std::isspace((unsigned char)t[i])) ++i;
    t = t.substr(i);
    if (t.empty()) return 0.0f;
    char *e = nullptr;
    double v = std::strtod(t.c_str(), &e);
    if (*e != '\0') return 0.0f;
    return (float)v;
}

unsigned to_uint(const std::string *s, unsigned def = 0)
{
    if (!s) return def;
    std::string t = *s;
    t.erase(0, t.find_first_not_of(" \t\r\n"));
    t.erase(t.find_last_not_of(" \t\r\n") + 1);
    if (t.empty()) return 0;
    char *e = nullptr;
    unsigned long v = std::strtoul(t.c_str(), &e, 10);
    if (*e != '\0' || t[0] == '-') return 0;
    return (unsigned)v;
}

int to_int(const std::string *s, int def = 0)
{
    if (!s) return def;
    char *e = nullptr;
    long v = std::strtol(s->c_str(), &e, 10);
    if (*e != '\0') return 0;
    return (int)v;
}

}  // namespace

extern "C" void mdsm_free_survey(SURVEY *s)
{
    if (!s) return;
    std::free(s->pass_parameters);
    std::free(s->gpu_ids);
    std::free(s);
}

extern "C" SURVEY *mdsm_parse_survey_xml(const char *xml, size_t len)
{
    if (!xml) {
        fail(MDSM_ERR_INVALID, "NULL xml");
        return nullptr;
    }
    XParser P{xml, xml + len, {}};
    XNode root;
    if (!P.skip_misc() || !P.element(root)) {
        fail(MDSM_ERR_PARSE, "observation XML parse error: " + P.err);
        return nullptr;
    }
    if (lower(root.tag) != "observation") {
        fail(MDSM_ERR_PARSE, "Invalid root element of observation parameter xml file, should be "
                             "'observation'");
        return nullptr;
    }
    auto *s = (SURVEY *)std::calloc(1, sizeof(SURVEY));
    // defaults (src/mdsm/dedispersion_manager.cpp:84-95)
    s->useBruteForce = false;
    s->useL1Cache = true;
    s->detection_threshold = 5.0f;
    std::strcpy(s->fileprefix, "output");
    std::strcpy(s->basedir, ".");
    s->secs_per_file = 600;
    s->use_pc_time = 1;
    s->single_file_mode = 0;
    std::vector<SUBBAND_PASSES> passes;
    for (const XNode &e : root.kids) {
        const std::string t = lower(e.tag);
        if (t == "frequencies") {
            s->fch1 = to_float(e.attr("top"));
            s->foff = to_float(e.attr("offset"));
        } else if (t == "dm") {
            s->lowdm = to_float(e.attr("lowDM"));
            s->tdms = to_uint(e.attr("numDMs"));
            s->dmstep = to_float(e.attr("dmStep"));
            s->useBruteForce = to_int(e.attr("useBruteForce")) != 0;
            s->useL1Cache = to_int(e.attr("useL1Cache")) != 0;
        } else if (t == "channels") {
            s->nchans = to_uint(e.attr("number"));
            s->nsubs = to_uint(e.attr("subbands"));
        } else if (t == "timing") {
            s->tsamp = to_float(e.attr("tsamp"));
        } else if (t == "samples") {
            s->nsamp = to_uint(e.attr("number"));
        } else if (t == "detection") {
            s->detection_threshold = to_float(e.attr("threshold"));
        } else if (t == "output") {
            const std::string *fp = e.attr("filePrefix"), *bd = e.attr("baseDirectory");
            std::snprintf(s->fileprefix, sizeof(s->fileprefix), "%s", fp ? fp->c_str() : "output");
            std::snprintf(s->basedir, sizeof(s->basedir), "%s", bd ? bd->c_str() : ".");
            s->secs_per_file = to_uint(e.attr("secondsPerFile"), 600);
            s->use_pc_time = (char)to_uint(e.attr("usePCTime"), 1);
            s->single_file_mode = (char)to_uint(e.attr("singleFileMode"), 0);
        } else if (t == "gpus") {
            const std::string *ids = e.attr("ids");
            std::vector<unsigned> v;
            if (ids) {
                std::stringstream ss(*ids);
                std::string tok;
                while (std::getline(ss, tok, ','))
                    if (!tok.empty()) v.push_back(to_uint(&tok));
            }
            if (!v.empty()) {
                s->gpu_ids = (unsigned *)std::malloc(v.size() * sizeof(unsigned));
                std::copy(v.begin(), v.end(), s->gpu_ids);
            }
            s->num_gpus = (unsigned)v.size();
        } else if (t == "passes") {
            for (const XNode &pn : e.kids) {
                SUBBAND_PASSES sp{};
                for (const XNode &f : pn.kids) {
                    const std::string ft = lower(f.tag);
                    const std::string *tx = &f.text;
                    if (ft == "lowdm") sp.lowdm = to_float(tx);
                    else if (ft == "highdm") sp.highdm = to_float(tx);
                    else if (ft == "deltadm") sp.dmstep = to_float(tx);
                    else if (ft == "downsample") sp.binsize = (int)to_uint(tx);
                    else if (ft == "subdm") sp.sub_dmstep = to_float(tx);
                    else if (ft == "numdms") sp.ndms = (int)to_uint(tx);
                    else if (ft == "dmspercall") sp.calldms = (int)to_uint(tx);
                    else if (ft == "ncalls") sp.ncalls = (int)to_uint(tx);
                }
                passes.push_back(sp);
            }
        }
    }
    s->num_passes = (unsigned)passes.size();
    if (!passes.empty()) {
        s->pass_parameters = (SUBBAND_PASSES *)std::malloc(passes.size() * sizeof(SUBBAND_PASSES));
        std::copy(passes.begin(), passes.end(), s->pass_parameters);
    }
    if (!s->useBruteForce) {
        s->tdms = 0;
        for (auto &q : passes) s->tdms += (unsigned)q.ndms;
    }
    return s;
}

extern "C" SURVEY *processSurveyParameters(const char *filepath)
{
    if (!filepath) {
        fail(MDSM_ERR_INVALID, "NULL path");
        return nullptr;
    }
    std::ifstream f(filepath, std::ios::binary);
    if (!f) {
        fail(MDSM_ERR_PARSE, std::string("Cannot open observation file '") + filepath + "'");
        return nullptr;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    std::string x = ss.str();
    return mdsm_parse_survey_xml(x.data(), x.size());
}

// ---------------------------------------------------------------------------
// The pipeline
// ---------------------------------------------------------------------------
namespace {

constexpr int kSlots = 3;          // host output slots (MDSM_STAGES)
constexpr uint32_t kCandCap = 1u << 20;

struct Slot {
    float *out_host = nullptr;       // reference layout
    mdsm_cand *cand_host = nullptr;  // per range, kCandCap each
    uint32_t *count_host = nullptr;  // per range
    hipEvent_t done = nullptr;
    unsigned nsamp = 0;
    double timestamp = 0, rate = 0;
    float *mean = nullptr, *rms = nullptr;
    unsigned spc = 0;                // samples per chunk of mean/rms, 0 = not used
    bool valid = false;
};

struct Compat {
    SURVEY *sv = nullptr;
    mdsm_plan *plan = nullptr;
    hipStream_t stream = nullptr;
    unsigned C = 0, nsamp = 0, M = 0;
    float *in_host = nullptr;
    float *d_in = nullptr, *d_tail = nullptr, *d_out = nullptr;
    float *d_dmv = nullptr, *d_cmean = nullptr, *d_crms = nullptr;
    double *d_stats = nullptr;
    mdsm_cand *d_cands = nullptr;
    uint32_t *d_count = nullptr;
    void *d_work = nullptr;
    std::vector<float> dmv;
    uint64_t out_floats_ref = 0;     // reference-layout output floats
    Slot slot[kSlots];
    long started = 0;                // blocks enqueued
    long calls = 0;                  // next_chunk calls
    bool flushed = false;
    // published by next_chunk for the next start_processing
    unsigned pend_nsamp = 0;
    double pend_ts = 0, pend_rate = 0;
    float *pend_mean = nullptr, *pend_rms = nullptr;
    // candidate file
    FILE *fp = nullptr;
    long written_samples = 0;
    // new samples arrive in d_in, written on `stream` by the caller (the
    // module adapter's GPU corner turn), not in the host input buffer
    bool dev_input = false;
};

Compat *g = nullptr;

bool hipok(hipError_t e, const char *what)
{
    if (e == hipSuccess) return true;
    hip_fail(e, what);
    return false;
}

void close_compat(Compat *c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &s : c->slot) {
        if (s.out_host) (void)hipHostFree(s.out_host);
        if (s.cand_host) (void)hipHostFree(s.cand_host);
        if (s.count_host) (void)hipHostFree(s.count_host);
        if (s.done) (void)hipEventDestroy(s.done);
    }
    if (c->in_host) (void)hipHostFree(c->in_host);
    for (void *p : {(void *)c->d_in, (void *)c->d_tail, (void *)c->d_out, (void *)c->d_dmv,
                    (void *)c->d_cmean, (void *)c->d_crms, (void *)c->d_stats, (void *)c->d_cands,
                    (void *)c->d_count, c->d_work})
        if (p) (void)hipFree(p);
    if (c->plan) mdsm_plan_destroy(c->plan);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->fp) std::fclose(c->fp);
    delete c;
}

void open_output_file(Compat *c, double ts)
{
    SURVEY *s = c->sv;
    std::string path = std::string(s->basedir) + "/" + s->fileprefix;
    if (s->single_file_mode) {
        path += ".dat";
    } else {
        // prefix_<%F_%T>_<secs>.dat (src/mdsm/dedispersion_output.cpp:254-281)
        time_t t = s->use_pc_time ? time(nullptr) : (time_t)ts;
        struct tm tmv;
        localtime_r(&t, &tmv);
        char buf[64];
        std::strftime(buf, sizeof(buf), "%F_%T", &tmv);
        path += std::string("_") + buf + "_" + std::to_string(s->secs_per_file) + ".dat";
    }
    c->fp = std::fopen(path.c_str(), "w");
}

// Write the candidates of a completed slot in the reference's text format.
void write_candidates(Compat *c, Slot &sl)
{
    SURVEY *s = c->sv;
    if (!c->fp) open_output_file(c, sl.timestamp);
    if (!c->fp) return;
    const uint32_t nr = c->plan->info.nranges;
    for (uint32_t r = 0; r < nr; ++r) {
        const mdsm_range_info &ri = c->plan->ranges[r].info;
        uint32_t n = std::min(sl.count_host[r], kCandCap);
        std::vector<mdsm_cand> v(sl.cand_host + (size_t)r * kCandCap,
                                 sl.cand_host + (size_t)r * kCandCap + n);
        std::sort(v.begin(), v.end(), [](const mdsm_cand &a, const mdsm_cand &b) {
            return a.dm != b.dm ? a.dm < b.dm : a.t < b.t;
        });
        const double step = sl.rate * (double)ri.downsample;
        for (const mdsm_cand &cd : v) {
            const float dm = c->dmv[ri.first_dm + cd.dm];
            if (sl.spc) {
                const unsigned ch = cd.t / sl.spc;
                std::fprintf(c->fp, "%lf, %f, %f, %f, %f  \n", sl.timestamp + cd.t * step, dm,
                             cd.snr, sl.mean[ch], sl.rms[ch]);
            } else {
                std::fprintf(c->fp, "%lf, %f, %f\n", sl.timestamp + cd.t * step, dm, cd.snr);
            }
        }
    }
    std::fflush(c->fp);
    if (!s->single_file_mode) {
        c->written_samples += sl.nsamp;
        if (c->written_samples * sl.rate > s->secs_per_file) {
            c->written_samples = 0;
            std::fclose(c->fp);
            c->fp = nullptr;
        }
    }
}

}  // namespace

extern "C" float *initialiseMDSM(SURVEY *sv)
{
    if (!sv) {
        fail(MDSM_ERR_INVALID, "NULL survey");
        return nullptr;
    }
    if (g) {
        fail(MDSM_ERR_STATE, "initialiseMDSM called twice without tearDownMDSM");
        return nullptr;
    }
    auto *c = new Compat();
    c->sv = sv;
    c->C = sv->nchans;
    int dev = (sv->gpu_ids && sv->num_gpus) ? (int)sv->gpu_ids[0] : 0;
    if (!hipok(hipSetDevice(dev), "hipSetDevice")) {
        delete c;
        return nullptr;
    }
    // maxshift: the largest delay of the plan (dedispersion_manager.cpp:291-297)
    mdsm_obs obs{};
    obs.nchans = sv->nchans;
    obs.fch1 = sv->fch1;
    obs.foff = sv->foff;
    obs.tsamp = sv->tsamp;
    obs.dtype = MDSM_F32;
    obs.shift_rule = MDSM_SHIFT_REF;
    std::vector<mdsm_range> ranges;
    std::vector<mdsm_pass> passes;
    unsigned kmax = 1;
    if (sv->useBruteForce) {
        ranges.push_back({sv->lowdm, sv->dmstep, sv->tdms, 1});
    } else {
        if (!sv->num_passes || !sv->nsubs) {
            fail(MDSM_ERR_INVALID, "subband survey without passes/subbands");
            delete c;
            return nullptr;
        }
        for (unsigned i = 0; i < sv->num_passes; ++i) {
            const SUBBAND_PASSES &q = sv->pass_parameters[i];
            passes.push_back({q.lowdm, q.highdm, q.dmstep, q.sub_dmstep, (uint32_t)q.binsize,
                              (uint32_t)q.ndms, (uint32_t)q.ncalls, (uint32_t)q.calldms});
            kmax = std::max(kmax, (unsigned)std::max(1, q.binsize));
        }
    }
    // First plan at a nominal length to learn maxshift, then the real one.
    auto make = [&](uint32_t S, mdsm_plan **p) {
        obs.nsamp_in = S;
        return sv->useBruteForce
                   ? mdsm_plan_create(&obs, ranges.data(), (uint32_t)ranges.size(), p)
                   : mdsm_plan_create_subband(&obs, sv->nsubs, passes.data(),
                                              (uint32_t)passes.size(), p);
    };
    mdsm_plan *probe = nullptr;
    if (make(1u << 30, &probe)) {
        delete c;
        return nullptr;
    }
    uint32_t M = probe->info.maxshift;
    mdsm_plan_destroy(probe);
    M = (M + kmax - 1) / kmax * kmax;  // the overlap spans whole decimated samples
    if (sv->nsamp == 0) {
        // Size a block to a quarter of free device memory (the reference fills
        // the GPU: dedispersion_manager.cpp:240-250).
        size_t fr = 0, tot = 0;
        (void)hipMemGetInfo(&fr, &tot);
        double per_sample = 4.0 * sv->nchans + 4.0 * std::max(1u, sv->tdms) * 1.25;
        double n = (double)fr * 0.25 / per_sample - M;
        n = std::min(n, (double)(1u << 22));
        sv->nsamp = (unsigned)std::max(1024.0, n);
    }
    sv->nsamp -= sv->nsamp % kmax;
    sv->maxshift = M;
    sv->num_threads = 1;
    c->nsamp = sv->nsamp;
    c->M = M;
    if (make(c->nsamp + M, &c->plan)) {
        delete c;
        return nullptr;
    }
    const uint64_t S = (uint64_t)c->nsamp + M;
    const mdsm_plan_info &pi = c->plan->info;
    c->out_floats_ref = 0;
    for (auto &rh : c->plan->ranges)
        c->out_floats_ref += (uint64_t)rh.info.ndms * (c->nsamp / rh.info.downsample);
    c->dmv.resize(pi.ndms_total);
    mdsm_plan_dm_values(c->plan, c->dmv.data());
    const uint32_t nr = pi.nranges;
    bool ok = hipok(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "stream") &&
              hipok(hipHostMalloc((void **)&c->in_host, S * c->C * 4), "host input") &&
              hipok(hipMalloc((void **)&c->d_in, S * c->C * 4), "device input") &&
              hipok(hipMalloc((void **)&c->d_tail, (size_t)std::max(1u, M) * c->C * 4), "tail") &&
              hipok(hipMalloc((void **)&c->d_out, pi.out_elems * 4), "device output") &&
              hipok(hipMalloc((void **)&c->d_dmv, pi.ndms_total * 4), "dm values") &&
              hipok(hipMalloc((void **)&c->d_stats,
                                  mdsm_median5_stats_len(pi.ndms_total, c->nsamp) * 8),
                        "stats") &&
              hipok(hipMalloc((void **)&c->d_cands, (size_t)nr * kCandCap * sizeof(mdsm_cand)),
                    "candidates") &&
              hipok(hipMalloc((void **)&c->d_count, nr * 4), "count") &&
              hipok(hipMalloc(&c->d_work, std::max<uint64_t>(pi.work_bytes, 256)), "workspace") &&
              hipok(hipMemcpy(c->d_dmv, c->dmv.data(), pi.ndms_total * 4, hipMemcpyHostToDevice),
                    "dm values");
    for (auto &s : c->slot) {
        ok = ok && hipok(hipHostMalloc((void **)&s.out_host, c->out_floats_ref * 4), "host output") &&
             hipok(hipHostMalloc((void **)&s.cand_host, (size_t)nr * kCandCap * sizeof(mdsm_cand)),
                   "host candidates") &&
             hipok(hipHostMalloc((void **)&s.count_host, nr * 4), "host counts") &&
             hipok(hipEventCreateWithFlags(&s.done, hipEventDisableTiming), "event");
    }
    if (!ok) {
        close_compat(c);
        return nullptr;
    }
    std::memset(c->in_host, 0, S * c->C * 4);
    g = c;
    return c->in_host;
}

extern "C" float *next_chunk(unsigned int data_read, unsigned *samples, double timestamp,
                             double blockRate, float *mean, float *rms)
{
    if (!g) {
        fail(MDSM_ERR_STATE, "next_chunk before initialiseMDSM");
        if (samples) *samples = 0;
        return nullptr;
    }
    Compat *c = g;
    // round a short block down to a multiple of the largest decimation
    // (dedispersion_manager.cpp:452-460)
    unsigned kmax = 1;
    for (auto &rh : c->plan->ranges) kmax = std::max(kmax, rh.info.downsample);
    if (data_read > c->nsamp) data_read = c->nsamp;
    data_read -= data_read % kmax;
    c->pend_nsamp = data_read;
    c->pend_ts = timestamp;
    c->pend_rate = blockRate;
    c->pend_mean = mean;
    c->pend_rms = rms;
    const long want = c->calls - 2;  // block delivered by this call
    c->calls++;
    if (want < 0 || want >= c->started) {
        if (samples) *samples = (unsigned)-1;
        return nullptr;
    }
    Slot &sl = c->slot[want % kSlots];
    if (!hipok(hipEventSynchronize(sl.done), "waiting for block")) {
        if (samples) *samples = 0;
        return nullptr;
    }
    write_candidates(c, sl);
    if (samples) *samples = sl.nsamp;
    return sl.out_host;
}

extern "C" int start_processing(unsigned int data_read)
{
    if (!g) {
        fail(MDSM_ERR_STATE, "start_processing before initialiseMDSM");
        return 0;
    }
    Compat *c = g;
    if (data_read == 0) {
        // the first end-of-stream call keeps the loop alive for the last
        // buffered block (dedispersion_manager.cpp:489-493)
        if (!c->flushed) {
            c->flushed = true;
            return (int)(c->started + 1);
        }
        return 0;
    }
    const unsigned nb = c->pend_nsamp ? c->pend_nsamp : std::min(data_read, c->nsamp);
    const long k = c->started;
    Slot &sl = c->slot[k % kSlots];
    const size_t S = (size_t)c->nsamp + c->M, C = c->C, M = c->M;
    hipStream_t st = c->stream;
    bool ok = true;
    if (k == 0) {
        if (!c->dev_input)
            ok = hipok(hipMemcpyAsync(c->d_in, c->in_host, S * C * 4, hipMemcpyHostToDevice, st),
                       "H2D");
    } else {
        // overlap: last M samples of the previous block, kept on the device
        if (M)
            ok = hipok(hipMemcpy2DAsync(c->d_in, S * 4, c->d_tail, M * 4, M * 4, C,
                                        hipMemcpyDeviceToDevice, st), "overlap restore");
        if (!c->dev_input)
            ok = ok && hipok(hipMemcpy2DAsync(c->d_in + M, S * 4, c->in_host + M, S * 4,
                                              (size_t)c->nsamp * 4, C, hipMemcpyHostToDevice, st),
                             "H2D");
    }
    // a short last block: zero the samples past the end
    if (ok && nb < c->nsamp)
        ok = hipok(hipMemset2DAsync(c->d_in + M + nb, S * 4, 0, (size_t)(c->nsamp - nb) * 4, C, st),
                   "zero tail");
    // keep this block's tail for the next one
    if (ok && M)
        ok = hipok(hipMemcpy2DAsync(c->d_tail, M * 4, c->d_in + c->nsamp, S * 4, M * 4, C,
                                    hipMemcpyDeviceToDevice, st), "overlap save");
    // the caller may rewrite the host input buffer once the H2D has completed
    if (!c->dev_input) ok = ok && hipok(hipStreamSynchronize(st), "H2D sync");
    if (!ok) return 0;
    if (mdsm_dedisperse(c->plan, c->d_in, (int64_t)S, c->d_out, c->d_work, st)) return 0;
    // detection and D2H per range, in the reference's output layout
    SURVEY *s = c->sv;
    uint64_t ho = 0;
    const uint32_t nr = c->plan->info.nranges;
    for (uint32_t r = 0; r < nr && ok; ++r) {
        const mdsm_range_info &ri = c->plan->ranges[r].info;
        const uint32_t n = nb / ri.downsample;
        const uint32_t ncols = c->nsamp / ri.downsample;
        const float *src = c->d_out + ri.out_offset;
        const float *cm = nullptr, *cr = nullptr;
        if (c->pend_mean && c->pend_rms && s->useBruteForce && s->samplesPerChunk) {
            const size_t nchunk = (S + s->samplesPerChunk - 1) / s->samplesPerChunk + 1;
            if (!c->d_cmean) {
                ok = hipok(hipMalloc((void **)&c->d_cmean, nchunk * 4), "chunk mean") &&
                     hipok(hipMalloc((void **)&c->d_crms, nchunk * 4), "chunk rms");
            }
            ok = ok && hipok(hipMemcpyAsync(c->d_cmean, c->pend_mean, nchunk * 4,
                                            hipMemcpyHostToDevice, st), "chunk mean") &&
                 hipok(hipMemcpyAsync(c->d_crms, c->pend_rms, nchunk * 4, hipMemcpyHostToDevice, st),
                       "chunk rms");
            cm = c->d_cmean;
            cr = c->d_crms;
        }
        if (ok && mdsm_detect_median5(src, ri.ndms, n, ri.nout, c->d_dmv + ri.first_dm,
                                      s->detection_threshold, s->useBruteForce ? 0 : 1, cm, cr,
                                      s->samplesPerChunk, c->d_cands + (size_t)r * kCandCap,
                                      kCandCap, c->d_count + r, c->d_stats, st))
            return 0;
        ok = ok && hipok(hipMemcpy2DAsync(sl.out_host + ho, (size_t)ncols * 4, src,
                                          (size_t)ri.nout * 4, (size_t)ncols * 4, ri.ndms,
                                          hipMemcpyDeviceToHost, st), "D2H output");
        ho += (uint64_t)ri.ndms * ncols;
    }
    ok = ok &&
         hipok(hipMemcpyAsync(sl.count_host, c->d_count, nr * 4, hipMemcpyDeviceToHost, st),
               "D2H counts") &&
         hipok(hipMemcpyAsync(sl.cand_host, c->d_cands, (size_t)nr * kCandCap * sizeof(mdsm_cand),
                              hipMemcpyDeviceToHost, st), "D2H candidates") &&
         hipok(hipEventRecord(sl.done, st), "event");
    if (!ok) return 0;
    sl.nsamp = nb;
    sl.timestamp = c->pend_ts;
    sl.rate = c->pend_rate;
    sl.mean = c->pend_mean;
    sl.rms = c->pend_rms;
    sl.spc = (c->pend_mean && c->pend_rms && s->useBruteForce) ? s->samplesPerChunk : 0;
    sl.valid = true;
    c->pend_nsamp = 0;
    c->started++;
    return (int)c->started;
}

extern "C" void tearDownMDSM(void)
{
    close_compat(g);
    g = nullptr;
}

extern "C" const float *mdsm_compat_device_output(void) { return g ? g->d_out : nullptr; }

int mdsm::compat_device_input(float **d_in, int64_t *stride, hipStream_t *stream)
{
    if (!g) return fail(MDSM_ERR_STATE, "no manager instance (initialiseMDSM)");
    g->dev_input = true;
    *d_in = g->d_in;
    *stride = (int64_t)g->nsamp + g->M;
    *stream = g->stream;
    return MDSM_OK;
}
