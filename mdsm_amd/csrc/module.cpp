// The Pelican MdsmModule adapter without Pelican (include/mdsm_module.h).
//
// MdsmModule::run (src/pelican-mdsm/modules/MdsmModule.cpp:49-257) fills
// the manager's input buffer blob by blob, corner-turning on the host, and
// calls next_chunk / start_processing when a block is complete.  Here a
// blob's spectra are copied as they are (time-major) into a pinned staging
// block; a complete block is uploaded and corner-turned on the GPU into the
// manager's device input rows, on the manager's stream, so the host never
// touches the channel-major layout.  The control flow -- required samples,
// the first block's extra maxshift, timestamps, the weighted blobs'
// per-chunk mean/rms arrays rotated over three stages, end of stream -- is
// the reference's.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/mdsm_module.h"
#include "mdsm_internal.h"

using namespace mdsm;

namespace {
constexpr int kStages = 3;  // MDSM_STAGES: input, processing, output (MdsmModule.h:2-6)
}

struct mdsm_module {
    SURVEY *survey = nullptr;
    bool create_output = false, invert = true;
    unsigned samples = 0, gettime = 0, counter = 0;
    double timestamp = 0, block_rate = 0;
    std::vector<float> means[kStages], rmss[kStages];
    // block staging: pinned time-major blocks (two, alternating), one device copy
    float *d_in = nullptr;
    int64_t stride = 0;
    hipStream_t st = nullptr;
    float *h_stage[2] = {nullptr, nullptr};
    hipEvent_t h_done[2] = {nullptr, nullptr};
    bool h_pending[2] = {false, false};
    float *d_stage = nullptr;
    int cur = 0;
    size_t rows = 0, width = 0;
    // the output of the last run
    std::vector<float> dmv;
    std::vector<uint32_t> lens;
    std::vector<const float *> ptrs;
};

namespace {

void free_module(mdsm_module *m)
{
    if (m->st) (void)hipStreamSynchronize(m->st);
    for (int i = 0; i < 2; ++i) {
        if (m->h_stage[i]) (void)hipHostFree(m->h_stage[i]);
        if (m->h_done[i]) (void)hipEventDestroy(m->h_done[i]);
    }
    if (m->d_stage) (void)hipFree(m->d_stage);
    delete m;
}

// Wait until the staging block `i` is no longer read by an upload.
int stage_ready(mdsm_module *m, int i)
{
    if (!m->h_pending[i]) return MDSM_OK;
    const hipError_t e = hipEventSynchronize(m->h_done[i]);
    if (e != hipSuccess) return hip_fail(e, "waiting for a staging upload");
    m->h_pending[i] = false;
    return MDSM_OK;
}

// Copy spectra [t0, t1) of the blob to staging rows starting at `row`.
void stage_spectra(mdsm_module *m, const float *spectra, uint32_t t0, uint32_t t1,
                   uint32_t nsub, uint32_t nchan, int64_t t_stride, int64_t s_stride, size_t row)
{
    float *dst = m->h_stage[m->cur] + row * m->width;
    for (uint32_t t = t0; t < t1; ++t, dst += m->width) {
        const float *src = spectra + (int64_t)t * t_stride;
        if (s_stride == (int64_t)nchan) {
            std::memcpy(dst, src, m->width * sizeof(float));
        } else {
            for (uint32_t s = 0; s < nsub; ++s)
                std::memcpy(dst + (size_t)s * nchan, src + (int64_t)s * s_stride,
                            nchan * sizeof(float));
        }
    }
}

// Upload the first `n` staging rows and corner-turn them into the device
// input rows at column col0 (channel inversion of the whole band: subband s,
// channel c of the blob is row (nsub-1-s)*nchan + (nchan-1-c) when inverting,
// MdsmModule.cpp:137,145).
int upload_block(mdsm_module *m, size_t n, int64_t col0)
{
    if (n == 0) return MDSM_OK;
    const int i = m->cur;
    hipError_t e = hipMemcpyAsync(m->d_stage, m->h_stage[i], n * m->width * sizeof(float),
                                  hipMemcpyHostToDevice, m->st);
    if (e == hipSuccess) e = hipEventRecord(m->h_done[i], m->st);
    if (e != hipSuccess) return hip_fail(e, "staging upload");
    m->h_pending[i] = true;
    m->cur ^= 1;
    return mdsm_corner_turn(m->d_stage, MDSM_F32, (uint32_t)n, (uint32_t)m->width,
                            (int64_t)m->width, m->invert ? 1 : 0, m->d_in, MDSM_F32, m->stride,
                            col0, m->st);
}

// The DedispersedTimeSeriesF32 of a returned output buffer
// (MdsmModule.cpp:172-217): DM labels as the reference sets them, series of
// `samples` (per pass: samples/binsize) samples; rows keep the manager's
// pitch of survey->nsamp/binsize floats.
void describe_output(mdsm_module *m, const float *buf, unsigned samples)
{
    const SURVEY *s = m->survey;
    m->dmv.clear();
    m->lens.clear();
    m->ptrs.clear();
    if (s->useBruteForce) {
        for (unsigned d = 0; d < s->tdms; ++d) {
            m->dmv.push_back(s->lowdm + s->dmstep * (float)d);
            m->lens.push_back(samples);
            m->ptrs.push_back(buf + (size_t)d * s->nsamp);
        }
        return;
    }
    const unsigned nt = std::max(1u, s->num_threads);
    size_t off = 0;
    for (unsigned th = 0; th < nt; ++th)
        for (unsigned p = 0; p < s->num_passes; ++p) {
            const SUBBAND_PASSES &q = s->pass_parameters[p];
            const unsigned ncalls = (unsigned)q.ncalls / nt;
            const unsigned ndms = ncalls * (unsigned)q.calldms;
            const float startdm = q.lowdm + q.sub_dmstep * (float)ncalls * (float)th;
            const unsigned k = (unsigned)std::max(1, q.binsize);
            for (unsigned d = 0; d < ndms; ++d) {
                m->dmv.push_back(startdm + (float)d * q.dmstep);
                m->lens.push_back(samples / k);
                m->ptrs.push_back(buf + off);
                off += s->nsamp / k;
            }
        }
}

}  // namespace

extern "C" int mdsm_module_create(const char *observation_file, int create_output_blob,
                                  int invert_channels, mdsm_module **module)
{
    if (!observation_file || !module) return fail(MDSM_ERR_INVALID, "NULL argument");
    *module = nullptr;
    SURVEY *s = processSurveyParameters(observation_file);
    if (!s) return MDSM_ERR_PARSE;
    if (!initialiseMDSM(s)) {
        mdsm_free_survey(s);
        return MDSM_ERR_DEVICE;
    }
    auto *m = new mdsm_module();
    m->survey = s;
    m->create_output = create_output_blob != 0;
    m->invert = invert_channels != 0;
    m->rows = (size_t)s->nsamp + s->maxshift;
    m->width = s->nchans;
    int rc = compat_device_input(&m->d_in, &m->stride, &m->st);
    hipError_t e = hipSuccess;
    for (int i = 0; i < 2 && rc == MDSM_OK && e == hipSuccess; ++i) {
        e = hipHostMalloc((void **)&m->h_stage[i], m->rows * m->width * sizeof(float));
        if (e == hipSuccess) e = hipEventCreateWithFlags(&m->h_done[i], hipEventDisableTiming);
    }
    if (rc == MDSM_OK && e == hipSuccess)
        e = hipMalloc((void **)&m->d_stage, m->rows * m->width * sizeof(float));
    if (rc == MDSM_OK && e != hipSuccess) rc = hip_fail(e, "module staging buffers");
    if (rc != MDSM_OK) {
        free_module(m);
        tearDownMDSM();
        mdsm_free_survey(s);
        return rc;
    }
    *module = m;
    return MDSM_OK;
}

extern "C" const SURVEY *mdsm_module_survey(const mdsm_module *module)
{
    return module ? module->survey : nullptr;
}

extern "C" int mdsm_module_run(mdsm_module *m, const float *spectra, uint32_t ntime,
                               uint32_t nsubbands, uint32_t nchannels, int64_t t_stride,
                               int64_t s_stride, int weighted, float mean, float rms,
                               double timestamp, double block_rate, mdsm_module_output *out)
{
    if (!m || !out) return fail(MDSM_ERR_INVALID, "NULL argument");
    *out = mdsm_module_output{0, nullptr, nullptr, nullptr};
    SURVEY *s = m->survey;
    const unsigned maxshift = s->maxshift;
    if (ntime) {
        if (!spectra) return fail(MDSM_ERR_INVALID, "NULL spectra");
        if ((uint64_t)nsubbands * nchannels != s->nchans)
            return fail(MDSM_ERR_INVALID, "nsubbands * nchannels differs from the observation's channels");
        if (s_stride < (int64_t)nchannels || t_stride < (int64_t)(nsubbands - 1) * s_stride + nchannels)
            return fail(MDSM_ERR_INVALID, "spectra strides overlap");
        if (ntime > s->nsamp) return fail(MDSM_ERR_INVALID, "a blob longer than a block");
    }

    // weighted blobs: the blob's mean and rms at its chunk of the block
    // (MdsmModule.cpp:57-96)
    if (weighted && ntime) {
        if (m->means[0].empty()) {
            const size_t chunks = (s->nsamp + maxshift) / ntime;
            s->samplesPerChunk = ntime;
            // the manager reads ceil((nsamp+maxshift)/samplesPerChunk) + 1 entries
            const size_t n = std::max(chunks + 1, ((size_t)s->nsamp + maxshift + ntime - 1) / ntime + 1);
            for (int i = 0; i < kStages; ++i) {
                m->means[i].assign(n, 0.0f);
                m->rmss[i].assign(n, 0.0f);
            }
        }
        const size_t index = m->counter == 0
                                 ? m->samples / ntime
                                 : m->samples / ntime + (size_t)std::floor(maxshift / (float)ntime);
        if (index < m->means[0].size()) {
            m->means[m->counter % kStages][index] = mean;
            m->rmss[m->counter % kStages][index] = rms;
        }
    } else if (!weighted && ntime) {
        // an empty blob (the end of the stream) keeps the chunk length
        s->samplesPerChunk = ntime;
    }

    // the first blob of a block gives its timestamp (MdsmModule.cpp:116-122)
    if (m->gettime == 0) {
        m->timestamp = timestamp;
        m->block_rate = block_rate;
        if (m->counter > 0) m->timestamp = timestamp - block_rate * maxshift;
    }

    unsigned req = m->counter == 0 ? s->nsamp + maxshift : s->nsamp;
    unsigned copy = ntime <= req - m->samples ? ntime : req - m->samples;
    if (ntime == 0) req = copy = 0;

    int rc = stage_ready(m, m->cur);
    if (rc) return rc;
    stage_spectra(m, spectra, 0, copy, nsubbands, nchannels, t_stride, s_stride, m->samples);
    m->samples += copy;
    m->gettime = m->samples;

    if (m->samples == req || ntime == 0) {
        const unsigned num = m->counter == 0 ? m->samples - maxshift : m->samples;
        rc = upload_block(m, m->samples, m->counter == 0 ? 0 : maxshift);
        if (rc) return rc;
        const bool w = !m->means[0].empty();
        unsigned got = 0;
        float *buf = next_chunk(num, &got, m->timestamp, m->block_rate,
                                w ? m->means[m->counter % kStages].data() : nullptr,
                                w ? m->rmss[m->counter % kStages].data() : nullptr);
        m->gettime = m->samples;
        const bool more = start_processing(m->samples) != 0;
        // the reference returns before describing the output once the stream
        // has drained, which drops the last block; it is delivered here
        if (buf && m->create_output) {
            describe_output(m, buf, got);
            out->nseries = (uint32_t)m->dmv.size();
            out->dm_values = m->dmv.data();
            out->nsamp = m->lens.data();
            out->data = m->ptrs.data();
        }
        if (!more) return MDSM_OK;
        ++m->counter;
        m->samples = 0;
        m->gettime = 0;
        // the maxshift part's statistics open the next block (MdsmModule.cpp:226-237)
        if (w && ntime) {
            const size_t from = s->nsamp / ntime, n = maxshift / ntime;
            std::vector<float> &dm = m->means[m->counter % kStages], &dr = m->rmss[m->counter % kStages];
            const std::vector<float> &sm = m->means[(m->counter - 1) % kStages],
                                     &sr = m->rmss[(m->counter - 1) % kStages];
            for (size_t i = 0; i < n && from + i < sm.size() && i < dm.size(); ++i) {
                dm[i] = sm[from + i];
                dr[i] = sr[from + i];
            }
        }
        // the rest of the blob starts the next block
        if (copy < ntime) {
            rc = stage_ready(m, m->cur);
            if (rc) return rc;
            stage_spectra(m, spectra, copy, ntime, nsubbands, nchannels, t_stride, s_stride, 0);
            m->samples += ntime - copy;
        }
    }
    return MDSM_OK;
}

extern "C" void mdsm_module_destroy(mdsm_module *module)
{
    if (!module) return;
    SURVEY *s = module->survey;
    free_module(module);
    tearDownMDSM();
    mdsm_free_survey(s);
}
