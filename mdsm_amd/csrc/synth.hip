// Seeded synthetic filterbank on the device: the generator of
// mdsm_synth_filterbank (host) run as one thread per sample, so that
// benchmark inputs of tens of GB are made in HBM at memory speed.  The value
// of every sample is a function of (seed, channel, sample) only: a
// counter-based hash (splitmix64 finaliser) gives two 53-bit uniforms, the
// Box-Muller transform a standard normal; pulses follow the nearest-sample
// FP64 delay track; a per-channel gain U[lo, hi] scales the noise.  The host
// and device libm may round log/cos differently in the last bit, so the two
// generators agree to rounding, not bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mdsm_internal.h"

using namespace mdsm;

namespace {

constexpr int kMaxSynthPulses = 16;

struct SynthArgs {
    uint64_t seed;
    int32_t dtype;
    uint32_t nchans, nsamp;
    int64_t stride;
    double fch1, foff, tsamp, mean, sigma, gain_lo, gain_hi, itop;
    uint32_t npulses;
    double pdm[kMaxSynthPulses], pamp[kMaxSynthPulses];
    int64_t pt0[kMaxSynthPulses], pw[kMaxSynthPulses];
};

__device__ __forceinline__ uint64_t mix64_d(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void __launch_bounds__(256) synth_kernel(SynthArgs a, void *out)
{
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    const uint32_t c = blockIdx.y;
    if (t >= a.nsamp) return;
    const uint64_t idx = (uint64_t)c * a.nsamp + t;
    const uint64_t h1 = mix64_d(a.seed ^ mix64_d(idx));
    const uint64_t h2 = mix64_d(h1 ^ 0xD1B54A32D192ED03ull);
    const double u1 = ((double)(h1 >> 11) + 1.0) * (1.0 / 9007199254740992.0);
    const double u2 = (double)(h2 >> 11) * (1.0 / 9007199254740992.0);
    double v = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    const double f = a.fch1 + a.foff * (double)c;
    const double dinv = 1.0 / (f * f) - a.itop;
    for (uint32_t k = 0; k < a.npulses; ++k) {
        const double d = 4.148808e3 * a.pdm[k] * dinv;
        const int64_t sh = llround(d / a.tsamp);
        const int64_t s0 = a.pt0[k] + sh;
        if ((int64_t)t >= s0 && (int64_t)t < s0 + a.pw[k]) v += a.pamp[k];
    }
    const double g = a.gain_lo + (a.gain_hi - a.gain_lo) *
                                     ((double)(mix64_d(a.seed ^ mix64_d((uint64_t)c ^
                                                                        0xA5A5A5A5DEADBEEFull)) >>
                                               11) *
                                      (1.0 / 9007199254740992.0));
    const double x = a.mean + a.sigma * g * v;
    const int64_t o = (int64_t)c * a.stride + t;
    if (a.dtype == MDSM_U8) {
        const double r = floor(x + 0.5);
        ((uint8_t *)out)[o] = (uint8_t)fmin(255.0, fmax(0.0, r));
    } else if (a.dtype == MDSM_F16) {
        ((_Float16 *)out)[o] = (_Float16)(float)x;
    } else {
        ((float *)out)[o] = (float)x;
    }
}

}  // namespace

extern "C" int mdsm_synth_filterbank_device(const mdsm_synth_spec *sp, void *d_out,
                                            int64_t out_stride, void *stream)
{
    if (!sp || !d_out) return fail(MDSM_ERR_INVALID, "NULL argument");
    if (sp->dtype < MDSM_U8 || sp->dtype > MDSM_F32) return fail(MDSM_ERR_INVALID, "bad dtype");
    if (out_stride < (int64_t)sp->nsamp) return fail(MDSM_ERR_INVALID, "out_stride < nsamp");
    if (sp->npulses > (uint32_t)kMaxSynthPulses)
        return fail(MDSM_ERR_INVALID, "at most 16 pulses on the device");
    if (sp->npulses && !sp->pulses) return fail(MDSM_ERR_INVALID, "pulses is NULL");
    if (!(sp->tsamp > 0.0f)) return fail(MDSM_ERR_INVALID, "tsamp must be > 0");
    if (sp->nchans > 65535) return fail(MDSM_ERR_INVALID, "nchans > 65535");
    if (sp->nchans == 0 || sp->nsamp == 0) return MDSM_OK;
    SynthArgs a{};
    // the host generator's seed derivation and per-pulse fields
    a.seed = sp->seed;
    a.seed += 0x9E3779B97F4A7C15ull;
    a.seed = (a.seed ^ (a.seed >> 30)) * 0xBF58476D1CE4E5B9ull;
    a.seed = (a.seed ^ (a.seed >> 27)) * 0x94D049BB133111EBull;
    a.seed = a.seed ^ (a.seed >> 31);
    a.dtype = sp->dtype;
    a.nchans = sp->nchans;
    a.nsamp = sp->nsamp;
    a.stride = out_stride;
    a.fch1 = sp->fch1;
    a.foff = sp->foff;
    a.tsamp = sp->tsamp;
    a.mean = sp->mean;
    a.sigma = sp->sigma;
    a.gain_lo = sp->gain_lo;
    a.gain_hi = sp->gain_hi;
    a.itop = 1.0 / ((double)sp->fch1 * (double)sp->fch1);
    a.npulses = sp->npulses;
    for (uint32_t k = 0; k < sp->npulses; ++k) {
        a.pdm[k] = sp->pulses[k].dm;
        a.pamp[k] = sp->pulses[k].amp;
        a.pt0[k] = sp->pulses[k].t0;
        a.pw[k] = sp->pulses[k].width;
    }
    hipLaunchKernelGGL(synth_kernel, dim3((sp->nsamp + 255) / 256, sp->nchans), dim3(256), 0,
                       (hipStream_t)stream, a, d_out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "synth launch");
    return MDSM_OK;
}
