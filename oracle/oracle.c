/*
 * oracle.c -- CPU restatement of MDSM's incoherent-dedispersion path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker).  Built with
 * -ffp-contract=off and without -ffast-math so that every float/double
 * operation below is one IEEE operation, in the order written.
 *
 * Reference paths are relative to the reference repository root.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Shift tables                                                        */
/* ------------------------------------------------------------------ */

/* Per-channel dispersion delay in seconds per unit DM.
 * Restates src/prototypes/cpu_dedisp.c:52-55 and
 * src/mdsm/dedispersion_manager.cpp:262-265: float arguments promoted to
 * double, the constant 4148.741601 s MHz^2, the result narrowed to float. */
float orc_dmdelay(float f1, float f2)
{
    double inv1 = 1.0 / (double)f1;
    inv1 = inv1 / (double)f1;
    double inv2 = 1.0 / (double)f2;
    inv2 = inv2 / (double)f2;
    double d = 4148.741601 * (inv1 - inv2);
    return (float)d;
}

/* delay[c] = dmdelay(fch1 + foff*c, fch1) with the channel frequency formed
 * in float (src/mdsm/dedispersion_manager.cpp:287-289,
 * src/prototypes/cpu_dedisp.c:72). */
void orc_delay_table(float fch1, float foff, int nchans, float *delay)
{
    for (int c = 0; c < nchans; c++) {
        float f = foff * (float)c;
        f = fch1 + f;
        delay[c] = orc_dmdelay(f, fch1);
    }
}

/* REF rule: shift[n][c] = (int)(delay[c] * (startdm + n*dmstep) / tsamp),
 * every operation in float, left to right, truncated toward zero
 * (src/prototypes/cpu_dedisp.c:67-73). */
void orc_shift_table_ref(const float *delay, int nchans, float startdm,
                         float dmstep, int ndms, float tsamp, int32_t *shift)
{
    for (int n = 0; n < ndms; n++) {
        float dm = (float)n * dmstep;
        dm = startdm + dm;
        for (int c = 0; c < nchans; c++) {
            float v = delay[c] * dm;
            v = v / tsamp;
            shift[(size_t)n * nchans + c] = (int32_t)v;
        }
    }
}

/* NS rule: FP64, constant 4.148808e3 s MHz^2, round to nearest sample.
 * Used for pulse injection and physics; no reference counterpart (the
 * reference only truncates). */
void orc_shift_table_ns(double fch1, double foff, int nchans, double startdm,
                        double dmstep, int ndms, double tsamp, int32_t *shift)
{
    double itop = 1.0 / (fch1 * fch1);
    for (int n = 0; n < ndms; n++) {
        double dm = startdm + (double)n * dmstep;
        for (int c = 0; c < nchans; c++) {
            double f = fch1 + foff * (double)c;
            double delay = 4.148808e3 * dm * (1.0 / (f * f) - itop);
            shift[(size_t)n * nchans + c] = (int32_t)llround(delay / tsamp);
        }
    }
}

/* ------------------------------------------------------------------ */
/* Element access                                                      */
/* ------------------------------------------------------------------ */

float orc_half_to_float(uint16_t h)
{
    uint32_t sign = (uint32_t)(h >> 15) << 31;
    uint32_t exp = (h >> 10) & 0x1f;
    uint32_t man = h & 0x3ff;
    uint32_t bits;
    if (exp == 0) {
        if (man == 0) {
            bits = sign;
        } else { /* subnormal: value = man * 2^-24, exact in float */
            float v = (float)man * (1.0f / 16777216.0f);
            memcpy(&bits, &v, 4);
            bits |= sign;
        }
    } else if (exp == 31) {
        bits = sign | 0x7f800000u | (man << 13);
    } else {
        bits = sign | ((exp + 112u) << 23) | (man << 13);
    }
    float out;
    memcpy(&out, &bits, 4);
    return out;
}

/* uint8 samples are their plain value (src/mdsm/file_handler.cpp:255-264). */
double orc_load(const void *in, int dtype, size_t idx)
{
    switch (dtype) {
    case ORC_U8:
        return (double)((const uint8_t *)in)[idx];
    case ORC_F16:
        return (double)orc_half_to_float(((const uint16_t *)in)[idx]);
    default:
        return (double)((const float *)in)[idx];
    }
}

static float load_f(const void *in, int dtype, size_t idx)
{
    switch (dtype) {
    case ORC_U8:
        return (float)((const uint8_t *)in)[idx];
    case ORC_F16:
        return orc_half_to_float(((const uint16_t *)in)[idx]);
    default:
        return ((const float *)in)[idx];
    }
}

/* ------------------------------------------------------------------ */
/* Brute force                                                         */
/* ------------------------------------------------------------------ */

/* The reference CPU dedisperser's loop, layouts as in the reference:
 * time-major input in[(s + shift)*nchans + c] and time-major output
 * out[s*ndms + n], float accumulation in channel order
 * (src/prototypes/cpu_dedisp.c:89-96). */
void orc_dedisperse_ref_tm(const float *in_tm, int nchans, int nsamp,
                           int ndms, const int32_t *shift, float *out_tm)
{
    for (int n = 0; n < ndms; n++) {
        const int32_t *sh = shift + (size_t)n * nchans;
        for (int s = 0; s < nsamp; s++) {
            float value = 0.0f;
            for (int c = 0; c < nchans; c++)
                value += in_tm[((size_t)s + sh[c]) * nchans + c];
            out_tm[(size_t)s * ndms + n] = value;
        }
    }
}

/* The same loop over (trial, 1024-sample run) pairs shared among nthreads
 * OpenMP threads; every output is the same channel-order float sum, so the
 * result equals orc_dedisperse_ref_tm bit for bit. */
void orc_dedisperse_ref_tm_mt(const float *in_tm, int nchans, int nsamp,
                              int ndms, const int32_t *shift, float *out_tm, int nthreads)
{
    const int runs = (nsamp + 1023) / 1024;
    const long jobs = (long)ndms * runs;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
    for (long q = 0; q < jobs; q++) {
        const int n = (int)(q / runs), s0 = (int)(q % runs) * 1024;
        const int s1 = s0 + 1024 < nsamp ? s0 + 1024 : nsamp;
        const int32_t *sh = shift + (size_t)n * nchans;
        for (int s = s0; s < s1; s++) {
            float value = 0.0f;
            for (int c = 0; c < nchans; c++)
                value += in_tm[((size_t)s + sh[c]) * nchans + c];
            out_tm[(size_t)s * ndms + n] = value;
        }
    }
}

/* Same sum with the channel-major input [c][S] and DM-major output [d][t]
 * that the GPU path uses (src/mdsm/dedispersion_kernel.cu:51-77 layout),
 * float accumulation in channel order as cpu_dedisp.c:92-93. */
void orc_dedisperse_f32(const void *in, int dtype, int nchans,
                        int64_t in_stride, int nout, int ndms,
                        const int32_t *shift, float *out, int64_t out_stride)
{
    for (int d = 0; d < ndms; d++) {
        const int32_t *sh = shift + (size_t)d * nchans;
        float *o = out + (size_t)d * out_stride;
        for (int t = 0; t < nout; t++) {
            float v = 0.0f;
            for (int c = 0; c < nchans; c++)
                v += load_f(in, dtype, (size_t)c * in_stride + t + sh[c]);
            o[t] = v;
        }
    }
}

void orc_dedisperse_f64(const void *in, int dtype, int nchans,
                        int64_t in_stride, int nout, int ndms,
                        const int32_t *shift, double *out, int64_t out_stride)
{
    for (int d = 0; d < ndms; d++) {
        const int32_t *sh = shift + (size_t)d * nchans;
        double *o = out + (size_t)d * out_stride;
        for (int t = 0; t < nout; t++) {
            double v = 0.0;
            for (int c = 0; c < nchans; c++)
                v += orc_load(in, dtype, (size_t)c * in_stride + t + sh[c]);
            o[t] = v;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Decimation and multi-range                                          */
/* ------------------------------------------------------------------ */

/* Time decimation by k: out[c][t'] = sum_{j<k} x[c][k t' + j], t' < S/k.
 * The reference divides by k (src/mdsm/dedispersion_kernel.cu:203-246);
 * here the sum is kept and the division by k is applied once to each
 * dedispersed value (orc_range_*), which keeps integer data exact. */
void orc_decimate_f64(const void *in, int dtype, int nchans, int64_t in_stride,
                      int nsamp_in, int k, double *out)
{
    int sk = nsamp_in / k;
    for (int c = 0; c < nchans; c++)
        for (int t = 0; t < sk; t++) {
            double v = 0.0;
            for (int j = 0; j < k; j++)
                v += orc_load(in, dtype, (size_t)c * in_stride + (size_t)k * t + j);
            out[(size_t)c * sk + t] = v;
        }
}

/* One DM range of a multi-range plan (src/mdsm/dedispersion_thread.cu:
 * 159-171,197-205 pattern): dedisperse the k-decimated block with the
 * range's shift table (built for tsamp*k) and scale by 1/k.
 * Float order: k-sum (j ascending), channel sum (c ascending), one
 * division by k.  For uint8 input every intermediate is an exact integer. */
void orc_range_f32(const void *in, int dtype, int nchans, int64_t in_stride,
                   int nsamp_in, int k, int nout, int ndms,
                   const int32_t *shift, float *out, int64_t out_stride)
{
    (void)nsamp_in;
    for (int d = 0; d < ndms; d++) {
        const int32_t *sh = shift + (size_t)d * nchans;
        float *o = out + (size_t)d * out_stride;
        for (int t = 0; t < nout; t++) {
            float v = 0.0f;
            for (int c = 0; c < nchans; c++) {
                size_t base = (size_t)c * in_stride + (size_t)k * ((size_t)t + sh[c]);
                float s = 0.0f;
                for (int j = 0; j < k; j++)
                    s += load_f(in, dtype, base + j);
                v += s;
            }
            o[t] = (k > 1) ? v / (float)k : v;
        }
    }
}

void orc_range_f64(const void *in, int dtype, int nchans, int64_t in_stride,
                   int nsamp_in, int k, int nout, int ndms,
                   const int32_t *shift, double *out, int64_t out_stride)
{
    int sk = nsamp_in / k;
    double *dec = (double *)malloc((size_t)nchans * sk * sizeof(double));
    orc_decimate_f64(in, dtype, nchans, in_stride, nsamp_in, k, dec);
    for (int d = 0; d < ndms; d++) {
        const int32_t *sh = shift + (size_t)d * nchans;
        double *o = out + (size_t)d * out_stride;
        for (int t = 0; t < nout; t++) {
            double v = 0.0;
            for (int c = 0; c < nchans; c++)
                v += dec[(size_t)c * sk + t + sh[c]];
            o[t] = v / (double)k;
        }
    }
    free(dec);
}

/* ------------------------------------------------------------------ */
/* Two-stage subband                                                   */
/* ------------------------------------------------------------------ */

/* Shift tables of one subband pass.
 *   coarse[j][c]: REF shift of channel c at coarse DM lowdm + j*sub_dmstep
 *     (stage 1, src/mdsm/dedispersion_kernel.cu:144-175 with
 *      startdm/dmstep of src/mdsm/dedispersion_thread.cu:159-168);
 *   fine[j*calldms + i][b]: REF shift of the top channel (b*C/nsubs) of
 *     subband b at fine DM start_j + i*dmstep, where
 *     start_j = max(lowdm + j*sub_dmstep - calldms*dmstep/2, 0)
 *     (src/mdsm/dedispersion_thread.cu:195-205, kernel :51-77 with
 *      chanfactor = C/nsubs).
 * Float association of the REF rule (cpu_dedisp.c:72) is used throughout. */
void orc_subband_tables(const float *delay, int nchans, int nsubs,
                        float lowdm, float sub_dmstep, float dmstep,
                        int ncalls, int calldms, float tsamp_k,
                        int32_t *coarse, int32_t *fine)
{
    int cps = nchans / nsubs;
    for (int j = 0; j < ncalls; j++) {
        float cdm = (float)j * sub_dmstep;
        cdm = lowdm + cdm;
        for (int c = 0; c < nchans; c++) {
            float v = delay[c] * cdm;
            v = v / tsamp_k;
            coarse[(size_t)j * nchans + c] = (int32_t)v;
        }
        float half = (float)calldms * dmstep;
        half = half / 2.0f;
        float start = cdm - half;
        if (!(start > 0.0f))
            start = 0.0f;
        for (int i = 0; i < calldms; i++) {
            float fdm = (float)i * dmstep;
            fdm = start + fdm;
            for (int b = 0; b < nsubs; b++) {
                float v = delay[b * cps] * fdm;
                v = v / tsamp_k;
                fine[((size_t)j * calldms + i) * nsubs + b] = (int32_t)v;
            }
        }
    }
}

/* Two-stage subband dedispersion of one pass, FP64.
 * Stage 1: sb[j][b][t] = sum_{c in b} x_k[c][t + coarse[j][c] - coarse[j][top(b)]]
 *          for t < len1.
 * Stage 2: out[j*calldms+i][t] = (1/k) sum_b sb[j][b][t + fine[j*calldms+i][b]]
 *          for t < nout. */
void orc_subband_f64(const void *in, int dtype, int nchans, int64_t in_stride,
                     int nsamp_in, int k, int nsubs, int ncalls, int calldms,
                     const int32_t *coarse, const int32_t *fine, int len1,
                     int nout, double *out, int64_t out_stride)
{
    int sk = nsamp_in / k;
    int cps = nchans / nsubs;
    double *dec = (double *)malloc((size_t)nchans * sk * sizeof(double));
    double *sb = (double *)malloc((size_t)nsubs * len1 * sizeof(double));
    orc_decimate_f64(in, dtype, nchans, in_stride, nsamp_in, k, dec);
    for (int j = 0; j < ncalls; j++) {
        const int32_t *co = coarse + (size_t)j * nchans;
        for (int b = 0; b < nsubs; b++) {
            int top = co[b * cps];
            for (int t = 0; t < len1; t++) {
                double v = 0.0;
                for (int c = b * cps; c < (b + 1) * cps; c++)
                    v += dec[(size_t)c * sk + t + (co[c] - top)];
                sb[(size_t)b * len1 + t] = v;
            }
        }
        for (int i = 0; i < calldms; i++) {
            size_t d = (size_t)j * calldms + i;
            const int32_t *fi = fine + d * nsubs;
            double *o = out + d * out_stride;
            for (int t = 0; t < nout; t++) {
                double v = 0.0;
                for (int b = 0; b < nsubs; b++)
                    v += sb[(size_t)b * len1 + t + fi[b]];
                o[t] = v / (double)k;
            }
        }
    }
    free(sb);
    free(dec);
}

/* ------------------------------------------------------------------ */
/* RFI clipper (uint8)                                                 */
/* ------------------------------------------------------------------ */

static int rank_from_hist(const int64_t *hist, int nbins, int64_t rank)
{
    int64_t acc = 0;
    for (int v = 0; v < nbins; v++) {
        acc += hist[v];
        if (acc > rank)
            return v;
    }
    return nbins - 1;
}

/* Per-channel robust clip over samples [t0, t0+n) of each channel:
 *   med = lower median (element of rank (n-1)/2 in ascending order),
 *   mad = lower median of |x - med|,
 *   thr = (nsigma * 1.4826f) * (float)mad   (float arithmetic),
 *   x -> med where (float)|x - med| > thr.
 * The reference's prototype clipper (src/prototypes/rfi_clipper.cu:8-67)
 * uses mean/sigma and replaces outliers with the centre; the median/MAD
 * centre and scale here are this implementation's definition. */
void orc_clip_u8(const uint8_t *in, int nchans, int64_t in_stride, int64_t t0,
                 int n, float nsigma, uint8_t *out, int64_t out_stride,
                 int32_t *med, int32_t *mad, int32_t *nclipped)
{
    int64_t rank = (n - 1) / 2;
    float scale = nsigma * 1.4826f;
    for (int c = 0; c < nchans; c++) {
        const uint8_t *x = in + (size_t)c * in_stride + t0;
        int64_t hist[256] = {0}, dev[256] = {0};
        for (int t = 0; t < n; t++)
            hist[x[t]]++;
        int m = rank_from_hist(hist, 256, rank);
        for (int t = 0; t < n; t++)
            dev[abs((int)x[t] - m)]++;
        int a = rank_from_hist(dev, 256, rank);
        float thr = scale * (float)a;
        int32_t clipped = 0;
        uint8_t *o = out + (size_t)c * out_stride;
        for (int t = 0; t < n; t++) {
            int dv = abs((int)x[t] - m);
            if ((float)dv > thr) {
                o[t] = (uint8_t)m;
                clipped++;
            } else {
                o[t] = x[t];
            }
        }
        med[c] = m;
        mad[c] = a;
        nclipped[c] = clipped;
    }
}

/* ------------------------------------------------------------------ */
/* Boxcar single-pulse detector                                        */
/* ------------------------------------------------------------------ */

/* Row statistics: mean = S1/n, sigma = sqrt(S2/n - mean*mean), with
 * S1 = sum x and S2 = sum x*x accumulated in double in index order. */
void orc_row_stats(const float *row, int n, double *mean, double *sigma)
{
    double s1 = 0.0, s2 = 0.0;
    for (int t = 0; t < n; t++) {
        double v = (double)row[t];
        s1 += v;
        double v2 = v * v;
        s2 += v2;
    }
    double m = s1 / (double)n;
    double ex2 = s2 / (double)n;
    double mm = m * m;
    double var = ex2 - mm;
    if (var < 0.0)
        var = 0.0;
    *mean = m;
    *sigma = sqrt(var);
}

/* Boxcar search over widths w (powers of two).  Boxcar sums are formed as
 * a pairwise tree in double: S_1(t) = x[t], S_2w(t) = S_w(t) + S_w(t+w).
 *   snr_w(t) = (S_w(t) - w*mean) / (sigma * sqrt(w)),  0 <= t <= n-w.
 * A candidate (d, t, w) is emitted when snr_w(t) > thr and S_w has a peak
 * at t: (t == 0 or S_w(t) >= S_w(t-1)) and (t == n-w or S_w(t) > S_w(t+1)).
 * Order: d ascending, then w in the order given, then t ascending.
 * The reference's boxcar prototype (src/prototypes/singlepulse.cpp:41-146)
 * convolves centred 1/sqrt(w) kernels via FFTW; this is the direct sum.
 * Returns the number of candidates found (stores at most cap). */
int orc_boxcar_detect(const float *x, int ndms, int n, int64_t stride,
                      const int *widths, int nwidths, double thr,
                      orc_cand *cands, int cap, double *means, double *sigmas)
{
    int maxw = 1;
    for (int i = 0; i < nwidths; i++)
        if (widths[i] > maxw)
            maxw = widths[i];
    double *cur = (double *)malloc((size_t)n * sizeof(double));
    double *nxt = (double *)malloc((size_t)n * sizeof(double));
    double *byw = (double *)malloc((size_t)n * sizeof(double));
    int count = 0;
    for (int d = 0; d < ndms; d++) {
        const float *row = x + (size_t)d * stride;
        double mean, sigma;
        orc_row_stats(row, n, &mean, &sigma);
        if (means)
            means[d] = mean;
        if (sigmas)
            sigmas[d] = sigma;
        for (int wi = 0; wi < nwidths; wi++) {
            int w = widths[wi];
            if (w > n || sigma <= 0.0)
                continue;
            /* build S_w by doubling from S_1 */
            for (int t = 0; t < n; t++)
                cur[t] = (double)row[t];
            int cw = 1;
            while (cw < w) {
                for (int t = 0; t + 2 * cw <= n; t++)
                    nxt[t] = cur[t] + cur[t + cw];
                double *tmp = cur;
                cur = nxt;
                nxt = tmp;
                cw *= 2;
            }
            memcpy(byw, cur, (size_t)(n - w + 1) * sizeof(double));
            double wm = (double)w * mean;
            double den = sigma * sqrt((double)w);
            int last = n - w;
            for (int t = 0; t <= last; t++) {
                double s = byw[t];
                double num = s - wm;
                double snr = num / den;
                if (!(snr > thr))
                    continue;
                if (t > 0 && !(s >= byw[t - 1]))
                    continue;
                if (t < last && !(s > byw[t + 1]))
                    continue;
                if (count < cap) {
                    cands[count].dm = (uint32_t)d;
                    cands[count].t = (int32_t)t;
                    cands[count].width = (uint32_t)w;
                    cands[count].snr = (float)snr;
                }
                count++;
            }
        }
    }
    free(cur);
    free(nxt);
    free(byw);
    (void)maxw;
    return count;
}

/* Streaming boxcar search over the concatenation X of nblk blocks of n
 * samples per row.  Statistics are per (row, block): mean_k and sigma_k of
 * X[k n, (k+1) n) as orc_row_stats forms them.  Window (t, w) ends at
 * e = t + w - 1 and is owned by block k = (e + 1) / n, i.e. by the block
 * holding e + 1, the last sample its right neighbour S_w(t+1) reads; it is
 * searched when that neighbour exists (t + 1 + w <= nblk n), with block k's
 * statistics:  snr = (S_w(t) - w mean_k) / (sigma_k sqrt(w)) > thr and
 * (t == 0 or S_w(t) >= S_w(t-1)) and S_w(t) > S_w(t+1).  Boxcar sums are the
 * pairwise tree of orc_boxcar_detect over X.  So every window is searched
 * exactly once, the ones straddling a block edge included, and a stream
 * searched block by block with the previous block's last samples carried
 * over finds the same candidates.  No counterpart in the reference, whose
 * detectors work block by block (src/mdsm/dedispersion_output.cpp:176-181;
 * src/prototypes/singlepulse.cpp:41-146 convolves each 8192-sample chunk
 * on its own). */
int orc_boxcar_stream(const float *x, int ndms, int n, int nblk, int64_t stride,
                      const int *widths, int nwidths, double thr, orc_scand *cands, int cap)
{
    const int64_t T = (int64_t)n * nblk;
    double *cur = (double *)malloc((size_t)T * sizeof(double));
    double *nxt = (double *)malloc((size_t)T * sizeof(double));
    double *mean = (double *)malloc((size_t)nblk * sizeof(double));
    double *sig = (double *)malloc((size_t)nblk * sizeof(double));
    int count = 0;
    for (int d = 0; d < ndms; d++) {
        const float *row = x + (size_t)d * stride;
        for (int k = 0; k < nblk; k++)
            orc_row_stats(row + (int64_t)k * n, n, &mean[k], &sig[k]);
        for (int wi = 0; wi < nwidths; wi++) {
            const int w = widths[wi];
            if ((int64_t)w + 1 > T)
                continue;
            for (int64_t t = 0; t < T; t++)
                cur[t] = (double)row[t];
            int cw = 1;
            while (cw < w) {
                for (int64_t t = 0; t + 2 * cw <= T; t++)
                    nxt[t] = cur[t] + cur[t + cw];
                double *tmp = cur;
                cur = nxt;
                nxt = tmp;
                cw *= 2;
            }
            const double sw = sqrt((double)w);
            for (int64_t t = 0; t + 1 + w <= T; t++) {
                const int64_t k = (t + w) / n; /* block of e + 1 */
                if (!(sig[k] > 0.0))
                    continue;
                const double s = cur[t];
                const double snr = (s - (double)w * mean[k]) / (sig[k] * sw);
                if (!(snr > thr))
                    continue;
                if (t > 0 && !(s >= cur[t - 1]))
                    continue;
                if (!(s > cur[t + 1]))
                    continue;
                if (count < cap) {
                    cands[count].dm = (uint32_t)d;
                    cands[count].t = (int32_t)(t - k * n);
                    cands[count].width = (uint32_t)w;
                    cands[count].snr = (float)snr;
                    cands[count].blk = (uint32_t)k;
                }
                count++;
            }
        }
    }
    free(cur);
    free(nxt);
    free(mean);
    free(sig);
    return count;
}

/* ------------------------------------------------------------------ */
/* Reference in-path detector                                          */
/* ------------------------------------------------------------------ */

/* Block statistics of the reference's brute-force output thread when no
 * noise arrays are given (src/mdsm/dedispersion_output.cpp:143-167): the
 * ndms x n values, rows read one after another as one series, are taken in
 * chunks of mf = min(vals, 32768); each chunk's sum and sum of squares are
 * accumulated in double in index order; the chunk means are added to a
 * float running mean (each addition formed in double, then rounded to
 * float) and the mean squares to a double; the loop stops before the chunk
 * that would reach the end of the values, so the last whole chunk and any
 * partial one are never read (iters = max(1, ceil(vals/mf) - 1)).  mean =
 * float(mean_sum / iters), rms = float(sqrt(msq_sum/iters - mean*mean)) with
 * mean*mean a float product. */
void orc_brute_stats(const float *x, int ndms, int n, int64_t stride, float *mean, float *rms)
{
    const uint64_t vals = (uint64_t)ndms * (uint64_t)n;
    const uint64_t mf = vals < 32768 ? vals : 32768;
    float localmean = 0.0f;
    double mean2 = 0.0;
    unsigned iters = 0;
    for (;;) {
        double total = 0.0, total2 = 0.0;
        for (uint64_t j = 0; j < mf; j++) {
            const uint64_t f = (uint64_t)iters * mf + j;
            const double v = (double)x[(f / (uint64_t)n) * (uint64_t)stride + f % (uint64_t)n];
            total += v;
            total2 += v * v;
        }
        localmean = (float)((double)localmean + total / (double)mf);
        mean2 += total2 / (double)mf;
        iters++;
        if ((uint64_t)iters * mf + mf >= vals)
            break;
    }
    localmean = localmean / (float)iters;
    const float lm2 = localmean * localmean;
    *mean = localmean;
    *rms = (float)sqrt(mean2 / (double)iters - (double)lm2);
}

/* Median of five values (src/mdsm/dedispersion_output.cpp:14-53), stated
 * here as an insertion sort and the middle element. */
float orc_median_of_five(float a, float b, float c, float d, float e)
{
    float v[5] = {a, b, c, d, e};
    for (int i = 1; i < 5; i++) {
        float key = v[i];
        int j = i - 1;
        while (j >= 0 && v[j] > key) {
            v[j + 1] = v[j];
            j--;
        }
        v[j + 1] = key;
    }
    return v[2];
}

/* ------------------------------------------------------------------ */
/* Input corner turn                                                   */
/* ------------------------------------------------------------------ */

/* out[c][t] = in[t][invert ? C-1-c : c]: the copy loop of
 * src/pelican-mdsm/modules/MdsmModule.cpp:135-151 (subband and channel
 * indices reversed together when invertChannels is set, :137,145) and
 * src/MDSM.cpp:100-120. */
void orc_corner_turn_f32(const float *in, int nsamp, int nchans, int64_t in_stride,
                         int invert, float *out, int64_t out_stride)
{
    for (int t = 0; t < nsamp; t++)
        for (int c = 0; c < nchans; c++) {
            int src = invert ? nchans - 1 - c : c;
            out[(int64_t)c * out_stride + t] = in[(int64_t)t * in_stride + src];
        }
}

/* ------------------------------------------------------------------ */
/* Detrend and normalise                                               */
/* ------------------------------------------------------------------ */

/* Per row, per segment of seg_len samples (a shorter last segment uses its
 * own length L): x_i = -L/2 + 0.5 + i, a = sum(y)/L, b = sum(x y)/sum(x x),
 * r_i = y_i - (a + b x_i), out_i = r_i / sqrt(sum(r^2)/L)
 * (src/prototypes/detrend.cu:19-31,52-56,62-75,82-96), in FP64.  A zero
 * residual RMS gives zeros (the reference divides by it). */
void orc_detrend_f64(const float *x, int nrows, int n, int64_t stride, int seg_len,
                     double *out)
{
    for (int r = 0; r < nrows; r++) {
        const float *row = x + (int64_t)r * stride;
        double *o = out + (int64_t)r * n;
        for (int s0 = 0; s0 < n; s0 += seg_len) {
            int L = n - s0 < seg_len ? n - s0 : seg_len;
            double sy = 0.0, sxy = 0.0, sxx = 0.0;
            for (int i = 0; i < L; i++) {
                double xi = -(double)L / 2.0 + 0.5 + (double)i;
                double y = row[s0 + i];
                sy += y;
                sxy += xi * y;
                sxx += xi * xi;
            }
            double a = sy / (double)L;
            double b = sxx > 0.0 ? sxy / sxx : 0.0;
            double ss = 0.0;
            for (int i = 0; i < L; i++) {
                double xi = -(double)L / 2.0 + 0.5 + (double)i;
                double v = (double)row[s0 + i] - (a + b * xi);
                o[s0 + i] = v;
                ss += v * v;
            }
            double rms = sqrt(ss / (double)L);
            for (int i = 0; i < L; i++) o[s0 + i] = rms > 0.0 ? o[s0 + i] / rms : 0.0;
        }
    }
}

/* ------------------------------------------------------------------ */
/* SIGPROC sample unpacking                                            */
/* ------------------------------------------------------------------ */

/* read_block's unpack (src/mdsm/file_handler.cpp:203-285): 1-bit samples
 * least significant bit first (:211-214), 2-bit with char4ints (:18-24,
 * :225-231), 4-bit with char2ints (low nibble first, :11-16, :245-249),
 * 8- and 16-bit unsigned (:257-272), 32-bit float (:276-278). */
void orc_unpack_bits(const uint8_t *raw, int nbits, int64_t nvals, float *out)
{
    for (int64_t i = 0; i < nvals; i++) {
        switch (nbits) {
        case 1: out[i] = (float)((raw[i / 8] >> (i % 8)) & 1); break;
        case 2: out[i] = (float)((raw[i / 4] >> (2 * (i % 4))) & 3); break;
        case 4: out[i] = (float)((raw[i / 2] >> (4 * (i % 2))) & 15); break;
        case 8: out[i] = (float)raw[i]; break;
        case 16: {
            uint16_t v;
            memcpy(&v, raw + 2 * i, 2);
            out[i] = (float)v;
            break;
        }
        default: memcpy(&out[i], raw + 4 * i, 4); break;
        }
    }
}
