/*
 * oracle.h -- CPU restatement of MDSM's incoherent-dedispersion path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * gfx950 implementation in mdsm_amd/.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product path never does.
 *
 * Every function cites the reference file:line whose algorithm it restates
 * (paths relative to the reference repository root).
 *
 * Pinning (how each function is checked independently of this restatement):
 *   orc_dmdelay / orc_delay_table / orc_shift_table_ref
 *       bit-exact against the reference's own cpu_dedisp.c main() built
 *       under oracle/_ref (its shift rows are captured through a recording
 *       allocator), against committed golden vectors, and against the
 *       disagreement counts of independent FP64 rules.
 *   orc_shift_table_ns
 *       closed form in FP64; impulse placement checked end to end.
 *   orc_dedisperse_ref_tm / orc_dedisperse_f32
 *       bit-exact against the reference's own cpu_dedisp.c loop (oracle/_ref)
 *       on seeded data; constant-per-channel known answer
 *       (src/prototypes/brute_dedisp_imp_tex.cu:112-114,167-175).
 *   orc_dedisperse_f64
 *       impulse response, DM-0 channel sum, linearity, integer exactness.
 *   orc_decimate_f64 / orc_range_f32
 *       closed form (k-sample sums), reduces to brute force at k = 1.
 *   orc_subband_tables / orc_subband_f64
 *       reduces to brute force with one channel per subband; integer
 *       exactness; hand-computed tiny case.
 *   orc_clip_u8
 *       against numpy sort-based median/MAD on tiny cases (ties, even and
 *       odd lengths, constant channels).
 *   orc_boxcar_detect
 *       closed forms (spike: snr_w = h/sqrt(w) pattern; top hat), brute
 *       force numpy on tiny inputs.
 *   orc_boxcar_stream
 *       with one block it is orc_boxcar_detect less the windows without a
 *       right neighbour; numpy brute force over tiny multi-block series
 *       (per-block statistics, window ownership by end sample); a pulse
 *       straddling a block edge is found once, at its concatenated S/N.
 *   orc_median_of_five
 *       all 5^5 tuples against the reference's own medianOfFive built under
 *       oracle/_ref and against sort.
 *   orc_brute_stats
 *       the candidate file of the reference's own process_brute (oracle/_ref)
 *       on multi-chunk blocks, line for line, and a committed golden vector
 *       of its statistics.
 *   orc_corner_turn_f32
 *       numpy transpose (and channel reversal) of the same block.
 *   orc_detrend_f64
 *       numpy least-squares line fit (polyfit) on tiny inputs; invariants
 *       of the output (zero mean, zero slope, unit RMS per segment); a pure
 *       line detrends to zero.
 *   orc_unpack_bits
 *       the reference's own read_block (oracle/_ref) on the same file bytes;
 *       numpy unpackbits (little bit order) for 1-, 2- and 4-bit samples.
 */
#ifndef MDSM_ORACLE_H
#define MDSM_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_U8 = 0, ORC_F16 = 1, ORC_F32 = 2 };

/* --- shift tables --------------------------------------------------- */
float orc_dmdelay(float f1, float f2);
void orc_delay_table(float fch1, float foff, int nchans, float *delay);
void orc_shift_table_ref(const float *delay, int nchans, float startdm,
                         float dmstep, int ndms, float tsamp, int32_t *shift);
void orc_shift_table_ns(double fch1, double foff, int nchans, double startdm,
                        double dmstep, int ndms, double tsamp, int32_t *shift);

/* --- element access -------------------------------------------------- */
float orc_half_to_float(uint16_t h);
double orc_load(const void *in, int dtype, size_t idx);

/* --- brute force ----------------------------------------------------- */
void orc_dedisperse_ref_tm(const float *in_tm, int nchans, int nsamp,
                           int ndms, const int32_t *shift, float *out_tm);
void orc_dedisperse_ref_tm_mt(const float *in_tm, int nchans, int nsamp,
                              int ndms, const int32_t *shift, float *out_tm, int nthreads);
void orc_dedisperse_f32(const void *in, int dtype, int nchans,
                        int64_t in_stride, int nout, int ndms,
                        const int32_t *shift, float *out, int64_t out_stride);
void orc_dedisperse_f64(const void *in, int dtype, int nchans,
                        int64_t in_stride, int nout, int ndms,
                        const int32_t *shift, double *out, int64_t out_stride);

/* --- decimation and multi-range -------------------------------------- */
void orc_decimate_f64(const void *in, int dtype, int nchans, int64_t in_stride,
                      int nsamp_in, int k, double *out);
void orc_range_f32(const void *in, int dtype, int nchans, int64_t in_stride,
                   int nsamp_in, int k, int nout, int ndms,
                   const int32_t *shift, float *out, int64_t out_stride);
void orc_range_f64(const void *in, int dtype, int nchans, int64_t in_stride,
                   int nsamp_in, int k, int nout, int ndms,
                   const int32_t *shift, double *out, int64_t out_stride);

/* --- two-stage subband ----------------------------------------------- */
void orc_subband_tables(const float *delay, int nchans, int nsubs,
                        float lowdm, float sub_dmstep, float dmstep,
                        int ncalls, int calldms, float tsamp_k,
                        int32_t *coarse, int32_t *fine);
void orc_subband_f64(const void *in, int dtype, int nchans, int64_t in_stride,
                     int nsamp_in, int k, int nsubs, int ncalls, int calldms,
                     const int32_t *coarse, const int32_t *fine, int len1,
                     int nout, double *out, int64_t out_stride);

/* --- RFI clipper ----------------------------------------------------- */
void orc_clip_u8(const uint8_t *in, int nchans, int64_t in_stride, int64_t t0,
                 int n, float nsigma, uint8_t *out, int64_t out_stride,
                 int32_t *med, int32_t *mad, int32_t *nclipped);

/* --- boxcar single-pulse detector ------------------------------------ */
typedef struct {
    uint32_t dm;
    int32_t t;
    uint32_t width;
    float snr;
} orc_cand;

void orc_row_stats(const float *row, int n, double *mean, double *sigma);
int orc_boxcar_detect(const float *x, int ndms, int n, int64_t stride,
                      const int *widths, int nwidths, double thr,
                      orc_cand *cands, int cap, double *means, double *sigmas);

/* Streaming form over the concatenation of nblk blocks of n samples per row
 * (row d: x[d*stride + 0 .. nblk*n)): every window is searched once, by the
 * block its end falls in, with that block's statistics.  Candidate t is the
 * window start relative to the first sample of that block (negative: the
 * window starts in the previous block); blk receives the block index. */
typedef struct {
    uint32_t dm;
    int32_t t;
    uint32_t width;
    float snr;
    uint32_t blk;
} orc_scand;
int orc_boxcar_stream(const float *x, int ndms, int n, int nblk, int64_t stride,
                      const int *widths, int nwidths, double thr, orc_scand *cands, int cap);

/* --- reference in-path detector (median of five) --------------------- */
float orc_median_of_five(float a, float b, float c, float d, float e);
/* Whole-block statistics of the brute-force output thread without noise
 * arrays (chunks of 32768 values; src/mdsm/dedispersion_output.cpp:143-167). */
void orc_brute_stats(const float *x, int ndms, int n, int64_t stride, float *mean, float *rms);

/* Corner turn of a time-major float block (MdsmModule.cpp:135-151). */
void orc_corner_turn_f32(const float *in, int nsamp, int nchans, int64_t in_stride,
                         int invert, float *out, int64_t out_stride);

/* Unpack nvals SIGPROC samples of nbits (1, 2, 4, 8, 16, 32) into floats
 * (src/mdsm/file_handler.cpp:193-290). */
void orc_unpack_bits(const uint8_t *raw, int nbits, int64_t nvals, float *out);

/* Detrend and normalise, FP64 (src/prototypes/detrend.cu:12-100). */
void orc_detrend_f64(const float *x, int nrows, int n, int64_t stride, int seg_len,
                     double *out);

#ifdef __cplusplus
}
#endif
#endif
