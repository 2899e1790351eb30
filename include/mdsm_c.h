/*
 * mdsm_c.h -- C ABI of the MI355X-native MDSM dedispersion path
 * (libmdsm_hip.so, gfx950).
 *
 * Plain C types only: plain pointers, sizes and status codes.  Device
 * buffers are caller-owned device pointers (hipMalloc or any allocator on
 * the plan's device); `stream` is a hipStream_t passed as void* (NULL =
 * the null stream).  No call exits the process; every entry point returns
 * an mdsm_status and mdsm_last_error() holds the message of the most recent
 * failure on the calling thread.
 *
 * Each entry point cites the reference interface it replaces (paths
 * relative to the reference repository root).
 */
#ifndef MDSM_C_H
#define MDSM_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Status, types                                                       */
/* ------------------------------------------------------------------ */

/* Replaces the reference's fprintf(stderr)+exit(0) error convention
 * (src/mdsm/dedispersion_thread.cu:11-12,58-59,329-330). */
typedef enum {
    MDSM_OK = 0,
    MDSM_ERR_INVALID = -1,  /* malformed plan, argument or size */
    MDSM_ERR_NOMEM = -2,    /* host or device allocation failed */
    MDSM_ERR_DEVICE = -3,   /* HIP runtime error */
    MDSM_ERR_CAPACITY = -4, /* candidate buffer too small (count still valid) */
    MDSM_ERR_BUDGET = -5,   /* request can never fit the HBM budget */
    MDSM_ERR_STATE = -6,    /* call out of order / object closed */
    MDSM_ERR_PARSE = -7,    /* observation XML could not be parsed */
    MDSM_ERR_AGAIN = -8,    /* nothing ready yet / admission deferred */
    MDSM_ERR_CANCELLED = -9 /* block dropped by mdsm_beam_cancel before it ran */
} mdsm_status;

/* Sample storage of the filterbank.  uint8 is the plain byte value
 * (src/mdsm/file_handler.cpp:255-264); every path accumulates in float32. */
typedef enum { MDSM_U8 = 0, MDSM_F16 = 1, MDSM_F32 = 2 } mdsm_dtype;

/* Integer shift rule.
 *   MDSM_SHIFT_REF: shift = (int)(delay_c * dm / tsamp) in float32, with
 *     delay_c = (float)(4148.741601 * (f_c^-2 - f_top^-2)), f_c = fch1 + foff*c
 *     formed in float -- bit-exact with src/prototypes/cpu_dedisp.c:52-55,72
 *     and src/mdsm/dedispersion_manager.cpp:262-265,287-289.
 *   MDSM_SHIFT_NS: shift = llround(4.148808e3 * dm * (f_c^-2 - f_top^-2) / tsamp)
 *     in FP64 (physical delay, nearest sample). */
typedef enum { MDSM_SHIFT_REF = 0, MDSM_SHIFT_NS = 1 } mdsm_shift_rule;

/* Observation of one beam: the subset of SURVEY (src/mdsm/survey.h:13-57)
 * the dedispersion path reads. */
typedef struct {
    uint32_t nchans;   /* C; channel 0 is the highest frequency */
    float fch1;        /* MHz, channel 0 */
    float foff;        /* MHz per channel (negative) */
    float tsamp;       /* seconds */
    uint32_t nsamp_in; /* S: samples per channel per block, overlap included */
    int32_t dtype;     /* mdsm_dtype */
    int32_t shift_rule;/* mdsm_shift_rule */
} mdsm_obs;

/* One brute-force DM range of a multi-range plan.  With downsample k the
 * range runs on the k-sample-decimated block at tsamp*k and its output is
 * the channel sum of k-sample means (exact for uint8 input).  Replaces the
 * brute-force fields of SURVEY (lowdm, dmstep, tdms; survey.h:23-24) and
 * the per-pass downsample of SUBBAND_PASSES (survey.h:6-11). */
typedef struct {
    float lowdm;
    float dmstep;
    uint32_t ndms;
    uint32_t downsample;
} mdsm_range;

/* Two-stage subband pass, field for field SUBBAND_PASSES
 * (src/mdsm/survey.h:6-11; parsed at dedispersion_manager.cpp:158-184). */
typedef struct {
    float lowdm, highdm, dmstep, sub_dmstep;
    uint32_t binsize, ndms, ncalls, calldms;
} mdsm_pass;

typedef struct mdsm_plan mdsm_plan;

/* ------------------------------------------------------------------ */
/* Plans                                                               */
/* ------------------------------------------------------------------ */

/* Build a brute-force multi-range plan on the current HIP device: shift
 * tables, maxshift, output layout and the device work tables.
 * Replaces the plan set-up half of initialiseMDSM
 * (src/mdsm/dedispersion_manager.cpp:286-314) and the shift/maxshift code of
 * src/prototypes/cpu_dedisp.c:67-76. */
int mdsm_plan_create(const mdsm_obs *obs, const mdsm_range *ranges,
                     uint32_t nranges, mdsm_plan **out);

/* Build a two-stage subband plan (nsubs subbands, one or more passes).
 * Replaces the subband set-up of initialiseMDSM and
 * src/mdsm/dedispersion_thread.cu:150-210. */
int mdsm_plan_create_subband(const mdsm_obs *obs, uint32_t nsubs,
                             const mdsm_pass *passes, uint32_t npasses,
                             mdsm_plan **out);

void mdsm_plan_destroy(mdsm_plan *plan);

/* Upload the plan's device tables now (set-up) instead of on the first
 * mdsm_dedisperse.  Required before mdsm_dedisperse_noalloc.  Part of the
 * plan set-up of initialiseMDSM (src/mdsm/dedispersion_manager.cpp:286-314). */
int mdsm_plan_upload(mdsm_plan *plan);

typedef struct {
    uint32_t nranges;        /* ranges (brute) or passes (subband) */
    uint32_t ndms_total;     /* rows of the packed output */
    uint32_t maxshift;       /* full-rate overlap samples the plan needs */
    uint32_t nchans;
    uint64_t out_elems;      /* floats in the packed output */
    uint64_t work_bytes;     /* device workspace (decimated copies, stage 1) */
    double dm_chan_samples;  /* sum_r D_r*C*N_r (brute-force equivalent) */
    double adds;             /* float adds the plan performs */
    double bytes;            /* algorithmic HBM bytes per block */
    uint32_t is_subband;
    uint32_t nblocks;        /* DM blocks in the device work table */
} mdsm_plan_info;

int mdsm_plan_get_info(const mdsm_plan *plan, mdsm_plan_info *info);

typedef struct {
    float lowdm, dmstep;
    uint32_t ndms, downsample;
    uint32_t nout;        /* output samples per DM row */
    uint32_t maxshift;    /* in decimated samples */
    uint64_t out_offset;  /* float offset of row 0 in the packed output */
    uint32_t first_dm;    /* global row index of row 0 */
    uint32_t ncalls;      /* subband: calls of the pass (stage-1 table rows); brute: ndms */
} mdsm_range_info;

int mdsm_plan_get_range(const mdsm_plan *plan, uint32_t r, mdsm_range_info *info);

/* Host copy of range r's integer shift table [ndms][nchans] (brute force),
 * or of pass r's stage-1 table [ncalls][nchans] (subband). */
int mdsm_plan_shift_table(const mdsm_plan *plan, uint32_t r, int32_t *host_out);

/* Subband pass r's stage-2 table [ncalls*calldms][nsubs]. */
int mdsm_plan_fine_table(const mdsm_plan *plan, uint32_t r, int32_t *host_out);

/* DM value of every output row (float32), as the reference labels them
 * (src/pelican-mdsm/modules/MdsmModule.cpp:189,202-210). */
int mdsm_plan_dm_values(const mdsm_plan *plan, float *host_out);

/* DM plan generation: passes (DM range, step, downsampling, and with nsubs
 * > 0 the subband DM step and calls) chosen from the smearing budget, as the
 * reference's python/ddplan.py computes them (observation, dedisp_method,
 * dm_steps; :6-253).  dt in seconds, f_ctr and bw in MHz, ok_smearing in ms
 * (0: the data's own).  Brute-force passes (nsubs = 0) have ncalls = 1 and
 * calldms = ndms and map to mdsm_range {lowdm, dmstep, ndms, binsize};
 * subband passes feed mdsm_plan_create_subband.  *npasses is set even when
 * cap is too small (MDSM_ERR_CAPACITY). */
typedef struct {
    double lodm, hidm;
    double f_ctr, bw, dt;
    uint32_t nchans, nsubs;
    double ok_smearing;
} mdsm_ddplan_params;

int mdsm_ddplan(const mdsm_ddplan_params *params, mdsm_pass *passes, uint32_t cap,
                uint32_t *npasses);

/* Launch shape of a plan's dedispersion kernels: the autotuner's search
 * space.  Results do not depend on it (integer sums are exact, float sums
 * keep their channel order).  No counterpart in the reference, whose launch
 * shapes are compile-time constants (src/mdsm/cache_brute_force.h:4-6,
 * src/mdsm/dedispersion_thread.cu:69-79). */
typedef struct {
    uint32_t chunk_rows;  /* channels per double-buffered LDS chunk, 2..8;
                             0 = 8, fewer where a window is too wide */
    uint32_t lds_budget;  /* LDS bytes per workgroup for staged channels;
                             0 = 79872 (two workgroups per CU); at most 162816 */
    uint32_t persistent;  /* 1: persistent workgroups over a device work queue */
    uint32_t grid;        /* persistent workgroups; 0 = 8 per compute unit (a launch whose
                             items fit in the grid runs one workgroup per item) */
    uint32_t tile_group;  /* work order: consecutive items cover this many time
                             tiles of one DM block before the next block
                             (0 or 1: every block of a tile, then the next tile) */
} mdsm_tuning;

/* Occupancy of the dedispersion kernel stage `stage` of the plan launches
 * with its current tuning (theoretical, from the kernel's register and LDS
 * use): resident workgroups per compute unit, waves per SIMD (4 SIMDs per
 * CU), VGPRs per lane, static and dynamic LDS bytes per workgroup, scratch
 * bytes per lane.  No counterpart in the reference. */
typedef struct {
    int32_t workgroups_per_cu, waves_per_simd, vgprs, static_lds, dynamic_lds, scratch_bytes;
} mdsm_occupancy;

int mdsm_plan_occupancy(mdsm_plan *plan, uint32_t stage, mdsm_occupancy *out);

/* LDS bytes one launch of dedispersion stage `stage` moves: *reads, the
 * trial reads (per trial slot, row and 256-sample tile: 2 bytes per sample
 * for 16-bit staged rows, 4 for float32; the padding slots of a block's last
 * trials and the whole last tile included),
 * and *writes, the staging stores (every row's phase layout, once per tile).
 * With the LDS read and write roofs (mdsm_probe_lds_read/_write) they give
 * the time the LDS needs for the launch.  No counterpart in the reference. */
int mdsm_plan_lds_bytes(const mdsm_plan *plan, uint32_t stage, uint64_t *reads,
                        uint64_t *writes);

int mdsm_plan_set_tuning(mdsm_plan *plan, const mdsm_tuning *t);
int mdsm_plan_get_tuning(const mdsm_plan *plan, mdsm_tuning *t);

/* Time the candidate launch shapes on the caller's block (reps timed runs
 * each, after one warm-up), keep the fastest, and remember it for every
 * later plan of the same shape in this process.  d_out/d_work as for
 * mdsm_dedisperse; the block is dedispersed several times.  best may be
 * NULL. */
int mdsm_plan_autotune(mdsm_plan *plan, const void *d_in, int64_t in_stride,
                       float *d_out, void *d_work, void *stream, uint32_t reps,
                       mdsm_tuning *best);

/* ------------------------------------------------------------------ */
/* Dedispersion                                                        */
/* ------------------------------------------------------------------ */

/* Dedisperse one block.  d_in: channel-major [nchans][in_stride] samples of
 * obs.dtype (in_stride >= nsamp_in).  d_out: out_elems floats, ranges packed
 * row-major ([ndms_r][nout_r] each, at out_offset_r).  d_work: work_bytes of
 * device memory, or NULL to use a workspace the plan allocates once.
 * Replaces the kernel launches of src/mdsm/dedispersion_thread.cu:64-243
 * (opt_dedisperse_loop, global_for_time_dedisperse_loop, the binning kernels
 * and opt_dedisperse_subband). */
int mdsm_dedisperse(mdsm_plan *plan, const void *d_in, int64_t in_stride,
                    float *d_out, void *d_work, void *stream);

/* mdsm_dedisperse without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid it).  No counterpart in the reference.  The workspace and the
 * work-queue counters live in a static per-GPU array of 256 MiB held until
 * the process ends; plans needing more workspace than
 * mdsm_noalloc_max_work() bytes are MDSM_ERR_INVALID.  The plan's tables must
 * be on the device (mdsm_plan_upload at set-up; MDSM_ERR_STATE otherwise).
 * All calls on one GPU share the array: calls that overlap on one GPU race.
 * Results equal mdsm_dedisperse's. */
uint64_t mdsm_noalloc_max_work(void);
int mdsm_dedisperse_noalloc(mdsm_plan *plan, const void *d_in, int64_t in_stride,
                            float *d_out, void *stream);

/* The same float32 brute-force result from a plain kernel in the shape of
 * the reference's opt_dedisperse_loop (src/mdsm/dedispersion_kernel.cu:51-77):
 * one thread per (DM, sample), channel loop of cached global gathers.  A
 * performance baseline (float32 plans with undecimated ranges only); the
 * sums are in the same channel order, so results equal mdsm_dedisperse. */
int mdsm_dedisperse_plain(mdsm_plan *plan, const float *d_in, int64_t in_stride,
                          float *d_out, void *stream);

/* k-sample time decimation (sum) of a channel-major block:
 * d_out[c][t] = sum_{j<k} d_in[c][k t + j], t < nsamp/k.  uint8 input gives
 * uint16 output, float16/float32 input gives float32.
 * Replaces binning_kernel/inplace_binning_kernel
 * (src/mdsm/dedispersion_kernel.cu:203-246). */
int mdsm_decimate(const void *d_in, int32_t dtype, uint32_t nchans,
                  int64_t in_stride, uint32_t nsamp, uint32_t k, void *d_out,
                  int64_t out_stride, void *stream);

/* ------------------------------------------------------------------ */
/* RFI clipper and single-pulse detector                               */
/* ------------------------------------------------------------------ */

/* Corner turn of a time-major block into channel-major rows:
 * d_out[c][out_col0 + t] = d_in[t][invert ? nchans-1-c : c], t < nsamp,
 * c < nchans (in_stride / out_stride in elements).  dtype_out is dtype_in
 * (storage kept, e.g. uint8 for the integer path) or MDSM_F32 (unpacked to
 * float as the reference does, (float)byte for uint8).  Replaces the host
 * corner-turn loops of src/pelican-mdsm/modules/MdsmModule.cpp:135-151
 * (invertChannels: :27,137,145) and src/MDSM.cpp:100-120, and the uint8
 * unpack of src/mdsm/file_handler.cpp:255-264. */
int mdsm_corner_turn(const void *d_in, int32_t dtype_in, uint32_t nsamp,
                     uint32_t nchans, int64_t in_stride, int32_t invert,
                     void *d_out, int32_t dtype_out, int64_t out_stride,
                     int64_t out_col0, void *stream);

/* ------------------------------------------------------------------ */
/* SIGPROC filterbank files                                            */
/* ------------------------------------------------------------------ */

/* Header of a SIGPROC filterbank file, the fields src/mdsm/file_handler.h:6-28
 * keeps (keywords and value types as read_header parses them,
 * src/mdsm/file_handler.cpp:56-177). */
typedef struct {
    char source_name[80];
    char rawdatafile[80];
    int32_t machine_id, telescope_id, data_type, nchans, nbits, nifs;
    int32_t nbeams, ibeam, barycentric, pulsarcentric, nbins, nsamples;
    double tstart, tsamp, fch1, foff, refdm, az_start, za_start, src_raj,
        src_dej, period;
    int64_t npuls;
    uint64_t header_bytes; /* offset of the first spectrum */
    uint64_t nspectra;     /* whole spectra after the header */
} mdsm_sigproc_header;

/* Parse a file's header.  An unknown keyword is MDSM_ERR_PARSE (the
 * reference exits).  Replaces read_header (file_handler.cpp:56-177). */
int mdsm_sigproc_read_header(const char *path, mdsm_sigproc_header *h);

/* Sequential reader of raw packed spectra (nchans*nifs*nbits/8 bytes each,
 * time-major); *got = whole spectra read (fewer at the end of the file).
 * The samples are unpacked on the GPU by mdsm_unpack_spectra.  Replaces the
 * file side of read_block (file_handler.cpp:193-290) and readBinaryData
 * (src/MDSM.cpp:100-120). */
typedef struct mdsm_sigproc_reader mdsm_sigproc_reader;
int mdsm_sigproc_open(const char *path, mdsm_sigproc_reader **r, mdsm_sigproc_header *h);
int mdsm_sigproc_read(mdsm_sigproc_reader *r, uint64_t nspectra, void *host_raw,
                      uint64_t *got);
int mdsm_sigproc_seek(mdsm_sigproc_reader *r, uint64_t spectrum);
void mdsm_sigproc_close(mdsm_sigproc_reader *r);

/* Unpack time-major packed spectra (nbits = 1, 2, 4, 8 unsigned, 16
 * unsigned, 32 float; sub-byte samples least significant bits first, as
 * read_block does) into channel-major rows with the corner turn:
 * d_out[c][out_col0 + t] = sample c of spectrum t (channel order reversed
 * when invert != 0).  dtype_out: MDSM_U8 (nbits <= 8) or MDSM_F32.
 * in_stride_bytes: bytes between spectra.  Replaces read_block's unpack
 * (file_handler.cpp:193-290) and the corner turn of src/MDSM.cpp:100-120. */
int mdsm_unpack_spectra(const void *d_raw, int32_t nbits, uint32_t nsamp,
                        uint32_t nchans, int64_t in_stride_bytes, int32_t invert,
                        void *d_out, int32_t dtype_out, int64_t out_stride,
                        int64_t out_col0, void *stream);

/* Dedispersed-series writer: one SIGPROC file <prefix>_<dm>.fil (float32,
 * one channel) per DM value, header as DedispersedDataWriter writes it; each
 * write appends, for every file, the row whose DM value is within 1e-5 of
 * the file's.  Replaces src/pelican-mdsm/modules/DedispersedDataWriter.cpp:
 * 17-135. */
typedef struct mdsm_series_writer mdsm_series_writer;
int mdsm_series_writer_open(const char *prefix, const float *dm_values, uint32_t ndm,
                            double fch1, double foff, double tsamp, double tstart,
                            mdsm_series_writer **w);
int mdsm_series_writer_write(mdsm_series_writer *w, const float *host_rows,
                             const float *row_dm_values, uint32_t nrows,
                             uint32_t nsamp, int64_t stride);
void mdsm_series_writer_close(mdsm_series_writer *w);

/* In-place detrend and normalise of nrows rows of n floats: each segment of
 * seg_len samples (the last one may be shorter) has its least-squares line
 * a + b*x removed, x = i - L/2 + 0.5 centred on the segment, and is divided
 * by the RMS of the residual.  Sums are formed in FP64.  A segment whose
 * residual RMS is zero is set to zero (the reference divides by it).
 * Replaces detrend_normalise (src/prototypes/detrend.cu:12-100). */
int mdsm_detrend(float *d_x, uint32_t nrows, uint32_t n, int64_t stride,
                 uint32_t seg_len, void *stream);

/* Per-channel robust clip of uint8 samples [t0, t0+n) of each channel:
 * med = lower median, mad = lower median of |x-med|, samples with
 * |x-med| > (nsigma*1.4826f)*mad are replaced by med.  d_out may equal d_in.
 * d_med/d_mad/d_nclipped: per-channel int32 outputs (may be NULL).
 * Replaces the Pelican-LOFAR RFI_Clipper stage that precedes MdsmModule
 * (src/pelican-mdsm/pipelines/src/MdsmPipeline.cpp:57-59) and the prototype
 * src/prototypes/rfi_clipper.cu:8-67. */
int mdsm_clip_u8(const uint8_t *d_in, int64_t in_stride, uint32_t nchans,
                 int64_t t0, uint32_t n, float nsigma, uint8_t *d_out,
                 int64_t out_stride, int32_t *d_med, int32_t *d_mad,
                 int32_t *d_nclipped, void *stream);

typedef struct {
    uint32_t dm;     /* output row */
    int32_t t;       /* first sample of the boxcar, relative to the block's first
                        sample (negative in a stream search: the window starts
                        in the previous block) */
    uint32_t width;  /* boxcar width in samples */
    float snr;
} mdsm_cand;

/* Boxcar single-pulse search of a DM x time block d_x [ndms][stride]
 * (n valid samples per row).  Widths must be powers of two (<= 1024).
 * Per row: mean and sigma over the n samples (FP64); boxcar sums as a
 * pairwise FP64 tree; a candidate (d,t,w,snr) where
 * snr = (S_w(t) - w*mean)/(sigma*sqrt(w)) > thr and S_w peaks at t
 * (S_w(t) >= S_w(t-1), S_w(t) > S_w(t+1)).  The library zeroes *d_count,
 * then appends candidates in arbitrary order to d_cands (capacity cap);
 * *d_count receives the total found, which may exceed cap (the caller
 * checks it after the stream completes).  d_mean/d_sigma: required [ndms]
 * FP64 outputs, the row statistics.
 * Replaces the detector thread's threshold search
 * (src/mdsm/dedispersion_output.cpp:135-213) with the boxcar search of
 * src/prototypes/singlepulse.cpp:41-146. */
int mdsm_detect_boxcar(const float *d_x, uint32_t ndms, uint32_t n,
                       int64_t stride, const uint32_t *widths, uint32_t nwidths,
                       double thr, mdsm_cand *d_cands, uint32_t cap,
                       uint32_t *d_count, double *d_mean, double *d_sigma,
                       void *stream);

/* Streaming form of mdsm_detect_boxcar for consecutive blocks of one
 * series: every boxcar window is searched exactly once, windows that
 * straddle two blocks included.  d_carry ([ndms][carry_stride] floats, at
 * least mdsm_boxcar_carry_len(widths) = maxw + 1 rounded up to a multiple of
 * 4 per row) holds the previous block's last samples; carry_valid = 0 for the
 * first block of a stream (nothing carried).  The searched series of row d is
 * the carried samples followed by the block's n samples; window (t, w) is
 * searched by the block holding sample t + w, the last one its right
 * neighbour reads (t >= -carried), with this block's mean and sigma (over its
 * n samples; d_mean/d_sigma receive them).  Candidate t is relative to the
 * block's first sample.  The peak rule is mdsm_detect_boxcar's; the window
 * ending at the block's last sample is left to the next block.  The call then
 * stores the block's last samples in d_carry (requires n >= the carry
 * length).  No counterpart in the reference, whose detectors search each
 * block on its own (src/mdsm/dedispersion_output.cpp:176-181). */
uint32_t mdsm_boxcar_carry_len(uint32_t widths_mask);
int mdsm_detect_boxcar_stream(const float *d_x, uint32_t ndms, uint32_t n, int64_t stride,
                              float *d_carry, int64_t carry_stride, int32_t carry_valid,
                              const uint32_t *widths, uint32_t nwidths, double thr,
                              mdsm_cand *d_cands, uint32_t cap, uint32_t *d_count,
                              double *d_mean, double *d_sigma, void *stream);

/* Median-of-five threshold detector of the reference's output thread
 * (src/mdsm/dedispersion_output.cpp:14-53,56-213): per row d, for l < n-4,
 * m = median(x[l..l+4]); a candidate (d, l, 5, snr) where dm_values[d] > 1.0
 * and m - mean >= rms*thr (float arithmetic, as the reference).  Statistics:
 *   chunk arrays given  : mean = d_chunk_mean[l/spc], rms = d_chunk_rms[l/spc],
 *                         snr = (m-mean)/rms   (process_brute with noise arrays);
 *   per_row_stats != 0  : per-row mean/rms of the n samples (process_subband);
 *   otherwise           : the whole-block statistics of process_brute without
 *                         noise arrays (:143-167): the ndms x n values, rows one
 *                         after another, in chunks of min(values, 32768); each
 *                         chunk summed in double in index order, the chunk
 *                         means added to a float mean and the mean squares to a
 *                         double; the last whole chunk and any partial one are
 *                         not read; snr = m/rms.
 * d_stats: device scratch of mdsm_median5_stats_len(ndms, n) doubles
 * (d_stats[0..1] = block mean and rms in the whole-block mode).  *d_count is
 * zeroed by the call. */
uint64_t mdsm_median5_stats_len(uint32_t ndms, uint32_t n);
int mdsm_detect_median5(const float *d_x, uint32_t ndms, uint32_t n,
                        int64_t stride, const float *d_dm_values, double thr,
                        int32_t per_row_stats, const float *d_chunk_mean,
                        const float *d_chunk_rms, uint32_t samples_per_chunk,
                        mdsm_cand *d_cands, uint32_t cap, uint32_t *d_count,
                        double *d_stats, void *stream);

/* ------------------------------------------------------------------ */
/* Batches: many plans (beams) per launch, HIP-graph replay            */
/* ------------------------------------------------------------------ */

/* A batch runs up to 64 plans of any shapes, dtypes and decimations as one
 * step: each plan's decimation kernels, then one dedispersion launch per
 * stage index over the union of the plans' work (longest items first, a
 * persistent work queue when the plans' tuning asks for it).  Plan i reads
 * d_ins[i] (row stride in_strides[i]) and writes d_outs[i]; d_works[i] is
 * its workspace (d_works or an entry NULL: the batch allocates one).  Each
 * plan's results equal mdsm_dedisperse on that plan alone.  The plans must
 * outlive the batch.  Replaces the per-thread launch loop of
 * src/mdsm/dedispersion_thread.cu:215-243 (one GPU, one beam per call);
 * the batching itself has no counterpart in the reference. */
typedef struct mdsm_batch mdsm_batch;

int mdsm_batch_create(mdsm_plan *const *plans, uint32_t nplans, mdsm_batch **out);
void mdsm_batch_destroy(mdsm_batch *batch);
int mdsm_batch_run(mdsm_batch *batch, const void *const *d_ins,
                   const int64_t *in_strides, float *const *d_outs,
                   void *const *d_works, void *stream);

/* Capture one step on `stream` (non-NULL) into a HIP graph; replay it with
 * mdsm_batch_replay.  The graph holds the buffers given here.  Changing a
 * plan's tuning afterwards requires a new capture. */
int mdsm_batch_capture(mdsm_batch *batch, const void *const *d_ins,
                       const int64_t *in_strides, float *const *d_outs,
                       void *const *d_works, void *stream);
int mdsm_batch_replay(mdsm_batch *batch, void *stream);

/* Autotune the batch as a whole (same candidates as mdsm_plan_autotune,
 * applied to every plan); drops a captured graph. */
int mdsm_batch_autotune(mdsm_batch *batch, const void *const *d_ins,
                        const int64_t *in_strides, float *const *d_outs,
                        void *const *d_works, void *stream, uint32_t reps,
                        mdsm_tuning *best);

/* ------------------------------------------------------------------ */
/* Streaming server: resident beams, continuous batching               */
/* ------------------------------------------------------------------ */
/* Replaces the reference's per-GPU worker loop and barrier protocol
 * (src/mdsm/dedispersion_thread.cu:246-401, src/mdsm/dedispersion_manager.cpp:
 * 324-501) and the host-side overlap copy (dedispersion_thread.cu:266-303):
 * each beam keeps its max-delay overlap on the device; blocks from many beams
 * are admitted by HBM budget and dedispersed together in one launch. */

typedef struct mdsm_server mdsm_server;

typedef struct {
    uint64_t hbm_budget;    /* device bytes the server may hold (0: 90% of free memory) */
    uint32_t max_inflight;  /* blocks on the device at once (0: 16) */
    uint32_t max_batch;     /* blocks per dedispersion launch (0: 16; <= 64) */
    uint32_t persistent;    /* 1: persistent workgroups (4 per CU) pulling from a device work
                               counter; > 1: that many workgroups.  A launch whose items fit
                               in the grid runs one workgroup per item instead */
    uint32_t queue_depth;   /* pushed-but-unadmitted blocks per beam (0: 4) */
} mdsm_server_config;

typedef struct {
    float clip_nsigma;      /* > 0: median/MAD clip of each block's new samples (uint8) */
    double snr_thr;         /* boxcar S/N threshold */
    uint32_t widths_mask;   /* OR of boxcar widths (powers of two <= 1024); 0: no detection */
    uint32_t cand_cap;      /* candidates kept per block (0: 65536) */
    uint32_t retain_output; /* 1: keep each block's DM x time output until mdsm_server_release */
    uint32_t pad;
} mdsm_detect_config;

/* mdsm_result.flags */
enum {
    MDSM_RESULT_OVERFLOW = 1,    /* a range found more than cand_cap candidates: the kept
                                    subset of that range is arbitrary (atomic order) */
    MDSM_RESULT_ZERO_FILLED = 2, /* the block never arrived (a sequence gap) and was
                                    processed as zeros, as the reference's UDP chunker
                                    fills lost packets (src/standalone/UdpChunker.cpp:
                                    104-128) */
    MDSM_RESULT_SKIPPED = 4      /* zero-filled block whose whole history is zeros: its
                                    output is exactly zero and was not computed */
};

typedef struct {
    uint32_t beam;
    int32_t status;         /* MDSM_OK, or the error that ended the block */
    uint64_t seq;           /* block number within the beam */
    uint32_t ncand;         /* candidates found (may exceed those kept and returned) */
    uint32_t nreturned;     /* candidates copied to poll's cands (<= its cap) */
    uint32_t nclipped;      /* samples replaced by the clipper */
    uint32_t batch;         /* blocks in the launch that carried this block */
    uint32_t flags;         /* MDSM_RESULT_* */
    uint32_t pad;
    double t_arrival;       /* seconds, steady clock (as passed to push, or push time) */
    double t_admit;         /* admitted to the device */
    double t_done;          /* results on the host */
} mdsm_result;

typedef struct {
    uint64_t pushed, admitted, completed, deferred_budget, deferred_inflight, launches;
    uint64_t bytes_in_use, bytes_peak, budget;
    uint32_t beams, inflight;
} mdsm_server_stats;

int mdsm_server_create(const mdsm_server_config *cfg, mdsm_server **out);
void mdsm_server_destroy(mdsm_server *srv);

/* Open a beam.  obs->nsamp_in is N, the new samples per block.  The beam's
 * overlap is M = the plan's maxshift (full-rate samples); output sample t of
 * block k (range r, decimation k_r) starts at stream sample k*N - M + t*k_r,
 * so the first n_r = N/k_r samples of each row of consecutive blocks tile the
 * range's series without gap or repeat.  Detection searches that series as
 * mdsm_detect_boxcar_stream does: each beam keeps the last
 * mdsm_boxcar_carry_len(widths) samples of every DM row resident on the
 * device, so boxcar windows straddling two blocks are searched once (by the
 * later block; candidate t may be negative), with statistics per row and
 * block.  Requires n_r >= that carry length.
 * MDSM_ERR_BUDGET when the beam's resident buffers plus one block can never
 * fit the budget. */
int mdsm_beam_open(mdsm_server *srv, const mdsm_obs *obs, const mdsm_range *ranges,
                   uint32_t nranges, const mdsm_detect_config *det, uint32_t *beam_id);
/* Close a beam: every block already pushed is processed and reported first
 * (also while admissions are held), then its resources are freed. */
int mdsm_beam_close(mdsm_server *srv, uint32_t beam_id);
/* Close a beam mid-stream: its queued blocks are dropped and reported with
 * status MDSM_ERR_CANCELLED; a block already on the device completes and is
 * reported as usual. */
int mdsm_beam_cancel(mdsm_server *srv, uint32_t beam_id);
/* overlap M, plan info of the beam (nsamp_in of the plan is M + N) */
int mdsm_beam_info(mdsm_server *srv, uint32_t beam_id, uint32_t *overlap,
                   mdsm_plan_info *info);

/* Queue one block: host channel-major [nchans][stride] samples of obs.dtype,
 * N per channel; copied before return.  arrival_s < 0: stamped now.
 * *seq receives the block's number (the next in the beam's sequence).
 * MDSM_ERR_AGAIN when the beam's queue is full (back-pressure). */
int mdsm_beam_push(mdsm_server *srv, uint32_t beam_id, const void *host, int64_t stride,
                   double arrival_s, uint64_t *seq);

/* Push with the caller's block number, as a packet stream numbers its
 * blocks.  seq equal to the beam's next expected number: as mdsm_beam_push.
 * seq below it (a duplicate or a block that arrives after its gap was
 * filled): MDSM_ERR_STATE, nothing queued.  seq above it: each missing block
 * is queued first as a block of zeros and reported with
 * MDSM_RESULT_ZERO_FILLED (the reference's chunker fills lost packets with
 * empty ones and drops duplicates, src/standalone/UdpChunker.cpp:100-128);
 * missing blocks whose whole history is zeros (more than M/N of them in a
 * row) are not computed -- their output is exactly zero -- and are reported
 * with MDSM_RESULT_SKIPPED as well.  Missing blocks do not count against the
 * queue depth; gaps longer than 2^20 blocks are MDSM_ERR_INVALID. */
int mdsm_beam_push_seq(mdsm_server *srv, uint32_t beam_id, const void *host, int64_t stride,
                       uint64_t seq, double arrival_s);

/* Queue one block that is already in device memory (channel-major
 * [nchans][stride] samples of obs.dtype on the server's device), for callers
 * whose upstream stage runs on the GPU.  The server reads d_block when the
 * block runs and does not copy it first: it must stay allocated and unchanged
 * until the block's result has been polled. */
int mdsm_beam_push_device(mdsm_server *srv, uint32_t beam_id, const void *d_block,
                          int64_t stride, double arrival_s, uint64_t *seq);

/* Zero-copy push: mdsm_beam_acquire lends the beam's next free pinned
 * staging block ([nchans][N] samples of obs.dtype, row stride N); the caller
 * fills it and queues it with mdsm_beam_commit (arrival_s as for push).
 * MDSM_ERR_AGAIN when no block is free.  A lent block counts against the
 * beam's queue depth until committed.  Closing or cancelling the beam frees
 * its staging blocks, lent ones included; a commit while the beam is closing
 * or exporting its state is refused (MDSM_ERR_STATE) and the block returned
 * to the beam.  (The reference's caller writes the
 * library-owned input buffer in place the same way,
 * src/pelican-mdsm/modules/MdsmModule.cpp:135-151.) */
int mdsm_beam_acquire(mdsm_server *srv, uint32_t beam_id, void **host_block);
int mdsm_beam_commit(mdsm_server *srv, uint32_t beam_id, void *host_block,
                     double arrival_s, uint64_t *seq);

/* Next completed block.  Its kept candidates (sorted by dm, width, t; at
 * most cand_cap per range) are copied into cands[0..res->nreturned),
 * nreturned <= cap.  wait_ms < 0 waits indefinitely.  MDSM_ERR_AGAIN when
 * nothing completed in time. */
int mdsm_server_poll(mdsm_server *srv, mdsm_result *res, mdsm_cand *cands, uint32_t cap,
                     int32_t wait_ms);

/* Retained output of block (beam, seq) (retain_output = 1): device pointer to
 * the packed plan output, and its release back to the pool. */
int mdsm_server_output(mdsm_server *srv, uint32_t beam_id, uint64_t seq, const float **d_out);
int mdsm_server_release(mdsm_server *srv, uint32_t beam_id, uint64_t seq);

/* Pause (hold != 0) or resume admissions: while held, pushed blocks queue
 * and nothing new is launched, so that a burst pushed in the meantime is
 * admitted as one batch on resume (blocks already on the GPU run on).  No
 * reference counterpart: the reference's manager has no batching.
 * mdsm_server_drain resumes admissions. */
int mdsm_server_hold(mdsm_server *srv, int hold);

/* Block until every pushed block has completed (results stay queued).
 * Resumes admissions if they were held. */
int mdsm_server_drain(mdsm_server *srv);

/* Beam state for re-admission: the device-resident overlap (the last M input
 * samples of every channel, after clipping), the detector's carried samples
 * of every DM row and the next sequence number.  mdsm_beam_export waits until
 * the beam's pushed blocks have completed (admitting them while held), then
 * writes the state to host (*bytes = its size; with host NULL or cap too
 * small only *bytes is set and MDSM_ERR_CAPACITY returned).
 * mdsm_beam_import loads it into a beam opened with the same observation,
 * plan and widths before any block was pushed to it; the stream then goes
 * on as if it had never left.  The reference carries the overlap on the host
 * (src/mdsm/dedispersion_thread.cu:266-303) and the per-chunk noise
 * statistics in the module (src/pelican-mdsm/modules/MdsmModule.cpp:226-237). */
int mdsm_beam_export(mdsm_server *srv, uint32_t beam_id, void *host, uint64_t cap,
                     uint64_t *bytes);
int mdsm_beam_import(mdsm_server *srv, uint32_t beam_id, const void *host, uint64_t bytes);
int mdsm_server_get_stats(mdsm_server *srv, mdsm_server_stats *st);
/* Steady clock in seconds (the clock of t_arrival/t_admit/t_done). */
double mdsm_now(void);

/* ------------------------------------------------------------------ */
/* Synthetic filterbanks (seeded, host side)                           */
/* ------------------------------------------------------------------ */

typedef struct {
    double dm;        /* pc cm^-3 */
    uint32_t t0;      /* arrival sample at channel 0 (top of band) */
    uint32_t width;   /* samples */
    float amp;        /* in units of the noise sigma, per channel */
    uint32_t pad;
} mdsm_pulse;

typedef struct {
    uint64_t seed;
    int32_t dtype;      /* uint8: clip(round(mean + sigma*(z+pulse)),0,255) */
    uint32_t nchans;
    uint32_t nsamp;
    float fch1, foff, tsamp;
    float mean, sigma;  /* noise N(mean, sigma) before gain */
    float gain_lo, gain_hi; /* per-channel gain U[lo,hi] (both 1: none) */
    uint32_t npulses;
    const mdsm_pulse *pulses;
    uint32_t nthreads;  /* host threads (0 = hardware concurrency) */
} mdsm_synth_spec;

/* Fill host_out ([nchans][out_stride] of dtype) with seeded Gaussian noise
 * plus pulses placed on the MDSM_SHIFT_NS track (nearest sample).  The value
 * of every sample depends only on (seed, channel, sample). */
int mdsm_synth_filterbank(const mdsm_synth_spec *spec, void *host_out,
                          int64_t out_stride);

/* The same generator on the device: d_out is [nchans][out_stride] of dtype
 * in device memory, written on `stream` (at most 16 pulses, nchans <= 65535;
 * nthreads is ignored).  Agrees with mdsm_synth_filterbank to the rounding of
 * the device's log and cos (a uint8 sample can differ by one where the value
 * falls on a rounding boundary).  For benchmark inputs too large to make on
 * the host. */
int mdsm_synth_filterbank_device(const mdsm_synth_spec *spec, void *d_out, int64_t out_stride,
                                 void *stream);

/* ------------------------------------------------------------------ */
/* Device plumbing                                                     */
/* ------------------------------------------------------------------ */
/* Memory, copies and streams, so that hosts need no second HIP runtime.
 * They replace the reference worker's direct CUDA runtime calls:
 * cudaSetDevice/cudaMalloc (src/mdsm/dedispersion_thread.cu:258-262), the
 * H2D/D2H copies (:285,297,353-355), cudaMemset (:324) and the device
 * enumeration of initialise_devices (:4-62). */

typedef struct {
    char name[256];
    char arch[64];
    uint64_t total_mem, free_mem;
    int32_t cus, clock_khz, mem_clock_khz, mem_bus_bits;
    uint32_t lds_per_block;
    uint32_t pad;
} mdsm_device_props;

enum { MDSM_H2D = 1, MDSM_D2H = 2, MDSM_D2D = 3 };

int mdsm_device_count(int *count);
int mdsm_device_set(int device);
int mdsm_device_query(int device, mdsm_device_props *props);
int mdsm_device_malloc(void **ptr, uint64_t bytes);
int mdsm_device_free(void *ptr);
int mdsm_host_malloc_pinned(void **ptr, uint64_t bytes);
int mdsm_host_free_pinned(void *ptr);
/* stream == NULL: synchronous copy. */
int mdsm_memcpy(void *dst, const void *src, uint64_t bytes, int kind, void *stream);
int mdsm_memcpy_2d(void *dst, uint64_t dpitch, const void *src, uint64_t spitch,
                   uint64_t width, uint64_t rows, int kind, void *stream);
/* Fill count uint32 words with value by a kernel (graph-capturable). */
int mdsm_fill_u32(void *dst, uint32_t value, uint64_t count, void *stream);
int mdsm_stream_create(void **stream);
int mdsm_stream_destroy(void *stream);
int mdsm_stream_synchronize(void *stream);
int mdsm_device_synchronize(void);

/* ------------------------------------------------------------------ */
/* Kernel timing and errors                                            */
/* ------------------------------------------------------------------ */

/* When enabled, every kernel the library launches is bracketed by a pair
 * of hipEvents on its launch stream.  mdsm_timing_get synchronises those
 * events and returns the summed duration (ms) and launch count of kernels
 * named `kernel` ("dedisperse", "decimate", "clip", "detect", ...). */
int mdsm_timing_enable(int on);
int mdsm_timing_reset(void);
int mdsm_timing_get(const char *kernel, double *total_ms, uint64_t *launches);

/* LDS read bandwidth of this GPU (TB/s) for reads of bytes_per_lane = 4, 8
 * or 16 bytes per lane (ds_read_b32/_b64/_b128, 64 contiguous lanes, a
 * wave-uniform offset moving on every read: the trial reads of the
 * dedispersion kernel) at waves_per_simd (1..8) resident waves per SIMD.
 * The roof the dedispersion kernel's LDS rate is measured against.  No
 * counterpart in the reference. */
int mdsm_probe_lds_read(uint32_t bytes_per_lane, uint32_t waves_per_simd, double *tbs);

/* LDS write bandwidth (TB/s), the same way: ds_write_b32/_b64/_b128 of 64
 * contiguous lanes at a wave-uniform moving offset (the dedispersion
 * kernel's staging stores).  No counterpart in the reference. */
int mdsm_probe_lds_write(uint32_t bytes_per_lane, uint32_t waves_per_simd, double *tbs);

/* LDS read bandwidth (TB/s) of the read forms the dedispersion kernel's row
 * layouts chose between, the same way, at waves_per_simd (1..8): form 1 two
 * dwords 256 bytes apart per lane (ds_read2st64_b32: the float16 rows'
 * reads), 2 two 8-byte entries 512 bytes apart (ds_read2st64_b64: the
 * float32 rows'), 3 8 bytes at a 4-byte-aligned, not 8-byte-aligned,
 * address (ds_read_b64), 4 the same 8 bytes as two adjacent dwords
 * (ds_read2_b32).  No counterpart in the reference. */
int mdsm_probe_lds_read_form(uint32_t form, uint32_t waves_per_simd, double *tbs);

const char *mdsm_last_error(void);
const char *mdsm_version(void);

#ifdef __cplusplus
}
#endif
#endif /* MDSM_C_H */
