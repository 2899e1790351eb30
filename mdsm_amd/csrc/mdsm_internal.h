// Internal structures shared by the host plan builder and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "../../include/mdsm_c.h"

namespace mdsm {

// Dedispersion tiling (the kernel's compile-time shape).
#ifndef MDSM_WAVES
#define MDSM_WAVES 8
#endif
#ifndef MDSM_RPW
#define MDSM_RPW 8
#endif
constexpr int kWaves = MDSM_WAVES;   // waves per workgroup
constexpr int kRPW = MDSM_RPW;       // DM trials per wave (registers)
constexpr int kThreads = 64 * kWaves;
constexpr int kDB = kWaves * kRPW;   // DM trials per block
constexpr int kTile = 256;          // output samples per tile (4 per lane)

// Buffers a block can read or write.
enum BufId : int32_t { BUF_IN = 0, BUF_OUT = 1, BUF_WORK = 2 };

// Internal sample types (the public three plus uint16 decimated sums).
enum DType : int32_t { DT_U8 = 0, DT_F16 = 1, DT_F32 = 2, DT_U16 = 3 };

inline size_t dtype_size(int dt)
{
    return dt == DT_U8 ? 1 : (dt == DT_F32 ? 4 : 2);
}

// One block of up to kDB trials that share one set of input rows.
// Trial j of the block computes, for t < nout,
//   out[out_off + j*out_trial_stride + t] =
//       (1/k) * sum_{i<nrows} in[row0+i][abase[i] + t + rel[i][j]]
// where in[row][col] = buf[src][row*row_stride + col] (row_stride of BUF_IN
// is the caller's in_stride, given at launch).
struct BlockDesc {
    int64_t row0;
    int64_t row_stride;
    int64_t in_off;           // element offset of row 0 col 0 within buf[src]
    int64_t out_off;          // float offset of trial 0 within buf[dst]
    int64_t out_trial_stride;
    int32_t in_len;           // valid columns per row
    int32_t nrows;
    int32_t ntrials;
    int32_t W;                // widest row window of the block (samples), multiple of 64
    int32_t nout;
    int32_t k;                // divisor applied at store (1 = none)
    int32_t src, dst;
    int32_t dtype;
    int32_t tab_off;          // int offset of rel[nrows][kDB] in the rel table
    int32_t base_off;         // int offset of the block's (window start, row info) pairs in
                              // the base table: info bits 0-9 entries per phase of the row's
                              // window, bits 10- its byte place in the chunk buffer
    int32_t ccr;              // rows staged per LDS chunk
    int32_t flush;            // integer path: rows summed in 16-bit lanes between flushes
    int32_t rpw;              // trials per wave: kRPW, or kRPW/2 for float blocks of <= kDB/2 trials
    int32_t bufsz;            // bytes of one chunk buffer (its rows at their own widths)
    // Row statistics of a uint8 range at full rate (launches with stats on):
    // item (trial j, tile i < stat_tiles) writes the FP64 pair (sum, sum of
    // squares) of its samples t < stat_n to work + stat_off + 16*(j*stat_tiles + i).
    // stat_tiles = 0: the block writes none.
    int32_t stat_tiles;
    int32_t stat_n;
    int64_t stat_off;
};

// Staged LDS rows of W samples:
//  * uint8 / uint16: four phases of 8-byte entries of 16-bit lanes (8 bytes
//    per sample; one ds_read_b64 per lane and offset);
//  * float16: two planes of 16-bit pairs, dword j of plane p holding samples
//    2j+p and 2j+p+1 (4 bytes per sample; one ds_read2st64_b32);
//  * float32: two phases of float2 entries, entry j of phase p holding
//    samples 2j+p and 2j+p+1 (8 bytes per sample; one ds_read2st64_b64).
constexpr bool two_planes16(int dtype) { return dtype == DT_F16; }
inline int lds_row_bytes(int dtype, int W)
{
    if (dtype == DT_F32) return 8 * W;
    return two_planes16(dtype) ? 4 * W : 8 * W;
}
// Byte offset of shift o (relative to the row base) in a staged row of W samples.
inline int32_t lds_offset(int dtype, int W, int32_t o)
{
    if (two_planes16(dtype)) return (o & 1) * (2 * W) + (o >> 1) * 4;
    if (dtype == DT_F32) return (o & 1) * (4 * W) + (o >> 1) * 8;
    return (o & 3) * (2 * W) + (o >> 2) * 8;  // four phases of W/4 entries
}
inline bool int_path(int dtype) { return dtype == DT_U8 || dtype == DT_U16; }
// Widest staged row (samples) of a block: bounded by the kernel's register
// prefetch, 256 entry columns (1024 samples), or 384 (1536 samples) for
// float16, whose wide subband windows would otherwise split blocks (float32
// rows carry twice the prefetch registers and stay at 1024).
constexpr int max_window(int dtype) { return dtype == DT_F16 ? 1536 : 1024; }
constexpr int kLdsBudget = 79872;  // default LDS bytes per workgroup (two per CU: 2 x 78 KiB + static)
constexpr int kLdsMax = 163840 - 1024;  // gfx950: 160 KiB per CU, less the kernel's static LDS
constexpr int kChunkRows = 8;      // default channels per LDS chunk (2..8)

// A run of blocks tiled over the same number of output tiles; flat work
// items of a segment are ordered tile-major (all blocks of tile 0 first).
// Segments are self-describing so that one launch can mix plans: each names
// its plan's device tables and the job (buffer set) it reads and writes.
struct SegDesc {
    int64_t item_start;
    const BlockDesc *blocks;  // device: the plan's block table
    const int32_t *rel;       // device: the plan's offset table
    const int32_t *abase;     // device: the plan's base table
    int32_t first_block;
    int32_t nblocks;
    int32_t ntiles;
    int32_t job;              // index into the launch's job buffers
};

// Buffers of one job (one block of one beam) in a launch.
constexpr int kMaxJobs = 64;
struct JobBufs {
    const void *in;
    float *out;
    char *work;
    int64_t in_stride;
};

// One dedispersion launch: a contiguous run of segments.
struct Stage {
    int32_t seg_begin = 0, seg_end = 0;
    int64_t items = 0;
    int32_t wmax = 0;
    int32_t lds_bytes = 0;   // dynamic LDS of the launch
    int32_t dtypes = 0;      // bit d set when a block of DType d is in the launch
};

struct DecimStep {
    int32_t k = 1;
    int32_t dtype_in = DT_U8, dtype_out = DT_U8;
    int64_t work_off = 0;     // byte offset of the decimated copy in workspace
    int64_t stride = 0;       // elements per row of the copy
    int32_t nsamp_out = 0;
};

struct RangeHost {
    mdsm_range_info info{};
    std::vector<int32_t> shift;   // [ndms][nchans] brute; [ncalls][nchans] subband coarse
    std::vector<int32_t> fine;    // subband stage-2 [ncalls*calldms][nsubs]
    std::vector<float> dm_values;
    int32_t decim = -1;           // DecimStep index, -1 = raw input
    mdsm_pass pass{};
    int64_t sb_off = 0;           // subband stage-1 buffer (byte offset)
    int32_t len1 = 0;
    // Row statistics written by the dedispersion launch (uint8 at full rate,
    // sums exact in FP64): per DM row stat_tiles (sum, sum of squares) pairs
    // of its first stat_n samples at work + stat_off.  stat_tiles = 0: none.
    int64_t stat_off = 0;
    int32_t stat_tiles = 0;
    int32_t stat_n = 0;
};

struct KernelTables {
    BlockDesc *blocks = nullptr;
    SegDesc *segs = nullptr;
    int32_t *rel = nullptr;
    int32_t *abase = nullptr;
    uint32_t *counter = nullptr;  // persistent work-queue counters, one per stage
    int32_t *plain_shift = nullptr;  // raw shift tables of every range, for the plain kernel
};

// A dedispersion launch over segments already resident on the device.
struct LaunchSpec {
    const SegDesc *segs = nullptr;  // device
    int32_t seg_begin = 0, nseg = 0;
    int64_t items = 0;
    int32_t lds_bytes = 0;
    uint32_t *counter = nullptr;    // persistent mode: device work counter (zeroed here)
    int32_t grid = 0;               // persistent mode: workgroups (0 = one per item)
    int32_t group = 1;              // tiles per item group
    int32_t dtypes = 0;             // bit d: blocks of DType d in the launch (0 = any)
    int32_t stats = 0;              // 1: blocks with stat_tiles > 0 write their row statistics
    int32_t njobs = 0;
    JobBufs jobs[kMaxJobs];
};

}  // namespace mdsm

struct mdsm_plan {
    mdsm_obs obs{};
    int device = 0;
    int cus = 0;            // compute units of the device
    bool uploaded = false;  // device tables are uploaded on first use
    bool subband = false;
    uint32_t nsubs = 0;
    std::vector<mdsm::RangeHost> ranges;
    std::vector<mdsm::DecimStep> decim;
    std::vector<mdsm::BlockDesc> blocks;
    std::vector<mdsm::SegDesc> segs;
    std::vector<int32_t> rel0;         // per-(row, trial) phase offsets, row place not included
    std::vector<int32_t> rel, abase;   // rel = rel0 + (row mod chunk rows) * row bytes (device copy)
    std::vector<mdsm::Stage> stages;   // launched in order
    mdsm::KernelTables dev{};
    mdsm_tuning tune{};
    std::string shape_key;             // autotuner cache key
    void *own_work = nullptr;
    uint64_t work_bytes = 0;
    uint64_t stat_bytes = 0;           // row-statistics partials after the workspace (ranges' stat_off)
    mdsm_plan_info info{};
};

namespace mdsm {

int upload_plan(mdsm_plan *p);
// Row statistics of every uint8 full-rate range over its rows' first n
// samples, written by launches with LaunchSpec::stats (plan.cpp).
int plan_row_stats(mdsm_plan *p, uint32_t n);

// Switch the manager instance (initialiseMDSM) to device input: from now on
// start_processing takes each block's new samples from d_in (rows of
// `stride` floats; block 0 at column 0, later blocks at column maxshift),
// written by the caller on `stream` before the call, instead of the host
// input buffer.
int compat_device_input(float **d_in, int64_t *stride, hipStream_t *stream);

// Launch-shape defaults and the autotuner's search (plan.cpp): `apply` sets a
// candidate, `step` enqueues one timed run on stream s.
mdsm_tuning default_tuning();
int tune_search(const mdsm_tuning &start, const std::function<int(const mdsm_tuning &)> &apply,
                const std::function<int()> &step, hipStream_t s, uint32_t reps,
                mdsm_tuning *best_out);

// Error plumbing (capi.cpp).
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);

// Timing hooks around launches (capi.cpp).  timing_swap sets the switch and
// returns its previous state.
bool timing_swap(bool on);
struct TimedLaunch {
    TimedLaunch(const char *name, hipStream_t s);
    ~TimedLaunch();
    const char *name;
    hipStream_t stream;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

// Launchers (kernels.hip).
hipError_t launch_copy2d(void *dst, int64_t dpitch, const void *src, int64_t spitch,
                         int64_t width, uint32_t height, hipStream_t s);
hipError_t launch_zero2d(void *dst, int64_t dpitch, int64_t width, uint32_t height, hipStream_t s);
hipError_t launch_plain(const float *in, int64_t in_stride, uint32_t nchans,
                        const int32_t *shift, uint32_t ndms, uint32_t nout, float *out,
                        hipStream_t s);
hipError_t launch_dedisperse(const mdsm_plan *p, const Stage &st, const void *in,
                             int64_t in_stride, float *out, void *work, hipStream_t s);
hipError_t launch_dedisperse_spec(const LaunchSpec &L, hipStream_t s);
// launch_dedisperse with the persistent work-queue counter given (one uint32
// per launch, zeroed by the launch) instead of the plan's.
hipError_t launch_dedisperse_ctr(const mdsm_plan *p, const Stage &st, const void *in,
                                 int64_t in_stride, float *out, void *work, uint32_t *counter,
                                 hipStream_t s);
// Resident workgroups per CU (and waves per SIMD) of the dedispersion kernel
// a launch with this dynamic LDS and these dtypes takes, with its registers.
struct DedispOccupancy {
    int32_t workgroups_per_cu, waves_per_simd, vgprs, static_lds, dynamic_lds, scratch_bytes;
};
hipError_t dedisperse_occupancy(int32_t lds_bytes, int32_t dtypes, DedispOccupancy *o);
hipError_t launch_decimate(const void *in, int dtype_in, uint32_t nchans, int64_t in_stride,
                           uint32_t nsamp, uint32_t k, void *out, int dtype_out,
                           int64_t out_stride, hipStream_t s);
// Every decimation of a plan (steps[0..n)), reading `in` once for the
// uint8 -> uint16 steps with k in {2, 4, 8, 16}; outputs at work + work_off.
hipError_t launch_decimate_steps(const void *in, uint32_t nchans, int64_t in_stride,
                                 uint32_t nsamp, const DecimStep *steps, size_t nsteps, char *work,
                                 hipStream_t s);
hipError_t launch_clip_u8(const uint8_t *in, int64_t in_stride, uint32_t nchans, int64_t t0,
                          uint32_t n, float nsigma, uint8_t *out, int64_t out_stride,
                          int32_t *med, int32_t *mad, int32_t *ncl, hipStream_t s);
struct SqrtW {
    double v[11];  // sqrt(2^i), i = 0..10, computed on the host
};
hipError_t launch_fill_u32(uint32_t *p, uint32_t n, uint32_t v, hipStream_t s);
// Streaming boxcar search: the searched series of row d is H carried samples
// carry[d*carry_stride + 0..H) followed by the block's n samples; windows
// start at t >= H - w (t >= 0), and tail = 1 leaves the window without a
// right neighbour to the next block (H = 0, tail = 0: one block on its own).
// Candidate t is relative to the block's first sample.
struct BoxStream {
    const float *carry;
    int64_t carry_stride;
    int32_t H;
    int32_t tail;
};
uint32_t boxcar_maxw(uint32_t wmask);
// Carried samples per row a stream with these widths needs: maxw + 1, rounded
// up to a multiple of 4 (so that 16-byte loads never straddle the two sources).
uint32_t boxcar_carry_len(uint32_t wmask);
// stats = false: mean/sigma already hold the row statistics of the block.
// int_exact: the caller guarantees the series (carry included) are integers
// whose sums over maxw samples stay below 2^24 (then the tree runs exactly in
// float32; same candidates).
hipError_t launch_boxcar_stream(const float *x, uint32_t ndms, uint32_t n, int64_t stride,
                                const BoxStream &bs, uint32_t wmask, double thr,
                                const SqrtW &sqrtw, mdsm_cand *cands, uint32_t cap,
                                uint32_t *count, double *mean, double *sigma, bool stats,
                                hipStream_t s, bool int_exact = false);
// mean/sigma of ndms rows of n samples from the dedispersion launch's
// per-tile partials (tiles (sum, sum of squares) FP64 pairs per row).
hipError_t launch_row_partials(const void *part, uint32_t ndms, uint32_t tiles, uint32_t n,
                               double *mean, double *sigma, hipStream_t s);
hipError_t launch_boxcar(const float *x, uint32_t ndms, uint32_t n, int64_t stride,
                         uint32_t wmask, double thr, const SqrtW &sqrtw, mdsm_cand *cands,
                         uint32_t cap, uint32_t *count, double *mean, double *sigma,
                         hipStream_t s);
// Chunks the brute-force block statistics read (mf: values per chunk).
uint32_t brute_stat_chunks(uint32_t ndms, uint32_t n, uint64_t *mf);
hipError_t launch_median5(const float *x, uint32_t ndms, uint32_t n, int64_t stride,
                          const float *dmv, double thr, int mode, const float *cmean,
                          const float *crms, uint32_t spc, mdsm_cand *cands, uint32_t cap,
                          uint32_t *count, double *stats, hipStream_t s);

}  // namespace mdsm
