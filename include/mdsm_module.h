/*
 * mdsm_module.h -- the Pelican MdsmModule (src/pelican-mdsm/modules/
 * MdsmModule.h:16-45, MdsmModule.cpp:19-257) without Pelican: the adapter
 * that turns a stream of spectra blobs into MDSM blocks, with the corner
 * turn on the GPU and the per-chunk mean/rms plumbing of weighted blobs.
 * It drives the manager of mdsm_compat.h, so, like the reference module,
 * there is one per process.
 *
 * A blob is what MdsmModule::run reads from SpectrumDataSetStokes /
 * WeightedSpectrumDataSet: ntime spectra of the first Stokes parameter, each
 * nsubbands x nchannels floats (spectrumData(t, s, 0)), plus, for a weighted
 * blob, the blob's mean and rms.  The output is what run() puts in its
 * DedispersedTimeSeriesF32: one series per DM trial, or none.
 *
 * Differences from the reference:
 *   - the blob is copied time-major into a pinned staging block; when a block
 *     is complete it is uploaded and corner-turned on the GPU straight into
 *     the device input rows (channel inversion included), where the
 *     reference corner-turns sample by sample into the host input buffer
 *     (MdsmModule.cpp:135-151);
 *   - the samples of a blob that straddles two blocks go to the start of the
 *     next block; the reference writes them at maxshift + t instead of
 *     maxshift + t - copySamp (MdsmModule.cpp:238-249), which leaves stale
 *     columns and overlaps the next blob -- with blocks that are multiples of
 *     the blob length the two agree;
 *   - once the stream has drained the reference returns before filling the
 *     output (MdsmModule.cpp:168), which drops the last block; it is
 *     delivered here;
 *   - a blob longer than a block (ntime > nsamp) is rejected;
 *   - a series points at its row of the manager's output buffer (valid until
 *     the next run) instead of a copy;
 *   - errors return a status; the reference throws a QString.
 */
#ifndef MDSM_MODULE_H_
#define MDSM_MODULE_H_

#include <stdint.h>

#include "mdsm_compat.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mdsm_module mdsm_module;

/* The reference's configuration options (MdsmModule.cpp:23-27):
 * observationfile/filepath, createOutputBlob/value (reference default 0),
 * invertChannels/value (reference default 1).  Parses the observation file
 * and starts the manager (processSurveyParameters + initialiseMDSM).
 * Replaces MdsmModule::MdsmModule. */
int mdsm_module_create(const char *observation_file, int create_output_blob,
                       int invert_channels, mdsm_module **module);

/* The DedispersedTimeSeriesF32 of one run: nseries == 0 is resize(0). */
typedef struct {
    uint32_t nseries;
    const float *dm_values;   /* [nseries], setDmValue of each series */
    const uint32_t *nsamp;    /* [nseries] samples of each series */
    const float *const *data; /* [nseries] pointers to each series' samples */
} mdsm_module_output;

/* One blob: spectrum t, subband s starts at spectra + t*t_stride + s*s_stride
 * (floats) and holds nchannels floats; nsubbands*nchannels must equal the
 * observation's channel count.  ntime == 0 is the end-of-stream blob.
 * weighted != 0 marks a WeightedSpectrumDataSet with the given mean and rms.
 * timestamp and block_rate are the blob's getLofarTimestamp() and
 * getBlockRate().  *out describes the output, owned by the module.
 * Replaces MdsmModule::run. */
int mdsm_module_run(mdsm_module *module, const float *spectra, uint32_t ntime,
                    uint32_t nsubbands, uint32_t nchannels, int64_t t_stride,
                    int64_t s_stride, int weighted, float mean, float rms,
                    double timestamp, double block_rate, mdsm_module_output *out);

/* The observation the module runs, owned by the module. */
const SURVEY *mdsm_module_survey(const mdsm_module *module);

/* Stops the manager (tearDownMDSM) and frees the module.  Replaces
 * MdsmModule::~MdsmModule. */
void mdsm_module_destroy(mdsm_module *module);

#ifdef __cplusplus
}
#endif

#endif /* MDSM_MODULE_H_ */
