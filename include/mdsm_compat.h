/*
 * mdsm_compat.h -- drop-in replacement of MDSM's dedispersion manager API
 * (src/mdsm/dedispersion_manager.h:7-12 and src/mdsm/survey.h), backed by
 * libmdsm_hip.so.
 *
 * Same five calls, same buffer ownership, same two-block output latency as
 * the reference's three-stage pipeline; differences:
 *   - processSurveyParameters takes a const char* path (no Qt) and returns
 *     NULL on failure (mdsm_last_error() has the reason) instead of throwing
 *     a QString;
 *   - errors never exit the process; initialiseMDSM returns NULL,
 *     next_chunk returns NULL with samples = 0, start_processing returns 0.
 * One instance per process, not re-entrant -- as in the reference.
 */
#ifndef MDSM_COMPAT_H
#define MDSM_COMPAT_H

#include <stdint.h>
#include <stdio.h>

#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* src/mdsm/survey.h:6-11, field for field. */
typedef struct {
    float lowdm, highdm, dmstep, sub_dmstep;
    int binsize, ndms, ncalls, calldms, mean, stddev;
} SUBBAND_PASSES;

/* src/mdsm/survey.h:13-57, field for field. */
typedef struct {
    unsigned int nsamp, nchans, tdms, maxshift, nbits;
    float tsamp, foff, fch1;
    bool useBruteForce;
    bool useL1Cache;
    float lowdm, dmstep;
    SUBBAND_PASSES *pass_parameters;
    unsigned num_passes, nsubs;
    double timestamp;
    double blockRate;
    float *noiseMean;
    float *noiseRMS;
    unsigned samplesPerChunk;
    FILE *fp;
    char fileprefix[80], basedir[120];
    unsigned secs_per_file;
    char use_pc_time, single_file_mode;
    float detection_threshold;
    unsigned num_threads;
    unsigned *gpu_ids;
    unsigned num_gpus;
} SURVEY;

/* Parse an MDSM observation XML file (schema of
 * src/mdsm/dedispersion_manager.cpp:40-199: <frequencies top offset>,
 * <dm lowDM numDMs dmStep useBruteForce useL1Cache>, <channels number
 * subbands>, <timing tsamp>, <samples number>, <detection threshold>,
 * <output filePrefix baseDirectory secondsPerFile usePCTime singleFileMode>,
 * <gpus ids>, <passes><pass><lowDm>...</pass></passes>).  Tag names are
 * matched case-insensitively; missing attributes read as 0; defaults as the
 * reference.  The SURVEY is malloc'ed; free it with mdsm_free_survey. */
SURVEY *processSurveyParameters(const char *filepath);

/* Same, from an XML string. */
SURVEY *mdsm_parse_survey_xml(const char *xml, size_t len);
void mdsm_free_survey(SURVEY *survey);

/* Set up the plan, device buffers and the library-owned host input buffer
 * [nchans][nsamp + maxshift] float32, which is returned.  Sets
 * survey->nsamp (when 0: sized from device memory), survey->maxshift and
 * survey->num_threads (= 1 device).
 * (src/mdsm/dedispersion_manager.cpp:269-382) */
float *initialiseMDSM(SURVEY *input_survey);

/* Publish the block just written into the input buffer (data_read new
 * samples per channel; 0 = end of stream) and return the dedispersed output
 * of the block two calls earlier ([tdms][nsamp] for brute force; passes
 * concatenated, nsamp/binsize samples per DM, for subband), or NULL while
 * fewer than two blocks have been processed.  *samples receives that
 * block's sample count.  The buffer stays valid until the next call.
 * Candidates of the returned block are appended to the survey's output
 * file in the reference's text format.
 * (src/mdsm/dedispersion_manager.cpp:416-484) */
float *next_chunk(unsigned int data_read, unsigned *samples, double timestamp,
                  double blockRate, float *mean, float *rms);

/* Start processing the published block; the input buffer may be rewritten
 * as soon as this returns (block k >= 1: samples [maxshift, maxshift+nsamp)
 * of each channel; the overlap is kept on the device).  Returns the block
 * count, or 0 once the stream has drained (second call with data_read == 0).
 * (src/mdsm/dedispersion_manager.cpp:487-501) */
int start_processing(unsigned int data_read);

/* Release everything.  (src/mdsm/dedispersion_manager.cpp:385-413) */
void tearDownMDSM(void);

/* Device output of the last processed block (for callers that stay on the
 * GPU): packed plan layout, valid until the next start_processing. */
const float *mdsm_compat_device_output(void);

#ifdef __cplusplus
}

/* C++ callers of the reference pass `samples` by reference. */
inline float *next_chunk(unsigned int data_read, unsigned &samples, double timestamp = 0,
                         double blockRate = 0, float *mean = NULL, float *rms = NULL)
{
    return next_chunk(data_read, &samples, timestamp, blockRate, mean, rms);
}
#endif

#endif /* MDSM_COMPAT_H */
