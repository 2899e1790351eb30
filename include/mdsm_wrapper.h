/*
 * mdsm_wrapper.h -- the reference's host/device ABI, for callers that keep
 * the reference's dedispersion manager (its pthreads, barriers, rw lock and
 * output thread) and link this library instead of its CUDA wrapper
 * (src/mdsm/dedispersion_wrapper.cu:4-13).  DEVICE_INFO, DEVICES and
 * THREAD_PARAMS are field for field those of src/mdsm/dedispersion_thread.h:9-57;
 * SURVEY is the one of mdsm_compat.h (src/mdsm/survey.h).
 *
 * Differences from the reference: errors are printed and never exit the
 * process (call_initialise_devices returns NULL; a worker that cannot set up
 * still takes part in every barrier and writes zeros); THREAD_PARAMS.dmshifts
 * is not read -- the worker derives the same delay table from the survey
 * (fch1, foff) by the manager's formula (dedispersion_manager.cpp:262-265,
 * 286-289); input rows keep the stride they were allocated with when a short
 * last subband block lowers survey->nsamp.
 */
#ifndef MDSM_WRAPPER_H
#define MDSM_WRAPPER_H

#include <pthread.h>
#include <stddef.h>
#include <time.h>

#include "mdsm_compat.h"

#ifdef __cplusplus
extern "C" {
#endif

/* src/mdsm/dedispersion_thread.h:9-20, field for field. */
typedef struct {
    int multiprocessor_count;
    int constant_memory;
    int shared_memory;
    int register_count;
    int thread_count;
    int clock_rate;
    int device_id;
} DEVICE_INFO;

/* src/mdsm/dedispersion_thread.h:22-29, field for field (minTotalGlobalMem in KiB). */
typedef struct {
    int num_devices;
    unsigned long int minTotalGlobalMem;
    DEVICE_INFO *devices;
} DEVICES;

/* src/mdsm/dedispersion_thread.h:31-57, field for field. */
typedef struct {
    int iterations, maxiters;
    int maxshift, binsize;
    float *dmshifts;
    size_t inputsize, outputsize, dm_output, dedispersed_size;
    float *output;
    float *input;
    int device_id;
    unsigned short thread_num, num_threads, stop;
    pthread_rwlock_t *rw_lock;
    pthread_barrier_t *input_barrier;
    pthread_barrier_t *output_barrier;
    time_t start;
    SURVEY *survey;
} THREAD_PARAMS;

/* Enumerate the GPUs with more than 2.5 GiB (the ones in survey->gpu_ids when
 * given) into a malloc'ed DEVICES (free devices->devices, then devices).
 * NULL when there is none.  Replaces initialise_devices
 * (src/mdsm/dedispersion_thread.cu:4-58). */
DEVICES *call_initialise_devices(SURVEY *survey);

/* One GPU's worker, run on its own thread by the manager: per iteration,
 * upload the input block ([nchans][nsamp + maxshift] float32, the previous
 * block's last maxshift samples carried to the front of each row on the host,
 * as the reference does), wait the input barrier, dedisperse this thread's
 * share of the trials (tdms/num_threads brute-force trials from
 * lowdm + dmstep*tdms/num_threads*thread_num; subband: ncalls/num_threads
 * calls of every pass), wait the output barrier, download
 * [trials][nsamp/binsize] rows into params->output; stop as the reference's
 * stopping clause.  Returns thread_params.  Replaces dedisperse
 * (src/mdsm/dedispersion_thread.cu:246-401). */
void *call_dedisperse(void *thread_params);

#ifdef __cplusplus
}
#endif

#endif /* MDSM_WRAPPER_H */
