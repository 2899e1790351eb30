/*
 * ref_shim.cpp -- calls into the reference's own output-thread detector,
 * linked into oracle/_ref/libmdsm_ref.so.
 *
 * TEST INFRASTRUCTURE ONLY.  process_brute (src/mdsm/dedispersion_output.cpp)
 * is compiled unmodified from the reference's source; this shim fills the
 * reference's own SURVEY (src/mdsm/survey.h, included from the reference
 * tree) for one brute-force thread without noise arrays and writes its
 * candidate lines to `path`.
 */
#include <cstdio>
#include <cstring>
#include <ctime>

#include "survey.h"

void process_brute(float *buffer, FILE *output, SURVEY *survey, int read_nsamp, size_t size,
                   double timestamp, double blockRate, time_t start_time, float *mean,
                   float *rms);

extern "C" int mdsm_ref_process_brute(float *buffer, unsigned nsamp, unsigned tdms,
                                      int read_nsamp, float lowdm, float dmstep, float thr,
                                      double timestamp, double block_rate, const char *path)
{
    SURVEY sv;
    std::memset(&sv, 0, sizeof(sv));
    sv.nsamp = nsamp;
    sv.tdms = tdms;
    sv.useBruteForce = true;
    sv.lowdm = lowdm;
    sv.dmstep = dmstep;
    sv.detection_threshold = thr;
    sv.num_threads = 1;
    FILE *fp = std::fopen(path, "w");
    if (!fp) return -1;
    process_brute(buffer, fp, &sv, read_nsamp, (size_t)nsamp * tdms, timestamp, block_rate,
                  std::time(nullptr), nullptr, nullptr);
    std::fclose(fp);
    return 0;
}
