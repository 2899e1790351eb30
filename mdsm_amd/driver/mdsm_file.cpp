// mdsm_file -- the standalone file pipeline of the reference's src/MDSM.cpp
// (:19-60 argument handling, :100-170 read loop) on libmdsm_hip.so, without
// Qt or Pelican:
//
//     mdsm_file <data.fil> -obs <observation.xml> [-nsamp N]
//
// A SIGPROC filterbank is read spectrum-block by block (mdsm_sigproc_*; the
// reference's read_header/read_block), unpacked and corner-turned on the GPU
// (mdsm_unpack_spectra: 1/2/4/8-bit, 16-bit and float samples into
// channel-major float rows; the reference does this on the host in
// readBinaryData) into the manager's input buffer, and driven through the
// manager calls exactly as MDSM.cpp does: processSurveyParameters ->
// initialiseMDSM -> loop { next_chunk(data_read, unsigned&, timestamp,
// blockRate); start_processing(data_read) } -> tearDownMDSM.  The manager
// writes the candidate file (the survey's output prefix and directory).
// It is also the link check of the C++ next_chunk overload of mdsm_compat.h.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "mdsm_c.h"
#include "mdsm_compat.h"

namespace {

int die(const char *what)
{
    std::fprintf(stderr, "mdsm_file: %s: %s\n", what, mdsm_last_error());
    return 1;
}

}  // namespace

int main(int argc, char *argv[])
{
    if (argc < 4 || std::strcmp(argv[2], "-obs") != 0) {
        std::fprintf(stderr, "usage: mdsm_file <data.fil> -obs <observation.xml> [-nsamp N]\n");
        return 2;
    }
    SURVEY *survey = processSurveyParameters(argv[3]);
    if (!survey) return die("observation file");
    mdsm_sigproc_reader *reader = nullptr;
    mdsm_sigproc_header hdr;
    if (mdsm_sigproc_open(argv[1], &reader, &hdr)) return die("data file");
    survey->nbits = (unsigned)hdr.nbits;
    for (int i = 4; i < argc; ++i) {
        if (!std::strcmp(argv[i], "-nsamp") && i + 1 < argc) {
            survey->nsamp = (unsigned)std::atoi(argv[++i]);
        } else {
            std::fprintf(stderr, "mdsm_file: invalid parameter %s\n", argv[i]);
            return 2;
        }
    }
    if ((unsigned)hdr.nchans != survey->nchans) {
        std::fprintf(stderr, "mdsm_file: file has %d channels, observation %u\n", hdr.nchans,
                     survey->nchans);
        return 2;
    }

    float *input = initialiseMDSM(survey);
    if (!input) return die("initialiseMDSM");
    const unsigned C = survey->nchans, N = survey->nsamp, Mx = survey->maxshift;
    const uint64_t row = (uint64_t)N + Mx;  // input row: [maxshift overlap | N new samples]
    const uint64_t spec_bytes = (uint64_t)C * hdr.nbits / 8;
    std::printf("maxshift %u nsamp %u\n", Mx, N);

    // staging: raw spectra (host), raw and unpacked block (device)
    std::vector<unsigned char> raw((size_t)(N + Mx) * spec_bytes);
    void *d_raw = nullptr, *d_rows = nullptr;
    if (mdsm_device_malloc(&d_raw, raw.size()) ||
        mdsm_device_malloc(&d_rows, (uint64_t)C * row * sizeof(float)))
        return die("device buffers");

    unsigned counter = 0, total = 0;
    for (;;) {
        // MDSM.cpp:139-153: the first read also fills the overlap
        const unsigned to_read = counter == 0 ? N + Mx : N;
        const unsigned col0 = counter == 0 ? 0 : Mx;
        uint64_t got = 0;
        if (mdsm_sigproc_read(reader, to_read, raw.data(), &got)) return die("read");
        if (got) {
            if (mdsm_memcpy(d_raw, raw.data(), got * spec_bytes, MDSM_H2D, nullptr) ||
                mdsm_unpack_spectra(d_raw, hdr.nbits, (uint32_t)got, C, (int64_t)spec_bytes, 0,
                                    d_rows, MDSM_F32, (int64_t)row, col0, nullptr) ||
                mdsm_memcpy_2d(input + col0, row * sizeof(float), (char *)d_rows + col0 * 4,
                               row * sizeof(float), got * sizeof(float), C, MDSM_D2H, nullptr))
                return die("unpack");
        }
        unsigned data_read = (unsigned)got;
        if (counter == 0) {
            if (data_read < Mx) {
                std::fprintf(stderr, "mdsm_file: not enough samples in file to dedisperse\n");
                data_read = 0;
            } else {
                data_read -= Mx;
            }
        }
        unsigned samples = 0;  // the reference's C++ call: samples by reference
        next_chunk(data_read, samples, (double)survey->tsamp * N * counter, survey->tsamp);
        std::printf("block %u data_read %u\n", counter, data_read);
        if (!start_processing(data_read)) break;
        total += data_read;
        counter++;
    }
    std::printf("samples %u blocks %u\n", total, counter);
    tearDownMDSM();
    mdsm_device_free(d_raw);
    mdsm_device_free(d_rows);
    mdsm_sigproc_close(reader);
    mdsm_free_survey(survey);
    return 0;
}
