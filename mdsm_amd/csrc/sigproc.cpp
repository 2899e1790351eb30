// SIGPROC filterbank files on either side of the path: the header and raw
// spectra reader (the unpack to samples runs on the GPU, prep.hip) and the
// dedispersed-series writer (one float32 SIGPROC file per selected DM).
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mdsm_internal.h"

using namespace mdsm;

struct mdsm_sigproc_reader {
    FILE *f = nullptr;
    mdsm_sigproc_header h{};
    uint64_t spectrum_bytes = 0;
};

struct mdsm_series_writer {
    std::vector<FILE *> files;
    std::vector<float> dms;
};

namespace {

// One length-prefixed header string (int32 length, 1..80 bytes).
int read_string(FILE *f, std::string &s)
{
    int32_t n = 0;
    if (std::fread(&n, sizeof n, 1, f) != 1) return fail(MDSM_ERR_PARSE, "SIGPROC header truncated");
    if (n < 1 || n > 80) return fail(MDSM_ERR_PARSE, "SIGPROC header string length out of range");
    char buf[81];
    if (std::fread(buf, 1, (size_t)n, f) != (size_t)n)
        return fail(MDSM_ERR_PARSE, "SIGPROC header truncated");
    buf[n] = '\0';
    s.assign(buf, (size_t)n);
    return MDSM_OK;
}

template <typename T>
int read_value(FILE *f, T *v)
{
    if (std::fread(v, sizeof(T), 1, f) != 1) return fail(MDSM_ERR_PARSE, "SIGPROC header truncated");
    return MDSM_OK;
}

// Keywords and value types of the SIGPROC header as the reference reads them
// (src/mdsm/file_handler.cpp:56-177).
int parse_header(FILE *f, mdsm_sigproc_header *h)
{
    std::memset(h, 0, sizeof *h);
    std::string s;
    int rc = read_string(f, s);
    if (rc) return rc;
    if (s != "HEADER_START") return fail(MDSM_ERR_PARSE, "not a SIGPROC file (no HEADER_START)");
    bool want_raw = false, want_source = false;
    for (;;) {
        if ((rc = read_string(f, s))) return rc;
        if (s == "HEADER_END") break;
        if (want_raw) {
            std::snprintf(h->rawdatafile, sizeof h->rawdatafile, "%s", s.c_str());
            want_raw = false;
        } else if (want_source) {
            std::snprintf(h->source_name, sizeof h->source_name, "%s", s.c_str());
            want_source = false;
        } else if (s == "rawdatafile") {
            want_raw = true;
        } else if (s == "source_name") {
            want_source = true;
        } else if (s == "FREQUENCY_START" || s == "FREQUENCY_END") {
        } else if (s == "fchannel") {
            double v;
            if ((rc = read_value(f, &v))) return rc;
        } else if (s == "Telescope") {
            std::string name;
            if ((rc = read_string(f, name))) return rc;
        } else if (s == "az_start") rc = read_value(f, &h->az_start);
        else if (s == "za_start") rc = read_value(f, &h->za_start);
        else if (s == "src_raj") rc = read_value(f, &h->src_raj);
        else if (s == "src_dej") rc = read_value(f, &h->src_dej);
        else if (s == "tstart") rc = read_value(f, &h->tstart);
        else if (s == "tsamp") rc = read_value(f, &h->tsamp);
        else if (s == "period") rc = read_value(f, &h->period);
        else if (s == "fch1") rc = read_value(f, &h->fch1);
        else if (s == "foff") rc = read_value(f, &h->foff);
        else if (s == "refdm") rc = read_value(f, &h->refdm);
        else if (s == "nchans") rc = read_value(f, &h->nchans);
        else if (s == "telescope_id") rc = read_value(f, &h->telescope_id);
        else if (s == "machine_id") rc = read_value(f, &h->machine_id);
        else if (s == "data_type") rc = read_value(f, &h->data_type);
        else if (s == "ibeam") rc = read_value(f, &h->ibeam);
        else if (s == "nbeams") rc = read_value(f, &h->nbeams);
        else if (s == "nbits") rc = read_value(f, &h->nbits);
        else if (s == "barycentric") rc = read_value(f, &h->barycentric);
        else if (s == "pulsarcentric") rc = read_value(f, &h->pulsarcentric);
        else if (s == "nbins") rc = read_value(f, &h->nbins);
        else if (s == "nsamples") rc = read_value(f, &h->nsamples);
        else if (s == "nifs") rc = read_value(f, &h->nifs);
        else if (s == "npuls") rc = read_value(f, &h->npuls);
        else
            return fail(MDSM_ERR_PARSE, "SIGPROC header: unknown parameter " + s);
        if (rc) return rc;
    }
    const long pos = std::ftell(f);
    if (pos < 0) return fail(MDSM_ERR_PARSE, "SIGPROC: ftell failed");
    h->header_bytes = (uint64_t)pos;
    return MDSM_OK;
}

int spectrum_bytes(const mdsm_sigproc_header &h, uint64_t *out)
{
    const int nb = h.nbits;
    if (nb != 1 && nb != 2 && nb != 4 && nb != 8 && nb != 16 && nb != 32)
        return fail(MDSM_ERR_PARSE, "SIGPROC: nbits must be 1, 2, 4, 8, 16 or 32");
    if (h.nchans < 1) return fail(MDSM_ERR_PARSE, "SIGPROC: nchans < 1");
    const uint64_t nifs = h.nifs > 0 ? (uint64_t)h.nifs : 1;
    const uint64_t bits = (uint64_t)h.nchans * nifs * (uint64_t)nb;
    if (bits % 8) return fail(MDSM_ERR_PARSE, "SIGPROC: a spectrum is not a whole number of bytes");
    *out = bits / 8;
    return MDSM_OK;
}

void write_string(FILE *f, const char *s)
{
    const int32_t n = (int32_t)std::strlen(s);
    std::fwrite(&n, sizeof n, 1, f);
    std::fwrite(s, 1, (size_t)n, f);
}

void write_int(FILE *f, const char *k, int32_t v)
{
    write_string(f, k);
    std::fwrite(&v, sizeof v, 1, f);
}

void write_double(FILE *f, const char *k, double v)
{
    write_string(f, k);
    std::fwrite(&v, sizeof v, 1, f);
}

}  // namespace

extern "C" int mdsm_sigproc_read_header(const char *path, mdsm_sigproc_header *h)
{
    if (!path || !h) return fail(MDSM_ERR_INVALID, "NULL argument");
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(MDSM_ERR_INVALID, std::string("cannot open ") + path + ": " + strerror(errno));
    int rc = parse_header(f, h);
    uint64_t sb = 0;
    if (!rc) rc = spectrum_bytes(*h, &sb);
    if (!rc && std::fseek(f, 0, SEEK_END) == 0) {
        const long end = std::ftell(f);
        if (end >= (long)h->header_bytes) h->nspectra = ((uint64_t)end - h->header_bytes) / sb;
    }
    std::fclose(f);
    return rc;
}

extern "C" int mdsm_sigproc_open(const char *path, mdsm_sigproc_reader **out, mdsm_sigproc_header *h)
{
    if (!path || !out) return fail(MDSM_ERR_INVALID, "NULL argument");
    *out = nullptr;
    auto *r = new mdsm_sigproc_reader();
    int rc = mdsm_sigproc_read_header(path, &r->h);
    if (!rc) rc = spectrum_bytes(r->h, &r->spectrum_bytes);
    if (!rc) {
        r->f = std::fopen(path, "rb");
        if (!r->f || std::fseek(r->f, (long)r->h.header_bytes, SEEK_SET) != 0)
            rc = fail(MDSM_ERR_INVALID, std::string("cannot open ") + path);
    }
    if (rc) {
        if (r->f) std::fclose(r->f);
        delete r;
        return rc;
    }
    if (h) *h = r->h;
    *out = r;
    return MDSM_OK;
}

extern "C" int mdsm_sigproc_read(mdsm_sigproc_reader *r, uint64_t nspectra, void *host_raw,
                                 uint64_t *got)
{
    if (!r || !host_raw || !got) return fail(MDSM_ERR_INVALID, "NULL argument");
    const size_t n = std::fread(host_raw, (size_t)r->spectrum_bytes, (size_t)nspectra, r->f);
    *got = n;
    if (n < nspectra && std::ferror(r->f)) return fail(MDSM_ERR_INVALID, "SIGPROC read error");
    return MDSM_OK;
}

extern "C" int mdsm_sigproc_seek(mdsm_sigproc_reader *r, uint64_t spectrum)
{
    if (!r) return fail(MDSM_ERR_INVALID, "NULL argument");
    const uint64_t off = r->h.header_bytes + spectrum * r->spectrum_bytes;
    if (std::fseek(r->f, (long)off, SEEK_SET) != 0) return fail(MDSM_ERR_INVALID, "SIGPROC seek failed");
    return MDSM_OK;
}

extern "C" void mdsm_sigproc_close(mdsm_sigproc_reader *r)
{
    if (!r) return;
    if (r->f) std::fclose(r->f);
    delete r;
}

extern "C" int mdsm_series_writer_open(const char *prefix, const float *dm_values, uint32_t ndm,
                                       double fch1, double foff, double tsamp, double tstart,
                                       mdsm_series_writer **out)
{
    if (!prefix || !out || (ndm && !dm_values)) return fail(MDSM_ERR_INVALID, "NULL argument");
    *out = nullptr;
    auto *w = new mdsm_series_writer();
    for (uint32_t i = 0; i < ndm; ++i) {
        char path[1024];
        std::snprintf(path, sizeof path, "%s_%g.fil", prefix, (double)dm_values[i]);
        FILE *f = std::fopen(path, "wb");
        if (!f) {
            for (FILE *g : w->files) std::fclose(g);
            delete w;
            return fail(MDSM_ERR_INVALID, std::string("cannot create ") + path);
        }
        write_string(f, "HEADER_START");
        write_int(f, "machine_id", 0);
        write_int(f, "telescope_id", 0);
        write_int(f, "data_type", 1);
        write_double(f, "fch1", fch1);
        write_double(f, "foff", foff);
        write_int(f, "nchans", 1);
        write_double(f, "tsamp", tsamp);
        write_int(f, "nbits", 32);
        write_double(f, "tstart", tstart);
        write_int(f, "nifs", 1);
        write_string(f, "HEADER_END");
        std::fflush(f);
        w->files.push_back(f);
        w->dms.push_back(dm_values[i]);
    }
    *out = w;
    return MDSM_OK;
}

extern "C" int mdsm_series_writer_write(mdsm_series_writer *w, const float *host_rows,
                                        const float *row_dm_values, uint32_t nrows, uint32_t nsamp,
                                        int64_t stride)
{
    if (!w || (nrows && (!host_rows || !row_dm_values))) return fail(MDSM_ERR_INVALID, "NULL argument");
    if (stride < (int64_t)nsamp) return fail(MDSM_ERR_INVALID, "stride < nsamp");
    for (size_t i = 0; i < w->files.size(); ++i)
        for (uint32_t j = 0; j < nrows; ++j)
            if (std::fabs(row_dm_values[j] - w->dms[i]) < 0.00001f) {
                const float *row = host_rows + (int64_t)j * stride;
                if (std::fwrite(row, sizeof(float), nsamp, w->files[i]) != nsamp)
                    return fail(MDSM_ERR_INVALID, "series write failed");
                std::fflush(w->files[i]);
                break;
            }
    return MDSM_OK;
}

extern "C" void mdsm_series_writer_close(mdsm_series_writer *w)
{
    if (!w) return;
    for (FILE *f : w->files) std::fclose(f);
    delete w;
}
