// sigproc.cpp -- SIGPROC filterbank input on the host: the FilterbankReader
// that transientx_fil opens, reads the header of and reads block by block
// (src/singlepulse/dedisperse_fil.cpp:167-195, 214-217, 278-283, 324), and a
// header writer for files of the same form.  The reader lives in XLibs, which
// the TransientX tree does not hold, so this follows the public SIGPROC
// format as declared in include/txgpu.h.  No device code: the blocks it reads
// go to tx_beam_push like any host block.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/txgpu.h"
#include "sigproc.h"

namespace txgpu {
int set_error(int code, const std::string &msg);
}
using txgpu::set_error;

namespace {

struct FilFile {
    std::string path;
    tx_fil_header h{};
    std::vector<double> freqs;
};

bool rd(FILE *f, void *p, size_t n) { return std::fread(p, 1, n, f) == n; }

// a length-prefixed string of at most maxlen bytes; false on a short file or
// a length outside [0, maxlen]
bool read_str(FILE *f, std::string &s, int maxlen)
{
    int32_t n = 0;
    if (!rd(f, &n, 4) || n < 0 || n > maxlen) return false;
    s.assign((size_t)n, '\0');
    return n == 0 || rd(f, &s[0], (size_t)n);
}

bool copy_name(char *dst, const std::string &v)
{
    if (v.size() >= 128) return false;
    std::memcpy(dst, v.data(), v.size());
    dst[v.size()] = '\0';
    return true;
}

struct FileCloser {
    FILE *f;
    ~FileCloser()
    {
        if (f) std::fclose(f);
    }
};

int parse_header(const std::string &path, tx_fil_header *h, std::vector<double> &freqs)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return set_error(TX_EIO, "cannot open " + path + ": " + std::strerror(errno));
    FileCloser fc{f};
    *h = tx_fil_header{};
    h->nifs = 1;
    h->nbeams = 1;
    const std::string bad = path + ": not a SIGPROC filterbank header";
    std::string key, sval;
    if (!read_str(f, key, 64) || key != "HEADER_START") return set_error(TX_EINVAL, bad + " (no HEADER_START)");
    bool have_nbits = false, have_nchans = false, have_tsamp = false, have_fch1 = false, have_foff = false;
    bool in_freqs = false;
    int8_t sign = 0;
    std::vector<double> fchannel;
    int32_t skip_i = 0;
    // keywords with a 32-bit integer or a 64-bit double value, and where it goes
    struct IntKey {
        const char *key;
        int32_t *v;
        bool *seen;
    } ints[] = {{"telescope_id", &h->telescope_id, nullptr}, {"machine_id", &h->machine_id, nullptr},
                {"data_type", &h->data_type, nullptr},       {"barycentric", &h->barycentric, nullptr},
                {"pulsarcentric", &h->pulsarcentric, nullptr}, {"nbits", &h->nbits, &have_nbits},
                {"nchans", &h->nchans, &have_nchans},        {"nifs", &h->nifs, nullptr},
                {"nbeams", &h->nbeams, nullptr},             {"ibeam", &h->ibeam, nullptr},
                {"nsamples", &skip_i, nullptr}}; // the file's size decides nsamples
    struct DblKey {
        const char *key;
        double *v;
        bool *seen;
    } dbls[] = {{"tstart", &h->tstart, nullptr},   {"tsamp", &h->tsamp, &have_tsamp},
                {"fch1", &h->fch1, &have_fch1},    {"foff", &h->foff, &have_foff},
                {"refdm", &h->refdm, nullptr},     {"az_start", &h->az_start, nullptr},
                {"za_start", &h->za_start, nullptr}, {"src_raj", &h->src_raj, nullptr},
                {"src_dej", &h->src_dej, nullptr}, {"period", &h->period, nullptr}};
    for (long item = 0;; item++) {
        if (item > (1L << 22)) return set_error(TX_EINVAL, bad + " (no HEADER_END)");
        if (!read_str(f, key, 64)) return set_error(TX_EINVAL, bad + " (truncated)");
        if (key == "HEADER_END") break;
        const std::string trunc = bad + " (truncated value of " + key + ")";
        bool known = false;
        for (const IntKey &k : ints)
            if (key == k.key) {
                if (!rd(f, k.v, 4)) return set_error(TX_EINVAL, trunc);
                if (k.seen) *k.seen = true;
                known = true;
            }
        for (const DblKey &k : dbls)
            if (key == k.key) {
                if (!rd(f, k.v, 8)) return set_error(TX_EINVAL, trunc);
                if (k.seen) *k.seen = true;
                known = true;
            }
        if (known) continue;
        if (key == "signed") {
            if (!rd(f, &sign, 1)) return set_error(TX_EINVAL, trunc);
        } else if (key == "FREQUENCY_START" || key == "FREQUENCY_END") {
            in_freqs = key == "FREQUENCY_START";
        } else if (key == "fchannel") {
            double fc = 0.;
            if (!in_freqs) return set_error(TX_EINVAL, bad + " (fchannel outside FREQUENCY_START/END)");
            if (!rd(f, &fc, 8)) return set_error(TX_EINVAL, trunc);
            fchannel.push_back(fc);
        } else if (key == "source_name" || key == "rawdatafile") {
            if (!read_str(f, sval, 4096)) return set_error(TX_EINVAL, trunc);
            if (!copy_name(key == "source_name" ? h->source_name : h->rawdatafile, sval))
                return set_error(TX_EINVAL, path + ": " + key + " longer than 127 bytes");
        } else {
            return set_error(TX_EINVAL, path + ": unknown SIGPROC header keyword '" + key +
                                            "' (the size of its value is not known)");
        }
    }
    const long hb = std::ftell(f);
    if (hb < 0 || std::fseek(f, 0, SEEK_END) != 0) return set_error(TX_EIO, "cannot seek in " + path);
    const long size = std::ftell(f);
    if (!have_nbits || !have_nchans || !have_tsamp)
        return set_error(TX_EINVAL, path + ": the header lacks nbits, nchans or tsamp");
    if (h->nbits != 8 && h->nbits != 32)
        return set_error(TX_EINVAL, path + ": " + std::to_string(h->nbits) +
                                        "-bit samples; 8-bit unsigned and 32-bit float are read");
    if (sign != 0 && h->nbits == 8) return set_error(TX_EINVAL, path + ": signed 8-bit samples are not read");
    if (h->nifs != 1) return set_error(TX_EINVAL, path + ": " + std::to_string(h->nifs) + " IFs; one is read");
    if (h->nchans < 1 || h->nchans > (1 << 20) || !(h->tsamp > 0.))
        return set_error(TX_EINVAL, path + ": nchans or tsamp out of range");
    if (!fchannel.empty()) {
        if ((long)fchannel.size() != h->nchans)
            return set_error(TX_EINVAL, path + ": fchannel count differs from nchans");
        freqs = fchannel;
        if (!have_fch1) h->fch1 = fchannel[0];
        if (!have_foff) h->foff = h->nchans > 1 ? fchannel[1] - fchannel[0] : 0.;
    } else {
        if (!have_fch1 || !have_foff) return set_error(TX_EINVAL, path + ": the header lacks fch1 or foff");
        freqs.resize((size_t)h->nchans);
        for (int j = 0; j < h->nchans; j++) freqs[j] = h->fch1 + (double)j * h->foff;
    }
    const long bps = (long)h->nchans * h->nifs * h->nbits / 8;
    h->header_bytes = hb;
    h->nsamples = size > hb ? (size - hb) / bps : 0;
    return TX_OK;
}

bool close_to(double a, double b) { return std::fabs(a - b) <= 1e-9 * std::fmax(std::fabs(a), std::fabs(b)) + 1e-300; }

} // namespace

struct tx_fil {
    std::vector<FilFile> files;
    std::vector<long> first; // stream sample of each file's first sample
    long bps = 0;            // bytes per sample (all channels)
    long total = 0;
    long skip_start = 0, skip_end = 0;
    long pos = 0; // stream sample of the next read
    int cur = -1; // file open in fp
    FILE *fp = nullptr;
};

extern "C" int tx_fil_read_header(const char *path, tx_fil_header *out)
{
    if (!path || !out) return set_error(TX_EINVAL, "null argument");
    std::vector<double> freqs;
    return parse_header(path, out, freqs);
}

extern "C" int tx_fil_write_header(const char *path, const tx_fil_header *hd)
{
    using namespace txgpu::sig;
    if (!path || !hd) return set_error(TX_EINVAL, "null argument");
    if ((hd->nbits != 8 && hd->nbits != 32) || hd->nifs != 1 || hd->nchans < 1 || !(hd->tsamp > 0.))
        return set_error(TX_EINVAL, "a filterbank header needs nbits 8 or 32, nifs 1, nchans >= 1 and tsamp > 0");
    if (::strnlen(hd->source_name, 128) == 128 || ::strnlen(hd->rawdatafile, 128) == 128)
        return set_error(TX_EINVAL, "source_name and rawdatafile must be NUL-terminated within 128 bytes");
    std::string h;
    put_str(h, "HEADER_START");
    if (hd->rawdatafile[0]) {
        put_str(h, "rawdatafile");
        put_str(h, hd->rawdatafile);
    }
    if (hd->source_name[0]) {
        put_str(h, "source_name");
        put_str(h, hd->source_name);
    }
    put_int(h, "telescope_id", hd->telescope_id);
    put_int(h, "machine_id", hd->machine_id);
    put_int(h, "data_type", hd->data_type);
    put_int(h, "barycentric", hd->barycentric);
    put_int(h, "pulsarcentric", hd->pulsarcentric);
    put_double(h, "src_raj", hd->src_raj);
    put_double(h, "src_dej", hd->src_dej);
    put_double(h, "az_start", hd->az_start);
    put_double(h, "za_start", hd->za_start);
    put_double(h, "fch1", hd->fch1);
    put_double(h, "foff", hd->foff);
    put_int(h, "nchans", hd->nchans);
    put_int(h, "nbeams", hd->nbeams);
    put_int(h, "ibeam", hd->ibeam);
    put_int(h, "nbits", hd->nbits);
    put_double(h, "tstart", hd->tstart);
    put_double(h, "tsamp", hd->tsamp);
    put_int(h, "nifs", hd->nifs);
    if (hd->refdm != 0.) put_double(h, "refdm", hd->refdm);
    if (hd->period != 0.) put_double(h, "period", hd->period);
    put_str(h, "HEADER_END");
    FILE *f = std::fopen(path, "wb");
    if (!f) return set_error(TX_EIO, std::string("cannot open ") + path + ": " + std::strerror(errno));
    const bool ok = std::fwrite(h.data(), 1, h.size(), f) == h.size();
    if (std::fclose(f) != 0 || !ok) return set_error(TX_EIO, std::string("cannot write ") + path);
    return TX_OK;
}

extern "C" int tx_fil_open(const char *const *paths, int npaths, int check_continuity, tx_fil **out,
                           tx_fil_header *hdr)
{
    if (!paths || npaths < 1 || !out) return set_error(TX_EINVAL, "null argument or no file");
    *out = nullptr;
    tx_fil *r = new tx_fil();
    for (int k = 0; k < npaths; k++) {
        if (!paths[k]) {
            delete r;
            return set_error(TX_EINVAL, "null path");
        }
        FilFile ff;
        ff.path = paths[k];
        const int rc = parse_header(ff.path, &ff.h, ff.freqs);
        if (rc) {
            delete r;
            return rc;
        }
        if (k > 0) {
            const tx_fil_header &a = r->files[0].h, &p = r->files.back().h, &b = ff.h;
            std::string why;
            if (a.nbits != b.nbits || a.nchans != b.nchans || a.nifs != b.nifs || !close_to(a.tsamp, b.tsamp) ||
                !close_to(a.fch1, b.fch1) || !close_to(a.foff, b.foff))
                why = ff.path + ": nbits, nchans, nifs, tsamp, fch1 or foff differ from " + r->files[0].path;
            const double gap = (b.tstart - (p.tstart + (double)p.nsamples * p.tsamp / 86400.)) * 86400.;
            if (why.empty() && check_continuity && std::fabs(gap) > 0.5 * a.tsamp)
                why = ff.path + " does not continue " + r->files.back().path + " (" + std::to_string(gap) +
                      " s apart; pass check_continuity 0 to read the files as one stream anyway)";
            if (!why.empty()) {
                delete r;
                return set_error(TX_EINVAL, why);
            }
        }
        r->first.push_back(r->total);
        r->total += ff.h.nsamples;
        r->files.push_back(ff);
    }
    const tx_fil_header &h0 = r->files[0].h;
    r->bps = (long)h0.nchans * h0.nifs * h0.nbits / 8;
    if (hdr) {
        *hdr = h0;
        hdr->nsamples = r->total;
    }
    *out = r;
    return TX_OK;
}

extern "C" int tx_fil_skip(tx_fil *f, long skip_start, long skip_end)
{
    if (!f) return set_error(TX_EINVAL, "null argument");
    if (skip_start < 0 || skip_end < 0 || skip_start > f->total - skip_end)
        return set_error(TX_EINVAL, "skips must be >= 0 and leave a non-negative number of samples");
    f->skip_start = skip_start;
    f->skip_end = skip_end;
    f->pos = skip_start;
    return TX_OK;
}

extern "C" int tx_fil_frequencies(const tx_fil *f, double *freqs)
{
    if (!f || !freqs) return set_error(TX_EINVAL, "null argument");
    const std::vector<double> &v = f->files[0].freqs;
    std::memcpy(freqs, v.data(), v.size() * sizeof(double));
    return TX_OK;
}

extern "C" long tx_fil_read(tx_fil *f, void *dst, long nsamp)
{
    if (!f || (!dst && nsamp > 0) || nsamp < 0) return set_error(TX_EINVAL, "bad argument");
    const long end = f->total - f->skip_end;
    long want = std::min(nsamp, end - f->pos), got = 0;
    char *out = (char *)dst;
    while (want > 0) {
        int k = -1; // the file holding stream sample pos
        for (int i = 0; i < (int)f->files.size() && k < 0; i++)
            if (f->pos >= f->first[i] && f->pos < f->first[i] + f->files[i].h.nsamples) k = i;
        if (k < 0) break;
        const FilFile &ff = f->files[k];
        const long in_file = f->first[k] + ff.h.nsamples - f->pos; // samples left in file k
        if (f->cur != k) {
            if (f->fp) std::fclose(f->fp);
            f->fp = std::fopen(ff.path.c_str(), "rb");
            f->cur = f->fp ? k : -1;
            if (!f->fp) return set_error(TX_EIO, "cannot open " + ff.path + ": " + std::strerror(errno));
        }
        const long n = std::min(want, in_file);
        const long off = ff.h.header_bytes + (f->pos - f->first[k]) * f->bps;
        if (std::fseek(f->fp, off, SEEK_SET) != 0 ||
            std::fread(out + got * f->bps, 1, (size_t)(n * f->bps), f->fp) != (size_t)(n * f->bps))
            return set_error(TX_EIO, "short read from " + ff.path);
        f->pos += n;
        got += n;
        want -= n;
    }
    return got;
}

extern "C" int tx_fil_close(tx_fil *f)
{
    if (!f) return TX_OK;
    if (f->fp) std::fclose(f->fp);
    delete f;
    return TX_OK;
}
