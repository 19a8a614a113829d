// dump.cpp -- --savetim files of the dedispersed series: SubbandDedispersion's
// preparedump / rundump / makeinf / modifynblock, which transientx_fil calls
// at src/singlepulse/singlepulse.cpp:238,340 and
// src/singlepulse/dedisperse_fil.cpp:356-369 with the options of
// src/singlepulse/dedisperse_fil.cpp:95-99.  Those methods live in XLibs, which
// the TransientX tree does not hold, so the files follow the rule declared in
// include/txgpu.h.  Host C++ over the public C ABI: the values come from
// tx_dump_series (one kernel and one copy per run); this file only writes
// them out.
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

struct tx_dump {
    tx_plan *plan = nullptr;
    tx_dump_desc d{};
    std::string root, source;
    tx_info info{};
    std::vector<std::string> paths; // one per DM trial, or one per row (PULSARX)
    std::vector<char> buf;          // [ndm][ndump] values of one run
    long per_dm = 0;                // samples written per DM trial
    long nblock = 0;
    long tau_first = 0;
    bool started = false;
    long nblock_pos = -1; // byte offset of the PULSARX header's nblock value
};

namespace {

using txgpu::sig::put_byte;
using txgpu::sig::put_double;
using txgpu::sig::put_int;
using txgpu::sig::put_str;

std::string fmt_dm(double dm)
{
    char b[64];
    std::snprintf(b, sizeof b, "%.2f", dm);
    return b;
}

double dm_of(const tx_dump *dp, long i) { return dp->info.dms + (double)i * dp->info.ddm; }

int write_file(const std::string &path, const char *mode, const void *a, size_t na, const void *b, size_t nb)
{
    FILE *f = std::fopen(path.c_str(), mode);
    if (!f) return set_error(TX_EIO, "cannot open " + path + ": " + std::strerror(errno));
    bool ok = (na == 0 || std::fwrite(a, 1, na, f) == na) && (nb == 0 || std::fwrite(b, 1, nb, f) == nb);
    ok = (std::fclose(f) == 0) && ok;
    return ok ? TX_OK : set_error(TX_EIO, "cannot write " + path);
}

// the header of a file whose first sample is observation sample tau
std::string header(const tx_dump *dp, long tau, double refdm, bool row)
{
    std::string h;
    put_str(h, "HEADER_START");
    put_str(h, "source_name");
    put_str(h, dp->source);
    put_int(h, "data_type", 2);
    put_int(h, "nbits", dp->d.nbits);
    if (dp->d.nbits < 32) put_byte(h, "signed", 1);
    put_double(h, "tstart", dp->d.tstart + (double)tau * dp->info.tsamp / 86400.);
    put_double(h, "tsamp", dp->info.tsamp);
    put_double(h, "fch1", dp->d.fch1);
    put_double(h, "foff", dp->d.foff);
    put_int(h, "nifs", 1);
    put_int(h, "ibeam", dp->d.ibeam);
    put_double(h, "refdm", refdm);
    if (row) {
        put_int(h, "nchans", dp->d.nchans);
        put_double(h, "ddm", dp->info.ddm);
        put_int(h, "ndm", (int32_t)dp->info.ndm);
        put_int(h, "ndump", (int32_t)dp->info.ndump);
        put_str(h, "nblock");
    } else {
        put_int(h, "nchans", 1);
    }
    return h;
}

} // namespace

extern "C" int tx_dump_open(tx_plan *plan, const tx_dump_desc *desc, tx_dump **out)
{
    if (!plan || !desc || !out || !desc->rootname) return set_error(TX_EINVAL, "null argument");
    *out = nullptr;
    if (desc->format != TX_DUMP_PULSARX && desc->format != TX_DUMP_SIGPROC && desc->format != TX_DUMP_PRESTO)
        return set_error(TX_EINVAL, "dump format must be TX_DUMP_PULSARX, TX_DUMP_SIGPROC or TX_DUMP_PRESTO");
    if (desc->nbits != 8 && desc->nbits != 16 && desc->nbits != 32)
        return set_error(TX_EINVAL, "dump nbits must be 8, 16 or 32");
    if (desc->format == TX_DUMP_PRESTO && desc->nbits != 32)
        return set_error(TX_EINVAL, "PRESTO dumps are float32: nbits must be 32");
    tx_dump *dp = new tx_dump();
    dp->plan = plan;
    dp->d = *desc;
    dp->root = desc->rootname;
    dp->source = desc->source_name ? desc->source_name : "";
    int rc = tx_get_info(plan, &dp->info);
    if (rc != TX_OK) {
        delete dp;
        return rc;
    }
    if (desc->format == TX_DUMP_PULSARX) {
        dp->paths.push_back(dp->root + "_DM" + fmt_dm(dm_of(dp, 0)) + "_" + fmt_dm(dm_of(dp, dp->info.ndm - 1)) +
                            ".dedisp");
    } else {
        const char *ext = desc->format == TX_DUMP_SIGPROC ? ".tim" : ".dat";
        for (long i = 0; i < dp->info.ndm; i++) dp->paths.push_back(dp->root + "_DM" + fmt_dm(dm_of(dp, i)) + ext);
    }
    // create (or empty) every file now, so that a bad path fails at preparedump
    for (const std::string &p : dp->paths) {
        rc = write_file(p, "wb", nullptr, 0, nullptr, 0);
        if (rc != TX_OK) {
            delete dp;
            return rc;
        }
    }
    dp->buf.resize((size_t)dp->info.ndm * dp->info.ndump * (desc->nbits / 8));
    *out = dp;
    return TX_OK;
}

extern "C" int tx_dump_block(tx_dump *dp, float mean, float std)
{
    if (!dp) return set_error(TX_EINVAL, "null argument");
    long tau0 = 0, n = 0;
    int rc = tx_dump_series(dp->plan, mean, std, dp->d.nbits, dp->buf.data(), dp->buf.size(), &tau0, &n);
    if (rc != TX_OK || n == 0) return rc;
    const size_t es = (size_t)dp->d.nbits / 8, row = (size_t)n * es;
    const long ndm = dp->info.ndm;
    if (!dp->started) {
        dp->started = true;
        dp->tau_first = tau0;
        if (dp->d.format == TX_DUMP_PULSARX) {
            std::string h = header(dp, tau0, dm_of(dp, 0), true);
            dp->nblock_pos = (long)h.size();
            const int32_t zero = 0;
            h.append((const char *)&zero, 4);
            put_str(h, "HEADER_END");
            rc = write_file(dp->paths[0], "wb", h.data(), h.size(), nullptr, 0);
            if (rc != TX_OK) return rc;
        } else if (dp->d.format == TX_DUMP_SIGPROC) {
            for (long i = 0; i < ndm; i++) {
                std::string h = header(dp, tau0, dm_of(dp, i), false);
                put_str(h, "HEADER_END");
                rc = write_file(dp->paths[i], "wb", h.data(), h.size(), nullptr, 0);
                if (rc != TX_OK) return rc;
            }
        }
    }
    if (dp->d.format == TX_DUMP_PULSARX) {
        rc = write_file(dp->paths[0], "ab", dp->buf.data(), row * ndm, nullptr, 0);
        if (rc != TX_OK) return rc;
    } else {
        for (long i = 0; i < ndm; i++) {
            rc = write_file(dp->paths[i], "ab", dp->buf.data() + (size_t)i * row, row, nullptr, 0);
            if (rc != TX_OK) return rc;
        }
    }
    dp->per_dm += n;
    dp->nblock++;
    return TX_OK;
}

extern "C" int tx_dump_finish(tx_dump *dp)
{
    if (!dp) return set_error(TX_EINVAL, "null argument");
    if (dp->d.format == TX_DUMP_PULSARX && dp->started) {
        // modifynblock: the block count into the header
        FILE *f = std::fopen(dp->paths[0].c_str(), "r+b");
        if (!f) return set_error(TX_EIO, "cannot open " + dp->paths[0] + ": " + std::strerror(errno));
        const int32_t nb = (int32_t)dp->nblock;
        bool ok = std::fseek(f, dp->nblock_pos, SEEK_SET) == 0 && std::fwrite(&nb, 4, 1, f) == 1;
        ok = (std::fclose(f) == 0) && ok;
        if (!ok) return set_error(TX_EIO, "cannot write " + dp->paths[0]);
    } else if (dp->d.format == TX_DUMP_PRESTO) {
        // makeinf: PRESTO's .inf beside each .dat
        const double tstart = dp->d.tstart + (double)dp->tau_first * dp->info.tsamp / 86400.;
        const double bw = std::fabs(dp->d.foff) * dp->d.nchans;
        const double flo = dp->d.foff < 0 ? dp->d.fch1 + dp->d.foff * (dp->d.nchans - 1) : dp->d.fch1;
        for (long i = 0; i < dp->info.ndm; i++) {
            std::string base = dp->paths[i].substr(0, dp->paths[i].size() - 4);
            const std::string name = base.substr(base.find_last_of('/') == std::string::npos
                                                     ? 0
                                                     : base.find_last_of('/') + 1);
            char t[2048];
            std::snprintf(t, sizeof t,
                          " Data file name without suffix          =  %s\n"
                          " Telescope used                         =  Unknown\n"
                          " Instrument used                        =  Unknown\n"
                          " Object being observed                  =  %s\n"
                          " J2000 Right Ascension (hh:mm:ss.ssss)  =  00:00:00.0000\n"
                          " J2000 Declination     (dd:mm:ss.ssss)  =  00:00:00.0000\n"
                          " Data observed by                       =  Unknown\n"
                          " Epoch of observation (MJD)             =  %.15f\n"
                          " Barycentered?           (1=yes, 0=no)  =  0\n"
                          " Number of bins in the time series      =  %ld\n"
                          " Width of each time series bin (sec)    =  %.17g\n"
                          " Any breaks in the data? (1=yes, 0=no)  =  0\n"
                          " Type of observation (EM band)          =  Radio\n"
                          " Beam diameter (arcsec)                 =  0\n"
                          " Dispersion measure (cm-3 pc)           =  %.2f\n"
                          " Central freq of low channel (Mhz)      =  %.12g\n"
                          " Total bandwidth (Mhz)                  =  %.12g\n"
                          " Number of channels                     =  %d\n"
                          " Channel bandwidth (Mhz)                =  %.12g\n"
                          " Data analyzed by                       =  Unknown\n"
                          " Any additional notes:\n"
                          "    Dedispersed series of a TransientX search\n\n",
                          name.c_str(), dp->source.empty() ? "Unknown" : dp->source.c_str(), tstart, dp->per_dm,
                          dp->info.tsamp, dm_of(dp, i), flo, bw, dp->d.nchans, std::fabs(dp->d.foff));
            const int rc = write_file(base + ".inf", "wb", t, std::strlen(t), nullptr, 0);
            if (rc != TX_OK) return rc;
        }
    }
    return TX_OK;
}

extern "C" int tx_dump_close(tx_dump *dp)
{
    delete dp;
    return TX_OK;
}
