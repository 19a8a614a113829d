// transientx_gpu.hpp -- header-only C++ drop-in for TransientX's
// RealTime::SubbandDedispersion and Boxcar, over the C ABI of txgpu.h.
//
// The classes keep the members and methods TransientX's callers use
// (src/singlepulse/singlepulse.cpp:230-254,328-338, src/singlepulse/boxcar.cpp:144-183,
// src/singlepulse/cluster.cpp:84-112, src/singlepulse/candplot.cpp:47-48,122-164),
// so SinglePulse, Cluster and CandPlot compile against them unchanged:
//
//   dedisp.dms/ddm/ndm/overlap/ndump/rootname = ...;  dedisp.prepare(databuffer);
//   boxcar.prepare(dedisp);
//   dedisp.run(databuffer, ns);  if (boxcar.run(dedisp, vwn, iqr)) { ... boxcar.mxS ... }
//   dedisp.cache();
//   --savetim: dedisp.preparedump(fil, nbits, format) after prepare,
//   dedisp.rundump(mean, std, nbits, format) after each cache(), and at the
//   end dedisp.makeinf(fil) (presto) or dedisp.modifynblock() (pulsarx).
//
// DataBuffer is any type with `long nsamples, nchans; double tsamp;
// std::vector<double> frequencies; std::vector<float> buffer` (time-major
// [nsamples][nchans]), as XLibs' DataBuffer<float> has.  Filterbank is any
// type with the fields transientx_fil sets on SinglePulse::fildedisp
// (tstart, ibeam, fch1, foff, nchans; src/singlepulse/dedisperse_fil.cpp:306-310).
//
// Device resources are taken in prepare(), not in the constructor: SinglePulse
// copies its dedisperser and boxcar into a vector before prepare
// (src/singlepulse/singlepulse.cpp:455,543).  Copies made after prepare share
// the device plan.  Errors end the process with exit(-1) after printing the
// reason, as the reference does on fatal errors (src/singlepulse/singlepulse.cpp:414-415).
#ifndef TRANSIENTX_GPU_HPP
#define TRANSIENTX_GPU_HPP

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "txgpu.h"

namespace txgpu_dropin {

inline void check(int rc, const char *what)
{
    if (rc < 0) {
        std::fprintf(stderr, "txgpu: %s: %s\n", what, tx_last_error());
        std::exit(-1);
    }
}

// One beam + one plan: a ddplan row on its own input stream, and the
// --savetim files of that row once preparedump has opened them.
struct Handles {
    tx_beam *beam = nullptr;
    tx_plan *plan = nullptr;
    tx_dump *dump = nullptr;
    int dump_format = -1;
    Handles() = default;
    Handles(const Handles &) = delete;
    Handles &operator=(const Handles &) = delete;
    ~Handles()
    {
        tx_dump_close(dump);
        tx_plan_destroy(plan);
        tx_beam_destroy(beam);
    }
};

// --format of src/singlepulse/dedisperse_fil.cpp:96
inline int dump_format(const std::string &format)
{
    if (format == "pulsarx") return TX_DUMP_PULSARX;
    if (format == "sigproc") return TX_DUMP_SIGPROC;
    if (format == "presto") return TX_DUMP_PRESTO;
    std::fprintf(stderr, "txgpu: unknown dump format %s (pulsarx, sigproc, presto)\n", format.c_str());
    std::exit(-1);
}

} // namespace txgpu_dropin

namespace RealTime {

class SubbandDedispersion {
public:
    // written by the caller before prepare (src/singlepulse/singlepulse.cpp:230-235)
    double dms = 0.;
    double ddm = 0.;
    int ndm = 0;
    double overlap = 0.;
    long ndump = 0;
    std::string rootname;
    // additions: sub-band count and fine DMs per coarse DM (0 = defaults),
    // HIP device, full S/N maps kept for Boxcar::mxS, reference FP32 order
    int nsub = 0;
    int group = 0;
    int device = 0;
    bool store_maps = true;
    bool exact_order = false;

    // read back by Boxcar, Cluster and CandPlot
    long counter = 0;
    long offset = 0;
    long noverlap = 0;
    long nchans = 0;
    double tsamp = 0.;
    std::vector<double> frequencies;
    std::vector<float> fcnt; // live channels per sub-band (candplot.cpp:164)
    struct {
        long nchans = 0;
        double tsamp = 0.;
        std::vector<double> frequencies; // sub-band centre frequencies
    } sub;

    // include/candidate.h:116-119
    static double dmdelay(double dm, double fh, double fl)
    {
        return 4.148741601e3 * dm * (1. / (fl * fl) - 1. / (fh * fh));
    }

    template <class DataBuffer>
    void prepare(DataBuffer &databuffer)
    {
        nchans = databuffer.nchans;
        tsamp = databuffer.tsamp;
        frequencies.assign(databuffer.frequencies.begin(), databuffer.frequencies.end());
        if (ndump <= 0) ndump = databuffer.nsamples;
        h = std::make_shared<txgpu_dropin::Handles>();
        tx_beam_desc bd;
        std::memset(&bd, 0, sizeof(bd));
        bd.nchans = (int)nchans;
        bd.dtype = TX_DTYPE_F32;
        bd.ndump = ndump;
        bd.device = device;
        txgpu_dropin::check(tx_beam_create(&bd, &h->beam), "beam");
        tx_plan_desc pd;
        std::memset(&pd, 0, sizeof(pd));
        pd.frequencies = frequencies.data();
        pd.tsamp = tsamp;
        pd.dms = dms;
        pd.ddm = ddm;
        pd.ndm = ndm;
        pd.overlap = overlap;
        pd.nsub = nsub;
        pd.group = group;
        pd.nbox = 0; // Boxcar::run brings the ladder
        pd.store_maps = store_maps ? 1 : 0;
        pd.exact_order = exact_order ? 1 : 0;
        txgpu_dropin::check(tx_plan_create(h->beam, &pd, &h->plan), "plan");
        tx_info in;
        txgpu_dropin::check(tx_get_info(h->plan, &in), "info");
        offset = in.offset;
        noverlap = in.noverlap;
        counter = in.counter;
        std::vector<double> fs(in.nsub), fc(in.nsub);
        txgpu_dropin::check(tx_get_tables(h->plan, nullptr, nullptr, nullptr, nullptr, fs.data(), fc.data()),
                            "tables");
        fcnt.assign(fc.begin(), fc.end());
        sub.nchans = in.nsub;
        sub.tsamp = tsamp;
        sub.frequencies = fs;
    }

    // src/singlepulse/singlepulse.cpp:328: the block's first ns == ndump samples
    template <class DataBuffer>
    void run(DataBuffer &databuffer, long ns)
    {
        txgpu_dropin::check(tx_beam_push(h->beam, databuffer.buffer.data(), ns, 0), "push");
        txgpu_dropin::check(tx_dedisp_run(h->plan), "dedisperse");
        tx_info in;
        txgpu_dropin::check(tx_get_info(h->plan, &in), "info");
        counter = in.counter;
    }

    // the device ring already slides with every run (src/singlepulse/singlepulse.cpp:338)
    void cache() {}

    void get_timdata(std::vector<float> &timdata, int idm, bool with_overlap = true) const
    {
        timdata.resize(with_overlap ? ndump + noverlap : ndump);
        txgpu_dropin::check(tx_get_timdata(h->plan, idm, with_overlap ? 1 : 0, timdata.data()), "timdata");
    }

    void get_subdata(std::vector<float> &subdata, int idm, bool = true) const
    {
        subdata.resize((size_t)sub.nchans * (ndump + noverlap));
        txgpu_dropin::check(tx_get_subdata(h->plan, idm, subdata.data()), "subdata");
    }

    // --savetim (src/singlepulse/singlepulse.cpp:237-238): open this row's
    // dump files with the header fields transientx_fil set on fildedisp
    // (src/singlepulse/dedisperse_fil.cpp:306-310).  Formats and files: the
    // declared rule of tx_dump_open in txgpu.h.
    template <class Filterbank>
    void preparedump(Filterbank &fil, int nbits, const std::string &format)
    {
        if (!h) {
            std::fprintf(stderr, "txgpu: preparedump before prepare\n");
            std::exit(-1);
        }
        tx_dump_desc d;
        std::memset(&d, 0, sizeof(d));
        d.rootname = rootname.c_str();
        d.format = txgpu_dropin::dump_format(format);
        d.nbits = d.format == TX_DUMP_PRESTO ? 32 : nbits;
        d.tstart = fil.tstart;
        d.fch1 = fil.fch1;
        d.foff = fil.foff;
        d.nchans = (int)fil.nchans;
        d.ibeam = (int)fil.ibeam;
        tx_dump_close(h->dump);
        h->dump = nullptr;
        txgpu_dropin::check(tx_dump_open(h->plan, &d, &h->dump), "preparedump");
        h->dump_format = d.format;
    }

    // src/singlepulse/singlepulse.cpp:339-340: this run's new samples of
    // every DM trial, scaled to mean / std and quantized to nbits
    void rundump(float mean, float std, int, const std::string &)
    {
        if (!h || !h->dump) {
            std::fprintf(stderr, "txgpu: rundump before preparedump\n");
            std::exit(-1);
        }
        txgpu_dropin::check(tx_dump_block(h->dump, mean, std), "rundump");
    }

    // src/singlepulse/dedisperse_fil.cpp:356-362: PRESTO's .inf files.
    // transientx_fil calls it whenever --format is presto, with or without
    // --savetim; without dump files there is nothing to do.
    template <class Filterbank>
    void makeinf(Filterbank &)
    {
        if (h && h->dump && h->dump_format == TX_DUMP_PRESTO)
            txgpu_dropin::check(tx_dump_finish(h->dump), "makeinf");
    }

    // src/singlepulse/dedisperse_fil.cpp:363-369: the PULSARX file's block
    // count.  Called whenever --format is not sigproc, with or without
    // --savetim; without a dump file there is nothing to do.
    void modifynblock()
    {
        if (h && h->dump && h->dump_format == TX_DUMP_PULSARX)
            txgpu_dropin::check(tx_dump_finish(h->dump), "modifynblock");
    }

    tx_plan *plan() const { return h ? h->plan : nullptr; }
    tx_beam *beam() const { return h ? h->beam : nullptr; }

private:
    std::shared_ptr<txgpu_dropin::Handles> h;
};

} // namespace RealTime

// include/boxcar.h:17-43 over the device plan of a SubbandDedispersion.
class Boxcar {
public:
    long counter = 0;
    double tsamp = 0.;
    long nsamples = 0;
    double dms = 0.;
    double ddm = 0.;
    long ndm = 0;
    int maxwn = 0;
    double fmin = 0.;
    double fmax = 0.;
    float *mxS = nullptr; // [ndm][nsamples], filled by run() when maps is set
    int *mxwn = nullptr;
    // additions: the threshold of the fused compaction (points()), and
    // whether run() copies the full maps to mxS/mxwn
    float thre = INFINITY;
    bool maps = true;

    Boxcar() = default;
    Boxcar(const Boxcar &o) { copy_from(o); }
    Boxcar &operator=(const Boxcar &o)
    {
        if (this != &o) {
            release();
            copy_from(o);
        }
        return *this;
    }
    ~Boxcar() { release(); }

    // src/singlepulse/boxcar.cpp:115-156
    void prepare(RealTime::SubbandDedispersion &dedisp)
    {
        plan_ = dedisp.plan();
        tsamp = dedisp.tsamp;
        nsamples = dedisp.ndump + dedisp.noverlap;
        dms = dedisp.dms;
        ddm = dedisp.ddm;
        ndm = dedisp.ndm;
        fmin = 1e6;
        fmax = 0.;
        for (double f : dedisp.frequencies) {
            fmax = std::max(fmax, f);
            fmin = std::min(fmin, f);
        }
        release();
        mxS = new float[(size_t)ndm * nsamples];
        mxwn = new int[(size_t)ndm * nsamples];
        std::fill(mxS, mxS + (size_t)ndm * nsamples, 0.f);
        std::fill(mxwn, mxwn + (size_t)ndm * nsamples, 0);
    }

    // src/singlepulse/boxcar.cpp:158-178: false until the delay buffer is full
    bool run(RealTime::SubbandDedispersion &dedisp, std::vector<int> &vwn, bool iqr)
    {
        plan_ = dedisp.plan();
        txgpu_dropin::check(tx_boxcar_widths(plan_, vwn.data(), (int)vwn.size(), iqr ? 1 : 0), "widths");
        const int rc = tx_boxcar_run(plan_, thre);
        txgpu_dropin::check(rc, "boxcar");
        if (rc == TX_NOT_READY) return false;
        counter = dedisp.counter;
        maxwn = vwn.back();
        if (maps) txgpu_dropin::check(tx_boxcar_rows(plan_, 0, (int)ndm, mxS, mxwn), "maps");
        return true;
    }

    // points with mxS > thre in row-major order (the scan of src/singlepulse/cluster.cpp:89-112)
    std::vector<tx_point> points() const
    {
        size_t n = 0;
        const int rc = tx_boxcar_points(plan_, nullptr, 0, &n); // count only
        if (rc != TX_EOVERFLOW) txgpu_dropin::check(rc, "points");
        std::vector<tx_point> out(n);
        if (n) txgpu_dropin::check(tx_boxcar_points(plan_, out.data(), n, &n), "points");
        return out;
    }

private:
    tx_plan *plan_ = nullptr;
    void release()
    {
        delete[] mxS;
        delete[] mxwn;
        mxS = nullptr;
        mxwn = nullptr;
    }
    void copy_from(const Boxcar &o)
    {
        counter = o.counter;
        tsamp = o.tsamp;
        nsamples = o.nsamples;
        dms = o.dms;
        ddm = o.ddm;
        ndm = o.ndm;
        maxwn = o.maxwn;
        fmin = o.fmin;
        fmax = o.fmax;
        thre = o.thre;
        maps = o.maps;
        plan_ = o.plan_;
        if (o.mxS) {
            const size_t n = (size_t)ndm * nsamples;
            mxS = new float[n];
            mxwn = new int[n];
            std::copy(o.mxS, o.mxS + n, mxS);
            std::copy(o.mxwn, o.mxwn + n, mxwn);
        }
    }
};

#endif // TRANSIENTX_GPU_HPP
