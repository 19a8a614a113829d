// Every expression TransientX's own callers write on RealTime::SubbandDedispersion
// and Boxcar, compiled against include/transientx_gpu.hpp with the reference's
// language level (-std=c++11, configure.ac:14) and -Wall -Werror.  Each
// function below names the reference lines whose expressions it repeats.
// The program makes no GPU call unless it gets an argument (a rootname):
// the CPU test compiles and links it, the GPU test runs it with --savetim.
#include <cstdio>
#include <string>
#include <vector>

#include "transientx_gpu.hpp"

// the XLibs types the calls take, with the members TransientX uses
struct DataBuffer {
    long nsamples = 0, nchans = 0, counter = 0;
    double tsamp = 0.;
    std::vector<double> frequencies;
    std::vector<float> buffer;
    bool isbusy = false, closable = false;
};
struct Filterbank {
    double tstart = 0., fch1 = 0., foff = 0.;
    int ibeam = 0;
    long nchans = 0;
};

struct Row {
    RealTime::SubbandDedispersion dedisp;
    Boxcar boxcar;
    Filterbank fildedisp;
    std::vector<int> vwn;
    double dms = 0., ddm = 1., overlap = 0.;
    int ndm = 8, outnbits = 8;
    float outmean = 0.f, outstd = 3.f, minw = 1e-4f, maxw = 0.1f, snrloss = 0.1f;
    bool savetim = true, iqr = false;
    std::string rootname = "J0000-00", format = "pulsarx";

    // src/singlepulse/singlepulse.cpp:228-254 (SinglePulse::prepare)
    void prepare(DataBuffer &rfi)
    {
        if (format == "presto") outnbits = 32;
        dedisp.dms = dms;
        dedisp.ddm = ddm;
        dedisp.ndm = ndm;
        dedisp.overlap = overlap;
        dedisp.ndump = rfi.nsamples;
        dedisp.rootname = rootname;
        dedisp.prepare(rfi);
        if (savetim) dedisp.preparedump(fildedisp, outnbits, format);
        boxcar.prepare(dedisp);
        minw = minw < boxcar.tsamp ? boxcar.tsamp : minw;
        vwn.assign(1, (int)(minw / boxcar.tsamp + 0.5));
        while ((vwn.back() + 1) * boxcar.tsamp <= maxw) vwn.push_back(vwn.back() + 1);
    }

    // src/singlepulse/singlepulse.cpp:287-341 (SinglePulse::run)
    bool run(DataBuffer &databuffer)
    {
        DataBuffer *data = &databuffer;
        dedisp.run(*data, data->nsamples);
        const bool ready = boxcar.run(dedisp, vwn, iqr);
        dedisp.cache();
        if (savetim) dedisp.rundump(outmean, outstd, outnbits, format);
        return ready;
    }
};

// src/singlepulse/boxcar.cpp:142-161,183,318: Boxcar's reads of the dedisperser
static double boxcar_reads(const RealTime::SubbandDedispersion &dedisp)
{
    double acc = (double)(dedisp.ndump + dedisp.noverlap) + dedisp.tsamp + dedisp.dms + dedisp.ddm +
                 (double)dedisp.ndm + (double)dedisp.nchans + dedisp.frequencies.front() +
                 dedisp.frequencies.back() + (double)dedisp.counter + (double)dedisp.offset;
    std::vector<float> tim, sub;
    dedisp.get_timdata(tim, 0, true);
    dedisp.get_subdata(sub, 0);
    acc += (double)dedisp.sub.nchans;
    return acc + tim[0] + sub[0];
}

// src/singlepulse/candplot.cpp:47-48,122-164,259-278,426-427: CandPlot's reads
static double candplot_reads(const RealTime::SubbandDedispersion &dedisp, const Boxcar &boxcar, long isamp,
                             int idm, int j)
{
    double tstart = 60000.;
    double mjd = tstart + (-dedisp.offset + boxcar.counter - dedisp.ndump + isamp) * boxcar.tsamp / 86400.;
    double bmjd = tstart + (-dedisp.offset + boxcar.counter - dedisp.ndump) * boxcar.tsamp / 86400.;
    std::vector<float> tim, sub;
    dedisp.get_timdata(tim, idm, true);
    dedisp.get_subdata(sub, idm, true);
    double vt = (isamp + 1) * dedisp.sub.tsamp;
    double vf = dedisp.sub.frequencies[j] + dedisp.sub.frequencies.front() + dedisp.sub.frequencies.back();
    float mxft = sub[j * (dedisp.noverlap + dedisp.ndump) + isamp] / dedisp.fcnt[j];
    long int ndm = boxcar.ndm;
    double vdm = boxcar.dms + idm * boxcar.ddm;
    float mxdmt = boxcar.mxS[(j + idm) * boxcar.nsamples + isamp];
    return mjd + bmjd + vt + vf + mxft + (double)ndm + vdm + mxdmt + tim[0];
}

// src/singlepulse/cluster.cpp:65-112,150: Cluster's reads of the boxcar
static double cluster_reads(const Boxcar &boxcar)
{
    long counter = boxcar.counter;
    double tsamp = boxcar.tsamp;
    double acc = 0.;
    float *pS = boxcar.mxS;
    int *pwn = boxcar.mxwn;
    for (long int j = 0; j < boxcar.ndm; j++)
        for (long int i = 0; i < boxcar.nsamples; i++) {
            double dm = boxcar.dms + j * boxcar.ddm;
            double smear = RealTime::SubbandDedispersion::dmdelay(dm, boxcar.fmax, boxcar.fmin) * 0.5;
            acc += *pS++ + *pwn++ + smear;
        }
    return acc + counter + tsamp + (pwn[-1] != boxcar.maxwn);
}

// src/singlepulse/dedisperse_fil.cpp:306-310: the dump header fields
static void set_fildedisp(std::vector<Row> &search1, const DataBuffer &prep)
{
    for (auto sp = search1.begin(); sp != search1.end(); ++sp) {
        sp->fildedisp.tstart = 60000.;
        sp->fildedisp.ibeam = 1;
        sp->fildedisp.fch1 = prep.frequencies.front();
        sp->fildedisp.foff = prep.frequencies.back() - prep.frequencies.front();
        sp->fildedisp.nchans = (long)prep.frequencies.size();
    }
}

// src/singlepulse/dedisperse_fil.cpp:345-369: flush blocks, then the dump files' ends
static void flush_and_finish(std::vector<Row> &search1, DataBuffer &prep, const std::string &format)
{
    for (auto sp = search1.begin(); sp != search1.end(); ++sp) {
        int nleft = sp->dedisp.offset / sp->dedisp.ndump + 1;
        for (long int k = 0; k < nleft; k++) {
            prep.counter += prep.nsamples;
            prep.isbusy = true;
            sp->run(prep);
        }
    }
    if (format == "presto") {
        for (auto sp = search1.begin(); sp != search1.end(); ++sp) (*sp).dedisp.makeinf((*sp).fildedisp);
    } else if (format != "sigproc") {
        for (auto sp = search1.begin(); sp != search1.end(); ++sp) (*sp).dedisp.modifynblock();
    }
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::printf("usage: dropin_calls <rootname>  (runs a 64-channel search with --savetim on device 0)\n");
        return 0;
    }
    DataBuffer prep;
    prep.nchans = 64;
    prep.nsamples = 256;
    prep.tsamp = 1e-3;
    for (long j = 0; j < prep.nchans; j++) prep.frequencies.push_back(1500. - 300. * j / (prep.nchans - 1.));
    prep.buffer.assign((size_t)prep.nsamples * prep.nchans, 0.f);
    std::vector<Row> search1(1);
    search1[0].rootname = argv[1];
    set_fildedisp(search1, prep);
    search1[0].prepare(prep);
    double acc = boxcar_reads(search1[0].dedisp);
    for (int b = 0; b < 3; b++) {
        for (size_t k = 0; k < prep.buffer.size(); k++) prep.buffer[k] = (float)((k * 7 + b) % 13) - 6.f;
        if (search1[0].run(prep))
            acc += candplot_reads(search1[0].dedisp, search1[0].boxcar, 10, 0, 0) + cluster_reads(search1[0].boxcar);
    }
    flush_and_finish(search1, prep, search1[0].format);
    std::printf("dropin_calls ok %g\n", acc * 0.);
    return 0;
}
