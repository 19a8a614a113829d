// dropin_search.cpp -- drives the C++ drop-in (include/transientx_gpu.hpp) the
// way SinglePulse::prepare / SinglePulse::run do
// (src/singlepulse/singlepulse.cpp:206-254,287-341), on float32 blocks read
// from a file, and writes every ready block's S/N maps and points.
//
// usage: dropin_search data.f32 freqs.f64 out.bin nchans ndump nblocks tsamp
//                      dms ddm ndm overlap nsub minw maxw snrloss thre exact
// out.bin, per ready block: int64 block, int64 counter, int64 ndm, int64 N,
//   float mxS[ndm*N], int32 mxwn[ndm*N], int64 npts, tx_point pts[npts]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "transientx_gpu.hpp"

struct Buffer { // the members of DataBuffer<float> the drop-in reads
    long nsamples = 0, nchans = 0;
    double tsamp = 0.;
    std::vector<double> frequencies;
    std::vector<float> buffer;
};

int main(int argc, char **argv)
{
    if (argc != 18) {
        std::fprintf(stderr, "usage: %s data.f32 freqs.f64 out.bin nchans ndump nblocks tsamp dms ddm ndm "
                             "overlap nsub minw maxw snrloss thre exact\n", argv[0]);
        return 2;
    }
    const long nchans = std::atol(argv[4]), ndump = std::atol(argv[5]), nblocks = std::atol(argv[6]);
    Buffer db;
    db.nsamples = ndump;
    db.nchans = nchans;
    db.tsamp = std::atof(argv[7]);
    db.frequencies.resize(nchans);
    db.buffer.resize((size_t)ndump * nchans);
    FILE *ff = std::fopen(argv[2], "rb");
    if (!ff || std::fread(db.frequencies.data(), 8, nchans, ff) != (size_t)nchans) return 3;
    std::fclose(ff);

    RealTime::SubbandDedispersion dedisp;
    Boxcar boxcar;
    dedisp.dms = std::atof(argv[8]);
    dedisp.ddm = std::atof(argv[9]);
    dedisp.ndm = std::atoi(argv[10]);
    dedisp.overlap = std::atof(argv[11]);
    dedisp.ndump = ndump;
    dedisp.nsub = std::atoi(argv[12]);
    dedisp.exact_order = std::atoi(argv[17]) != 0;
    dedisp.prepare(db);
    boxcar.thre = (float)std::atof(argv[16]);
    boxcar.prepare(dedisp);

    // the width ladder of SinglePulse::prepare (src/singlepulse/singlepulse.cpp:242-254)
    float minw = (float)std::atof(argv[13]), maxw = (float)std::atof(argv[14]);
    float snrloss = (float)std::atof(argv[15]);
    minw = minw < boxcar.tsamp ? boxcar.tsamp : minw;
    float wfactor = 1. / ((1. - snrloss) * (1. - snrloss));
    std::vector<int> vwn;
    vwn.push_back((int)std::round(minw / boxcar.tsamp));
    while (true) {
        int w1 = vwn.back();
        int w2 = w1 * wfactor;
        w2 = w2 < w1 + 1 ? w1 + 1 : w2;
        if (w2 * boxcar.tsamp > maxw) break;
        vwn.push_back(w2);
    }

    FILE *fd = std::fopen(argv[1], "rb");
    FILE *fo = std::fopen(argv[3], "wb");
    if (!fd || !fo) return 3;
    for (long b = 1; b <= nblocks; b++) {
        if (std::fread(db.buffer.data(), 4, db.buffer.size(), fd) != db.buffer.size()) return 4;
        dedisp.run(db, db.nsamples);
        if (boxcar.run(dedisp, vwn, false)) {
            const long ndm = boxcar.ndm, N = boxcar.nsamples;
            std::vector<tx_point> pts = boxcar.points();
            const long hdr[4] = {b, boxcar.counter, ndm, N};
            const long np = (long)pts.size();
            std::fwrite(hdr, 8, 4, fo);
            std::fwrite(boxcar.mxS, 4, (size_t)ndm * N, fo);
            std::fwrite(boxcar.mxwn, 4, (size_t)ndm * N, fo);
            std::fwrite(&np, 8, 1, fo);
            if (np) std::fwrite(pts.data(), sizeof(tx_point), pts.size(), fo);
            // the copy SinglePulse makes keeps its own maps and shares the plan
            Boxcar copy = boxcar;
            if (copy.mxS == boxcar.mxS || copy.mxS[0] != boxcar.mxS[0] || copy.points().size() != pts.size())
                return 5;
        }
        dedisp.cache();
    }
    std::fclose(fd);
    std::fclose(fo);
    std::printf("dropin_search: %ld blocks, %zu widths\n", nblocks, vwn.size());
    return 0;
}
