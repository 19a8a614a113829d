// ref_hough.cpp -- TEST INFRASTRUCTURE ONLY.
// C-ABI harness around TransientX's include/hough_transform.h, compiled from
// the reference tree where it lies (oracle/build.py adds it to the include
// path).  It lets the tests check the oracle's restatement of the dedispersion
// tree transform (include/hough_transform.h:20-125) against the reference's
// own code on the same inputs.
#include <cmath>
#include <functional>
#include <vector>

#include "hough_transform.h"

static double ref_dmdelay(double dm, double fh, double fl)
{
    return 4.148741601e3 * dm * (1. / (fl * fl) - 1. / (fh * fh));
}

// The get_shift lambda of src/singlepulse/candidate.cpp:970-974.
extern "C" void ref_hough(const float *in, float *out, int nrow, int ncol, const double *freqsP,
                          double dms, double ddm, double tbin, double dm0, int isdedispersed)
{
    std::vector<float> vin(in, in + (size_t)nrow * ncol), vout;
    std::function<int(int, int, int)> get_shift;
    if (!isdedispersed)
        get_shift = [&](int x0, int x1, int p) {
            return (int)std::round(ref_dmdelay(dms + p * ddm, freqsP[x0], freqsP[x1]) / tbin);
        };
    else
        get_shift = [&](int x0, int x1, int p) {
            return (int)std::round(ref_dmdelay(dms + p * ddm - dm0, freqsP[x0], freqsP[x1]) / tbin);
        };
    hough_transform(vin, vout, nrow, ncol, get_shift);
    for (size_t i = 0; i < vout.size(); i++) out[i] = vout[i];
}
