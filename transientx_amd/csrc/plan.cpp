// plan.cpp -- the declared two-stage sub-band dedispersion rule (see plan.h).
#include "plan.h"

#include <algorithm>
#include <cmath>

namespace txgpu {

double dmdelay(double dm, double fh, double fl)
{
    // include/candidate.h:116-119
    return 4.148741601e3 * dm * (1. / (fl * fl) - 1. / (fh * fh));
}

std::string build_plan(PlanMath &pm, const double *freqs, int nchans, double tsamp, long ndump,
                       double overlap, double dms, double ddm, long ndm, int nsub, int group)
{
    if (nchans <= 0 || !freqs) return "nchans must be positive";
    if (!(tsamp > 0)) return "tsamp must be positive";
    if (ndump <= 0) return "ndump must be positive";
    if (ndm <= 0) return "ndm must be positive";
    if (!(ddm >= 0) || !(dms >= 0)) return "dms and ddm must be non-negative";
    if (!(overlap >= 0)) return "overlap must be non-negative";
    pm = PlanMath();
    pm.nchans = nchans;
    pm.tsamp = tsamp;
    pm.dms = dms;
    pm.ddm = ddm;
    pm.overlap = overlap;
    pm.ndm = ndm;
    pm.ndump = ndump;
    pm.freqs.assign(freqs, freqs + nchans);
    pm.noverlap = (long)(overlap * (double)ndump);
    pm.nsamples = ndump + pm.noverlap;
    pm.nsub = nsub > 0 ? nsub : std::max(1, nchans / 32);
    if (nchans % pm.nsub != 0) return "nsub must divide nchans";
    pm.cps = nchans / pm.nsub;
    pm.fmax = *std::max_element(pm.freqs.begin(), pm.freqs.end());
    pm.fmin = *std::min_element(pm.freqs.begin(), pm.freqs.end());
    if (!(pm.fmin > 0)) return "channel frequencies must be positive";

    pm.fref.resize(pm.nsub);
    pm.fsub.resize(pm.nsub);
    for (int s = 0; s < pm.nsub; s++) {
        double hi = 0, sum = 0;
        for (int j = s * pm.cps; j < (s + 1) * pm.cps; j++) {
            hi = std::max(hi, pm.freqs[j]);
            sum += pm.freqs[j];
        }
        pm.fref[s] = hi;
        pm.fsub[s] = sum / pm.cps;
    }

    if (group > 0) {
        pm.group = (int)std::min<long>(group, ndm);
    } else {
        int jlo = (int)(std::min_element(pm.freqs.begin(), pm.freqs.end()) - pm.freqs.begin());
        double fr = pm.fref[jlo / pm.cps];
        double dsub = 1. / (pm.fmin * pm.fmin) - 1. / (fr * fr);
        double dband = 1. / (pm.fmin * pm.fmin) - 1. / (pm.fmax * pm.fmax);
        long g = dsub > 0 ? 1 + (long)std::floor(dband / dsub) : ndm;
        pm.group = (int)std::max<long>(1, std::min<long>(g, ndm));
    }
    pm.ndmc = (int)((ndm + pm.group - 1) / pm.group);

    pm.dmc.resize(pm.ndmc);
    pm.d1.assign((size_t)pm.ndmc * nchans, 0);
    for (int k = 0; k < pm.ndmc; k++) {
        long i0 = (long)k * pm.group;
        long gk = std::min<long>(pm.group, ndm - i0);
        double dmc = dms + ddm * ((double)i0 + (double)(gk - 1) / 2.);
        pm.dmc[k] = dmc;
        for (int s = 0; s < pm.nsub; s++) {
            long r = std::lround(dmdelay(dmc, pm.fmax, pm.fref[s]) / tsamp);
            for (int j = s * pm.cps; j < (s + 1) * pm.cps; j++) {
                long d = std::lround(dmdelay(dmc, pm.fmax, pm.freqs[j]) / tsamp) - r;
                pm.d1[(size_t)k * nchans + j] = (int)d;
                pm.D1 = std::max(pm.D1, d);
            }
        }
    }
    pm.d2.assign((size_t)ndm * pm.nsub, 0);
    for (long i = 0; i < ndm; i++) {
        double dm = dms + (double)i * ddm;
        for (int s = 0; s < pm.nsub; s++) {
            long d = std::lround(dmdelay(dm, pm.fmax, pm.fref[s]) / tsamp);
            pm.d2[(size_t)i * pm.nsub + s] = (int)d;
            pm.D2 = std::max(pm.D2, d);
        }
    }
    pm.maxdelay = pm.D1 + pm.D2;
    pm.offset = pm.maxdelay + pm.noverlap;
    if (pm.maxdelay > (1L << 30)) return "delays exceed 2^30 samples";
    return "";
}

} // namespace txgpu
