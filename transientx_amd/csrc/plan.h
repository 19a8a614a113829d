// plan.h -- host-side dedispersion plan: the declared two-stage sub-band rule.
//
// TransientX configures RealTime::SubbandDedispersion (XLibs, not part of the
// TransientX tree) with dms/ddm/ndm/overlap/ndump at
// src/singlepulse/singlepulse.cpp:230-236 and reads offset/noverlap back at
// src/singlepulse/boxcar.cpp:144,161.  The sub-band grouping, reference
// frequencies and rounding below are this implementation's declared rule:
//   noverlap   = (long)(overlap*ndump), N = ndump + noverlap
//   nsub       = given or nchans/32, dividing nchans; cps = nchans/nsub
//   f_ref(s)   = highest channel frequency of sub-band s
//   group g    = given, or 1 + floor((1/fmin^2 - 1/fmax^2) / (1/fmin^2 - 1/f_ref(s_lo)^2))
//   dmc[k]     = dms + ddm*(k*g + (g_k - 1)/2)
//   d1[k][j]   = round(delay(dmc_k, f_j)/tsamp) - round(delay(dmc_k, f_ref(s(j)))/tsamp)
//   d2[i][s]   = round(delay(dm_i, f_ref(s))/tsamp)
//   maxdelay   = max d1 + max d2, offset = maxdelay + noverlap
// with delay(dm, f) = 4.148741601e3*dm*(1/f^2 - 1/fmax^2) (include/candidate.h:116-119)
// and round() half away from zero (src/singlepulse/candidate.cpp:113).
#pragma once
#include <string>
#include <vector>

namespace txgpu {

double dmdelay(double dm, double fh, double fl);

struct PlanMath {
    int nchans = 0;
    double tsamp = 0, dms = 0, ddm = 0, overlap = 0;
    long ndm = 0, ndump = 0;
    long noverlap = 0, nsamples = 0;
    int nsub = 0, cps = 0, group = 0, ndmc = 0;
    long D1 = 0, D2 = 0, maxdelay = 0, offset = 0;
    double fmin = 0, fmax = 0;
    std::vector<double> freqs, fref, fsub, dmc;
    std::vector<int> d1; // [ndmc][nchans]
    std::vector<int> d2; // [ndm][nsub]
};

// Returns an empty string on success, else the reason the row is invalid.
std::string build_plan(PlanMath &pm, const double *freqs, int nchans, double tsamp, long ndump,
                       double overlap, double dms, double ddm, long ndm, int nsub, int group);

} // namespace txgpu
