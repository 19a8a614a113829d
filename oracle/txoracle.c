/*
 * txoracle.c -- CPU oracle for the TransientX single-pulse search hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke
 * entry point and the CPU-baseline leg of the benchmark load it; the product
 * (transientx_amd/) never does.  It restates, in plain C, the arithmetic of
 * the reference path (file:line citations are relative to the TransientX
 * source tree) plus the dedispersion rule this implementation declares,
 * because TransientX takes RealTime::SubbandDedispersion from the
 * un-vendored, unpinned XLibs library.
 *
 * Floating point follows the reference build: `g++ -std=c++11 -O3 -mavx2
 * -mfma` contracts `var += x*x` and `var -= mean*mean` into FMA, so those two
 * statements use fmaf() here; everything else is compiled with
 * -ffp-contract=off so that no other contraction happens.
 *
 * Parity status per function (what pins it; see tests/test_oracle_*.py):
 *   txo_dmdelay         closed form + published dispersion constant
 *   txo_ladder          integer sequences worked out independently
 *   txo_plan_*          declared rule; pinned by reductions to direct
 *                       dedispersion (nsub=C, or nsub=1 with group 1)
 *   txo_dedisp_window   brute-force FP64 direct dedispersion, impulse
 *                       response, linearity, DM=0 channel sum
 *   txo_dedisp_window_f64  equal to txo_dedisp_window on integer data
 *                       (exact sums) and to FP64 direct dedispersion with
 *                       one channel per sub-band
 *   txo_boxcar_f64      the single-spike closed form; within FP32 rounding
 *                       of txo_boxcar's maximum
 *   txo_boxcar_*        closed forms (constant row, spike, wn=1), FP64
 *                       window sums, AVX2 tail rule, tie rule
 *   txo_boxcar_runnerup its best equals txo_boxcar's map at the point;
 *                       its second equals a brute-force FP64 maximum
 *                       over the other widths on integer series
 *   txo_threshold       brute-force numpy scan
 *   txo_dump            declared (XLibs' rundump is not in the TransientX
 *                       tree): pinned to float64 arithmetic with one
 *                       rounding per operation, half-away rounding and
 *                       the signed clamp
 *   txo_cluster         scipy connected components on the radius graph
 *   txo_kadane          brute-force O(n^2) maximum subarray
 *   txo_cand_flfh       the reference's vpf window restated; checked
 *                       end to end against the GPU's .cands lines
 *   txo_cand_dedisperse closed form (numpy roll by the rounded delay)
 *   txo_hough           bit-identical to the reference's own
 *                       include/hough_transform.h compiled from source
 *                       with the reference's flags (oracle/_ref), DM-0
 *                       channel sums, linearity, linear sweep
 *   txo_refine_optimize band-order invariance, padding, DM range
 *   txo_mf_ladder       integer sequence worked out by hand
 *   txo_refine_mf       closed-form spike (incl. circular wrap), FP64
 *                       circular boxcar, constant map
 *   txo_mean_var        declared (XLibs get_mean_var is not in the
 *                       TransientX tree): parity unpinned against XLibs
 *   txo_stream_*        composition of the above, checked against the
 *                       window functions
 *   txo_zdot            FP64 least squares (numpy lstsq) on the
 *                       normalised data, residual orthogonal to the
 *                       channel mean and to 1, a channel proportional to
 *                       the channel mean removed, constant channels
 *   txo_kadanef         brute-force maximum runs of integer-valued data,
 *                       planted positive and negative runs, the wnlimit
 *                       and threshold rules
 *   txo_frontend_u8     declared (XLibs Downsample / Equalize are not in
 *                       the TransientX tree): pinned to numpy integer
 *                       block sums and float32 affine, constant input,
 *                       Candidate::downsample's float sums on 8-bit data
 *   txo_frontend_stats  declared; pinned to numpy float64 mean/variance
 *                       (exact on integer data) and the normalised
 *                       output's zero mean / unit variance
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TXO_OK 0
#define TXO_EINVAL (-1)
#define TXO_ENOMEM (-2)
#define TXO_ERANGE (-3)

/* ------------------------------------------------------------------ a1 */
/* include/candidate.h:116-119 (same in include/archivewriter_tx.h:33-36) */
double txo_dmdelay(double dm, double fh, double fl)
{
    return 4.148741601e3 * dm * (1. / (fl * fl) - 1. / (fh * fh));
}

/* ------------------------------------------------------------------ a3 */
/* src/singlepulse/singlepulse.cpp:242-254.  minw/maxw/snrloss are float
 * members of SinglePulse; tsamp is the double Boxcar::tsamp. */
int txo_ladder(double tsamp, float minw, float maxw, float snrloss, int *vwn, int cap)
{
    if (tsamp <= 0. || cap <= 0) return TXO_EINVAL;
    minw = minw < tsamp ? (float)tsamp : minw;
    float wfactor = (float)(1. / ((1. - (double)snrloss) * (1. - (double)snrloss)));
    int n = 0;
    vwn[n++] = (int)round((double)minw / tsamp);
    for (;;) {
        int w1 = vwn[n - 1];
        int w2 = (int)((float)w1 * wfactor);
        w2 = w2 < w1 + 1 ? w1 + 1 : w2;
        if ((double)w2 * tsamp > (double)maxw) break;
        if (n >= cap) return TXO_ERANGE;
        vwn[n++] = w2;
    }
    return n;
}

/* ------------------------------------------------------- a4 (declared) */
/* The dedispersion plan.  TransientX configures SubbandDedispersion with
 * dms/ddm/ndm/overlap/ndump (src/singlepulse/singlepulse.cpp:230-236) and
 * reads back offset/noverlap/ndump (src/singlepulse/boxcar.cpp:144,161).
 * Everything below that is not in the TransientX tree is this
 * implementation's declared rule:
 *   noverlap  = (long)(overlap * ndump);  N = ndump + noverlap
 *   nsub      = given, or nchans/32; must divide nchans
 *   subband s = channels [s*cps, (s+1)*cps), f_ref(s) = max freq in s
 *   group g   = given, or 1 + floor((1/fmin^2 - 1/fmax^2) /
 *               (1/fmin^2 - 1/f_ref(s_min)^2)) (sub-band smearing over
 *               +-(g-1)/2 steps <= half the full-band smearing of one step)
 *   coarse DM k = dms + ddm*(k*g + (g_k-1)/2), g_k = size of group k
 *   d1[k][j] = round(delay(dmc_k, f_j)/tsamp) - round(delay(dmc_k, f_ref(s(j)))/tsamp)
 *   d2[i][s] = round(delay(dm_i, f_ref(s))/tsamp)
 *   D1 = max d1, D2 = max d2, maxdelay = D1 + D2, offset = maxdelay + noverlap
 * delay(dm, f) = txo_dmdelay(dm, fmax, f); round = C round() (half away
 * from zero), as src/singlepulse/candidate.cpp:113 does for its shifts. */
typedef struct {
    int nchans;
    const double *freqs;
    double tsamp;
    long ndump;
    double overlap;
    double dms;
    double ddm;
    long ndm;
    int nsub;  /* 0 = default */
    int group; /* 0 = default */
} txo_plan_in;

typedef struct {
    long noverlap;
    long nsamples; /* N */
    int nsub;
    int cps;
    int group;
    int ndmc;
    long D1;
    long D2;
    long maxdelay;
    long offset;
    double fmax;
    double fmin;
} txo_plan_info;

static int plan_basic(const txo_plan_in *p, txo_plan_info *o)
{
    if (!p || !o || p->nchans <= 0 || !p->freqs || !(p->tsamp > 0.) || p->ndump <= 0 ||
        p->ndm <= 0 || !(p->ddm >= 0.) || !(p->dms >= 0.) || !(p->overlap >= 0.))
        return TXO_EINVAL;
    memset(o, 0, sizeof(*o));
    o->noverlap = (long)(p->overlap * (double)p->ndump);
    o->nsamples = p->ndump + o->noverlap;
    int nsub = p->nsub > 0 ? p->nsub : p->nchans / 32;
    if (nsub <= 0) nsub = 1;
    if (p->nchans % nsub != 0) return TXO_EINVAL;
    o->nsub = nsub;
    o->cps = p->nchans / nsub;
    double fmax = 0., fmin = 1e300;
    for (int j = 0; j < p->nchans; j++) {
        fmax = p->freqs[j] > fmax ? p->freqs[j] : fmax;
        fmin = p->freqs[j] < fmin ? p->freqs[j] : fmin;
    }
    if (!(fmin > 0.)) return TXO_EINVAL;
    o->fmax = fmax;
    o->fmin = fmin;
    return TXO_OK;
}

static double subband_fref(const txo_plan_in *p, int cps, int s)
{
    double f = 0.;
    for (int j = s * cps; j < (s + 1) * cps; j++) f = p->freqs[j] > f ? p->freqs[j] : f;
    return f;
}

static int plan_group(const txo_plan_in *p, const txo_plan_info *o)
{
    if (p->group > 0) return (int)(p->group < p->ndm ? p->group : p->ndm);
    /* the sub-band holding the lowest frequency */
    int smin = 0;
    for (int j = 0; j < p->nchans; j++)
        if (p->freqs[j] == o->fmin) { smin = j / o->cps; break; }
    double fr = subband_fref(p, o->cps, smin);
    double dsub = 1. / (o->fmin * o->fmin) - 1. / (fr * fr);
    double dband = 1. / (o->fmin * o->fmin) - 1. / (o->fmax * o->fmax);
    long g;
    if (!(dsub > 0.)) g = p->ndm;
    else g = 1 + (long)floor(dband / dsub);
    if (g > p->ndm) g = p->ndm;
    if (g < 1) g = 1;
    return (int)g;
}

int txo_plan_info_get(const txo_plan_in *p, txo_plan_info *o)
{
    int rc = plan_basic(p, o);
    if (rc) return rc;
    o->group = plan_group(p, o);
    o->ndmc = (int)((p->ndm + o->group - 1) / o->group);
    long D1 = 0, D2 = 0;
    for (int k = 0; k < o->ndmc; k++) {
        long i0 = (long)k * o->group;
        long gk = p->ndm - i0 < o->group ? p->ndm - i0 : o->group;
        double dmc = p->dms + p->ddm * ((double)i0 + (double)(gk - 1) / 2.);
        for (int s = 0; s < o->nsub; s++) {
            double fr = subband_fref(p, o->cps, s);
            long r = lround(txo_dmdelay(dmc, o->fmax, fr) / p->tsamp);
            for (int j = s * o->cps; j < (s + 1) * o->cps; j++) {
                long d = lround(txo_dmdelay(dmc, o->fmax, p->freqs[j]) / p->tsamp) - r;
                D1 = d > D1 ? d : D1;
            }
        }
    }
    for (long i = 0; i < p->ndm; i++) {
        double dm = p->dms + (double)i * p->ddm;
        for (int s = 0; s < o->nsub; s++) {
            long d = lround(txo_dmdelay(dm, o->fmax, subband_fref(p, o->cps, s)) / p->tsamp);
            D2 = d > D2 ? d : D2;
        }
    }
    o->D1 = D1;
    o->D2 = D2;
    o->maxdelay = D1 + D2;
    o->offset = o->maxdelay + o->noverlap;
    return TXO_OK;
}

/* d1: [ndmc * nchans], d2: [ndm * nsub], dmc: [ndmc], fref: [nsub] (any may be NULL) */
int txo_plan_tables(const txo_plan_in *p, int *d1, int *d2, double *dmc_out, double *fref_out)
{
    txo_plan_info o;
    int rc = txo_plan_info_get(p, &o);
    if (rc) return rc;
    for (int s = 0; s < o.nsub; s++)
        if (fref_out) fref_out[s] = subband_fref(p, o.cps, s);
    for (int k = 0; k < o.ndmc; k++) {
        long i0 = (long)k * o.group;
        long gk = p->ndm - i0 < o.group ? p->ndm - i0 : o.group;
        double dmc = p->dms + p->ddm * ((double)i0 + (double)(gk - 1) / 2.);
        if (dmc_out) dmc_out[k] = dmc;
        if (!d1) continue;
        for (int s = 0; s < o.nsub; s++) {
            long r = lround(txo_dmdelay(dmc, o.fmax, subband_fref(p, o.cps, s)) / p->tsamp);
            for (int j = s * o.cps; j < (s + 1) * o.cps; j++)
                d1[(long)k * p->nchans + j] =
                    (int)(lround(txo_dmdelay(dmc, o.fmax, p->freqs[j]) / p->tsamp) - r);
        }
    }
    if (d2) {
        for (long i = 0; i < p->ndm; i++) {
            double dm = p->dms + (double)i * p->ddm;
            for (int s = 0; s < o.nsub; s++)
                d2[i * o.nsub + s] =
                    (int)lround(txo_dmdelay(dm, o.fmax, subband_fref(p, o.cps, s)) / p->tsamp);
        }
    }
    return TXO_OK;
}

/* ------------------------------------------------------- a5-a7 (declared) */
/* Output window of block b (1-based; counter = b*ndump after the block is
 * ingested): series times tau in [w0, w0+N) with
 * w0 = counter - maxdelay - N, so that output sample i is the pulse arrival
 * time at fmax of  -offset + counter - ndump + i, the time candplot assigns
 * to it (src/singlepulse/candplot.cpp:47).  Input before time 0 is zero.
 *   sub[k][s](tau) = sum_{j in s} x[j](tau + d1[k][j])   (FP32, j ascending)
 *   tim[i](tau)    = sum_s sub[k(i)][s](tau + d2[i][s])   (FP32, s ascending)
 * x is the front-end output, time-major [ntot][nchans] (the DataBuffer
 * layout).  Returns TXO_ERANGE if the window needs input at or past ntot. */
static inline float xat(const float *x, long ntot, int nchans, long t, int j)
{
    return t < 0 ? 0.f : x[t * nchans + j];
}

int txo_dedisp_window(const txo_plan_in *p, const float *x, long ntot, long w0, float *tim,
                      float *subout /* optional [ndm][nsub][N]: get_subdata of every DM */,
                      int nthreads)
{
    txo_plan_info o;
    int rc = txo_plan_info_get(p, &o);
    if (rc) return rc;
    if (w0 + o.nsamples + o.maxdelay > ntot) return TXO_ERANGE;
    const int C = p->nchans, nsub = o.nsub, cps = o.cps;
    const long N = o.nsamples;
    int *d1 = (int *)malloc(sizeof(int) * (size_t)o.ndmc * C);
    int *d2 = (int *)malloc(sizeof(int) * (size_t)p->ndm * nsub);
    if (!d1 || !d2) { free(d1); free(d2); return TXO_ENOMEM; }
    txo_plan_tables(p, d1, d2, NULL, NULL);
    /* stage 1 on the union of times stage 2 reads: [w0, w0 + N + D2) */
    const long L = N + o.D2;
    float *sub = (float *)malloc(sizeof(float) * (size_t)o.ndmc * nsub * L);
    if (!sub) { free(d1); free(d2); return TXO_ENOMEM; }
    (void)nthreads;
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) collapse(2) schedule(static)
    for (int k = 0; k < o.ndmc; k++) {
        for (int s = 0; s < nsub; s++) {
            float *ps = sub + ((size_t)k * nsub + s) * L;
            const int *dk = d1 + (size_t)k * C;
            for (long t = 0; t < L; t++) {
                float acc = 0.f;
                for (int j = s * cps; j < (s + 1) * cps; j++)
                    acc += xat(x, ntot, C, w0 + t + dk[j], j);
                ps[t] = acc;
            }
        }
    }
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(static)
    for (long i = 0; i < p->ndm; i++) {
        int k = (int)(i / o.group);
        const int *di = d2 + i * nsub;
        float *out = tim + i * N;
        for (long t = 0; t < N; t++) {
            float acc = 0.f;
            for (int s = 0; s < nsub; s++) acc += sub[((size_t)k * nsub + s) * L + t + di[s]];
            out[t] = acc;
        }
        if (subout) {
            for (int s = 0; s < nsub; s++)
                for (long t = 0; t < N; t++)
                    subout[((size_t)i * nsub + s) * N + t] =
                        sub[((size_t)k * nsub + s) * L + t + di[s]];
        }
    }
    free(sub);
    free(d1);
    free(d2);
    return TXO_OK;
}

/* txo_dedisp_window's rule (the same integer delays and sums) accumulated
 * in FP64: the reference the FP32 series' rounding error is stated against. */
int txo_dedisp_window_f64(const txo_plan_in *p, const float *x, long ntot, long w0, double *tim, int nthreads)
{
    txo_plan_info o;
    int rc = txo_plan_info_get(p, &o);
    if (rc) return rc;
    if (w0 + o.nsamples + o.maxdelay > ntot) return TXO_ERANGE;
    const int C = p->nchans, nsub = o.nsub, cps = o.cps;
    const long N = o.nsamples;
    int *d1 = (int *)malloc(sizeof(int) * (size_t)o.ndmc * C);
    int *d2 = (int *)malloc(sizeof(int) * (size_t)p->ndm * nsub);
    if (!d1 || !d2) { free(d1); free(d2); return TXO_ENOMEM; }
    txo_plan_tables(p, d1, d2, NULL, NULL);
    const long L = N + o.D2;
    double *sub = (double *)malloc(sizeof(double) * (size_t)o.ndmc * nsub * L);
    if (!sub) { free(d1); free(d2); return TXO_ENOMEM; }
    (void)nthreads;
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) collapse(2) schedule(static)
    for (int k = 0; k < o.ndmc; k++) {
        for (int s = 0; s < nsub; s++) {
            double *ps = sub + ((size_t)k * nsub + s) * L;
            const int *dk = d1 + (size_t)k * C;
            for (long t = 0; t < L; t++) {
                double acc = 0.;
                for (int j = s * cps; j < (s + 1) * cps; j++) acc += (double)xat(x, ntot, C, w0 + t + dk[j], j);
                ps[t] = acc;
            }
        }
    }
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(static)
    for (long i = 0; i < p->ndm; i++) {
        int k = (int)(i / o.group);
        const int *di = d2 + i * nsub;
        double *out = tim + i * N;
        for (long t = 0; t < N; t++) {
            double acc = 0.;
            for (int s = 0; s < nsub; s++) acc += sub[((size_t)k * nsub + s) * L + t + di[s]];
            out[t] = acc;
        }
    }
    free(sub);
    free(d1);
    free(d2);
    return TXO_OK;
}

/* Boxcar::match's S/N maximum (src/singlepulse/boxcar.cpp:201-281, default
 * statistics, AVX2 output ranges, negative S/N never kept) entirely in FP64:
 * mean and variance of the row, running sum, S = box / sqrt(wn * var). */
int txo_boxcar_f64(const double *tim, long ndm, long N, const int *vwn, int nbox, double *mxS, int nthreads)
{
    if (N <= 0 || nbox <= 0) return TXO_EINVAL;
    int rc = TXO_OK;
    (void)nthreads;
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(dynamic, 4)
    for (long r = 0; r < ndm; r++) {
        const double *x = tim + r * N;
        double *vS = mxS + r * N;
        double *cs = (double *)malloc(sizeof(double) * (size_t)N);
        if (!cs) {
            rc = TXO_ENOMEM;
            continue;
        }
        double mean = 0., var = 0.;
        for (long i = 0; i < N; i++) mean += x[i];
        mean /= (double)N;
        for (long i = 0; i < N; i++) var += (x[i] - mean) * (x[i] - mean);
        var /= (double)N;
        double c = 0.;
        for (long i = 0; i < N; i++) {
            c += x[i] - mean;
            cs[i] = c;
        }
        for (long i = 0; i < N; i++) vS[i] = 0.;
        if (var > 0.) {
            for (int k = 0; k < nbox; k++) {
                const int wn = vwn[k], wl = wn / 2 + 1, wh = (wn - 1) / 2;
                const long len = N - wh - wl, hi = wl + (len > 0 ? (len / 8) * 8 : 0);
                const double sc = 1. / sqrt((double)wn * var);
                for (long i = wl; i < hi; i++) {
                    const double S = (cs[i + wh] - cs[i - wl]) * sc;
                    if (S > vS[i]) vS[i] = S;
                }
            }
        }
        free(cs);
    }
    return rc;
}

/* Brute-force direct dedispersion in FP64 with the ideal integer delays
 * round(delay(dm, f_j)/tsamp) of src/singlepulse/candidate.cpp:110-114. */
int txo_dedisp_direct_f64(int nchans, const double *freqs, double tsamp, double dms, double ddm,
                          long ndm, const float *x, long ntot, long w0, long N, double *out)
{
    double fmax = 0.;
    for (int j = 0; j < nchans; j++) fmax = freqs[j] > fmax ? freqs[j] : fmax;
    for (long i = 0; i < ndm; i++) {
        double dm = dms + (double)i * ddm;
        for (long t = 0; t < N; t++) out[i * N + t] = 0.;
        for (int j = 0; j < nchans; j++) {
            long d = lround(txo_dmdelay(dm, fmax, freqs[j]) / tsamp);
            for (long t = 0; t < N; t++) {
                long tt = w0 + t + d;
                if (tt >= ntot) return TXO_ERANGE;
                if (tt >= 0) out[i * N + t] += (double)x[tt * nchans + j];
            }
        }
    }
    return TXO_OK;
}

/* ------------------------------------------------------------ a11-a14 */
static int cmp_float(const void *a, const void *b)
{
    float x = *(const float *)a, y = *(const float *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* Boxcar::match, src/singlepulse/boxcar.cpp:180-296.
 * variant 0 = the AVX2 branch (:263-280), the one the reference's configure
 * selects on AVX2 hosts (configure.ac:33): outputs only at
 * [wl, wl + 8*floor((N-wh-wl)/8)), blend = (S>=vS ? S : 0) + (S<vS ? vS : 0).
 * variant 1 = the scalar branch (:252-262): outputs at [wl, N-wh).
 * fma = 1 reproduces the FMA contraction of the -O3 -mfma build. */
int txo_boxcar_row(const float *tim, long N, const int *vwn, int nbox, int iqr, int variant,
                   int use_fma, float *vS, int *vwnS, float *mean_out, float *var_out)
{
    if (N <= 0 || nbox <= 0) return TXO_EINVAL;
    for (long i = 0; i < N; i++) { vS[i] = 0.f; vwnS[i] = 0; }
    float mean = 0.f, var = 0.f;
    if (!iqr) {
        for (long i = 0; i < N; i++) {
            mean += tim[i];
            var = use_fma ? fmaf(tim[i], tim[i], var) : var + tim[i] * tim[i];
        }
        mean /= (float)N;
        var /= (float)N;
        var = use_fma ? fmaf(-mean, mean, var) : var - mean * mean;
    } else {
        /* nth_element positions N/4, N/2 ascending and N/4 descending
         * (:217-226) are order statistics; a full sort gives the same values */
        float *c = (float *)malloc(sizeof(float) * (size_t)N);
        if (!c) return TXO_ENOMEM;
        memcpy(c, tim, sizeof(float) * (size_t)N);
        qsort(c, (size_t)N, sizeof(float), cmp_float);
        float Q1 = c[N / 4], Q2 = c[N / 2], Q3 = c[N - 1 - N / 4];
        free(c);
        mean = Q2;
        var = (float)(((double)(Q3 - Q1) / 1.349) * ((double)(Q3 - Q1) / 1.349));
    }
    if (mean_out) *mean_out = mean;
    if (var_out) *var_out = var;
    float *csump = (float *)malloc(sizeof(float) * (size_t)N);
    if (!csump) return TXO_ENOMEM;
    float csum = 0.f;
    for (long i = 0; i < N; i++) {
        csum += tim[i] - mean;
        csump[i] = csum;
    }
    if (var > 0) {
        for (int k = 0; k < nbox; k++) {
            int wn = vwn[k];
            int wl = wn / 2 + 1;
            int wh = (wn - 1) / 2;
            float temp = (float)sqrt(1. / (double)((float)wn * var));
            long lo = wl, hi;
            if (variant == 0) {
                long len = N - wh - wl;
                hi = wl + (len > 0 ? (len / 8) * 8 : 0);
            } else {
                hi = N - wh;
            }
            for (long i = lo; i < hi; i++) {
                float S = (csump[i + wh] - csump[i - wl]) * temp;
                if (variant == 0) {
                    int ge = S >= vS[i], lt = S < vS[i];
                    float nv = (ge ? S : 0.f) + (lt ? vS[i] : 0.f);
                    float nw = (ge ? (float)wn : 0.f) + (lt ? (float)vwnS[i] : 0.f);
                    vS[i] = nv;
                    vwnS[i] = (int)nw;
                } else if (S >= vS[i]) {
                    vS[i] = S;
                    vwnS[i] = wn;
                }
            }
        }
    }
    free(csump);
    return TXO_OK;
}

/* Boxcar::run, src/singlepulse/boxcar.cpp:158-178: every DM row, rows
 * independent (OpenMP over DMs as :166-175). */
int txo_boxcar(const float *tim, long ndm, long N, const int *vwn, int nbox, int iqr, int variant,
               int use_fma, float *mxS, int *mxwn, int nthreads)
{
    int rc = TXO_OK;
    (void)nthreads;
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(dynamic, 4)
    for (long k = 0; k < ndm; k++) {
        int r = txo_boxcar_row(tim + k * N, N, vwn, nbox, iqr, variant, use_fma, mxS + k * N,
                               mxwn + k * N, NULL, NULL);
        if (r) rc = r;
    }
    return rc;
}

/* The width margin of thresholded points, in the arithmetic of
 * txo_boxcar_row (default statistics, variant 0 ranges): for point p at
 * (idm[p], isamp[p]), best[p] = the S/N Boxcar::match keeps there and
 * second[p] = the largest S/N of any other width whose output range holds
 * the sample (-FLT_MAX if none).  A parallel boxcar whose sums differ from
 * the sequential FP32 ones by d can pick another width only where
 * best - second < 2d.  Points in any order. */
int txo_boxcar_runnerup(const float *tim, long ndm, long N, const int *vwn, int nbox, int use_fma,
                        long npts, const int *idm, const int *isamp, float *best, float *second)
{
    if (N <= 0 || nbox <= 0 || npts < 0) return TXO_EINVAL;
    float *csump = (float *)malloc(sizeof(float) * (size_t)N);
    if (!csump) return TXO_ENOMEM;
    long row = -1;
    float var = 0.f;
    for (long p = 0; p < npts; p++) {
        if (idm[p] < 0 || idm[p] >= ndm || isamp[p] < 0 || isamp[p] >= N) {
            free(csump);
            return TXO_EINVAL;
        }
        if (idm[p] != row) {
            row = idm[p];
            const float *x = tim + row * N;
            float mean = 0.f;
            var = 0.f;
            for (long i = 0; i < N; i++) {
                mean += x[i];
                var = use_fma ? fmaf(x[i], x[i], var) : var + x[i] * x[i];
            }
            mean /= (float)N;
            var /= (float)N;
            var = use_fma ? fmaf(-mean, mean, var) : var - mean * mean;
            float csum = 0.f;
            for (long i = 0; i < N; i++) {
                csum += x[i] - mean;
                csump[i] = csum;
            }
        }
        const long i = isamp[p];
        float vS = 0.f, Sk[256];
        int kwin = -1, ok[256];
        if (nbox > 256) {
            free(csump);
            return TXO_EINVAL;
        }
        for (int k = 0; k < nbox; k++) {
            const int wn = vwn[k], wl = wn / 2 + 1, wh = (wn - 1) / 2;
            const long len = N - wh - wl, hi = wl + (len > 0 ? (len / 8) * 8 : 0);
            ok[k] = var > 0 && i >= wl && i < hi;
            if (!ok[k]) continue;
            const float temp = (float)sqrt(1. / (double)((float)wn * var));
            Sk[k] = (csump[i + wh] - csump[i - wl]) * temp;
            if (Sk[k] >= vS) {
                vS = Sk[k];
                kwin = k;
            }
        }
        float s2 = -FLT_MAX;
        for (int k = 0; k < nbox; k++)
            if (ok[k] && k != kwin && Sk[k] > s2) s2 = Sk[k];
        best[p] = vS;
        second[p] = s2;
    }
    free(csump);
    return TXO_OK;
}

/* --savetim values, the rule declared for SubbandDedispersion::rundump
 * (called at src/singlepulse/singlepulse.cpp:340; XLibs holds the method):
 * the ndump new samples tim[i][noverlap ...] of every DM row, y = x*scale +
 * mean in float with scale = (float)(std/sqrt(nlive)); 8 and 16 bits round
 * half away from zero and clamp to the signed range; 32 bits keep y. */
void txo_dump(const float *tim, long ndm, long N, long noverlap, double nlive, float mean, float std, int nbits,
              void *out)
{
    const float scale = nlive > 0. ? (float)((double)std / sqrt(nlive)) : 0.f;
    const long n = N - noverlap;
    for (long i = 0; i < ndm; i++)
        for (long k = 0; k < n; k++) {
            const float p = tim[i * N + noverlap + k] * scale;
            const float y = p + mean;
            if (nbits == 32) {
                ((float *)out)[i * n + k] = y;
            } else {
                const float lo = nbits == 8 ? -128.f : -32768.f, hi = nbits == 8 ? 127.f : 32767.f;
                float r = roundf(y);
                r = r < lo ? lo : (r > hi ? hi : r);
                if (nbits == 8)
                    ((int8_t *)out)[i * n + k] = (int8_t)(int)r;
                else
                    ((int16_t *)out)[i * n + k] = (int16_t)(int)r;
            }
        }
}

/* ------------------------------------------------------------------ a15 */
/* src/singlepulse/cluster.cpp:89-112: points with mxS > thre (strict), in
 * row-major (DM, sample) order.  Returns the count (which may exceed cap;
 * only the first cap are stored). */
long txo_threshold(const float *mxS, const int *mxwn, long ndm, long N, float thre, int *idm,
                   int *isamp, float *S, int *wn, long cap)
{
    long n = 0;
    for (long j = 0; j < ndm; j++)
        for (long i = 0; i < N; i++) {
            float v = mxS[j * N + i];
            if (v > thre) {
                if (n < cap) {
                    idm[n] = (int)j;
                    isamp[n] = (int)i;
                    S[n] = v;
                    wn[n] = mxwn[j * N + i];
                }
                n++;
            }
        }
    return n;
}

/* ------------------------------------------------------------------ a16 */
/* Cluster<double>::run (src/singlepulse/cluster.cpp:63-161) with the
 * clustering() loop of include/clustering.h:26-100.  The KD-tree radius
 * query of XLibs is replaced by a uniform grid of cell size r; neighbours
 * are the points with squared distance <= r^2 (declared: inclusive).
 * Point coordinates: (dmdelay(dm, fmax, fmin)*0.5, i*tsamp) in seconds.
 * Outputs the kept candidates as point indices in candlist order and the
 * per-point cluster label (0 = not clustered).  Reproduces the index shift
 * of :145-158, where candcluster.erase() runs inside the loop over k. */
typedef struct {
    long *cell_start; /* hash table heads */
    long *next;
    long nbucket;
    double r;
    const double *x, *y;
} grid_t;

static long cell_of(double v, double r) { return (long)floor(v / r); }
static long hash2(long a, long b, long nb)
{
    uint64_t h = (uint64_t)a * 0x9E3779B97F4A7C15ull ^ ((uint64_t)b + 0x7F4A7C159E3779B9ull) * 0xC2B2AE3D27D4EB4Full;
    h ^= h >> 29;
    return (long)(h % (uint64_t)nb);
}

int txo_cluster(long npts, const int *idm, const int *isamp, const float *S, const int *wn,
                double dms, double ddm, double fmax, double fmin, double tsamp, double radius,
                int maxncand, int minpts, int drop, int maxwn, long *cand_pt, int *ncand_out,
                int *labels)
{
    *ncand_out = 0;
    for (long i = 0; i < npts; i++) labels[i] = 0;
    if (radius < 1.5 * tsamp) radius = 1.5 * tsamp; /* cluster.cpp:117 */
    double r2 = radius * radius;
    if (npts == 0) return TXO_OK;
    double *px = (double *)malloc(sizeof(double) * npts);
    double *py = (double *)malloc(sizeof(double) * npts);
    long nb = 2 * npts + 1;
    long *head = (long *)malloc(sizeof(long) * nb);
    long *next = (long *)malloc(sizeof(long) * npts);
    long *queue = (long *)malloc(sizeof(long) * 2 * npts);
    char *alive = (char *)malloc(npts);
    if (!px || !py || !head || !next || !queue || !alive) {
        free(px); free(py); free(head); free(next); free(queue); free(alive);
        return TXO_ENOMEM;
    }
    for (long i = 0; i < npts; i++) {
        double dm = dms + idm[i] * ddm;
        px[i] = txo_dmdelay(dm, fmax, fmin) * 0.5;
        py[i] = isamp[i] * tsamp;
        alive[i] = 1;
    }
    for (long b = 0; b < nb; b++) head[b] = -1;
    for (long i = npts - 1; i >= 0; i--) {
        long h = hash2(cell_of(px[i], radius), cell_of(py[i], radius), nb);
        next[i] = head[h];
        head[h] = i;
    }
    int icluster = 0;
    long remaining = npts;
    /* include/clustering.h:217: while ((icluster++ < maxncluster) && !idxs1.empty()) */
    while ((icluster++ < maxncand) && remaining > 0) {
        float maxsnr = -1;
        long seed = -1;
        for (long i = 0; i < npts; i++) /* idxs1 is kept sorted ascending */
            if (alive[i] && S[i] > maxsnr) { maxsnr = S[i]; seed = i; }
        if (seed < 0) break;
        /* clustering.h:44 keeps the seed index in a float: above 2^24
         * points the seed position is that of the rounded index */
        long pos = (long)(float)seed;
        long qh = 0, qt = 0;
        /* the seed's neighbours join whatever their state (clustering.h:56-63) */
        {
            long cx = cell_of(px[pos], radius), cy = cell_of(py[pos], radius);
            for (long dx = -1; dx <= 1; dx++)
                for (long dy = -1; dy <= 1; dy++) {
                    long h = hash2(cx + dx, cy + dy, nb);
                    for (long q = head[h]; q >= 0; q = next[q]) {
                        if (cell_of(px[q], radius) != cx + dx || cell_of(py[q], radius) != cy + dy) continue;
                        double ddx = px[q] - px[pos], ddy = py[q] - py[pos];
                        if (ddx * ddx + ddy * ddy <= r2) {
                            labels[q] = icluster;
                            if (alive[q]) { alive[q] = 0; remaining--; }
                            queue[qt++] = q;
                        }
                    }
                }
        }
        /* then the region grows through unclustered points (clustering.h:65-88) */
        while (qh < qt) {
            long p = queue[qh++];
            long cx = cell_of(px[p], radius), cy = cell_of(py[p], radius);
            for (long dx = -1; dx <= 1; dx++)
                for (long dy = -1; dy <= 1; dy++) {
                    long h = hash2(cx + dx, cy + dy, nb);
                    for (long q = head[h]; q >= 0; q = next[q]) {
                        if (!alive[q]) continue;
                        if (cell_of(px[q], radius) != cx + dx || cell_of(py[q], radius) != cy + dy) continue;
                        double ddx = px[q] - px[p], ddy = py[q] - py[p];
                        if (ddx * ddx + ddy * ddy <= r2) {
                            alive[q] = 0;
                            labels[q] = icluster;
                            remaining--;
                            queue[qt++] = q;
                        }
                    }
                }
        }
    }
    int ncluster = icluster - 1; /* tree.ncluster, clustering.h:275 */
    long *best = (long *)malloc(sizeof(long) * (ncluster > 0 ? ncluster : 1));
    float *bestS = (float *)malloc(sizeof(float) * (ncluster > 0 ? ncluster : 1));
    long *size = (long *)malloc(sizeof(long) * (ncluster > 0 ? ncluster : 1));
    for (int k = 0; k < ncluster; k++) { best[k] = -1; bestS[k] = 0.f; size[k] = 0; }
    for (long i = 0; i < npts; i++) {
        int c = labels[i];
        if (c == 0) continue;
        size[c - 1]++;
        if (S[i] > bestS[c - 1]) { bestS[c - 1] = S[i]; best[c - 1] = i; }
    }
    /* cluster.cpp:145-158: candcluster.erase(begin()+k) shifts the sizes
     * seen by later k.  An index past the shrunken vector reads as size 0. */
    long cur = ncluster;
    int nc = 0;
    for (int k = 0; k < ncluster; k++) {
        long bi = best[k];
        int keep;
        if (bi < 0) { keep = 0; }
        else if (drop && wn[bi] == maxwn) { keep = 0; }
        else {
            unsigned long sz = (k < cur) ? (unsigned long)size[k] : 0ul;
            keep = sz >= (unsigned long)(long)minpts;
        }
        if (keep) cand_pt[nc++] = bi;
        else if (k < cur) {
            memmove(size + k, size + k + 1, sizeof(long) * (size_t)(cur - k - 1));
            cur--;
        }
    }
    *ncand_out = nc;
    free(best); free(bestS); free(size);
    free(px); free(py); free(head); free(next); free(queue); free(alive);
    return TXO_OK;
}

/* Maximum-sum contiguous run, the XLibs kadane<float>(data, n, &start, &end)
 * that src/singlepulse/candplot.cpp:229 calls (declared: strict '>' keeps the
 * earliest maximum; the running sum restarts after it drops below zero). */
void txo_kadane(const float *a, long n, long *start, long *end)
{
    float best = -INFINITY, cur = 0.f;
    long s = 0, bs = 0, be = 0;
    for (long i = 0; i < n; i++) {
        cur += a[i];
        if (cur > best) { best = cur; bs = s; be = i; }
        if (cur < 0.f) { cur = 0.f; s = i + 1; }
    }
    *start = bs;
    *end = be;
}

/* fl/fh of a .cands line, src/singlepulse/candplot.cpp:141-231: vpf[j] is
 * the sum over samples isamp-wn < i < isamp+wn of sub[j][i]/fcnt[j]; the
 * kadane run over vpf picks (fl, fh) on the linear frequency axis vfnorm.
 * sub is get_subdata at the candidate's DM: [nsub][N]. */
void txo_cand_flfh(const float *sub, const double *fcnt, const double *fsub, int nsub, long N,
                   long isamp, int wn, double *fl, double *fh)
{
    float *vpf = (float *)calloc((size_t)nsub, sizeof(float));
    float *vfnorm = (float *)malloc(sizeof(float) * (size_t)nsub);
    long a0 = isamp - wn + 1, a1 = isamp + wn - 1;
    if (a0 < 0) a0 = 0;
    if (a1 > N - 1) a1 = N - 1;
    for (int j = 0; j < nsub; j++) {
        vfnorm[j] = (float)(fsub[0] + j / (nsub - 1.) * (fsub[nsub - 1] - fsub[0]));
        for (long i = a0; i <= a1; i++) vpf[j] += sub[(long)j * N + i] / (float)fcnt[j];
    }
    long fs = 0, fe = 0;
    txo_kadane(vpf, nsub, &fs, &fe);
    *fl = vfnorm[fs];
    *fh = vfnorm[fe];
    free(vpf);
    free(vfnorm);
}

/* ------------------------------------------------- streaming CPU search */
/* The CPU counterpart of one ddplan row run block by block, as
 * transientx_fil does (src/singlepulse/dedisperse_fil.cpp:319-338): the
 * dedisperser keeps rings of input (D1 history) and sub-band series
 * (D2 + N history), so each block costs stage 1 on its ndump new samples,
 * stage 2 on the N-sample window, then Boxcar::match (AVX2 variant, OpenMP
 * over DMs as src/singlepulse/boxcar.cpp:166-175) and the threshold scan.
 * Same declared rule and FP32 order as txo_dedisp_window. */
typedef struct {
    txo_plan_in in;
    txo_plan_info o;
    double *freqs;
    int *d1, *d2, *vwn;
    int nbox, zp, nthreads;
    long Lin, Lsub;
    float *xin;  /* [C][Lin] channel-major input ring */
    float *sub;  /* [ndmc][nsub][Lsub] */
    float *tim;  /* [ndm][N] */
    float *mxS;  /* [ndm][N] */
    int *mxwn;   /* [ndm][N] */
    long counter;
    long npoints;
} txo_stream;

void txo_stream_destroy(txo_stream *s)
{
    if (!s) return;
    free(s->freqs); free(s->d1); free(s->d2); free(s->vwn);
    free(s->xin); free(s->sub); free(s->tim); free(s->mxS); free(s->mxwn);
    free(s);
}

txo_stream *txo_stream_create(const txo_plan_in *p, const int *vwn, int nbox, int zp, int nthreads)
{
    txo_stream *s = (txo_stream *)calloc(1, sizeof(txo_stream));
    if (!s) return NULL;
    s->in = *p;
    s->freqs = (double *)malloc(sizeof(double) * p->nchans);
    memcpy(s->freqs, p->freqs, sizeof(double) * p->nchans);
    s->in.freqs = s->freqs;
    if (txo_plan_info_get(&s->in, &s->o)) { txo_stream_destroy(s); return NULL; }
    const txo_plan_info *o = &s->o;
    const long C = p->nchans, N = o->nsamples;
    s->d1 = (int *)malloc(sizeof(int) * o->ndmc * C);
    s->d2 = (int *)malloc(sizeof(int) * p->ndm * o->nsub);
    txo_plan_tables(&s->in, s->d1, s->d2, NULL, NULL);
    s->vwn = (int *)malloc(sizeof(int) * nbox);
    memcpy(s->vwn, vwn, sizeof(int) * nbox);
    s->nbox = nbox;
    s->zp = zp;
    s->nthreads = nthreads > 0 ? nthreads : 1;
    s->Lin = p->ndump + o->D1 + 4;
    s->Lsub = N + o->D2 + 4;
    s->xin = (float *)calloc((size_t)C * s->Lin, sizeof(float));
    s->sub = (float *)calloc((size_t)o->ndmc * o->nsub * s->Lsub, sizeof(float));
    s->tim = (float *)malloc(sizeof(float) * p->ndm * N);
    s->mxS = (float *)malloc(sizeof(float) * p->ndm * N);
    s->mxwn = (int *)malloc(sizeof(int) * p->ndm * N);
    if (!s->xin || !s->sub || !s->tim || !s->mxS || !s->mxwn) { txo_stream_destroy(s); return NULL; }
    return s;
}

static inline long wrap(long a, long m) { long r = a % m; return r < 0 ? r + m : r; }

/* acc[0..n) += ring[(base + t) mod L], as two contiguous runs (n <= L) */
static inline void ring_add(float *restrict acc, const float *restrict row, long L, long base, long n)
{
    const long st = wrap(base, L);
    const long n1 = n < L - st ? n : L - st;
    for (long t = 0; t < n1; t++) acc[t] += row[st + t];
    for (long t = n1; t < n; t++) acc[t] += row[t - n1];
}

static inline void ring_put(float *restrict row, long L, long base, const float *restrict v, long n)
{
    const long st = wrap(base, L);
    const long n1 = n < L - st ? n : L - st;
    for (long t = 0; t < n1; t++) row[st + t] = v[t];
    for (long t = n1; t < n; t++) row[t - n1] = v[t];
}

/* One block of ndump u8 samples, time-major.  Returns the number of points
 * above thre once the boxcar is ready (0 before), < 0 on error. */
long txo_stream_run_u8(txo_stream *s, const uint8_t *blk, float thre)
{
    const txo_plan_info *o = &s->o;
    const long C = s->in.nchans, nd = s->in.ndump, N = o->nsamples;
    const int nsub = o->nsub, cps = o->cps, nt = s->nthreads;
    const long c0 = s->counter;
    /* ingest: centre and transpose into the input ring */
#pragma omp parallel for num_threads(nt) schedule(static)
    for (long j = 0; j < C; j++) {
        float *row = s->xin + j * s->Lin;
        for (long t = 0; t < nd; t++) row[wrap(c0 + t, s->Lin)] = (float)((int)blk[t * C + j] - s->zp);
    }
    s->counter += nd;
    const long c = s->counter;
    /* stage 1 on the new sub-band times */
    const long tb = c - nd - o->D1;
#pragma omp parallel for num_threads(nt) collapse(2) schedule(static)
    for (int k = 0; k < o->ndmc; k++)
        for (int sb = 0; sb < nsub; sb++) {
            float acc[1024];
            float *out = s->sub + ((size_t)k * nsub + sb) * s->Lsub;
            for (long t0 = 0; t0 < nd; t0 += 1024) {
                const long n = nd - t0 < 1024 ? nd - t0 : 1024;
                for (long t = 0; t < n; t++) acc[t] = 0.f;
                for (int j = sb * cps; j < (sb + 1) * cps; j++)
                    ring_add(acc, s->xin + (size_t)j * s->Lin, s->Lin,
                             tb + t0 + s->d1[(size_t)k * C + j], n);
                ring_put(out, s->Lsub, tb + t0, acc, n);
            }
        }
    /* stage 2 on the window */
    const long w0 = c - o->maxdelay - N;
#pragma omp parallel for num_threads(nt) schedule(static)
    for (long i = 0; i < s->in.ndm; i++) {
        const int k = (int)(i / o->group);
        float *out = s->tim + i * N;
        for (long t = 0; t < N; t++) out[t] = 0.f;
        for (int sb = 0; sb < nsub; sb++)
            ring_add(out, s->sub + ((size_t)k * nsub + sb) * s->Lsub, s->Lsub,
                     w0 + s->d2[i * nsub + sb], N);
    }
    if (c < o->offset - o->noverlap + nd) return 0; /* src/singlepulse/boxcar.cpp:161 */
    txo_boxcar(s->tim, s->in.ndm, N, s->vwn, s->nbox, 0, 0, 1, s->mxS, s->mxwn, nt);
    long np = 0;
    for (long q = 0; q < s->in.ndm * N; q++) np += s->mxS[q] > thre;
    s->npoints = np;
    return np;
}

const float *txo_stream_tim(const txo_stream *s) { return s->tim; }
const float *txo_stream_mxS(const txo_stream *s) { return s->mxS; }
const int *txo_stream_mxwn(const txo_stream *s) { return s->mxwn; }

/* ----------------------------------------------------------------- a17 */
/* The refinement grid of the candidate replot (src/singlepulse/replot_fil.cpp:419-458):
 * Candidate::dedisperse (incoherent), Candidate::optimize with the
 * dedispersion tree of include/hough_transform.h, Candidate::matched_filter.
 * The reference build contracts a + b*c into FMA (configure.ac:14 plus
 * config/ax_check_x86_features.m4:76 add -O3 -mavx2 -mfma), so the two such
 * expressions below use fma(). */

/* Candidate::dedisperse, incoherent branch (src/singlepulse/candidate.cpp:108-140):
 * channel j is rotated left by round(dmdelay(dm, fmax, f_j)/tbin) mod nbin. */
void txo_cand_dedisperse(float *data, int nchan, long nbin, const double *freqs, double dm,
                         double tbin)
{
    double fmax = freqs[0];
    for (int j = 1; j < nchan; j++) fmax = freqs[j] > fmax ? freqs[j] : fmax;
    float *tmp = (float *)malloc(sizeof(float) * (size_t)nbin);
    for (int j = 0; j < nchan; j++) {
        int d = (int)round(txo_dmdelay(dm, fmax, freqs[j]) / tbin);
        d = (int)(d % nbin);
        if (d < 0) d += (int)nbin;
        float *row = data + (size_t)j * nbin;
        memcpy(tmp, row, sizeof(float) * (size_t)nbin);
        for (long i = 0; i < nbin; i++) row[i] = tmp[(i + d) % nbin];
    }
    free(tmp);
}

static long bitrev(long k, int nbits)
{
    long r = 0;
    for (int b = 0; b < nbits; b++) r |= ((k >> b) & 1) << (nbits - 1 - b);
    return r;
}

/* Shift of the tree at level l for the channel pair (x0, x1) and DM trial p:
 * round(dmdelay(dms + p*ddm - dm0, f[x0], f[x1])/tbin) mod ncol
 * (the get_shift lambda of src/singlepulse/candidate.cpp:970-974 and the
 * modulo of include/hough_transform.h:88-89). */
static int tree_shift(const double *fP, int x0, int x1, long p, double dms, double ddm, double dm0,
                      double tbin, int ncol)
{
    const double dm = fma((double)p, ddm, dms) - dm0;
    int s = (int)round(txo_dmdelay(dm, fP[x0], fP[x1]) / tbin);
    s %= ncol;
    if (s < 0) s += ncol;
    return s;
}

/* The dedispersion tree of include/hough_transform.h:20-125, level by level.
 * After level l (l = 0..n-1, n = log2 nrow) the rows are grouped in nodes of
 * 2^(l+1) channels; node j (channels [j*2^(l+1), (j+1)*2^(l+1))) holds
 * 2^(l+1) rows, row k standing for DM trial bitrev_n(k).  With half = 2^l,
 *   row k of node j = A[k mod half][i] + B[k mod half][(i + s) mod ncol]
 * where A, B are the two child nodes (channels x0 = j*2^(l+1) and
 * x1 = x0 + half lead them) and s = tree_shift(x0, x1, bitrev_n(k)).
 * out row bitrev_n(k) = root row k.  dm0 = 0 unless the data were already
 * dedispersed at dm0 (Candidate::isdedispersed). */
int txo_hough(const float *in, float *out, int nrow, int ncol, const double *fP, double dms,
              double ddm, double tbin, double dm0)
{
    int n = 0;
    while ((1 << n) < nrow) n++;
    if ((1 << n) != nrow || ncol <= 0) return TXO_EINVAL;
    const size_t sz = (size_t)nrow * ncol;
    float *cur = (float *)malloc(sizeof(float) * sz);
    float *nxt = (float *)malloc(sizeof(float) * sz);
    if (!cur || !nxt) {
        free(cur);
        free(nxt);
        return TXO_ENOMEM;
    }
    memcpy(cur, in, sizeof(float) * sz);
    for (int l = 0; l < n; l++) {
        const int half = 1 << l, width = 2 * half, nodes = nrow / width;
        for (int j = 0; j < nodes; j++) {
            const int x0 = j * width, x1 = x0 + half;
            for (int k = 0; k < width; k++) {
                const int s = tree_shift(fP, x0, x1, bitrev(k, n), dms, ddm, dm0, tbin, ncol);
                const float *A = cur + ((size_t)(2 * j) * half + (k % half)) * ncol;
                const float *B = cur + ((size_t)(2 * j + 1) * half + (k % half)) * ncol;
                float *o = nxt + ((size_t)j * width + k) * ncol;
                for (int i = 0; i < ncol; i++) o[i] = A[i] + B[(i + s) % ncol];
            }
        }
        float *t = cur;
        cur = nxt;
        nxt = t;
    }
    for (long k = 0; k < nrow; k++)
        memcpy(out + (size_t)bitrev(k, n) * ncol, cur + (size_t)k * ncol, sizeof(float) * ncol);
    free(cur);
    free(nxt);
    return TXO_OK;
}

/* Candidate::optimize (src/singlepulse/candidate.cpp:899-977) for one
 * polarisation: DM range nbin*tbin/dmdelay(1, fmax, fmin) centred on dm
 * (clipped at 0), nrow = 2^ceil(log2 nchan) trials, channels ordered from
 * the highest frequency, zero rows past nchan at extrapolated frequencies.
 * dmtmap: [nrow][nbin]. */
int txo_refine_optimize(const float *data, int nchan, int nbin, const double *freqs, double tbin,
                        double dm, int isdedispersed, float *dmtmap, int nrow_cap, double *dms_out,
                        double *ddm_out, int *ndm_out)
{
    double fmax = freqs[0], fmin = freqs[0];
    for (int j = 1; j < nchan; j++) {
        fmax = freqs[j] > fmax ? freqs[j] : fmax;
        fmin = freqs[j] < fmin ? freqs[j] : fmin;
    }
    const double tmp = txo_dmdelay(1., fmax, fmin);
    const double dmrange = (nbin * tbin) / tmp;
    const double dms = dm - dmrange / 2. > 0. ? dm - dmrange / 2. : 0.;
    const double dme = dm + dmrange / 2.;
    const int nrow = (int)pow(2, ceil(log2((double)nchan)));
    if (nrow > nrow_cap) return TXO_ERANGE;
    const double ddm = (dme - dms) / nrow;
    float *buf = (float *)calloc((size_t)nrow * nbin, sizeof(float));
    double *fP = (double *)malloc(sizeof(double) * nrow);
    const double foff = freqs[1] - freqs[0];
    for (int j = 0; j < nchan; j++) {
        const int src = foff < 0. ? j : nchan - 1 - j;
        for (int i = 0; i < nbin; i++) buf[(size_t)j * nbin + i] += data[(size_t)src * nbin + i];
        fP[j] = freqs[src];
    }
    for (int j = nchan; j < nrow; j++)
        fP[j] = foff < 0. ? fma((double)j, foff, freqs[0]) : fma((double)(nchan - 1 - j), foff, freqs[0]);
    int rc = txo_hough(buf, dmtmap, nrow, nbin, fP, dms, ddm, tbin, isdedispersed ? dm : 0.);
    free(buf);
    free(fP);
    *dms_out = dms;
    *ddm_out = ddm;
    *ndm_out = nrow;
    return rc;
}

/* get_mean_var (XLibs, called at src/singlepulse/candidate.cpp:1010).
 * Declared: sequential FP64 sums of x and x*x, var = E[x^2] - mean^2, the
 * last step contracted to an FMA as the reference's -mfma build contracts
 * `var -= mean*mean` (x*x of a float is exact in FP64, so that sum is the
 * same either way). */
void txo_mean_var(const float *x, long n, double *mean, double *var)
{
    double m = 0., v = 0.;
    for (long i = 0; i < n; i++) {
        m += (double)x[i];
        v += (double)x[i] * (double)x[i];
    }
    m /= n;
    v /= n;
    v = fma(-m, m, v);
    *mean = m;
    *var = v;
}

/* The width ladder of Candidate::matched_filter (src/singlepulse/candidate.cpp:981-993):
 * from 1 up to nbin/2, held in floats. */
int txo_mf_ladder(long nbin, float snrloss, int *vwn, int cap)
{
    const float wfactor = (float)(1. / ((1. - (double)snrloss) * (1. - (double)snrloss)));
    int n = 0;
    float last = 1.f;
    vwn[n++] = 1;
    for (;;) {
        int w1 = (int)last;
        int w2 = (int)((float)w1 * wfactor);
        w2 = w2 < w1 + 1 ? w1 + 1 : w2;
        if (w2 > nbin / 2) break;
        if (n >= cap) return TXO_ERANGE;
        last = (float)w2;
        vwn[n++] = w2;
    }
    return n;
}

/* Candidate::matched_filter (src/singlepulse/candidate.cpp:979-1087): per DM
 * row a circular boxcar (running sum over 2*nbin samples, each step
 * (double)csum + (x - mean) rounded to float), S >= keeps the wider width;
 * then the first maximum of the map in row-major order. */
int txo_refine_mf(const float *dmtmap, int ndm, long nbin, float snrloss, float *snrmap,
                  int *widmap, int *maxdmid, long *maxtid, float *snr)
{
    int vwn[64];
    const int nbox = txo_mf_ladder(nbin, snrloss, vwn, 64);
    if (nbox < 0) return nbox;
    float *cs = (float *)malloc(sizeof(float) * 2 * (size_t)nbin);
    for (int j = 0; j < ndm; j++) {
        const float *tim = dmtmap + (size_t)j * nbin;
        float *vS = snrmap + (size_t)j * nbin;
        int *vw = widmap + (size_t)j * nbin;
        double mean, var;
        txo_mean_var(tim, nbin, &mean, &var);
        float csum = 0.f;
        for (long i = 0; i < 2 * nbin; i++) {
            csum = (float)((double)csum + ((double)tim[i % nbin] - mean));
            cs[i] = csum;
        }
        for (long i = 0; i < nbin; i++) {
            vS[i] = 0.f;
            vw[i] = 0;
        }
        if (var > 0) {
            for (int k = 0; k < nbox; k++) {
                const int wn = vwn[k], wl = wn / 2 + 1, wh = (wn - 1) / 2;
                const float temp = (float)sqrt(1. / (wn * var));
                for (long i = wl; i < nbin + wl; i++) {
                    const float S = (cs[i + wh] - cs[i - wl]) * temp;
                    const long o = i < nbin ? i : i - nbin;
                    if (S >= vS[o]) {
                        vS[o] = S;
                        vw[o] = wn;
                    }
                }
            }
        }
    }
    free(cs);
    int bj = 0;
    long bi = 0;
    float best = snrmap[0];
    for (int j = 0; j < ndm; j++)
        for (long i = 0; i < nbin; i++)
            if (snrmap[(size_t)j * nbin + i] > best) {
                best = snrmap[(size_t)j * nbin + i];
                bj = j;
                bi = i;
            }
    *maxdmid = bj;
    *maxtid = bi;
    *snr = best;
    return nbox;
}

/* ---- RFI cleaning of one frame [nchan][nbin] (channel-major), in place ----
 * The reference's build (g++ -O3 -mavx2 -mfma) contracts a*b + c and
 * c - a*b into FMA, and in a*b - c*d rounds c*d first and fuses a*b; those
 * contractions are written out with fma() below, nothing else fuses. */

/* Candidate::zdot, src/singlepulse/candidate.cpp:831-897. */
void txo_zdot(float *data, long nchan, long nbin)
{
    double *s = (double *)calloc((size_t)nbin, sizeof(double));
    for (long j = 0; j < nchan; j++) {
        float *x = data + j * nbin;
        double m = 0., sd = 0.;
        for (long i = 0; i < nbin; i++) {
            m += x[i];
            sd += x[i] * x[i];          /* a float product */
        }
        m /= nbin;
        sd /= nbin;
        sd = fma(-m, m, sd);            /* tmpstd -= tmpmean*tmpmean */
        sd = sqrt(sd);
        if (sd == 0.) sd = 1.;
        for (long i = 0; i < nbin; i++) {
            x[i] -= m;
            x[i] /= sd;
            s[i] += x[i];
        }
    }
    for (long i = 0; i < nbin; i++) s[i] /= (int)nchan;
    for (long j = 0; j < nchan; j++) {
        float *x = data + j * nbin;
        double xe = 0., xs = 0., se = 0., ss = 0., alpha = 0., beta = 0.;
        for (long i = 0; i < nbin; i++) {
            xe += x[i];
            xs = fma((double)x[i], s[i], xs);
            se += s[i];
            ss = fma(s[i], s[i], ss);
        }
        const double nb = (double)nbin;
        const double tmp = fma(se, se, -(ss * nb));
        if (tmp != 0) {
            alpha = fma(xe, se, -(xs * nb)) / tmp;
            beta = fma(xs, se, -(xe * ss)) / tmp;
        }
        for (long i = 0; i < nbin; i++) x[i] -= fma(alpha, s[i], beta);
    }
    free(s);
}

/* kadane<float> as the kadane filter calls it, returning the run's sum
 * (declared as txo_kadane). */
static float txo_kadane_sum(const float *a, long n, long *start, long *end)
{
    float best = -INFINITY, cur = 0.f;
    long s = 0, bs = 0, be = 0;
    for (long i = 0; i < n; i++) {
        cur += a[i];
        if (cur > best) { best = cur; bs = s; be = i; }
        if (cur < 0.f) { cur = 0.f; s = i + 1; }
    }
    *start = bs;
    *end = be;
    return best;
}

/* Candidate::kadaneF, src/singlepulse/candidate.cpp:710-829, with wnlimit
 * the caller's int(nwidth*width/tbin/td). */
void txo_kadanef(float *data, long nchan, long nbin, int td, int fd, float threshold, int wnlimit)
{
    const long nsd = nbin / td, ncd = nchan / fd;
    if (nsd == 0 || ncd == 0) return;
    float *ds = (float *)calloc((size_t)(ncd * nsd), sizeof(float));
    float *ch = (float *)malloc((size_t)nsd * sizeof(float));
    for (long j = 0; j < ncd; j++)
        for (long k = 0; k < fd; k++)
            for (long n = 0; n < td; n++)
                for (long i = 0; i < nsd; i++) ds[j * nsd + i] += data[(j * fd + k) * nbin + (i * td + n)];
    for (long j = 0; j < ncd; j++) {
        for (long i = 0; i < nsd; i++) ch[i] = ds[j * nsd + i];
        for (int pass = 0; pass < 2; pass++) {
            if (pass == 1)
                for (long i = 0; i < nsd; i++) ch[i] = -ch[i];
            long start, end;
            const float box = txo_kadane_sum(ch, nsd, &start, &end);
            const long wn = end - start + 1;
            const float var = (float)(td * fd);
            float snr2 = 0.f;
            if (wn > wnlimit) snr2 = box * box / (wn * var);
            start *= td;
            end = (end + 1) * td;
            if (snr2 > threshold * threshold)
                for (long k = 0; k < fd; k++)
                    for (long i = start; i < end; i++) data[(j * fd + k) * nbin + i] = 0.f;
        }
    }
    free(ch);
    free(ds);
}

/* ----------------------------------------------- front end (8-bit ingest) */
/* The front end's Downsample then Equalize on one 8-bit block, as declared
 * (XLibs' classes, called at src/singlepulse/singlepulse.cpp:289-291, are
 * not in the TransientX tree).  Output channel c < nchi/fd, sample
 * t < nsi/td: y = the td x fd neighbours summed (docs/pipeline/
 * preprocess.rst:5-9), in the loop order of Candidate::downsample
 * (src/singlepulse/candidate.cpp:381-405) -- an integer sum of 8-bit values,
 * exact in any order, equal to that float sum while y < 2^24 -- then
 * (y - off[c]) * gain[c] in float (two roundings).  Masked output channels
 * (mask[c] == 0) are 0.  out is time-major [nsi/td][nchi/fd], the
 * DataBuffer layout. */
void txo_frontend_u8(const uint8_t *in, long nsi, int nchi, int td, int fd, const float *off,
                     const float *gain, const uint8_t *mask, float *out)
{
    const long nso = nsi / td;
    const int ncho = nchi / fd;
    for (long t = 0; t < nso; t++)
        for (int c = 0; c < ncho; c++) {
            long y = 0;
            for (int jj = 0; jj < fd; jj++)
                for (int ii = 0; ii < td; ii++) y += in[(t * td + ii) * nchi + (long)c * fd + jj];
            const float d = (float)y - off[c];
            const float v = d * gain[c];
            out[t * ncho + c] = (mask && !mask[c]) ? 0.f : v;
        }
}

/* Equalize's statistics, as declared: per output channel, over the block's
 * nsi/td downsampled samples y, the exact integer sums S1 = sum y and
 * S2 = sum y^2; mean = S1/n, var = S2/n - mean*mean (double, no
 * contraction); off = (float)mean, gain = var > 0 ? (float)(1/sqrt(var)) : 1
 * (a constant channel is left unscaled, as zdot does with sd 0). */
void txo_frontend_stats(const uint8_t *in, long nsi, int nchi, int td, int fd, float *off, float *gain)
{
    const long nso = nsi / td;
    const int ncho = nchi / fd;
    for (int c = 0; c < ncho; c++) {
        unsigned long long s1 = 0, s2 = 0;
        for (long t = 0; t < nso; t++) {
            unsigned long long y = 0;
            for (int jj = 0; jj < fd; jj++)
                for (int ii = 0; ii < td; ii++) y += in[(t * td + ii) * nchi + (long)c * fd + jj];
            s1 += y;
            s2 += y * y;
        }
        const double n = (double)nso;
        const double mean = (double)s1 / n;
        const double m2 = mean * mean;
        const double var = (double)s2 / n - m2;
        off[c] = (float)mean;
        gain[c] = var > 0. ? (float)(1. / sqrt(var)) : 1.f;
    }
}
