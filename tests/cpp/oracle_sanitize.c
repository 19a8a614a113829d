/* The oracle (oracle/txoracle.c, test infrastructure) built with
 * AddressSanitizer and UndefinedBehaviorSanitizer and driven over small
 * random cases of every family of functions: plan tables, windowed and FP64
 * dedispersion, the boxcar variants (AVX2 tail, IQR, FMA or not), the
 * threshold scan, clustering, kadane, the streaming search, --savetim
 * values, the refinement tree and matched filter, zdot, kadaneF and the
 * front end.  Sizes cover ragged and degenerate shapes (one sub-band, N
 * smaller than the widest box, constant and NaN rows). */
#include "../../oracle/txoracle.c"

#include <stdio.h>

static unsigned long long rs = 12345;
static double urand(void)
{
    rs = rs * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(rs >> 11) / 9007199254740992.0;
}

#define CHECK(c)                                                                                   \
    do {                                                                                           \
        if (!(c)) {                                                                                \
            fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c);                          \
            return 1;                                                                              \
        }                                                                                          \
    } while (0)

static int plans(void)
{
    const int shapes[][4] = {{64, 8, 256, 20}, {32, 32, 300, 7}, {48, 1, 128, 5}, {16, 16, 64, 3}};
    for (int c = 0; c < 4; c++) {
        const int C = shapes[c][0], nsub = shapes[c][1];
        const long ndump = shapes[c][2], ndm = shapes[c][3];
        double *f = malloc(sizeof(double) * C);
        for (int j = 0; j < C; j++) f[j] = 1500. - j * (300. / C);
        txo_plan_in p = {C, f, 1e-3, ndump, 0.25, 10., 2., ndm, nsub, 0};
        txo_plan_info o;
        CHECK(txo_plan_info_get(&p, &o) == TXO_OK);
        int *d1 = malloc(sizeof(int) * o.ndmc * C), *d2 = malloc(sizeof(int) * ndm * o.nsub);
        double *dmc = malloc(sizeof(double) * o.ndmc), *fref = malloc(sizeof(double) * o.nsub);
        CHECK(txo_plan_tables(&p, d1, d2, dmc, fref) == TXO_OK);
        const long ntot = o.nsamples + o.maxdelay + 17;
        float *x = malloc(sizeof(float) * ntot * C);
        for (long i = 0; i < ntot * C; i++) x[i] = (float)(urand() * 20. - 10.);
        float *tim = malloc(sizeof(float) * ndm * o.nsamples);
        float *sub = malloc(sizeof(float) * ndm * o.nsub * o.nsamples);
        double *t64 = malloc(sizeof(double) * ndm * o.nsamples);
        CHECK(txo_dedisp_window(&p, x, ntot, 5, tim, sub, 1) == TXO_OK);
        CHECK(txo_dedisp_window(&p, x, ntot, -7, tim, NULL, 1) == TXO_OK); /* window before the start */
        CHECK(txo_dedisp_window_f64(&p, x, ntot, 5, t64, 1) == TXO_OK);
        int vwn[40];
        const int nbox = txo_ladder(1e-3, 1e-3, 0.05f, 0.1f, vwn, 40);
        CHECK(nbox > 0);
        float *mxS = malloc(sizeof(float) * ndm * o.nsamples);
        int *mxwn = malloc(sizeof(int) * ndm * o.nsamples);
        double *m64 = malloc(sizeof(double) * ndm * o.nsamples);
        for (int iqr = 0; iqr < 2; iqr++)
            for (int variant = 0; variant < 2; variant++)
                CHECK(txo_boxcar(tim, ndm, o.nsamples, vwn, nbox, iqr, variant, variant, mxS, mxwn, 1) == TXO_OK);
        CHECK(txo_boxcar_f64(t64, ndm, o.nsamples, vwn, nbox, m64, 1) == TXO_OK);
        const long cap = ndm * o.nsamples;
        int *idm = malloc(sizeof(int) * cap), *isamp = malloc(sizeof(int) * cap), *wn = malloc(sizeof(int) * cap);
        float *S = malloc(sizeof(float) * cap);
        const long np = txo_threshold(mxS, mxwn, ndm, o.nsamples, 1.5f, idm, isamp, S, wn, cap);
        CHECK(np >= 0);
        long *cand = malloc(sizeof(long) * (np + 1));
        int *labels = malloc(sizeof(int) * (np + 1));
        int nc = 0;
        CHECK(txo_cluster(np, idm, isamp, S, wn, 10., 2., o.fmax, o.fmin, 1e-3, 2e-3, 50, 2, 1, vwn[nbox - 1], cand,
                          &nc, labels) == TXO_OK);
        unsigned char dump8[4096];
        if (ndm * ndump <= 4096) txo_dump(tim, ndm, o.nsamples, o.noverlap, (double)C, 0.f, 1.f, 8, dump8);
        free(f); free(d1); free(d2); free(dmc); free(fref); free(x); free(tim); free(sub); free(t64);
        free(mxS); free(mxwn); free(m64); free(idm); free(isamp); free(wn); free(S); free(cand); free(labels);
    }
    return 0;
}

static int boxcar_edges(void)
{
    /* N shorter than the widest box, constant rows, a NaN row */
    int vwn[40];
    const int nbox = txo_ladder(1e-3, 1e-3, 0.05f, 0.1f, vwn, 40);
    for (long N = 1; N < 80; N += 13) {
        float tim[3 * 80];
        for (long i = 0; i < N; i++) {
            tim[i] = 3.f;
            tim[N + i] = (float)urand();
            tim[2 * N + i] = i == N / 2 ? NAN : (float)urand();
        }
        float mxS[3 * 80];
        int mxwn[3 * 80];
        CHECK(txo_boxcar(tim, 3, N, vwn, nbox, 0, 1, 1, mxS, mxwn, 1) == TXO_OK);
        CHECK(txo_boxcar(tim, 3, N, vwn, nbox, 1, 0, 0, mxS, mxwn, 1) == TXO_OK);
    }
    long s, e;
    float a[5] = {-1, 2, -1, 3, -9};
    txo_kadane(a, 5, &s, &e);
    txo_kadane(a, 0, &s, &e);
    return 0;
}

static int streams(void)
{
    const int C = 32;
    double f[32];
    for (int j = 0; j < C; j++) f[j] = 1500. - j * 10.;
    txo_plan_in p = {C, f, 1e-3, 200, 0.1, 0., 3., 12, 8, 0};
    int vwn[40];
    const int nbox = txo_ladder(1e-3, 1e-3, 0.03f, 0.1f, vwn, 40);
    txo_stream *st = txo_stream_create(&p, vwn, nbox, 128, 1);
    CHECK(st != NULL);
    unsigned char blk[200 * 32];
    for (int b = 0; b < 9; b++) {
        for (int i = 0; i < 200 * 32; i++) blk[i] = (unsigned char)(urand() * 256.);
        CHECK(txo_stream_run_u8(st, blk, 4.f) >= 0);
    }
    txo_stream_destroy(st);
    return 0;
}

static int refinement(void)
{
    const int nchan = 16, nbin = 64;
    double f[16];
    for (int j = 0; j < nchan; j++) f[j] = 1500. - j * 20.;
    float data[16 * 64];
    for (int i = 0; i < nchan * nbin; i++) data[i] = (float)(urand() - 0.5);
    txo_cand_dedisperse(data, nchan, nbin, f, 50., 1e-3);
    float map[64 * 64];
    double dms = 0., ddm = 0.;
    int ndm = 0;
    CHECK(txo_refine_optimize(data, nchan, nbin, f, 1e-3, 50., 1, map, 64, &dms, &ddm, &ndm) == TXO_OK);
    float snrmap[64 * 64], snr = 0.f;
    int widmap[64 * 64], maxdmid = 0;
    long maxtid = 0;
    CHECK(txo_refine_mf(map, ndm, nbin, 0.1f, snrmap, widmap, &maxdmid, &maxtid, &snr) > 0); /* its ladder length */
    /* zdot and kadaneF in place, the front end and its statistics */
    float fr[16 * 64];
    for (int i = 0; i < nchan * nbin; i++) fr[i] = (float)(urand() * 4.);
    for (int i = 0; i < nbin; i++) fr[3 * nbin + i] = 1.f; /* a constant channel */
    txo_zdot(fr, nchan, nbin);
    txo_kadanef(fr, nchan, nbin, 4, 2, 3.f, 0);
    txo_kadanef(fr, nchan, nbin, 64, 16, 0.5f, 2);
    unsigned char raw[37 * 12];
    for (int i = 0; i < 37 * 12; i++) raw[i] = (unsigned char)(urand() * 256.);
    float off[12], gain[12], out[37 * 12];
    txo_frontend_stats(raw, 37, 12, 3, 2, off, gain);
    txo_frontend_u8(raw, 37, 12, 3, 2, off, gain, NULL, out);
    return 0;
}

int main(void)
{
    if (plans() || boxcar_edges() || streams() || refinement()) return 1;
    printf("oracle_sanitize ok\n");
    return 0;
}
