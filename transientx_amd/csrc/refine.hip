// refine.hip -- the candidate refinement grid on gfx950: Candidate::dedisperse
// (incoherent), Candidate::optimize with the dedispersion tree of
// include/hough_transform.h, and Candidate::matched_filter
// (src/singlepulse/candidate.cpp:108-140,899-1087), for a batch of cutouts.
//
//   R1 k_tree_pass   a run of tree levels [a, b) in LDS.  After level l the
//                    rows form nodes of 2^(l+1) channels; row k of node j is
//                    A[k mod 2^l][i] + B[k mod 2^l][(i + s) mod ncol] with s
//                    the rounded delay between the nodes' leading channels at
//                    DM trial bitrev(k).  Row k of a node after level b-1
//                    only needs row (k mod 2^a) of each of its 2^(b-a)
//                    descendants at level a-1, so one workgroup takes such a
//                    group of rows through all levels of the pass.  The
//                    first pass loads the cutout (band order, padding and the
//                    incoherent shift fused in); the last one writes the DM
//                    rows in trial order.
//   R2 k_mf          one lane per DM row for the sequential parts (FP64
//                    mean/variance, the FP32 circular running sum, in the
//                    reference's order), then the whole wavefront per row for
//                    the width loop; per-row first maximum.
//   R3 k_argmax      first maximum of each cutout's S/N map (row-major).
//
// All sums are elementwise FP32 adds (tree) or sequential in the reference's
// order (matched filter), so the maps are bit-identical to the restatement.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/txgpu.h"
#include "device_util.h"

namespace txgpu {
int set_error(int code, const std::string &msg);
}
using txgpu::lds_fill;
using txgpu::lds_fill_done;
using txgpu::set_error;

namespace {

constexpr int kTreeLdsFloats = 16384; // 64 KB of rows per tree workgroup
constexpr int kMaxItems = 16;         // pair items per thread per chunk (256 threads)
constexpr int kMaxMfBox = 64;

// include/candidate.h:116-119, no contraction (the reference computes it as written)
__device__ __host__ inline double dmdelay(double dm, double fh, double fl)
{
#pragma clang fp contract(off)
    return 4.148741601e3 * dm * (1. / (fl * fl) - 1. / (fh * fh));
}

// get_shift of src/singlepulse/candidate.cpp:970-974 (dms + p*ddm contracted
// to an FMA as the reference's -mfma build does) and the modulo of
// include/hough_transform.h:88-89.
__device__ inline int tree_shift(double fh, double fl, int p, double dms, double ddm, double dm0,
                                 double tbin, int ncol)
{
#pragma clang fp contract(off)
    const double dm = fma((double)p, ddm, dms) - dm0;
    int s = (int)round(dmdelay(dm, fh, fl) / tbin);
    s %= ncol;
    if (s < 0) s += ncol;
    return s;
}

__device__ inline int bitrev(int k, int n) { return (int)(__builtin_bitreverse32((unsigned)k) >> (32 - n)); }

struct TreeArgs {
    const float *src;   // first pass: cutouts [ncand][nchan][ncol]; else rows [ncand][nrow][ncol]
    float *dst;         // rows [ncand][nrow][ncol] (node layout, or DM order in the last pass)
    const double *fP;   // [nrow] frequency of buffer row j (from the highest)
    const int *srcrow;  // [nrow] cutout channel of buffer row j, -1 for padding
    const double *fsrc; // [nchan] cutout channel frequencies
    const double *dms, *ddm, *dm0, *dmc; // per cutout
    double fmax, tbin;
    int nchan, nrow, n, ncol;
    int a, b;
    int first, last, dedisperse;
};

__global__ __launch_bounds__(256) void k_tree_pass(TreeArgs A)
{
    extern __shared__ float rows[];
    __shared__ int sh[2048]; // G <= 1024: shifts of <= 512 pairs, or G row delays
    const int tid = threadIdx.x;
    const int c = blockIdx.y;
    const int G = 1 << (A.b - A.a);
    const int g = (int)(blockIdx.x >> A.a), rho = (int)(blockIdx.x & ((1u << A.a) - 1));
    const int ncol = A.ncol;
    const double dms = A.dms[c], ddm = A.ddm[c], dm0 = A.dm0[c];

    // ---- load the group's G rows
    if (A.first) {
        // Candidate::dedisperse: cutout channel q rotated left by its delay at dmc
        for (int t = tid; t < G; t += 256) {
            const int q = A.srcrow[g * G + t];
            int d = 0;
            if (q >= 0 && A.dedisperse) {
                d = (int)round(dmdelay(A.dmc[c], A.fmax, A.fsrc[q]) / A.tbin);
                d %= ncol;
                if (d < 0) d += ncol;
            }
            sh[t] = d;
        }
        __syncthreads();
        const float *cut = A.src + (size_t)c * A.nchan * ncol;
        const int wave = tid >> 6, lane = tid & 63;
        for (int t = wave; t < G; t += 4) {
            const int q = A.srcrow[g * G + t];
            float *dst = rows + (size_t)t * ncol;
            if (q < 0) {
                for (int i = lane; i < ncol; i += 64) dst[i] = 0.f;
                continue;
            }
            // the row rotated left by d: [d, ncol) then [0, d)
            const int d = sh[t];
            const uint32_t *srow = (const uint32_t *)(cut + (size_t)q * ncol);
            lds_fill(srow + d, (uint32_t *)dst, ncol - d, 0, 1, lane);
            if (d) lds_fill(srow, (uint32_t *)dst + (ncol - d), d, 0, 1, lane);
        }
    } else {
        const float *in = A.src + (size_t)c * A.nrow * ncol;
        const int wave = tid >> 6, lane = tid & 63;
        for (int t = wave; t < G; t += 4)
            lds_fill((const uint32_t *)(in + ((size_t)(g * G + t) * (1 << A.a) + rho) * ncol),
                     (uint32_t *)(rows + (size_t)t * ncol), ncol, 0, 1, lane);
    }
    lds_fill_done();

    // ---- levels a .. b-1, in place
    const int pairs = G / 2;
    const int P = max(1, 4096 / ncol); // pairs per chunk: <= 16 items per thread
    // item -> (pair, column) and pair -> (row pair, offset) by shifts where
    // the sizes are powers of two (half always is; ncol usually)
    const bool p2 = (ncol & (ncol - 1)) == 0;
    const int lc = __builtin_ctz((unsigned)ncol);
    for (int l = A.a; l < A.b; l++) {
        const int half = 1 << (l - A.a), hs = l - A.a;
        for (int q = tid; q < pairs; q += 256) {
            const int jl = q >> hs, t = q & (half - 1);
            const int jg = (g << (A.b - l - 1)) + jl;
            const int x0 = jg << (l + 1), x1 = x0 + (1 << l);
            const int k0 = rho + (t << A.a), k1 = k0 + (1 << l);
            sh[2 * q] = tree_shift(A.fP[x0], A.fP[x1], bitrev(k0, A.n), dms, ddm, dm0, A.tbin, ncol);
            sh[2 * q + 1] = tree_shift(A.fP[x0], A.fP[x1], bitrev(k1, A.n), dms, ddm, dm0, A.tbin, ncol);
        }
        __syncthreads();
        for (int q0 = 0; q0 < pairs; q0 += P) {
            const int items = min(P, pairs - q0) * ncol;
            float o0[kMaxItems], o1[kMaxItems];
#pragma unroll
            for (int u = 0; u < kMaxItems; u++) {
                const int idx = tid + 256 * u;
                if (idx < items) {
                    const int q = q0 + (p2 ? idx >> lc : idx / ncol), i = p2 ? idx & (ncol - 1) : idx % ncol;
                    const int jl = q >> hs, t = q & (half - 1);
                    const float *Ar = rows + (size_t)(2 * jl * half + t) * ncol;
                    const float *Br = Ar + (size_t)half * ncol;
                    int i0 = i + sh[2 * q], i1 = i + sh[2 * q + 1];
                    if (i0 >= ncol) i0 -= ncol;
                    if (i1 >= ncol) i1 -= ncol;
                    const float av = Ar[i];
                    o0[u] = av + Br[i0];
                    o1[u] = av + Br[i1];
                }
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < kMaxItems; u++) {
                const int idx = tid + 256 * u;
                if (idx < items) {
                    const int q = q0 + (p2 ? idx >> lc : idx / ncol), i = p2 ? idx & (ncol - 1) : idx % ncol;
                    const int jl = q >> hs, t = q & (half - 1);
                    float *Ar = rows + (size_t)(2 * jl * half + t) * ncol;
                    Ar[i] = o0[u];
                    Ar[(size_t)half * ncol + i] = o1[u];
                }
            }
            __syncthreads();
        }
    }

    // ---- store: node layout, or DM order after the last level
    float *out = A.dst + (size_t)c * A.nrow * ncol;
    for (int idx = tid; idx < G * ncol; idx += 256) {
        const int t = p2 ? idx >> lc : idx / ncol, i = idx - t * ncol;
        const int k = rho + (t << A.a);
        const int r = A.last ? bitrev(k, A.n) : g * (1 << A.b) + k;
        out[(size_t)r * ncol + i] = rows[idx];
    }
}

struct MfArgs {
    const float *dmt; // [rows][ncol]
    float *cs;        // [rows][2*ncol] circular running sums
    float *snr;       // [rows][ncol] or null
    int *wid;         // [rows][ncol] or null
    float *rowmax;
    int *rowarg, *rowwid;
    double *var;      // [rows] variance of each row (get_mean_var)
    const int *vwn;
    int nbox;
    long nrows;
    int ncol;
};

// Sequential parts of matched_filter: one 64-lane workgroup per 64 DM rows,
// one lane per row.
__global__ __launch_bounds__(64) void k_mf_scan(MfArgs M)
{
    extern __shared__ float smem[];
    float *stage = smem; // [64][65]
    const int lane = threadIdx.x;
    const long R0 = (long)blockIdx.x * 64;
    const long row = R0 + lane;
    const bool live = row < M.nrows;
    const int ncol = M.ncol;

    // get_mean_var: sequential FP64 sums (one lane per row)
    double m = 0., v = 0.;
    const int nr = (int)min((long)64, M.nrows - R0);
    for (int i0 = 0; i0 < ncol; i0 += 64) {
        for (int cc = 0; cc < nr; cc++)
            lds_fill((const uint32_t *)(M.dmt + (R0 + cc) * ncol + i0), (uint32_t *)(stage + cc * 65),
                     min(64, ncol - i0), 0, 1, lane);
        lds_fill_done();
        const int nn = min(64, ncol - i0);
        for (int ii = 0; ii < nn; ii++) {
            const double x = (double)stage[lane * 65 + ii];
            m += x;
            v = fma(x, x, v);
        }
        __syncthreads();
    }
    m /= ncol;
    v /= ncol;
    v = fma(-m, m, v);

    // csump over 2*ncol samples: (double)csum + (x - mean), rounded to float each step
    float csum = 0.f;
    for (int i0 = 0; i0 < 2 * ncol; i0 += 64) {
        const int e = i0 + lane;
        // samples e in [i0, i0 + 64) of the doubled row are x[e mod ncol]: up to two pieces
        const int end = min(i0 + 64, 2 * ncol);
        const int cut = (i0 < ncol && end > ncol) ? ncol : end;
        for (int cc = 0; cc < nr; cc++) {
            const float *xr = M.dmt + (R0 + cc) * ncol;
            const int s0 = i0 < ncol ? i0 : i0 - ncol;
            lds_fill((const uint32_t *)(xr + s0), (uint32_t *)(stage + cc * 65), cut - i0, 0, 1, lane);
            if (cut < end) lds_fill((const uint32_t *)xr, (uint32_t *)(stage + cc * 65 + (cut - i0)), end - cut, 0, 1, lane);
        }
        lds_fill_done();
        const int nn = min(64, 2 * ncol - i0);
        for (int ii = 0; ii < nn; ii++) {
            csum = (float)((double)csum + ((double)stage[lane * 65 + ii] - m));
            stage[lane * 65 + ii] = csum;
        }
        __syncthreads();
        for (int cc = 0; cc < 64; cc++) {
            const long r = R0 + cc;
            if (r < M.nrows && e < 2 * ncol) M.cs[r * 2 * ncol + e] = stage[cc * 65 + lane];
        }
        __syncthreads();
    }

    if (live) M.var[row] = v;
}

// Width loop of matched_filter: one wavefront per DM row (4 rows per
// workgroup), the row's running sums staged in LDS; U samples per lane.
template <int U>
__global__ __launch_bounds__(256) void k_mf_box(MfArgs M)
{
    extern __shared__ float smem[];
    const int ncol = M.ncol;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long r = (long)blockIdx.x * 4 + w;
    const bool live = r < M.nrows;
    float *crow = smem + (size_t)w * 2 * ncol;
    if (live) lds_fill((const uint32_t *)(M.cs + r * 2 * ncol), (uint32_t *)crow, 2 * ncol, 0, 1, lane);
    lds_fill_done();
    if (!live) return;
    const double var = M.var[r];
    float vS[U];
    int vW[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        vS[u] = 0.f;
        vW[u] = 0;
    }
    if (var > 0) {
        // each width's scale computed once per row, by lane k (FP64 sqrt and
        // divide, the same expression), and read back with v_readlane
        const bool lanes = M.nbox <= 64;
        const float tp = lanes && lane < M.nbox ? (float)sqrt(1. / (M.vwn[lane] * var)) : 0.f;
        for (int k = 0; k < M.nbox; k++) {
            const int wn = M.vwn[k], wl = wn / 2 + 1, wh = (wn - 1) / 2;
            const float temp = lanes ? __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tp), k))
                                     : (float)sqrt(1. / (wn * var));
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int o = lane + 64 * u;
                if (o < ncol) {
                    const int i = o >= wl ? o : o + ncol;
                    const float S = (crow[i + wh] - crow[i - wl]) * temp;
                    if (S >= vS[u]) {
                        vS[u] = S;
                        vW[u] = wn;
                    }
                }
            }
        }
    }
    float best = -INFINITY;
    int bo = 0x7fffffff, bw = 0;
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int o = lane + 64 * u;
        if (o < ncol) {
            if (M.snr) M.snr[r * ncol + o] = vS[u];
            if (M.wid) M.wid[r * ncol + o] = vW[u];
            if (vS[u] > best) {
                best = vS[u];
                bo = o;
                bw = vW[u];
            }
        }
    }
    // first maximum of the row: larger value, then smaller sample
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const float ob = __shfl_xor(best, off, 64);
        const int oo = __shfl_xor(bo, off, 64);
        const int ow = __shfl_xor(bw, off, 64);
        if (ob > best || (ob == best && oo < bo)) {
            best = ob;
            bo = oo;
            bw = ow;
        }
    }
    if (lane == 0) {
        M.rowmax[r] = best;
        M.rowarg[r] = bo;
        M.rowwid[r] = bw;
    }
}

__global__ __launch_bounds__(256) void k_argmax(const float *rowmax, const int *rowarg,
                                                const int *rowwid, int nrow, int *odm, int *ot,
                                                float *os, int *ow)
{
    __shared__ float sv[256];
    __shared__ int sj[256];
    const int c = blockIdx.x, tid = threadIdx.x;
    float best = -INFINITY;
    int bj = 0x7fffffff;
    for (int j = tid; j < nrow; j += 256) {
        const float v = rowmax[(size_t)c * nrow + j];
        if (v > best) {
            best = v;
            bj = j;
        }
    }
    sv[tid] = best;
    sj[tid] = bj;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) {
            const float ov = sv[tid + s];
            const int oj = sj[tid + s];
            if (ov > sv[tid] || (ov == sv[tid] && oj < sj[tid])) {
                sv[tid] = ov;
                sj[tid] = oj;
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        const int j = sj[0];
        odm[c] = j;
        ot[c] = rowarg[(size_t)c * nrow + j];
        os[c] = sv[0];
        ow[c] = rowwid[(size_t)c * nrow + j];
    }
}

} // namespace

// ------------------------------------------------------------------- host
struct tx_refiner {
    tx_refine_desc d{};
    int device = 0;
    int nrow = 0, n = 0, batch = 0;
    std::vector<double> freqs, fP;
    std::vector<int> srcrow, vwn;
    std::vector<std::pair<int, int>> passes;
    double fmax = 0, fmin = 0;
    hipStream_t stream = nullptr;
    double *d_fP = nullptr, *d_fsrc = nullptr, *d_par = nullptr; // par: dms, ddm, dm0, dmc [4][batch]
    int *d_srcrow = nullptr, *d_vwn = nullptr;
    float *d_t1 = nullptr, *d_t2 = nullptr, *d_cs = nullptr, *d_rowmax = nullptr, *d_res_s = nullptr;
    int *d_rowarg = nullptr, *d_rowwid = nullptr, *d_res_i = nullptr; // res_i: dm, t, w [3][batch]
    double *d_var = nullptr;
};

#define RCHK(x)                                                                                    \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) return set_error(TX_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

static void refiner_free(tx_refiner *r)
{
    hipFree(r->d_fP);
    hipFree(r->d_fsrc);
    hipFree(r->d_par);
    hipFree(r->d_srcrow);
    hipFree(r->d_vwn);
    hipFree(r->d_t1);
    hipFree(r->d_t2);
    hipFree(r->d_cs);
    hipFree(r->d_rowmax);
    hipFree(r->d_res_s);
    hipFree(r->d_rowarg);
    hipFree(r->d_rowwid);
    hipFree(r->d_res_i);
    hipFree(r->d_var);
    if (r->stream) hipStreamDestroy(r->stream);
}

// The ladder of Candidate::matched_filter (src/singlepulse/candidate.cpp:981-993),
// held in floats as the reference holds it.
static std::vector<int> mf_ladder(long nbin, float snrloss)
{
    const float wfactor = (float)(1. / ((1. - (double)snrloss) * (1. - (double)snrloss)));
    std::vector<float> v{1.f};
    for (;;) {
        const int w1 = (int)v.back();
        int w2 = (int)((float)w1 * wfactor);
        w2 = w2 < w1 + 1 ? w1 + 1 : w2;
        if (w2 > nbin / 2) break;
        v.push_back((float)w2);
    }
    std::vector<int> out;
    for (float x : v) out.push_back((int)x);
    return out;
}

extern "C" int tx_refine_create(const tx_refine_desc *d, int max_batch, tx_refiner **out)
{
    if (!d || !out || !d->frequencies) return set_error(TX_EINVAL, "null argument");
    if (d->nchan < 2 || d->nbin < 2 || d->nbin > 4096 || d->tbin <= 0. || max_batch <= 0 || max_batch > 65535)
        return set_error(TX_EINVAL, "refine: need nchan >= 2, 2 <= nbin <= 4096, tbin > 0, 1 <= max_batch <= 65535");
    if (d->snrloss < 0.f || d->snrloss >= 1.f) return set_error(TX_EINVAL, "refine: snrloss in [0, 1)");
    RCHK(hipSetDevice(d->device));
    tx_refiner *r = new tx_refiner();
    r->d = *d;
    r->device = d->device;
    r->batch = max_batch;
    r->freqs.assign(d->frequencies, d->frequencies + d->nchan);
    r->fmax = *std::max_element(r->freqs.begin(), r->freqs.end());
    r->fmin = *std::min_element(r->freqs.begin(), r->freqs.end());
    // Candidate::optimize: 2^ceil(log2 nchan) rows from the highest frequency,
    // padding rows at f[0] + j*foff (contracted to an FMA in the reference build)
    r->nrow = (int)std::pow(2, std::ceil(std::log2((double)d->nchan)));
    r->n = 0;
    while ((1 << r->n) < r->nrow) r->n++;
    const int nchan = d->nchan, nrow = r->nrow;
    const double foff = r->freqs[1] - r->freqs[0];
    r->fP.assign(nrow, 0.);
    r->srcrow.assign(nrow, -1);
    for (int j = 0; j < nchan; j++) {
        const int q = foff < 0. ? j : nchan - 1 - j;
        r->srcrow[j] = q;
        r->fP[j] = r->freqs[q];
    }
    for (int j = nchan; j < nrow; j++)
        r->fP[j] = foff < 0. ? std::fma((double)j, foff, r->freqs[0])
                             : std::fma((double)(nchan - 1 - j), foff, r->freqs[0]);
    r->vwn = mf_ladder(d->nbin, d->snrloss);
    if ((int)r->vwn.size() > kMaxMfBox) {
        delete r;
        return set_error(TX_EINVAL, "refine: ladder too long");
    }
    // tree passes of at most Lmax levels (G rows of nbin floats in 64 KB)
    int Lmax = 0;
    while ((2 << Lmax) * d->nbin <= kTreeLdsFloats && (1 << (Lmax + 1)) <= 1024) Lmax++;
    if (Lmax < 1) {
        delete r;
        return set_error(TX_EINVAL, "refine: nbin too large for one LDS tree tile");
    }
    const int npass = (r->n + Lmax - 1) / Lmax;
    for (int p = 0, a = 0; p < npass; p++) {
        const int len = (r->n - a + (npass - p) - 1) / (npass - p);
        r->passes.push_back({a, a + len});
        a += len;
    }
    (void)hipFuncSetAttribute((const void *)k_tree_pass, hipFuncAttributeMaxDynamicSharedMemorySize,
                              kTreeLdsFloats * 4);
    const size_t B = (size_t)max_batch, rowsz = (size_t)nrow * d->nbin;
    bool ok = hipStreamCreate(&r->stream) == hipSuccess &&
              hipMalloc(&r->d_fP, nrow * sizeof(double)) == hipSuccess &&
              hipMalloc(&r->d_fsrc, nchan * sizeof(double)) == hipSuccess &&
              hipMalloc(&r->d_par, 4 * B * sizeof(double)) == hipSuccess &&
              hipMalloc(&r->d_srcrow, nrow * sizeof(int)) == hipSuccess &&
              hipMalloc(&r->d_vwn, kMaxMfBox * sizeof(int)) == hipSuccess &&
              hipMalloc(&r->d_t1, B * rowsz * 4) == hipSuccess &&
              hipMalloc(&r->d_t2, B * rowsz * 4) == hipSuccess &&
              hipMalloc(&r->d_cs, 2 * B * rowsz * 4) == hipSuccess &&
              hipMalloc(&r->d_rowmax, B * nrow * 4) == hipSuccess &&
              hipMalloc(&r->d_rowarg, B * nrow * 4) == hipSuccess &&
              hipMalloc(&r->d_rowwid, B * nrow * 4) == hipSuccess &&
              hipMalloc(&r->d_var, B * nrow * 8) == hipSuccess &&
              hipMalloc(&r->d_res_s, B * 4) == hipSuccess &&
              hipMalloc(&r->d_res_i, 3 * B * 4) == hipSuccess;
    if (!ok) {
        refiner_free(r);
        delete r;
        return set_error(TX_ENOMEM, "refine: device allocation");
    }
    ok = hipMemcpy(r->d_fP, r->fP.data(), nrow * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(r->d_fsrc, r->freqs.data(), nchan * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(r->d_srcrow, r->srcrow.data(), nrow * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(r->d_vwn, r->vwn.data(), r->vwn.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        refiner_free(r);
        delete r;
        return set_error(TX_EHIP, "refine: table upload");
    }
    *out = r;
    return TX_OK;
}

extern "C" int tx_refine_destroy(tx_refiner *r)
{
    if (!r) return TX_OK;
    hipSetDevice(r->device);
    if (r->stream) hipStreamSynchronize(r->stream);
    refiner_free(r);
    delete r;
    return TX_OK;
}

extern "C" int tx_refine_info(const tx_refiner *r, int *nrow, int *nbox, int *npass)
{
    if (!r) return set_error(TX_EINVAL, "null argument");
    if (nrow) *nrow = r->nrow;
    if (nbox) *nbox = (int)r->vwn.size();
    if (npass) *npass = (int)r->passes.size();
    return TX_OK;
}

extern "C" int tx_refine_stream(tx_refiner *r, void **stream)
{
    if (!r || !stream) return set_error(TX_EINVAL, "null argument");
    *stream = (void *)r->stream;
    return TX_OK;
}

extern "C" int tx_refine_run(tx_refiner *r, int ncand, const float *cutouts, const double *dm,
                             tx_refine_result *res, float *dmtmap, float *snrmap, int *widmap)
{
    if (!r || ncand < 0 || (ncand && (!cutouts || !dm || !res))) return set_error(TX_EINVAL, "null argument");
    RCHK(hipSetDevice(r->device));
    const int nrow = r->nrow, nbin = r->d.nbin, nchan = r->d.nchan;
    const size_t rowsz = (size_t)nrow * nbin;
    // Candidate::optimize's DM grid (src/singlepulse/candidate.cpp:901-911)
    const double tmp = 4.148741601e3 * 1. * (1. / (r->fmin * r->fmin) - 1. / (r->fmax * r->fmax));
    const double dmrange = (nbin * r->d.tbin) / tmp;
    for (int c0 = 0; c0 < ncand; c0 += r->batch) {
        const int nc = std::min(r->batch, ncand - c0);
        std::vector<double> par(4 * (size_t)r->batch, 0.);
        for (int c = 0; c < nc; c++) {
            const double dmc = dm[c0 + c];
            const double dms = std::max(0., dmc - dmrange / 2.);
            const double dme = dmc + dmrange / 2.;
            par[c] = dms;
            par[r->batch + c] = (dme - dms) / nrow;
            par[2 * r->batch + c] = r->d.isdedispersed ? dmc : 0.;
            par[3 * r->batch + c] = dmc;
        }
        RCHK(hipMemcpyAsync(r->d_par, par.data(), par.size() * sizeof(double), hipMemcpyHostToDevice,
                            r->stream));
        TreeArgs A;
        A.fP = r->d_fP;
        A.srcrow = r->d_srcrow;
        A.fsrc = r->d_fsrc;
        A.dms = r->d_par;
        A.ddm = r->d_par + r->batch;
        A.dm0 = r->d_par + 2 * r->batch;
        A.dmc = r->d_par + 3 * r->batch;
        A.fmax = r->fmax;
        A.tbin = r->d.tbin;
        A.nchan = nchan;
        A.nrow = nrow;
        A.n = r->n;
        A.ncol = nbin;
        A.dedisperse = r->d.dedisperse != 0;
        float *dmt_out = dmtmap ? dmtmap + (size_t)c0 * rowsz : r->d_t2;
        const float *src = cutouts + (size_t)c0 * nchan * nbin;
        float *bufs[2] = {r->d_t1, r->d_t2};
        const int np = (int)r->passes.size();
        for (int p = 0; p < np; p++) {
            A.a = r->passes[p].first;
            A.b = r->passes[p].second;
            A.first = p == 0;
            A.last = p == np - 1;
            A.src = src;
            // the last pass never writes the buffer it reads
            float *dst = A.last ? dmt_out : bufs[p & 1];
            if (A.last && !dmtmap && src == r->d_t2) dst = r->d_t1, dmt_out = r->d_t1;
            A.dst = dst;
            const int G = 1 << (A.b - A.a);
            dim3 grid((unsigned)((nrow >> A.b) << A.a), (unsigned)nc);
            hipLaunchKernelGGL(k_tree_pass, grid, dim3(256), (size_t)G * nbin * 4, r->stream, A);
            RCHK(hipGetLastError());
            src = dst;
        }
        MfArgs M;
        M.dmt = dmt_out;
        M.cs = r->d_cs;
        M.snr = snrmap ? snrmap + (size_t)c0 * rowsz : nullptr;
        M.wid = widmap ? widmap + (size_t)c0 * rowsz : nullptr;
        M.rowmax = r->d_rowmax;
        M.rowarg = r->d_rowarg;
        M.rowwid = r->d_rowwid;
        M.var = r->d_var;
        M.vwn = r->d_vwn;
        M.nbox = (int)r->vwn.size();
        M.nrows = (long)nc * nrow;
        M.ncol = nbin;
        hipLaunchKernelGGL(k_mf_scan, dim3((unsigned)((M.nrows + 63) / 64)), dim3(64), 64 * 65 * 4,
                           r->stream, M);
        RCHK(hipGetLastError());
        const size_t shm = 4 * 2 * (size_t)nbin * 4;
        const dim3 mg((unsigned)((M.nrows + 3) / 4));
        if (nbin <= 512)
            hipLaunchKernelGGL(k_mf_box<8>, mg, dim3(256), shm, r->stream, M);
        else if (nbin <= 1024)
            hipLaunchKernelGGL(k_mf_box<16>, mg, dim3(256), shm, r->stream, M);
        else if (nbin <= 2048)
            hipLaunchKernelGGL(k_mf_box<32>, mg, dim3(256), shm, r->stream, M);
        else
            hipLaunchKernelGGL(k_mf_box<64>, mg, dim3(256), shm, r->stream, M);
        RCHK(hipGetLastError());
        hipLaunchKernelGGL(k_argmax, dim3((unsigned)nc), dim3(256), 0, r->stream, r->d_rowmax,
                           r->d_rowarg, r->d_rowwid, nrow, r->d_res_i, r->d_res_i + r->batch, r->d_res_s,
                           r->d_res_i + 2 * r->batch);
        RCHK(hipGetLastError());
        std::vector<int> ri(3 * (size_t)r->batch);
        std::vector<float> rs(r->batch);
        RCHK(hipMemcpyAsync(ri.data(), r->d_res_i, ri.size() * 4, hipMemcpyDeviceToHost, r->stream));
        RCHK(hipMemcpyAsync(rs.data(), r->d_res_s, rs.size() * 4, hipMemcpyDeviceToHost, r->stream));
        RCHK(hipStreamSynchronize(r->stream));
        for (int c = 0; c < nc; c++) {
            tx_refine_result &o = res[c0 + c];
            o.dms = par[c];
            o.ddm = par[r->batch + c];
            o.ndm = nrow;
            o.maxdmid = ri[c];
            o.maxtid = ri[r->batch + c];
            o.snr = rs[c];
            o.wn = ri[2 * r->batch + c];
            o.width = (float)(o.wn * r->d.tbin);
            o.dm_maxsnr = std::fma((double)o.maxdmid, o.ddm, o.dms); // dms + maxdmid*ddm, contracted
        }
    }
    return TX_OK;
}
