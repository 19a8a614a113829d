// diff.hip -- differentiable candidate S/N for gradient refinement.
//
// For a cutout x [nchan][nbin] with channel weights w, a DM, an arrival time
// t0 and a width W (samples):
//   tau_j   = K*DM*(1/f_j^2 - 1/fmax^2)/tbin          fractional delay (samples)
//   d_j     = floor(tau_j), a_j = tau_j - d_j
//   p(t)    = sum_j w_j*((1-a_j)*x_j[t+d_j] + a_j*x_j[t+d_j+1])   (indices mod nbin)
//   g(t)    = sig((delta+W/2)/s) - sig((delta-W/2)/s),  delta = t - t0 wrapped to [-nbin/2, nbin/2)
//   S       = sum_t gc(t)*p(t) / (|gc| * |w|),  gc = g - mean(g)
// S is the matched-filter S/N of the soft boxcar on unit-variance channels.
// The adjoint recomputes the forward (no shifted cutouts are stored) and
// returns dS/dDM, dS/dt0, dS/dW, dS/dw_j and, optionally, dS/dx (the
// interpolation's transpose, written as a per-element gather: deterministic,
// no atomics).  Channel sums run in a fixed order: channel j is handled by
// wave j mod 4, the four wave partials are added in wave order; every
// reduction is FP64 in a fixed butterfly order.
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
using txgpu::set_error;
using txgpu::wave_sum_d;

namespace {

constexpr double kDM = 4.148741601e3;

struct DiffArgs {
    const int *cand;      // [neval] cutout of each evaluation (null: e)
    const float *x;       // [ncut][nchan][nbin]
    const float *w;       // [ncut][nchan]
    const float *dm, *t0, *width; // [neval]
    const float *gsnr;    // [neval] upstream gradient (adjoint)
    const double *kf;     // [nchan] K*(1/f_j^2 - 1/fmax^2)/tbin: samples per unit DM
    float *snr;           // [neval]
    float *gdm, *gt0, *gwidth; // [neval]
    float *gw;            // [neval][nchan] or null
    float *gx;            // [neval][nchan][nbin] or null
    int nchan, nbin;
    float soft;
    int run; // k_diff_run: evaluation e = c*run + i belongs to cutout c
    float *gtt, *gdt; // k_diff_run: d2S/dt0^2 and d2S/dDM dt0 [neval] (upstream-scaled), or null
};

__device__ inline float sigm(float z) { return 1.f / (1.f + expf(-z)); }

// Dynamic LDS of k_diff / k_dgrid: 4 wave partials of p, and per channel its
// whole-sample delay and fraction.
inline size_t diff_lds(int nbin, int nchan) { return (size_t)16 * nbin + (size_t)8 * nchan; }
constexpr size_t kDiffMaxLds = 160 * 1024;

// Sum of a double over the 256-thread block in fixed order (waves, then wave 0..3).
__device__ inline double block_sum_d(double v, double *sh)
{
    v = wave_sum_d(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    return ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// One 256-thread workgroup per evaluation; U samples per lane (nbin <= 64*U).
template <int U, bool ADJ>
__global__ __launch_bounds__(256) void k_diff(DiffArgs A)
{
    extern __shared__ float smem[];
    float *pw = smem;                  // [4][nbin] wave partials of p, then p, then G
    __shared__ double shd[4];
    int *sd = (int *)(smem + 4 * A.nbin);   // [nchan] whole-sample delay mod nbin
    float *sa = (float *)(sd + A.nchan);    // [nchan] its fraction
    // logical order: consecutive evaluations (which share a cutout) on one
    // XCD, so the cutout is fetched into one L2 rather than eight
    const int e = (int)txgpu::xcd_logical(blockIdx.x, gridDim.x);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nbin = A.nbin, nchan = A.nchan;
    const int c = A.cand ? A.cand[e] : e;
    const float *X = A.x + (size_t)c * nchan * nbin;
    const float *W = A.w + (size_t)c * nchan;
    const double dm = (double)A.dm[e];
    // the channels' delays at this DM, once per workgroup, all lanes at once
    for (int j = tid; j < nchan; j += 256) {
        const double tau = dm * A.kf[j];
        const double fl = floor(tau);
        sa[j] = (float)(tau - fl);
        int d = (int)fmod(fl, (double)nbin);
        if (d < 0) d += nbin;
        sd[j] = d;
    }
    __syncthreads();

    // ---- forward: p(t), channel j on wave j % 4
    float acc[U];
#pragma unroll
    for (int u = 0; u < U; u++) acc[u] = 0.f;
    for (int j = wave; j < nchan; j += 4) {
        const float wj = W[j];
        if (wj == 0.f) continue;
        const float al = sa[j];
        const int d = sd[j];
        const float *row = X + (size_t)j * nbin;
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = lane + 64 * u;
            if (t < nbin) {
                int i0 = t + d;
                if (i0 >= nbin) i0 -= nbin;
                const int i1 = i0 + 1 < nbin ? i0 + 1 : 0;
                acc[u] += wj * ((1.f - al) * row[i0] + al * row[i1]);
            }
        }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int t = lane + 64 * u;
        if (t < nbin) pw[wave * nbin + t] = acc[u];
    }
    __syncthreads();
    // p(t) in wave order; the boxcar and its derivatives per sample
    const float t0 = A.t0[e], Wd = A.width[e], s = A.soft;
    double sg = 0., sw2 = 0.;
    for (int j = tid; j < nchan; j += 256) sw2 += (double)W[j] * (double)W[j];
    // each thread owns samples tid, tid+256, ...
    constexpr int V = (64 * U + 255) / 256;
    float p2[V], g2[V], dt2[V], dw2[V];
#pragma unroll
    for (int v = 0; v < V; v++) {
        const int t = tid + 256 * v;
        p2[v] = g2[v] = dt2[v] = dw2[v] = 0.f;
        if (t < nbin) {
            p2[v] = ((pw[t] + pw[nbin + t]) + pw[2 * nbin + t]) + pw[3 * nbin + t];
            float dl = (float)t - t0 + 0.5f * nbin;
            dl = dl - floorf(dl / nbin) * nbin - 0.5f * nbin;
            const float za = (dl + 0.5f * Wd) / s, zb = (dl - 0.5f * Wd) / s;
            const float sa = sigm(za), sb = sigm(zb);
            g2[v] = sa - sb;
            const float da = sa * (1.f - sa) / s, db = sb * (1.f - sb) / s;
            dt2[v] = -(da - db);          // dg/dt0
            dw2[v] = 0.5f * (da + db);    // dg/dW
            sg += (double)g2[v];
        }
    }
    sg = block_sum_d(sg, shd);
    sw2 = block_sum_d(sw2, shd);
    const double gm = sg / nbin;
    double n1 = 0., gg = 0., ndt = 0., ndw = 0., gdt = 0., gdw = 0., pm = 0.;
#pragma unroll
    for (int v = 0; v < V; v++) {
        const int t = tid + 256 * v;
        if (t < nbin) {
            const double gc = (double)g2[v] - gm;
            n1 += gc * (double)p2[v];
            gg += gc * gc;
            pm += (double)p2[v];
        }
    }
    n1 = block_sum_d(n1, shd);
    gg = block_sum_d(gg, shd);
    pm = block_sum_d(pm, shd) / nbin;
    const double D = sqrt(gg) * sqrt(sw2);
    const double S = D > 0. ? n1 / D : 0.;
    if (tid == 0) A.snr[e] = (float)S;
    if (!ADJ) return;

    // ---- adjoint
    const double gs = (double)A.gsnr[e];
    // dS/dt0, dS/dW: S = N/D, N = sum g*(p - pm), D = |gc|*|w|
#pragma unroll
    for (int v = 0; v < V; v++) {
        const int t = tid + 256 * v;
        if (t < nbin) {
            const double pc = (double)p2[v] - pm, gc = (double)g2[v] - gm;
            ndt += (double)dt2[v] * pc;
            ndw += (double)dw2[v] * pc;
            gdt += gc * (double)dt2[v];
            gdw += gc * (double)dw2[v];
        }
    }
    ndt = block_sum_d(ndt, shd);
    ndw = block_sum_d(ndw, shd);
    gdt = block_sum_d(gdt, shd);
    gdw = block_sum_d(gdw, shd);
    const double gn = sqrt(gg), wn = sqrt(sw2);
    if (tid == 0) {
        A.gt0[e] = D > 0. ? (float)(gs * (ndt - S * wn * gdt / gn) / D) : 0.f;
        A.gwidth[e] = D > 0. ? (float)(gs * (ndw - S * wn * gdw / gn) / D) : 0.f;
    }
    // G(t) = dS/dp(t) = gc(t)/D, into LDS
    __syncthreads();
#pragma unroll
    for (int v = 0; v < V; v++) {
        const int t = tid + 256 * v;
        if (t < nbin) pw[t] = D > 0. ? (float)(((double)g2[v] - gm) / D) : 0.f;
    }
    __syncthreads();
    float G[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int t = lane + 64 * u;
        G[u] = t < nbin ? pw[t] : 0.f;
    }
    // per channel: A_j = sum G*xs_j, B_j = sum G*(x1 - x0); dS/dw_j, dS/dDM.
    // Without dS/dw (and dS/dx) only dS/dDM = sum_j w_j kf_j B_j is needed:
    // each lane accumulates its own share over the channels and the wave
    // reduces once (fixed order: lanes' channel sums, then the butterfly).
    double gdm_w = 0.;
    if (!A.gw && !A.gx) {
        double lacc = 0.;
        for (int j = wave; j < nchan; j += 4) {
            const float wj = W[j];
            if (wj == 0.f) continue;
            const int d = sd[j];
            const float *row = X + (size_t)j * nbin;
            double b = 0.;
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int t = lane + 64 * u;
                if (t < nbin) {
                    int i0 = t + d;
                    if (i0 >= nbin) i0 -= nbin;
                    const int i1 = i0 + 1 < nbin ? i0 + 1 : 0;
                    b += (double)G[u] * (double)(row[i1] - row[i0]);
                }
            }
            lacc += (double)wj * A.kf[j] * b;
        }
        gdm_w = wave_sum_d(lacc);
    } else
    for (int j = wave; j < nchan; j += 4) {
        const float wj = W[j];
        const float al = sa[j];
        const int d = sd[j];
        const float *row = X + (size_t)j * nbin;
        double a = 0., b = 0.;
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = lane + 64 * u;
            if (t < nbin) {
                int i0 = t + d;
                if (i0 >= nbin) i0 -= nbin;
                const int i1 = i0 + 1 < nbin ? i0 + 1 : 0;
                const float x0 = row[i0], x1 = row[i1];
                a += (double)G[u] * (double)((1.f - al) * x0 + al * x1);
                b += (double)G[u] * (double)(x1 - x0);
            }
        }
        a = wave_sum_d(a);
        b = wave_sum_d(b);
        if (A.gw && lane == 0) A.gw[(size_t)e * nchan + j] = (float)(gs * (a - S * (double)wj / sw2));
        gdm_w += (double)wj * A.kf[j] * b; // identical on every lane
        if (A.gx) {
            // dS/dx_j[k] = w_j*((1-a_j)*G(k-d_j) + a_j*G(k-d_j-1))
            float *gxr = A.gx + ((size_t)e * nchan + j) * nbin;
            for (int k = lane; k < nbin; k += 64) {
                int t0i = k - d;
                if (t0i < 0) t0i += nbin;
                const int t1i = t0i - 1 >= 0 ? t0i - 1 : nbin - 1;
                gxr[k] = (float)gs * wj * ((1.f - al) * pw[t0i] + al * pw[t1i]);
            }
        }
    }
    __syncthreads();
    if (lane == 0) shd[wave] = gdm_w;
    __syncthreads();
    if (tid == 0) A.gdm[e] = (float)(gs * (((shd[0] + shd[1]) + shd[2]) + shd[3]));
}

// ---- runs of evaluations of one cutout (the refinement's pattern).  A
// workgroup takes up to 4*EW consecutive evaluations of one cutout, EW per
// wave.  The cutout's rows stream through LDS in chunks of cj rows, each row
// stored twice over (a shift by d < nbin then needs no wrap), once per pass
// for all the workgroup's evaluations: two chunk buffers, the next chunk's
// direct global->LDS loads in flight while the current one is summed.  The
// channels' weights and delay factors sit in LDS; a chunk's delays of
// (evaluation v, row jj) sit in lane v*cj + jj.  Each wave carries its
// evaluations through every channel in order (channel sums in channel order;
// FP64 reductions as wave butterflies).  Forward S and the adjoint without
// dS/dw and dS/dx: the same quantities as k_diff up to FP32 rounding order.
struct RunChunk {
    int dl;   // whole-sample delay mod nbin of (v, jj) = (lane / cj, lane % cj)
    float al; // its fraction
};

template <int U, int EW, bool SECOND>
__global__ __launch_bounds__(256) void k_diff_run(DiffArgs A, int cj)
{
    extern __shared__ float smem[];
    const int nbin = A.nbin, nchan = A.nchan;
    float *buf0 = smem, *buf1 = smem + (size_t)cj * 2 * nbin; // [2][cj][2*nbin]
    double *skf = (double *)(smem + (size_t)4 * cj * nbin);   // [nchan]
    float *sw = (float *)(skf + nchan);                        // [nchan]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int per = (A.run + 4 * EW - 1) / (4 * EW); // workgroups per cutout
    const int c = (int)blockIdx.x / per;
    const int i0 = ((int)blockIdx.x % per) * 4 * EW + wave * EW; // the wave's first evaluation in the run
    const int nv = __builtin_amdgcn_readfirstlane(max(0, min(EW, A.run - i0)));
    const long e0 = (long)c * A.run + i0;
    const float *X = A.x + (size_t)c * nchan * nbin;
    const float *W = A.w + (size_t)c * nchan;
    for (int j = tid; j < nchan; j += 256) {
        skf[j] = A.kf[j];
        sw[j] = W[j];
    }
    const int lv = lane / cj; // the evaluation whose delays this lane holds
    const double dmv = lv < nv ? (double)A.dm[e0 + lv] : 0.;
    const int nk = (nchan + cj - 1) / cj;

    auto fill = [&](int k) {
        const int j0 = k * cj, nj = min(cj, nchan - j0);
        float *b = (k & 1) ? buf1 : buf0;
        for (int jj = 0; jj < nj; jj++) {
            const uint32_t *src = (const uint32_t *)(X + (size_t)(j0 + jj) * nbin);
            uint32_t *d = (uint32_t *)(b + (size_t)jj * 2 * nbin);
            txgpu::lds_fill16(src, d, nbin, wave, 4, lane);
            txgpu::lds_fill16(src, d + nbin, nbin, wave, 4, lane);
        }
    };
    // delays of chunk k from the LDS copy of the delay factors (no global load)
    auto delays = [&](int k, RunChunk &ch) {
        const int j0 = k * cj, jj = lane - lv * cj;
        ch.dl = 0;
        ch.al = 0.f;
        if (lv < nv && jj < min(cj, nchan - j0)) {
            const double tau = dmv * skf[j0 + jj];
            const double fl = floor(tau);
            ch.al = (float)(tau - fl);
            int d = (int)fmod(fl, (double)nbin);
            if (d < 0) d += nbin;
            ch.dl = d;
        }
    };
    // one pass over the rows; body(jj, j, row) per live channel
    auto sweep = [&](auto body) {
        fill(0);
        __syncthreads(); // skf, sw written
        RunChunk cur, nxt;
        delays(0, cur);
        for (int k = 0; k < nk; k++) {
            txgpu::lds_fill_done(); // chunk k landed; every wave is done with chunk k-1's buffer
            if (k + 1 < nk) {
                fill(k + 1);
                delays(k + 1, nxt);
            }
            const int j0 = k * cj, nj = min(cj, nchan - j0);
            const float *b = (k & 1) ? buf1 : buf0;
            for (int jj = 0; jj < nj; jj++) {
                const float wj = sw[j0 + jj];
                if (wj == 0.f) continue;
                body(cur, jj, j0 + jj, wj, b + (size_t)jj * 2 * nbin + lane);
            }
            cur = nxt;
        }
        __syncthreads(); // every wave is done with the buffers
    };

    // ---- forward
    float acc[EW][U];
#pragma unroll
    for (int v = 0; v < EW; v++)
#pragma unroll
        for (int u = 0; u < U; u++) acc[v][u] = 0.f;
    sweep([&](const RunChunk &ch, int jj, int j, float wj, const float *rr) {
#pragma unroll
        for (int v = 0; v < EW; v++) {
            if (v >= nv) break;
            const int d = __builtin_amdgcn_readlane(ch.dl, v * cj + jj);
            const float a = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ch.al), v * cj + jj));
            const float *q = rr + d;
#pragma unroll
            for (int u = 0; u < U; u++)
                if (lane + 64 * u < nbin) acc[v][u] += wj * ((1.f - a) * q[64 * u] + a * q[64 * u + 1]);
        }
    });
    double sw2 = 0.;
    for (int j = lane; j < nchan; j += 64) sw2 += (double)sw[j] * (double)sw[j];
    sw2 = wave_sum_d(sw2);
    const double wn = sqrt(sw2);
    const float s = A.soft;
    float G[EW][U], Gt[SECOND ? EW : 1][SECOND ? U : 1]; // dS/dp(t) and its t0 derivative
#pragma unroll
    for (int v = 0; v < EW; v++) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            G[v][u] = 0.f;
            if (SECOND) Gt[v][u] = 0.f;
        }
        if (v >= nv) continue;
        const long e = e0 + v;
        const float t0 = A.t0[e], Wd = A.width[e];
        float g[U], gt[U], gw[U], gtt[U];
        double sg = 0.;
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = lane + 64 * u;
            g[u] = gt[u] = gw[u] = gtt[u] = 0.f;
            if (t < nbin) {
                float dl = (float)t - t0 + 0.5f * nbin;
                dl = dl - floorf(dl / nbin) * nbin - 0.5f * nbin;
                const float za = (dl + 0.5f * Wd) / s, zb = (dl - 0.5f * Wd) / s;
                const float sa = sigm(za), sb = sigm(zb);
                g[u] = sa - sb;
                const float da = sa * (1.f - sa) / s, db = sb * (1.f - sb) / s;
                gt[u] = -(da - db);
                gw[u] = 0.5f * (da + db);
                // d2g/dt0^2 = (sig''(za) - sig''(zb)) / s^2, sig'' = sig (1-sig) (1-2 sig)
                if (SECOND) gtt[u] = (da * (1.f - 2.f * sa) - db * (1.f - 2.f * sb)) / s;
                sg += (double)g[u];
            }
        }
        sg = wave_sum_d(sg);
        const double gm = sg / nbin;
        double n1 = 0., gg = 0., pm = 0.;
#pragma unroll
        for (int u = 0; u < U; u++)
            if (lane + 64 * u < nbin) {
                const double gc = (double)g[u] - gm;
                n1 += gc * (double)acc[v][u];
                gg += gc * gc;
                pm += (double)acc[v][u];
            }
        n1 = wave_sum_d(n1);
        gg = wave_sum_d(gg);
        pm = wave_sum_d(pm) / nbin;
        const double D = sqrt(gg) * wn;
        const double S = D > 0. ? n1 / D : 0.;
        if (lane == 0) A.snr[e] = (float)S;
        double ndt = 0., ndw = 0., gdt = 0., gdw = 0.;
        double ntt = 0., gtg = 0., gtm = 0., gcgtt = 0.; // second-derivative sums
#pragma unroll
        for (int u = 0; u < U; u++)
            if (lane + 64 * u < nbin) {
                const double pc = (double)acc[v][u] - pm, gc = (double)g[u] - gm;
                ndt += (double)gt[u] * pc;
                ndw += (double)gw[u] * pc;
                gdt += gc * (double)gt[u];
                gdw += gc * (double)gw[u];
                G[v][u] = D > 0. ? (float)(gc / D) : 0.f;
                if (SECOND) {
                    ntt += (double)gtt[u] * pc;
                    gtg += (double)gt[u] * (double)gt[u];
                    gtm += (double)gt[u];
                    gcgtt += gc * (double)gtt[u];
                }
            }
        ndt = wave_sum_d(ndt);
        ndw = wave_sum_d(ndw);
        gdt = wave_sum_d(gdt);
        gdw = wave_sum_d(gdw);
        const double gs = (double)A.gsnr[e], gn = sqrt(gg);
        const double S1 = D > 0. ? (ndt - S * wn * gdt / gn) / D : 0.; // dS/dt0
        if (lane == 0) {
            A.gt0[e] = (float)(gs * S1);
            A.gwidth[e] = D > 0. ? (float)(gs * (ndw - S * wn * gdw / gn) / D) : 0.f;
        }
        if (SECOND) {
            // S = N/D, N = sum g (p - pm), D = |gc| |w|, all in t0 through g:
            // D' = |w| A/|gc| with A = sum gc g'; D'' = |w| (B/|gc| - A^2/|gc|^3)
            // with B = sum g'^2 - n mean(g')^2 + sum gc g''; S'' = (N'' - 2 S' D' - S D'')/D;
            // dG/dt0 = (g' - mean g')/D - gc D'/D^2
            ntt = wave_sum_d(ntt);
            gtg = wave_sum_d(gtg);
            gtm = wave_sum_d(gtm) / nbin;
            gcgtt = wave_sum_d(gcgtt);
            double S2 = 0., D1 = 0.;
            if (D > 0.) {
                D1 = wn * gdt / gn;
                const double B = gtg - nbin * gtm * gtm + gcgtt;
                const double D2 = wn * (B / gn - gdt * gdt / (gn * gn * gn));
                S2 = (ntt - 2. * S1 * D1 - S * D2) / D;
            }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (lane + 64 * u < nbin && D > 0.)
                    Gt[v][u] = (float)(((double)gt[u] - gtm) / D - ((double)g[u] - gm) * D1 / (D * D));
            if (lane == 0) A.gtt[e] = (float)(gs * S2);
        }
    }

    // ---- adjoint: dS/dDM = sum_j w_j kf_j sum_t G(t) (x1 - x0), per lane in
    // FP64 over the channels, one butterfly per evaluation at the end
    double lacc[EW], lacc2[EW];
#pragma unroll
    for (int v = 0; v < EW; v++) lacc[v] = lacc2[v] = 0.;
    sweep([&](const RunChunk &ch, int jj, int j, float wj, const float *rr) {
        const double wk = (double)wj * skf[j];
#pragma unroll
        for (int v = 0; v < EW; v++) {
            if (v >= nv) break;
            const int d = __builtin_amdgcn_readlane(ch.dl, v * cj + jj);
            const float *q = rr + d;
            double b = 0., b2 = 0.;
#pragma unroll
            for (int u = 0; u < U; u++)
                if (lane + 64 * u < nbin) {
                    const double dx = (double)(q[64 * u + 1] - q[64 * u]);
                    b += (double)G[v][u] * dx;
                    if (SECOND) b2 += (double)Gt[v][u] * dx;
                }
            lacc[v] += wk * b;
            if (SECOND) lacc2[v] += wk * b2;
        }
    });
#pragma unroll
    for (int v = 0; v < EW; v++) {
        if (v >= nv) break;
        const double gd = wave_sum_d(lacc[v]);
        if (lane == 0) A.gdm[e0 + v] = (float)((double)A.gsnr[e0 + v] * gd);
        if (SECOND) {
            const double gd2 = wave_sum_d(lacc2[v]);
            if (lane == 0) A.gdt[e0 + v] = (float)((double)A.gsnr[e0 + v] * gd2);
        }
    }
}

// rows per chunk buffer: two buffers of rows stored twice over within 24 KB
inline int diff_run_rows(int nbin) { return std::max(1, std::min(8, 1536 / nbin)); }
inline size_t diff_run_lds(int nbin, int nchan)
{
    return (size_t)16 * diff_run_rows(nbin) * nbin + (size_t)12 * nchan;
}

// ---- grids of the same S/N: a DM trial's profile p(t) is shared by every
// (t0, W) pair, and t0 trials a whole number of samples apart share one
// boxcar template g shifted (S = (sum g(t-n) p(t) - mean(g) sum p) / (|gc| |w|)).
struct GridArgs {
    const float *x, *w;       // [ncut][nchan][nbin], [ncut][nchan]
    const float *dm0, *t00;   // [ncut] first DM trial, first t0 trial
    const float *widths;      // [nwidth]
    const double *kf;         // [nchan]
    float *best;              // [ncut][ndm] highest S over (t0, W)
    int *best_idx;            // [ncut][ndm] its it0*nwidth + iw (first maximum)
    int nchan, nbin, ndm, nt0, nwidth;
    int q;                    // t0 phases: trial i = n*q + r sits r*dt + n*istep after t00
    int istep;                // whole-sample step between phases' shifts (1 when q > 1)
    float ddm, dt, soft;
};

// One 256-thread workgroup per (cutout, DM trial), logical order DM fastest
// (the trials of a cutout share its rows in one L2).  p(t) is formed exactly
// as k_diff forms it; each wave then takes templates (phase r, width k) in
// turn and correlates them with p at every shift of that phase.
template <int U>
__global__ __launch_bounds__(256) void k_dgrid(GridArgs A)
{
    extern __shared__ float smem[];
    float *pw = smem; // [4][nbin] wave partials, then p in [0, nbin)
    __shared__ double shd[4];
    __shared__ float sbS[4];
    __shared__ int sbI[4];
    int *sd = (int *)(smem + 4 * A.nbin);   // [nchan] whole-sample delay mod nbin
    float *sa = (float *)(sd + A.nchan);    // [nchan] its fraction
    const unsigned lid = txgpu::xcd_logical(blockIdx.x, gridDim.x);
    const int c = (int)(lid / A.ndm), m = (int)(lid % A.ndm);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nbin = A.nbin, nchan = A.nchan;
    const float *X = A.x + (size_t)c * nchan * nbin;
    const float *W = A.w + (size_t)c * nchan;
    const double dm = (double)(A.dm0[c] + (float)m * A.ddm);
    for (int j = tid; j < nchan; j += 256) {
        const double tau = dm * A.kf[j];
        const double fl = floor(tau);
        sa[j] = (float)(tau - fl);
        int d = (int)fmod(fl, (double)nbin);
        if (d < 0) d += nbin;
        sd[j] = d;
    }
    __syncthreads();

    float acc[U];
#pragma unroll
    for (int u = 0; u < U; u++) acc[u] = 0.f;
    for (int j = wave; j < nchan; j += 4) {
        const float wj = W[j];
        if (wj == 0.f) continue;
        const float al = sa[j];
        const int d = sd[j];
        const float *row = X + (size_t)j * nbin;
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = lane + 64 * u;
            if (t < nbin) {
                int i0 = t + d;
                if (i0 >= nbin) i0 -= nbin;
                const int i1 = i0 + 1 < nbin ? i0 + 1 : 0;
                acc[u] += wj * ((1.f - al) * row[i0] + al * row[i1]);
            }
        }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int t = lane + 64 * u;
        if (t < nbin) pw[wave * nbin + t] = acc[u];
    }
    __syncthreads();
    constexpr int V = (64 * U + 255) / 256;
    float p2[V];
    double P = 0., sw2 = 0.;
#pragma unroll
    for (int v = 0; v < V; v++) {
        const int t = tid + 256 * v;
        p2[v] = t < nbin ? ((pw[t] + pw[nbin + t]) + pw[2 * nbin + t]) + pw[3 * nbin + t] : 0.f;
        P += (double)p2[v];
    }
    for (int j = tid; j < nchan; j += 256) sw2 += (double)W[j] * (double)W[j];
    P = block_sum_d(P, shd);
    sw2 = block_sum_d(sw2, shd);
    __syncthreads();
#pragma unroll
    for (int v = 0; v < V; v++) {
        const int t = tid + 256 * v;
        if (t < nbin) pw[t] = p2[v];
    }
    __syncthreads();
    const double wn = sqrt(sw2);
    const float s = A.soft;
    float bS = -INFINITY;
    int bI = 0x7fffffff;
    const int ntpl = A.q * A.nwidth;
    for (int tpl = wave; tpl < ntpl; tpl += 4) {
        const int r = tpl / A.nwidth, k = tpl % A.nwidth;
        if (r >= A.nt0) continue;
        const float t0 = A.t00[c] + (float)r * A.dt, Wd = A.widths[k];
        float g[U];
        double sg = 0.;
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = lane + 64 * u;
            g[u] = 0.f;
            if (t < nbin) {
                float dl = (float)t - t0 + 0.5f * nbin;
                dl = dl - floorf(dl / nbin) * nbin - 0.5f * nbin;
                g[u] = sigm((dl + 0.5f * Wd) / s) - sigm((dl - 0.5f * Wd) / s);
                sg += (double)g[u];
            }
        }
        sg = wave_sum_d(sg);
        const double gm = sg / nbin;
        double gg = 0.;
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = lane + 64 * u;
            if (t < nbin) gg += ((double)g[u] - gm) * ((double)g[u] - gm);
        }
        gg = wave_sum_d(gg);
        const double D = sqrt(gg) * wn;
        for (int i = r, n = 0; i < A.nt0; i += A.q, n += A.istep) {
            float cs = 0.f;
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int t = lane + 64 * u;
                if (t < nbin) {
                    int ip = t + n;
                    if (ip >= nbin) ip -= nbin;
                    cs += g[u] * pw[ip];
                }
            }
            const double corr = wave_sum_d((double)cs);
            const float S = D > 0. ? (float)((corr - gm * P) / D) : 0.f;
            const int idx = i * A.nwidth + k;
            if (S > bS || (S == bS && idx < bI)) {
                bS = S;
                bI = idx;
            }
        }
    }
    if (lane == 0) {
        sbS[wave] = bS;
        sbI[wave] = bI;
    }
    __syncthreads();
    if (tid == 0) {
        float S = sbS[0];
        int I = sbI[0];
        for (int q = 1; q < 4; q++)
            if (sbS[q] > S || (sbS[q] == S && sbI[q] < I)) {
                S = sbS[q];
                I = sbI[q];
            }
        A.best[(size_t)c * A.ndm + m] = S;
        A.best_idx[(size_t)c * A.ndm + m] = I;
    }
}

template <bool ADJ>
int launch_diff(hipStream_t st, const DiffArgs &A, int neval)
{
    const size_t shm = diff_lds(A.nbin, A.nchan);
    const dim3 g((unsigned)neval), b(256);
    auto go = [&](const void *f) {
        (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        void *p[] = {(void *)&A};
        (void)hipLaunchKernel(f, g, b, p, shm, st);
    };
    if (A.nbin <= 512)
        go((const void *)k_diff<8, ADJ>);
    else if (A.nbin <= 1024)
        go((const void *)k_diff<16, ADJ>);
    else if (A.nbin <= 2048)
        go((const void *)k_diff<32, ADJ>);
    else
        go((const void *)k_diff<64, ADJ>);
    return 0;
}

int launch_diff_run(hipStream_t st, const DiffArgs &A, int ncut)
{
    constexpr int EW = 2;
    const int cj = diff_run_rows(A.nbin);
    const size_t shm = diff_run_lds(A.nbin, A.nchan);
    const long per = (A.run + 4 * EW - 1) / (4 * EW);
    const dim3 g((unsigned)(ncut * per)), b(256);
    auto go = [&](const void *f) {
        (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        void *p[] = {(void *)&A, (void *)&cj};
        (void)hipLaunchKernel(f, g, b, p, shm, st);
    };
    const bool second = A.gtt && A.gdt;
    if (A.nbin <= 512)
        go(second ? (const void *)k_diff_run<8, EW, true> : (const void *)k_diff_run<8, EW, false>);
    else
        go(second ? (const void *)k_diff_run<16, EW, true> : (const void *)k_diff_run<16, EW, false>);
    return 0;
}

} // namespace

struct tx_diff {
    tx_diff_desc d{};
    int device = 0;
    hipStream_t stream = nullptr;
    double *d_kf = nullptr;
};

extern "C" int tx_diff_create(const tx_diff_desc *d, tx_diff **out)
{
    if (!d || !out || !d->frequencies) return set_error(TX_EINVAL, "null argument");
    if (d->nchan < 1 || d->nbin < 2 || d->nbin > 4096 || d->tbin <= 0. || !(d->softness > 0.f))
        return set_error(TX_EINVAL, "diff: need nchan >= 1, 2 <= nbin <= 4096, tbin > 0, softness > 0");
    if (diff_lds(d->nbin, d->nchan) > kDiffMaxLds)
        return set_error(TX_EINVAL, "diff: 16*nbin + 8*nchan bytes must fit in 160 KB of LDS");
    if (hipSetDevice(d->device) != hipSuccess) return set_error(TX_EHIP, "hipSetDevice");
    tx_diff *r = new tx_diff();
    r->d = *d;
    r->device = d->device;
    const double fmax = *std::max_element(d->frequencies, d->frequencies + d->nchan);
    std::vector<double> kf(d->nchan);
    for (int j = 0; j < d->nchan; j++) {
        const double f = d->frequencies[j];
        kf[j] = kDM * (1. / (f * f) - 1. / (fmax * fmax)) / d->tbin;
    }
    if (hipStreamCreate(&r->stream) != hipSuccess || hipMalloc(&r->d_kf, kf.size() * 8) != hipSuccess ||
        hipMemcpy(r->d_kf, kf.data(), kf.size() * 8, hipMemcpyHostToDevice) != hipSuccess) {
        if (r->stream) hipStreamDestroy(r->stream);
        hipFree(r->d_kf);
        delete r;
        return set_error(TX_ENOMEM, "diff: device setup");
    }
    *out = r;
    return TX_OK;
}

extern "C" int tx_diff_destroy(tx_diff *r)
{
    if (!r) return TX_OK;
    hipSetDevice(r->device);
    hipStreamSynchronize(r->stream);
    hipFree(r->d_kf);
    hipStreamDestroy(r->stream);
    delete r;
    return TX_OK;
}

extern "C" int tx_diff_stream(tx_diff *r, void **stream)
{
    if (!r || !stream) return set_error(TX_EINVAL, "null argument");
    *stream = (void *)r->stream;
    return TX_OK;
}

static int diff_run(tx_diff *r, bool adj, int neval, const int *cand, const float *x, const float *w,
                    const float *dm, const float *t0, const float *width, const float *gsnr, float *snr,
                    float *gdm, float *gt0, float *gwidth, float *gw, float *gx)
{
    if (!r || neval < 0 || (neval && (!x || !w || !dm || !t0 || !width || !snr)))
        return set_error(TX_EINVAL, "null argument");
    if (adj && neval && (!gsnr || !gdm || !gt0 || !gwidth)) return set_error(TX_EINVAL, "null gradient output");
    if (neval == 0) return TX_OK;
    if (hipSetDevice(r->device) != hipSuccess) return set_error(TX_EHIP, "hipSetDevice");
    DiffArgs A;
    A.cand = cand;
    A.x = x;
    A.w = w;
    A.dm = dm;
    A.t0 = t0;
    A.width = width;
    A.gsnr = gsnr;
    A.kf = r->d_kf;
    A.snr = snr;
    A.gdm = gdm;
    A.gt0 = gt0;
    A.gwidth = gwidth;
    A.gw = gw;
    A.gx = gx;
    A.nchan = r->d.nchan;
    A.nbin = r->d.nbin;
    A.soft = r->d.softness;
    A.run = 0;
    A.gtt = A.gdt = nullptr;
    if (adj)
        launch_diff<true>(r->stream, A, neval);
    else
        launch_diff<false>(r->stream, A, neval);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(TX_EHIP, std::string("diff launch: ") + hipGetErrorString(e));
    e = hipStreamSynchronize(r->stream);
    if (e != hipSuccess) return set_error(TX_EHIP, std::string("diff: ") + hipGetErrorString(e));
    return TX_OK;
}

extern "C" int tx_refine_fwd(tx_diff *r, int neval, const int *cand, const float *x, const float *w,
                             const float *dm, const float *t0, const float *width, float *snr)
{
    return diff_run(r, false, neval, cand, x, w, dm, t0, width, nullptr, snr, nullptr, nullptr, nullptr,
                    nullptr, nullptr);
}

extern "C" int tx_refine_adj(tx_diff *r, int neval, const int *cand, const float *x, const float *w,
                             const float *dm, const float *t0, const float *width, const float *gsnr,
                             float *snr, float *gdm, float *gt0, float *gwidth, float *gw, float *gx)
{
    return diff_run(r, true, neval, cand, x, w, dm, t0, width, gsnr, snr, gdm, gt0, gwidth, gw, gx);
}

extern "C" int tx_refine_adj_runs(tx_diff *r, int ncut, int run, const float *x, const float *w,
                                  const float *dm, const float *t0, const float *width, const float *gsnr,
                                  float *snr, float *gdm, float *gt0, float *gwidth, float *gtt, float *gdt)
{
    if (!r || ncut < 0 || run < 1) return set_error(TX_EINVAL, "bad argument");
    if (ncut && (!x || !w || !dm || !t0 || !width || !gsnr || !snr || !gdm || !gt0 || !gwidth))
        return set_error(TX_EINVAL, "null argument");
    if (r->d.nbin > 1024 || r->d.nbin % 4) return set_error(TX_EINVAL, "runs: need nbin <= 1024, a multiple of 4");
    if ((uintptr_t)x % 16) return set_error(TX_EINVAL, "runs: x must be 16-byte aligned");
    if (diff_run_lds(r->d.nbin, r->d.nchan) > 64 * 1024)
        return set_error(TX_EINVAL, "runs: 16*nbin*rows + 12*nchan bytes of LDS exceed 64 KB");
    if ((long)ncut * run > 0x7fffffffL) return set_error(TX_EINVAL, "runs: too many evaluations");
    if (ncut == 0) return TX_OK;
    if (hipSetDevice(r->device) != hipSuccess) return set_error(TX_EHIP, "hipSetDevice");
    DiffArgs A{};
    A.cand = nullptr;
    A.x = x;
    A.w = w;
    A.dm = dm;
    A.t0 = t0;
    A.width = width;
    A.gsnr = gsnr;
    A.kf = r->d_kf;
    A.snr = snr;
    A.gdm = gdm;
    A.gt0 = gt0;
    A.gwidth = gwidth;
    A.gw = nullptr;
    A.gx = nullptr;
    A.nchan = r->d.nchan;
    A.nbin = r->d.nbin;
    A.soft = r->d.softness;
    A.run = run;
    if ((gtt == nullptr) != (gdt == nullptr)) return set_error(TX_EINVAL, "runs: gtt and gdt go together");
    A.gtt = gtt;
    A.gdt = gdt;
    launch_diff_run(r->stream, A, ncut);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(TX_EHIP, std::string("diff runs launch: ") + hipGetErrorString(e));
    e = hipStreamSynchronize(r->stream);
    if (e != hipSuccess) return set_error(TX_EHIP, std::string("diff runs: ") + hipGetErrorString(e));
    return TX_OK;
}

extern "C" int tx_refine_grid(tx_diff *r, int ncut, const float *x, const float *w, const float *dm0,
                              const float *t00, int ndm, float ddm, int nt0, float dt, int nwidth,
                              const float *widths, float *best, int *best_idx)
{
    if (!r || ncut < 0 || (ncut && (!x || !w || !dm0 || !t00 || !widths || !best || !best_idx)))
        return set_error(TX_EINVAL, "null argument");
    if (ndm < 1 || nt0 < 1 || nwidth < 1 || nwidth > 4096 || !(dt > 0.f))
        return set_error(TX_EINVAL, "grid: need ndm, nt0, nwidth >= 1 and dt > 0");
    int q = 1, istep = 1;
    const double rd = std::round((double)dt), rq = std::round(1. / (double)dt);
    if (rd >= 1. && std::fabs((double)dt - rd) < 1e-6) {
        istep = (int)rd;
    } else if (rq >= 2. && rq <= 64. && std::fabs(1. / (double)dt - rq) < 1e-4) {
        q = (int)rq;
    } else {
        return set_error(TX_EINVAL, "grid: the t0 step must be a whole number of samples or 1/q, q <= 64");
    }
    if ((double)(nt0 - 1) * dt >= (double)r->d.nbin)
        return set_error(TX_EINVAL, "grid: the t0 trials must span less than one cutout");
    if ((long)ncut * ndm > 0x7fffffffL) return set_error(TX_EINVAL, "grid: too many (cutout, DM) pairs");
    if (ncut == 0) return TX_OK;
    if (hipSetDevice(r->device) != hipSuccess) return set_error(TX_EHIP, "hipSetDevice");
    GridArgs A;
    A.x = x;
    A.w = w;
    A.dm0 = dm0;
    A.t00 = t00;
    A.widths = widths;
    A.kf = r->d_kf;
    A.best = best;
    A.best_idx = best_idx;
    A.nchan = r->d.nchan;
    A.nbin = r->d.nbin;
    A.ndm = ndm;
    A.nt0 = nt0;
    A.nwidth = nwidth;
    A.q = q;
    A.istep = istep;
    A.ddm = ddm;
    A.dt = dt;
    A.soft = r->d.softness;
    const size_t shm = diff_lds(A.nbin, A.nchan);
    const dim3 g((unsigned)((long)ncut * ndm)), b(256);
    auto go = [&](const void *f) {
        (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        void *p[] = {(void *)&A};
        (void)hipLaunchKernel(f, g, b, p, shm, r->stream);
    };
    if (A.nbin <= 512)
        go((const void *)k_dgrid<8>);
    else if (A.nbin <= 1024)
        go((const void *)k_dgrid<16>);
    else if (A.nbin <= 2048)
        go((const void *)k_dgrid<32>);
    else
        go((const void *)k_dgrid<64>);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(TX_EHIP, std::string("grid launch: ") + hipGetErrorString(e));
    e = hipStreamSynchronize(r->stream);
    if (e != hipSuccess) return set_error(TX_EHIP, std::string("grid: ") + hipGetErrorString(e));
    return TX_OK;
}
