// rfi.hip -- RFI cleaning of channel x time frames on the device:
// Candidate::zdot (zero-DM matched filter) and Candidate::kadaneF
// (src/singlepulse/candidate.cpp:710-897), the formulas behind the front
// end's "zdot" and "kadaneF" items (src/singlepulse/singlepulse.cpp:305-317).
//
// A frame is nchan x nsamp floats; element (channel j, sample i) of frame f
// is x[f*fstride + j*cstride + i*tstride], so one kernel serves the
// Candidate layout [nchan][nbin] (cstride = nsamp, tstride = 1) and a
// DataBuffer block [nsamp][nchans] (cstride = 1, tstride = nchans).
//
// The arithmetic is the reference's, operation by operation: FP64
// accumulations in its sequential order (one thread per channel or per
// sample), float products where it multiplies floats, float stores where it
// assigns to float, and fused multiply-adds exactly where its build
// (g++ -O3 -mavx2 -mfma, config/ax_check_x86_features.m4:76) contracts:
// a*b + c and c - a*b fuse; in a*b - c*d the second product is rounded
// first and the first is fused; a float product added to a double is not
// fused.  Nothing else is contracted (fp contract off).  Results are
// bit-identical to the oracle's restatement.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "../../include/txgpu.h"

#pragma clang fp contract(off)

namespace txgpu {
int set_error(int code, const std::string &msg);
}
using txgpu::set_error;

namespace {

struct Frames {
    float *x;
    long fstride, cstride, tstride;
    int nframes, nchan;
    long nsamp;
    __device__ float &at(long f, long j, long i) const { return x[f * fstride + j * cstride + i * tstride]; }
};

// zdot 1: per (frame, channel) mean and standard deviation over the samples
// (candidate.cpp:837-851), then the channel normalised in place (:853-859)
__global__ void k_zdot_norm(Frames F)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)F.nframes * F.nchan) return;
    const long f = g / F.nchan, j = g % F.nchan;
    double m = 0., sd = 0.;
    for (long i = 0; i < F.nsamp; i++) {
        const float v = F.at(f, j, i);
        m += v;
        sd += v * v; // a float product, as data[]*data[] is
    }
    m /= F.nsamp;
    sd /= F.nsamp;
    sd = fma(-m, m, sd); // tmpstd -= tmpmean*tmpmean
    sd = sqrt(sd);
    if (sd == 0.) sd = 1.;
    for (long i = 0; i < F.nsamp; i++) {
        float v = F.at(f, j, i);
        v = (float)((double)v - m);
        v = (float)((double)v / sd);
        F.at(f, j, i) = v;
    }
}

// zdot 2: per (frame, sample) the channel mean of the normalised data,
// summed in channel order (candidate.cpp:858, 862-865)
__global__ void k_zdot_s(Frames F, double *s)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)F.nframes * F.nsamp) return;
    const long f = g / F.nsamp, i = g % F.nsamp;
    double a = 0.;
    long j = 0;
    for (; j + 16 <= F.nchan; j += 16) { // 16 loads in flight, summed in channel order
        float v[16];
#pragma unroll
        for (int u = 0; u < 16; u++) v[u] = F.at(f, j + u, i);
#pragma unroll
        for (int u = 0; u < 16; u++) a += v[u];
    }
    for (; j < F.nchan; j++) a += F.at(f, j, i);
    s[g] = a / F.nchan;
}

// zdot 3: per frame sum s and sum s^2 (the same for every channel,
// candidate.cpp:879-880)
__global__ void k_zdot_ss(Frames F, const double *s, double *sess)
{
    const long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F.nframes) return;
    double se = 0., ss = 0.;
    const double *sf = s + f * F.nsamp;
    long i = 0;
    for (; i + 32 <= F.nsamp; i += 32) { // 32 loads in flight (one thread per frame)
        double v[32];
#pragma unroll
        for (int u = 0; u < 32; u++) v[u] = sf[i + u];
#pragma unroll
        for (int u = 0; u < 32; u++) {
            se += v[u];
            ss = fma(v[u], v[u], ss);
        }
    }
    for (; i < F.nsamp; i++) {
        se += sf[i];
        ss = fma(sf[i], sf[i], ss);
    }
    sess[2 * f] = se;
    sess[2 * f + 1] = ss;
}

// zdot 4: per (frame, channel) the least-squares fit alpha*s + beta of the
// channel, subtracted (candidate.cpp:870-896)
__global__ void k_zdot_fit(Frames F, const double *s, const double *sess)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)F.nframes * F.nchan) return;
    const long f = g / F.nchan, j = g % F.nchan;
    const double *sf = s + f * F.nsamp;
    double xe = 0., xs = 0.;
    for (long i = 0; i < F.nsamp; i++) {
        const float v = F.at(f, j, i);
        xe += v;
        xs = fma((double)v, sf[i], xs);
    }
    const double se = sess[2 * f], ss = sess[2 * f + 1], nb = (double)F.nsamp;
    const double tmp = fma(se, se, -(ss * nb));
    double alpha = 0., beta = 0.;
    if (tmp != 0) {
        alpha = fma(xe, se, -(xs * nb)) / tmp;
        beta = fma(xs, se, -(xe * ss)) / tmp;
    }
    for (long i = 0; i < F.nsamp; i++) {
        const double fit = fma(alpha, sf[i], beta);
        F.at(f, j, i) = (float)((double)F.at(f, j, i) - fit);
    }
}

// ---- the same passes with memory access shaped for the two common layouts;
// every channel's sums still run sample by sample in order, so results are
// those of the kernels above.

// Channel-major frames (tstride 1): a wave takes 64 channels; 64-sample
// tiles of their rows pass through LDS (coalesced row loads, a padded
// [64][65] tile read one row per lane without bank conflicts).
constexpr int kTile = 64;

template <bool FIT>
__global__ __launch_bounds__(256) void k_zdot_rows(Frames F, const double *s, const double *sess)
{
    __shared__ float tile[4][kTile][kTile + 1];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long g0 = ((long)blockIdx.x * 4 + w) * 64; // first (frame, channel) of the wave
    const long ntot = (long)F.nframes * F.nchan;
    const long g = g0 + lane;
    const bool live = g < ntot;
    const long f = live ? g / F.nchan : 0, j = live ? g % F.nchan : 0;
    float (*T)[kTile + 1] = tile[w];
    // rows of the wave's channels: row c is (frame, channel) g0 + c
    auto rowp = [&](int c) -> float * {
        const long gc = g0 + c;
        return F.x + (gc / F.nchan) * F.fstride + (gc % F.nchan) * F.cstride;
    };
    const int nrow = (int)min((long)64, ntot - g0);
    // 16 row loads in flight at a time (addresses clamped into the frame,
    // only rows and samples that exist are stored)
    auto load = [&](long t0) {
        const int n = (int)min((long)kTile, F.nsamp - t0);
        const long ti = t0 + min(lane, n - 1);
        for (int c0 = 0; c0 < nrow; c0 += 16) {
            float v[16];
#pragma unroll
            for (int u = 0; u < 16; u++) v[u] = rowp(min(c0 + u, nrow - 1))[ti];
#pragma unroll
            for (int u = 0; u < 16; u++)
                if (c0 + u < nrow && lane < n) T[c0 + u][lane] = v[u];
        }
        __builtin_amdgcn_wave_barrier();
        return n;
    };
    auto store = [&](long t0, int n) {
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        for (int c = 0; c < nrow; c++) {
            float *r = rowp(c) + t0;
            if (lane < n) r[lane] = T[c][lane];
        }
        __builtin_amdgcn_wave_barrier();
    };
    const double *sf = s + f * F.nsamp; // read only when FIT
    double a = 0., b = 0.;
    for (long t0 = 0; t0 < F.nsamp; t0 += kTile) {
        const int n = load(t0);
        if (live) {
            int i = 0;
            for (; FIT && i + 16 <= n; i += 16) {
                double sv[16]; // the channel mean's values, loaded ahead
#pragma unroll
                for (int u = 0; u < 16; u++) sv[u] = sf[t0 + i + u];
#pragma unroll
                for (int u = 0; u < 16; u++) {
                    const float v = T[lane][i + u];
                    a += v;
                    b = fma((double)v, sv[u], b);
                }
            }
            for (; i < n; i++) {
                const float v = T[lane][i];
                a += v;
                if (FIT) b = fma((double)v, sf[t0 + i], b);
                else b += v * v;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    double m = 0., sd = 1., alpha = 0., beta = 0.;
    if (!FIT) {
        m = a / F.nsamp;
        sd = b / F.nsamp;
        sd = fma(-m, m, sd);
        sd = sqrt(sd);
        if (sd == 0.) sd = 1.;
    } else if (live) {
        const double se = sess[2 * f], ss = sess[2 * f + 1], nb = (double)F.nsamp;
        const double tmp = fma(se, se, -(ss * nb));
        if (tmp != 0) {
            alpha = fma(a, se, -(b * nb)) / tmp;
            beta = fma(b, se, -(a * ss)) / tmp;
        }
    }
    for (long t0 = 0; t0 < F.nsamp; t0 += kTile) {
        const int n = load(t0);
        if (live) {
            int i = 0;
            for (; FIT && i + 16 <= n; i += 16) {
                double sv[16];
#pragma unroll
                for (int u = 0; u < 16; u++) sv[u] = sf[t0 + i + u];
#pragma unroll
                for (int u = 0; u < 16; u++)
                    T[lane][i + u] = (float)((double)T[lane][i + u] - fma(alpha, sv[u], beta));
            }
            for (; i < n; i++) {
                float v = T[lane][i];
                if (FIT) {
                    v = (float)((double)v - fma(alpha, sf[t0 + i], beta));
                } else {
                    v = (float)((double)v - m);
                    v = (float)((double)v / sd);
                }
                T[lane][i] = v;
            }
        }
        store(t0, n);
    }
    (void)j;
}

// Time-major frames (cstride 1): lanes are consecutive channels (coalesced
// at every sample); each lane keeps U loads in flight ahead of its
// sequential sums.  A block has only nchan such threads (64 wavefronts at
// 4096 channels), so memory parallelism comes from U.
template <bool FIT>
__global__ __launch_bounds__(256) void k_zdot_cols(Frames F, const double *s, const double *sess)
{
    constexpr int U = 32;
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)F.nframes * F.nchan) return;
    const long f = g / F.nchan, j = g % F.nchan;
    float *col = F.x + f * F.fstride + j;
    const long ts = F.tstride, N = F.nsamp;
    const double *sf = s + f * N; // read only when FIT
    // One sweep over the column in batches of U samples; the next batch's
    // loads are issued before the current batch is processed, so they
    // overlap it.  body(i, v, sv) handles samples i..i+U-1 in order.
    auto sweep = [&](auto body, auto tail) {
        long i = 0;
        if (N >= U) {
            float nx[U];
            double snx[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                nx[u] = col[u * ts];
                if (FIT) snx[u] = sf[u];
            }
            for (; i + U <= N; i += U) {
                float v[U];
                double sv[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    v[u] = nx[u];
                    if (FIT) sv[u] = snx[u];
                }
                if (i + 2 * U <= N) {
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        nx[u] = col[(i + U + u) * ts];
                        if (FIT) snx[u] = sf[i + U + u];
                    }
                }
                body(i, v, sv);
            }
        }
        for (; i < N; i++) tail(i);
    };
    double a = 0., b = 0.;
    sweep(
        [&](long i, const float *v, const double *sv) {
#pragma unroll
            for (int u = 0; u < U; u++) {
                a += v[u];
                if (FIT) b = fma((double)v[u], sv[u], b);
                else b += v[u] * v[u];
            }
        },
        [&](long i) {
            const float v = col[i * ts];
            a += v;
            if (FIT) b = fma((double)v, sf[i], b);
            else b += v * v;
        });
    double m = 0., sd = 1., alpha = 0., beta = 0.;
    if (!FIT) {
        m = a / N;
        sd = b / N;
        sd = fma(-m, m, sd);
        sd = sqrt(sd);
        if (sd == 0.) sd = 1.;
    } else {
        const double se = sess[2 * f], ss = sess[2 * f + 1], nb = (double)N;
        const double tmp = fma(se, se, -(ss * nb));
        if (tmp != 0) {
            alpha = fma(a, se, -(b * nb)) / tmp;
            beta = fma(b, se, -(a * ss)) / tmp;
        }
    }
    auto update = [&](float v, double sv) {
        if (FIT) return (float)((double)v - fma(alpha, sv, beta));
        return (float)((double)((float)((double)v - m)) / sd);
    };
    sweep(
        [&](long i, const float *v, const double *sv) {
#pragma unroll
            for (int u = 0; u < U; u++) col[(i + u) * ts] = update(v[u], FIT ? sv[u] : 0.);
        },
        [&](long i) { col[i * ts] = update(col[i * ts], FIT ? sf[i] : 0.); });
}

// The channel mean for time-major frames: a workgroup takes 64 samples;
// 64-channel tiles of their rows pass through LDS, each lane (sample) then
// sums its row's channels in order.
__global__ __launch_bounds__(64) void k_zdot_s_rows(Frames F, double *s)
{
    __shared__ float T[64][65];
    const int lane = threadIdx.x;
    const long g0 = (long)blockIdx.x * 64; // first (frame, sample)
    const long ntot = (long)F.nframes * F.nsamp;
    const int nrow = (int)min((long)64, ntot - g0);
    double a = 0.;
    auto rowp = [&](int r) -> const float * {
        const long gr = g0 + r;
        return F.x + (gr / F.nsamp) * F.fstride + (gr % F.nsamp) * F.tstride;
    };
    for (int c0 = 0; c0 < F.nchan; c0 += 64) {
        const int n = min(64, F.nchan - c0);
        const int cl = c0 + min(lane, n - 1);
        // 16 row loads in flight (addresses clamped to rows and channels that exist)
        for (int r0 = 0; r0 < nrow; r0 += 16) {
            float v[16];
#pragma unroll
            for (int u = 0; u < 16; u++) v[u] = rowp(min(r0 + u, nrow - 1))[cl];
#pragma unroll
            for (int u = 0; u < 16; u++)
                if (r0 + u < nrow && lane < n) T[r0 + u][lane] = v[u];
        }
        __syncthreads();
        if (lane < nrow)
            for (int c = 0; c < n; c++) a += T[lane][c];
        __syncthreads();
    }
    if (lane < nrow) s[g0 + lane] = a / F.nchan;
}

// kadaneF 1: the td x fd downsampled frame, each output summed in float
// over k < fd (outer) and n < td (inner) (candidate.cpp:724-739)
__global__ void k_kad_down(Frames F, int td, int fd, int ncd, long nsd, float *ds)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)F.nframes * ncd * nsd) return;
    const long f = g / ((long)ncd * nsd), r = g % ((long)ncd * nsd);
    const long jd = r / nsd, id = r % nsd;
    float a = 0.f;
    for (int k = 0; k < fd; k++)
        for (int n = 0; n < td; n++) a += F.at(f, jd * fd + k, id * td + n);
    ds[g] = a;
}

// The maximum-sum contiguous run of sg*a[0..n) (the XLibs kadane<float>;
// declared: the running sum restarts after it drops below zero, strict '>'
// keeps the earliest maximum); returns the sum.
__device__ inline float kadane_run(const float *a, long n, float sg, long *start, long *end)
{
    float best = -INFINITY, cur = 0.f;
    long s = 0, bs = 0, be = 0;
    auto step = [&](long i, float v) {
        cur += sg * v;
        if (cur > best) {
            best = cur;
            bs = s;
            be = i;
        }
        if (cur < 0.f) {
            cur = 0.f;
            s = i + 1;
        }
    };
    long i = 0;
    for (; i + 16 <= n; i += 16) { // 16 loads in flight
        float v[16];
#pragma unroll
        for (int u = 0; u < 16; u++) v[u] = a[i + u];
#pragma unroll
        for (int u = 0; u < 16; u++) step(i + u, v[u]);
    }
    for (; i < n; i++) step(i, a[i]);
    *start = bs;
    *end = be;
    return best;
}

// kadaneF 2: per (frame, downsampled channel) the strongest positive and
// negative runs; a run longer than wnlimit whose S/N^2 = sum^2/(wn*td*fd)
// exceeds threshold^2 is zapped over its samples [start*td, (end+1)*td)
// (candidate.cpp:752-827).  zap[4*g..]: positive then negative range, empty
// when not zapped.
__global__ void k_kad_scan(int nframes, int ncd, long nsd, int td, int fd, float thre, const int *wnlimit,
                           const float *ds, long *zap)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)nframes * ncd) return;
    const long f = g / ncd;
    const float *ch = ds + g * nsd;
    const int wl = wnlimit[f];
    const float var = (float)(td * fd);
    for (int p = 0; p < 2; p++) {
        long s = 0, e = 0;
        const float box = kadane_run(ch, nsd, p ? -1.f : 1.f, &s, &e);
        const long wn = e - s + 1;
        float snr2 = 0.f;
        if (wn > wl) snr2 = box * box / ((float)wn * var);
        const bool hit = snr2 > thre * thre;
        zap[4 * g + 2 * p] = hit ? s * td : 0;
        zap[4 * g + 2 * p + 1] = hit ? (e + 1) * td : 0;
    }
}

// kadaneF 3: zero the zapped runs of every channel of each downsampled
// channel (channels past ncd*fd are left as they are); one thread per
// channel, which writes only its zapped samples (zaps are rare)
__global__ void k_kad_zero(Frames F, int fd, int ncd, const long *zap)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long per = (long)ncd * fd;
    if (g >= (long)F.nframes * per) return;
    const long f = g / per, j = g % per;
    const long *z = zap + 4 * (f * ncd + j / fd);
    for (int p = 0; p < 2; p++)
        for (long i = z[2 * p]; i < z[2 * p + 1]; i++) F.at(f, j, i) = 0.f;
}

inline unsigned blocks_for(long n) { return (unsigned)((n + 255) / 256); }

int check_frames(const Frames &F)
{
    if (!F.x || F.nframes < 0 || F.nchan < 1 || F.nsamp < 1 || F.cstride < 1 || F.tstride < 1 ||
        F.fstride < 0)
        return set_error(TX_EINVAL, "rfi: bad frame description");
    if ((long)F.nframes * F.nchan * F.nsamp > (1L << 40)) return set_error(TX_EINVAL, "rfi: frames too large");
    return TX_OK;
}

} // namespace

extern "C" size_t tx_rfi_work_bytes(int op, int nframes, int nchan, long nsamp, int td, int fd)
{
    if (nframes < 0 || nchan < 1 || nsamp < 1) return 0;
    if (op == TX_RFI_ZDOT) return (size_t)nframes * (nsamp + 2) * sizeof(double);
    if (op == TX_RFI_KADANEF) {
        if (td < 1 || fd < 1) return 0;
        const size_t ncd = nchan / fd, nsd = nsamp / td;
        const size_t ds = ((size_t)nframes * ncd * nsd * sizeof(float) + 255) / 256 * 256;
        return ds + (size_t)nframes * ncd * 4 * sizeof(long);
    }
    return 0;
}

extern "C" int tx_rfi_zdot(void *stream, float *data, int nframes, long fstride, int nchan, long nsamp,
                           long cstride, long tstride, void *work, size_t work_bytes)
{
    Frames F{data, fstride, cstride, tstride, nframes, nchan, nsamp};
    int rc = check_frames(F);
    if (rc) return rc;
    if (nframes == 0) return TX_OK;
    if (!work || work_bytes < tx_rfi_work_bytes(TX_RFI_ZDOT, nframes, nchan, nsamp, 1, 1))
        return set_error(TX_EINVAL, "rfi: work buffer smaller than tx_rfi_work_bytes");
    hipStream_t st = (hipStream_t)stream;
    double *s = (double *)work, *sess = s + (size_t)nframes * nsamp;
    const long nfc = (long)nframes * nchan, nfs = (long)nframes * nsamp;
    const double *cs = s, *css = sess;
    if (tstride == 1 && fstride == (long)nchan * nsamp && cstride == nsamp) {
        // the Candidate layout [nframes][nchan][nsamp]
        const unsigned nb = (unsigned)((nfc + 255) / 256);
        hipLaunchKernelGGL(k_zdot_rows<false>, dim3(nb), dim3(256), 0, st, F, cs, css);
        hipLaunchKernelGGL(k_zdot_s, dim3(blocks_for(nfs)), dim3(256), 0, st, F, s);
        hipLaunchKernelGGL(k_zdot_ss, dim3(blocks_for(nframes)), dim3(256), 0, st, F, cs, sess);
        hipLaunchKernelGGL(k_zdot_rows<true>, dim3(nb), dim3(256), 0, st, F, cs, css);
    } else if (cstride == 1 && fstride == (long)nchan * nsamp && tstride == nchan) {
        // the DataBuffer layout [nframes][nsamp][nchan]
        // one wavefront per workgroup: the few channel threads spread over
        // as many CUs as there are wavefronts
        const unsigned nw = (unsigned)((nfc + 63) / 64);
        hipLaunchKernelGGL(k_zdot_cols<false>, dim3(nw), dim3(64), 0, st, F, cs, css);
        hipLaunchKernelGGL(k_zdot_s_rows, dim3((unsigned)((nfs + 63) / 64)), dim3(64), 0, st, F, s);
        hipLaunchKernelGGL(k_zdot_ss, dim3(blocks_for(nframes)), dim3(256), 0, st, F, cs, sess);
        hipLaunchKernelGGL(k_zdot_cols<true>, dim3(nw), dim3(64), 0, st, F, cs, css);
    } else {
        hipLaunchKernelGGL(k_zdot_norm, dim3(blocks_for(nfc)), dim3(256), 0, st, F);
        hipLaunchKernelGGL(k_zdot_s, dim3(blocks_for(nfs)), dim3(256), 0, st, F, s);
        hipLaunchKernelGGL(k_zdot_ss, dim3(blocks_for(nframes)), dim3(256), 0, st, F, cs, sess);
        hipLaunchKernelGGL(k_zdot_fit, dim3(blocks_for(nfc)), dim3(256), 0, st, F, cs, css);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(TX_EHIP, std::string("rfi zdot: ") + hipGetErrorString(e));
    return TX_OK;
}

extern "C" int tx_rfi_kadanef(void *stream, float *data, int nframes, long fstride, int nchan, long nsamp,
                              long cstride, long tstride, int td, int fd, float threshold,
                              const int *wnlimit, void *work, size_t work_bytes)
{
    Frames F{data, fstride, cstride, tstride, nframes, nchan, nsamp};
    int rc = check_frames(F);
    if (rc) return rc;
    if (td < 1 || fd < 1) return set_error(TX_EINVAL, "rfi: td and fd must be >= 1");
    if (nframes == 0) return TX_OK;
    const int ncd = nchan / fd;
    const long nsd = nsamp / td;
    if (ncd == 0 || nsd == 0) return TX_OK; // nothing to downsample into
    if (!wnlimit) return set_error(TX_EINVAL, "rfi: null wnlimit");
    if (!work || work_bytes < tx_rfi_work_bytes(TX_RFI_KADANEF, nframes, nchan, nsamp, td, fd))
        return set_error(TX_EINVAL, "rfi: work buffer smaller than tx_rfi_work_bytes");
    hipStream_t st = (hipStream_t)stream;
    float *ds = (float *)work;
    long *zap = (long *)((char *)work + ((size_t)nframes * ncd * nsd * sizeof(float) + 255) / 256 * 256);
    hipLaunchKernelGGL(k_kad_down, dim3(blocks_for((long)nframes * ncd * nsd)), dim3(256), 0, st, F, td, fd,
                       ncd, nsd, ds);
    hipLaunchKernelGGL(k_kad_scan, dim3(blocks_for((long)nframes * ncd)), dim3(256), 0, st, nframes, ncd, nsd,
                       td, fd, threshold, wnlimit, (const float *)ds, zap);
    hipLaunchKernelGGL(k_kad_zero, dim3(blocks_for((long)nframes * ncd * fd)), dim3(256), 0, st, F, fd, ncd,
                       (const long *)zap);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(TX_EHIP, std::string("rfi kadaneF: ") + hipGetErrorString(e));
    return TX_OK;
}
