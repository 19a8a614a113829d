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
    for (long j = 0; j < F.nchan; j++) a += F.at(f, j, i);
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
    for (long i = 0; i < F.nsamp; i++) {
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
    for (long i = 0; i < n; i++) {
        cur += sg * a[i];
        if (cur > best) {
            best = cur;
            bs = s;
            be = i;
        }
        if (cur < 0.f) {
            cur = 0.f;
            s = i + 1;
        }
    }
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
// channel (channels past ncd*fd are left as they are)
__global__ void k_kad_zero(Frames F, int fd, int ncd, const long *zap)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long per = (long)ncd * fd * F.nsamp;
    if (g >= (long)F.nframes * per) return;
    const long f = g / per, r = g % per;
    const long j = r / F.nsamp, i = r % F.nsamp;
    const long *z = zap + 4 * (f * ncd + j / fd);
    if ((i >= z[0] && i < z[1]) || (i >= z[2] && i < z[3])) F.at(f, j, i) = 0.f;
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
    hipLaunchKernelGGL(k_zdot_norm, dim3(blocks_for(nfc)), dim3(256), 0, st, F);
    hipLaunchKernelGGL(k_zdot_s, dim3(blocks_for(nfs)), dim3(256), 0, st, F, s);
    hipLaunchKernelGGL(k_zdot_ss, dim3(blocks_for(nframes)), dim3(256), 0, st, F, (const double *)s, sess);
    hipLaunchKernelGGL(k_zdot_fit, dim3(blocks_for(nfc)), dim3(256), 0, st, F, (const double *)s,
                       (const double *)sess);
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
    hipLaunchKernelGGL(k_kad_zero, dim3(blocks_for((long)nframes * ncd * fd * nsamp)), dim3(256), 0, st, F,
                       fd, ncd, (const long *)zap);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(TX_EHIP, std::string("rfi kadaneF: ") + hipGetErrorString(e));
    return TX_OK;
}
