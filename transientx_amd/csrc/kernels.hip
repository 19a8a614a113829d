// kernels.hip -- gfx950 kernels of the TransientX search hot path.
//
//   K0 k_transpose_*   time-major input block -> channel-major input ring
//   K1 k_stage1_*      sub-band sums at the coarse DMs (stage 1)
//   K2 k_stage2        fine-DM sums of the sub-band series (stage 2)
//   K3 k_boxcar        Boxcar::match for every DM row + threshold compaction
//   k_subdata / k_cand_spectra   gathers for get_subdata and candplot's vpf
//
// Everything is VALU/LDS/HBM work; nothing here is a dense contraction, so
// no MFMA.  Wavefronts are 64 wide.
#include "kernels.h"

#include <cmath>

namespace txgpu {

// ---------------------------------------------------------------- K0
__global__ __launch_bounds__(256) void k_transpose_u8(const uint8_t *__restrict__ blk, long ns,
                                                      int C, const uint8_t *__restrict__ mask,
                                                      uint8_t zp, uint8_t *__restrict__ ring,
                                                      long pitch, long L, long W, long t0)
{
    __shared__ uint8_t tile[64][65];
    const long tb = (long)blockIdx.x * 64;
    const int cb = blockIdx.y * 64;
    const int tid = threadIdx.x, lane = tid & 63;
    {
        const int c = cb + lane;
        const bool live = c < C && (!mask || mask[c]);
        for (int r = tid >> 6; r < 64; r += 4) {
            const long t = tb + r;
            uint8_t v = zp;
            if (t < ns && live) v = blk[t * C + c];
            tile[r][lane] = v;
        }
    }
    __syncthreads();
    const long t = tb + lane;
    if (t >= ns) return;
    const long slot = lmod(t0 + t, L);
    for (int cc = tid >> 6; cc < 64; cc += 4) {
        const int ch = cb + cc;
        if (ch >= C) break;
        const uint8_t v = tile[lane][cc];
        uint8_t *row = ring + (long)ch * pitch;
        row[slot] = v;
        if (slot < W) row[slot + L] = v;
    }
}

__global__ __launch_bounds__(256) void k_transpose_f32(const float *__restrict__ blk, long ns,
                                                       int C, const uint8_t *__restrict__ mask,
                                                       float *__restrict__ ring, long pitch,
                                                       long L, long W, long t0)
{
    __shared__ float tile[64][65];
    const long tb = (long)blockIdx.x * 64;
    const int cb = blockIdx.y * 64;
    const int tid = threadIdx.x, lane = tid & 63;
    {
        const int c = cb + lane;
        const bool live = c < C && (!mask || mask[c]);
        for (int r = tid >> 6; r < 64; r += 4) {
            const long t = tb + r;
            float v = 0.f;
            if (t < ns && live) v = blk[t * C + c];
            tile[r][lane] = v;
        }
    }
    __syncthreads();
    const long t = tb + lane;
    if (t >= ns) return;
    const long slot = lmod(t0 + t, L);
    for (int cc = tid >> 6; cc < 64; cc += 4) {
        const int ch = cb + cc;
        if (ch >= C) break;
        const float v = tile[lane][cc];
        float *row = ring + (long)ch * pitch;
        row[slot] = v;
        if (slot < W) row[slot + L] = v;
    }
}

void launch_transpose_u8(hipStream_t st, const uint8_t *blk, long ns, int C, const uint8_t *mask,
                         uint8_t zp, uint8_t *ring, long pitch, long L, long W, long t0)
{
    dim3 grid((unsigned)((ns + 63) / 64), (unsigned)((C + 63) / 64));
    hipLaunchKernelGGL(k_transpose_u8, grid, dim3(256), 0, st, blk, ns, C, mask, zp, ring, pitch,
                       L, W, t0);
}

void launch_transpose_f32(hipStream_t st, const float *blk, long ns, int C, const uint8_t *mask,
                          float *ring, long pitch, long L, long W, long t0)
{
    dim3 grid((unsigned)((ns + 63) / 64), (unsigned)((C + 63) / 64));
    hipLaunchKernelGGL(k_transpose_f32, grid, dim3(256), 0, st, blk, ns, C, mask, ring, pitch, L,
                       W, t0);
}

// ---------------------------------------------------------------- K1
// One workgroup: sub-band blockIdx.y, coarse DMs [NCC*blockIdx.x, +NCC),
// output times tau0 + [0, 1024*Q).  Channel rows of the input ring are
// staged in LDS CH at a time, each row covering exactly the input times its
// NCC coarse DMs touch (per-channel base d1min).  Each thread owns Q quads of
// 4 consecutive output times.  The 8-bit samples are summed two at a time in
// 16-bit lanes of 32-bit registers (SWAR): v_perm_b32 pulls bytes
// (sh, sh+2) and (sh+1, sh+3) of the unaligned 4-byte window out of two
// aligned LDS words into the low bytes of the two 16-bit lanes.  A lane holds
// at most cps*255 <= 65280 for cps <= 256, so lanes never carry.
template <int NCC, int Q>
__global__ __launch_bounds__(256) void k_stage1_u8(Stage1Args a)
{
    extern __shared__ uint32_t lds_u[];
    const int kc = blockIdx.x, s = blockIdx.y;
    const long tau0 = a.tau_base + (long)blockIdx.z * (1024 * Q);
    const int k0 = kc * NCC;
    const int nk = min(NCC, a.ndmc - k0);
    const int tid = threadIdx.x;
    const uint8_t *ring = (const uint8_t *)a.ring;
    const int *d1min = a.d1min + (long)kc * a.C;
    uint32_t lo[NCC][Q], hi[NCC][Q];
#pragma unroll
    for (int k = 0; k < NCC; k++)
#pragma unroll
        for (int q = 0; q < Q; q++) lo[k][q] = hi[k][q] = 0u;

    for (int jc = 0; jc < a.cps; jc += a.CH) {
        const int nch = min(a.CH, a.cps - jc);
        __syncthreads();
        for (int jj = 0; jj < nch; jj++) {
            const int j = s * a.cps + jc + jj;
            const long start = lmod(tau0 + d1min[j], a.Lin);
            const uint32_t *src = (const uint32_t *)(ring + (long)j * a.pitch) + (start >> 2);
            uint32_t *dst = lds_u + jj * a.RS;
            for (int r = tid; r < a.RS; r += 256) dst[r] = src[r];
        }
        __syncthreads();
        for (int jj = 0; jj < nch; jj++) {
            const int j = s * a.cps + jc + jj;
            const int dmn = d1min[j];
            const int al = (int)(lmod(tau0 + dmn, a.Lin) & 3);
            const uint32_t *row = lds_u + jj * a.RS + tid;
#pragma unroll
            for (int k = 0; k < NCC; k++) {
                if (k < nk) {
                    const int e = al + a.d1[(long)(k0 + k) * a.C + j] - dmn;
                    const uint32_t sh = (uint32_t)(e & 3);
                    const uint32_t sel_lo = sh | 0x0C000C00u | ((sh + 2u) << 16);
                    const uint32_t sel_hi = (sh + 1u) | 0x0C000C00u | ((sh + 3u) << 16);
                    const uint32_t *p = row + (e >> 2);
#pragma unroll
                    for (int q = 0; q < Q; q++) {
                        const uint32_t w0 = p[q * 256], w1 = p[q * 256 + 1];
                        lo[k][q] += __builtin_amdgcn_perm(w1, w0, sel_lo);
                        hi[k][q] += __builtin_amdgcn_perm(w1, w0, sel_hi);
                    }
                }
            }
        }
    }

#pragma unroll
    for (int k = 0; k < NCC; k++) {
        if (k >= nk) continue;
        float *row = a.sub + ((long)(k0 + k) * a.nsub + s) * a.spitch;
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const long tau = tau0 + 4L * (tid + 256 * q);
            float v[4];
            v[0] = (float)((int)(lo[k][q] & 0xFFFFu) - a.zp_cps);
            v[1] = (float)((int)(hi[k][q] & 0xFFFFu) - a.zp_cps);
            v[2] = (float)((int)(lo[k][q] >> 16) - a.zp_cps);
            v[3] = (float)((int)(hi[k][q] >> 16) - a.zp_cps);
            const long slot = lmod(tau, a.Lsub);
            if (tau >= a.tau_begin && tau + 4 <= a.tau_end) {
                const float4 f = make_float4(v[0], v[1], v[2], v[3]);
                *(float4 *)(row + slot) = f;
                if (slot < a.Wsub) *(float4 *)(row + slot + a.Lsub) = f;
            } else {
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    if (tau + m >= a.tau_begin && tau + m < a.tau_end) {
                        row[slot + m] = v[m];
                        if (slot + m < a.Wsub) row[slot + m + a.Lsub] = v[m];
                    }
                }
            }
        }
    }
}

// f32 input: same tiling; each thread owns U = 4*Q output times strided by
// 256 (consecutive lanes read consecutive LDS words).  Channels are summed
// in ascending order, one FP32 add each, per output time.
template <int NCC, int Q>
__global__ __launch_bounds__(256) void k_stage1_f32(Stage1Args a)
{
    constexpr int U = 4 * Q;
    extern __shared__ float lds_f[];
    const int kc = blockIdx.x, s = blockIdx.y;
    const long tau0 = a.tau_base + (long)blockIdx.z * (1024 * Q);
    const int k0 = kc * NCC;
    const int nk = min(NCC, a.ndmc - k0);
    const int tid = threadIdx.x;
    const float *ring = (const float *)a.ring;
    const int *d1min = a.d1min + (long)kc * a.C;
    float acc[NCC][U];
#pragma unroll
    for (int k = 0; k < NCC; k++)
#pragma unroll
        for (int u = 0; u < U; u++) acc[k][u] = 0.f;

    for (int jc = 0; jc < a.cps; jc += a.CH) {
        const int nch = min(a.CH, a.cps - jc);
        __syncthreads();
        for (int jj = 0; jj < nch; jj++) {
            const int j = s * a.cps + jc + jj;
            const long start = lmod(tau0 + d1min[j], a.Lin);
            const float *src = ring + (long)j * a.pitch + start;
            float *dst = lds_f + jj * a.RS;
            for (int r = tid; r < a.RS; r += 256) dst[r] = src[r];
        }
        __syncthreads();
        for (int jj = 0; jj < nch; jj++) {
            const int j = s * a.cps + jc + jj;
            const int dmn = d1min[j];
            const float *row = lds_f + jj * a.RS + tid;
#pragma unroll
            for (int k = 0; k < NCC; k++) {
                if (k < nk) {
                    const float *p = row + (a.d1[(long)(k0 + k) * a.C + j] - dmn);
#pragma unroll
                    for (int u = 0; u < U; u++) acc[k][u] += p[u * 256];
                }
            }
        }
    }

#pragma unroll
    for (int k = 0; k < NCC; k++) {
        if (k >= nk) continue;
        float *row = a.sub + ((long)(k0 + k) * a.nsub + s) * a.spitch;
#pragma unroll
        for (int u = 0; u < U; u++) {
            const long tau = tau0 + tid + 256L * u;
            if (tau >= a.tau_begin && tau < a.tau_end) {
                const long slot = lmod(tau, a.Lsub);
                row[slot] = acc[k][u];
                if (slot < a.Wsub) row[slot + a.Lsub] = acc[k][u];
            }
        }
    }
}

template <int NCC, int Q>
static int launch_s1(hipStream_t st, const Stage1Args &a, bool u8)
{
    const long span = a.tau_end - a.tau_base;
    dim3 grid((unsigned)((a.ndmc + NCC - 1) / NCC), (unsigned)a.nsub,
              (unsigned)((span + 1024 * Q - 1) / (1024 * Q)));
    const size_t shm = (size_t)a.CH * a.RS * 4;
    if (u8) {
        (void)hipFuncSetAttribute((const void *)k_stage1_u8<NCC, Q>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        hipLaunchKernelGGL((k_stage1_u8<NCC, Q>), grid, dim3(256), shm, st, a);
    } else {
        (void)hipFuncSetAttribute((const void *)k_stage1_f32<NCC, Q>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        hipLaunchKernelGGL((k_stage1_f32<NCC, Q>), grid, dim3(256), shm, st, a);
    }
    return 0;
}

#define S1_CASE(n, q)                                                                              \
    if (NCC == n && Q == q) return launch_s1<n, q>(st, a, u8);

static int launch_stage1(hipStream_t st, const Stage1Args &a, int NCC, int Q, bool u8)
{
    S1_CASE(16, 2) S1_CASE(16, 1) S1_CASE(8, 2) S1_CASE(8, 1) S1_CASE(4, 2) S1_CASE(4, 1)
    S1_CASE(2, 2) S1_CASE(2, 1) S1_CASE(1, 2) S1_CASE(1, 1)
    return -1;
}

int launch_stage1_u8(hipStream_t st, const Stage1Args &a, int NCC, int Q)
{
    return launch_stage1(st, a, NCC, Q, true);
}
int launch_stage1_f32(hipStream_t st, const Stage1Args &a, int NCC, int Q)
{
    return launch_stage1(st, a, NCC, Q, false);
}

// ---------------------------------------------------------------- K2
// One workgroup: coarse group blockIdx.z, DM block blockIdx.x of that group
// (4*DPW fine DMs, DPW per wave), output samples t0 + [0, 64*R).  Sub-band
// rows of the group are staged SB at a time, each covering the times its
// DM block touches; lanes own consecutive samples, so LDS reads are
// conflict-free.  Sub-bands are added in ascending order (FP32).
template <int DPW, int R>
__global__ __launch_bounds__(256) void k_stage2(Stage2Args a)
{
    extern __shared__ float lds2[];
    const int bx = blockIdx.x;
    const long t0 = (long)blockIdx.y * (64 * R);
    const int k = blockIdx.z;
    const int blk = k * a.nblk + bx;
    const long gend = min((long)(k + 1) * a.group, a.ndm);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const long ibase = (long)k * a.group + (long)bx * 4 * DPW + wave; // DM ibase + 4*d
    const int *d2min = a.d2min + (long)blk * a.nsub;
    float acc[DPW][R];
#pragma unroll
    for (int d = 0; d < DPW; d++)
#pragma unroll
        for (int r = 0; r < R; r++) acc[d][r] = 0.f;

    for (int sc = 0; sc < a.nsub; sc += a.SB) {
        const int nsb = min(a.SB, a.nsub - sc);
        __syncthreads();
        for (int ss = 0; ss < nsb; ss++) {
            const int s = sc + ss;
            const long start = lmod(a.w0 + t0 + d2min[s], a.Lsub);
            const float *src = a.sub + ((long)k * a.nsub + s) * a.spitch + start;
            float *dst = lds2 + ss * a.RS2;
            for (int r = tid; r < a.RS2; r += 256) dst[r] = src[r];
        }
        __syncthreads();
        for (int ss = 0; ss < nsb; ss++) {
            const int s = sc + ss;
            const float *row = lds2 + ss * a.RS2 + lane - d2min[s];
#pragma unroll
            for (int d = 0; d < DPW; d++) {
                const long i = ibase + 4 * d;
                if (i < gend) {
                    const float *p = row + a.d2[i * a.nsub + s];
#pragma unroll
                    for (int r = 0; r < R; r++) acc[d][r] += p[64 * r];
                }
            }
        }
    }
#pragma unroll
    for (int d = 0; d < DPW; d++) {
        const long i = ibase + 4 * d;
        if (i >= gend) continue;
        float *out = a.tim + i * a.N;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const long t = t0 + lane + 64 * r;
            if (t < a.N) out[t] = acc[d][r];
        }
    }
}

template <int DPW, int R>
static int launch_s2(hipStream_t st, const Stage2Args &a)
{
    dim3 grid((unsigned)a.nblk, (unsigned)((a.N + 64 * R - 1) / (64 * R)), (unsigned)a.ndmc);
    const size_t shm = (size_t)a.SB * a.RS2 * 4;
    (void)hipFuncSetAttribute((const void *)k_stage2<DPW, R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)shm);
    hipLaunchKernelGGL((k_stage2<DPW, R>), grid, dim3(256), shm, st, a);
    return 0;
}

#define S2_CASE(d, r)                                                                              \
    if (DPW == d && R == r) return launch_s2<d, r>(st, a);

int launch_stage2(hipStream_t st, const Stage2Args &a, int DPW, int R)
{
    S2_CASE(1, 8) S2_CASE(2, 8) S2_CASE(4, 8) S2_CASE(6, 8) S2_CASE(8, 8) S2_CASE(12, 4)
    S2_CASE(16, 4)
    return -1;
}

// ---------------------------------------------------------------- K3
__device__ inline double wave_sum_d(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Deterministic workgroup sum: per-wave butterflies, then every thread adds
// the wave partials in wave order.
__device__ inline double block_sum_d(double v, double *red)
{
    v = wave_sum_d(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.;
    const int nw = blockDim.x >> 6;
    for (int w = 0; w < nw; w++) t += red[w];
    return t;
}

__device__ inline uint32_t float_key(float f)
{
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ inline float key_float(uint32_t k)
{
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

// k-th smallest (0-based) of cs[0..N) by 8-bit MSB-first radix selection.
__device__ float select_kth(const float *cs, long N, long kth, uint32_t *hist, uint32_t *sh)
{
    uint32_t prefix = 0, pmask = 0;
    long k = kth;
    for (int pass = 0; pass < 4; pass++) {
        const int shift = 24 - 8 * pass;
        __syncthreads();
        for (int b = threadIdx.x; b < 256; b += blockDim.x) hist[b] = 0;
        __syncthreads();
        for (long i = threadIdx.x; i < N; i += blockDim.x) {
            const uint32_t key = float_key(cs[i]);
            if ((key & pmask) == prefix) atomicAdd(&hist[(key >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            long c = 0;
            int b = 0;
            for (; b < 256; b++) {
                if (c + (long)hist[b] > k) break;
                c += hist[b];
            }
            sh[0] = (uint32_t)b;
            sh[1] = (uint32_t)(k - c);
        }
        __syncthreads();
        prefix |= sh[0] << shift;
        pmask |= 255u << shift;
        k = sh[1];
    }
    return key_float(prefix);
}

// One workgroup of 1024 threads per DM row.  The row's running sum lives in
// LDS (USE_LDS) or, for rows longer than the LDS of one workgroup, in a
// global scratch row.
template <int R, bool USE_LDS>
__global__ __launch_bounds__(1024) void k_boxcar(BoxcarArgs a)
{
    extern __shared__ float lds_cs[];
    float *cs = USE_LDS ? lds_cs : a.scratch + (long)blockIdx.x * a.N;
    __shared__ double red[16];
    __shared__ float s_temp[256];
    __shared__ float s_stats[2];
    __shared__ uint32_t s_hist[256];
    __shared__ uint32_t s_sel[2];
    const long row = blockIdx.x;
    const long N = a.N;
    const float *x = a.tim + row * N;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    float mean_f, var_f;
    for (long i = tid; i < N; i += 1024) cs[i] = x[i];
    __syncthreads();
    if (a.iqr) {
        // src/singlepulse/boxcar.cpp:217-226: Q1, Q2 ascending order
        // statistics at N/4, N/2; Q3 the N/4-th from the top.
        const float Q1 = select_kth(cs, N, N / 4, s_hist, s_sel);
        const float Q2 = select_kth(cs, N, N / 2, s_hist, s_sel);
        const float Q3 = select_kth(cs, N, N - 1 - N / 4, s_hist, s_sel);
        mean_f = Q2;
        const double q = (double)(Q3 - Q1) / 1.349;
        var_f = (float)(q * q);
    } else if (a.exact) {
        // src/singlepulse/boxcar.cpp:206-213 in the order the -O3 -mfma
        // build evaluates it (FMA-contracted, sequential)
        if (tid == 0) {
            float m = 0.f, v = 0.f;
            for (long i = 0; i < N; i++) {
                const float t = cs[i];
                m += t;
                v = fmaf(t, t, v);
            }
            m /= (float)N;
            v /= (float)N;
            v = fmaf(-m, m, v);
            s_stats[0] = m;
            s_stats[1] = v;
        }
        __syncthreads();
        mean_f = s_stats[0];
        var_f = s_stats[1];
    } else {
        double sum = 0.;
        for (long i = tid; i < N; i += 1024) sum += (double)cs[i];
        sum = block_sum_d(sum, red);
        const double mean_d = sum / (double)N;
        double ss = 0.;
        for (long i = tid; i < N; i += 1024) {
            const double d = (double)cs[i] - mean_d;
            ss += d * d;
        }
        ss = block_sum_d(ss, red);
        mean_f = (float)mean_d;
        var_f = (float)(ss / (double)N);
    }
    __syncthreads();

    // prefix sum of (x - mean) (src/singlepulse/boxcar.cpp:235-240), in place
    if (a.exact) {
        if (tid == 0) {
            float c = 0.f;
            for (long i = 0; i < N; i++) {
                c += cs[i] - mean_f;
                cs[i] = c;
            }
        }
    } else {
        const long seg = ((N + 15) / 16 + 63) / 64 * 64;
        const long s0 = (long)wave * seg, s1 = min(N, s0 + seg);
        double tot = 0.;
        for (long i = s0 + lane; i < s1; i += 64) tot += (double)(cs[i] - mean_f);
        tot = wave_sum_d(tot);
        if (lane == 0) red[wave] = tot;
        __syncthreads();
        double carry = 0.;
        for (int w = 0; w < wave; w++) carry += red[w];
        for (long base = s0; base < s1; base += 64) {
            const long i = base + lane;
            double v = i < s1 ? (double)(cs[i] - mean_f) : 0.;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const double u = __shfl_up(v, o, 64);
                if (lane >= o) v += u;
            }
            if (i < s1) cs[i] = (float)(carry + v);
            carry += __shfl(v, 63, 64);
        }
    }
    if (tid < a.nbox) {
        const int wn = a.vwn[tid];
        s_temp[tid] = (float)sqrt(1. / (double)((float)wn * var_f));
    }
    __syncthreads();

    bool finite = true;
    for (int k = 0; k < a.nbox; k++) finite = finite && isfinite(s_temp[k]);
    const bool do_box = var_f > 0.f;
    const long drow = row;
    for (long c0 = 0; c0 < N; c0 += 1024L * R) {
        float bS[R];
        int bw[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            bS[r] = 0.f;
            bw[r] = 0;
        }
        if (do_box) {
            const bool fast = finite && c0 >= a.lo_all && c0 + 1024L * R <= a.hi_all;
            for (int k = 0; k < a.nbox; k++) {
                const int wn = a.vwn[k];
                const int wl = wn / 2 + 1, wh = (wn - 1) / 2;
                const float temp = s_temp[k];
                if (fast) {
                    const float *ph = cs + (c0 + tid + wh);
                    const float *pl = cs + (c0 + tid - wl);
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        const float S = (ph[1024 * r] - pl[1024 * r]) * temp;
                        const bool ge = S >= bS[r];
                        bS[r] = ge ? S : bS[r];
                        bw[r] = ge ? wn : bw[r];
                    }
                } else {
                    const long hik = a.whi[k];
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        const long i = c0 + tid + 1024L * r;
                        const bool valid = i >= wl && i < hik;
                        const long ia = valid ? i + wh : 0, ib = valid ? i - wl : 0;
                        const float S = (cs[ia] - cs[ib]) * temp;
                        if (valid) {
                            // the AVX2 blend (boxcar.cpp:273-277):
                            // (S>=vS ? S : 0) + (S<vS ? vS : 0)
                            const bool ge = S >= bS[r], lt = S < bS[r];
                            bS[r] = ge ? S : (lt ? bS[r] : 0.f);
                            bw[r] = ge ? wn : (lt ? bw[r] : 0);
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            const long i = c0 + tid + 1024L * r;
            const float v = bS[r] + 0.f;
            const bool in = i < N;
            if (in && a.store_maps) {
                a.mxS[drow * N + i] = v;
                a.mxwn[drow * N + i] = bw[r];
            }
            const bool hit = in && v > a.thre;
            const unsigned long long m = __ballot(hit);
            if (m) {
                const int leader = __ffsll((long long)m) - 1;
                unsigned long long base = 0;
                if (lane == leader) base = atomicAdd(a.npts, (unsigned long long)__popcll(m));
                base = __shfl(base, leader, 64);
                if (hit) {
                    const unsigned long long idx =
                        base + (unsigned long long)__popcll(m & ((1ull << lane) - 1ull));
                    if (idx < (unsigned long long)a.cap) {
                        tx_point p;
                        p.idm = (int32_t)drow;
                        p.isamp = (int32_t)i;
                        p.S = v;
                        p.wn = bw[r];
                        a.pts[idx] = p;
                    }
                }
            }
        }
    }
}

bool boxcar_fits_lds(long N, int max_lds_bytes) { return N * 4 + 4096 <= (long)max_lds_bytes; }

int launch_boxcar(hipStream_t st, const BoxcarArgs &a, int max_lds_bytes)
{
    constexpr int R = 8;
    if (a.nbox > 256) return -1;
    if (boxcar_fits_lds(a.N, max_lds_bytes)) {
        const size_t shm = (size_t)a.N * 4;
        (void)hipFuncSetAttribute((const void *)k_boxcar<R, true>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        hipLaunchKernelGGL((k_boxcar<R, true>), dim3((unsigned)a.ndm), dim3(1024), shm, st, a);
    } else {
        if (!a.scratch) return -1;
        hipLaunchKernelGGL((k_boxcar<R, false>), dim3((unsigned)a.ndm), dim3(1024), 0, st, a);
    }
    return 0;
}

// ---------------------------------------------------------------- gathers
__global__ void k_subdata(const float *__restrict__ sub, long spitch, long Lsub,
                          const int *__restrict__ d2, int nsub, int group, long idm, long w0,
                          long N, float *__restrict__ out)
{
    const int s = blockIdx.y;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const long k = idm / group;
    const float *row = sub + (k * nsub + s) * spitch;
    out[(long)s * N + t] = row[lmod(w0 + t + d2[idm * nsub + s], Lsub)];
}

void launch_subdata(hipStream_t st, const float *sub, long spitch, long Lsub, const int *d2,
                    int nsub, int group, long idm, long w0, long N, float *out)
{
    dim3 grid((unsigned)((N + 255) / 256), (unsigned)nsub);
    hipLaunchKernelGGL(k_subdata, grid, dim3(256), 0, st, sub, spitch, Lsub, d2, nsub, group, idm,
                       w0, N, out);
}

// candplot.cpp:160-170: vpf[j] += sub/fcnt over isamp-wn < i < isamp+wn,
// sequential in i like the reference loop.
__global__ void k_cand_spectra(const float *__restrict__ sub, long spitch, long Lsub,
                               const int *__restrict__ d2, const float *__restrict__ fcnt, int nsub,
                               int group, long w0, long N, int ncand, const int *__restrict__ idm,
                               const int *__restrict__ isamp, const int *__restrict__ wn,
                               float *__restrict__ out)
{
    const int c = blockIdx.x;
    const long i = idm[c];
    const long k = i / group;
    long a0 = (long)isamp[c] - wn[c] + 1, a1 = (long)isamp[c] + wn[c] - 1;
    a0 = a0 < 0 ? 0 : a0;
    a1 = a1 > N - 1 ? N - 1 : a1;
    for (int s = threadIdx.x; s < nsub; s += blockDim.x) {
        const float *row = sub + (k * nsub + s) * spitch;
        const long d = d2[i * nsub + s];
        const float f = fcnt[s];
        float acc = 0.f;
        for (long t = a0; t <= a1; t++) acc += row[lmod(w0 + t + d, Lsub)] / f;
        out[(long)c * nsub + s] = acc;
    }
}

void launch_cand_spectra(hipStream_t st, const float *sub, long spitch, long Lsub, const int *d2,
                         const float *fcnt, int nsub, int group, long w0, long N, int ncand,
                         const int *idm, const int *isamp, const int *wn, float *out)
{
    if (ncand <= 0) return;
    hipLaunchKernelGGL(k_cand_spectra, dim3((unsigned)ncand), dim3(128), 0, st, sub, spitch, Lsub,
                       d2, fcnt, nsub, group, w0, N, ncand, idm, isamp, wn, out);
}

} // namespace txgpu
