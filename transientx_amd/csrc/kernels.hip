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
#include "device_util.h"

#include <cmath>
#include <type_traits>

namespace txgpu {

// ---------------------------------------------------------------- K0
// Tiles of 64 ring slots x 128 channels, the slots aligned to 64 (L and W
// are multiples of 64): tile bx covers slots sa + 64*bx + [0, 64) with sa =
// s0 rounded down to 64, i.e. block times 64*bx - lead + [0, 64).  Loads:
// lane = 2 channels 64 apart, one whole 128-byte line of a time row per pair
// of byte loads.  Stores: 16 lanes write one channel's 64
// slots as aligned dwords, a whole 64-byte line; only the first and last
// tile of a block hold slots of other blocks, and write those bytes singly.
__global__ __launch_bounds__(256) void k_transpose_u8(const uint8_t *__restrict__ blk, long ns,
                                                      int C, const uint8_t *__restrict__ mask,
                                                      uint8_t zp, uint8_t *__restrict__ ring,
                                                      long pitch, long L, long W, long s0,
                                                      const long *clock)
{
    if (clock) s0 = lmod(*clock, L);
    __shared__ uint32_t tile32[128 * 17]; // [channel][64 bytes + 4 pad]
    uint8_t *tile = (uint8_t *)tile32;
    const long sa = s0 & ~63L;
    const int lead = (int)(s0 - sa);
    const long tb = (long)blockIdx.x * 64 - lead; // block time of the tile's first slot
    const int cb = blockIdx.y * 128;
    const int tid = threadIdx.x, lane = tid & 63;
    {
        // lane = channels cb + lane and cb + 64 + lane: the two byte loads of
        // a time row cover one whole 128-byte line, and each LDS store hits
        // 64 distinct tile rows 17 dwords apart (no bank conflict)
        const int c0 = cb + lane, c1 = cb + 64 + lane;
        const bool live0 = c0 < C && (!mask || mask[c0]);
        const bool live1 = c1 < C && (!mask || mask[c1]);
        const int r0 = tid >> 6;
        uint8_t v0[16], v1[16];
#pragma unroll
        for (int u = 0; u < 16; u++) { // all 32 loads in flight before the LDS stores
            const long t = tb + r0 + 4 * u;
            const bool in = t >= 0 && t < ns;
            v0[u] = (in && live0) ? blk[t * C + c0] : zp;
            v1[u] = (in && live1) ? blk[t * C + c1] : zp;
        }
#pragma unroll
        for (int u = 0; u < 16; u++) {
            tile[lane * 68 + r0 + 4 * u] = v0[u];
            tile[(64 + lane) * 68 + r0 + 4 * u] = v1[u];
        }
    }
    __syncthreads();
    long base = sa + (long)blockIdx.x * 64; // L is a multiple of 64: the tile wraps whole
    if (base >= L) base -= L;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int idx = tid + 256 * k;
        const int cc = idx >> 4, q = idx & 15;
        const int ch = cb + cc;
        if (ch >= C) continue;
        const uint32_t w = tile32[cc * 17 + q];
        uint8_t *row = ring + (long)ch * pitch;
        const long slot = base + 4 * q;
        const long t = tb + 4 * q;
        if (t >= 0 && t + 4 <= ns) {
            *(uint32_t *)(row + slot) = w;
            if (slot < W) *(uint32_t *)(row + slot + L) = w;
        } else {
#pragma unroll
            for (int m = 0; m < 4; m++)
                if (t + m >= 0 && t + m < ns) {
                    const uint8_t b = (uint8_t)(w >> (8 * m));
                    row[slot + m] = b;
                    if (slot + m < W) row[slot + m + L] = b;
                }
        }
    }
}

__global__ __launch_bounds__(256) void k_transpose_f32(const float *__restrict__ blk, long ns,
                                                       int C, const uint8_t *__restrict__ mask,
                                                       float *__restrict__ ring, long pitch,
                                                       long L, long W, long s0, const long *clock)
{
    if (clock) s0 = lmod(*clock, L);
    __shared__ float tile[64][65];
    const long tb = (long)blockIdx.x * 64;
    const int cb = blockIdx.y * 64;
    const int tid = threadIdx.x, lane = tid & 63;
    {
        const int c = cb + lane;
        const bool live = c < C && (!mask || mask[c]);
        const int r0 = tid >> 6;
        float v[16];
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const long t = tb + r0 + 4 * u;
            v[u] = (t < ns && live) ? blk[t * C + c] : 0.f;
        }
#pragma unroll
        for (int u = 0; u < 16; u++) tile[r0 + 4 * u][lane] = v[u];
    }
    __syncthreads();
    const long t = tb + lane;
    if (t >= ns) return;
    long slot = s0 + (long)blockIdx.x * 64 + lane;
    while (slot >= L) slot -= L;
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
                         uint8_t zp, uint8_t *ring, long pitch, long L, long W, long t0,
                         const long *clock, LaunchSink *sink)
{
    // tiles start on 64-aligned ring slots: up to 63 slots of lead, so one
    // more tile than ns/64 (a graph pipeline does not know t0: always one)
    dim3 grid((unsigned)((ns + 63) / 64 + 1), (unsigned)((C + 127) / 128));
    const long s0 = lmod(t0, L);
    kemit(sink, st, k_transpose_u8, grid, dim3(256), 0, blk, ns, C, mask, zp, ring, pitch,
          L, W, s0, clock);
}

void launch_transpose_f32(hipStream_t st, const float *blk, long ns, int C, const uint8_t *mask,
                          float *ring, long pitch, long L, long W, long t0, const long *clock,
                          LaunchSink *sink)
{
    dim3 grid((unsigned)((ns + 63) / 64), (unsigned)((C + 63) / 64));
    const long s0 = lmod(t0, L);
    kemit(sink, st, k_transpose_f32, grid, dim3(256), 0, blk, ns, C, mask, ring, pitch, L,
          W, s0, clock);
}

// K0 with the front end (8-bit ingest): the tiles of k_transpose_f32, each
// value computed from the raw block.  Lane = output channel c, 16 output
// times per thread; for each of the td x fd neighbours (ii, jj) the 16
// bytes blk[((t*td + ii)*Cin + c*fd + jj] are loaded at once (lanes read
// fd-strided bytes of one input row) and added as integers.  TD/FD > 0 fix
// td/fd at compile time (the common td = fd = 1).
template <int TD, int FD>
__global__ __launch_bounds__(256) void k_frontend_u8(const uint8_t *__restrict__ blk, long ns, int Cin,
                                                     int td, int fd, const float *__restrict__ off,
                                                     const float *__restrict__ gain,
                                                     const uint8_t *__restrict__ mask,
                                                     float *__restrict__ ring, long pitch, long L,
                                                     long W, long s0, const long *clock)
{
    if (clock) s0 = lmod(*clock, L);
    if (TD) td = TD;
    if (FD) fd = FD;
    const int C = Cin / fd;
    __shared__ float tile[64][65];
    const long tb = (long)blockIdx.x * 64;
    const int cb = blockIdx.y * 64;
    const int tid = threadIdx.x, lane = tid & 63;
    {
        const int c = cb + lane;
        const bool live = c < C && (!mask || mask[c]);
        const int r0 = tid >> 6;
        int acc[16];
#pragma unroll
        for (int u = 0; u < 16; u++) acc[u] = 0;
        if (live) {
            const long rs = (long)td * Cin; // bytes per output sample
            for (int ii = 0; ii < td; ii++)
                for (int jj = 0; jj < fd; jj++) {
                    const uint8_t *q = blk + (long)ii * Cin + (long)c * fd + jj;
                    uint8_t v[16];
#pragma unroll
                    for (int u = 0; u < 16; u++) {
                        const long t = tb + r0 + 4 * u;
                        v[u] = t < ns ? q[t * rs] : (uint8_t)0;
                    }
#pragma unroll
                    for (int u = 0; u < 16; u++) acc[u] += v[u];
                }
        }
        const float o = live ? off[c] : 0.f, g = live ? gain[c] : 0.f;
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const float d = (float)acc[u] - o;
            tile[r0 + 4 * u][lane] = live ? d * g : 0.f;
        }
    }
    __syncthreads();
    const long t = tb + lane;
    if (t >= ns) return;
    long slot = s0 + (long)blockIdx.x * 64 + lane;
    while (slot >= L) slot -= L;
    for (int cc = tid >> 6; cc < 64; cc += 4) {
        const int ch = cb + cc;
        if (ch >= C) break;
        const float v = tile[lane][cc];
        float *row = ring + (long)ch * pitch;
        row[slot] = v;
        if (slot < W) row[slot + L] = v;
    }
}

// The front end into a time-major [ns][Cin/fd] block: thread = output
// channel (coalesced stores), rows of the grid stride over time.
__global__ __launch_bounds__(256) void k_frontend_tm(const uint8_t *__restrict__ blk, long ns, int Cin,
                                                     int td, int fd, const float *__restrict__ off,
                                                     const float *__restrict__ gain,
                                                     const uint8_t *__restrict__ mask,
                                                     float *__restrict__ out)
{
    const int C = Cin / fd;
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= C) return;
    const bool live = !mask || mask[c];
    const float o = off[c], g = gain[c];
    for (long t = blockIdx.y; t < ns; t += gridDim.y) {
        int y = 0;
        for (int ii = 0; ii < td; ii++) {
            const uint8_t *q = blk + (t * td + ii) * Cin + (long)c * fd;
            for (int jj = 0; jj < fd; jj++) y += q[jj];
        }
        const float d = (float)y - o;
        out[t * C + c] = live ? d * g : 0.f;
    }
}

__global__ void k_frontend_zero(unsigned long long *acc, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        acc[i] = 0ull;
}

// Exact sums of y and y^2 per output channel: workgroup = 64 channels x a
// chunk of 256 output samples (4 rows of 64 lanes), one atomic add per
// channel and chunk.  Integer sums, so the order of the adds does not
// matter.
__global__ __launch_bounds__(256) void k_frontend_sums(const uint8_t *__restrict__ blk, long ns, int Cin,
                                                       int td, int fd, unsigned long long *acc)
{
    const int C = Cin / fd;
    const int tid = threadIdx.x, lane = tid & 63, r0 = tid >> 6;
    const int c = blockIdx.y * 64 + lane;
    const long t0 = (long)blockIdx.x * 256;
    __shared__ unsigned long long red[2][4][64];
    unsigned long long s1 = 0, s2 = 0;
    if (c < C) {
        for (int k = 0; k < 4; k++) {
            int y[16];
#pragma unroll
            for (int u = 0; u < 16; u++) y[u] = 0;
            for (int ii = 0; ii < td; ii++)
                for (int jj = 0; jj < fd; jj++) {
                    const uint8_t *q = blk + (long)ii * Cin + (long)c * fd + jj;
                    uint8_t v[16];
#pragma unroll
                    for (int u = 0; u < 16; u++) {
                        const long t = t0 + 64 * k + r0 + 4 * u;
                        v[u] = t < ns ? q[t * td * Cin] : (uint8_t)0;
                    }
#pragma unroll
                    for (int u = 0; u < 16; u++) y[u] += v[u];
                }
#pragma unroll
            for (int u = 0; u < 16; u++) {
                s1 += (unsigned)y[u];
                s2 += (unsigned long long)y[u] * (unsigned)y[u];
            }
        }
    }
    red[0][r0][lane] = s1;
    red[1][r0][lane] = s2;
    __syncthreads();
    if (r0 == 0 && c < C) {
        s1 = red[0][0][lane] + red[0][1][lane] + red[0][2][lane] + red[0][3][lane];
        s2 = red[1][0][lane] + red[1][1][lane] + red[1][2][lane] + red[1][3][lane];
        atomicAdd(&acc[2 * (long)c], s1);
        atomicAdd(&acc[2 * (long)c + 1], s2);
    }
}

// mean = S1/n, var = S2/n - mean*mean in double (no contraction);
// off = (float)mean, gain = var > 0 ? (float)(1/sqrt(var)) : 1.
__global__ void k_frontend_fin(const unsigned long long *acc, int C, long ns, float *off, float *gain)
{
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const double n = (double)ns;
    const double mean = (double)acc[2 * (long)c] / n;
    const double m2 = mean * mean;
    const double var = (double)acc[2 * (long)c + 1] / n - m2;
    off[c] = (float)mean;
    gain[c] = var > 0. ? (float)(1. / sqrt(var)) : 1.f;
}

void launch_frontend_u8(hipStream_t st, const uint8_t *blk, long ns, int Cin, int td, int fd,
                        const float *off, const float *gain, const uint8_t *mask, float *ring,
                        long pitch, long L, long W, long t0, const long *clock, LaunchSink *sink)
{
    const int C = Cin / fd;
    dim3 grid((unsigned)((ns + 63) / 64), (unsigned)((C + 63) / 64));
    const long s0 = lmod(t0, L);
    auto f = (td == 1 && fd == 1) ? k_frontend_u8<1, 1> : k_frontend_u8<0, 0>;
    kemit(sink, st, f, grid, dim3(256), 0, blk, ns, Cin, td, fd, off, gain, mask, ring, pitch, L, W, s0,
          clock);
}

// Zeroes the masked channels of a time-major f32 block [ns][C].
__global__ __launch_bounds__(256) void k_mask_tm(float *blk, long ns, int C, const uint8_t *mask)
{
    const long n = ns * (long)C;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256)
        if (!mask[i % C]) blk[i] = 0.f;
}

void launch_mask_tm(hipStream_t st, float *blk, long ns, int C, const uint8_t *mask)
{
    const long n = ns * (long)C;
    hipLaunchKernelGGL(k_mask_tm, dim3((unsigned)std::min<long>((n + 255) / 256, 65536)), dim3(256), 0, st, blk,
                       ns, C, mask);
}

void launch_frontend_tm(hipStream_t st, const uint8_t *blk, long ns, int Cin, int td, int fd,
                        const float *off, const float *gain, const uint8_t *mask, float *out)
{
    const int C = Cin / fd;
    dim3 grid((unsigned)((C + 255) / 256), (unsigned)std::min<long>(std::max<long>(ns, 1), 8192));
    hipLaunchKernelGGL(k_frontend_tm, grid, dim3(256), 0, st, blk, ns, Cin, td, fd, off, gain, mask, out);
}

void launch_frontend_stats(hipStream_t st, const uint8_t *blk, long ns, int Cin, int td, int fd,
                           unsigned long long *acc, float *off, float *gain)
{
    const int C = Cin / fd;
    hipLaunchKernelGGL(k_frontend_zero, dim3((unsigned)((2L * C + 255) / 256)), dim3(256), 0, st, acc, 2L * C);
    dim3 grid((unsigned)((ns + 255) / 256), (unsigned)((C + 63) / 64));
    hipLaunchKernelGGL(k_frontend_sums, grid, dim3(256), 0, st, blk, ns, Cin, td, fd, acc);
    hipLaunchKernelGGL(k_frontend_fin, dim3((unsigned)((C + 255) / 256)), dim3(256), 0, st,
                       (const unsigned long long *)acc, C, ns, off, gain);
}

__global__ void k_tick(long *clock, long ndump) { *clock += ndump; }

void launch_tick(hipStream_t st, long *clock, long ndump, LaunchSink *sink)
{
    kemit(sink, st, k_tick, dim3(1), dim3(1), 0, clock, ndump);
}

// The block-dependent values of a graph pipeline, from the beam's clock.
__device__ inline void stage1_clock(Stage1Args &a)
{
    if (!a.clock) return;
    const long c = *a.clock + a.ndump;
    a.tau_begin = c - a.ndump - a.D1;
    a.tau_end = c - a.D1;
    a.tau_base = a.tau_begin - lmod(a.tau_begin, 4);
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
// Time tiles start on multiples of 4 and each channel's LDS row starts at a
// multiple of 4 samples (d1min rounded down), so the byte offset of a
// (channel, coarse DM) inside its row is fixed at plan time: the table d1sel
// holds its perm selector and word-aligned byte offset, and the inner loop
// spends one scalar add per (channel, coarse DM) instead of recomputing them.
template <int NCC, int Q>
__global__ __launch_bounds__(256) void k_stage1_u8(Stage1Args a)
{
    extern __shared__ uint32_t lds_u[];
    stage1_clock(a);
    // logical order: coarse-DM chunk fastest, then time tile, then sub-band
    const unsigned lid = xcd_logical(blockIdx.x, gridDim.x);
    const int kc = (int)(lid % a.nkc), tz = (int)((lid / a.nkc) % a.ntz), s = (int)(lid / a.nkc / a.ntz);
    const long tau0 = a.tau_base + (long)tz * (1024 * Q);
    const int k0 = kc * NCC;
    const int nk = min(NCC, a.ndmc - k0);
    const int tid = threadIdx.x;
    const uint8_t *ring = (const uint8_t *)a.ring;
    const int *d1min = a.d1min + (long)kc * a.C;
    // ring slots of tau0 (input and sub-band rings): one 64-bit modulo per
    // workgroup, every later slot is a bounded offset from these
    const long in_base = lmod(tau0, a.Lin);
    const long sub_base = lmod(tau0, a.Lsub);
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
            long start = in_base + ldc(d1min, j);
            if (start >= a.Lin) start -= a.Lin;
            const uint32_t *src = (const uint32_t *)(ring + (long)j * a.pitch) + (start >> 2);
            // RS is a multiple of 4 words; the ring's mirror covers the source
            lds_fill16(src, lds_u + jj * a.RS, a.RS, tid >> 6, 4, tid & 63);
        }
        lds_fill_done();
        // channels two at a time: one v_add3_u32 adds both channels' 16-bit
        // pairs to an accumulator (integer sums, so the order is free); an
        // odd last channel goes alone
        auto chans = [&](auto two_c, int jj) {
            constexpr int NJ = decltype(two_c)::value ? 2 : 1;
            const char *rows[NJ];
            long dj[NJ];
#pragma unroll
            for (int c = 0; c < NJ; c++) {
                rows[c] = (const char *)(lds_u + (jj + c) * a.RS + tid);
                // (selector, byte offset) pairs [channel][coarse DM], padded to whole chunks of NCC
                dj[c] = 2 * ((long)(s * a.cps + jc + jj + c) * a.d1pitch + k0);
            }
            constexpr int KB = NCC < 4 ? NCC : 4; // coarse DMs whose LDS reads are in flight together
#pragma unroll
            for (int kb = 0; kb < NCC; kb += KB) {
                uint32_t w[NJ][KB][Q][2], sl[NJ][KB];
#pragma unroll
                for (int c = 0; c < NJ; c++)
#pragma unroll
                    for (int u = 0; u < KB; u++) {
                        sl[c][u] = (uint32_t)ldc(a.d1sel, dj[c] + 2 * (kb + u));
                        const uint32_t *p = (const uint32_t *)(rows[c] + ldc(a.d1sel, dj[c] + 2 * (kb + u) + 1));
#pragma unroll
                        for (int q = 0; q < Q; q++) {
                            w[c][u][q][0] = p[q * 256];
                            w[c][u][q][1] = p[q * 256 + 1];
                        }
                    }
#pragma unroll
                for (int u = 0; u < KB; u++)
#pragma unroll
                    for (int q = 0; q < Q; q++) {
                        uint32_t l = 0u, h = 0u;
#pragma unroll
                        for (int c = 0; c < NJ; c++) {
                            l += __builtin_amdgcn_perm(w[c][u][q][1], w[c][u][q][0], sl[c][u]);
                            h += __builtin_amdgcn_perm(w[c][u][q][1], w[c][u][q][0], sl[c][u] + 0x00010001u);
                        }
                        lo[kb + u][q] += l;
                        hi[kb + u][q] += h;
                    }
            }
        };
        int jj = 0;
        for (; jj + 1 < nch; jj += 2) chans(std::true_type{}, jj);
        if (jj < nch) chans(std::false_type{}, jj);
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
            long slot = sub_base + 4L * (tid + 256 * q);
            if (slot >= a.Lsub) slot -= a.Lsub;
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
// FE: the ring holds 8-bit samples y (the front end's td = fd = 1 input) and
// the LDS rows their values (y - fe_off[j]) * fe_gain[j]: each pass stages the
// raw rows (RS bytes each, behind the CH f32 rows) and forms the floats once,
// so that HBM and the fills move 1 byte per sample instead of 4.  The rows
// start on multiples of 4 samples (d1min rounded down), so the raw rows are
// whole 32-bit words of the ring.
template <int NCC, int Q, bool FE>
__global__ __launch_bounds__(256) void k_stage1_f32(Stage1Args a)
{
    constexpr int U = 4 * Q;
    extern __shared__ float lds_f[];
    stage1_clock(a);
    const unsigned lid = xcd_logical(blockIdx.x, gridDim.x);
    const int kc = (int)(lid % a.nkc), tz = (int)((lid / a.nkc) % a.ntz), s = (int)(lid / a.nkc / a.ntz);
    const long tau0 = a.tau_base + (long)tz * (1024 * Q);
    const int k0 = kc * NCC;
    const int nk = min(NCC, a.ndmc - k0);
    const int tid = threadIdx.x;
    const float *ring = (const float *)a.ring;
    const int *d1min = a.d1min + (long)kc * a.C;
    const long in_base = lmod(tau0, a.Lin);
    const long sub_base = lmod(tau0, a.Lsub);
    float acc[NCC][U];
#pragma unroll
    for (int k = 0; k < NCC; k++)
#pragma unroll
        for (int u = 0; u < U; u++) acc[k][u] = 0.f;

    const int RW = a.RS >> 2;                            // FE: words per raw row
    uint32_t *raw = (uint32_t *)(lds_f + (long)a.CH * a.RS); // FE: the raw rows
    for (int jc = 0; jc < a.cps; jc += a.CH) {
        const int nch = min(a.CH, a.cps - jc);
        __syncthreads();
        for (int jj = 0; jj < nch; jj++) {
            const int j = s * a.cps + jc + jj;
            long start = in_base + ldc(d1min, j);
            if (start >= a.Lin) start -= a.Lin;
            if (FE) {
                // RS is a multiple of 16: whole 16-byte fills of RS bytes
                const uint32_t *src = (const uint32_t *)((const uint8_t *)a.ring + (long)j * a.pitch + start);
                lds_fill16(src, raw + jj * RW, RW, tid >> 6, 4, tid & 63);
            } else {
                const float *src = ring + (long)j * a.pitch + start;
                // RS is a multiple of 4 floats: 16-byte fills, rows on 16-byte lines
                lds_fill16((const uint32_t *)src, (uint32_t *)(lds_f + jj * a.RS), a.RS, tid >> 6, 4,
                           tid & 63);
            }
        }
        lds_fill_done();
        if (FE) {
            for (int jj = 0; jj < nch; jj++) {
                const int j = s * a.cps + jc + jj;
                const float o = __int_as_float(ldc((const int *)a.fe_off, j));
                const float g = __int_as_float(ldc((const int *)a.fe_gain, j));
                const long t0r = tau0 + ldc(d1min, j); // stream time of the row's first sample
                const uint32_t *rr = raw + jj * RW;
                float *fr = lds_f + jj * a.RS;
                for (int w = tid; w < RW; w += 256) {
                    const uint32_t x = rr[w];
                    f4v v;
                    v.x = ((float)(x & 0xFFu) - o) * g;
                    v.y = ((float)((x >> 8) & 0xFFu) - o) * g;
                    v.z = ((float)((x >> 16) & 0xFFu) - o) * g;
                    v.w = ((float)(x >> 24) - o) * g;
                    if (t0r < 0) { // rows reaching before the stream start read 0 there
                        const long t = t0r + 4L * w;
                        if (t < 0) v.x = 0.f;
                        if (t + 1 < 0) v.y = 0.f;
                        if (t + 2 < 0) v.z = 0.f;
                        if (t + 3 < 0) v.w = 0.f;
                    }
                    *(f4v *)__builtin_assume_aligned(fr + 4 * w, 16) = v;
                }
            }
            __syncthreads();
        }
        for (int jj = 0; jj < nch; jj++) {
            const int j = s * a.cps + jc + jj;
            const int dmn = ldc(d1min, j);
            const float *row = lds_f + jj * a.RS + tid;
            const long dj = (long)j * a.d1pitch + k0;
            // every LDS read of the channel issued before the adds (one
            // wait per channel instead of one per read); each accumulator
            // still takes the channels in ascending order
            float v[NCC][U];
#pragma unroll
            for (int k = 0; k < NCC; k++) {
                const float *p = row + (ldc(a.d1, dj + k) - dmn);
#pragma unroll
                for (int u = 0; u < U; u++) v[k][u] = p[u * 256];
            }
#pragma unroll
            for (int k = 0; k < NCC; k++)
#pragma unroll
                for (int u = 0; u < U; u++) acc[k][u] += v[k][u];
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
                long slot = sub_base + tid + 256L * u;
                if (slot >= a.Lsub) slot -= a.Lsub;
                row[slot] = acc[k][u];
                if (slot < a.Wsub) row[slot + a.Lsub] = acc[k][u];
            }
        }
    }
}

enum { S1_U8, S1_F32, S1_FE };

template <int NCC, int Q>
static int launch_s1(hipStream_t st, const Stage1Args &a, int mode, LaunchSink *sink)
{
    // a graph pipeline's window starts up to 3 samples before its block
    const long span = a.clock ? a.ndump + 3 : a.tau_end - a.tau_base;
    Stage1Args b = a;
    b.nkc = (a.ndmc + NCC - 1) / NCC;
    b.ntz = (int)((span + 1024 * Q - 1) / (1024 * Q));
    dim3 grid((unsigned)b.nkc * b.ntz * a.nsub);
    // FE: the CH f32 rows, then the CH raw rows of RS bytes
    const size_t shm = (size_t)a.CH * a.RS * (mode == S1_FE ? 5 : 4);
    if (mode == S1_U8) {
        (void)hipFuncSetAttribute((const void *)k_stage1_u8<NCC, Q>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        kemit(sink, st, k_stage1_u8<NCC, Q>, grid, dim3(256), shm, b);
        return 0;
    }
    // f32 sums: at most 8 coarse DMs per workgroup
    if constexpr (NCC <= 8) {
        if (mode == S1_FE) {
            if (!a.fe_off || !a.fe_gain || a.RS % 16) return -1;
            (void)hipFuncSetAttribute((const void *)k_stage1_f32<NCC, Q, true>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
            kemit(sink, st, k_stage1_f32<NCC, Q, true>, grid, dim3(256), shm, b);
        } else {
            (void)hipFuncSetAttribute((const void *)k_stage1_f32<NCC, Q, false>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
            kemit(sink, st, k_stage1_f32<NCC, Q, false>, grid, dim3(256), shm, b);
        }
        return 0;
    }
    return -1;
}

#define S1_CASE(n, q)                                                                              \
    if (NCC == n && Q == q) return launch_s1<n, q>(st, a, mode, sink);

static int launch_stage1(hipStream_t st, const Stage1Args &a, int NCC, int Q, int mode, LaunchSink *sink)
{
    S1_CASE(16, 2) S1_CASE(16, 1) S1_CASE(8, 2) S1_CASE(8, 1) S1_CASE(4, 2) S1_CASE(4, 1)
    S1_CASE(2, 2) S1_CASE(2, 1) S1_CASE(1, 2) S1_CASE(1, 1)
    return -1;
}

int launch_stage1_u8(hipStream_t st, const Stage1Args &a, int NCC, int Q, LaunchSink *sink)
{
    return launch_stage1(st, a, NCC, Q, S1_U8, sink);
}
int launch_stage1_f32(hipStream_t st, const Stage1Args &a, int NCC, int Q, LaunchSink *sink)
{
    return launch_stage1(st, a, NCC, Q, S1_F32, sink);
}
int launch_stage1_fe(hipStream_t st, const Stage1Args &a, int NCC, int Q, LaunchSink *sink)
{
    return launch_stage1(st, a, NCC, Q, S1_FE, sink);
}

// ---------------------------------------------------------------- K2
// One workgroup (8 waves): DM block bx of coarse group k (8*DPW fine DMs, DPW
// per wave, interleaved), output samples t0 + [0, 64*R).  Each sub-band row of
// the group is staged in LDS over exactly the times the block's DMs read
// (d2min .. d2max + 64*R), rows packed back to back; the sub-bands are cut
// into passes that fit half the LDS, and the rows of pass p+1 stream in
// (direct global->LDS loads) while pass p is added.  Lanes own consecutive
// samples, so LDS reads are conflict-free and each read feeds one FP32 add.
// Sub-bands are added in ascending order.

// Per-pass lane tables: chunk j holds, lane l, the delay offset of DM slot
// l % DPW for sub-band s0 + j*(64/DPW) + l / DPW, and the LDS row start of
// sub-band s0 + j*(64/DPW) + l.  A pass spans at most 4 chunks.
struct S2Tab {
    int off[4], row[4];
};

template <int DPW>
__device__ __forceinline__ void stage2_tab(const Stage2Args &a, long bs, int s0, int s1, int wave, int lane,
                                           S2Tab &t)
{
    constexpr int SC = 64 / DPW;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int c0 = s0 + j * SC, cn = min(SC, s1 - c0);
        t.off[j] = lane < cn * DPW ? a.d2off[((bs + c0 + lane / DPW) * 8 + wave) * DPW + lane % DPW] : 0;
        t.row[j] = lane < cn ? a.ldsoff[bs + c0 + lane] : 0;
    }
}

// Adds the staged sub-band rows [s0, s1) of one pass into the wave's first
// NA DM slots, taking offsets and row starts from the lane tables with
// v_readlane (no memory wait between sub-bands).  All NA*R LDS reads of a
// sub-band are issued before its adds.
template <int DPW, int NA, int R>
__device__ __forceinline__ void stage2_pass(const float *buf, int lane, int s0, int s1, const S2Tab &t,
                                            float (&acc)[DPW][R])
{
    constexpr int SC = 64 / DPW;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int cn = min(SC, s1 - (s0 + j * SC));
        for (int q = 0; q < cn; q++) {
            const float *row = buf + __builtin_amdgcn_readlane(t.row[j], q) + lane;
            float v[NA][R];
#pragma unroll
            for (int d = 0; d < NA; d++) {
                const float *pd = row + __builtin_amdgcn_readlane(t.off[j], q * DPW + d);
#pragma unroll
                for (int r = 0; r < R; r++) v[d][r] = pd[64 * r];
            }
#pragma unroll
            for (int d = 0; d < NA; d++)
#pragma unroll
                for (int r = 0; r < R; r++) acc[d][r] += v[d][r];
        }
    }
}

template <int DPW, int R>
__global__ __launch_bounds__(512) void k_stage2(Stage2Args a)
{
    extern __shared__ float lds2[]; // two buffers of a.half floats
    if (a.clock) a.w0 = *a.clock + a.ndump - a.maxdelay - a.N;
    // logical order: DM block fastest, then time tile, then coarse group
    const unsigned lid = xcd_logical(blockIdx.x, gridDim.x);
    const int bx = (int)(lid % a.nblk);
    const long t0 = (long)((lid / a.nblk) % a.ntile) * (64 * R);
    const int k = (int)(lid / a.nblk / a.ntile);
    const int blk = k * a.nblk + bx;
    const long gend = min((long)(k + 1) * a.group, a.ndm);
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6); // uniform per wavefront
    const long ibase = (long)k * a.group + (long)bx * 8 * DPW + wave; // DMs ibase + 8*d
    const long bs = (long)blk * a.nsub;
    const long base = lmod(a.w0 + t0, a.Lsub);
    const float *grp = a.sub + (long)k * a.nsub * a.spitch;
    const long pb = (long)blk * a.pstride;
    const int np = ldc(a.npass, blk);
    float acc[DPW][R];
#pragma unroll
    for (int d = 0; d < DPW; d++)
#pragma unroll
        for (int r = 0; r < R; r++) acc[d][r] = 0.f;

    int nact = 0; // live DM slots of this wave (wave-uniform): slot d is DM ibase + 8*d
#pragma unroll
    for (int d = 0; d < DPW; d++) nact += (ibase + 8 * d < gend) ? 1 : 0;

    // rows of pass p into buffer p&1
    auto fill = [&](int p) {
        float *buf = lds2 + (p & 1) * a.half;
        const int s0 = ldc(a.pstart, pb + p), s1 = ldc(a.pstart, pb + p + 1);
        for (int s = s0; s < s1; s++) {
            long start = base + ldc(a.d2min, bs + s);
            if (start >= a.Lsub) start -= a.Lsub;
            const float *src = grp + (long)s * a.spitch + start;
            const int len = ldc(a.d2span, bs + s) + 64 * R; // spread of the block's delays + tile
            // 16 bytes per lane: rows start 16-byte aligned with room for
            // len rounded up to 4, and the ring's mirror covers the source
            lds_fill16((const uint32_t *)src, (uint32_t *)(buf + ldc(a.ldsoff, bs + s)), len, wave, 8, lane);
        }
    };
    S2Tab cur, nxt;
    stage2_tab<DPW>(a, bs, ldc(a.pstart, pb), ldc(a.pstart, pb + 1), wave, lane, cur);
    fill(0);
    for (int p = 0; p < np; p++) {
        // pass p (rows and lane tables) has landed; every wave is done with
        // pass p-1's buffer, which pass p+1 now streams into
        lds_fill_done();
        const int s0 = ldc(a.pstart, pb + p), s1 = ldc(a.pstart, pb + p + 1);
        if (p + 1 < np) {
            stage2_tab<DPW>(a, bs, s1, ldc(a.pstart, pb + p + 2), wave, lane, nxt);
            fill(p + 1);
        }
        const float *buf = lds2 + (p & 1) * a.half;
        // only the wave's live DM slots (a prefix of d) read and add
        if (nact == DPW) {
            stage2_pass<DPW, DPW, R>(buf, lane, s0, s1, cur, acc);
        } else if (DPW > 1 && nact == DPW - 1) {
            stage2_pass<DPW, (DPW > 1 ? DPW - 1 : 1), R>(buf, lane, s0, s1, cur, acc);
        } else if (nact > 0) { // ragged group ends
            for (int s = s0; s < s1; s++) {
                const float *row = buf + ldc(a.ldsoff, bs + s) + lane;
#pragma unroll
                for (int d = 0; d < DPW; d++)
                    if (d < nact) {
                        const float *q = row + ldc(a.d2off, ((bs + s) * 8 + wave) * DPW + d);
#pragma unroll
                        for (int r = 0; r < R; r++) acc[d][r] += q[64 * r];
                    }
            }
        }
        cur = nxt;
    }
    const int tile = (int)(t0 / (64 * R));
#pragma unroll
    for (int d = 0; d < DPW; d++) {
        const long i = ibase + 8 * d;
        if (i >= gend) continue;
        float *out = a.tim + i * a.N;
        double s1 = 0., s2 = 0.;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const long t = t0 + lane + 64 * r;
            if (t < a.N) {
                out[t] = acc[d][r];
                s1 += (double)acc[d][r];
                s2 += (double)acc[d][r] * (double)acc[d][r];
            }
        }
        // this tile's sum and sum of squares of the series (FP64, fixed
        // butterfly order): the boxcar's mean and variance come from these
        s1 = wave_sum_d(s1);
        s2 = wave_sum_d(s2);
        if (lane == 0 && a.tstat) {
            a.tstat[(i * a.ntile + tile) * 2] = s1;
            a.tstat[(i * a.ntile + tile) * 2 + 1] = s2;
        }
    }
}

template <int DPW, int R>
static int launch_s2(hipStream_t st, const Stage2Args &a, LaunchSink *sink)
{
    Stage2Args b = a;
    b.ntile = (int)((a.N + 64 * R - 1) / (64 * R));
    dim3 grid((unsigned)a.nblk * b.ntile * a.ndmc);
    const size_t shm = (size_t)2 * a.half * 4;
    (void)hipFuncSetAttribute((const void *)k_stage2<DPW, R>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    kemit(sink, st, k_stage2<DPW, R>, grid, dim3(512), shm, b);
    return 0;
}

#define S2_CASE(d, r)                                                                              \
    if (DPW == d && R == r) return launch_s2<d, r>(st, a, sink);

// (DPW, R) pairs with DPW*R <= 32 accumulators per lane
const int kStage2Shapes[][2] = {{1, 32}, {1, 16}, {2, 16}, {2, 8}, {4, 8}, {8, 4}};
const int kNumStage2Shapes = 6;

int launch_stage2(hipStream_t st, const Stage2Args &a, int DPW, int R, LaunchSink *sink)
{
    S2_CASE(1, 32) S2_CASE(1, 16) S2_CASE(2, 16) S2_CASE(2, 8) S2_CASE(4, 8) S2_CASE(8, 4)
    return -1;
}

// ---------------------------------------------------------------- K3

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

// Writes one output sample of Boxcar::mxS / mxwn (src/singlepulse/boxcar.cpp:284-285)
// and appends it to the point list when S > thre (src/singlepulse/cluster.cpp:95):
// one atomic per wavefront, lanes placed by their rank among the hits.
// Every lane of the wavefront must call it (it ballots).
__device__ inline void box_emit(const BoxcarArgs &a, long row, int i, float bS, int bw, bool in,
                                int lane)
{
    const float v = bS + 0.f; // the AVX2 blend adds +0.0: -0.0 is stored as +0.0
    if (in && a.store_maps) {
        a.mxS[row * a.N + i] = v;
        a.mxwn[row * a.N + i] = bw;
    }
    const bool hit = in && v > a.thre;
    const unsigned long long m = __ballot(hit);
    if (!m) return;
    const int leader = __ffsll((long long)m) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(a.npts, (unsigned long long)__popcll(m));
    base = __shfl(base, leader, 64);
    if (!hit) return;
    const unsigned long long idx = base + (unsigned long long)__popcll(m & ((1ull << lane) - 1ull));
    if (idx >= (unsigned long long)a.cap) return;
    if (a.pts_ex) {
        tx_point_ex p;
        p.block = a.block;
        p.idm = (int32_t)row;
        p.isamp = (int32_t)i;
        p.S = v;
        p.wn = bw;
        p.reserved = 0;
        a.pts_ex[idx] = p;
    } else {
        tx_point p;
        p.idm = (int32_t)row;
        p.isamp = (int32_t)i;
        p.S = v;
        p.wn = bw;
        a.pts[idx] = p;
    }
}

// Graph pipelines: the block tag from the beam's clock, and false while
// Boxcar::run's gate is closed (src/singlepulse/boxcar.cpp:161).
__device__ inline bool boxcar_clock(BoxcarArgs &a)
{
    if (!a.clock) return true;
    const long c = *a.clock + a.ndump;
    a.block = (int)(c / a.ndump);
    return c >= a.ready_at;
}

// Exact order, the sequential part: the two FP32 passes of
// src/singlepulse/boxcar.cpp over a row, the statistics (:206-213, in the
// order the -O3 -mfma build evaluates them: sequential sums, FMA-contracted
// squares) and the running sum of (x - mean) (:235-240).  Each pass is a
// chain of N dependent adds that no regrouping may split without changing
// the rounding, so one lane runs one row's chains: wave 0 of the workgroup
// runs kXRows rows side by side and does nothing else.  The rows stream
// through LDS in chunks of kXChunk samples, double-buffered: the helper
// waves load the next chunk (direct global->LDS loads; a wave keeps only a
// few of these in flight, so each helper takes a few rows) while wave 0
// runs the current one, and in the second pass also subtract the row's
// mean in place, so that wave 0's second chain is one add per sample.  Wave 0
// stores each row's (mean, var) and the running sum before every kXSeg-th
// sample; from those k_boxcar rebuilds the running sum with one thread per
// segment, bit for bit the sequential result.
constexpr int kXChunk = kXSeg * kXSpc; // 1056 samples, a multiple of 4
constexpr int kXRS = kXChunk + 4;      // LDS row stride in words: 16-byte aligned, the 16 rows on distinct banks
constexpr int kXHelpers = 7;           // helper wave h + 1 takes rows h, h + 7, h + 14
constexpr size_t kXLds = (size_t)2 * kXRows * kXRS * 4;

// Wave 0 loads the next segment right after its chain has used the first
// sample of the current one: that first use waits for every LDS load still
// in flight (the current segment's, issued a segment earlier, so the wait
// costs nothing), and the new loads land while the rest of the chain runs.
// The scheduling barriers keep that order.
__device__ __forceinline__ void xchain_load(float (&r)[kXSeg], const float *p)
{
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < kXSeg; j++) r[j] = p[j];
    __builtin_amdgcn_sched_barrier(0);
}

template <int J0, int J1>
__device__ __forceinline__ void xchain_stats(const float (&r)[kXSeg], float &m, float &v)
{
#pragma unroll
    for (int j = J0; j < J1; j++) {
        m += r[j];
        v = fmaf(r[j], r[j], v);
    }
}

// the second pass: r holds x - mean already
template <int J0, int J1>
__device__ __forceinline__ void xchain_sum(const float (&r)[kXSeg], float &c)
{
#pragma unroll
    for (int j = J0; j < J1; j++) c += r[j];
}

__global__ __launch_bounds__(64 * (1 + kXHelpers)) void k_boxcar_chain(BoxcarArgs a)
{
    if (!boxcar_clock(a)) return;
    extern __shared__ float lds_x[]; // [2][kXRows][kXRS]
    __shared__ float s_mean[kXRows];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int own = wave - 1; // a helper's first row
    const long r0 = (long)blockIdx.x * kXRows;
    const int nr = (int)min((long)kXRows, a.ndm - r0);
    const long N = a.N;
    const int nchunk = (int)((N + kXChunk - 1) / kXChunk);
    const int nq = 2 * nchunk; // step q < nchunk: statistics over chunk q; then the running sum over q - nchunk

    auto chunk0 = [&](int q) { return (long)(q < nchunk ? q : q - nchunk) * kXChunk; };
    auto buffer = [&](int q) { return lds_x + (q & 1) * (kXRows * kXRS); };
    // a helper's rows of chunk q into LDS, then waits for them
    auto fill = [&](int q) {
        const long s0 = chunk0(q);
        const int len = (int)min((long)kXChunk, N - s0);
        // len rounds up to whole 16-byte lines: at most 3 floats past the
        // chunk, inside the row pitch or the series' tail padding
        for (int r = own; r < nr; r += kXHelpers)
            lds_fill16((const uint32_t *)(a.tim + (r0 + r) * N + s0), (uint32_t *)(buffer(q) + r * kXRS), len, 0,
                       1, lane);
        __builtin_amdgcn_s_waitcnt(0);
    };
    // x - mean in place over a helper's rows of chunk q, as 16-byte lines
    // (a ragged last line's words past len are never read), all of a lane's
    // lines in flight at once
    auto subtract = [&](int q) {
        constexpr int F = kXChunk / 4, PR = (F + 63) / 64, NK = PR * ((kXRows + kXHelpers - 1) / kXHelpers);
        const int nf = (int)((min((long)kXChunk, N - chunk0(q)) + 3) >> 2);
        float *x = buffer(q);
        float4 t[NK];
        float mu[NK];
        int off[NK];
#pragma unroll
        for (int k = 0; k < NK; k++) {
            const int r = own + kXHelpers * (k / PR), f = lane + 64 * (k % PR);
            off[k] = r < nr && f < nf ? r * kXRS + 4 * f : -1;
            if (off[k] >= 0) {
                t[k] = *(const float4 *)(x + off[k]);
                mu[k] = s_mean[r];
            }
        }
#pragma unroll
        for (int k = 0; k < NK; k++)
            if (off[k] >= 0) {
                t[k].x = t[k].x - mu[k];
                t[k].y = t[k].y - mu[k];
                t[k].z = t[k].z - mu[k];
                t[k].w = t[k].w - mu[k];
                *(float4 *)(x + off[k]) = t[k];
            }
    };

    const bool live = wave == 0 && lane < nr;
    const long row = r0 + (live ? lane : 0);
    float *ck = a.xck + row * a.xpitch;
    float m = 0.f, v = 0.f, c = 0.f;
    if (wave > 0) fill(0);
    __syncthreads();
    for (int q = 0; q < nq; q++) {
        if (wave == 0) {
            const bool stats = q < nchunk;
            const long s0 = chunk0(q);
            const int len = (int)min((long)kXChunk, N - s0);
            const float *x = buffer(q) + (live ? lane : 0) * kXRS;
            const long g0 = s0 / kXSeg; // first segment of the chunk
            if (live) {
                if (len == kXChunk && stats) {
                    // the statistics need no checkpoints: 16-byte reads over
                    // blocks of 32 samples (33 per chunk), same order
                    const float4 *x4 = (const float4 *)x;
                    float4 A4[8], B4[8];
                    auto ld = [&](float4(&r)[8], const float4 *p) {
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int j = 0; j < 8; j++) r[j] = p[j];
                        __builtin_amdgcn_sched_barrier(0);
                    };
                    auto st = [&](float t) {
                        m += t;
                        v = fmaf(t, t, v);
                    };
                    auto rest = [&](const float4(&r)[8]) {
                        st(r[0].y);
                        st(r[0].z);
                        st(r[0].w);
#pragma unroll
                        for (int j = 1; j < 8; j++) {
                            st(r[j].x);
                            st(r[j].y);
                            st(r[j].z);
                            st(r[j].w);
                        }
                    };
                    static_assert(kXChunk == 33 * 32, "33 blocks of 32 samples");
                    ld(A4, x4);
#pragma unroll 1
                    for (int s = 0; s < 32; s += 2) {
                        st(A4[0].x);
                        ld(B4, x4 + (s + 1) * 8);
                        rest(A4);
                        st(B4[0].x);
                        ld(A4, x4 + (s + 2) * 8);
                        rest(B4);
                    }
                    st(A4[0].x);
                    rest(A4);
                } else if (len == kXChunk) {
                    float A[kXSeg], B[kXSeg];
                    xchain_load(A, x);
                    // the last pair reloads segment 0 into A, unused
                    {
#pragma unroll 1
                        for (int s = 0; s < kXSpc; s += 2) {
                            ck[g0 + s] = c;
                            xchain_sum<0, 1>(A, c);
                            xchain_load(B, x + (s + 1) * kXSeg);
                            xchain_sum<1, kXSeg>(A, c);
                            ck[g0 + s + 1] = c;
                            xchain_sum<0, 1>(B, c);
                            xchain_load(A, x + ((s + 2) & (kXSpc - 1)) * kXSeg);
                            xchain_sum<1, kXSeg>(B, c);
                        }
                    }
                } else {
                    // the row's last, partial chunk
                    for (int s = 0; s * kXSeg < len; s++) {
                        if (!stats) ck[g0 + s] = c;
                        const int e = min(len, (s + 1) * kXSeg);
                        for (int i = s * kXSeg; i < e; i++) {
                            const float t = x[i];
                            if (stats) {
                                m += t;
                                v = fmaf(t, t, v);
                            } else {
                                c += t;
                            }
                        }
                    }
                }
            }
            if (q == nchunk - 1 && live) {
                m /= (float)N;
                v /= (float)N;
                v = fmaf(-m, m, v);
                s_mean[lane] = m;
                a.xstat[2 * row] = m;
                a.xstat[2 * row + 1] = v;
            }
        } else if (q + 1 < nq) {
            // chunk q + 1 into the buffer wave 0 released at the last
            // barrier (the first chunk of the second pass waits for the means)
            fill(q + 1);
            if (q + 1 > nchunk) subtract(q + 1);
        }
        __syncthreads();
        if (q + 1 == nchunk) {
            // the first chunk of the second pass: the means are known now
            if (wave > 0) subtract(nchunk);
            __syncthreads();
        }
    }
}

// One workgroup of 1024 threads per DM row.  The row's running sum lives in
// LDS (USE_LDS) or, for rows longer than the LDS of one workgroup, in a
// global scratch row.
template <int R, bool USE_LDS>
__global__ __launch_bounds__(1024) void k_boxcar(BoxcarArgs a)
{
    if (!boxcar_clock(a)) return;
    extern __shared__ float lds_cs[];
    float *cs = USE_LDS ? lds_cs : a.scratch + (long)blockIdx.x * a.N;
    __shared__ double red[16];
    __shared__ float s_temp[256];
    __shared__ int s_wn[256], s_whi[256];
    __shared__ float s_stats[2];
    __shared__ uint32_t s_hist[256];
    __shared__ uint32_t s_sel[2];
    const long row = blockIdx.x;
    const long N = a.N;
    const float *x = a.tim + row * N;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    float mean_f, var_f;
    if (USE_LDS) {
        lds_fill((const uint32_t *)x, (uint32_t *)cs, (int)N, wave, 16, lane);
        lds_fill_done();
    } else {
        for (long i = tid; i < N; i += 1024) cs[i] = x[i];
        __syncthreads();
    }
    if (a.iqr) {
        // src/singlepulse/boxcar.cpp:217-226: Q1, Q2 ascending order
        // statistics at N/4, N/2; Q3 the N/4-th from the top.
        const float Q1 = select_kth(cs, N, N / 4, s_hist, s_sel);
        const float Q2 = select_kth(cs, N, N / 2, s_hist, s_sel);
        const float Q3 = select_kth(cs, N, N - 1 - N / 4, s_hist, s_sel);
        mean_f = Q2;
        const double q = (double)(Q3 - Q1) / 1.349;
        var_f = (float)(q * q);
    } else if (a.exact && a.xstat) {
        mean_f = a.xstat[2 * row]; // from k_boxcar_chain
        var_f = a.xstat[2 * row + 1];
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
    if (a.exact && a.xck) {
        // one thread per kXSeg-sample segment from k_boxcar_chain's running
        // sum before it: the same adds in the same order as one sequential
        // pass.  Threads kXSeg (odd) words apart read distinct LDS banks.
        const long nseg = boxcar_xsegments(N);
        const float *ckr = a.xck + row * a.xpitch;
        for (long j = tid; j < nseg; j += 1024) {
            float c = ckr[j];
            float *p = cs + j * kXSeg;
            if ((j + 1) * kXSeg <= N) {
                float r[kXSeg];
#pragma unroll
                for (int k = 0; k < kXSeg; k++) r[k] = p[k];
#pragma unroll
                for (int k = 0; k < kXSeg; k++) {
                    c += r[k] - mean_f;
                    r[k] = c;
                }
#pragma unroll
                for (int k = 0; k < kXSeg; k++) p[k] = r[k];
            } else {
                for (long i = j * kXSeg; i < N; i++) {
                    c += cs[i] - mean_f;
                    cs[i] = c;
                }
            }
        }
    } else if (a.exact) {
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
        s_wn[tid] = wn;
        s_whi[tid] = a.whi[tid];
        s_temp[tid] = (float)sqrt(1. / (double)((float)wn * var_f));
    }
    __syncthreads();

    bool finite = true;
    for (int k = 0; k < a.nbox; k++) finite = finite && isfinite(s_temp[k]);
    const bool do_box = var_f > 0.f;
    const int Ni = (int)N;
    const int lo_all = (int)a.lo_all, hi_all = (int)a.hi_all;

    // Interior: chunks of 1024*R samples inside [lo_all, hi_all), where every
    // width's output range covers every sample, so the body needs no checks.
    // Straight-line per width: 2R LDS reads in flight, then R compare-selects.
    const int nfast =
        (do_box && finite && hi_all > lo_all) ? (hi_all - lo_all) / (1024 * R) : 0;
    // width k's (wn, temp) sit in lane k of one VGPR each; v_readlane puts
    // them in SGPRs without an LDS round trip (ladders of up to 64 widths)
    const bool lanes = a.nbox <= 64;
    const int wn_lane = lane < a.nbox ? s_wn[lane] : 0;
    const int tp_lane = lane < a.nbox ? __float_as_int(s_temp[lane]) : 0;
    for (int m = 0; m < nfast; m++) {
        const int c0 = lo_all + m * 1024 * R;
        float bS[R];
        int bw[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            bS[r] = 0.f;
            bw[r] = 0;
        }
        for (int k = 0; k < a.nbox; k++) {
            const int wn = lanes ? __builtin_amdgcn_readlane(wn_lane, k) : s_wn[k];
            const int wl = wn / 2 + 1, wh = (wn - 1) / 2;
            const float temp =
                lanes ? __int_as_float(__builtin_amdgcn_readlane(tp_lane, k)) : s_temp[k];
            const float *ph = cs + (c0 + tid + wh);
            const float *pl = cs + (c0 + tid - wl);
            float S[R];
#pragma unroll
            for (int r = 0; r < R; r++) S[r] = (ph[1024 * r] - pl[1024 * r]) * temp;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const bool ge = S[r] >= bS[r];
                bS[r] = ge ? S[r] : bS[r];
                bw[r] = ge ? wn : bw[r];
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++) box_emit(a, row, c0 + tid + 1024 * r, bS[r], bw[r], true, lane);
    }

    // Edges [0, lo_e) and [hi_e, N): one slice of 1024 samples at a time with
    // each width's range check (src/singlepulse/boxcar.cpp:264-267).
    const int lo_e = nfast ? lo_all : 0;
    const int hi_e = nfast ? lo_all + nfast * 1024 * R : 0;
    for (int pass = 0; pass < 2; pass++) {
        const int a0 = pass == 0 ? 0 : hi_e;
        const int a1 = pass == 0 ? lo_e : Ni;
        for (int c0 = a0; c0 < a1; c0 += 1024) {
            const int i = c0 + tid;
            const bool in = i < a1;
            float bS = 0.f;
            int bw = 0;
            if (do_box) {
                for (int k = 0; k < a.nbox; k++) {
                    const int wn = s_wn[k];
                    const int wl = wn / 2 + 1, wh = (wn - 1) / 2;
                    const float temp = s_temp[k];
                    const bool valid = in && i >= wl && i < s_whi[k];
                    const int ia = min(i + wh, Ni - 1), ib = max(i - wl, 0);
                    const float S = (cs[ia] - cs[ib]) * temp;
                    if (finite) {
                        const bool ge = valid && S >= bS;
                        bS = ge ? S : bS;
                        bw = ge ? wn : bw;
                    } else if (valid) {
                        // the AVX2 blend (boxcar.cpp:273-277), which turns an
                        // unordered compare into 0: (S>=vS ? S : 0) + (S<vS ? vS : 0)
                        const bool ge = S >= bS, lt = S < bS;
                        bS = ge ? S : (lt ? bS : 0.f);
                        bw = ge ? wn : (lt ? bw : 0);
                    }
                }
            }
            box_emit(a, row, i, bS, bw, in, lane);
        }
    }
}

// Exact order with the chain kernel's statistics and checkpoints: one
// 256-thread workgroup per (DM row, segment of outputs).  Segment 0 holds
// outputs [0, lo_all + L), segment k > 0 [lo_all + kL, lo_all + (k+1)L),
// the last one up to N, so the whole-chunk body starts on every segment's
// first sample after the left edge.  A segment's box sums read the running
// sum over [o0 - Hl, o1 + Hr); the workgroup loads that window from the
// checkpoint at or before its start and rebuilds the running sum in LDS,
// one thread per kXSeg-sample segment, bit for bit the sequential values.
// The width loop is k_boxcar's (src/singlepulse/boxcar.cpp:242-281) on
// those values, so the maps are the same bits; small windows let several
// workgroups share a CU.
template <int R>
__global__ __launch_bounds__(256) void k_boxcar_xseg(BoxcarArgs a, int L, int nseg)
{
    if (!boxcar_clock(a)) return;
    extern __shared__ float cs[]; // running sum over [w0, w1)
    __shared__ float s_temp[256];
    __shared__ int s_wn[256], s_whi[256];
    const long row = blockIdx.y;
    const int k = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int Ni = (int)a.N;
    const int lo_all = (int)a.lo_all, hi_all = (int)a.hi_all;
    const int o0 = k == 0 ? 0 : lo_all + k * L;
    const int o1 = k == nseg - 1 ? Ni : min(Ni, lo_all + (k + 1) * L);
    if (o0 >= o1) return; // the whole workgroup
    const int w0 = max(0, o0 - a.Hl) / kXSeg * kXSeg;
    const int w1 = min(Ni, o1 + a.Hr);
    lds_fill((const uint32_t *)(a.tim + row * a.N + w0), (uint32_t *)cs, w1 - w0, wave, 4, lane);
    const float mean_f = a.xstat[2 * row], var_f = a.xstat[2 * row + 1];
    if (tid < a.nbox) {
        const int wn = a.vwn[tid];
        s_wn[tid] = wn;
        s_whi[tid] = a.whi[tid];
        s_temp[tid] = (float)sqrt(1. / (double)((float)wn * var_f));
    }
    lds_fill_done();
    {
        const float *ckr = a.xck + row * a.xpitch;
        for (int j = w0 / kXSeg + tid; j * kXSeg < w1; j += 256) {
            float c = ckr[j];
            float *p = cs + (j * kXSeg - w0);
            if ((j + 1) * kXSeg <= w1) {
                float r[kXSeg];
#pragma unroll
                for (int q = 0; q < kXSeg; q++) r[q] = p[q];
#pragma unroll
                for (int q = 0; q < kXSeg; q++) {
                    c += r[q] - mean_f;
                    r[q] = c;
                }
#pragma unroll
                for (int q = 0; q < kXSeg; q++) p[q] = r[q];
            } else {
                for (int i = j * kXSeg; i < w1; i++) {
                    c += cs[i - w0] - mean_f;
                    cs[i - w0] = c;
                }
            }
        }
    }
    __syncthreads();

    bool finite = true;
    for (int q = 0; q < a.nbox; q++) finite = finite && isfinite(s_temp[q]);
    const bool do_box = var_f > 0.f;
    // whole chunks of 256*R outputs inside [lo_all, hi_all): no range checks
    const int f0 = max(o0, lo_all), f1 = min(o1, hi_all);
    const int nfast = (do_box && finite && f1 > f0) ? (f1 - f0) / (256 * R) : 0;
    const bool lanes = a.nbox <= 64;
    const int wn_lane = lane < a.nbox ? s_wn[lane] : 0;
    const int tp_lane = lane < a.nbox ? __float_as_int(s_temp[lane]) : 0;
    for (int m = 0; m < nfast; m++) {
        const int c0 = f0 + m * 256 * R;
        float bS[R];
        int bw[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            bS[r] = 0.f;
            bw[r] = 0;
        }
        for (int q = 0; q < a.nbox; q++) {
            const int wn = lanes ? __builtin_amdgcn_readlane(wn_lane, q) : s_wn[q];
            const int wl = wn / 2 + 1, wh = (wn - 1) / 2;
            const float temp = lanes ? __int_as_float(__builtin_amdgcn_readlane(tp_lane, q)) : s_temp[q];
            const float *ph = cs + (c0 + tid + wh - w0);
            const float *pl = cs + (c0 + tid - wl - w0);
            float S[R];
#pragma unroll
            for (int r = 0; r < R; r++) S[r] = (ph[256 * r] - pl[256 * r]) * temp;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const bool ge = S[r] >= bS[r];
                bS[r] = ge ? S[r] : bS[r];
                bw[r] = ge ? wn : bw[r];
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++) box_emit(a, row, c0 + tid + 256 * r, bS[r], bw[r], true, lane);
    }

    // the rest of the segment, 256 outputs at a time, with each width's
    // range check (src/singlepulse/boxcar.cpp:264-267)
    const int lo_e = nfast ? f0 : o0;
    const int hi_e = nfast ? f0 + nfast * 256 * R : o0;
    for (int pass = 0; pass < 2; pass++) {
        const int a0 = pass == 0 ? o0 : hi_e;
        const int a1 = pass == 0 ? lo_e : o1;
        for (int c0 = a0; c0 < a1; c0 += 256) {
            const int i = c0 + tid;
            const bool in = i < a1;
            float bS = 0.f;
            int bw = 0;
            if (do_box) {
                for (int q = 0; q < a.nbox; q++) {
                    const int wn = s_wn[q];
                    const int wl = wn / 2 + 1, wh = (wn - 1) / 2;
                    const float temp = s_temp[q];
                    const bool valid = in && i >= wl && i < s_whi[q];
                    const int ia = min(i + wh, Ni - 1), ib = max(i - wl, 0);
                    // lanes past a1 read inside the window too: ia, ib clamp
                    // to [w0, w1)
                    const float S = (cs[min(max(ia, w0), w1 - 1) - w0] - cs[min(max(ib, w0), w1 - 1) - w0]) * temp;
                    if (finite) {
                        const bool ge = valid && S >= bS;
                        bS = ge ? S : bS;
                        bw = ge ? wn : bw;
                    } else if (valid) {
                        // the AVX2 blend (boxcar.cpp:273-277), which turns an
                        // unordered compare into 0: (S>=vS ? S : 0) + (S<vS ? vS : 0)
                        const bool ge = S >= bS, lt = S < bS;
                        bS = ge ? S : (lt ? bS : 0.f);
                        bw = ge ? wn : (lt ? bw : 0);
                    }
                }
            }
            box_emit(a, row, i, bS, bw, in, lane);
        }
    }
}

// K3 segmented form (the default statistics): one 256-thread workgroup per
// (DM row, segment of TS = NC chunks of T = 256*R samples).  A box sum is the
// difference of two running-sum values of the same buffer, so a segment only
// needs the running sum over its own samples plus the widest box's reach,
// [c0 - Hl, c0 + TS + Hr), started from zero: no carry from the rest of the
// row.  The row's mean and variance come from the FP64 per-tile sums stage 2
// wrote (fixed order).  Each thread sums R consecutive samples in FP64, one
// block scan joins them; the width loop then runs chunk by chunk.  Longer
// segments scan the reach (shared with the neighbouring segments) fewer
// times.  The work of one (row, segment starting at sample c0); B is the
// dynamic LDS window [TS + Hl + Hr].  Every thread of the workgroup calls it.
// Eight consecutive floats of LDS from p, where p sits sh floats (the same
// on every lane of the wavefront) past a 16-byte line: three aligned 16-byte
// reads (lanes 32 bytes apart conflict at most 2-way, where eight 4-byte
// reads 32 bytes apart conflict 8-way) and a wave-uniform shift.  Reads up to
// 12 bytes before p (from the 16-byte line holding p) and up to 12 after p + 8.
__device__ __forceinline__ void lds_load8(const float *p, int sh, float (&o)[8])
{
    const f4v q0 = lds_f4(p - sh), q1 = lds_f4(p - sh + 4), q2 = lds_f4(p - sh + 8);
    const float w[12] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
    switch (sh) {
    case 0:
#pragma unroll
        for (int i = 0; i < 8; i++) o[i] = w[i];
        break;
    case 1:
#pragma unroll
        for (int i = 0; i < 8; i++) o[i] = w[i + 1];
        break;
    case 2:
#pragma unroll
        for (int i = 0; i < 8; i++) o[i] = w[i + 2];
        break;
    default:
#pragma unroll
        for (int i = 0; i < 8; i++) o[i] = w[i + 3];
        break;
    }
}

// The points-only path's block tables (largest and smallest running sum of
// each 8-sample block of a segment window of TS outputs and reach H): their
// offset (floats) after the window, and the entry count of each.
__host__ __device__ constexpr int boxcar_seg_xoff(int TS, int H) { return (TS + H + 8 + 3) & ~3; }
__host__ __device__ constexpr int boxcar_seg_nb8(int TS, int H) { return (TS + H) / 8 + 3; }

// kNarrowWidths[k] as a constant expression for unrolled device loops
__device__ constexpr int narrow_width(int k) { return k < 9 ? k + 1 : (k == 9 ? 11 : (k == 10 ? 13 : 16)); }
static_assert(narrow_width(11) == kNarrowWidths[11] && narrow_width(9) == kNarrowWidths[9], "narrow widths");

template <int R, int NC>
__device__ inline void boxcar_seg_body(const BoxcarArgs &a, long row, int c0, float *B)
{
    constexpr int NT = 256, T = NT * R, TS = NC * T;
    __shared__ double red[16];
    __shared__ float s_temp[256];
    __shared__ int s_wn[256], s_whi[256];
    const int N = (int)a.N, Hl = a.Hl, Hr = a.Hr, H = Hl + Hr;
    const float *x = a.tim + row * a.N;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    // the running sums start 0-3 floats in, so that sample c0 + 4m of every
    // chunk sits on a 16-byte line (the narrow widths' window reads)
    float *Bs = B + ((4 - (Hl & 3)) & 3);

    // mean and variance of the row from the per-tile sums and sums of squares
    double sum = 0., sq = 0.;
    {
        const double *ts = a.tstat + row * a.ntile * 2;
        for (int t = lane; t < a.ntile; t += 64) {
            sum += ts[2 * t];
            sq += ts[2 * t + 1];
        }
        sum = wave_sum_d(sum);
        sq = wave_sum_d(sq);
    }
    const double mean_d = sum / (double)N;
    const float mean_f = (float)mean_d;
    const float var_f = (float)(sq / (double)N - mean_d * mean_d);

    if (tid < a.nbox) {
        const int wn = a.vwn[tid];
        s_wn[tid] = wn;
        s_whi[tid] = a.whi[tid];
        s_temp[tid] = (float)sqrt(1. / (double)((float)wn * var_f));
    }
    __syncthreads();
    // every width's scale finite: one LDS read and a ballot per 64 widths
    bool finite = true;
    for (int k0 = 0; k0 < a.nbox; k0 += 64)
        finite = finite && !__ballot(k0 + lane < a.nbox && !isfinite(s_temp[min(k0 + lane, 255)]));
    const bool do_box = var_f > 0.f;
    const int lo_all = (int)a.lo_all, hi_all = (int)a.hi_all;
    const bool lanes = a.nbox <= 64;
    const int tp_lane = lane < a.nbox ? __float_as_int(s_temp[lane]) : 0;
    const int wn_lane = lane < a.nbox ? s_wn[lane] : 0;
    const int whi_lane = lane < a.nbox ? s_whi[lane] : 0;

    double carry = 0.; // running sum through the last sample written to B (every thread)
    // Extends the running sum over absolute samples [p0, p0+n), n <= kPiece,
    // into Bs[p0 - base ...].  Wave w takes the 1024 samples p0 + 1024*w +
    // [0, 1024): it loads them straight into their LDS positions (16 bytes
    // per lane, whole 1 KB lines), then lane l sums its 16 consecutive
    // samples in FP64 and one DPP scan joins the 64 lanes; the four wave
    // totals join through red.  The sums overwrite the samples in place (each
    // thread reads and writes only its own 16 positions).  p0 - base + Bs
    // offset is a multiple of 4 floats, so every 16-byte access is aligned;
    // a piece's last line may read up to 3 floats past the row end (the
    // series buffer has slack) and write them past the window (the LDS
    // allocation has room).
    constexpr int kPiece = NT * 16;
    auto scan_piece = [&](int p0, int n, int base) {
        const int end = p0 + n;
        const int w0 = p0 + 1024 * wave; // this wave's first sample
        {
            uint32_t *dst = (uint32_t *)(Bs + (w0 - base));
            const uint32_t *src = (const uint32_t *)(x + w0);
#pragma unroll
            for (int m = 0; m < 4; m++)
                if (w0 + 256 * m + 4 * lane < end)
                    __builtin_amdgcn_global_load_lds(src + 256 * m + 4 * lane, dst + 256 * m, 16, 0, 0);
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads(); // the piece is in LDS; red is free again
        const int q = w0 + 16 * lane; // this thread's first sample
        // every sample of the wave in the piece (all but the last wave of
        // the last piece): no per-sample range checks
        const bool whole = w0 + 1024 <= end;
        double pre[16], r = 0.;
        if (whole) {
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const f4v v = lds_f4(Bs + (q + 4 * m - base));
                r += (double)(v.x - mean_f);
                pre[4 * m] = r;
                r += (double)(v.y - mean_f);
                pre[4 * m + 1] = r;
                r += (double)(v.z - mean_f);
                pre[4 * m + 2] = r;
                r += (double)(v.w - mean_f);
                pre[4 * m + 3] = r;
            }
        } else {
            float d[16];
#pragma unroll
            for (int m = 0; m < 4; m++) {
                f4v v = {0.f, 0.f, 0.f, 0.f};
                if (q + 4 * m < end) v = lds_f4(Bs + (q + 4 * m - base));
                d[4 * m] = v.x;
                d[4 * m + 1] = v.y;
                d[4 * m + 2] = v.z;
                d[4 * m + 3] = v.w;
            }
#pragma unroll
            for (int u = 0; u < 16; u++) {
                r += (double)((q + u < end) ? d[u] - mean_f : 0.f);
                pre[u] = r;
            }
        }
        const double incl = wave_scan_incl_d(r);
        const double excl = wave_shr1_d(incl);
        if (lane == 63) red[wave] = incl;
        __syncthreads();
        double wpre = 0., tot = 0.;
#pragma unroll
        for (int w = 0; w < NT / 64; w++) {
            wpre += w < wave ? red[w] : 0.;
            tot += red[w];
        }
        const double off = carry + wpre + excl;
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const int qm = q + 4 * m;
            if (whole || qm + 4 <= end) {
                f4v o;
                o.x = (float)(off + pre[4 * m]);
                o.y = (float)(off + pre[4 * m + 1]);
                o.z = (float)(off + pre[4 * m + 2]);
                o.w = (float)(off + pre[4 * m + 3]);
                *(f4v *)__builtin_assume_aligned(Bs + (qm - base), 16) = o;
            } else {
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (qm + u < end) Bs[qm + u - base] = (float)(off + pre[4 * m + u]);
            }
        }
        carry += tot;
    };

    {
        // running sum of this segment's samples and reach, started at zero
        // up to 3 samples early so that every piece starts 16-byte aligned
        const int base = c0 - Hl; // Bs[j] holds sample base + j
        const int b0 = max(c0 - 4 * ((Hl + 3) / 4), 0), b1 = min(N, c0 + TS + Hr);
        for (int p0 = b0; p0 < b1; p0 += kPiece) scan_piece(p0, min(kPiece, b1 - p0), base);
        __syncthreads();

        // the segment's chunks of T outputs, one after another
        for (int ch = 0; ch < NC; ch++) {
        const int cc = c0 + ch * T; // first output of the chunk
        if (cc >= N) break;
        const float *Bc = Bs + ch * T; // Bc[Hl + j] holds sample cc + j

        float bS[R];
        int bw[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            bS[r] = 0.f;
            bw[r] = 0;
        }
        if (do_box && finite && lanes && a.window && NC == 1 && !a.store_maps) {
            // Points only (no maps), every chunk of the row.  Lane tid takes
            // the 8 consecutive outputs i0 + r, i0 = cc + 8*tid.  Only
            // outputs whose maximum over the widths exceeds thre are
            // emitted, so a width may be skipped wherever its S/N provably
            // stays <= thre: with Hmax >= every running sum of the box ends
            // above and Lmin <= every one below, (Hmax - Lmin) * temp >=
            // (B[i+wh] - B[i-wl]) * temp in FP32 (rounding is monotonic), so
            // a skipped width can neither exceed thre nor equal the maximum of
            // an emitted output.  The emitted points, their S/N and widths are
            // those of the full evaluation.
            //
            // (0) the largest and smallest running sum of each 8-sample block
            // of the window: entry e covers Bs positions [q0 - 8 + 8e, q0 + 8e),
            // q0 = Hl mod 8, so that entry jo = (Hl - q0)/8 + 1 + tid holds this
            // lane's outputs.  Positions outside the scanned range do not count.
            const int q0 = Hl & 7;
            const int nb8 = boxcar_seg_nb8(TS, H);
            float *Tmx = B + boxcar_seg_xoff(TS, H);
            float *Tmn = Tmx + nb8;
            float *Tmx2 = Tmn + nb8; // entries e and e + 1 together (16 samples)
            float *Tmn2 = Tmx2 + nb8;
            {
                const int plo = (c0 - Hl > 0 ? c0 - Hl : 0) - (c0 - Hl);
                const int phi = min(N, c0 + TS + Hr) - (c0 - Hl);
                for (int e = tid; e < nb8; e += NT) {
                    const int p0 = q0 - 8 + 8 * e;
                    float mx = -INFINITY, mn = INFINITY;
                    if (p0 >= plo && p0 + 8 <= phi) {
                        const float4 u = *(const float4 *)(Bs + p0);
                        const float4 v = *(const float4 *)(Bs + p0 + 4);
                        mx = fmaxf(fmaxf(fmaxf(u.x, u.y), fmaxf(u.z, u.w)), fmaxf(fmaxf(v.x, v.y), fmaxf(v.z, v.w)));
                        mn = fminf(fminf(fminf(u.x, u.y), fminf(u.z, u.w)), fminf(fminf(v.x, v.y), fminf(v.z, v.w)));
                    } else {
                        for (int q = max(p0, plo); q < min(p0 + 8, phi); q++) {
                            mx = fmaxf(mx, Bs[q]);
                            mn = fminf(mn, Bs[q]);
                        }
                    }
                    Tmx[e] = mx;
                    Tmn[e] = mn;
                }
                __syncthreads();
                for (int e = tid; e < nb8; e += NT) {
                    const bool last = e + 1 == nb8;
                    Tmx2[e] = last ? Tmx[e] : fmaxf(Tmx[e], Tmx[e + 1]);
                    Tmn2[e] = last ? Tmn[e] : fminf(Tmn[e], Tmn[e + 1]);
                }
            }
            __syncthreads(); // the block tables are complete; no barrier below
            const int i0 = cc + 8 * tid;
            const float *P = Bc + Hl + 8 * tid; // P[r] = running sum at output i0 + r
            float M[R];
#pragma unroll
            for (int r = 0; r < R; r++) M[r] = 0.f;
            // lanes whose outputs every width covers, lanes at the row's
            // edges, lanes past its end (nothing to do)
            const bool full = i0 >= lo_all && i0 + 8 <= hi_all;
            unsigned wk[2] = {0u, 0u}; // per output r, byte r & 3 of wk[r >> 2]: ladder index + 1 (0: none)
            if (full) {
                // (1) the narrow widths from the register window, exactly
                float W[28];
#pragma unroll
                for (int m = 0; m < 7; m++) {
                    const f4v q = lds_f4(P - 12 + 4 * m);
                    W[4 * m] = q.x;
                    W[4 * m + 1] = q.y;
                    W[4 * m + 2] = q.z;
                    W[4 * m + 3] = q.w;
                }
#pragma unroll
                for (int k = 0; k < kNarrow; k += 2) {
                    const int wa = narrow_width(k), wb = narrow_width(k + 1);
                    const float ta = __int_as_float(__builtin_amdgcn_readlane(tp_lane, k));
                    const float tb = __int_as_float(__builtin_amdgcn_readlane(tp_lane, k + 1));
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        const float Sa = (W[12 + r + (wa - 1) / 2] - W[12 + r - (wa / 2 + 1)]) * ta;
                        const float Sb = (W[12 + r + (wb - 1) / 2] - W[12 + r - (wb / 2 + 1)]) * tb;
                        M[r] = fmaxf(fmaxf(M[r], Sa), Sb);
                    }
                }
                // (2) the wider widths: the bound of the lane's 8 outputs, and
                // the exact S/N only where the bound exceeds thre; there the
                // running maximum is kept as Boxcar::match keeps it (S >= vS,
                // the narrow widths first), with the ladder index of the last
                // width that reached it
                // The bound reads one table entry each side: the 8-sample
                // block when the box end is block-aligned, else the 16-sample
                // pair.  A width is evaluated for the whole wavefront when
                // any lane's bound exceeds thre (an evaluation where the
                // bound holds finds S <= thre, which changes neither an
                // emitted maximum nor the width that last reached it).
                const float *T0 = Tmx + (Hl - q0) / 8 + 1 + tid; // this lane's entry of Tmx
                for (int k = kNarrow; k < a.nbox; k++) {
                    const int wn = __builtin_amdgcn_readlane(wn_lane, k);
                    const int wl = wn / 2 + 1, wh = (wn - 1) / 2;
                    const float t = __int_as_float(__builtin_amdgcn_readlane(tp_lane, k));
                    // entries from T0 (uniform offsets): Tmx2 = Tmx + 2 nb8, Tmn = Tmx + nb8
                    const float hmax = T0[(wh >> 3) + ((wh & 7) ? 2 * nb8 : 0)];
                    const float lmin = T0[nb8 - ((wl + 7) >> 3) + ((wl & 7) ? 2 * nb8 : 0)];
                    const float U = (hmax - lmin) * t;
                    if (__ballot(!(U <= a.thre))) {
                        float h8[8], l8[8];
                        lds_load8(P + wh, wh & 3, h8);
                        lds_load8(P - wl, -wl & 3, l8);
#pragma unroll
                        for (int r = 0; r < R; r++) {
                            const float S = (h8[r] - l8[r]) * t;
                            const bool ge = S >= M[r];
                            M[r] = ge ? S : M[r];
                            wk[r >> 2] = ge ? (wk[r >> 2] & ~(0xFFu << (8 * (r & 3)))) | ((unsigned)(k + 1) << (8 * (r & 3)))
                                            : wk[r >> 2];
                        }
                    }
                }
            } else if (i0 < N) {
                // edge lanes: every width where it has an output
                // (src/singlepulse/boxcar.cpp:264-267, the AVX2 range), in
                // ladder order with S >= vS
                for (int k = 0; k < a.nbox; k++) {
                    const int wn = __builtin_amdgcn_readlane(wn_lane, k);
                    const int wl = wn / 2 + 1, wh = (wn - 1) / 2;
                    const int hik = __builtin_amdgcn_readlane(whi_lane, k);
                    const float t = __int_as_float(__builtin_amdgcn_readlane(tp_lane, k));
                    float h8[8], l8[8];
                    lds_load8(P + wh, wh & 3, h8);
                    lds_load8(P - wl, -wl & 3, l8);
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        const int i = i0 + r;
                        const float S = (h8[r] - l8[r]) * t;
                        const bool ge = i >= wl && i < hik && i < N && S >= M[r];
                        M[r] = ge ? S : M[r];
                        wk[r >> 2] = ge ? (wk[r >> 2] & ~(0xFFu << (8 * (r & 3)))) | ((unsigned)(k + 1) << (8 * (r & 3)))
                                        : wk[r >> 2];
                    }
                }
            }
            // (3) emitted outputs: the width whose S/N last reached the
            // maximum.  A wide width (or any width of an edge lane) that did
            // is recorded; otherwise it is a narrow width: the last one whose
            // S/N equals the maximum, recomputed from the register window
            // with the same two operations on the same running sums.
            bool need = false;
#pragma unroll
            for (int r = 0; r < R; r++) {
                bS[r] = M[r];
                const int kr = (int)((wk[r >> 2] >> (8 * (r & 3))) & 0xFFu);
                const bool hit = i0 + r < N && M[r] + 0.f > a.thre;
                bw[r] = kr ? s_wn[kr - 1] : 0;
                need = need || (full && hit && kr == 0);
            }
            if (need) {
                float W[28];
#pragma unroll
                for (int m = 0; m < 7; m++) {
                    const f4v q = lds_f4(P - 12 + 4 * m);
                    W[4 * m] = q.x;
                    W[4 * m + 1] = q.y;
                    W[4 * m + 2] = q.z;
                    W[4 * m + 3] = q.w;
                }
                int nb[R];
#pragma unroll
                for (int r = 0; r < R; r++) nb[r] = 0;
#pragma unroll
                for (int k = 0; k < kNarrow; k++) {
                    const int wn = narrow_width(k);
                    const float t = __int_as_float(__builtin_amdgcn_readlane(tp_lane, k));
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        const float S = (W[12 + r + (wn - 1) / 2] - W[12 + r - (wn / 2 + 1)]) * t;
                        nb[r] = S == M[r] ? wn : nb[r];
                    }
                }
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int kr = (int)((wk[r >> 2] >> (8 * (r & 3))) & 0xFFu);
                    bw[r] = kr ? bw[r] : nb[r];
                }
            }
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int i = i0 + r;
                box_emit(a, row, i, bS[r], bw[r], i < N, lane);
            }
            continue; // the chunk's points are out (register layout)
        } else if (do_box && finite && cc >= lo_all && cc + T <= hi_all && lanes && a.window && NC == 1) {
            // every width's output range covers the whole chunk: no checks.
            // (1) The ladder starts with kNarrowWidths: lane tid takes the 8
            // consecutive outputs cc + 8*tid + r from a register window of
            // running sums [8*tid - 12, 8*tid + 16) (7 aligned 128-bit LDS
            // reads instead of 16 reads per width).
            float W[28];
#pragma unroll
            for (int m = 0; m < 7; m++) {
                const f4v q = lds_f4(Bc + Hl + 8 * tid - 12 + 4 * m);
                W[4 * m] = q.x;
                W[4 * m + 1] = q.y;
                W[4 * m + 2] = q.z;
                W[4 * m + 3] = q.w;
            }
            float nS[R];
            int nW[R];
#pragma unroll
            for (int r = 0; r < R; r++) {
                nS[r] = 0.f;
                nW[r] = 0;
            }
#pragma unroll
            for (int k = 0; k < kNarrow; k++) {
                const int wn = narrow_width(k);
                const int wl = wn / 2 + 1, wh = (wn - 1) / 2;
                const float temp = __int_as_float(__builtin_amdgcn_readlane(tp_lane, k));
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const float S = (W[12 + r + wh] - W[12 + r - wl]) * temp;
                    const bool ge = S >= nS[r];
                    nS[r] = ge ? S : nS[r];
                    nW[r] = ge ? wn : nW[r];
                }
            }
            // (2) the wider widths in the strided layout (outputs cc + tid + 256*r)
            for (int k = kNarrow; k < a.nbox; k++) {
                const int wn = __builtin_amdgcn_readlane(wn_lane, k);
                const int wl = wn / 2 + 1, wh = (wn - 1) / 2;
                const float temp = __int_as_float(__builtin_amdgcn_readlane(tp_lane, k));
                const float *ph = Bc + (Hl + tid + wh);
                const float *pl = Bc + (Hl + tid - wl);
                float S[R];
#pragma unroll
                for (int r = 0; r < R; r++) S[r] = (ph[NT * r] - pl[NT * r]) * temp;
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const bool ge = S[r] >= bS[r];
                    bS[r] = ge ? S[r] : bS[r];
                    bw[r] = ge ? wn : bw[r];
                }
            }
            // (3) the narrow results into the strided layout through LDS (the
            // running sums are no longer read), then merged: the sequential
            // S >= vS scan over narrow then wider widths keeps the wider
            // result when its maximum is larger, or equal and reached by one
            // of them (bw != 0)
            __syncthreads();
            float4 *tp = (float4 *)(B + 8 * tid);
            tp[0] = make_float4(nS[0], nS[1], nS[2], nS[3]);
            tp[1] = make_float4(nS[4], nS[5], nS[6], nS[7]);
            __syncthreads();
#pragma unroll
            for (int r = 0; r < R; r++) nS[r] = B[tid + NT * r];
            __syncthreads();
            tp[0] = make_float4(__int_as_float(nW[0]), __int_as_float(nW[1]), __int_as_float(nW[2]),
                                __int_as_float(nW[3]));
            tp[1] = make_float4(__int_as_float(nW[4]), __int_as_float(nW[5]), __int_as_float(nW[6]),
                                __int_as_float(nW[7]));
            __syncthreads();
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int w2 = __float_as_int(B[tid + NT * r]);
                const bool wide = bS[r] > nS[r] || (bS[r] == nS[r] && bw[r] != 0);
                bS[r] = wide ? bS[r] : nS[r];
                bw[r] = wide ? bw[r] : w2;
            }
        } else if (do_box && finite && cc >= lo_all && cc + T <= hi_all && lanes) {
            // every width's output range covers the whole chunk: no checks.
            // The width and its scale come from lane k of wn_lane and
            // tp_lane (no memory wait per width); S >= bS keeps the wider
            // width on ties.  (Two widths per step, for more LDS reads in
            // flight, measured slower: the loop is bound by LDS bandwidth.)
            for (int k = 0; k < a.nbox; k++) {
                const int wn = __builtin_amdgcn_readlane(wn_lane, k);
                const int wl = wn / 2 + 1, wh = (wn - 1) / 2;
                const float temp = __int_as_float(__builtin_amdgcn_readlane(tp_lane, k));
                const float *ph = Bc + (Hl + tid + wh);
                const float *pl = Bc + (Hl + tid - wl);
                float S[R];
#pragma unroll
                for (int r = 0; r < R; r++) S[r] = (ph[NT * r] - pl[NT * r]) * temp;
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const bool ge = S[r] >= bS[r];
                    bS[r] = ge ? S[r] : bS[r];
                    bw[r] = ge ? wn : bw[r];
                }
            }
        } else if (do_box && finite && cc >= lo_all && cc + T <= hi_all) {
            for (int k = 0; k < a.nbox; k++) {
                const int wn = s_wn[k];
                const int wl = wn / 2 + 1, wh = (wn - 1) / 2;
                const float temp = s_temp[k];
                const float *ph = Bc + (Hl + tid + wh);
                const float *pl = Bc + (Hl + tid - wl);
                float S[R];
#pragma unroll
                for (int r = 0; r < R; r++) S[r] = (ph[NT * r] - pl[NT * r]) * temp;
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const bool ge = S[r] >= bS[r];
                    bS[r] = ge ? S[r] : bS[r];
                    bw[r] = ge ? wn : bw[r];
                }
            }
        } else if (do_box) {
            // range checks (src/singlepulse/boxcar.cpp:264-267) and the AVX2
            // blend (:273-277): (S>=vS ? S : 0) + (S<vS ? vS : 0)
            const int L = TS + H;
            for (int k = 0; k < a.nbox; k++) {
                const int wn = s_wn[k], hik = s_whi[k];
                const int wl = wn / 2 + 1, wh = (wn - 1) / 2;
                const float temp = s_temp[k];
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int i = cc + tid + NT * r;
                    const bool valid = i < N && i >= wl && i < hik;
                    const int ia = min(max(i + wh - base, 0), L - 1);
                    const int ib = min(max(i - wl - base, 0), L - 1);
                    const float S = (Bs[ia] - Bs[ib]) * temp;
                    const bool ge = S >= bS[r], lt = S < bS[r];
                    const float nb = ge ? S : (lt ? bS[r] : 0.f);
                    const int nw = ge ? wn : (lt ? bw[r] : 0);
                    bS[r] = valid ? nb : bS[r];
                    bw[r] = valid ? nw : bw[r];
                }
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int i = cc + tid + NT * r;
            box_emit(a, row, i, bS[r], bw[r], i < N, lane);
        }
        }
    }
}

template <int R, int NC>
__global__ __launch_bounds__(256) void k_boxcar_seg(BoxcarArgs a)
{
    if (!boxcar_clock(a)) return;
    constexpr int TS = 256 * R * NC;
    extern __shared__ __align__(16) float B[]; // [TS + Hl + Hr + slack]
    const int nseg = ((int)a.N + TS - 1) / TS;
    // logical order: segment fastest, so a row's segments share an L2
    const unsigned lid = xcd_logical(blockIdx.x, gridDim.x);
    boxcar_seg_body<R, NC>(a, lid / nseg, (int)(lid % nseg) * TS, B);
}

// The segmented boxcar of many rows (of many beams) in one persistent
// launch.  Job j owns items [first[j], first[j+1]) = its rows x segments, row
// by row; every workgroup claims one item at a time from the device counter
// *next (zero before the launch) and leaves once the counter has passed the
// last item, so the launch ends after total + grid claims.
template <int R, int NC>
__global__ __launch_bounds__(256) void k_boxcar_queue(BoxcarQueue q)
{
    constexpr int TS = 256 * R * NC;
    extern __shared__ __align__(16) float B[];
    __shared__ unsigned s_item;
    for (;;) {
        __syncthreads(); // the previous item is done with the shared arrays
        if (threadIdx.x == 0) s_item = atomicAdd(q.next, 1u);
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane((int)s_item);
        if ((long)item >= q.total) return;
        int lo = 0, hi = q.njobs - 1; // the last job whose first item <= item
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (q.first[mid] <= item) lo = mid;
            else hi = mid - 1;
        }
        // the job's arguments by scalar loads (uniform, read-only here)
        BoxcarArgs a;
        {
            static_assert(sizeof(BoxcarArgs) % 4 == 0, "BoxcarArgs in dwords");
            const int *src = (const int *)(q.jobs + __builtin_amdgcn_readfirstlane(lo));
            int *dst = (int *)&a;
#pragma unroll
            for (int k = 0; k < (int)(sizeof(BoxcarArgs) / 4); k++) dst[k] = ldc(src, k);
        }
        const int local = item - (int)q.first[lo];
        const int nseg = ((int)a.N + TS - 1) / TS;
        boxcar_seg_body<R, NC>(a, local / nseg, (local % nseg) * TS, B);
    }
}

bool boxcar_fits_lds(long N, int max_lds_bytes) { return N * 4 + 4096 <= (long)max_lds_bytes; }

// segmented boxcar: 8 outputs per thread, segments of kSegChunks chunks of
// 2048.  One chunk measured fastest on MI355X: two chunks scan less reach but
// their larger LDS window cuts residency from 8 to 6 waves per SIMD.
constexpr int kSegR = 8, kSegChunks = 1, kSegLen = 256 * kSegR * kSegChunks;

size_t boxcar_seg_lds(const BoxcarArgs &a)
{
    if (a.exact || a.iqr || !a.tstat) return 0;
    // the window (+ the 16-byte alignment shift and the 3 floats a piece's
    // last 16-byte line may land past it); without maps also the block
    // tables of the points-only path
    if (a.store_maps) return (size_t)(kSegLen + a.Hl + a.Hr + 8) * 4;
    return (size_t)(boxcar_seg_xoff(kSegLen, a.Hl + a.Hr) + 4 * boxcar_seg_nb8(kSegLen, a.Hl + a.Hr)) * 4;
}

long boxcar_seg_items(const BoxcarArgs &a) { return a.ndm * ((a.N + kSegLen - 1) / kSegLen); }

int boxcar_queue_grid(size_t shm, int ncu)
{
    int nb = 0;
    if (hipFuncSetAttribute((const void *)k_boxcar_queue<kSegR, kSegChunks>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm) != hipSuccess)
        return -1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void *)k_boxcar_queue<kSegR, kSegChunks>, 256,
                                                     shm) !=
            hipSuccess ||
        nb < 1)
        return -1;
    return nb * ncu;
}

void launch_boxcar_queue(hipStream_t st, const BoxcarQueue &q, size_t shm, int grid)
{
    hipLaunchKernelGGL((k_boxcar_queue<kSegR, kSegChunks>), dim3((unsigned)grid), dim3(256), shm, st, q);
}

int launch_boxcar(hipStream_t st, const BoxcarArgs &a, int max_lds_bytes, LaunchSink *sink)
{
    constexpr int R = 8;
    if (a.nbox > 256) return -1;
    if (!a.exact && !a.iqr && a.tstat) {
        // segmented form: running-sum window of one segment plus the widest box's reach
        const size_t shm = boxcar_seg_lds(a);
        if ((long)shm + 4096 > max_lds_bytes) return -1;
        auto k = k_boxcar_seg<kSegR, kSegChunks>;
        (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        kemit(sink, st, k, dim3((unsigned)boxcar_seg_items(a)), dim3(256), shm, a);
        return 0;
    }
    if (a.xck) {
        if ((long)kXLds > max_lds_bytes) return -1;
        (void)hipFuncSetAttribute((const void *)k_boxcar_chain, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)kXLds);
        kemit(sink, st, k_boxcar_chain, dim3((unsigned)((a.ndm + kXRows - 1) / kXRows)),
              dim3(64 * (1 + kXHelpers)), kXLds, a);
        // the width loop over segments of kXOut outputs (a multiple of the
        // 256*R chunk); a window holds the segment and the widest box's
        // reach.  Ladders too wide for such a window take the whole-row
        // kernel below, which rebuilds from the same checkpoints.
        constexpr int kXR = 8, kXOut = 4096;
        const long rest = a.N - a.lo_all;
        const long nseg = rest > 0 ? (rest + kXOut - 1) / kXOut : 1;
        const size_t shm = (size_t)(kXOut + a.Hl + a.Hr + kXSeg + 4) * 4;
        if ((long)shm + 8192 <= max_lds_bytes && nseg <= 65535 && a.ndm <= 65535) { // grid y: the rows
            auto kx = k_boxcar_xseg<kXR>;
            (void)hipFuncSetAttribute((const void *)kx, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
            kemit(sink, st, kx, dim3((unsigned)nseg, (unsigned)a.ndm), dim3(256), shm, a, kXOut, (int)nseg);
            return 0;
        }
    }
    if (boxcar_fits_lds(a.N, max_lds_bytes)) {
        const size_t shm = (size_t)a.N * 4;
        (void)hipFuncSetAttribute((const void *)k_boxcar<R, true>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        kemit(sink, st, k_boxcar<R, true>, dim3((unsigned)a.ndm), dim3(1024), shm, a);
    } else {
        if (!a.scratch) return -1;
        kemit(sink, st, k_boxcar<R, false>, dim3((unsigned)a.ndm), dim3(1024), 0, a);
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

// --savetim values (declared rule of the dump): y = x * scale + mean in float
// with two roundings; 8 and 16 bits round half away from zero and clamp to
// the signed range.  Rows of n values from tim[row][n0 ...]; 4 outputs per
// thread, every store within the row.
template <typename T>
__global__ __launch_bounds__(256) void k_dump(const float *__restrict__ tim, long N, long n0, long n, float scale,
                                              float mean, T *__restrict__ out)
{
#pragma clang fp contract(off)
    const long row = blockIdx.y;
    const long k0 = ((long)blockIdx.x * 256 + threadIdx.x) * 4;
    const float *x = tim + row * N + n0;
    T *o = out + row * n;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const long k = k0 + u;
        if (k >= n) break;
        const float y = x[k] * scale + mean;
        if constexpr (sizeof(T) == 4) {
            o[k] = y;
        } else {
            constexpr float lo = sizeof(T) == 1 ? -128.f : -32768.f;
            constexpr float hi = sizeof(T) == 1 ? 127.f : 32767.f;
            const float r = fminf(fmaxf(roundf(y), lo), hi);
            o[k] = (T)(int)r;
        }
    }
}

void launch_dump(hipStream_t st, const float *tim, long N, long n0, long n, int ndm, float scale, float mean,
                 int nbits, void *out)
{
    if (n <= 0 || ndm <= 0) return;
    dim3 grid((unsigned)((n + 1023) / 1024), (unsigned)ndm);
    if (nbits == 8)
        hipLaunchKernelGGL(k_dump<int8_t>, grid, dim3(256), 0, st, tim, N, n0, n, scale, mean, (int8_t *)out);
    else if (nbits == 16)
        hipLaunchKernelGGL(k_dump<int16_t>, grid, dim3(256), 0, st, tim, N, n0, n, scale, mean, (int16_t *)out);
    else
        hipLaunchKernelGGL(k_dump<float>, grid, dim3(256), 0, st, tim, N, n0, n, scale, mean, (float *)out);
}

} // namespace txgpu
