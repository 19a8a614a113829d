// kernels.h -- gfx950 kernels of the search hot path and their launch helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/txgpu.h"
#include "launch_sink.h"

namespace txgpu {

// Floor modulo for ring slots (m > 0).
__host__ __device__ inline long lmod(long a, long m)
{
    long r = a % m;
    return r < 0 ? r + m : r;
}


// K0: time-major block [ns][C] -> channel-major ring rows [C][pitch] at slot
// lmod(t0 + t, L), mirrored to slot + L when slot < W.  Masked channels get
// the zero-signal value.
// clock (graph pipelines): t0 is read from *clock, the beam's counter
// before the block, instead of the argument.
void launch_transpose_u8(hipStream_t st, const uint8_t *blk, long ns, int C, const uint8_t *mask,
                         uint8_t zp, uint8_t *ring, long pitch, long L, long W, long t0,
                         const long *clock = nullptr, LaunchSink *sink = nullptr);
void launch_transpose_f32(hipStream_t st, const float *blk, long ns, int C, const uint8_t *mask,
                          float *ring, long pitch, long L, long W, long t0,
                          const long *clock = nullptr, LaunchSink *sink = nullptr);

// K0 with the front end on the device (8-bit ingest): blk is a raw 8-bit
// block, time-major [ns*td][Cin]; output channel c < Cin/fd, sample t < ns
// gets y = the sum of the td x fd neighbours blk[(t*td+ii)*Cin + c*fd+jj]
// (an exact integer), then (y - off[c]) * gain[c] in float, masked channels
// 0, written into the f32 ring as launch_transpose_f32 writes it.
void launch_frontend_u8(hipStream_t st, const uint8_t *blk, long ns, int Cin, int td, int fd,
                        const float *off, const float *gain, const uint8_t *mask, float *ring,
                        long pitch, long L, long W, long t0, const long *clock = nullptr,
                        LaunchSink *sink = nullptr);
// The same values into a time-major [ns][Cin/fd] block (the DataBuffer
// layout, for RFI filters before the push).
// Zeroes the masked channels (mask[c] == 0) of a time-major f32 block [ns][C].
void launch_mask_tm(hipStream_t st, float *blk, long ns, int C, const uint8_t *mask);
void launch_frontend_tm(hipStream_t st, const uint8_t *blk, long ns, int Cin, int td, int fd,
                        const float *off, const float *gain, const uint8_t *mask, float *out);
// Per output channel of a raw block: exact sums of y and y^2 over ns
// samples (acc: [Cin/fd][2] unsigned 64-bit, zeroed here), then off = mean
// and gain = 1/sd (1 when the variance is not positive).
void launch_frontend_stats(hipStream_t st, const uint8_t *blk, long ns, int Cin, int td, int fd,
                           unsigned long long *acc, float *off, float *gain);

// Graph pipelines: the beam's device clock (its counter before the block)
// gives each kernel its block-dependent values, and this one-thread kernel
// advances it by ndump after the block's last kernel.
void launch_tick(hipStream_t st, long *clock, long ndump, LaunchSink *sink);

struct Stage1Args {
    const void *ring; // u8 or f32 channel-major rows
    long pitch;       // elements per ring row
    long Lin;
    const int *d1;    // [C][d1pitch]: delays per channel, coarse DMs padded to whole chunks
    int d1pitch;      // nkchunks * NCC
    const int *d1min; // [nkchunks][C] (u8: rounded down to a multiple of 4)
    const int *d1sel; // u8: [C][d1pitch][2] perm selector and byte offset of the delay in its LDS row
    int C, cps, nsub, ndmc;
    int zp_cps;       // cps * zero_point (u8 only)
    long tau_begin, tau_end, tau_base;
    int RS;           // LDS row stride in 32-bit words
    int CH;           // channels per LDS pass
    int nkc, ntz;     // coarse-DM chunks and time tiles (set by the launcher)
    float *sub;       // [ndmc][nsub][spitch]
    long spitch, Lsub, Wsub;
    // graph pipelines: c = *clock + ndump replaces the host's counter, and
    // tau_begin = c - ndump - D1, tau_end = c - D1 are formed in the kernel
    const long *clock;
    long ndump, D1;
    // front end in the load (launch_stage1_fe): the ring holds the raw 8-bit
    // samples y of an f32 beam and channel j takes (y - fe_off[j]) * fe_gain[j]
    // (0 and 0 for a masked channel); samples before the stream start are 0
    const float *fe_off, *fe_gain;
};
// NCC coarse DMs per workgroup, Q quads (4 samples) per thread.
int launch_stage1_u8(hipStream_t st, const Stage1Args &a, int NCC, int Q, LaunchSink *sink = nullptr);
int launch_stage1_f32(hipStream_t st, const Stage1Args &a, int NCC, int Q, LaunchSink *sink = nullptr);
// The f32 stage 1 over an 8-bit ring: each pass stages CH raw rows of RS
// bytes, forms their f32 values in LDS, then sums them as launch_stage1_f32.
int launch_stage1_fe(hipStream_t st, const Stage1Args &a, int NCC, int Q, LaunchSink *sink = nullptr);

struct Stage2Args {
    const float *sub;
    long spitch, Lsub;
    const int *d2;    // [ndm][nsub]
    const int *d2min;  // [ngroups*nblk][nsub] lowest delay of each DM block
    const int *d2span; // [ngroups*nblk][nsub] highest minus lowest delay of each DM block
    const int *d2off;  // [ngroups*nblk][nsub][8 waves][DPW] delay - d2min of each DM slot
    const int *ldsoff; // [ngroups*nblk][nsub] LDS offset (floats) of each sub-band row in its pass
    const int *npass;  // [ngroups*nblk] passes of each DM block
    const int *pstart; // [ngroups*nblk][pstride] first sub-band of each pass (+ end)
    int pstride;
    int half;          // floats per LDS buffer (two buffers)
    int nsub;
    long ndm;
    int group;        // fine DMs per coarse DM
    int nblk;         // DM blocks per group
    int ntile;        // time tiles (set by the launcher)
    int ndmc;
    long w0, N;
    float *tim;       // [ndm][N]
    double *tstat;    // [ndm][ntile][2]: per-tile sum and sum of squares of tim (optional)
    const long *clock; // graph pipelines: w0 = *clock + ndump - maxdelay - N
    long ndump, maxdelay;
};
// Workgroups of 8 waves; DPW fine DMs per wave, R samples per lane.
extern const int kStage2Shapes[][2];
extern const int kNumStage2Shapes;
int launch_stage2(hipStream_t st, const Stage2Args &a, int DPW, int R, LaunchSink *sink = nullptr);

struct BoxcarArgs {
    const float *tim;
    long N;
    long ndm;
    const int *vwn; // [nbox]
    const int *whi; // [nbox] end of each width's output range (AVX2 rule)
    int nbox;
    long lo_all, hi_all;
    float thre;
    float *mxS;
    int *mxwn;
    int store_maps;
    tx_point *pts;
    unsigned long long *npts;
    long cap;
    int exact;      // 1: sequential FP32 statistics and prefix sum, the reference's order
    int iqr;
    float *scratch; // [ndm][N] running sums when a row does not fit in LDS
    tx_point_ex *pts_ex; // session list (replaces pts when set)
    int block;           // block tag of session points
    int Hl, Hr;          // reach of the widest box left (wl) and right (wh) of its sample
    const double *tstat; // [ndm][ntile][2] per-tile sums of tim from stage 2
    int ntile;
    // graph pipelines: c = *clock + ndump; nothing is done while c <
    // ready_at (Boxcar::run's gate) and block = c / ndump
    const long *clock;
    long ndump, ready_at;
    // 1: the ladder starts with kNarrowWidths and the reach allows the
    // register window (segmented form)
    int window;
    // exact order without IQR: k_boxcar_chain's per-row (mean, var) and the
    // running sum before every kXSeg-th sample ([ndm][xpitch]); null: the
    // boxcar runs the sequential passes itself
    float *xstat, *xck;
    long xpitch;
};
// Exact-order chains: rows per workgroup, samples per checkpoint segment
// (odd, so that threads kXSeg samples apart hit distinct LDS banks), and
// segments per LDS chunk.
constexpr int kXRows = 16, kXSeg = 33, kXSpc = 32;
constexpr long boxcar_xsegments(long N) { return (N + kXSeg - 1) / kXSeg; }
// The start of every width ladder from one sample with snrloss 0.1
// (src/singlepulse/singlepulse.cpp:242-254): the segmented boxcar takes
// these from a register window of each lane's 8 consecutive outputs.
constexpr int kNarrow = 12;
constexpr int kNarrowWidths[kNarrow] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 13, 16};
bool boxcar_fits_lds(long N, int max_lds_bytes);
int launch_boxcar(hipStream_t st, const BoxcarArgs &a, int max_lds_bytes, LaunchSink *sink = nullptr);

// One persistent launch of the segmented boxcar over many rows: job j (a
// row's BoxcarArgs, clock unset) owns work items [first[j], first[j+1]),
// total = first[njobs]; *next is the claim counter, zero at launch.
struct BoxcarQueue {
    const BoxcarArgs *jobs;
    const long *first; // [njobs + 1]
    int njobs;
    long total;
    unsigned *next;
};
// LDS bytes of the segmented form for a, or 0 when a needs the row-wide
// form (exact order, IQR or no stage-2 tile sums).
size_t boxcar_seg_lds(const BoxcarArgs &a);
long boxcar_seg_items(const BoxcarArgs &a);
// Workgroups of a persistent queue launch that fit on the device at once
// (-1 when the occupancy query fails).
int boxcar_queue_grid(size_t shm, int ncu);
void launch_boxcar_queue(hipStream_t st, const BoxcarQueue &q, size_t shm, int grid);

// Gathers of one DM's sub-band series: out[s][t] = sub(k, s, w0 + t + d2[i][s]).
void launch_subdata(hipStream_t st, const float *sub, long spitch, long Lsub, const int *d2,
                    int nsub, int group, long idm, long w0, long N, float *out);
// vpf of candidates: out[c][s] = sum_{a in window} sub(...)/fcnt[s].
void launch_cand_spectra(hipStream_t st, const float *sub, long spitch, long Lsub, const int *d2,
                         const float *fcnt, int nsub, int group, long w0, long N, int ncand,
                         const int *idm, const int *isamp, const int *wn, float *out);
// --savetim: [ndm][n] dump values of tim[.][n0, n0 + n) (nbits 8, 16 or 32)
void launch_dump(hipStream_t st, const float *tim, long N, long n0, long n, int ndm, float scale, float mean,
                 int nbits, void *out);

} // namespace txgpu
