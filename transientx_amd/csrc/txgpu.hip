// txgpu.hip -- the C ABI (include/txgpu.h): beams, plans, launch shapes.
//
// A tx_plan is one ddplan row: SubbandDedispersion + Boxcar of one
// SinglePulse (include/singlepulse.h:44-46).  Device layout per plan:
//   sub  [ndmc][nsub][Lsub + Wsub] f32  sub-band series ring, slot = tau mod Lsub,
//                                       slots [0, Wsub) mirrored at [Lsub, Lsub+Wsub)
//   tim  [ndm][N] f32                   dedispersed series of the current window
//   mxS  [ndm][N] f32, mxwn [ndm][N] i32  Boxcar::mxS / mxwn (optional)
//   pts  [cap] tx_point                 thresholded points (cluster.cpp:89-112)
// Per beam:
//   ring [nchans][L + W] u8|f32         channel-major input ring, same mirroring
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/txgpu.h"
#include "kernels.h"
#include "plan.h"

using namespace txgpu;

static thread_local std::string g_err;

static int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

namespace txgpu {
// error reporting for the host-side entry points (host.cpp)
int set_error(int code, const std::string &msg) { return fail(code, msg); }
} // namespace txgpu

#define HIPCHK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess)                                                                      \
            return fail(TX_EHIP, std::string(#x) + ": " + hipGetErrorString(e_));                 \
    } while (0)

static long round_up(long v, long m) { return (v + m - 1) / m * m; }

struct tx_beam {
    int device = 0;
    int C = 0;
    int dtype = TX_DTYPE_U8;
    int zp = 0;
    long ndump = 0;
    uint8_t *d_mask = nullptr;
    std::vector<uint8_t> mask;
    hipStream_t stream = nullptr;
    void *d_ring = nullptr;
    long L = 0, W = 0, pitch = 0; // elements
    long counter = 0;
    long need_hist = 0, need_win = 0;
    void *d_stage = nullptr;
    size_t stage_bytes = 0;
    // front end on the device (tx_beam_set_frontend): raw 8-bit blocks of
    // fe_cin channels and ndump*fe_td samples
    int fe = 0, fe_cin = 0, fe_td = 1, fe_fd = 1;
    float *d_fe_off = nullptr, *d_fe_gain = nullptr;
    // the front end in the stage-1 load (tx_beam_set_frontend_in_load): with
    // td = fd = 1 and no RFI filters when the ring is allocated, the ring
    // keeps the raw 8-bit samples and stage 1 forms (y - off) * gain as it
    // stages them; masked channels take off = gain = 0 in these copies
    bool fe_in_load = false, fused = false;
    float *d_fe_offm = nullptr, *d_fe_gainm = nullptr;
    int max_lds = 65536;
    int nplans = 0;
    std::vector<tx_plan *> plans;
    // per-kernel HIP-event timing (tx_beam_set_timing)
    bool timing = false;
    std::vector<hipEvent_t> evfree;
    struct Pending {
        int cls;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    double ms[4] = {0, 0, 0, 0};
    long launches[4] = {0, 0, 0, 0};
    // per-block graph pipeline (tx_beam_graph)
    long *d_clock = nullptr;        // the beam's counter before the next graph block
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipGraphNode_t gtrans = nullptr; // the transpose node (its source changes per block)
    KNode trans;
    long gcounter = -1;              // b->counter the device clock matches (-1: none)
    std::vector<tx_point_ex *> gsess; // session lists the graph writes, per plan
    std::vector<size_t> gcap;         // and their capacities
    // TX_PUSH_HOST_ASYNC: two staging buffers filled on a copy stream, so
    // that one block crosses PCIe while the previous one is searched
    // the front end's RFI filters (tx_beam_set_rfi): ops in order on a
    // time-major f32 staging block before the transpose into the ring
    struct RfiOp {
        int op, td, fd;
        float threshold;
    };
    std::vector<RfiOp> rfi;
    float *d_rfi_blk = nullptr;
    void *d_rfi_work = nullptr;
    size_t rfi_work_bytes = 0;
    int *d_rfi_wnl = nullptr; // [nops] run-length limits
    hipStream_t cstream = nullptr;
    void *d_astage[2] = {nullptr, nullptr};
    hipEvent_t ev_full[2] = {nullptr, nullptr}, ev_free[2] = {nullptr, nullptr};
    bool astage_used[2] = {false, false};
    int astage_k = 0;
    size_t astage_bytes = 0;
    size_t esize() const { return dtype == TX_DTYPE_U8 || fused ? 1 : 4; }
};

// Brackets one kernel launch with events on the beam's stream when timing.
struct KernelTimer {
    tx_beam *b;
    int cls;
    hipEvent_t a = nullptr, e = nullptr;
    KernelTimer(tx_beam *beam, int c) : b(beam), cls(c)
    {
        if (!b->timing) return;
        a = take();
        e = take();
        if (a && e) hipEventRecord(a, b->stream);
    }
    hipEvent_t take()
    {
        if (!b->evfree.empty()) {
            hipEvent_t ev = b->evfree.back();
            b->evfree.pop_back();
            return ev;
        }
        hipEvent_t ev = nullptr;
        if (hipEventCreate(&ev) != hipSuccess) return nullptr;
        return ev;
    }
    ~KernelTimer()
    {
        if (!b->timing || !a || !e) return;
        hipEventRecord(e, b->stream);
        b->pending.push_back({cls, a, e});
    }
};

struct tx_plan {
    tx_beam *beam = nullptr;
    PlanMath pm;
    std::vector<int> vwn, whi;
    int iqr = 0, store_maps = 0, exact = 0, window_off = 0;
    long cap = 0;
    long lo_all = 0, hi_all = 0;
    // stage 1 shape
    int NCC = 0, Q = 0, CH = 0, RS = 0, nkc = 0;
    int CH_fe = 0;  // f32 plans: channels per LDS pass with the front end in the load
    long read1 = 0; // ring elements one stage-1 row load spans
    // stage 2 shape
    int DPW = 0, R = 0, nblk = 0, half2 = 0, pstride2 = 0, maxrow2 = 0;
    int *d_ldsoff = nullptr, *d_npass = nullptr, *d_pstart = nullptr;
    // device buffers
    int *d_d1 = nullptr, *d_d1min = nullptr, *d_d1sel = nullptr, *d_d2 = nullptr, *d_d2min = nullptr;
    int *d_d2span = nullptr, *d_d2off = nullptr;
    int *d_vwn = nullptr, *d_whi = nullptr;
    float *d_fcnt = nullptr;
    float *d_sub = nullptr;
    long Lsub = 0, Wsub = 0, spitch = 0;
    float *d_tim = nullptr, *d_mxS = nullptr, *d_scratch = nullptr;
    float *d_xstat = nullptr, *d_xck = nullptr; // exact order: k_boxcar_chain's statistics and checkpoints
    double *d_tstat = nullptr; // [ndm][ntile2][2] per-tile sums of tim (stage 2 -> boxcar)
    int ntile2 = 0;
    int *d_mxwn = nullptr;
    tx_point *d_pts = nullptr;
    unsigned long long *d_npts = nullptr;
    float *d_tmp = nullptr; // nsub*N gather buffer
    int *d_cand = nullptr;  // candidate triples for tx_cand_spectra
    float *d_spec = nullptr;
    int cand_cap = 0;
    void *d_dump = nullptr; // [ndm][ndump] --savetim values (tx_dump_series)
    // tx_boxcar_rows without stored maps: rows recomputed into these
    float *d_rowS = nullptr;
    int *d_rowW = nullptr;
    unsigned long long *d_rown = nullptr;
    size_t row_cap = 0; // elements
    std::vector<double> fcnt;
    long counter = 0;
    long w0 = 0;
    bool have_points = false;
    // session mode (tx_session_begin/end): the kernels append to d_sess, of
    // capacity sess_cap: the plan's own device list, or with
    // tx_session_begin_host the device address of the caller's pinned list
    tx_point_ex *d_sess = nullptr;
    unsigned long long *d_sess_n = nullptr;
    size_t sess_cap = 0;
    tx_point_ex *d_own = nullptr;
    size_t own_cap = 0;
    tx_point_ex *h_sess = nullptr;
    bool in_session = false;
};

extern "C" const char *tx_last_error(void) { return g_err.c_str(); }
extern "C" int tx_version(void) { return 1; }

// ------------------------------------------------------------------- beam
extern "C" int tx_beam_create(const tx_beam_desc *d, tx_beam **out)
{
    if (!d || !out) return fail(TX_EINVAL, "null argument");
    if (d->nchans <= 0) return fail(TX_EINVAL, "nchans must be positive");
    if (d->ndump <= 0) return fail(TX_EINVAL, "ndump must be positive");
    if (d->dtype != TX_DTYPE_U8 && d->dtype != TX_DTYPE_F32) return fail(TX_EINVAL, "bad dtype");
    if (d->dtype == TX_DTYPE_U8 && (d->zero_point < 0 || d->zero_point > 255))
        return fail(TX_EINVAL, "zero_point must be in [0,255]");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (d->device < 0 || d->device >= ndev) return fail(TX_EINVAL, "bad device ordinal");
    HIPCHK(hipSetDevice(d->device));
    tx_beam *b = new tx_beam();
    b->device = d->device;
    b->C = d->nchans;
    b->dtype = d->dtype;
    b->zp = d->zero_point;
    b->ndump = d->ndump;
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, d->device) ==
            hipSuccess &&
        lds > 0)
        b->max_lds = lds;
    if (d->chan_mask) {
        b->mask.assign(d->chan_mask, d->chan_mask + d->nchans);
        if (hipMalloc(&b->d_mask, d->nchans) != hipSuccess) {
            delete b;
            return fail(TX_ENOMEM, "mask allocation");
        }
        hipMemcpy(b->d_mask, d->chan_mask, d->nchans, hipMemcpyHostToDevice);
    }
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        hipFree(b->d_mask);
        delete b;
        return fail(TX_EHIP, "stream creation");
    }
    *out = b;
    return TX_OK;
}

extern "C" int tx_beam_destroy(tx_beam *b)
{
    if (!b) return TX_OK;
    hipSetDevice(b->device);
    if (b->stream) hipStreamSynchronize(b->stream);
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    if (b->graph) hipGraphDestroy(b->graph);
    if (b->cstream) hipStreamSynchronize(b->cstream);
    for (int k = 0; k < 2; k++) {
        hipFree(b->d_astage[k]);
        if (b->ev_full[k]) hipEventDestroy(b->ev_full[k]);
        if (b->ev_free[k]) hipEventDestroy(b->ev_free[k]);
    }
    if (b->cstream) hipStreamDestroy(b->cstream);
    hipFree(b->d_rfi_blk);
    hipFree(b->d_rfi_work);
    hipFree(b->d_rfi_wnl);
    hipFree(b->d_clock);
    hipFree(b->d_ring);
    hipFree(b->d_stage);
    hipFree(b->d_fe_off);
    hipFree(b->d_fe_gain);
    hipFree(b->d_fe_offm);
    hipFree(b->d_fe_gainm);
    hipFree(b->d_mask);
    for (auto &p : b->pending) {
        hipEventDestroy(p.a);
        hipEventDestroy(p.b);
    }
    for (auto ev : b->evfree) hipEventDestroy(ev);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
    return TX_OK;
}

// Drops the beam's per-block graph: its kernel nodes hold device pointers of
// the plans (widths, session lists, buffers) that are about to change.
static void beam_drop_graph(tx_beam *b)
{
    if (!b) return;
    if (b->gexec || b->graph) hipStreamSynchronize(b->stream);
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    b->gtrans = nullptr;
    b->gcounter = -1;
    b->gsess.clear();
    b->gcap.clear();
}

static int beam_alloc_ring(tx_beam *b)
{
    // the ring's format is fixed here: raw 8-bit samples when the front end
    // was asked into the stage-1 load and can go there for every plan
    b->fused = b->fe_in_load && b->fe && b->fe_td == 1 && b->fe_fd == 1 && b->rfi.empty() && b->d_fe_offm;
    for (const tx_plan *p : b->plans)
        if (p->CH_fe < 1) b->fused = false;
    b->L = round_up(b->ndump + b->need_hist + 64, 64);
    b->W = round_up(std::max<long>(b->need_win, 64) + 64, 64);
    b->pitch = b->L + b->W;
    const size_t bytes = (size_t)b->C * b->pitch * b->esize();
    if (hipMalloc(&b->d_ring, bytes) != hipSuccess) return fail(TX_ENOMEM, "input ring allocation");
    // the zero signal: x == zero_point contributes nothing
    HIPCHK(hipMemsetAsync(b->d_ring, b->dtype == TX_DTYPE_U8 ? b->zp : 0, bytes, b->stream));
    return TX_OK;
}

// Samples and bytes of one pushed block: ndump samples of the beam's
// channels, or with the front end ndump*td raw 8-bit samples of fe_cin.
static long beam_block_samples(const tx_beam *b) { return b->fe ? b->ndump * b->fe_td : b->ndump; }
static size_t beam_block_bytes(const tx_beam *b)
{
    return b->fe ? (size_t)b->ndump * b->fe_td * b->fe_cin : (size_t)b->ndump * b->C * b->esize();
}

static int beam_stage(tx_beam *b)
{
    const size_t bytes = beam_block_bytes(b);
    if (b->d_stage && b->stage_bytes >= bytes) return TX_OK;
    if (b->d_stage) {
        HIPCHK(hipStreamSynchronize(b->stream)); // a queued copy may still read it
        hipFree(b->d_stage);
        b->d_stage = nullptr;
        b->stage_bytes = 0;
    }
    if (hipMalloc(&b->d_stage, bytes) != hipSuccess) return fail(TX_ENOMEM, "staging allocation");
    b->stage_bytes = bytes;
    return TX_OK;
}

// TX_PUSH_HOST_ASYNC: queue the copy of a pinned host block into the next of
// two staging buffers on the beam's copy stream, once the search that last
// read that buffer is done (ev_free), and make the beam's stream wait for the
// copy (ev_full).  *src receives the staging buffer; the caller records
// ev_free[*k] on the beam's stream after the work that reads it.
static int beam_copy_async(tx_beam *b, const void *block, const void **src, int *k)
{
    const size_t bytes = beam_block_bytes(b);
    if (!b->cstream) {
        HIPCHK(hipStreamCreateWithFlags(&b->cstream, hipStreamNonBlocking));
        for (int j = 0; j < 2; j++) {
            HIPCHK(hipEventCreateWithFlags(&b->ev_full[j], hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&b->ev_free[j], hipEventDisableTiming));
        }
    }
    if (b->astage_bytes < bytes) {
        HIPCHK(hipStreamSynchronize(b->stream));
        HIPCHK(hipStreamSynchronize(b->cstream));
        for (int j = 0; j < 2; j++) {
            hipFree(b->d_astage[j]);
            b->d_astage[j] = nullptr;
            b->astage_used[j] = false;
        }
        b->astage_bytes = 0;
        for (int j = 0; j < 2; j++)
            if (hipMalloc(&b->d_astage[j], bytes) != hipSuccess) return fail(TX_ENOMEM, "staging allocation");
        b->astage_bytes = bytes;
    }
    const int j = b->astage_k;
    b->astage_k ^= 1;
    if (b->astage_used[j]) HIPCHK(hipStreamWaitEvent(b->cstream, b->ev_free[j], 0));
    HIPCHK(hipMemcpyAsync(b->d_astage[j], block, bytes, hipMemcpyHostToDevice, b->cstream));
    HIPCHK(hipEventRecord(b->ev_full[j], b->cstream));
    HIPCHK(hipStreamWaitEvent(b->stream, b->ev_full[j], 0));
    b->astage_used[j] = true;
    *src = b->d_astage[j];
    *k = j;
    return TX_OK;
}

// K0 of one block at ring time t0 (or the device clock), launched or recorded.
// The RFI chain of one block (tx_beam_set_rfi) on the beam's stream: the
// front end's (or the pushed f32 block's) values into the time-major staging
// block with masked channels zeroed, each filter in place, then the f32
// transpose into the ring.
static int beam_launch_rfi(tx_beam *b, const void *src, long t0)
{
    const long nd = b->ndump;
    const int C = b->C;
    float *blk = b->d_rfi_blk;
    if (b->fe) {
        // nd output samples from the raw block of nd*td samples
        launch_frontend_tm(b->stream, (const uint8_t *)src, nd, b->fe_cin, b->fe_td, b->fe_fd, b->d_fe_off,
                           b->d_fe_gain, b->d_mask, blk);
    } else {
        HIPCHK(hipMemcpyAsync(blk, src, (size_t)nd * C * 4, hipMemcpyDeviceToDevice, b->stream));
        if (b->d_mask) launch_mask_tm(b->stream, blk, nd, C, b->d_mask);
    }
    for (size_t k = 0; k < b->rfi.size(); k++) {
        const tx_beam::RfiOp &o = b->rfi[k];
        const int rc = o.op == TX_RFI_ZDOT
                           ? tx_rfi_zdot(b->stream, blk, 1, (long)nd * C, C, nd, 1, C, b->d_rfi_work,
                                         b->rfi_work_bytes)
                           : tx_rfi_kadanef(b->stream, blk, 1, (long)nd * C, C, nd, 1, C, o.td, o.fd, o.threshold,
                                            b->d_rfi_wnl + k, b->d_rfi_work, b->rfi_work_bytes);
        if (rc) return rc;
    }
    launch_transpose_f32(b->stream, blk, nd, C, b->d_mask, (float *)b->d_ring, b->pitch, b->L, b->W, t0, nullptr,
                         nullptr);
    HIPCHK(hipGetLastError());
    return TX_OK;
}

static void beam_launch_input(tx_beam *b, const void *src, long t0, const long *clock, LaunchSink *sink)
{
    if (b->fused) // raw samples into the 8-bit ring; stage 1 applies the front end
        launch_transpose_u8(b->stream, (const uint8_t *)src, b->ndump, b->C, b->d_mask, 0, (uint8_t *)b->d_ring,
                            b->pitch, b->L, b->W, t0, clock, sink);
    else if (b->fe)
        launch_frontend_u8(b->stream, (const uint8_t *)src, b->ndump, b->fe_cin, b->fe_td, b->fe_fd,
                           b->d_fe_off, b->d_fe_gain, b->d_mask, (float *)b->d_ring, b->pitch, b->L, b->W,
                           t0, clock, sink);
    else if (b->dtype == TX_DTYPE_U8)
        launch_transpose_u8(b->stream, (const uint8_t *)src, b->ndump, b->C, b->d_mask, (uint8_t)b->zp,
                            (uint8_t *)b->d_ring, b->pitch, b->L, b->W, t0, clock, sink);
    else
        launch_transpose_f32(b->stream, (const float *)src, b->ndump, b->C, b->d_mask, (float *)b->d_ring,
                             b->pitch, b->L, b->W, t0, clock, sink);
}

extern "C" int tx_beam_push(tx_beam *b, const void *block, long ns, int on_device)
{
    if (!b || !block) return fail(TX_EINVAL, "null argument");
    if (ns != beam_block_samples(b))
        return fail(TX_EINVAL, b->fe ? "a front-end block must hold exactly ndump*td raw samples"
                                     : "a block must hold exactly ndump samples");
    HIPCHK(hipSetDevice(b->device));
    if (!b->d_ring) {
        int rc = beam_alloc_ring(b);
        if (rc) return rc;
    }
    const void *src = block;
    int k = -1;
    if (on_device == TX_PUSH_HOST) {
        // the staging buffer is reused block after block: a copy is queued
        // behind the previous block's transpose on the same stream
        const int rc = beam_stage(b);
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(b->d_stage, block, beam_block_bytes(b), hipMemcpyHostToDevice, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
        src = b->d_stage;
    } else if (on_device == TX_PUSH_HOST_ASYNC) {
        const int rc = beam_copy_async(b, block, &src, &k);
        if (rc) return rc;
    }
    if (!b->rfi.empty()) {
        const int rc = beam_launch_rfi(b, src, b->counter);
        if (rc) return rc;
    } else {
        KernelTimer kt(b, 0);
        beam_launch_input(b, src, b->counter, nullptr, nullptr);
    }
    HIPCHK(hipGetLastError());
    if (k >= 0) HIPCHK(hipEventRecord(b->ev_free[k], b->stream));
    b->counter += b->ndump;
    return TX_OK;
}

extern "C" int tx_beam_set_frontend(tx_beam *b, int nchans_in, int td, int fd, const float *offset,
                                    const float *gain)
{
    if (!b) return fail(TX_EINVAL, "null argument");
    if (nchans_in == 0) {
        if (b->fused)
            return fail(TX_ESTATE, "the beam's input ring holds the front end's 8-bit samples since its first "
                                   "push: the front end stays on (create a new beam for f32 blocks)");
        HIPCHK(hipSetDevice(b->device));
        beam_drop_graph(b);
        HIPCHK(hipStreamSynchronize(b->stream));
        b->fe = 0;
        return TX_OK;
    }
    if (b->dtype != TX_DTYPE_F32) return fail(TX_EINVAL, "the front end needs a TX_DTYPE_F32 beam");
    if (td < 1 || fd < 1 || td * fd > 256) return fail(TX_EINVAL, "td and fd must be >= 1 with td*fd <= 256");
    if (nchans_in < fd || nchans_in / fd != b->C)
        return fail(TX_EINVAL, "nchans_in / fd must equal the beam's nchans");
    if (!offset || !gain) return fail(TX_EINVAL, "null offset or gain");
    if (b->fused && (td != 1 || fd != 1))
        return fail(TX_ESTATE, "the beam's input ring holds 8-bit samples since its first push: only td = fd = 1 "
                               "(new offsets and gains) can follow");
    HIPCHK(hipSetDevice(b->device));
    beam_drop_graph(b);
    HIPCHK(hipStreamSynchronize(b->stream)); // queued blocks still read the old values
    const size_t fb = (size_t)b->C * sizeof(float);
    if (!b->d_fe_off && (hipMalloc((void **)&b->d_fe_off, fb) != hipSuccess ||
                         hipMalloc((void **)&b->d_fe_gain, fb) != hipSuccess ||
                         hipMalloc((void **)&b->d_fe_offm, fb) != hipSuccess ||
                         hipMalloc((void **)&b->d_fe_gainm, fb) != hipSuccess))
        return fail(TX_ENOMEM, "front-end allocation");
    HIPCHK(hipMemcpy(b->d_fe_off, offset, fb, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->d_fe_gain, gain, fb, hipMemcpyHostToDevice));
    {
        std::vector<float> om(offset, offset + b->C), gm(gain, gain + b->C);
        for (int c = 0; c < b->C; c++)
            if (!b->mask.empty() && !b->mask[c]) om[c] = gm[c] = 0.f;
        HIPCHK(hipMemcpy(b->d_fe_offm, om.data(), fb, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(b->d_fe_gainm, gm.data(), fb, hipMemcpyHostToDevice));
    }
    b->fe = 1;
    b->fe_cin = nchans_in;
    b->fe_td = td;
    b->fe_fd = fd;
    return TX_OK;
}

extern "C" int tx_beam_set_frontend_in_load(tx_beam *b, int on)
{
    if (!b) return fail(TX_EINVAL, "null argument");
    if (b->dtype != TX_DTYPE_F32) return fail(TX_EINVAL, "the front end needs a TX_DTYPE_F32 beam");
    if (b->d_ring) return fail(TX_ESTATE, "tx_beam_set_frontend_in_load: call it before the first push");
    b->fe_in_load = on != 0;
    return TX_OK;
}

extern "C" int tx_beam_set_rfi(tx_beam *b, int nops, const int *ops, const int *td, const int *fd,
                               const float *threshold, const int *wnlimit)
{
    if (!b) return fail(TX_EINVAL, "null argument");
    if (nops < 0 || nops > 16) return fail(TX_EINVAL, "between 0 and 16 RFI filters");
    if (nops > 0 && !ops) return fail(TX_EINVAL, "null ops");
    HIPCHK(hipSetDevice(b->device));
    if (nops == 0) {
        HIPCHK(hipStreamSynchronize(b->stream));
        b->rfi.clear();
        return TX_OK;
    }
    if (b->dtype != TX_DTYPE_F32) return fail(TX_EINVAL, "the RFI filters need a TX_DTYPE_F32 beam");
    if (b->fused)
        return fail(TX_ESTATE, "the beam's input ring holds the front end's 8-bit samples since its first push: "
                               "RFI filters must be set before it");
    std::vector<tx_beam::RfiOp> v;
    std::vector<int> wnl(nops, 0);
    size_t work = 0;
    for (int k = 0; k < nops; k++) {
        tx_beam::RfiOp o{ops[k], 1, 1, 0.f};
        if (o.op == TX_RFI_KADANEF) {
            if (!td || !fd || !threshold || !wnlimit) return fail(TX_EINVAL, "kadaneF needs td, fd, threshold, wnlimit");
            o.td = td[k];
            o.fd = fd[k];
            o.threshold = threshold[k];
            wnl[k] = wnlimit[k];
            if (o.td < 1 || o.fd < 1 || o.td > b->ndump || o.fd > b->C)
                return fail(TX_EINVAL, "kadaneF: 1 <= td <= ndump and 1 <= fd <= nchans");
        } else if (o.op != TX_RFI_ZDOT) {
            return fail(TX_EINVAL, "RFI filters are TX_RFI_ZDOT or TX_RFI_KADANEF");
        }
        work = std::max(work, tx_rfi_work_bytes(o.op, 1, b->C, b->ndump, o.td, o.fd));
        v.push_back(o);
    }
    beam_drop_graph(b);
    HIPCHK(hipStreamSynchronize(b->stream)); // queued blocks may still use the old buffers
    if (!b->d_rfi_blk && hipMalloc((void **)&b->d_rfi_blk, (size_t)b->ndump * b->C * 4) != hipSuccess)
        return fail(TX_ENOMEM, "RFI staging allocation");
    if (work > b->rfi_work_bytes) {
        hipFree(b->d_rfi_work);
        b->d_rfi_work = nullptr;
        b->rfi_work_bytes = 0;
        if (hipMalloc(&b->d_rfi_work, work) != hipSuccess) return fail(TX_ENOMEM, "RFI work allocation");
        b->rfi_work_bytes = work;
    }
    if (!b->d_rfi_wnl && hipMalloc((void **)&b->d_rfi_wnl, 16 * sizeof(int)) != hipSuccess)
        return fail(TX_ENOMEM, "RFI allocation");
    HIPCHK(hipMemcpy(b->d_rfi_wnl, wnl.data(), nops * sizeof(int), hipMemcpyHostToDevice));
    b->rfi = v;
    return TX_OK;
}

extern "C" int tx_frontend_u8(void *stream, const uint8_t *in, long ns_in, int nchans_in, int td, int fd,
                              const float *offset, const float *gain, const uint8_t *chan_mask, float *out)
{
    if (!in || !offset || !gain || !out) return fail(TX_EINVAL, "null argument");
    if (ns_in < 0 || nchans_in < 1) return fail(TX_EINVAL, "bad block shape");
    if (td < 1 || fd < 1 || td * fd > 256) return fail(TX_EINVAL, "td and fd must be >= 1 with td*fd <= 256");
    if (nchans_in < fd) return fail(TX_EINVAL, "nchans_in must be >= fd");
    const long ns = ns_in / td;
    if (ns == 0) return TX_OK;
    launch_frontend_tm((hipStream_t)stream, in, ns, nchans_in, td, fd, offset, gain, chan_mask, out);
    HIPCHK(hipGetLastError());
    return TX_OK;
}

extern "C" size_t tx_frontend_stats_work_bytes(int nchans_in, int fd)
{
    return fd >= 1 && nchans_in >= fd ? (size_t)(nchans_in / fd) * 2 * sizeof(unsigned long long) : 0;
}

extern "C" int tx_frontend_stats(void *stream, const uint8_t *in, long ns_in, int nchans_in, int td, int fd,
                                 float *offset, float *gain, void *work, size_t work_bytes)
{
    if (!in || !offset || !gain || !work) return fail(TX_EINVAL, "null argument");
    if (nchans_in < 1 || td < 1 || fd < 1 || td * fd > 256 || nchans_in < fd)
        return fail(TX_EINVAL, "td and fd must be >= 1 with td*fd <= 256 and nchans_in >= fd");
    const long ns = ns_in / td;
    if (ns < 1) return fail(TX_EINVAL, "the block holds no downsampled sample");
    if (ns > (1L << 31)) return fail(TX_EINVAL, "block too long for exact 64-bit sums");
    if (work_bytes < tx_frontend_stats_work_bytes(nchans_in, fd)) return fail(TX_EINVAL, "work buffer too small");
    launch_frontend_stats((hipStream_t)stream, in, ns, nchans_in, td, fd, (unsigned long long *)work, offset,
                          gain);
    HIPCHK(hipGetLastError());
    return TX_OK;
}

// Scratch of tx_frontend_stats_noalloc: the exact per-channel sums (two
// 64-bit words per output channel) for callers that cannot allocate device
// memory.  Each GPU holds its own copy until the process ends.  In host code
// the name stands for no device address: kernels get the address that
// hipGetSymbolAddress returns for the device, kept per device below.
static constexpr int kNoAllocMaxChans = 65536;
static __device__ __align__(256) unsigned long long scratch[2 * kNoAllocMaxChans];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];

// The address of this device's scratch, looked up on the first call on a
// device (with that device current) and kept.
static int scratch_address(hipStream_t st, void **out)
{
    int dev = 0;
    if (st) {
        HIPCHK(hipStreamGetDevice(st, &dev));
    } else {
        HIPCHK(hipGetDevice(&dev));
    }
    if (dev < 0 || dev >= kMaxDevices) return fail(TX_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int cur = 0;
        HIPCHK(hipGetDevice(&cur));
        if (cur != dev) HIPCHK(hipSetDevice(dev));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) HIPCHK(hipSetDevice(cur));
        if (e != hipSuccess || !p)
            return fail(TX_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return TX_OK;
}

extern "C" int tx_frontend_stats_noalloc(void *stream, const uint8_t *in, long ns_in, int nchans_in, int td,
                                         int fd, float *offset, float *gain)
{
    if (!in || !offset || !gain) return fail(TX_EINVAL, "null argument");
    if (nchans_in < 1 || td < 1 || fd < 1 || td * fd > 256 || nchans_in < fd)
        return fail(TX_EINVAL, "td and fd must be >= 1 with td*fd <= 256 and nchans_in >= fd");
    if (nchans_in / fd > kNoAllocMaxChans)
        return fail(TX_EINVAL, "tx_frontend_stats_noalloc: more than 65536 output channels (the fixed scratch "
                               "holds 65536); use tx_frontend_stats with a work buffer");
    const long ns = ns_in / td;
    if (ns < 1) return fail(TX_EINVAL, "the block holds no downsampled sample");
    if (ns > (1L << 31)) return fail(TX_EINVAL, "block too long for exact 64-bit sums");
    void *work = nullptr;
    const int rc = scratch_address((hipStream_t)stream, &work);
    if (rc != TX_OK) return rc;
    // the kernels clear the sums they use first (no memset node), so the
    // previous call's contents do not matter
    launch_frontend_stats((hipStream_t)stream, in, ns, nchans_in, td, fd, (unsigned long long *)work, offset,
                          gain);
    HIPCHK(hipGetLastError());
    return TX_OK;
}

extern "C" int tx_beam_sync(tx_beam *b)
{
    if (!b) return fail(TX_EINVAL, "null argument");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    if (b->cstream) HIPCHK(hipStreamSynchronize(b->cstream));
    return TX_OK;
}

// ------------------------------------------------------------------- plan
static void plan_free(tx_plan *p)
{
    hipFree(p->d_d1);
    hipFree(p->d_d1min);
    hipFree(p->d_d1sel);
    hipFree(p->d_d2);
    hipFree(p->d_d2min);
    hipFree(p->d_d2span);
    hipFree(p->d_ldsoff);
    hipFree(p->d_npass);
    hipFree(p->d_pstart);
    hipFree(p->d_d2off);
    hipFree(p->d_vwn);
    hipFree(p->d_whi);
    hipFree(p->d_fcnt);
    hipFree(p->d_sub);
    hipFree(p->d_tim);
    hipFree(p->d_mxS);
    hipFree(p->d_dump);
    hipFree(p->d_rowS);
    hipFree(p->d_rowW);
    hipFree(p->d_rown);
    hipFree(p->d_mxwn);
    hipFree(p->d_scratch);
    hipFree(p->d_xstat);
    hipFree(p->d_xck);
    hipFree(p->d_tstat);
    hipFree(p->d_pts);
    hipFree(p->d_npts);
    hipFree(p->d_tmp);
    hipFree(p->d_cand);
    hipFree(p->d_spec);
    hipFree(p->d_own); // d_sess may be a host list's address: never freed here
    hipFree(p->d_sess_n);
}

// Stage-1 shape: NCC coarse DMs per workgroup (one input pass serves them
// all), Q quads per thread, CH channel rows per LDS pass within budget.
static bool choose_stage1(tx_plan *p, std::vector<int> &d1min, int lds_budget, int force_ncc,
                          int force_q)
{
    const PlanMath &pm = p->pm;
    const bool u8 = p->beam->dtype == TX_DTYPE_U8;
    const int maxncc = u8 ? 16 : 8;
    int ncc0 = 1;
    while (ncc0 < pm.ndmc && ncc0 < maxncc) ncc0 <<= 1;
    if (force_ncc > maxncc) return false;
    if (force_ncc > ncc0) ncc0 = force_ncc;
    for (int NCC = ncc0; NCC >= 1; NCC >>= 1) {
        if (force_ncc > 0 && NCC != force_ncc) continue;
        const int nkc = (pm.ndmc + NCC - 1) / NCC;
        std::vector<int> mn((size_t)nkc * pm.nchans);
        int spread = 0;
        for (int kc = 0; kc < nkc; kc++)
            for (int j = 0; j < pm.nchans; j++) {
                int lo = 1 << 30, hi = -(1 << 30);
                for (int k = kc * NCC; k < std::min(pm.ndmc, (kc + 1) * NCC); k++) {
                    const int d = pm.d1[(size_t)k * pm.nchans + j];
                    lo = std::min(lo, d);
                    hi = std::max(hi, d);
                }
                // rows start on whole 32-bit words of an 8-bit ring (tiles
                // start on multiples of 4 samples and the ring length is
                // one); f32 rows too, so that an f32 plan also serves the
                // front end in the stage-1 load over an 8-bit ring
                lo -= ((lo % 4) + 4) % 4;
                mn[(size_t)kc * pm.nchans + j] = lo;
                spread = std::max(spread, hi - lo);
            }
        for (int Q = 2; Q >= 1; Q--) {
            if (force_q > 0 && Q != force_q) continue;
            // u8: words covering spread (from the word-aligned row start) +
            // 4*256*Q bytes + the extra word each quad reads; f32: spread +
            // 1024*Q floats, a multiple of 16 so that the same rows of an
            // 8-bit ring are whole 16-byte fills
            // (rows a multiple of 4 words: 16-byte global->LDS fills)
            const int RS = u8 ? ((spread + 4 * 256 * Q + 8 + 3) / 4 + 3) & ~3 : (spread + 1024 * Q + 1 + 15) & ~15;
            int CH = 0, CH_fe = 0;
            for (int c = std::min(16, pm.cps); c >= 1; c--)
                if (pm.cps % c == 0 && (long)c * RS * 4 <= lds_budget) {
                    CH = c;
                    break;
                }
            // the front end in the load: RS floats plus RS raw bytes per row
            for (int c = std::min(16, pm.cps); c >= 1 && !u8; c--)
                if (pm.cps % c == 0 && (long)c * RS * 5 <= lds_budget) {
                    CH_fe = c;
                    break;
                }
            if (CH >= std::min(4, pm.cps) || (NCC == 1 && Q == 1 && CH >= 1)) {
                p->NCC = NCC;
                p->Q = Q;
                p->CH = CH;
                p->CH_fe = CH_fe;
                p->RS = RS;
                p->nkc = nkc;
                p->read1 = u8 ? (long)RS * 4 + 4 : (long)RS + 1;
                d1min.swap(mn);
                return true;
            }
        }
    }
    return false;
}

// Stage-2 delay extents of every DM block for blocks of `blk` fine DMs.
static int stage2_extents(const PlanMath &pm, int blk, int nblk, std::vector<int> &d2min,
                          std::vector<int> &d2span, std::vector<int> *nact)
{
    d2min.assign((size_t)pm.ndmc * nblk * pm.nsub, 0);
    d2span.assign(d2min.size(), 0);
    if (nact) nact->assign((size_t)pm.ndmc * nblk, 0);
    int span = 0;
    for (int k = 0; k < pm.ndmc; k++)
        for (int b = 0; b < nblk; b++) {
            const long i0 = (long)k * pm.group + (long)b * blk;
            const long i1 =
                std::min<long>(std::min<long>((long)(k + 1) * pm.group, pm.ndm), i0 + blk);
            if (nact) (*nact)[(size_t)k * nblk + b] = (int)std::max<long>(0, i1 - i0);
            for (int s = 0; s < pm.nsub; s++) {
                int lo = 0, hi = 0;
                if (i1 > i0) {
                    lo = 1 << 30;
                    hi = -(1 << 30);
                    for (long i = i0; i < i1; i++) {
                        const int d = pm.d2[(size_t)i * pm.nsub + s];
                        lo = std::min(lo, d);
                        hi = std::max(hi, d);
                    }
                }
                d2min[((size_t)k * nblk + b) * pm.nsub + s] = lo;
                d2span[((size_t)k * nblk + b) * pm.nsub + s] = hi - lo;
                span = std::max(span, hi - lo);
            }
        }
    return span;
}

// Pass plan of stage 2 for one shape: each DM block's sub-band rows (its
// delay extent + 64*R floats) packed back to back into passes of at most
// `half` floats and `maxsub` sub-bands (the kernel's lane tables hold 4 chunks
// of 64/DPW).  Returns false if one row alone does not fit.
struct Stage2Plan {
    std::vector<int> ldsoff, npass, pstart;
    int pstride = 0, maxrow = 0;
    long passes = 0;
};
static bool stage2_passes(const PlanMath &pm, int nblkall, int T, const std::vector<int> &sp,
                          const std::vector<int> &nact, int half, int maxsub, Stage2Plan &pl)
{
    pl.ldsoff.assign((size_t)nblkall * pm.nsub, 0);
    pl.npass.assign(nblkall, 0);
    std::vector<std::vector<int>> starts(nblkall);
    pl.maxrow = 0;
    pl.passes = 0;
    for (int w = 0; w < nblkall; w++) {
        int used = 0;
        starts[w].push_back(0);
        for (int s = 0; s < pm.nsub; s++) {
            const int len = sp[(size_t)w * pm.nsub + s] + T;
            pl.maxrow = std::max(pl.maxrow, len);
            if (len > half) return false;
            if (used + len > half || s - starts[w].back() >= maxsub) {
                starts[w].push_back(s);
                used = 0;
            }
            pl.ldsoff[(size_t)w * pm.nsub + s] = used;
            used += (len + 3) & ~3; // rows start 16-byte aligned
        }
        starts[w].push_back(pm.nsub);
        pl.npass[w] = (int)starts[w].size() - 1;
        if (nact[w]) pl.passes += pl.npass[w];
    }
    pl.pstride = 0;
    for (auto &v : starts) pl.pstride = std::max(pl.pstride, (int)v.size());
    pl.pstart.assign((size_t)nblkall * pl.pstride, pm.nsub);
    for (int w = 0; w < nblkall; w++)
        std::copy(starts[w].begin(), starts[w].end(), pl.pstart.begin() + (size_t)w * pl.pstride);
    return true;
}

// Stage-2 shape: 8 waves, DPW fine DMs per wave, R samples per lane.  The
// cost of a shape counts, per time tile and DM block: one LDS read per
// sample-add of its live DMs, the staging traffic of each sub-band row (the
// block's delay extent + the tile, read from L2/HBM), and a fixed cost per
// pass (barrier and load latency the double buffer cannot hide).
static bool choose_stage2(tx_plan *p, std::vector<int> &d2min, std::vector<int> &d2span,
                          Stage2Plan &plan, int half, int force_shape)
{
    const PlanMath &pm = p->pm;
    double best = 1e300;
    int bi = -1;
    for (int c = 0; c < kNumStage2Shapes; c++) {
        if (force_shape > 0 && force_shape - 1 != c) continue;
        const int DPW = kStage2Shapes[c][0], R = kStage2Shapes[c][1];
        const int blk = 8 * DPW, T = 64 * R;
        const int nblk = (pm.group + blk - 1) / blk;
        std::vector<int> mn, sp, nact;
        stage2_extents(pm, blk, nblk, mn, sp, &nact);
        Stage2Plan pl;
        if (!stage2_passes(pm, pm.ndmc * nblk, T, sp, nact, half, 4 * (64 / DPW), pl)) continue;
        const long ntile = (pm.nsamples + T - 1) / T;
        double cost = 0;
        for (size_t w = 0; w < nact.size(); w++) {
            if (!nact[w]) continue;
            double fill = 0;
            for (int s = 0; s < pm.nsub; s++) fill += sp[w * pm.nsub + s] + T;
            cost += (double)ntile * ((double)nact[w] * T * pm.nsub + 2.0 * fill);
        }
        cost += (double)ntile * pl.passes * 16384.0;
        const double nwg = (double)ntile * pm.ndmc * nblk;
        if (nwg < 1024) cost *= 1.0 + 0.5 * (1024 - nwg) / 1024; // too few workgroups to fill 256 CUs
        if (cost < best) {
            best = cost;
            bi = c;
        }
    }
    if (bi < 0) return false;
    p->DPW = kStage2Shapes[bi][0];
    p->R = kStage2Shapes[bi][1];
    const int blk = 8 * p->DPW;
    p->nblk = (pm.group + blk - 1) / blk;
    std::vector<int> nact;
    stage2_extents(pm, blk, p->nblk, d2min, d2span, &nact);
    return stage2_passes(pm, pm.ndmc * p->nblk, 64 * p->R, d2span, nact, half, 4 * (64 / p->DPW), plan);
}

template <typename T>
static int upload(T **dst, const std::vector<T> &src)
{
    if (src.empty()) return TX_OK;
    if (hipMalloc((void **)dst, src.size() * sizeof(T)) != hipSuccess)
        return fail(TX_ENOMEM, "table allocation");
    HIPCHK(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return TX_OK;
}

static const int kMaxBox = 256;

static std::string check_widths(const int *vwn, int nbox, long N)
{
    if (!vwn || nbox <= 0 || nbox > kMaxBox) return "nbox must be 1..256";
    for (int k = 0; k < nbox; k++)
        if (vwn[k] <= 0 || (k > 0 && vwn[k] < vwn[k - 1])) return "widths must be positive and non-decreasing";
    if (vwn[nbox - 1] >= N) return "widest boxcar must be shorter than ndump+noverlap";
    return "";
}

// The width ladder and the AVX2 output range of each width
// (src/singlepulse/boxcar.cpp:264-267); uploaded only when it changes.
static int set_widths(tx_plan *p, const int *vwn, int nbox)
{
    const long N = p->pm.nsamples;
    const std::string why = check_widths(vwn, nbox, N);
    if (!why.empty()) return fail(TX_EINVAL, why);
    if ((int)p->vwn.size() == nbox && std::equal(vwn, vwn + nbox, p->vwn.begin())) return TX_OK;
    beam_drop_graph(p->beam);
    p->vwn.assign(vwn, vwn + nbox);
    p->whi.resize(nbox);
    p->lo_all = 0;
    p->hi_all = N;
    for (int k = 0; k < nbox; k++) {
        const int wn = vwn[k], wl = wn / 2 + 1, wh = (wn - 1) / 2;
        const long len = N - wh - wl;
        p->whi[k] = (int)(wl + (len > 0 ? len / 8 * 8 : 0));
        p->lo_all = std::max<long>(p->lo_all, wl);
        p->hi_all = std::min<long>(p->hi_all, p->whi[k]);
    }
    // kernels of earlier blocks may still read the old ladder
    HIPCHK(hipStreamSynchronize(p->beam->stream));
    HIPCHK(hipMemcpy(p->d_vwn, p->vwn.data(), nbox * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(p->d_whi, p->whi.data(), nbox * sizeof(int), hipMemcpyHostToDevice));
    return TX_OK;
}

extern "C" int tx_plan_create(tx_beam *b, const tx_plan_desc *d, tx_plan **out)
{
    if (!b || !d || !out || !d->frequencies) return fail(TX_EINVAL, "null argument");
    if (b->d_ring) return fail(TX_ESTATE, "plans must be created before the first tx_beam_push");
    if (d->nbox < 0 || d->nbox > kMaxBox || (d->nbox > 0 && !d->vwn))
        return fail(TX_EINVAL, "nbox must be 0..256 (0: widths given later by tx_boxcar_widths)");
    HIPCHK(hipSetDevice(b->device));
    tx_plan *p = new tx_plan();
    p->beam = b;
    std::string why = build_plan(p->pm, d->frequencies, b->C, d->tsamp, b->ndump, d->overlap,
                                 d->dms, d->ddm, d->ndm, d->nsub, d->group);
    if (!why.empty()) {
        delete p;
        return fail(TX_EINVAL, why);
    }
    PlanMath &pm = p->pm;
    if (b->dtype == TX_DTYPE_U8 && pm.cps > 256) {
        delete p;
        return fail(TX_EINVAL, "8-bit input needs at most 256 channels per sub-band");
    }
    if (d->nbox > 0) {
        const std::string why_w = check_widths(d->vwn, d->nbox, pm.nsamples);
        if (!why_w.empty()) {
            delete p;
            return fail(TX_EINVAL, why_w);
        }
    }
    p->iqr = d->iqr != 0;
    p->store_maps = d->store_maps != 0;
    p->exact = d->exact_order != 0;
    p->window_off = d->s3_window != 0;
    p->cap = d->max_points > 0 ? d->max_points : (1L << 22);
    const long N = pm.nsamples;
    // live channels per sub-band (candplot divides by fcnt, candplot.cpp:164)
    p->fcnt.assign(pm.nsub, 0.);
    for (int j = 0; j < b->C; j++)
        if (b->mask.empty() || b->mask[j]) p->fcnt[j / pm.cps] += 1.;
    std::vector<float> fcntf(pm.nsub);
    for (int s = 0; s < pm.nsub; s++) fcntf[s] = (float)p->fcnt[s];

    // stage 1: 256-thread workgroups, ~6 per CU -> 24 KB of LDS each;
    // stage 2: 512-thread workgroups, 2 per CU -> 80 KB each
    const int budget1 = std::min(b->max_lds, 24576);
    int budget2 = std::min(b->max_lds, 81920);
    if (d->s2_lds > 0) budget2 = std::min(b->max_lds, d->s2_lds);
    if (d->s1_ncc < 0 || d->s1_ncc > 16 || (d->s1_ncc & (d->s1_ncc - 1)) || d->s1_q < 0 || d->s1_q > 2 ||
        d->s2_shape < 0 || d->s2_shape > kNumStage2Shapes || (d->s2_lds != 0 && d->s2_lds < 4096) ||
        d->s3_window < 0 || d->s3_window > 1) {
        delete p;
        return fail(TX_EINVAL, "launch shape override out of range");
    }
    std::vector<int> d1min, d2min, d2span;
    if (!choose_stage1(p, d1min, budget1, d->s1_ncc, d->s1_q)) {
        delete p;
        return fail(TX_EINVAL, "stage-1 tile does not fit in LDS (delay spread too large)");
    }
    Stage2Plan s2plan;
    bool fit2 = choose_stage2(p, d2min, d2span, s2plan, (budget2 / 8) & ~3, d->s2_shape);
    if (!fit2 && d->s2_lds == 0 && b->max_lds > budget2) {
        // rows whose DM step moves the low sub-bands by thousands of samples
        // (high time resolution, high DM): one workgroup per CU, all its LDS
        budget2 = b->max_lds;
        fit2 = choose_stage2(p, d2min, d2span, s2plan, (budget2 / 8) & ~3, d->s2_shape);
    }
    if (!fit2) {
        delete p;
        return fail(TX_EINVAL, "stage-2 tile does not fit in LDS (delay spread too large)");
    }
    // >= 4096 so that a stage-1 tile (<= 2048 samples) wraps the ring at most once
    p->Lsub = std::max<long>(round_up(N + pm.D2 + 64, 64), 4096);
    p->half2 = (budget2 / 8) & ~3; // floats per buffer (a multiple of 4): two buffers in budget2 bytes
    p->pstride2 = s2plan.pstride;
    p->maxrow2 = s2plan.maxrow;
    // a stage-2 row is read contiguously from its start slot
    p->Wsub = round_up(s2plan.maxrow + 64, 64);
    p->spitch = p->Lsub + p->Wsub;

    int rc = TX_OK;
#define TRY(x)                                                                                     \
    do {                                                                                           \
        rc = (x);                                                                                  \
        if (rc) {                                                                                  \
            plan_free(p);                                                                          \
            delete p;                                                                              \
            return rc;                                                                             \
        }                                                                                          \
    } while (0)
#define ALLOC(ptr, bytes)                                                                          \
    do {                                                                                           \
        if (hipMalloc((void **)&(ptr), (bytes)) != hipSuccess) {                                   \
            plan_free(p);                                                                          \
            delete p;                                                                              \
            return fail(TX_ENOMEM, "device allocation of " #ptr);                                  \
        }                                                                                          \
    } while (0)
    {
        // stage 1 reads the coarse-DM delays of one channel together:
        // [nchans][nkc*NCC], the tail of the last chunk padded with that
        // chunk's lowest delay (computed, never stored)
        const int pitch = p->nkc * p->NCC;
        std::vector<int> d1t((size_t)pm.nchans * pitch);
        for (int j = 0; j < pm.nchans; j++)
            for (int k = 0; k < pitch; k++)
                d1t[(size_t)j * pitch + k] =
                    k < pm.ndmc ? pm.d1[(size_t)k * pm.nchans + j]
                                : d1min[(size_t)(k / p->NCC) * pm.nchans + j];
        TRY(upload(&p->d_d1, d1t));
        if (b->dtype == TX_DTYPE_U8) {
            // per (channel, coarse DM): v_perm_b32 selector picking bytes
            // (e, e+2) of the 8-byte window at byte e&3, and the byte offset
            // of that window's first word, e = delay - row start (>= 0)
            std::vector<int> sel((size_t)pm.nchans * pitch * 2);
            for (int j = 0; j < pm.nchans; j++)
                for (int k = 0; k < pitch; k++) {
                    const int e = d1t[(size_t)j * pitch + k] - d1min[(size_t)(k / p->NCC) * pm.nchans + j];
                    const uint32_t bb = (uint32_t)(e & 3);
                    sel[((size_t)j * pitch + k) * 2] = (int)(0x0C020C00u + bb * 0x00010001u);
                    sel[((size_t)j * pitch + k) * 2 + 1] = e & ~3;
                }
            TRY(upload(&p->d_d1sel, sel));
        }
    }
    TRY(upload(&p->d_d1min, d1min));
    TRY(upload(&p->d_d2, pm.d2));
    {
        // stage 2 reads, per (DM block, sub-band, wave), the DPW delays of
        // its DM slots relative to the block's lowest delay; slots past the
        // end of the group hold 0 (computed, never stored)
        const int blk = 8 * p->DPW;
        std::vector<int> off((size_t)pm.ndmc * p->nblk * pm.nsub * blk, 0);
        for (int k = 0; k < pm.ndmc; k++) {
            const long gend = std::min<long>((long)(k + 1) * pm.group, pm.ndm);
            for (int b = 0; b < p->nblk; b++) {
                const size_t w = (size_t)k * p->nblk + b;
                for (int s = 0; s < pm.nsub; s++)
                    for (int wave = 0; wave < 8; wave++)
                        for (int d = 0; d < p->DPW; d++) {
                            const long i = (long)k * pm.group + (long)b * blk + wave + 8 * d;
                            if (i < gend)
                                off[((w * pm.nsub + s) * 8 + wave) * p->DPW + d] =
                                    pm.d2[(size_t)i * pm.nsub + s] - d2min[w * pm.nsub + s];
                        }
            }
        }
        TRY(upload(&p->d_d2off, off));
    }
    TRY(upload(&p->d_d2min, d2min));
    TRY(upload(&p->d_d2span, d2span));
    TRY(upload(&p->d_ldsoff, s2plan.ldsoff));
    TRY(upload(&p->d_npass, s2plan.npass));
    TRY(upload(&p->d_pstart, s2plan.pstart));
    ALLOC(p->d_vwn, kMaxBox * sizeof(int));
    ALLOC(p->d_whi, kMaxBox * sizeof(int));
    if (d->nbox > 0) TRY(set_widths(p, d->vwn, d->nbox));
    TRY(upload(&p->d_fcnt, fcntf));
    const size_t subbytes = (size_t)pm.ndmc * pm.nsub * p->spitch * 4;
    ALLOC(p->d_sub, subbytes);
    if (hipMemset(p->d_sub, 0, subbytes) != hipSuccess) TRY(fail(TX_EHIP, "sub ring clear"));
    // + 64 bytes: the boxcar loads a row's last samples as 16-byte lines,
    // which may reach 3 floats past the last row
    ALLOC(p->d_tim, (size_t)pm.ndm * N * 4 + 64);
    p->ntile2 = (int)((N + 64 * p->R - 1) / (64 * p->R));
    ALLOC(p->d_tstat, (size_t)pm.ndm * p->ntile2 * 2 * sizeof(double));
    if (p->store_maps) {
        ALLOC(p->d_mxS, (size_t)pm.ndm * N * 4);
        ALLOC(p->d_mxwn, (size_t)pm.ndm * N * 4);
    }
    if (!boxcar_fits_lds(N, b->max_lds)) ALLOC(p->d_scratch, (size_t)pm.ndm * N * 4);
    if (p->exact) {
        ALLOC(p->d_xstat, (size_t)pm.ndm * 2 * 4);
        ALLOC(p->d_xck, (size_t)pm.ndm * boxcar_xsegments(N) * 4);
    }
    ALLOC(p->d_pts, (size_t)p->cap * sizeof(tx_point));
    ALLOC(p->d_npts, sizeof(unsigned long long));
    ALLOC(p->d_tmp, (size_t)pm.nsub * N * 4);
#undef TRY
#undef ALLOC
    b->need_hist = std::max(b->need_hist, pm.D1 + 8);
    b->need_win = std::max(b->need_win, p->read1 + 8);
    b->nplans++;
    b->plans.push_back(p);
    if (getenv("TX_VERBOSE"))
        fprintf(stderr,
                "txgpu plan: ndm %ld nsub %d group %d ndmc %d D1 %ld D2 %ld | stage1 NCC %d Q %d "
                "CH %d RS %d | stage2 DPW %d R %d nblk %d passes/block %d maxrow %d | Lsub %ld Wsub %ld\n",
                pm.ndm, pm.nsub, pm.group, pm.ndmc, pm.D1, pm.D2, p->NCC, p->Q, p->CH, p->RS,
                p->DPW, p->R, p->nblk, p->pstride2 - 1, p->maxrow2, p->Lsub, p->Wsub);
    *out = p;
    return TX_OK;
}

extern "C" int tx_stage2_shapes(int *shapes, int n)
{
    if (shapes)
        for (int c = 0; c < std::min(n, kNumStage2Shapes); c++) {
            shapes[2 * c] = kStage2Shapes[c][0];
            shapes[2 * c + 1] = kStage2Shapes[c][1];
        }
    return kNumStage2Shapes;
}

extern "C" int tx_plan_destroy(tx_plan *p)
{
    if (!p) return TX_OK;
    hipSetDevice(p->beam->device);
    hipStreamSynchronize(p->beam->stream);
    beam_drop_graph(p->beam);
    plan_free(p);
    p->beam->nplans--;
    auto &v = p->beam->plans;
    v.erase(std::remove(v.begin(), v.end(), p), v.end());
    delete p;
    return TX_OK;
}

extern "C" int tx_get_info(const tx_plan *p, tx_info *o)
{
    if (!p || !o) return fail(TX_EINVAL, "null argument");
    const PlanMath &pm = p->pm;
    o->counter = p->counter;
    o->offset = pm.offset;
    o->noverlap = pm.noverlap;
    o->ndump = pm.ndump;
    o->nsamples = pm.nsamples;
    o->ndm = pm.ndm;
    o->nsub = pm.nsub;
    o->cps = pm.cps;
    o->group = pm.group;
    o->ndmc = pm.ndmc;
    o->D1 = pm.D1;
    o->D2 = pm.D2;
    o->maxdelay = pm.maxdelay;
    o->fmin = pm.fmin;
    o->fmax = pm.fmax;
    o->maxwn = p->vwn.empty() ? 0 : p->vwn.back();
    o->dms = pm.dms;
    o->ddm = pm.ddm;
    o->tsamp = pm.tsamp;
    return TX_OK;
}

extern "C" int tx_get_tables(const tx_plan *p, int *d1, int *d2, double *dmc, double *fref,
                             double *fsub, double *fcnt)
{
    if (!p) return fail(TX_EINVAL, "null argument");
    const PlanMath &pm = p->pm;
    if (d1) std::copy(pm.d1.begin(), pm.d1.end(), d1);
    if (d2) std::copy(pm.d2.begin(), pm.d2.end(), d2);
    if (dmc) std::copy(pm.dmc.begin(), pm.dmc.end(), dmc);
    if (fref) std::copy(pm.fref.begin(), pm.fref.end(), fref);
    if (fsub) std::copy(pm.fsub.begin(), pm.fsub.end(), fsub);
    if (fcnt) std::copy(p->fcnt.begin(), p->fcnt.end(), fcnt);
    return TX_OK;
}

// ------------------------------------------------------------------- run
// Kernel arguments of one block at counter c (after the block's push).
static void stage1_args(tx_plan *p, long c, Stage1Args &a1)
{
    tx_beam *b = p->beam;
    const PlanMath &pm = p->pm;
    const bool u8 = b->dtype == TX_DTYPE_U8;
    a1 = Stage1Args{};
    a1.ring = b->d_ring;
    a1.pitch = b->pitch;
    a1.Lin = b->L;
    a1.d1 = p->d_d1;
    a1.d1min = p->d_d1min;
    a1.d1sel = p->d_d1sel;
    a1.d1pitch = p->nkc * p->NCC;
    a1.C = pm.nchans;
    a1.cps = pm.cps;
    a1.nsub = pm.nsub;
    a1.ndmc = pm.ndmc;
    a1.zp_cps = u8 ? pm.cps * b->zp : 0;
    a1.tau_begin = c - pm.ndump - pm.D1;
    a1.tau_end = c - pm.D1;
    a1.tau_base = a1.tau_begin - lmod(a1.tau_begin, 4);
    a1.RS = p->RS;
    a1.CH = p->CH;
    if (b->fused) {
        a1.CH = p->CH_fe;
        a1.fe_off = b->d_fe_offm;
        a1.fe_gain = b->d_fe_gainm;
    }
    a1.sub = p->d_sub;
    a1.spitch = p->spitch;
    a1.Lsub = p->Lsub;
    a1.Wsub = p->Wsub;
    a1.clock = nullptr;
    a1.ndump = pm.ndump;
    a1.D1 = pm.D1;
}

static void stage2_args(tx_plan *p, long c, Stage2Args &a2)
{
    const PlanMath &pm = p->pm;
    a2 = Stage2Args{};
    a2.sub = p->d_sub;
    a2.spitch = p->spitch;
    a2.Lsub = p->Lsub;
    a2.d2 = p->d_d2;
    a2.d2min = p->d_d2min;
    a2.d2span = p->d_d2span;
    a2.d2off = p->d_d2off;
    a2.nsub = pm.nsub;
    a2.ndm = pm.ndm;
    a2.group = pm.group;
    a2.nblk = p->nblk;
    a2.ndmc = pm.ndmc;
    a2.w0 = c - pm.maxdelay - pm.nsamples;
    a2.N = pm.nsamples;
    a2.ldsoff = p->d_ldsoff;
    a2.npass = p->d_npass;
    a2.pstart = p->d_pstart;
    a2.pstride = p->pstride2;
    a2.half = p->half2;
    a2.tim = p->d_tim;
    a2.tstat = p->d_tstat;
    a2.clock = nullptr;
    a2.ndump = pm.ndump;
    a2.maxdelay = pm.maxdelay;
}

static void boxcar_args(tx_plan *p, float thre, BoxcarArgs &a)
{
    const PlanMath &pm = p->pm;
    a = BoxcarArgs{};
    a.tim = p->d_tim;
    a.N = pm.nsamples;
    a.ndm = pm.ndm;
    a.vwn = p->d_vwn;
    a.whi = p->d_whi;
    a.nbox = (int)p->vwn.size();
    a.lo_all = p->lo_all;
    a.tstat = p->d_tstat;
    a.ntile = p->ntile2;
    a.Hl = 0;
    a.Hr = 0;
    for (int wn : p->vwn) {
        a.Hl = std::max(a.Hl, wn / 2 + 1);
        a.Hr = std::max(a.Hr, (wn - 1) / 2);
    }
    a.hi_all = p->hi_all;
    a.thre = thre;
    a.mxS = p->d_mxS;
    a.mxwn = p->d_mxwn;
    a.store_maps = p->store_maps;
    a.pts = p->d_pts;
    a.npts = p->d_npts;
    a.cap = p->cap;
    a.exact = p->exact;
    a.iqr = p->iqr;
    a.scratch = p->d_scratch;
    // exact order with the mean/variance statistics: the sequential passes
    // run in k_boxcar_chain (IQR statistics keep them in the boxcar)
    const bool chain = a.exact && !a.iqr && p->d_xck;
    a.xstat = chain ? p->d_xstat : nullptr;
    a.xck = chain ? p->d_xck : nullptr;
    a.xpitch = boxcar_xsegments(pm.nsamples);
    a.pts_ex = nullptr;
    a.block = (int)(p->counter / pm.ndump);
    if (p->in_session) {
        a.pts_ex = p->d_sess;
        a.npts = p->d_sess_n;
        a.cap = (long)p->sess_cap;
    }
    a.clock = nullptr;
    a.ndump = pm.ndump;
    a.ready_at = pm.offset - pm.noverlap + pm.ndump;
    // the register window needs the ladder to start with kNarrowWidths and
    // 12 samples of reach left, 8 right, of every chunk
    a.window = !p->window_off && !a.exact && !a.iqr && a.nbox >= kNarrow && a.Hl >= 12 && a.Hr >= 8;
    for (int k = 0; a.window && k < kNarrow; k++)
        if (p->vwn[k] != kNarrowWidths[k]) a.window = 0;
}

extern "C" int tx_dedisp_run(tx_plan *p)
{
    if (!p) return fail(TX_EINVAL, "null argument");
    tx_beam *b = p->beam;
    const PlanMath &pm = p->pm;
    if (b->counter != p->counter + pm.ndump)
        return fail(TX_ESTATE, "tx_dedisp_run needs exactly one new block on the beam");
    HIPCHK(hipSetDevice(b->device));
    const long c = b->counter;
    const bool u8 = b->dtype == TX_DTYPE_U8;

    Stage1Args a1;
    stage1_args(p, c, a1);
    int rc;
    {
        KernelTimer kt(b, 1);
        rc = u8 ? launch_stage1_u8(b->stream, a1, p->NCC, p->Q)
             : b->fused ? launch_stage1_fe(b->stream, a1, p->NCC, p->Q)
                        : launch_stage1_f32(b->stream, a1, p->NCC, p->Q);
    }
    if (rc) return fail(TX_EINVAL, "no stage-1 kernel for this shape");
    HIPCHK(hipGetLastError());

    Stage2Args a2;
    stage2_args(p, c, a2);
    {
        KernelTimer kt(b, 2);
        if (launch_stage2(b->stream, a2, p->DPW, p->R)) return fail(TX_EINVAL, "no stage-2 kernel");
    }
    HIPCHK(hipGetLastError());
    p->w0 = a2.w0;
    p->counter = c;
    p->have_points = false;
    return TX_OK;
}

__global__ static void k_zero_u64(unsigned long long *p) { *p = 0ull; }

extern "C" int tx_boxcar_run(tx_plan *p, float thre)
{
    if (!p) return fail(TX_EINVAL, "null argument");
    const PlanMath &pm = p->pm;
    // src/singlepulse/boxcar.cpp:161
    if (p->counter < pm.offset - pm.noverlap + pm.ndump) return TX_NOT_READY;
    if (p->vwn.empty()) return fail(TX_ESTATE, "no width ladder: create the plan with vwn or call tx_boxcar_widths");
    tx_beam *b = p->beam;
    HIPCHK(hipSetDevice(b->device));
    if (!p->in_session) hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, b->stream, p->d_npts);
    BoxcarArgs a;
    boxcar_args(p, thre, a);
    {
        KernelTimer kt(b, 3);
        if (launch_boxcar(b->stream, a, b->max_lds)) return fail(TX_EINVAL, "boxcar launch shape");
    }
    HIPCHK(hipGetLastError());
    p->have_points = !p->in_session;
    return TX_OK;
}

extern "C" int tx_boxcar_widths(tx_plan *p, const int *vwn, int nbox, int iqr)
{
    if (!p) return fail(TX_EINVAL, "null argument");
    HIPCHK(hipSetDevice(p->beam->device));
    const int rc = set_widths(p, vwn, nbox);
    if (rc) return rc;
    p->iqr = iqr != 0;
    return TX_OK;
}

extern "C" int tx_boxcar_points(tx_plan *p, tx_point *out, size_t cap, size_t *n)
{
    if (!p || !n) return fail(TX_EINVAL, "null argument");
    if (!p->have_points) return fail(TX_ESTATE, "no tx_boxcar_run since the last tx_dedisp_run");
    tx_beam *b = p->beam;
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    unsigned long long tot = 0;
    HIPCHK(hipMemcpy(&tot, p->d_npts, sizeof(tot), hipMemcpyDeviceToHost));
    *n = (size_t)tot;
    const size_t have = (size_t)std::min<unsigned long long>(tot, (unsigned long long)p->cap);
    std::vector<tx_point> tmp(have);
    if (have) HIPCHK(hipMemcpy(tmp.data(), p->d_pts, have * sizeof(tx_point), hipMemcpyDeviceToHost));
    std::sort(tmp.begin(), tmp.end(), [](const tx_point &x, const tx_point &y) {
        return x.idm != y.idm ? x.idm < y.idm : x.isamp < y.isamp;
    });
    const size_t m = std::min(have, cap);
    if (out && m) std::memcpy(out, tmp.data(), m * sizeof(tx_point));
    if (tot > (unsigned long long)p->cap) return fail(TX_EOVERFLOW, "device point list overflowed");
    if (have > cap) return fail(TX_EOVERFLOW, "caller buffer too small");
    return TX_OK;
}

extern "C" int tx_get_timdata(tx_plan *p, int idm, int with_overlap, float *out)
{
    if (!p || !out) return fail(TX_EINVAL, "null argument");
    const PlanMath &pm = p->pm;
    if (idm < 0 || idm >= pm.ndm) return fail(TX_ERANGE, "idm out of range");
    HIPCHK(hipSetDevice(p->beam->device));
    HIPCHK(hipStreamSynchronize(p->beam->stream));
    const long off = with_overlap ? 0 : pm.noverlap;
    const long n = with_overlap ? pm.nsamples : pm.ndump;
    HIPCHK(hipMemcpy(out, p->d_tim + (long)idm * pm.nsamples + off, n * 4, hipMemcpyDeviceToHost));
    return TX_OK;
}

extern "C" int tx_get_subdata(tx_plan *p, int idm, float *out)
{
    if (!p || !out) return fail(TX_EINVAL, "null argument");
    const PlanMath &pm = p->pm;
    if (idm < 0 || idm >= pm.ndm) return fail(TX_ERANGE, "idm out of range");
    tx_beam *b = p->beam;
    HIPCHK(hipSetDevice(b->device));
    launch_subdata(b->stream, p->d_sub, p->spitch, p->Lsub, p->d_d2, pm.nsub, pm.group, idm, p->w0,
                   pm.nsamples, p->d_tmp);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(out, p->d_tmp, (size_t)pm.nsub * pm.nsamples * 4, hipMemcpyDeviceToHost));
    return TX_OK;
}

extern "C" int tx_boxcar_rows(tx_plan *p, int dm0, int dm1, float *mxS, int *mxwn)
{
    if (!p) return fail(TX_EINVAL, "null argument");
    const PlanMath &pm = p->pm;
    if (dm0 < 0 || dm1 > pm.ndm || dm0 > dm1) return fail(TX_ERANGE, "DM rows out of range");
    HIPCHK(hipSetDevice(p->beam->device));
    const size_t n = (size_t)(dm1 - dm0) * pm.nsamples;
    if (!p->store_maps) {
        // points-only plan: the rows of the current window recomputed by the
        // boxcar kernels with stored maps, on just these rows (the same
        // arithmetic per row, so the same values a stored map would hold)
        if (p->counter < pm.offset - pm.noverlap + pm.ndump)
            return fail(TX_ESTATE, "tx_boxcar_rows: the boxcar gate is still closed");
        if (p->vwn.empty()) return fail(TX_ESTATE, "no width ladder");
        if (n == 0) return TX_OK;
        tx_beam *b = p->beam;
        if (p->row_cap < n) {
            HIPCHK(hipStreamSynchronize(b->stream));
            hipFree(p->d_rowS);
            hipFree(p->d_rowW);
            p->d_rowS = nullptr;
            p->d_rowW = nullptr;
            p->row_cap = 0;
            if (hipMalloc((void **)&p->d_rowS, n * 4) != hipSuccess ||
                hipMalloc((void **)&p->d_rowW, n * 4) != hipSuccess)
                return fail(TX_ENOMEM, "row allocation");
            p->row_cap = n;
        }
        if (!p->d_rown && hipMalloc((void **)&p->d_rown, sizeof(unsigned long long)) != hipSuccess)
            return fail(TX_ENOMEM, "row allocation");
        BoxcarArgs a;
        boxcar_args(p, INFINITY, a); // no point passes
        a.tim = p->d_tim + (size_t)dm0 * pm.nsamples;
        a.tstat = p->d_tstat + (size_t)dm0 * p->ntile2 * 2;
        if (a.scratch) a.scratch = p->d_scratch + (size_t)dm0 * pm.nsamples;
        if (a.xck) {
            a.xstat += (size_t)dm0 * 2;
            a.xck += (size_t)dm0 * a.xpitch;
        }
        a.ndm = dm1 - dm0;
        a.mxS = p->d_rowS;
        a.mxwn = p->d_rowW;
        a.store_maps = 1;
        a.pts = nullptr;
        a.pts_ex = nullptr;
        a.npts = p->d_rown;
        a.cap = 0;
        if (launch_boxcar(b->stream, a, b->max_lds)) return fail(TX_EINVAL, "boxcar launch shape");
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(b->stream));
        if (mxS) HIPCHK(hipMemcpy(mxS, p->d_rowS, n * 4, hipMemcpyDeviceToHost));
        if (mxwn) HIPCHK(hipMemcpy(mxwn, p->d_rowW, n * 4, hipMemcpyDeviceToHost));
        return TX_OK;
    }
    HIPCHK(hipStreamSynchronize(p->beam->stream));
    if (mxS) HIPCHK(hipMemcpy(mxS, p->d_mxS + (size_t)dm0 * pm.nsamples, n * 4, hipMemcpyDeviceToHost));
    if (mxwn)
        HIPCHK(hipMemcpy(mxwn, p->d_mxwn + (size_t)dm0 * pm.nsamples, n * 4, hipMemcpyDeviceToHost));
    return TX_OK;
}

extern "C" int tx_cand_spectra(tx_plan *p, int ncand, const int *idm, const int *isamp,
                               const int *wn, float *out)
{
    if (!p || ncand < 0) return fail(TX_EINVAL, "bad argument");
    if (ncand == 0) return TX_OK;
    if (!idm || !isamp || !wn || !out) return fail(TX_EINVAL, "null argument");
    const PlanMath &pm = p->pm;
    for (int c = 0; c < ncand; c++)
        if (idm[c] < 0 || idm[c] >= pm.ndm || isamp[c] < 0 || isamp[c] >= pm.nsamples || wn[c] < 1)
            return fail(TX_ERANGE, "candidate out of range");
    tx_beam *b = p->beam;
    HIPCHK(hipSetDevice(b->device));
    if (ncand > p->cand_cap) {
        hipFree(p->d_cand);
        hipFree(p->d_spec);
        p->d_cand = nullptr;
        p->d_spec = nullptr;
        p->cand_cap = 0;
        if (hipMalloc((void **)&p->d_cand, (size_t)ncand * 3 * sizeof(int)) != hipSuccess ||
            hipMalloc((void **)&p->d_spec, (size_t)ncand * pm.nsub * 4) != hipSuccess)
            return fail(TX_ENOMEM, "candidate buffers");
        p->cand_cap = ncand;
    }
    std::vector<int> h((size_t)ncand * 3);
    std::copy(idm, idm + ncand, h.begin());
    std::copy(isamp, isamp + ncand, h.begin() + ncand);
    std::copy(wn, wn + ncand, h.begin() + 2 * ncand);
    HIPCHK(hipMemcpyAsync(p->d_cand, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice,
                          b->stream));
    launch_cand_spectra(b->stream, p->d_sub, p->spitch, p->Lsub, p->d_d2, p->d_fcnt, pm.nsub,
                        pm.group, p->w0, pm.nsamples, ncand, p->d_cand, p->d_cand + ncand,
                        p->d_cand + 2 * ncand, p->d_spec);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(out, p->d_spec, (size_t)ncand * pm.nsub * 4, hipMemcpyDeviceToHost));
    return TX_OK;
}

extern "C" int tx_device_buffers(tx_plan *p, void **tim, void **mxS, void **mxwn)
{
    if (!p) return fail(TX_EINVAL, "null argument");
    if (tim) *tim = p->d_tim;
    if (mxS) *mxS = p->d_mxS;
    if (mxwn) *mxwn = p->d_mxwn;
    return TX_OK;
}

extern "C" int tx_dump_series(tx_plan *p, float mean, float std, int nbits, void *out, size_t cap, long *tau0,
                              long *n)
{
    if (!p || !out || !tau0 || !n) return fail(TX_EINVAL, "null argument");
    if (nbits != 8 && nbits != 16 && nbits != 32) return fail(TX_EINVAL, "nbits must be 8, 16 or 32");
    if (!std::isfinite(mean) || !std::isfinite(std)) return fail(TX_EINVAL, "mean and std must be finite");
    const PlanMath &pm = p->pm;
    *tau0 = p->w0 + pm.noverlap;
    *n = 0;
    if (p->w0 + pm.nsamples <= 0) return TX_OK; // every new sample precedes the observation
    const size_t need = (size_t)pm.ndm * pm.ndump * (nbits / 8);
    if (cap < need) return fail(TX_EOVERFLOW, "tx_dump_series: out holds fewer than ndm * ndump samples");
    double nlive = 0.;
    for (double f : p->fcnt) nlive += f;
    const float scale = nlive > 0. ? (float)((double)std / std::sqrt(nlive)) : 0.f;
    tx_beam *b = p->beam;
    HIPCHK(hipSetDevice(b->device));
    if (!p->d_dump && hipMalloc(&p->d_dump, (size_t)pm.ndm * pm.ndump * 4) != hipSuccess) {
        p->d_dump = nullptr;
        return fail(TX_ENOMEM, "dump staging allocation");
    }
    launch_dump(b->stream, p->d_tim, pm.nsamples, pm.noverlap, pm.ndump, pm.ndm, scale, mean, nbits, p->d_dump);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, p->d_dump, need, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    *n = pm.ndump;
    return TX_OK;
}

// ------------------------------------------------------------ streaming extras
extern "C" int tx_beam_reset(tx_beam *b)
{
    if (!b) return fail(TX_EINVAL, "null argument");
    HIPCHK(hipSetDevice(b->device));
    if (b->d_ring) {
        const size_t bytes = (size_t)b->C * b->pitch * b->esize();
        HIPCHK(hipMemsetAsync(b->d_ring, b->dtype == TX_DTYPE_U8 ? b->zp : 0, bytes, b->stream));
    }
    for (tx_plan *p : b->plans) {
        const size_t subbytes = (size_t)p->pm.ndmc * p->pm.nsub * p->spitch * 4;
        HIPCHK(hipMemsetAsync(p->d_sub, 0, subbytes, b->stream));
        p->counter = 0;
        p->have_points = false;
    }
    b->counter = 0;
    if (b->d_clock) {
        HIPCHK(hipMemsetAsync(b->d_clock, 0, sizeof(long), b->stream));
        b->gcounter = 0;
    }
    return TX_OK;
}

// ------------------------------------------------------------ graph pipeline
extern "C" int tx_beam_graph(tx_beam *b, float thre)
{
    if (!b) return fail(TX_EINVAL, "null argument");
    if (b->plans.empty()) return fail(TX_ESTATE, "tx_beam_graph: the beam has no plans");
    if (!b->rfi.empty()) return fail(TX_ESTATE, "tx_beam_graph: the RFI filters run on direct pushes only");
    for (tx_plan *p : b->plans) {
        if (!p->in_session) return fail(TX_ESTATE, "tx_beam_graph: every plan needs a session (tx_session_begin)");
        if (p->vwn.empty()) return fail(TX_ESTATE, "tx_beam_graph: a plan has no width ladder");
        if (p->counter != b->counter) return fail(TX_ESTATE, "tx_beam_graph: a plan lags the beam");
    }
    HIPCHK(hipSetDevice(b->device));
    if (!b->d_ring) {
        const int rc = beam_alloc_ring(b);
        if (rc) return rc;
    }
    {
        const int rc = beam_stage(b);
        if (rc) return rc;
    }
    if (!b->d_clock && hipMalloc((void **)&b->d_clock, sizeof(long)) != hipSuccess)
        return fail(TX_ENOMEM, "clock allocation");
    beam_drop_graph(b);
    // the clock starts at the beam's counter
    const long c0 = b->counter;
    HIPCHK(hipMemcpyAsync(b->d_clock, &c0, sizeof(long), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));

    // nodes: transpose, then per plan stage 1 -> stage 2 -> boxcar (plans
    // are independent branches), then the clock tick after every boxcar
    LaunchSink sink;
    beam_launch_input(b, b->d_stage, 0, b->d_clock, &sink);
    const bool u8 = b->dtype == TX_DTYPE_U8;
    b->gsess.clear();
    b->gcap.clear();
    std::vector<size_t> span; // plan j's nodes: [span[j], span[j + 1])
    for (tx_plan *p : b->plans) {
        span.push_back(sink.nodes.size());
        Stage1Args a1;
        stage1_args(p, 0, a1);
        a1.clock = b->d_clock;
        if (u8 ? launch_stage1_u8(b->stream, a1, p->NCC, p->Q, &sink)
               : b->fused ? launch_stage1_fe(b->stream, a1, p->NCC, p->Q, &sink)
                          : launch_stage1_f32(b->stream, a1, p->NCC, p->Q, &sink))
            return fail(TX_EINVAL, "no stage-1 kernel for this shape");
        Stage2Args a2;
        stage2_args(p, 0, a2);
        a2.clock = b->d_clock;
        if (launch_stage2(b->stream, a2, p->DPW, p->R, &sink)) return fail(TX_EINVAL, "no stage-2 kernel");
        BoxcarArgs a;
        boxcar_args(p, thre, a);
        a.clock = b->d_clock;
        if (launch_boxcar(b->stream, a, b->max_lds, &sink)) return fail(TX_EINVAL, "boxcar launch shape");
        b->gsess.push_back(p->d_sess);
        b->gcap.push_back(p->sess_cap);
    }
    span.push_back(sink.nodes.size());
    launch_tick(b->stream, b->d_clock, b->ndump, &sink);
    const size_t np = b->plans.size();
    if (sink.nodes.size() != span.back() + 1 || span[0] != 1) return fail(TX_EHIP, "tx_beam_graph: unexpected nodes");

    HIPCHK(hipGraphCreate(&b->graph, 0));
    std::vector<hipGraphNode_t> h(sink.nodes.size());
    auto add = [&](size_t i, const hipGraphNode_t *deps, size_t ndeps) -> hipError_t {
        KNode &n = sink.nodes[i];
        std::vector<void *> prm = n.params();
        hipKernelNodeParams kp{};
        kp.func = (void *)n.func;
        kp.gridDim = n.grid;
        kp.blockDim = n.block;
        kp.sharedMemBytes = n.shm;
        kp.kernelParams = prm.data();
        kp.extra = nullptr;
        return hipGraphAddKernelNode(&h[i], b->graph, deps, ndeps, &kp);
    };
    HIPCHK(add(0, nullptr, 0));
    std::vector<hipGraphNode_t> ends;
    for (size_t j = 0; j < np; j++) {
        // a plan's kernels in order (stage 1, stage 2, the boxcar's one or two)
        if (span[j + 1] <= span[j]) return fail(TX_EHIP, "tx_beam_graph: a plan without nodes");
        HIPCHK(add(span[j], &h[0], 1));
        for (size_t i = span[j] + 1; i < span[j + 1]; i++) HIPCHK(add(i, &h[i - 1], 1));
        ends.push_back(h[span[j + 1] - 1]);
    }
    HIPCHK(add(sink.nodes.size() - 1, ends.data(), ends.size()));
    HIPCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
    b->gtrans = h[0];
    b->trans = sink.nodes[0];
    b->gcounter = c0;
    return TX_OK;
}

extern "C" int tx_beam_graph_push(tx_beam *b, const void *block, long ns, int on_device)
{
    if (!b || !block) return fail(TX_EINVAL, "null argument");
    if (!b->gexec) return fail(TX_ESTATE, "no graph: call tx_beam_graph");
    if (ns != beam_block_samples(b))
        return fail(TX_EINVAL, b->fe ? "a front-end block must hold exactly ndump*td raw samples"
                                     : "a block must hold exactly ndump samples");
    if (b->gcounter != b->counter)
        return fail(TX_ESTATE, "the beam moved outside its graph (tx_beam_push): call tx_beam_graph again");
    if (b->gsess.size() != b->plans.size()) return fail(TX_ESTATE, "the beam's plans changed: call tx_beam_graph again");
    for (size_t j = 0; j < b->plans.size(); j++) {
        tx_plan *p = b->plans[j];
        if (!p->in_session || p->d_sess != b->gsess[j] || p->sess_cap != b->gcap[j] || p->counter != b->counter)
            return fail(TX_ESTATE, "a plan's session changed since tx_beam_graph: call it again");
    }
    HIPCHK(hipSetDevice(b->device));
    const void *src = block;
    int k = -1;
    if (on_device == TX_PUSH_HOST) {
        HIPCHK(hipMemcpyAsync(b->d_stage, block, beam_block_bytes(b), hipMemcpyHostToDevice, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
        src = b->d_stage;
    } else if (on_device == TX_PUSH_HOST_ASYNC) {
        const int rc = beam_copy_async(b, block, &src, &k);
        if (rc) return rc;
    }
    // the block's source is the transpose's first argument
    const void *cur = nullptr;
    std::memcpy(&cur, b->trans.blob.data() + b->trans.offs[0], sizeof(cur));
    if (cur != src) {
        std::memcpy(b->trans.blob.data() + b->trans.offs[0], &src, sizeof(src));
        std::vector<void *> prm = b->trans.params();
        hipKernelNodeParams kp{};
        kp.func = (void *)b->trans.func;
        kp.gridDim = b->trans.grid;
        kp.blockDim = b->trans.block;
        kp.sharedMemBytes = b->trans.shm;
        kp.kernelParams = prm.data();
        kp.extra = nullptr;
        HIPCHK(hipGraphExecKernelNodeSetParams(b->gexec, b->gtrans, &kp));
    }
    HIPCHK(hipGraphLaunch(b->gexec, b->stream));
    if (k >= 0) HIPCHK(hipEventRecord(b->ev_free[k], b->stream));
    b->counter += b->ndump;
    b->gcounter = b->counter;
    for (tx_plan *p : b->plans) {
        p->counter = b->counter;
        p->w0 = p->counter - p->pm.maxdelay - p->pm.nsamples;
        p->have_points = false;
    }
    return TX_OK;
}

extern "C" int tx_beam_stream(tx_beam *b, void **stream)
{
    if (!b || !stream) return fail(TX_EINVAL, "null argument");
    *stream = (void *)b->stream;
    return TX_OK;
}

extern "C" int tx_beam_set_timing(tx_beam *b, int enable)
{
    if (!b) return fail(TX_EINVAL, "null argument");
    b->timing = enable != 0;
    return TX_OK;
}

extern "C" int tx_beam_timing(tx_beam *b, double *ms, long *launches)
{
    if (!b) return fail(TX_EINVAL, "null argument");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    for (auto &p : b->pending) {
        float t = 0.f;
        if (hipEventElapsedTime(&t, p.a, p.b) == hipSuccess) {
            b->ms[p.cls] += t;
            b->launches[p.cls] += 1;
        }
        b->evfree.push_back(p.a);
        b->evfree.push_back(p.b);
    }
    b->pending.clear();
    for (int k = 0; k < 4; k++) {
        if (ms) ms[k] = b->ms[k];
        if (launches) launches[k] = b->launches[k];
        b->ms[k] = 0;
        b->launches[k] = 0;
    }
    return TX_OK;
}

// The session counter, zeroed on the beam's stream (a kernel: graph-safe).
static int session_counter(tx_plan *p)
{
    if (!p->d_sess_n && hipMalloc((void **)&p->d_sess_n, sizeof(unsigned long long)) != hipSuccess)
        return fail(TX_ENOMEM, "session counter allocation");
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, p->beam->stream, p->d_sess_n);
    HIPCHK(hipGetLastError());
    return TX_OK;
}

extern "C" int tx_session_begin(tx_plan *p, size_t cap)
{
    if (!p || cap == 0) return fail(TX_EINVAL, "bad argument");
    tx_beam *b = p->beam;
    HIPCHK(hipSetDevice(b->device));
    if (cap > p->own_cap) {
        beam_drop_graph(b);
        hipFree(p->d_own);
        p->d_own = nullptr;
        p->own_cap = 0;
        if (hipMalloc((void **)&p->d_own, cap * sizeof(tx_point_ex)) != hipSuccess)
            return fail(TX_ENOMEM, "session list allocation");
        p->own_cap = cap;
    }
    const int rc = session_counter(p);
    if (rc) return rc;
    p->d_sess = p->d_own;
    p->sess_cap = p->own_cap;
    p->h_sess = nullptr;
    p->in_session = true;
    return TX_OK;
}

extern "C" int tx_session_begin_host(tx_plan *p, tx_point_ex *host, size_t cap)
{
    if (!p || !host || cap == 0) return fail(TX_EINVAL, "bad argument");
    tx_beam *b = p->beam;
    HIPCHK(hipSetDevice(b->device));
    // only page-locked host memory the device can address: anything else is
    // refused here, before a kernel could store to it
    hipPointerAttribute_t at{};
    void *dptr = nullptr;
    if (hipPointerGetAttributes(&at, host) != hipSuccess || at.type != hipMemoryTypeHost ||
        hipHostGetDevicePointer(&dptr, host, 0) != hipSuccess || !dptr) {
        (void)hipGetLastError();
        return fail(TX_EINVAL, "tx_session_begin_host: the list must be page-locked host memory the device can "
                               "address (hipHostMalloc)");
    }
    const int rc = session_counter(p);
    if (rc) return rc;
    p->d_sess = (tx_point_ex *)dptr;
    p->sess_cap = cap;
    p->h_sess = host;
    p->in_session = true;
    return TX_OK;
}

extern "C" int tx_session_end(tx_plan *p, tx_point_ex *out, size_t cap, size_t *n)
{
    if (!p || !n) return fail(TX_EINVAL, "null argument");
    if (!p->in_session) return fail(TX_ESTATE, "no session in progress");
    tx_beam *b = p->beam;
    HIPCHK(hipSetDevice(b->device));
    p->in_session = false;
    unsigned long long tot = 0;
    HIPCHK(hipMemcpyAsync(&tot, p->d_sess_n, sizeof(tot), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    *n = (size_t)tot;
    const size_t have = (size_t)std::min<unsigned long long>(tot, p->sess_cap);
    const size_t m = std::min(have, cap);
    if (out && m && p->h_sess) {
        if (out != p->h_sess) std::memcpy(out, p->h_sess, m * sizeof(tx_point_ex)); // already on the host
    } else if (out && m) {
        HIPCHK(hipMemcpy(out, p->d_sess, m * sizeof(tx_point_ex), hipMemcpyDeviceToHost));
    }
    if (tot > p->sess_cap) return fail(TX_EOVERFLOW, "session point list overflowed");
    if (out && have > cap) return fail(TX_EOVERFLOW, "caller buffer too small");
    return TX_OK;
}

// ------------------------------------------------------------ batch queue
// Boxcar::run of the rows of many beams as one persistent launch on the
// batch's own stream: the launch waits for every beam's queued work and
// every beam's later work waits for the launch.
struct tx_batch {
    int device = 0;
    int ncu = 0;
    int max_lds = 65536;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    std::vector<hipEvent_t> evs; // one per distinct beam of a call
    // job table: pinned staging (two slots, each reused once its last copy
    // has completed) and its device copy; layout [BoxcarArgs jobs][long first]
    char *h_tab[2] = {nullptr, nullptr};
    hipEvent_t copied[2] = {nullptr, nullptr};
    size_t h_cap = 0;
    int slot = 0;
    char *d_tab = nullptr;
    size_t d_cap = 0;
    unsigned *d_next = nullptr;
    std::vector<BoxcarArgs> jobs;
    std::vector<long> first;
};

__global__ static void k_zero_u32(unsigned *p) { *p = 0u; }

extern "C" int tx_batch_create(int device, tx_batch **out)
{
    if (!out) return fail(TX_EINVAL, "null argument");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(TX_EINVAL, "bad device ordinal");
    HIPCHK(hipSetDevice(device));
    tx_batch *q = new tx_batch();
    q->device = device;
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess) q->ncu = v;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && v > 0)
        q->max_lds = v;
    bool ok = q->ncu > 0 && hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&q->done, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&q->copied[0], hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&q->copied[1], hipEventDisableTiming) == hipSuccess &&
              hipMalloc((void **)&q->d_next, sizeof(unsigned)) == hipSuccess;
    if (!ok) {
        tx_batch_destroy(q);
        return fail(TX_EHIP, "batch stream, events or counter");
    }
    *out = q;
    return TX_OK;
}

extern "C" int tx_batch_destroy(tx_batch *q)
{
    if (!q) return TX_OK;
    hipSetDevice(q->device);
    if (q->stream) hipStreamSynchronize(q->stream);
    for (auto ev : q->evs) hipEventDestroy(ev);
    if (q->done) hipEventDestroy(q->done);
    for (int s = 0; s < 2; s++) {
        if (q->copied[s]) hipEventDestroy(q->copied[s]);
        if (q->h_tab[s]) hipHostFree(q->h_tab[s]);
    }
    hipFree(q->d_tab);
    hipFree(q->d_next);
    if (q->stream) hipStreamDestroy(q->stream);
    delete q;
    return TX_OK;
}

extern "C" int tx_batch_stream(tx_batch *q, void **stream)
{
    if (!q || !stream) return fail(TX_EINVAL, "null argument");
    *stream = (void *)q->stream;
    return TX_OK;
}

static int batch_boxcar_launch(tx_batch *q, tx_plan *const *plans, int nplans, float thre, int *ready);

extern "C" int tx_batch_boxcar(tx_batch *q, tx_plan *const *plans, int nplans, float thre, int *ready)
{
    if (!q || (nplans > 0 && !plans) || nplans < 0) return fail(TX_EINVAL, "bad argument");
    HIPCHK(hipSetDevice(q->device));
    std::vector<tx_beam *> beams;
    for (int i = 0; i < nplans; i++) {
        tx_plan *p = plans[i];
        if (!p) return fail(TX_EINVAL, "null plan");
        if (p->beam->device != q->device) return fail(TX_EINVAL, "plan on another device than the batch");
        if (p->vwn.empty())
            return fail(TX_ESTATE, "no width ladder: create the plan with vwn or call tx_boxcar_widths");
        for (int k = 0; k < i; k++)
            if (plans[k] == p) return fail(TX_EINVAL, "a plan appears twice");
        if (std::find(beams.begin(), beams.end(), p->beam) == beams.end()) beams.push_back(p->beam);
    }
    // the launch follows every beam's queued work (stage 2 wrote tim)
    while (q->evs.size() < beams.size()) {
        hipEvent_t ev = nullptr;
        HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        q->evs.push_back(ev);
    }
    for (size_t k = 0; k < beams.size(); k++) {
        HIPCHK(hipEventRecord(q->evs[k], beams[k]->stream));
        HIPCHK(hipStreamWaitEvent(q->stream, q->evs[k], 0));
    }
    const int rc = batch_boxcar_launch(q, plans, nplans, thre, ready);
    // every beam's next work follows the launch (its tim and point lists),
    // also after a failure part-way
    HIPCHK(hipEventRecord(q->done, q->stream));
    for (tx_beam *b : beams) HIPCHK(hipStreamWaitEvent(b->stream, q->done, 0));
    return rc;
}

static int batch_boxcar_launch(tx_batch *q, tx_plan *const *plans, int nplans, float thre, int *ready)
{
    q->jobs.clear();
    q->first.assign(1, 0);
    size_t shm = 0;
    for (int i = 0; i < nplans; i++) {
        tx_plan *p = plans[i];
        const PlanMath &pm = p->pm;
        // src/singlepulse/boxcar.cpp:161
        const bool go = p->counter >= pm.offset - pm.noverlap + pm.ndump;
        if (ready) ready[i] = go ? 1 : 0;
        if (!go) continue;
        if (!p->in_session) hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, q->stream, p->d_npts);
        BoxcarArgs a;
        boxcar_args(p, thre, a);
        const size_t need = boxcar_seg_lds(a);
        if (need && (long)need + 4096 <= q->max_lds) {
            q->jobs.push_back(a);
            q->first.push_back(q->first.back() + boxcar_seg_items(a));
            shm = std::max(shm, need);
        } else if (launch_boxcar(q->stream, a, q->max_lds)) {
            return fail(TX_EINVAL, "boxcar launch shape");
        }
        p->have_points = !p->in_session;
    }
    HIPCHK(hipGetLastError());
    const int nj = (int)q->jobs.size();
    if (nj) {
        if (q->first.back() > 0x7fffffffL) return fail(TX_EINVAL, "too many boxcar work items in one batch");
        const size_t bytes = nj * sizeof(BoxcarArgs) + (nj + 1) * sizeof(long);
        if (bytes > q->h_cap) {
            for (int s = 0; s < 2; s++) {
                HIPCHK(hipEventSynchronize(q->copied[s]));
                hipHostFree(q->h_tab[s]);
                q->h_tab[s] = nullptr;
            }
            q->h_cap = 0;
            const size_t cap = std::max<size_t>(bytes * 2, 64 * 1024);
            for (int s = 0; s < 2; s++)
                if (hipHostMalloc((void **)&q->h_tab[s], cap, hipHostMallocDefault) != hipSuccess)
                    return fail(TX_ENOMEM, "pinned job table");
            q->h_cap = cap;
        }
        if (bytes > q->d_cap) {
            HIPCHK(hipStreamSynchronize(q->stream));
            hipFree(q->d_tab);
            q->d_tab = nullptr;
            q->d_cap = 0;
            if (hipMalloc((void **)&q->d_tab, q->h_cap) != hipSuccess) return fail(TX_ENOMEM, "job table");
            q->d_cap = q->h_cap;
        }
        const int s = q->slot;
        q->slot ^= 1;
        HIPCHK(hipEventSynchronize(q->copied[s])); // the slot's last copy has been read
        char *h = q->h_tab[s];
        std::memcpy(h, q->jobs.data(), nj * sizeof(BoxcarArgs));
        std::memcpy(h + nj * sizeof(BoxcarArgs), q->first.data(), (nj + 1) * sizeof(long));
        HIPCHK(hipMemcpyAsync(q->d_tab, h, bytes, hipMemcpyHostToDevice, q->stream));
        HIPCHK(hipEventRecord(q->copied[s], q->stream));
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, q->stream, q->d_next);
        BoxcarQueue bq;
        bq.jobs = (const BoxcarArgs *)q->d_tab;
        bq.first = (const long *)(q->d_tab + nj * sizeof(BoxcarArgs));
        bq.njobs = nj;
        bq.total = q->first.back();
        bq.next = q->d_next;
        const int fit = boxcar_queue_grid(shm, q->ncu);
        if (fit < 1) return fail(TX_EHIP, "occupancy of the boxcar queue kernel");
        const long grid = std::min<long>(fit, bq.total);
        launch_boxcar_queue(q->stream, bq, shm, (int)grid);
        HIPCHK(hipGetLastError());
    }
    return TX_OK;
}
