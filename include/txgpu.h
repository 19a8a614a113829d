/*
 * txgpu.h -- C ABI of the MI355X (gfx950) single-pulse search hot path.
 *
 * Plain pointers and sizes only.  Every call returns an int status:
 * TX_OK (0) on success, a negative TX_E* code on error (tx_last_error()
 * gives the text; nothing throws across this boundary), or TX_NOT_READY
 * (1) where the reference returns `false` for "not ready"
 * (src/singlepulse/boxcar.cpp:161, src/singlepulse/cluster.cpp:66).
 *
 * Objects:
 *   tx_beam  one input stream (one filterbank beam): the device-resident,
 *            channel-major ring of the most recent input samples, shared by
 *            every ddplan row searched on that beam, and the HIP stream
 *            their work runs on.
 *   tx_plan  one ddplan row: RealTime::SubbandDedispersion + Boxcar of one
 *            SinglePulse (include/singlepulse.h:44-46).  Owns its delay
 *            tables, sub-band ring, dedispersed series and S/N maps.
 *
 * Threading: calls on one beam (and its plans) must not overlap; distinct
 * beams may be driven from distinct host threads.
 */
#ifndef TXGPU_H
#define TXGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TX_OK 0
#define TX_NOT_READY 1
#define TX_EINVAL (-1)
#define TX_ENOMEM (-2)
#define TX_ERANGE (-3)
#define TX_EHIP (-4)
#define TX_ESTATE (-5)
#define TX_EOVERFLOW (-6)
#define TX_EIO (-7)

#define TX_DTYPE_U8 0
#define TX_DTYPE_F32 1

typedef struct tx_beam tx_beam;
typedef struct tx_plan tx_plan;

/* Input stream description.  The block layout is the DataBuffer<float>
 * layout the reference hands to SubbandDedispersion::run
 * (src/singlepulse/singlepulse.cpp:328): time-major [nsamples][nchans].
 * TX_DTYPE_U8 blocks are 8-bit filterbank samples; each contributes
 * (x - zero_point).  TX_DTYPE_F32 blocks are front-end output floats.
 * chan_mask (optional, nchans bytes, 0 = zapped) zeroes whole channels. */
typedef struct {
    int nchans;
    int dtype;
    int zero_point;
    const uint8_t *chan_mask;
    long ndump; /* samples per block, DataBuffer::nsamples */
    int device; /* HIP device ordinal */
} tx_beam_desc;

/* One ddplan row.  Fields mirror what SinglePulse::prepare writes into the
 * dedisperser and boxcar (src/singlepulse/singlepulse.cpp:230-254). */
typedef struct {
    const double *frequencies; /* nchans channel frequencies, MHz */
    double tsamp;              /* s */
    double dms, ddm;           /* DM start and step, pc cm^-3 */
    long ndm;
    double overlap;            /* noverlap = (long)(overlap*ndump) */
    int nsub;                  /* sub-bands; 0 = nchans/32; must divide nchans */
    int group;                 /* fine DMs per coarse DM; 0 = smearing-matched default */
    const int *vwn;            /* boxcar width ladder (samples) */
    int nbox;                  /* 0: ladder given later by tx_boxcar_widths */
    int iqr;                   /* Boxcar::run iqr flag */
    int store_maps;            /* keep full mxS/mxwn on the device; without them the boxcar
                                  emits points only and tx_boxcar_rows recomputes rows */
    long max_points;           /* capacity of the compacted point list; 0 = default */
    int exact_order;           /* 1: boxcar mean/var and prefix sum in the reference's
                                  sequential FP32 order (bit-identical mxS; slow) */
    /* launch shapes (0 = chosen by the built-in cost model; an autotuner
     * measures alternatives and passes its choice here; results never change):
     * stage 1: coarse DMs per workgroup (1, 2, 4, 8, 16) and output quads
     * per thread (1, 2); stage 2: shape index + 1 (see tx_stage2_shapes)
     * and LDS bytes per workgroup; boxcar: s3_window 0 takes the widths
     * 1-9, 11, 13, 16 (the start of every ladder from one sample with
     * snrloss 0.1) from a register window where the ladder starts so,
     * 1 never does. */
    int s1_ncc, s1_q, s2_shape, s2_lds;
    int s3_window;
} tx_plan_desc;

/* The stage-2 launch shapes: (fine DMs per wave, samples per lane) pairs,
 * [n][2] into shapes; returns the count (n = 0 queries it). */
int tx_stage2_shapes(int *shapes, int n);

/* Read-back fields of RealTime::SubbandDedispersion / Boxcar
 * (src/singlepulse/boxcar.cpp:144-161, src/singlepulse/candplot.cpp:47-48). */
typedef struct {
    long counter;
    long offset;
    long noverlap;
    long ndump;
    long nsamples; /* ndump + noverlap */
    long ndm;
    int nsub;
    int cps;
    int group;
    int ndmc;
    long D1, D2, maxdelay;
    double fmin, fmax;
    int maxwn;
    double dms, ddm, tsamp; /* Boxcar::dms / ddm / tsamp */
} tx_info;

/* One thresholded point of the S/N map (src/singlepulse/cluster.cpp:95-107). */
typedef struct {
    int32_t idm;
    int32_t isamp;
    float S;
    int32_t wn;
} tx_point;

const char *tx_last_error(void);
int tx_version(void);

/* ---- beam: replaces the block hand-off of transientx_fil
 *      (src/singlepulse/dedisperse_fil.cpp:324-336) ---- */
int tx_beam_create(const tx_beam_desc *desc, tx_beam **out);
int tx_beam_destroy(tx_beam *beam);
/* Append one block of ns == ndump samples.  on_device: TX_PUSH_HOST (0)
 * host memory, copied before the call returns; TX_PUSH_DEVICE (1, or any
 * other value but 2) a device pointer (data already resident in HBM);
 * TX_PUSH_HOST_ASYNC (2) pinned host memory that the caller leaves unchanged
 * until the next tx_beam_sync: the copy is queued on the beam's stream and
 * the call returns at once, so one beam's copy overlaps other beams' kernels. */
#define TX_PUSH_HOST 0
#define TX_PUSH_DEVICE 1
#define TX_PUSH_HOST_ASYNC 2
int tx_beam_push(tx_beam *beam, const void *block, long ns, int on_device);
int tx_beam_sync(tx_beam *beam);

/* ---- front end on the device: replaces SinglePulse's Downsample and
 *      Equalize stages (src/singlepulse/singlepulse.cpp:208-216 prepare,
 *      :289-291 run; --td/--fd of src/singlepulse/dedisperse_fil.cpp:54-55)
 *      for raw 8-bit input.  The XLibs classes are not in the TransientX
 *      tree; the declared rule: output channel c < nchans_in/fd and sample t
 *      take y = the sum of the td x fd neighbouring 8-bit samples (exact),
 *      then (y - offset[c]) * gain[c] in float; masked channels are 0. ---- */
/* Makes beam (a TX_DTYPE_F32 beam whose nchans == nchans_in/fd) take raw
 * 8-bit blocks [ndump*td][nchans_in] in tx_beam_push / tx_beam_graph_push;
 * one kernel forms the front end's output straight into the beam's f32
 * input ring.  offset, gain: host arrays of nchans floats, copied.  Waits
 * for the beam's queued work and drops its graph (call tx_beam_graph
 * again).  nchans_in == 0 switches the front end off. */
int tx_beam_set_frontend(tx_beam *beam, int nchans_in, int td, int fd, const float *offset,
                         const float *gain);
/* The same front end inside stage 1's load (on != 0; call before the first
 * push): if at the first push the front end is on with td = fd = 1 and no
 * RFI filters are set, the beam's input ring keeps the raw 8-bit samples (a
 * quarter of the f32 ring's HBM) and stage 1 forms (y - offset) * gain as it
 * stages each row, with the same values and sums as the f32 ring; otherwise
 * the f32 ring is used.  Stage 1 then converts each sample once per chunk of
 * coarse DMs, which costs more than the 4-byte ring saves (the f32 ring is
 * the faster default).  Once the ring holds 8-bit samples, the front end
 * cannot be switched off or resampled and RFI filters cannot be added
 * (TX_ESTATE); new offsets and gains with td = fd = 1 can. */
int tx_beam_set_frontend_in_load(tx_beam *beam, int on);
/* The front end's RFI filters on the search path (SinglePulse::run applies
 * its "zdot" / "kadaneF" items between Equalize and the dedisperser,
 * src/singlepulse/singlepulse.cpp:298-325): nops filters (TX_RFI_ZDOT,
 * TX_RFI_KADANEF; at most 16) run in order on every block pushed with
 * tx_beam_push, in the DataBuffer layout [ndump][nchans] f32, after the
 * front end (tx_beam_set_frontend) or on the pushed f32 block (masked
 * channels zeroed first), then the block goes into the ring.  For kadaneF,
 * td[k], fd[k], threshold[k] and wnlimit[k] as tx_rfi_kadanef takes them
 * (other ops ignore their entries; the arrays may be null without kadaneF).
 * Needs a TX_DTYPE_F32 beam; graph pipelines refuse a beam with filters.
 * nops == 0 removes them.  Arithmetic: tx_rfi_zdot / tx_rfi_kadanef. */
int tx_beam_set_rfi(tx_beam *beam, int nops, const int *ops, const int *td, const int *fd,
                    const float *threshold, const int *wnlimit);
/* The same front end into a time-major f32 block [ns_in/td][nchans_in/fd]
 * (the DataBuffer the RFI filters take), on stream.  All pointers are
 * device memory; chan_mask may be null. */
int tx_frontend_u8(void *stream, const uint8_t *in, long ns_in, int nchans_in, int td, int fd,
                   const float *offset, const float *gain, const uint8_t *chan_mask, float *out);
/* Equalize's statistics of one raw block, declared: per output channel the
 * exact integer sums of y and y^2 over the ns_in/td samples, mean = S1/n,
 * var = S2/n - mean^2 (double); offset = (float)mean, gain = var > 0 ?
 * (float)(1/sqrt(var)) : 1.  Device pointers; work: device scratch of
 * tx_frontend_stats_work_bytes bytes. */
size_t tx_frontend_stats_work_bytes(int nchans_in, int fd);
int tx_frontend_stats(void *stream, const uint8_t *in, long ns_in, int nchans_in, int td, int fd,
                      float *offset, float *gain, void *work, size_t work_bytes);
/* No counterpart in the reference.  tx_frontend_stats for callers that
 * cannot allocate device memory (code under hipGraph capture, serving
 * loops, harnesses that forbid hipMalloc): the same kernels and results,
 * with the per-channel sums in a fixed scratch array inside the library
 * (1 MiB per GPU, held until the process ends) instead of a work buffer.
 * At most 65536 output channels (nchans_in/fd); more gives TX_EINVAL.
 * Allocates nothing.  Every call on one GPU uses that GPU's one scratch
 * array, so calls that overlap on a GPU (other streams, other threads)
 * race; calls queued on one stream do not. */
int tx_frontend_stats_noalloc(void *stream, const uint8_t *in, long ns_in, int nchans_in, int td, int fd,
                              float *offset, float *gain);

/* ---- plan: SubbandDedispersion::prepare + Boxcar::prepare
 *      (src/singlepulse/singlepulse.cpp:230-240, src/singlepulse/boxcar.cpp:142-156) ---- */
int tx_plan_create(tx_beam *beam, const tx_plan_desc *desc, tx_plan **out);
int tx_plan_destroy(tx_plan *plan);
int tx_get_info(const tx_plan *plan, tx_info *out);
/* Delay tables of the declared two-stage rule: d1 [ndmc][nchans], d2 [ndm][nsub];
 * dmc [ndmc] coarse DMs, fref [nsub] sub-band reference frequencies,
 * fsub [nsub] sub-band centre frequencies, fcnt [nsub] live channels per sub-band.
 * Any pointer may be NULL. */
int tx_get_tables(const tx_plan *plan, int *d1, int *d2, double *dmc, double *fref, double *fsub,
                  double *fcnt);

/* ---- SubbandDedispersion::run (src/singlepulse/singlepulse.cpp:328):
 *      consume the beam's latest block ---- */
int tx_dedisp_run(tx_plan *plan);
/* SubbandDedispersion::get_timdata(vector<float>&, int idm, bool)
 * (src/singlepulse/boxcar.cpp:183): with_overlap != 0 gives all
 * ndump+noverlap samples, 0 the last ndump.  out: host memory. */
int tx_get_timdata(tx_plan *plan, int idm, int with_overlap, float *out);
/* SubbandDedispersion::get_subdata (src/singlepulse/candplot.cpp:123):
 * sub-band series at DM idm, [nsub][ndump+noverlap] (host memory). */
int tx_get_subdata(tx_plan *plan, int idm, float *out);

/* ---- Boxcar::run (src/singlepulse/boxcar.cpp:158-178) plus the threshold
 *      scan of Cluster::run (src/singlepulse/cluster.cpp:89-112), fused:
 *      TX_NOT_READY until counter >= offset - noverlap + ndump. ---- */
int tx_boxcar_run(tx_plan *plan, float thre);
/* The width ladder and iqr flag of Boxcar::run(dedisp, vwn, iqr)
 * (src/singlepulse/boxcar.cpp:158): SinglePulse::prepare computes vwn after
 * Boxcar::prepare (src/singlepulse/singlepulse.cpp:240-254), so a plan may be
 * created with nbox = 0 and given its ladder here before the first
 * tx_boxcar_run; calling it again with another ladder replaces it. */
int tx_boxcar_widths(tx_plan *plan, const int *vwn, int nbox, int iqr);
/* Thresholded points of the last tx_boxcar_run in row-major (DM, sample)
 * order.  *n receives the total; TX_EOVERFLOW if it exceeds cap or the
 * device list capacity. */
int tx_boxcar_points(tx_plan *plan, tx_point *out, size_t cap, size_t *n);
/* Rows [dm0, dm1) of Boxcar::mxS / Boxcar::mxwn (src/singlepulse/candplot.cpp:278),
 * host memory, [dm1-dm0][ndump+noverlap].  Without store_maps the rows of the
 * current window are recomputed by the same kernels on those rows only (the
 * values stored maps would hold); the boxcar gate must be open. */
int tx_boxcar_rows(tx_plan *plan, int dm0, int dm1, float *mxS, int *mxwn);
/* Sum over samples isamp-wn < i < isamp+wn of each sub-band series at DM
 * idm divided by its live-channel count: vpf of src/singlepulse/candplot.cpp:160-170,
 * for ncand candidates at once; out [ncand][nsub] host memory. */
int tx_cand_spectra(tx_plan *plan, int ncand, const int *idm, const int *isamp, const int *wn,
                    float *out);

/* Device pointers for zero-copy consumers on the same device (torch, tests):
 * tim [ndm][N] f32, mxS [ndm][N] f32, mxwn [ndm][N] i32 (NULL without store_maps). */
int tx_device_buffers(tx_plan *plan, void **tim, void **mxS, void **mxwn);

/* ---- --savetim: dumps of the dedispersed series.  SubbandDedispersion's
 *      preparedump / rundump / makeinf / modifynblock
 *      (src/singlepulse/singlepulse.cpp:238,340, src/singlepulse/dedisperse_fil.cpp:356-369,
 *      options --savetim --format --mean --std --nbits at
 *      src/singlepulse/dedisperse_fil.cpp:95-99) live in XLibs, which the
 *      TransientX tree does not hold; the rule here is declared:
 *  * samples: each run's ndump new samples of every DM trial (the window
 *    without its overlap), from the first run whose new samples reach
 *    observation sample 0; the dump's first sample is tau0 (<= 0 samples
 *    before the observation's start when the first block is partial);
 *  * value: y = x * scale + mean in float (two roundings), with
 *    scale = (float)(std / sqrt(live channels)), so that a series of live
 *    unit-variance channels comes out with standard deviation std;
 *  * nbits 8 / 16: round half away from zero, clamped to the signed range
 *    [-2^(nbits-1), 2^(nbits-1)-1], stored as int8 / int16; nbits 32: y as
 *    float32. ---- */
#define TX_DUMP_PULSARX 0
#define TX_DUMP_SIGPROC 1
#define TX_DUMP_PRESTO 2

/* The current run's dump values, [ndm][n] host memory (n samples per DM,
 * nbits/8 bytes each) in out of cap bytes.  *tau0 = observation sample of
 * the first, *n = samples per DM (ndump, or 0 while the new samples all
 * precede the observation's start). */
int tx_dump_series(tx_plan *plan, float mean, float std, int nbits, void *out, size_t cap, long *tau0,
                   long *n);

typedef struct tx_dump tx_dump;
/* Filterbank fields of the dump header: the ones transientx_fil sets on
 * SinglePulse::fildedisp (src/singlepulse/dedisperse_fil.cpp:306-310). */
typedef struct {
    const char *rootname;    /* dedisp.rootname */
    const char *source_name; /* may be NULL */
    int format;              /* TX_DUMP_* */
    int nbits;               /* 8, 16 or 32 (PRESTO: 32) */
    double tstart;           /* MJD of observation sample 0 */
    double fch1, foff;       /* first channel and channel step (MHz) */
    int nchans, ibeam;
} tx_dump_desc;
/* Files (declared):
 *  PULSARX: <rootname>_DM<dms>_<dme>.dedisp, one per row: a SIGPROC-style
 *           keyword header (source_name, nbits, signed, tstart, tsamp, fch1,
 *           foff, nchans, ibeam, refdm = dms, ddm, ndm, ndump, nblock) and
 *           then nblock blocks of [ndm][ndump] samples; nblock is written
 *           as 0 and set by tx_dump_finish (modifynblock);
 *  SIGPROC: <rootname>_DM<dm>.tim per DM trial, a SIGPROC time series
 *           (data_type 2, nchans 1, refdm = its DM, signed 1 for 8 and 16
 *           bits) with the samples;
 *  PRESTO:  <rootname>_DM<dm>.dat per DM trial, float32 samples; its .inf
 *           is written by tx_dump_finish (makeinf).
 * <dm> is printed "%.2f".  tstart of a file is the MJD of its first
 * sample.  preparedump = tx_dump_open, rundump = tx_dump_block after each
 * run, makeinf / modifynblock = tx_dump_finish, then tx_dump_close. */
int tx_dump_open(tx_plan *plan, const tx_dump_desc *desc, tx_dump **out);
int tx_dump_block(tx_dump *dump, float mean, float std);
int tx_dump_finish(tx_dump *dump);
int tx_dump_close(tx_dump *dump);

/* ---- streaming extras (no counterpart in the reference; for callers that
 *      search a whole observation without a host round trip per block) ---- */

/* A thresholded point tagged with its block (counter/ndump, 1-based). */
typedef struct {
    int32_t block;
    int32_t idm;
    int32_t isamp;
    float S;
    int32_t wn;
    int32_t reserved;
} tx_point_ex;

/* Start a fresh stream on the beam: counter 0, input ring and every attached
 * plan's sub-band ring back to the zero signal (as a new transientx_fil run). */
int tx_beam_reset(tx_beam *beam);
/* The HIP stream (hipStream_t) every kernel of this beam runs on. */
int tx_beam_stream(tx_beam *beam, void **stream);
/* Per-kernel HIP-event timing on the beam's stream.  enable != 0 starts
 * recording; tx_beam_timing synchronises and returns, per kernel class
 * (0 transpose, 1 stage 1, 2 stage 2, 3 boxcar), the summed milliseconds and
 * launch counts since the last call, then clears them. */
int tx_beam_set_timing(tx_beam *beam, int enable);
int tx_beam_timing(tx_beam *beam, double *ms, long *launches);
/* Session mode: after tx_session_begin, tx_boxcar_run appends its points to
 * a device list of capacity cap instead of the per-block list;
 * tx_session_end synchronises, copies them in device order (sort by block,
 * idm, isamp for the row-major order of src/singlepulse/cluster.cpp:91-112)
 * and leaves session mode.  TX_EOVERFLOW if more than cap points were found. */
int tx_session_begin(tx_plan *plan, size_t cap);
/* The same with the list in the caller's page-locked host memory (from
 * hipHostMalloc, e.g. a pinned torch tensor; anything the device cannot
 * address is refused): the kernels store each point straight into host[]
 * as they find it, over PCIe, so tx_session_end copies nothing (out may be
 * host itself, or NULL) and the D2H of a whole search's points overlaps its
 * kernels.  host must stay allocated and unread until tx_session_end. */
int tx_session_begin_host(tx_plan *plan, tx_point_ex *host, size_t cap);
int tx_session_end(tx_plan *plan, tx_point_ex *out, size_t cap, size_t *n);

/* ---- per-block graph pipeline: the block hand-off of transientx_fil
 *      (src/singlepulse/dedisperse_fil.cpp:324-336: every SinglePulse of the
 *      beam runs SubbandDedispersion::run then Boxcar::run on the block),
 *      captured once as a hipGraph and replayed per block ----
 * tx_beam_graph builds the graph of one block for the beam's plans, all in
 * session mode with their widths set: the transpose, then per plan stage 1
 * -> stage 2 -> boxcar (threshold thre) as independent branches, then a
 * one-thread tick of the beam's device clock.  The kernels take the block's
 * counter from that clock, so a replay needs no new arguments but the
 * block's source.  tx_beam_graph_push(beam, block, ns, on_device) is
 * tx_beam_push + tx_dedisp_run + tx_boxcar_run of every plan in one
 * hipGraphLaunch (on_device as for tx_beam_push); points land in the
 * sessions exactly as those calls put them.  Call tx_beam_graph again after
 * tx_beam_push, tx_session_begin with a larger cap or new widths (a push
 * refuses with TX_ESTATE until then); tx_beam_reset keeps the graph. */
int tx_beam_graph(tx_beam *beam, float thre);
int tx_beam_graph_push(tx_beam *beam, const void *block, long ns, int on_device);

/* ---- many beams: Boxcar::run (src/singlepulse/boxcar.cpp:158-178) of the
 *      rows of every beam in one persistent launch over a device work queue
 *      (the reference runs one beam per process: src/singlepulse/dedisperse_fil.cpp:69) ----
 * A tx_batch owns a stream on one device.  tx_batch_boxcar(batch, plans, n,
 * thre, ready) is tx_boxcar_run(plans[i], thre) for every i: the plans may
 * belong to any beams on the batch's device; each plan's gate is checked as
 * tx_boxcar_run checks it and ready[i] (optional) is set to 1 when it ran, 0
 * when it is not ready yet.  The rows of all ready plans in the default
 * statistics mode become one work list of (row, 2048-sample segment) items
 * that one launch of as many workgroups as the device holds at once works
 * through, each workgroup claiming the next item from a device counter;
 * rows in exact-order or IQR mode get a launch of their own on the same
 * stream.  The launch waits for the work already queued on every plan's
 * beam, and every such beam's later work waits for it.  Points land where
 * tx_boxcar_run puts them (session lists, or tx_boxcar_points).  Not for
 * beams with a per-block graph (tx_beam_graph_push runs its own boxcar). */
typedef struct tx_batch tx_batch;
int tx_batch_create(int device, tx_batch **out);
int tx_batch_destroy(tx_batch *batch);
int tx_batch_stream(tx_batch *batch, void **stream);
int tx_batch_boxcar(tx_batch *batch, tx_plan *const *plans, int nplans, float thre, int *ready);

/* ---- candidate refinement grid: Candidate::dedisperse (incoherent),
 *      Candidate::optimize with the dedispersion tree of include/hough_transform.h,
 *      Candidate::matched_filter (src/singlepulse/candidate.cpp:108-140,899-1087),
 *      as the replot runs them (src/singlepulse/replot_fil.cpp:419-458) on a
 *      batch of cutouts of one shape ---- */
typedef struct tx_refiner tx_refiner;

typedef struct {
    int nchan;                 /* channels per cutout (Candidate::nchan), >= 2 */
    int nbin;                  /* samples per cutout (Candidate::nbin), 2..4096 */
    const double *frequencies; /* nchan channel frequencies, either order, MHz */
    double tbin;               /* s */
    float snrloss;             /* matched_filter ladder */
    int dedisperse;            /* 1: rotate each channel by its delay at the cutout's dm first */
    int isdedispersed;         /* 1: DM trials relative to dm (Candidate::isdedispersed) */
    int device;
} tx_refine_desc;

/* Fields Candidate::optimize / matched_filter set (src/singlepulse/candidate.cpp:901-911,1067-1086). */
typedef struct {
    double dms, ddm; /* DM grid: dms + i*ddm, i < ndm */
    int ndm;         /* 2^ceil(log2 nchan) */
    int maxdmid;     /* first maximum of the S/N map, row-major */
    int maxtid;
    float snr;
    int wn;          /* boxcar width (samples) at the maximum */
    float width;     /* wn * tbin, s */
    double dm_maxsnr;
} tx_refine_result;

int tx_refine_create(const tx_refine_desc *desc, int max_batch, tx_refiner **out);
int tx_refine_destroy(tx_refiner *r);
/* nrow (= ndm), matched-filter widths, tree passes */
int tx_refine_info(const tx_refiner *r, int *nrow, int *nbox, int *npass);
int tx_refine_stream(tx_refiner *r, void **stream);
/* cutouts: device [ncand][nchan][nbin] f32; dm: host [ncand] (Candidate::dm);
 * res: host [ncand].  dmtmap / snrmap (f32) / widmap (i32): optional device
 * outputs [ncand][ndm][nbin] (Candidate::dmtmap, dmtsnrmap, dmtwidmap). */
int tx_refine_run(tx_refiner *r, int ncand, const float *cutouts, const double *dm,
                  tx_refine_result *res, float *dmtmap, float *snrmap, int *widmap);

/* ---- differentiable S/N for gradient refinement (no counterpart in the
 *      reference: it replaces the replot's DM/t0/width grid search,
 *      src/singlepulse/candidate.cpp:899-1087, by gradient ascent, with the
 *      grid above kept as fallback and judge) ----
 * For cutout x [nchan][nbin], channel weights w, DM, arrival time t0 and
 * width W (samples):
 *   tau_j = K*DM*(1/f_j^2 - 1/fmax^2)/tbin (fractional samples),
 *   p(t)  = sum_j w_j * linear interpolation of x_j at t + tau_j (circular),
 *   g(t)  = sig((t-t0+W/2)/s) - sig((t-t0-W/2)/s), t-t0 wrapped to [-nbin/2, nbin/2),
 *   S     = sum_t (g - mean g)(t) p(t) / (|g - mean g| |w|). */
typedef struct tx_diff tx_diff;
typedef struct {
    int nchan, nbin;           /* cutout shape, nbin <= 4096 */
    const double *frequencies; /* nchan, MHz */
    double tbin;               /* s */
    float softness;            /* s: edge softness of the boxcar, samples */
    int device;
} tx_diff_desc;
int tx_diff_create(const tx_diff_desc *desc, tx_diff **out);
int tx_diff_destroy(tx_diff *df);
int tx_diff_stream(tx_diff *df, void **stream);
/* neval evaluations, all device pointers; evaluation e reads cutout cand[e]
 * (e if cand is NULL) of x [ncut][nchan][nbin] and w [ncut][nchan], and
 * dm/t0/width [neval] (f32); snr [neval].  Returns when done. */
int tx_refine_fwd(tx_diff *df, int neval, const int *cand, const float *x, const float *w,
                  const float *dm, const float *t0, const float *width, float *snr);
/* Adjoint: recomputes the forward and returns gsnr[e] times dS/dDM, dS/dt0,
 * dS/dW [neval], dS/dw [neval][nchan] (optional) and dS/dx
 * [neval][nchan][nbin] (optional; a gather, no atomics). */
int tx_refine_adj(tx_diff *df, int neval, const int *cand, const float *x, const float *w,
                  const float *dm, const float *t0, const float *width, const float *gsnr, float *snr,
                  float *gdm, float *gt0, float *gwidth, float *gw, float *gx);
/* The adjoint without dS/dw and dS/dx for runs of evaluations: evaluation
 * e = c*run + i (i < run) of cutout c, ncut*run evaluations in all.  A
 * workgroup carries up to 8 evaluations of one cutout through the cutout's
 * rows, which it stages in LDS once per pass: the cutout is read once per 8
 * evaluations instead of once per evaluation.  Same outputs as
 * tx_refine_adj with cand[e] = e / run, up to FP32 rounding order (channels
 * summed in channel order).  gtt, gdt (optional, both or neither): gsnr[e]
 * times d2S/dt0^2 and d2S/dDM dt0, for Newton steps in the arrival time.  Needs nbin <= 1024, a multiple of 4, and x
 * 16-byte aligned.  Returns when done. */
int tx_refine_adj_runs(tx_diff *df, int ncut, int run, const float *x, const float *w, const float *dm,
                       const float *t0, const float *width, const float *gsnr, float *snr, float *gdm,
                       float *gt0, float *gwidth, float *gtt, float *gdt);
/* The same S on a (DM, t0, W) grid around each of ncut cutouts (the grid
 * search this path replaces, and its judge): DM trials dm0[c] + m*ddm
 * (m < ndm), t0 trials t00[c] + i*dt (i < nt0; dt a whole number of samples
 * or 1/q, q <= 64; (nt0-1)*dt < nbin), widths [nwidth] (device).  Per
 * (cutout, DM trial): best [ncut][ndm] the highest S over (t0, W) and
 * best_idx [ncut][ndm] = i*nwidth + k of its first occurrence.  Each DM
 * trial's profile is formed once; S agrees with tx_refine_fwd up to FP32
 * rounding order.  Returns when done. */
int tx_refine_grid(tx_diff *df, int ncut, const float *x, const float *w, const float *dm0,
                   const float *t00, int ndm, float ddm, int nt0, float dt, int nwidth,
                   const float *widths, float *best, int *best_idx);

/* ---- RFI cleaning of channel x time frames on the device: Candidate::zdot
 *      and Candidate::kadaneF (src/singlepulse/candidate.cpp:710-897), the
 *      formulas behind the front end's "zdot" and "kadaneF" items
 *      (src/singlepulse/singlepulse.cpp:305-317; there XLibs' RFI class, not
 *      in the reference, applies them to each DataBuffer block) ----
 * A frame is nchan x nsamp floats in device memory; element (channel j,
 * sample i) of frame f is data[f*fstride + j*cstride + i*tstride]:
 * (cstride, tstride) = (nsamp, 1) for the Candidate layout [nchan][nbin],
 * (1, nchans) for a DataBuffer block [nsamp][nchans].  In place, queued on
 * stream (a hipStream_t; NULL = the default stream), returning at once.
 * work: device scratch of at least tx_rfi_work_bytes(op, ...) bytes, not
 * touched by the caller until the stream has passed the call.  Arithmetic
 * and order are the reference's (FP64 sums in its sequential order, no
 * fused multiply-adds): bit-identical to the scalar reference code.
 * tx_rfi_kadanef: td x fd downsampling, then per downsampled channel the
 * strongest positive and negative runs; a run of more than wnlimit[f]
 * (device, one per frame: Candidate's int(nwidth*width/tbin/td)) samples
 * with sum^2/(wn*td*fd) > threshold^2 is zeroed in its fd channels. */
#define TX_RFI_ZDOT 1
#define TX_RFI_KADANEF 2
size_t tx_rfi_work_bytes(int op, int nframes, int nchan, long nsamp, int td, int fd);
int tx_rfi_zdot(void *stream, float *data, int nframes, long fstride, int nchan, long nsamp, long cstride,
                long tstride, void *work, size_t work_bytes);
int tx_rfi_kadanef(void *stream, float *data, int nframes, long fstride, int nchan, long nsamp,
                   long cstride, long tstride, int td, int fd, float threshold, const int *wnlimit,
                   void *work, size_t work_bytes);

/* ---- host side around the kernels (CPU only) ---- */

/* One search row of the ddplan: the SinglePulse fields that
 * src/singlepulse/singlepulse.cpp:343-545 (parse, parse_json) set.
 * rfi: front-end items ("mask a b", "kadaneF a b", "kadaneT a b", "zero",
 * "zdot") joined by ';'; zap: "fl fh" pairs joined by ';'.  nsub and group
 * are optional additions of this implementation (0 = default). */
typedef struct {
    int td, fd;
    double dms, ddm;
    long ndm;
    double overlap;
    int nsub, group;
    float minw, snrloss, maxw;
    int iqr;
    float thre;
    double radius; /* s (the config's ms / 1000) */
    int neighbors, maxncand, minpts, drop;
    float bswidth;
    double bandlimit, bandlimitKT, widthlimit;
    float threMask, threKadaneF, threKadaneT;
    char rfi[512];
    char zap[512];
    int id;        /* SinglePulse::id, 1-based */
} tx_ddrow;

/* parse (ddplan text, is_json = 0) and parse_json (is_json = 1) of
 * src/singlepulse/singlepulse.cpp:343-545.  defaults holds the command-line
 * values; as in the reference one row object is reused, so rfi and zap items
 * accumulate from row to row.  Text without ddplan lines gives one row of
 * the defaults.  *n receives the row count; TX_EOVERFLOW if above cap. */
int tx_ddplan_parse(const char *text, int is_json, const tx_ddrow *defaults, tx_ddrow *rows, int cap,
                    int *n);
/* ndump = (int)(seglen/tsamp)/lcm(tds)*lcm(tds) (src/singlepulse/dedisperse_fil.cpp:224-244);
 * tds[i] = row td * preprocessing td.  Negative TX_E* code on error. */
long tx_ndump(float seglen, double tsamp, const int *tds, int n);

/* Parameters of Cluster<T>::run (src/singlepulse/cluster.cpp:63): threS is
 * applied by tx_boxcar_run; radius in s (radius_smearing). */
typedef struct {
    double dms, ddm, fmax, fmin, tsamp;
    double radius;
    int maxncand, minpts, drop, maxwn;
} tx_cluster_desc;

/* One entry of Cluster::candlist (src/singlepulse/cluster.cpp:148-152). */
typedef struct {
    int32_t idm;
    int32_t isamp;
    int32_t wn;
    float S;
    int32_t cluster; /* 1-based cluster label */
    int32_t npts;    /* points in the cluster */
} tx_cand;

/* Cluster<T>::run + clustering() (src/singlepulse/cluster.cpp:63-161,
 * include/clustering.h:26-100) over points in row-major (DM, sample) order,
 * as tx_boxcar_points returns them.  out: candlist in cluster order;
 * labels (optional, n entries): cluster label per point, 0 = none. */
int tx_cluster(const tx_point *pts, size_t n, const tx_cluster_desc *desc, tx_cand *out, size_t cap,
               size_t *ncand, int32_t *labels);
/* Maximum-sum contiguous run (the kadane of src/singlepulse/candplot.cpp:229). */
void tx_kadane(const float *a, long n, long *start, long *end);
/* fl/fh of candidates (src/singlepulse/candplot.cpp:141-231) from the
 * plan's current sub-band series. */
int tx_cand_flfh(tx_plan *plan, int ncand, const int *idm, const int *isamp, const int *wn, float *fl,
                 float *fh);

/* Context of CandPlot::plot's .cands lines (src/singlepulse/candplot.cpp:32-110). */
typedef struct {
    double tstart;        /* MJD of the first sample */
    const char *rootname;
    const char *beam;     /* obsinfo["Beam"], e.g. "cfbf00001" */
    int plan_id;          /* SinglePulse::id */
    const char *fname;    /* input file name */
    long *num_cand;       /* CandPlot::num_cand, shared by all rows; incremented per line */
} tx_cands_ctx;
/* The .cands lines of the plan's current block (src/singlepulse/candplot.cpp:420-433),
 * candidates stable-sorted by descending S (include/candplot.h:27-33), as
 * one NUL-terminated text.  *len receives its length; TX_EOVERFLOW if
 * len + 1 > cap.  The file is rootname_<tstart %.10f>_<beam>.cands. */
int tx_cands_format(tx_plan *plan, const tx_cand *cands, size_t n, const tx_cands_ctx *ctx, char *buf,
                    size_t cap, size_t *len);

/* ---- SIGPROC filterbank input (CPU only): the FilterbankReader that
 *      transientx_fil configures and reads block by block
 *      (src/singlepulse/dedisperse_fil.cpp:167-195 open and header,
 *      :214-217 nchans/tsamp/nifs/nsamples, :278-283 --jump, :324
 *      read_data).  The reader is XLibs' and not in the TransientX tree; the
 *      declared rule is the public SIGPROC format: a header of keywords (a
 *      32-bit length, then the bytes) from HEADER_START to HEADER_END, each
 *      followed by its value (32-bit int, 64-bit double, or a length-prefixed
 *      string; "signed" one byte; fchannel doubles between FREQUENCY_START and
 *      FREQUENCY_END), then nsamples x nifs x nchans samples of nbits, little
 *      endian.  Inputs of 8-bit unsigned or 32-bit float samples with one IF
 *      are read; anything else is refused.  ---- */
typedef struct {
    char source_name[128];
    char rawdatafile[128];
    int telescope_id, machine_id, data_type;
    int nbits, nchans, nifs, nbeams, ibeam;
    int barycentric, pulsarcentric;
    double tstart;              /* MJD of the first sample */
    double tsamp;               /* s */
    double fch1, foff;          /* MHz: channel j is at fch1 + j*foff */
    double src_raj, src_dej;    /* hhmmss.s, ddmmss.s */
    double az_start, za_start, refdm, period;
    long header_bytes;          /* bytes before the first sample (read only) */
    long nsamples;              /* whole samples the data holds (read only) */
} tx_fil_header;
/* The header of one file; nsamples from the file's size. */
int tx_fil_read_header(const char *path, tx_fil_header *out);
/* Creates (or empties) path and writes hdr's header (the keywords above
 * with non-empty strings; nbits 8 or 32, nifs 1); the caller appends the
 * samples.  hdr->header_bytes and nsamples are ignored. */
int tx_fil_write_header(const char *path, const tx_fil_header *hdr);
typedef struct tx_fil tx_fil;
/* Opens npaths files as one stream, in order.  They must agree in nbits,
 * nchans, nifs, tsamp, fch1 and foff; with check_continuity each file must
 * start where the previous one ends (tstart within half a sample), as
 * transientx_fil checks unless --cont.  *hdr receives the first file's
 * header with nsamples the total. */
int tx_fil_open(const char *const *paths, int npaths, int check_continuity, tx_fil **out, tx_fil_header *hdr);
/* --jump: skips skip_start samples at the beginning and reads nothing of
 * the last skip_end samples (dedisperse_fil.cpp:278-283 sets them from
 * seconds as (long)(j0/tsamp) and ntotal - (long)(ntotal - j1/tsamp)).
 * Rewinds the stream to sample skip_start. */
int tx_fil_skip(tx_fil *f, long skip_start, long skip_end);
/* Channel frequencies (nchans doubles, MHz). */
int tx_fil_frequencies(const tx_fil *f, double *freqs);
/* Reads up to nsamp whole samples [nsamp][nchans] into dst (host memory,
 * nsamp * nchans * nbits/8 bytes), across file boundaries.  Returns the
 * number read (fewer only at the end, 0 there), or a negative TX_ error. */
long tx_fil_read(tx_fil *f, void *dst, long nsamp);
int tx_fil_close(tx_fil *f);

#ifdef __cplusplus
}
#endif

#endif /* TXGPU_H */
