"""ctypes binding of libtxgpu.so (include/txgpu.h).

The library is loaded from the package directory only; there is no CPU
fallback.  :func:`lib` raises if the HIP library is missing.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIBPATH = Path(__file__).resolve().parent / "libtxgpu.so"

TX_OK = 0
TX_NOT_READY = 1
TX_EINVAL = -1
TX_ENOMEM = -2
TX_ERANGE = -3
TX_EHIP = -4
TX_ESTATE = -5
TX_EOVERFLOW = -6
TX_EIO = -7
TX_DUMP_PULSARX = 0
TX_DUMP_SIGPROC = 1
TX_DUMP_PRESTO = 2
TX_PUSH_HOST = 0
TX_PUSH_DEVICE = 1
TX_RFI_ZDOT = 1
TX_RFI_KADANEF = 2
TX_PUSH_HOST_ASYNC = 2
TX_DTYPE_U8 = 0
TX_DTYPE_F32 = 1


class TxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"txgpu error {code}: {msg}")
        self.code = code


class BeamDesc(C.Structure):
    _fields_ = [
        ("nchans", C.c_int),
        ("dtype", C.c_int),
        ("zero_point", C.c_int),
        ("chan_mask", C.POINTER(C.c_uint8)),
        ("ndump", C.c_long),
        ("device", C.c_int),
    ]


class FilHeader(C.Structure):
    _fields_ = [
        ("source_name", C.c_char * 128),
        ("rawdatafile", C.c_char * 128),
        ("telescope_id", C.c_int),
        ("machine_id", C.c_int),
        ("data_type", C.c_int),
        ("nbits", C.c_int),
        ("nchans", C.c_int),
        ("nifs", C.c_int),
        ("nbeams", C.c_int),
        ("ibeam", C.c_int),
        ("barycentric", C.c_int),
        ("pulsarcentric", C.c_int),
        ("tstart", C.c_double),
        ("tsamp", C.c_double),
        ("fch1", C.c_double),
        ("foff", C.c_double),
        ("src_raj", C.c_double),
        ("src_dej", C.c_double),
        ("az_start", C.c_double),
        ("za_start", C.c_double),
        ("refdm", C.c_double),
        ("period", C.c_double),
        ("header_bytes", C.c_long),
        ("nsamples", C.c_long),
    ]


class PlanDesc(C.Structure):
    _fields_ = [
        ("frequencies", C.POINTER(C.c_double)),
        ("tsamp", C.c_double),
        ("dms", C.c_double),
        ("ddm", C.c_double),
        ("ndm", C.c_long),
        ("overlap", C.c_double),
        ("nsub", C.c_int),
        ("group", C.c_int),
        ("vwn", C.POINTER(C.c_int)),
        ("nbox", C.c_int),
        ("iqr", C.c_int),
        ("store_maps", C.c_int),
        ("max_points", C.c_long),
        ("exact_order", C.c_int),
        ("s1_ncc", C.c_int),
        ("s1_q", C.c_int),
        ("s2_shape", C.c_int),
        ("s2_lds", C.c_int),
        ("s3_window", C.c_int),
    ]


class Info(C.Structure):
    _fields_ = [
        ("counter", C.c_long),
        ("offset", C.c_long),
        ("noverlap", C.c_long),
        ("ndump", C.c_long),
        ("nsamples", C.c_long),
        ("ndm", C.c_long),
        ("nsub", C.c_int),
        ("cps", C.c_int),
        ("group", C.c_int),
        ("ndmc", C.c_int),
        ("D1", C.c_long),
        ("D2", C.c_long),
        ("maxdelay", C.c_long),
        ("fmin", C.c_double),
        ("fmax", C.c_double),
        ("maxwn", C.c_int),
        ("dms", C.c_double),
        ("ddm", C.c_double),
        ("tsamp", C.c_double),
    ]


class DdRow(C.Structure):
    _fields_ = [
        ("td", C.c_int), ("fd", C.c_int),
        ("dms", C.c_double), ("ddm", C.c_double),
        ("ndm", C.c_long),
        ("overlap", C.c_double),
        ("nsub", C.c_int), ("group", C.c_int),
        ("minw", C.c_float), ("snrloss", C.c_float), ("maxw", C.c_float),
        ("iqr", C.c_int),
        ("thre", C.c_float),
        ("radius", C.c_double),
        ("neighbors", C.c_int), ("maxncand", C.c_int), ("minpts", C.c_int), ("drop", C.c_int),
        ("bswidth", C.c_float),
        ("bandlimit", C.c_double), ("bandlimitKT", C.c_double), ("widthlimit", C.c_double),
        ("threMask", C.c_float), ("threKadaneF", C.c_float), ("threKadaneT", C.c_float),
        ("rfi", C.c_char * 512),
        ("zap", C.c_char * 512),
        ("id", C.c_int),
    ]


class ClusterDesc(C.Structure):
    _fields_ = [
        ("dms", C.c_double), ("ddm", C.c_double), ("fmax", C.c_double), ("fmin", C.c_double),
        ("tsamp", C.c_double), ("radius", C.c_double),
        ("maxncand", C.c_int), ("minpts", C.c_int), ("drop", C.c_int), ("maxwn", C.c_int),
    ]


class CandsCtx(C.Structure):
    _fields_ = [
        ("tstart", C.c_double),
        ("rootname", C.c_char_p),
        ("beam", C.c_char_p),
        ("plan_id", C.c_int),
        ("fname", C.c_char_p),
        ("num_cand", C.POINTER(C.c_long)),
    ]


class DumpDesc(C.Structure):
    _fields_ = [
        ("rootname", C.c_char_p), ("source_name", C.c_char_p),
        ("format", C.c_int), ("nbits", C.c_int),
        ("tstart", C.c_double), ("fch1", C.c_double), ("foff", C.c_double),
        ("nchans", C.c_int), ("ibeam", C.c_int),
    ]


class RefineDesc(C.Structure):
    _fields_ = [
        ("nchan", C.c_int), ("nbin", C.c_int),
        ("frequencies", C.POINTER(C.c_double)),
        ("tbin", C.c_double),
        ("snrloss", C.c_float),
        ("dedisperse", C.c_int), ("isdedispersed", C.c_int),
        ("device", C.c_int),
    ]


class DiffDesc(C.Structure):
    _fields_ = [
        ("nchan", C.c_int), ("nbin", C.c_int),
        ("frequencies", C.POINTER(C.c_double)),
        ("tbin", C.c_double),
        ("softness", C.c_float),
        ("device", C.c_int),
    ]


REFINE_RESULT_DTYPE = np.dtype([("dms", "<f8"), ("ddm", "<f8"), ("ndm", "<i4"), ("maxdmid", "<i4"),
                                ("maxtid", "<i4"), ("snr", "<f4"), ("wn", "<i4"), ("width", "<f4"),
                                ("dm_maxsnr", "<f8")], align=True)

CAND_DTYPE = np.dtype([("idm", "<i4"), ("isamp", "<i4"), ("wn", "<i4"), ("S", "<f4"),
                       ("cluster", "<i4"), ("npts", "<i4")])

POINT_DTYPE = np.dtype([("idm", "<i4"), ("isamp", "<i4"), ("S", "<f4"), ("wn", "<i4")])
POINT_EX_DTYPE = np.dtype([("block", "<i4"), ("idm", "<i4"), ("isamp", "<i4"), ("S", "<f4"),
                           ("wn", "<i4"), ("reserved", "<i4")])

# every entry point declared in include/txgpu.h
EXPORTS = [
    "tx_last_error", "tx_version",
    "tx_beam_create", "tx_beam_destroy", "tx_beam_push", "tx_beam_sync",
    "tx_beam_set_frontend", "tx_beam_set_frontend_in_load", "tx_beam_set_rfi", "tx_frontend_u8",
    "tx_frontend_stats_work_bytes", "tx_frontend_stats",
    "tx_frontend_stats_noalloc",
    "tx_plan_create", "tx_plan_destroy", "tx_get_info", "tx_get_tables", "tx_stage2_shapes",
    "tx_dedisp_run", "tx_get_timdata", "tx_get_subdata",
    "tx_boxcar_run", "tx_boxcar_widths", "tx_boxcar_points", "tx_boxcar_rows", "tx_cand_spectra",
    "tx_device_buffers",
    "tx_dump_series", "tx_dump_open", "tx_dump_block", "tx_dump_finish", "tx_dump_close",
    "tx_beam_reset", "tx_beam_stream", "tx_beam_set_timing", "tx_beam_timing",
    "tx_session_begin", "tx_session_begin_host", "tx_session_end", "tx_beam_graph", "tx_beam_graph_push",
    "tx_batch_create", "tx_batch_destroy", "tx_batch_stream", "tx_batch_boxcar",
    "tx_refine_create", "tx_refine_destroy", "tx_refine_info", "tx_refine_stream", "tx_refine_run",
    "tx_diff_create", "tx_diff_destroy", "tx_diff_stream", "tx_refine_fwd", "tx_refine_adj",
    "tx_refine_adj_runs",
    "tx_refine_grid",
    "tx_rfi_work_bytes", "tx_rfi_zdot", "tx_rfi_kadanef",
    "tx_ddplan_parse", "tx_ndump", "tx_cluster", "tx_kadane", "tx_cand_flfh", "tx_cands_format",
    "tx_fil_read_header", "tx_fil_write_header", "tx_fil_open", "tx_fil_skip", "tx_fil_frequencies",
    "tx_fil_read", "tx_fil_close",
]

_lib = None


def _declare(L):
    vp, ip, i = C.c_void_p, C.c_int, C.c_int
    L.tx_last_error.restype = C.c_char_p
    L.tx_last_error.argtypes = []
    L.tx_version.restype = ip
    L.tx_beam_create.argtypes = [C.POINTER(BeamDesc), C.POINTER(vp)]
    L.tx_beam_destroy.argtypes = [vp]
    L.tx_beam_push.argtypes = [vp, vp, C.c_long, i]
    L.tx_beam_sync.argtypes = [vp]
    L.tx_beam_set_frontend.argtypes = [vp, i, i, i, vp, vp]
    L.tx_beam_set_frontend_in_load.argtypes = [vp, i]
    L.tx_beam_set_rfi.argtypes = [vp, i, vp, vp, vp, vp, vp]
    L.tx_frontend_u8.argtypes = [vp, vp, C.c_long, i, i, i, vp, vp, vp, vp]
    L.tx_frontend_stats_work_bytes.argtypes = [i, i]
    L.tx_frontend_stats.argtypes = [vp, vp, C.c_long, i, i, i, vp, vp, vp, C.c_size_t]
    L.tx_frontend_stats_noalloc.argtypes = [vp, vp, C.c_long, i, i, i, vp, vp]
    L.tx_plan_create.argtypes = [vp, C.POINTER(PlanDesc), C.POINTER(vp)]
    L.tx_plan_destroy.argtypes = [vp]
    L.tx_stage2_shapes.argtypes = [vp, i]
    L.tx_get_info.argtypes = [vp, C.POINTER(Info)]
    L.tx_get_tables.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.tx_dedisp_run.argtypes = [vp]
    L.tx_get_timdata.argtypes = [vp, i, i, vp]
    L.tx_get_subdata.argtypes = [vp, i, vp]
    L.tx_boxcar_run.argtypes = [vp, C.c_float]
    L.tx_boxcar_widths.argtypes = [vp, vp, i, i]
    L.tx_boxcar_points.argtypes = [vp, vp, C.c_size_t, C.POINTER(C.c_size_t)]
    L.tx_boxcar_rows.argtypes = [vp, i, i, vp, vp]
    L.tx_cand_spectra.argtypes = [vp, i, vp, vp, vp, vp]
    L.tx_device_buffers.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]
    L.tx_dump_series.argtypes = [vp, C.c_float, C.c_float, i, vp, C.c_size_t, C.POINTER(C.c_long),
                                 C.POINTER(C.c_long)]
    L.tx_dump_open.argtypes = [vp, C.POINTER(DumpDesc), C.POINTER(vp)]
    L.tx_dump_block.argtypes = [vp, C.c_float, C.c_float]
    L.tx_dump_finish.argtypes = [vp]
    L.tx_dump_close.argtypes = [vp]
    L.tx_beam_reset.argtypes = [vp]
    L.tx_beam_stream.argtypes = [vp, C.POINTER(vp)]
    L.tx_beam_set_timing.argtypes = [vp, i]
    L.tx_beam_timing.argtypes = [vp, vp, vp]
    L.tx_session_begin.argtypes = [vp, C.c_size_t]
    L.tx_session_begin_host.argtypes = [vp, vp, C.c_size_t]
    L.tx_session_end.argtypes = [vp, vp, C.c_size_t, C.POINTER(C.c_size_t)]
    L.tx_beam_graph.argtypes = [vp, C.c_float]
    L.tx_beam_graph_push.argtypes = [vp, vp, C.c_long, i]
    L.tx_batch_create.argtypes = [i, C.POINTER(vp)]
    L.tx_batch_destroy.argtypes = [vp]
    L.tx_batch_stream.argtypes = [vp, C.POINTER(vp)]
    L.tx_batch_boxcar.argtypes = [vp, C.POINTER(vp), i, C.c_float, vp]
    L.tx_refine_create.argtypes = [C.POINTER(RefineDesc), i, C.POINTER(vp)]
    L.tx_refine_destroy.argtypes = [vp]
    L.tx_refine_info.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.tx_refine_stream.argtypes = [vp, C.POINTER(vp)]
    L.tx_refine_run.argtypes = [vp, i, vp, vp, vp, vp, vp, vp]
    L.tx_diff_create.argtypes = [C.POINTER(DiffDesc), C.POINTER(vp)]
    L.tx_diff_destroy.argtypes = [vp]
    L.tx_diff_stream.argtypes = [vp, C.POINTER(vp)]
    L.tx_refine_fwd.argtypes = [vp, i, vp, vp, vp, vp, vp, vp, vp]
    L.tx_refine_adj.argtypes = [vp, i, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.tx_refine_adj_runs.argtypes = [vp, i, i, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.tx_rfi_work_bytes.restype = C.c_size_t
    L.tx_rfi_work_bytes.argtypes = [i, i, i, C.c_long, i, i]
    L.tx_rfi_zdot.argtypes = [vp, vp, i, C.c_long, i, C.c_long, C.c_long, C.c_long, vp, C.c_size_t]
    L.tx_rfi_kadanef.argtypes = [vp, vp, i, C.c_long, i, C.c_long, C.c_long, C.c_long, i, i, C.c_float, vp,
                                 vp, C.c_size_t]
    L.tx_refine_grid.argtypes = [vp, i, vp, vp, vp, vp, i, C.c_float, i, C.c_float, i, vp, vp, vp]
    L.tx_ddplan_parse.argtypes = [C.c_char_p, i, C.POINTER(DdRow), vp, i, C.POINTER(C.c_int)]
    L.tx_ndump.argtypes = [C.c_float, C.c_double, vp, i]
    L.tx_cluster.argtypes = [vp, C.c_size_t, C.POINTER(ClusterDesc), vp, C.c_size_t,
                             C.POINTER(C.c_size_t), vp]
    L.tx_kadane.argtypes = [vp, C.c_long, C.POINTER(C.c_long), C.POINTER(C.c_long)]
    L.tx_cand_flfh.argtypes = [vp, i, vp, vp, vp, vp, vp]
    L.tx_cands_format.argtypes = [vp, vp, C.c_size_t, C.POINTER(CandsCtx), vp, C.c_size_t,
                                  C.POINTER(C.c_size_t)]
    for name in EXPORTS:
        if name not in ("tx_last_error", "tx_version"):
            getattr(L, name).restype = ip
    L.tx_ndump.restype = C.c_long
    L.tx_fil_read_header.argtypes = [C.c_char_p, C.POINTER(FilHeader)]
    L.tx_fil_write_header.argtypes = [C.c_char_p, C.POINTER(FilHeader)]
    L.tx_fil_open.argtypes = [C.POINTER(C.c_char_p), i, i, C.POINTER(vp), C.POINTER(FilHeader)]
    L.tx_fil_skip.argtypes = [vp, C.c_long, C.c_long]
    L.tx_fil_frequencies.argtypes = [vp, vp]
    L.tx_fil_read.argtypes = [vp, vp, C.c_long]
    L.tx_fil_read.restype = C.c_long
    L.tx_fil_close.argtypes = [vp]
    L.tx_frontend_stats_work_bytes.restype = C.c_size_t
    L.tx_kadane.restype = None


def lib():
    """Load libtxgpu.so (the HIP path).  Raises if it was not built."""
    global _lib
    if _lib is None:
        if not LIBPATH.exists():
            raise RuntimeError(
                f"{LIBPATH} is missing: build it with `python -m transientx_amd.build`")
        # One HIP runtime per process: torch ships its own libamdhip64 (same
        # soname), which its libraries load by another name.  Loading torch
        # first makes this library bind to that copy instead of a second one.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(str(LIBPATH))
        _declare(L)
        _lib = L
    return _lib


def check(rc: int) -> int:
    if rc < 0:
        raise TxError(rc, lib().tx_last_error().decode())
    return rc


def ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)
