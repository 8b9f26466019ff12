"""ctypes binding of libfastdose_hip.so (the C-ABI in include/fastdose_hip.h).

The library is built in-tree by ``python -m fastdose_amd.build``; importing this module never
falls back to anything else: a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libfastdose_hip.so"

FD_OK = 0
FD_ERR_ARG = -1
FD_ERR_HIP = -2
FD_ERR_GEOM = -3
FD_ERR_IO = -4
FD_ERR_CAPACITY = -5
FD_ERR_UNSUPPORTED = -6

FD_OPT_EXACT_G = 1
FD_OPT_RPL_FP64 = 2
FD_OPT_TIMING = 4
FD_OPT_RPL_FP64_ALL = 8
FD_OPT_GRAPH = 16
FD_OPT_IEEE_DIV = 32
FD_OPT_TABLES = 64
FD_OPT_NO_TABLES = 128
FD_VAL_F32 = 0
FD_VAL_BF16 = 1
FD_VAL_FP16 = 2

TIMING_SLOTS = ("terma", "cccs", "cone_sum", "resample_count", "resample_values", "resample_compact",
                "cccs_tables")

MAX_TABLE = 24


class Volume(C.Structure):
    """DENSITY_h geometry (density.cuh:9-28)."""

    _fields_ = [("voxel_size", C.c_float * 3), ("dim", C.c_uint32 * 3), ("bbox_start", C.c_uint32 * 3),
                ("bbox_dim", C.c_uint32 * 3)]


class Beam(C.Structure):
    """BEAM_h (beam.cuh:10-34) without the fluence vector."""

    _fields_ = [("isocenter", C.c_float * 3), ("beamlet_size", C.c_float * 2), ("fmap_size", C.c_uint32 * 2),
                ("sad", C.c_float), ("angles", C.c_float * 3), ("long_spacing", C.c_float),
                ("lim_min", C.c_float), ("lim_max", C.c_float), ("long_dim", C.c_uint32),
                ("source", C.c_float * 3)]


class Spectrum(C.Structure):
    """SPECTRUM_h (spectrum.cuh:10-20)."""

    _fields_ = [("n", C.c_int32), ("energy", C.c_float * MAX_TABLE), ("fluence", C.c_float * MAX_TABLE),
                ("mu", C.c_float * MAX_TABLE), ("mu_en", C.c_float * MAX_TABLE)]


class Kernel(C.Structure):
    """KERNEL_h (kernel.cuh:10-25)."""

    _fields_ = [("n_theta", C.c_int32), ("n_phi", C.c_int32), ("theta_begin", C.c_float * MAX_TABLE),
                ("theta_end", C.c_float * MAX_TABLE), ("theta_mid", C.c_float * MAX_TABLE),
                ("A", C.c_float * MAX_TABLE), ("a", C.c_float * MAX_TABLE), ("B", C.c_float * MAX_TABLE),
                ("b", C.c_float * MAX_TABLE), ("phi", C.c_float * MAX_TABLE)]


class Bundle(C.Structure):
    """Scalar fields of IMRT::BeamBundle (IMRTBeamBundle.cuh:9-30)."""

    _fields_ = [("isocenter", C.c_float * 3), ("beamlet_size", C.c_float * 2), ("fluence_dim", C.c_int32 * 2),
                ("sad", C.c_float), ("angles", C.c_float * 3), ("long_spacing", C.c_float),
                ("sub_dim", C.c_int32 * 2), ("sub_on", C.c_int32 * 2)]


class CSR(C.Structure):
    _fields_ = [("rows", C.c_uint64), ("nnz", C.c_uint64), ("capacity", C.c_uint64),
                ("offsets", C.c_void_p), ("columns", C.c_void_p), ("values", C.c_void_p)]


class ServeOpts(C.Structure):
    _fields_ = [("n_slots", C.c_int), ("max_batch_beamlets", C.c_int), ("hbm_budget", C.c_uint64),
                ("extent", C.c_float), ("options", C.c_int), ("threshold", C.c_float)]


class ServeStats(C.Structure):
    _fields_ = [("requests_done", C.c_uint64), ("beams_done", C.c_uint64), ("beamlets_done", C.c_uint64),
                ("batches", C.c_uint64), ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double),
                ("max_ms", C.c_double), ("first_submit_s", C.c_double), ("last_done_s", C.c_double),
                ("peak_inflight_bytes", C.c_uint64), ("peak_inflight_batches", C.c_int)]


class Tuning(C.Structure):
    """fd_tuning: launch parameters of the batched paths (fd_autotune)."""

    _fields_ = [("batch_beamlets", C.c_int), ("graph", C.c_int), ("terma_segments", C.c_int),
                ("beamlets_per_s", C.c_float)]


P = C.POINTER
vp = C.c_void_p
u8p = P(C.c_uint8)
fp = P(C.c_float)

# name -> (restype, argtypes)
SIGNATURES = {
    "fd_last_error": (C.c_char_p, []),
    "fd_version": (C.c_char_p, []),
    "fd_calc_range": (C.c_int, [P(Beam), P(Volume)]),
    "fd_beamlet_init": (C.c_int, [P(Bundle), C.c_int, C.c_int, P(Volume), P(Beam), fp]),
    "fd_beamlets_init": (C.c_int, [P(Bundle), u8p, P(Volume), P(Beam), fp, P(C.c_int)]),
    "fd_ptv_isocenter": (C.c_int, [u8p, P(Volume), fp]),
    "fd_beamlet_flags_host": (C.c_int, [P(Bundle), u8p, P(Volume), u8p]),
    "fd_beamlet_flags": (C.c_int, [vp, vp, P(Bundle), C.c_int, u8p, vp]),
    "fd_read_spectrum": (C.c_int, [C.c_char_p, P(Spectrum)]),
    "fd_read_kernel": (C.c_int, [C.c_char_p, C.c_int, P(Kernel)]),
    "fd_ctx_create": (C.c_int, [C.c_int, P(Volume), fp, P(Spectrum), P(Kernel), P(vp)]),
    "fd_ctx_destroy": (C.c_int, [vp]),
    "fd_ctx_set_options": (C.c_int, [vp, C.c_int]),
    "fd_ctx_density": (C.c_int, [vp, P(vp)]),
    "fd_ctx_timings": (C.c_int, [vp, fp, C.c_int]),
    "fd_ctx_timing_launches": (C.c_int, [vp, P(C.c_int), C.c_int]),
    "fd_ctx_set_threshold": (C.c_int, [vp, C.c_float]),
    "fd_terma": (C.c_int, [vp, P(Beam), C.c_int, vp, vp, vp, C.c_uint64, vp]),
    "fd_cccs": (C.c_int, [vp, P(Beam), C.c_int, vp, vp, vp, C.c_uint64, vp]),
    "fd_resample_csr": (C.c_int, [vp, P(Beam), C.c_int, vp, C.c_uint64, C.c_float, P(CSR), vp]),
    "fd_dense_to_csr": (C.c_int, [vp, C.c_uint64, C.c_uint64, P(CSR), vp]),
    "fd_dense_to_csr_noalloc": (C.c_int, [vp, C.c_uint64, C.c_uint64, P(CSR), vp]),
    "fd_cccs_segments": (C.c_int, [vp, P(Beam), C.c_int, P(C.c_int64)]),
    "fd_dosemat_create": (C.c_int, [C.c_uint64, C.c_uint64, P(vp)]),
    "fd_dosemat_destroy": (C.c_int, [vp]),
    "fd_dosemat_add_beam": (C.c_int, [vp, vp, P(Beam), C.c_int, fp, C.c_float, vp]),
    "fd_dosemat_add_beams": (C.c_int, [vp, vp, P(Beam), fp, P(C.c_int), C.c_int, C.c_int, C.c_float, vp]),
    "fd_dosemat_add_csr": (C.c_int, [vp, P(CSR), vp]),
    "fd_dosemat_info": (C.c_int, [vp, P(C.c_uint64), P(C.c_uint64), P(C.c_uint64)]),
    "fd_dosemat_matrix_size": (C.c_int, [vp, C.c_uint64, P(C.c_uint64), P(C.c_uint64)]),
    "fd_dosemat_copy_matrix": (C.c_int, [vp, C.c_uint64, P(C.c_uint64), P(C.c_uint64), fp]),
    "fd_dosemat_tofile": (C.c_int, [vp, C.c_char_p]),
    "fd_dosemat_fromfile": (C.c_int, [vp, C.c_char_p]),
    "fd_dosemat_merge_folders": (C.c_int, [C.POINTER(C.c_char_p), C.c_int, C.c_char_p]),
    "fd_write_fluence_map": (C.c_int, [C.c_char_p, u8p, C.c_uint64, C.c_uint64]),
    "fd_dosemat_pack": (C.c_int, [vp, C.c_int, vp]),
    "fd_dosemat_copy_packed": (C.c_int, [vp, fp, P(C.c_uint16)]),
    "fd_spmm_create": (C.c_int, [vp, C.c_int, P(vp), vp]),
    "fd_spmm_info": (C.c_int, [vp, P(C.c_uint64), P(C.c_uint64), P(C.c_uint64)]),
    "fd_spmm_run": (C.c_int, [vp, vp, C.c_int, vp, vp]),
    "fd_spmm_destroy": (C.c_int, [vp]),
    "fd_ctx_get_tuning": (C.c_int, [vp, P(Tuning)]),
    "fd_ctx_set_tuning": (C.c_int, [vp, P(Tuning)]),
    "fd_autotune": (C.c_int, [vp, P(Beam), fp, P(C.c_int), C.c_int, C.c_float, P(Tuning), vp]),
    "fd_serve_start": (C.c_int, [vp, P(ServeOpts), P(vp)]),
    "fd_serve_submit": (C.c_int, [vp, P(Beam), fp, P(C.c_int), C.c_int, P(C.c_uint64)]),
    "fd_serve_poll": (C.c_int, [vp, C.c_uint64, P(C.c_int)]),
    "fd_serve_wait": (C.c_int, [vp, C.c_uint64, C.c_double]),
    "fd_serve_fetch": (C.c_int, [vp, C.c_uint64, vp]),
    "fd_serve_stats_get": (C.c_int, [vp, P(ServeStats)]),
    "fd_serve_stats_reset": (C.c_int, [vp]),
    "fd_serve_stop": (C.c_int, [vp]),
}

_lib = None


class FastDoseError(RuntimeError):
    def __init__(self, code: int, func: str, msg: str):
        super().__init__(f"{func} failed ({code}): {msg}")
        self.code = code


def load() -> C.CDLL:
    """Load the library (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m fastdose_amd.build` first")
    lib = C.CDLL(str(LIB_PATH), mode=C.RTLD_GLOBAL if os.name == "posix" else 0)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(code: int, func: str) -> int:
    if code != FD_OK:
        msg = load().fd_last_error().decode(errors="replace")
        raise FastDoseError(code, func, msg)
    return code


def call(name: str, *args) -> int:
    return check(getattr(load(), name)(*args), name)
