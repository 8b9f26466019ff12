"""Python host mirror of FastDose's dose-mode API over the C-ABI.

Names follow the reference: ``calc_range`` (BEAM_h::calc_range, beam.cu:134-221),
``beamlet_init`` / ``beamlets_init`` (IMRTBeamBundle.cu:319-569), ``beamlet_flag_init``
(IMRTBeamBundle.cu:130-316), ``Context.terma`` (TermaComputeCollective), ``Context.cccs``
(DoseComputeCollective), ``Context.resample_csr`` (BEV2PVCSInterp + MatCSREnsemble::addMat),
``DoseMat`` (MatCSREnsemble with ``tofile``/``fromfile``) and ``dose_mat_construction``
(IMRTDoseMat.cu:14-301).  Errors raise FastDoseError (the reference returns true = failure).
Device buffers are torch tensors on ``cuda``; torch provides memory and streams only.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from . import _lib
from ._lib import Beam, Bundle, CSR, Kernel, ServeOpts, ServeStats, Spectrum, Volume, call, load

NPIX = 256


def _u8(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def _f32(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def beam_array(beams) -> C.Array:
    if isinstance(beams, C.Array):
        return beams
    arr = (Beam * len(beams))()
    for i, b in enumerate(beams):
        arr[i] = b
    return arr


def calc_range(beam: Beam, vol: Volume) -> Beam:
    call("fd_calc_range", C.byref(beam), C.byref(vol))
    return beam


def beamlet_init(bundle: Bundle, ix: int, iy: int, vol: Volume):
    out = Beam()
    flu = np.zeros(bundle.sub_dim[0] * bundle.sub_dim[1], dtype=np.float32)
    call("fd_beamlet_init", C.byref(bundle), ix, iy, C.byref(vol), C.byref(out), _f32(flu))
    return out, flu


def beamlets_init(bundle: Bundle, flags: np.ndarray, vol: Volume):
    """Active beamlets of a bundle in the reference's row order; returns (Beam array, fluence)."""
    flags = np.ascontiguousarray(flags, dtype=np.uint8)
    nmax = bundle.fluence_dim[0] * bundle.fluence_dim[1]
    if flags.size != nmax:
        raise ValueError("flags must have fluence_dim^2 entries")
    arr = (Beam * max(nmax, 1))()
    npix = bundle.sub_dim[0] * bundle.sub_dim[1]
    flu = np.zeros((max(nmax, 1), npix), dtype=np.float32)
    n = C.c_int(0)
    call("fd_beamlets_init", C.byref(bundle), _u8(flags), C.byref(vol), arr, _f32(flu), C.byref(n))
    out = (Beam * n.value)()
    C.memmove(out, arr, C.sizeof(Beam) * n.value)
    return out, flu[: n.value].copy()


def ptv_isocenter(mask: np.ndarray, vol: Volume):
    mask = np.ascontiguousarray(mask, dtype=np.uint8)
    iso = np.zeros(3, dtype=np.float32)
    call("fd_ptv_isocenter", _u8(mask), C.byref(vol), _f32(iso))
    return iso


def beamlet_flag_init(bundle: Bundle, ptv_mask: np.ndarray, vol: Volume) -> np.ndarray:
    mask = np.ascontiguousarray(ptv_mask, dtype=np.uint8)
    flags = np.zeros(bundle.fluence_dim[0] * bundle.fluence_dim[1], dtype=np.uint8)
    call("fd_beamlet_flags_host", C.byref(bundle), _u8(mask), C.byref(vol), _u8(flags))
    return flags


def read_spectrum(path) -> Spectrum:
    s = Spectrum()
    call("fd_read_spectrum", str(path).encode(), C.byref(s))
    return s


def read_kernel(path, n_phi: int = 8) -> Kernel:
    k = Kernel()
    call("fd_read_kernel", str(path).encode(), n_phi, C.byref(k))
    return k


def _stream_arg(stream):
    if stream is None:
        return None
    return C.c_void_p(int(getattr(stream, "cuda_stream", stream)))


class Context:
    """Device-resident density + tables (density_h2d, bind_spectrum, bind_kernel)."""

    def __init__(self, vol: Volume, density: np.ndarray, spectrum: Spectrum, kernel: Kernel, device: int = 0,
                 options: int = 0):
        density = np.ascontiguousarray(density, dtype=np.float32)
        nvox = vol.dim[0] * vol.dim[1] * vol.dim[2]
        if density.size != nvox:
            raise ValueError("density size does not match the volume dims")
        self.vol = vol
        self.spectrum = spectrum
        self.kernel = kernel
        h = C.c_void_p()
        call("fd_ctx_create", device, C.byref(vol), _f32(density), C.byref(spectrum), C.byref(kernel), C.byref(h))
        self._h = h
        if options:
            self.set_options(options)

    @property
    def handle(self):
        return self._h

    def set_options(self, options: int) -> None:
        call("fd_ctx_set_options", self._h, options)

    def set_threshold(self, rel: float) -> None:
        """Keep CSR entries >= rel x (the beamlet's maximum); 0 drops exact zeros only."""
        call("fd_ctx_set_threshold", self._h, rel)

    @property
    def tuning(self) -> dict:
        """Launch parameters of the batched paths (batch_beamlets, graph, terma_segments, beamlets_per_s)."""
        t = _lib.Tuning()
        call("fd_ctx_get_tuning", self._h, C.byref(t))
        return {f: getattr(t, f) for f, _ in t._fields_}

    def set_tuning(self, batch_beamlets: int, graph: int, terma_segments: int) -> None:
        t = _lib.Tuning(batch_beamlets, graph, terma_segments, 0.0)
        call("fd_ctx_set_tuning", self._h, C.byref(t))

    def autotune(self, beamlets, fluence: np.ndarray, counts, extent: float = 2.0, stream=None) -> dict:
        """Time trial batches of these beams (fd_autotune) and keep the fastest launch setting."""
        arr = beam_array(beamlets)
        cnt = np.ascontiguousarray(counts, dtype=np.int32)
        if int(cnt.sum()) != len(arr):
            raise ValueError("counts must add up to the number of beamlets")
        flu = np.ascontiguousarray(fluence, dtype=np.float32)
        if flu.size != len(arr) * NPIX:
            raise ValueError("fluence must hold 256 floats per beamlet")
        t = _lib.Tuning()
        call("fd_autotune", self._h, arr, _f32(flu), cnt.ctypes.data_as(C.POINTER(C.c_int)), len(cnt), extent,
             C.byref(t), _stream_arg(stream))
        return {f: getattr(t, f) for f, _ in t._fields_}

    def timings(self) -> dict:
        """Per-kernel milliseconds of the last call that launched each kernel, summed over its
        launches (needs FD_OPT_TIMING)."""
        n = len(_lib.TIMING_SLOTS)
        ms = (C.c_float * n)()
        call("fd_ctx_timings", self._h, ms, n)
        return {k: float(ms[i]) for i, k in enumerate(_lib.TIMING_SLOTS) if ms[i] >= 0}

    def timing_launches(self) -> dict:
        """How many launches each kernel's timings() entry sums."""
        n = len(_lib.TIMING_SLOTS)
        cnt = (C.c_int * n)()
        call("fd_ctx_timing_launches", self._h, cnt, n)
        return {k: int(cnt[i]) for i, k in enumerate(_lib.TIMING_SLOTS) if cnt[i] > 0}

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            load().fd_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def beamlet_flags(self, mask_dev, bundles, stream=None) -> list:
        """d_beamletFlagInit on the device for many bundles; mask_dev: u8 device tensor (z, y, x)."""
        arr = (Bundle * len(bundles))(*bundles)
        sizes = [b.fluence_dim[0] * b.fluence_dim[1] for b in bundles]
        out = np.zeros(max(sum(sizes), 1), dtype=np.uint8)
        call("fd_beamlet_flags", self._h, C.c_void_p(mask_dev.data_ptr()), arr, len(bundles), _u8(out),
             _stream_arg(stream))
        res, o = [], 0
        for n in sizes:
            res.append(out[o:o + n].copy())
            o += n
        return res

    # ---- split stages on torch device tensors
    def terma(self, beams, fluence_dev, terma_dev, dens_dev, stride: int, stream=None) -> None:
        arr = beam_array(beams)
        call("fd_terma", self._h, arr, len(arr), C.c_void_p(fluence_dev.data_ptr()),
             C.c_void_p(terma_dev.data_ptr()), C.c_void_p(dens_dev.data_ptr()), stride, _stream_arg(stream))

    def cccs(self, beams, terma_dev, dens_dev, dose_dev, stride: int, stream=None) -> None:
        arr = beam_array(beams)
        call("fd_cccs", self._h, arr, len(arr), C.c_void_p(terma_dev.data_ptr()), C.c_void_p(dens_dev.data_ptr()),
             C.c_void_p(dose_dev.data_ptr()), stride, _stream_arg(stream))

    def resample_csr(self, beams, dose_dev, stride: int, extent: float = 2.0, stream=None):
        """Returns (offsets, columns, values) as device tensors."""
        import torch

        arr = beam_array(beams)
        n = len(arr)
        dev = dose_dev.device
        offsets = torch.zeros(n + 1, dtype=torch.int64, device=dev)
        csr = CSR(0, 0, 0, offsets.data_ptr(), None, None)
        rc = load().fd_resample_csr(self._h, arr, n, C.c_void_p(dose_dev.data_ptr()), stride, extent,
                                    C.byref(csr), _stream_arg(stream))
        if rc not in (_lib.FD_OK, _lib.FD_ERR_CAPACITY):
            _lib.check(rc, "fd_resample_csr")
        nnz = int(csr.nnz)
        cols = torch.empty(max(nnz, 1), dtype=torch.int64, device=dev)
        vals = torch.empty(max(nnz, 1), dtype=torch.float32, device=dev)
        if rc == _lib.FD_ERR_CAPACITY:
            csr = CSR(0, 0, nnz, offsets.data_ptr(), cols.data_ptr(), vals.data_ptr())
            call("fd_resample_csr", self._h, arr, n, C.c_void_p(dose_dev.data_ptr()), stride, extent, C.byref(csr),
                 _stream_arg(stream))
        return offsets, cols[:nnz], vals[:nnz]

    def cccs_segments(self, beams) -> np.ndarray:
        arr = beam_array(beams)
        out = np.zeros(len(arr), dtype=np.int64)
        call("fd_cccs_segments", self._h, arr, len(arr), out.ctypes.data_as(C.POINTER(C.c_int64)))
        return out


def dense_to_csr(dense_dev, stream=None):
    """cusparseDenseToSparse semantics on a (rows, cols) float32 device tensor."""
    import torch

    rows, cols = dense_dev.shape
    dev = dense_dev.device
    offsets = torch.zeros(rows + 1, dtype=torch.int64, device=dev)
    csr = CSR(0, 0, 0, offsets.data_ptr(), None, None)
    rc = load().fd_dense_to_csr(C.c_void_p(dense_dev.data_ptr()), rows, cols, C.byref(csr), _stream_arg(stream))
    if rc not in (_lib.FD_OK, _lib.FD_ERR_CAPACITY):
        _lib.check(rc, "fd_dense_to_csr")
    nnz = int(csr.nnz)
    c = torch.empty(max(nnz, 1), dtype=torch.int64, device=dev)
    v = torch.empty(max(nnz, 1), dtype=torch.float32, device=dev)
    if nnz:
        csr = CSR(0, 0, nnz, offsets.data_ptr(), c.data_ptr(), v.data_ptr())
        call("fd_dense_to_csr", C.c_void_p(dense_dev.data_ptr()), rows, cols, C.byref(csr), _stream_arg(stream))
    return offsets, c[:nnz], v[:nnz]


class DoseMat:
    """MatCSREnsemble: one CSR matrix per beam (rows = beamlets, columns = voxels)."""

    def __init__(self, n_cols: int, est_nnz: int = 0):
        h = C.c_void_p()
        call("fd_dosemat_create", n_cols, est_nnz, C.byref(h))
        self._h = h
        self.n_cols = n_cols

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            load().fd_dosemat_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_beam(self, ctx: Context, beamlets, fluence: np.ndarray, extent: float = 2.0, stream=None) -> None:
        arr = beam_array(beamlets)
        flu = np.ascontiguousarray(fluence, dtype=np.float32)
        if flu.size != len(arr) * NPIX:
            raise ValueError("fluence must hold 256 floats per beamlet")
        call("fd_dosemat_add_beam", ctx.handle, self._h, arr, len(arr), _f32(flu), extent, _stream_arg(stream))

    def add_beams(self, ctx: Context, beamlets, fluence: np.ndarray, counts, max_chunk: int = 0,
                  extent: float = 2.0, stream=None) -> None:
        """Batched: all beams' beamlets through one launch per stage (per chunk), one matrix per beam."""
        arr = beam_array(beamlets)
        cnt = np.ascontiguousarray(counts, dtype=np.int32)
        if int(cnt.sum()) != len(arr):
            raise ValueError("counts must add up to the number of beamlets")
        flu = np.ascontiguousarray(fluence, dtype=np.float32)
        if flu.size != len(arr) * NPIX:
            raise ValueError("fluence must hold 256 floats per beamlet")
        call("fd_dosemat_add_beams", ctx.handle, self._h, arr, _f32(flu), cnt.ctypes.data_as(C.POINTER(C.c_int)),
             len(cnt), max_chunk, extent, _stream_arg(stream))

    def info(self):
        a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        call("fd_dosemat_info", self._h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def matrix(self, k: int):
        rows, nnz = C.c_uint64(), C.c_uint64()
        call("fd_dosemat_matrix_size", self._h, k, C.byref(rows), C.byref(nnz))
        off = np.zeros(rows.value + 1, dtype=np.uint64)
        col = np.zeros(max(nnz.value, 1), dtype=np.uint64)
        val = np.zeros(max(nnz.value, 1), dtype=np.float32)
        call("fd_dosemat_copy_matrix", self._h, k, off.ctypes.data_as(C.POINTER(C.c_uint64)),
             col.ctypes.data_as(C.POINTER(C.c_uint64)), _f32(val))
        return off, col[: nnz.value], val[: nnz.value]

    def pack(self, value_type: int, stream=None) -> None:
        """16-bit values (FD_VAL_BF16 / FD_VAL_FP16) with a per-row FP32 scale; tofile then also writes
        valuesBuffer.{bf16,fp16}.bin and beamletScale.bin."""
        call("fd_dosemat_pack", self._h, value_type, _stream_arg(stream))

    def packed(self):
        """(per-row scales f32, packed values u16) of the last pack()."""
        _, rows, nnz = self.info()
        scale = np.zeros(max(rows, 1), dtype=np.float32)
        q = np.zeros(max(nnz, 1), dtype=np.uint16)
        call("fd_dosemat_copy_packed", self._h, _f32(scale), q.ctypes.data_as(C.POINTER(C.c_uint16)))
        return scale[:rows], q[:nnz]

    def tofile(self, folder) -> None:
        call("fd_dosemat_tofile", self._h, str(folder).encode())

    def fromfile(self, folder) -> None:
        call("fd_dosemat_fromfile", self._h, str(folder).encode())


def write_fluence_map(path, flags_per_beam) -> None:
    """beamletFlagSave: u8 flags of every beam, concatenated."""
    arr = np.ascontiguousarray(np.stack([np.asarray(f, dtype=np.uint8) for f in flags_per_beam]), dtype=np.uint8)
    call("fd_write_fluence_map", str(path).encode(), _u8(arr), arr.shape[0], arr.shape[1])


def merge_dosemat_folders(folders, out_folder) -> None:
    """doseMatMerge (scripts/doseMerge.py:9-63): the doseMatFolder file sets of `folders`, in order,
    concatenated into `out_folder` (fd_dosemat_merge_folders, host only)."""
    names = [str(f).encode() for f in folders]
    arr = (C.c_char_p * len(names))(*names)
    call("fd_dosemat_merge_folders", arr, len(names), str(out_folder).encode())


def read_dosemat_folder(folder):
    """readDose.m semantics (scripts/readDose.m:14-53): list of per-beam CSR (offsets, cols, vals)."""
    folder = Path(folder)
    rows = np.fromfile(folder / "numRowsPerMat.bin", dtype=np.uint64)
    offsets = np.fromfile(folder / "offsetsBuffer.bin", dtype=np.uint64)
    cols = np.fromfile(folder / "columnsBuffer.bin", dtype=np.uint64)
    vals = np.fromfile(folder / "valuesBuffer.bin", dtype=np.float32)
    mats = []
    io = 0
    jo = 0
    for r in rows:
        r = int(r)
        seg = offsets[io: io + r + 1]
        n = int(seg[-1])
        mats.append((seg.copy(), cols[jo: jo + n].copy(), vals[jo: jo + n].copy()))
        io += r + 1
        jo += n
    return mats


def bundle_beamlets(bundles, flags_list, vol: Volume):
    """Beamlets of every bundle, concatenated: (Beam array, fluence (n, 256), per-beam counts)."""
    all_beams, flus, counts = [], [], []
    for bundle, flags in zip(bundles, flags_list):
        beams, flu = beamlets_init(bundle, flags, vol)
        all_beams.extend(beams)
        flus.append(flu)
        counts.append(len(beams))
    flu = np.concatenate(flus, 0) if flus else np.zeros((0, NPIX), np.float32)
    return beam_array(all_beams), flu, np.array(counts, dtype=np.int32)


def dose_mat_construction(ctx: Context, bundles, flags_list, extent: float = 2.0, est_nnz: int = 0,
                          stream=None, batched: bool = False, max_chunk: int = 0) -> DoseMat:
    """DoseMatConstruction (IMRTDoseMat.cu:14-301): one matrix per beam, beams in order.  batched=True
    sends all beams through the kernels together (fd_dosemat_add_beams)."""
    nvox = ctx.vol.dim[0] * ctx.vol.dim[1] * ctx.vol.dim[2]
    dm = DoseMat(nvox, est_nnz)
    if batched:
        beams, flu, counts = bundle_beamlets(bundles, flags_list, ctx.vol)
        dm.add_beams(ctx, beams, flu, counts, max_chunk, extent, stream)
        return dm
    for bundle, flags in zip(bundles, flags_list):
        beams, flu = beamlets_init(bundle, flags, ctx.vol)
        dm.add_beam(ctx, beams, flu, extent, stream)
    return dm


class Spmm:
    """Voxel-major copy of a DoseMat for Y = D^T X (fluence vectors -> dose), FP32 accumulation."""

    def __init__(self, dm: DoseMat, value_type: int = 0, stream=None):
        h = C.c_void_p()
        call("fd_spmm_create", dm._h, value_type, C.byref(h), _stream_arg(stream))
        self._h = h

    def info(self):
        r, c, n = C.c_uint64(), C.c_uint64(), C.c_uint64()
        call("fd_spmm_info", self._h, C.byref(r), C.byref(c), C.byref(n))
        return r.value, c.value, n.value

    def run(self, x_dev, y_dev=None, stream=None):
        """x_dev: (rows, k) float32 device tensor; returns y (n_cols, k)."""
        import torch

        rows, cols, _ = self.info()
        if x_dev.dtype != torch.float32 or not x_dev.is_contiguous() or x_dev.dim() != 2 or x_dev.shape[0] != rows:
            raise ValueError("x must be a contiguous float32 (rows, k) tensor")
        k = int(x_dev.shape[1])
        if y_dev is None:
            y_dev = torch.empty((cols, k), dtype=torch.float32, device=x_dev.device)
        call("fd_spmm_run", self._h, C.c_void_p(x_dev.data_ptr()), k, C.c_void_p(y_dev.data_ptr()),
             _stream_arg(stream))
        return y_dev

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            load().fd_spmm_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Server:
    """Resident request loop with continuous batching over a Context (fd_serve_*).

    submit() queues a request of one or more beams and returns its id at once; slot threads pack
    queued beams into batches (admitted by their HBM footprint) that run concurrently on their own
    streams; fetch() returns the request's per-beam matrices."""

    def __init__(self, ctx: Context, n_slots: int = 3, max_batch_beamlets: int = 4096, hbm_budget: int = 0,
                 extent: float = 2.0, options: int = 0, threshold: float = 0.0):
        self.ctx = ctx
        opts = ServeOpts(n_slots, max_batch_beamlets, hbm_budget, extent, options, threshold)
        h = C.c_void_p()
        call("fd_serve_start", ctx.handle, C.byref(opts), C.byref(h))
        self._h = h

    def submit(self, beamlets, fluence: np.ndarray, counts) -> int:
        arr = beam_array(beamlets)
        cnt = np.ascontiguousarray(counts, dtype=np.int32)
        if int(cnt.sum()) != len(arr):
            raise ValueError("counts must add up to the number of beamlets")
        flu = np.ascontiguousarray(fluence, dtype=np.float32)
        if flu.size != len(arr) * NPIX:
            raise ValueError("fluence must hold 256 floats per beamlet")
        rid = C.c_uint64()
        call("fd_serve_submit", self._h, arr, _f32(flu), cnt.ctypes.data_as(C.POINTER(C.c_int)), len(cnt),
             C.byref(rid))
        return rid.value

    def poll(self, rid: int) -> int:
        st = C.c_int()
        call("fd_serve_poll", self._h, rid, C.byref(st))
        return st.value

    def wait(self, rid: int, timeout: float = 0.0) -> None:
        call("fd_serve_wait", self._h, rid, timeout)

    def fetch(self, rid: int, into: "DoseMat | None" = None) -> "DoseMat":
        vol = self.ctx.vol
        dm = into if into is not None else DoseMat(vol.dim[0] * vol.dim[1] * vol.dim[2])
        call("fd_serve_fetch", self._h, rid, dm._h)
        return dm

    def stats(self) -> dict:
        st = ServeStats()
        call("fd_serve_stats_get", self._h, C.byref(st))
        return {k: getattr(st, k) for k, _ in ServeStats._fields_}

    def reset_stats(self) -> None:
        call("fd_serve_stats_reset", self._h)

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            load().fd_serve_stop(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
