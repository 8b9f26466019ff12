"""Relative CSR threshold, 16-bit packing with per-beamlet scale, and D^T X products (needs an
MI355X).  Precision config: the thresholded BF16/FP16 matrix times 256 fluence vectors stays
within the stated error of the same product through the FP64 oracle's matrix."""
import numpy as np
import pytest

import fastdose_amd as fd
import oracle as orc
from fastdose_amd import synth

from gpu_util import beamlets_of, log

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def bf16_rne(x: np.ndarray) -> np.ndarray:
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def bf16_to_f32(h: np.ndarray) -> np.ndarray:
    return (h.astype(np.uint32) << 16).view(np.float32)


def ensemble_csr(dm):
    """All matrices of a DoseMat as one scipy CSR (rows in order)."""
    import scipy.sparse as sp

    nm = dm.info()[0]
    blocks = []
    for k in range(nm):
        off, col, val = dm.matrix(k)
        blocks.append(sp.csr_matrix((val.astype(np.float64), col.astype(np.int64), off.astype(np.int64)),
                                    shape=(len(off) - 1, dm.n_cols)))
    return sp.vstack(blocks).tocsr()


def test_threshold_is_a_filter_of_the_exact_rows(torch_cuda):
    cfg = synth.small_config(seed=15, dim=(26, 24, 22), fluence_dim=3)
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel) as ctx:
        full = fd.dose_mat_construction(ctx, cfg.bundles, cfg.flags, cfg.extent)
        ctx.set_threshold(1e-3)
        thr = fd.dose_mat_construction(ctx, cfg.bundles, cfg.flags, cfg.extent)
        thr_b = fd.dose_mat_construction(ctx, cfg.bundles, cfg.flags, cfg.extent, batched=True)
    off, col, val = full.matrix(0)
    o2, c2, v2 = thr.matrix(0)
    assert all(np.array_equal(x, y) for x, y in zip(thr.matrix(0), thr_b.matrix(0)))
    for r in range(len(off) - 1):
        c, v = col[off[r]:off[r + 1]], val[off[r]:off[r + 1]]
        keep = v >= np.float32(1e-3) * v.max()
        assert np.array_equal(c[keep], c2[o2[r]:o2[r + 1]]) and np.array_equal(v[keep], v2[o2[r]:o2[r + 1]])
    assert len(c2) < len(col)


def test_pack_and_spmm(torch_cuda, tmp_path):
    torch = torch_cuda
    cfg = synth.small_config(seed=16, dim=(28, 26, 24), fluence_dim=3)
    iso = cfg.bundles[0].isocenter
    cfg.bundles.append(synth.make_bundle(iso, (140.0, 30.0, 10.0), fluence_dim=3))
    cfg.flags.append(np.ones(9, dtype=np.uint8))
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel) as ctx:
        dm = fd.dose_mat_construction(ctx, cfg.bundles, cfg.flags, cfg.extent)
    D = ensemble_csr(dm)
    R, C = D.shape
    vals = np.concatenate([dm.matrix(k)[2] for k in range(dm.info()[0])])
    # packing: scale = row max, q = RNE(v / scale)
    for vt in (fd.FD_VAL_BF16, fd.FD_VAL_FP16):
        dm.pack(vt)
        scale, q = dm.packed()
        rows = np.repeat(np.arange(R), np.diff(D.indptr))
        rmax = np.array([vals[D.indptr[r]:D.indptr[r + 1]].max() for r in range(R)], dtype=np.float32)
        assert np.array_equal(scale, rmax)
        t = (vals / rmax[rows]).astype(np.float32)
        expect = bf16_rne(t) if vt == fd.FD_VAL_BF16 else t.astype(np.float16).view(np.uint16)
        assert np.array_equal(q, expect)
        folder = tmp_path / f"dm{vt}"
        dm.tofile(folder)
        name = "valuesBuffer.bf16.bin" if vt == fd.FD_VAL_BF16 else "valuesBuffer.fp16.bin"
        assert np.array_equal(np.fromfile(folder / name, dtype=np.uint16), q)
        assert np.array_equal(np.fromfile(folder / "beamletScale.bin", dtype=np.float32), scale)
    # D^T X for 256 and an odd k, three value types
    rng = np.random.default_rng(0)
    for k in (256, 7):
        X = rng.uniform(size=(R, k)).astype(np.float32)
        Xd = torch.from_numpy(X).cuda()
        for vt in (fd.FD_VAL_F32, fd.FD_VAL_BF16, fd.FD_VAL_FP16):
            with fd.Spmm(dm, vt) as p:
                assert p.info() == (R, C, D.nnz)
                Y = p.run(Xd).cpu().numpy()
                Y2 = p.run(Xd).cpu().numpy()
            assert np.array_equal(Y, Y2)  # fixed summation order
            if vt == fd.FD_VAL_F32:
                Dq = D
            else:
                rows = np.repeat(np.arange(R), np.diff(D.indptr))
                rmax = np.array([D.data[D.indptr[r]:D.indptr[r + 1]].max() for r in range(R)], dtype=np.float32)
                t = (D.data.astype(np.float32) / rmax[rows]).astype(np.float32)
                deq = bf16_to_f32(bf16_rne(t)) if vt == fd.FD_VAL_BF16 else t.astype(np.float16).astype(np.float32)
                Dq = D.copy()
                Dq.data = deq.astype(np.float64) * rmax[rows].astype(np.float64)
            ref = np.asarray(Dq.T @ X.astype(np.float64))
            err = np.abs(Y - ref).max() / np.abs(ref).max()
            log(f"spmm k={k} vt={vt} rel err vs same-precision FP64 product {err:.2e}")
            assert err < 2e-6


def test_precision_config_error_vs_oracle(torch_cuda):
    """Thresholded (1e-3) BF16 / FP16 dose-loading matrix x 256 U[0,1] fluence vectors vs the same
    product through the FP64 oracle's exact matrix: max error <= 0.5 % of the maximum dose."""
    torch = torch_cuda
    cfg = synth.small_config(seed=17, dim=(30, 28, 26), fluence_dim=3)
    beams, flu = beamlets_of(cfg)
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel) as ctx:
        ctx.set_threshold(1e-3)
        dm = fd.dose_mat_construction(ctx, cfg.bundles, cfg.flags, cfg.extent)
    import scipy.sparse as sp

    R = len(beams)
    nvox = dm.n_cols
    rows, cols, vals = [], [], []
    for i in range(R):
        ref = orc.beamlet_pipeline(orc.cast(orc.Beam, beams[i]), flu[i], orc.cast(orc.Volume, cfg.volume),
                                   cfg.density, orc.cast(orc.Spectrum, cfg.spectrum), orc.cast(orc.Kernel, cfg.kernel),
                                   cfg.extent, 0)
        rows.append(np.full(len(ref["cols"]), i))
        cols.append(ref["cols"])
        vals.append(ref["vals"])
    Dref = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(R, nvox))
    X = np.random.default_rng(1).uniform(size=(R, 256)).astype(np.float32)
    Yref = np.asarray(Dref.T @ X.astype(np.float64))
    Xd = torch.from_numpy(X).cuda()
    for vt in (fd.FD_VAL_F32, fd.FD_VAL_BF16, fd.FD_VAL_FP16):
        with fd.Spmm(dm, vt) as p:
            Y = p.run(Xd).cpu().numpy()
        err = np.abs(Y - Yref).max() / np.abs(Yref).max()
        log(f"precision config vt={vt}: max |Y - Y_fp64| / max Y_fp64 = {err:.2e}")
        assert err <= 0.005
