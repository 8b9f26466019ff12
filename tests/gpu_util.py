"""Shared helpers of the GPU tests: run the HIP stages through the C-ABI and the oracle beside them."""
from __future__ import annotations

import os
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np

import fastdose_amd as fd
import oracle as orc

NPIX = 256


def log(msg: str) -> None:
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def beamlets_of(cfg):
    beams, flus = [], []
    for b, f in zip(cfg.bundles, cfg.flags):
        bb, ff = fd.beamlets_init(b, f, cfg.volume)
        beams.extend(bb)
        flus.append(ff)
    arr = fd.beam_array(beams)
    return arr, np.concatenate(flus, 0)


def gpu_stages(ctx, beams, flu, extent=2.0):
    """Terma -> CCCS -> resample on the device; host copies of every output."""
    import torch

    n = len(beams)
    Lmax = max(b.long_dim for b in beams)
    stride = Lmax * NPIX
    dev = torch.device("cuda:0")
    flu_d = torch.from_numpy(np.ascontiguousarray(flu, dtype=np.float32)).to(dev)
    T = torch.zeros(n * stride, dtype=torch.float32, device=dev)
    R = torch.zeros_like(T)
    D = torch.zeros_like(T)
    log(f"terma n={n} Lmax={Lmax}")
    ctx.terma(beams, flu_d, T, R, stride)
    torch.cuda.synchronize()
    log("cccs")
    ctx.cccs(beams, T, R, D, stride)
    torch.cuda.synchronize()
    log("resample")
    off, cols, vals = ctx.resample_csr(beams, D, stride, extent)
    torch.cuda.synchronize()
    out = {
        "stride": stride,
        "terma": T.cpu().numpy().reshape(n, Lmax, NPIX),
        "dens": R.cpu().numpy().reshape(n, Lmax, NPIX),
        "dose": D.cpu().numpy().reshape(n, Lmax, NPIX),
        "offsets": off.cpu().numpy(),
        "cols": cols.cpu().numpy(),
        "vals": vals.cpu().numpy(),
        "dose_dev": D,
    }
    return out


def oracle_beamlet(cfg, beam, flu, flags=0):
    return orc.beamlet_pipeline(orc.cast(orc.Beam, beam), flu, orc.cast(orc.Volume, cfg.volume), cfg.density,
                                orc.cast(orc.Spectrum, cfg.spectrum), orc.cast(orc.Kernel, cfg.kernel), cfg.extent,
                                flags)


def oracle_many(cfg, items, flags=0):
    """beamlet_pipeline of the FP64 oracle for each (beam, fluence) of items, on the host cores (the
    oracle's ctypes calls release the interpreter lock)."""
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=max(1, min(16, threads))) as ex:
        return list(ex.map(lambda it: oracle_beamlet(cfg, it[0], it[1], flags), items))


def check_row(vol, cols, vals, ref, tol=0.005, tag=""):
    """One CSR row (a beamlet's dose column) against the oracle's: |D - D_fp64| <= tol x the beamlet's
    maximum over the union of columns (missing = 0), gamma 1 %/1 mm >= 99 %, ascending columns, and
    columns present on one side only carry at most 1e-6 of the maximum.  Returns the relative error."""
    cols = np.asarray(cols).astype(np.int64)
    vals = np.asarray(vals)
    assert np.all(np.diff(cols) > 0), (tag, "columns not ascending")
    err, sym = dense_diff(cols, vals, ref["cols"], ref["vals"])
    err /= max(float(ref["vals"].max()), 1e-30)
    assert err <= tol, (tag, err)
    gp = gamma_pass_rate(vol, ref["cols"], ref["vals"], cols, vals, dd=0.01, dta=0.1)
    assert gp >= 0.99, (tag, gp)
    if len(sym):
        tiny = 1e-6 * ref["vals"].max()
        vg = dict(zip(cols.tolist(), vals.tolist()))
        vr = dict(zip(ref["cols"].tolist(), ref["vals"].tolist()))
        for c in sym.tolist():
            assert max(vg.get(c, 0.0), vr.get(c, 0.0)) <= tiny, (tag, c)
        assert len(sym) <= 1e-4 * len(np.union1d(cols, ref["cols"])) + 2, tag
    return err


def same_matrix(a, b):
    """Two host CSR matrices (offsets, columns, values) equal bit for bit."""
    return all(x.dtype == y.dtype and np.array_equal(x, y) for x, y in zip(a, b))


def row(g, i):
    a, b = int(g["offsets"][i]), int(g["offsets"][i + 1])
    return g["cols"][a:b], g["vals"][a:b]


def dense_diff(c1, v1, c2, v2):
    """max |v1 - v2| over the union of columns (missing = 0) and the symmetric difference."""
    u = np.union1d(c1, c2)
    a = np.zeros(len(u))
    b = np.zeros(len(u))
    a[np.searchsorted(u, c1)] = v1
    b[np.searchsorted(u, c2)] = v2
    sym = np.setxor1d(c1, c2)
    return float(np.abs(a - b).max()) if len(u) else 0.0, sym


def gamma_pass_rate(vol, cols_ref, vals_ref, cols_test, vals_test, dd=0.01, dta=0.1, steps=8):
    """Global gamma index (Low et al. 1998): dose criterion dd x the maximum reference dose, distance
    criterion dta [cm].  Evaluated at every voxel where either column is nonzero (a missing entry is
    0): the smallest Gamma^2 = |d|^2 / dta^2 + (R(x + d) - T(x))^2 / (dd max R)^2 over displacements d
    with |d| <= dta on a lattice of pitch dta / steps, the reference R interpolated trilinearly between
    voxel centres (0 outside the volume), so that distances below one voxel count.  Returns the
    fraction of points with Gamma <= 1."""
    cols_ref = np.asarray(cols_ref, dtype=np.int64)
    cols_test = np.asarray(cols_test, dtype=np.int64)
    if len(cols_ref) == 0 and len(cols_test) == 0:
        return 1.0
    dims = np.array([int(vol.dim[0]), int(vol.dim[1]), int(vol.dim[2])])
    vs = np.array([vol.voxel_size[0], vol.voxel_size[1], vol.voxel_size[2]], dtype=np.float64)
    pts = np.union1d(cols_ref, cols_test)
    xyz = np.stack([pts % dims[0], (pts // dims[0]) % dims[1], pts // (dims[0] * dims[1])], 1)
    lo = np.maximum(xyz.min(0) - 2, 0)
    hi = np.minimum(xyz.max(0) + 3, dims)
    shape = tuple((hi - lo)[::-1] + 2)                    # (z, y, x) with a zero border of one voxel
    grid = np.zeros(shape)
    r = np.stack([cols_ref % dims[0], (cols_ref // dims[0]) % dims[1], cols_ref // (dims[0] * dims[1])], 1) - lo + 1
    grid[r[:, 2], r[:, 1], r[:, 0]] = np.asarray(vals_ref, dtype=np.float64)
    rmax = float(np.max(vals_ref)) if len(cols_ref) else float(np.max(vals_test))
    tval = np.zeros(len(pts))
    tval[np.searchsorted(pts, cols_test)] = np.asarray(vals_test, dtype=np.float64)
    base = (xyz - lo + 1).astype(np.float64)              # grid coordinates of the points

    def interp(q):                                        # trilinear, voxel centres at integers
        i0 = np.floor(q).astype(np.int64)
        w = q - i0
        out = np.zeros(len(q))
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    ix, iy, iz = i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz
                    ok = (ix >= 0) & (ix < shape[2]) & (iy >= 0) & (iy < shape[1]) & (iz >= 0) & (iz < shape[0])
                    wt = (w[:, 0] if dx else 1 - w[:, 0]) * (w[:, 1] if dy else 1 - w[:, 1]) * \
                         (w[:, 2] if dz else 1 - w[:, 2])
                    out[ok] += wt[ok] * grid[iz[ok], iy[ok], ix[ok]]
        return out

    g2 = ((interp(base) - tval) / (dd * rmax)) ** 2
    todo = np.nonzero(g2 > 1.0)[0]
    if len(todo):
        h = dta / steps
        k = np.arange(-steps, steps + 1) * h
        disp = np.stack(np.meshgrid(k, k, k, indexing="ij"), -1).reshape(-1, 3)
        r2 = (disp ** 2).sum(1) / dta ** 2
        keep = (r2 <= 1.0) & (r2 > 0)
        best = g2[todo]
        for d, rr in zip(disp[keep], r2[keep]):
            v = interp(base[todo] + d / vs)
            best = np.minimum(best, rr + ((v - tval[todo]) / (dd * rmax)) ** 2)
        g2[todo] = best
    return float(np.count_nonzero(g2 <= 1.0 + 1e-12)) / len(pts)
