"""Shared helpers of the GPU tests: run the HIP stages through the C-ABI and the oracle beside them."""
from __future__ import annotations

import sys

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


def gamma_pass_rate(vol, cols_ref, vals_ref, cols_test, vals_test, dd=0.01, dta=0.1):
    """Global gamma index (dose criterion dd x max reference dose, distance criterion dta in cm) of a
    test dose column against the reference, evaluated at every nonzero reference voxel; the search
    covers test voxels within 3 dta (voxel grid of the volume).  Returns the fraction with gamma <= 1."""
    cols_ref = np.asarray(cols_ref, dtype=np.int64)
    cols_test = np.asarray(cols_test, dtype=np.int64)
    if len(cols_ref) == 0:
        return 1.0
    X, Y = int(vol.dim[0]), int(vol.dim[1])
    vs = np.array([vol.voxel_size[0], vol.voxel_size[1], vol.voxel_size[2]], dtype=np.float64)
    test = dict(zip(cols_test.tolist(), np.asarray(vals_test, dtype=np.float64).tolist()))
    dmax = float(np.max(vals_ref))
    reach = [int(np.floor(3 * dta / vs[k])) for k in range(3)]
    offs = [(i, j, k) for i in range(-reach[0], reach[0] + 1) for j in range(-reach[1], reach[1] + 1)
            for k in range(-reach[2], reach[2] + 1)]
    passed = 0
    for c, v in zip(cols_ref.tolist(), np.asarray(vals_ref, dtype=np.float64).tolist()):
        x, y, z = c % X, (c // X) % Y, c // (X * Y)
        best = np.inf
        for i, j, k in offs:
            cc = (x + i) + X * ((y + j) + Y * (z + k))
            if not (0 <= x + i < X and 0 <= y + j < Y and 0 <= z + k < vol.dim[2]):
                continue
            r2 = ((i * vs[0]) ** 2 + (j * vs[1]) ** 2 + (k * vs[2]) ** 2) / dta ** 2
            d2 = ((test.get(cc, 0.0) - v) / (dd * dmax)) ** 2
            best = min(best, r2 + d2)
        passed += best <= 1.0
    return passed / len(cols_ref)
