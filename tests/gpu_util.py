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
