"""Synthetic inputs of the benchmark configurations (seeded, reproducible).

Spectrum: 12 bins E_k = 0.5 + 0.5 k MeV, monotonically decreasing seeded weights, mu/rho =
0.0707 E^-0.523 and mu_en/rho = 0.0310 E^-0.30 (power laws fitted to scripts/spec_6mv.spec).
Kernel: 8 zenith bins of pi/8 (4 forward cones) x n_phi azimuths, seeded (A, a, B, b) per cone.
Seeds are 1000 + config number.  Volumes are float32 arrays shaped (Z, Y, X), x fastest, as
density.raw is laid out.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from ._lib import Bundle, Kernel, Spectrum, Volume

PI_F = np.float32(3.141592654)  # CUDART_PI_F


def make_volume(dim, voxel=0.25, bbox_start=(0, 0, 0), bbox_dim=None) -> Volume:
    v = Volume()
    for i in range(3):
        v.voxel_size[i] = voxel
        v.dim[i] = int(dim[i])
        v.bbox_start[i] = int(bbox_start[i])
        v.bbox_dim[i] = int(dim[i] if bbox_dim is None else bbox_dim[i])
    return v


def synthetic_spectrum(seed: int, n: int = 12) -> Spectrum:
    rng = np.random.default_rng(seed)
    energy = 0.5 + 0.5 * np.arange(n)
    w = np.sort(rng.uniform(0.0, 1.0, n))[::-1]
    w = w / w.sum()
    s = Spectrum()
    s.n = n
    for i in range(n):
        s.energy[i] = energy[i]
        s.fluence[i] = w[i]
        s.mu[i] = 0.0707 * energy[i] ** -0.523
        s.mu_en[i] = 0.0310 * energy[i] ** -0.30
    return s


def synthetic_kernel(seed: int, n_phi: int = 8, n_theta: int = 8) -> Kernel:
    rng = np.random.default_rng(seed + 7919)
    k = Kernel()
    k.n_theta = n_theta
    k.n_phi = n_phi
    width = np.float32(math.pi / n_theta)
    A = np.sort(rng.uniform(0.5, 2.2, n_theta))[::-1]
    for i in range(n_theta):
        tb = np.float32(i * width)
        te = np.float32((i + 1) * width)
        k.theta_begin[i] = tb
        k.theta_end[i] = te
        k.theta_mid[i] = (tb + te) / np.float32(2)
        k.A[i] = A[i]
        k.a[i] = rng.uniform(2.0, 8.0)
        k.B[i] = rng.uniform(1e-3, 2.5e-2)
        k.b[i] = rng.uniform(0.05, 0.11)
    itv = np.float32(2) * PI_F / np.float32(n_phi)
    for i in range(n_phi):
        k.phi[i] = np.float32((i + 0.5) * float(itv))
    return k


def degrees_to_rad(deg: float) -> np.float32:
    """BeamBundleInit's conversion (IMRTBeamBundle.cu:90): FP32 deg * pi / 180."""
    return np.float32(np.float32(deg) * PI_F) / np.float32(180)


def make_bundle(iso, angles_deg, fluence_dim, beamlet_size=0.5, sad=100.0, long_spacing=0.25, sub_dim=16,
                sub_on=4) -> Bundle:
    b = Bundle()
    for i in range(3):
        b.isocenter[i] = iso[i]
        b.angles[i] = degrees_to_rad(angles_deg[i])
    b.beamlet_size[0] = b.beamlet_size[1] = beamlet_size
    b.fluence_dim[0] = b.fluence_dim[1] = fluence_dim
    b.sad = sad
    b.long_spacing = long_spacing
    b.sub_dim[0] = b.sub_dim[1] = sub_dim
    b.sub_on[0] = b.sub_on[1] = sub_on
    return b


@dataclass
class Config:
    name: str
    volume: Volume
    density: np.ndarray             # (Z, Y, X) float32
    spectrum: Spectrum
    kernel: Kernel
    bundles: list                   # [Bundle]
    flags: list                     # [np.ndarray uint8 (fluence_dim^2,)]
    extent: float = 2.0
    ptv: np.ndarray | None = None
    notes: dict = field(default_factory=dict)


def config1() -> Config:
    """64^3 water, one beam at gantry 0 through the volume centre, 8x8 beamlets of 0.5 cm."""
    seed = 1001
    dim = (64, 64, 64)
    vol = make_volume(dim)
    density = np.ones(dim[::-1], dtype=np.float32)
    iso = [np.float32(0.25 * d * 0.5) for d in dim]
    b = make_bundle(iso, (0.0, 0.0, 0.0), fluence_dim=8)
    return Config("plain-small-64^3-water", vol, density, synthetic_spectrum(seed), synthetic_kernel(seed), [b],
                  [np.ones(64, dtype=np.uint8)])


def slab_density(dim, seed: int) -> np.ndarray:
    """Slabs along y (beam depth at gantry 0): water / lung 0.25 / bone 1.8 / water."""
    rng = np.random.default_rng(seed)
    Y = dim[1]
    f = rng.dirichlet([4.0, 4.0, 2.0, 4.0])
    t = np.maximum(1, np.round(f * Y).astype(int))
    t[-1] = Y - t[:-1].sum()
    rho_y = np.concatenate([np.full(n, r, dtype=np.float32) for n, r in zip(t, (1.0, 0.25, 1.8, 1.0))])
    return np.broadcast_to(rho_y[None, :, None], dim[::-1]).astype(np.float32).copy()


def config2() -> Config:
    """128^3 slab phantom, one seeded non-coplanar beam, 16x16 beamlets of 0.5 cm."""
    seed = 1002
    dim = (128, 128, 128)
    vol = make_volume(dim)
    density = slab_density(dim, seed)
    rng = np.random.default_rng(seed + 1)
    gantry = float(rng.uniform(20.0, 60.0))
    couch = float(rng.uniform(15.0, 45.0))
    iso = [np.float32(0.25 * d * 0.5) for d in dim]
    b = make_bundle(iso, (gantry, couch, 0.0), fluence_dim=16)
    cfg = Config("plain-slab-128^3", vol, density, synthetic_spectrum(seed), synthetic_kernel(seed), [b],
                 [np.ones(256, dtype=np.uint8)])
    cfg.notes = {"gantry_deg": gantry, "couch_deg": couch}
    return cfg


def small_config(seed: int = 7, dim=(24, 28, 20), fluence_dim: int = 2, angles=(30.0, 20.0, 10.0),
                 hetero: bool = True) -> Config:
    """A tiny heterogeneous case the FP64 oracle finishes in well under a second."""
    rng = np.random.default_rng(seed)
    vol = make_volume(dim)
    if hetero:
        density = rng.uniform(0.2, 1.8, size=dim[::-1]).astype(np.float32)
    else:
        density = np.ones(dim[::-1], dtype=np.float32)
    iso = [np.float32(0.25 * d * 0.5) for d in dim]
    b = make_bundle(iso, angles, fluence_dim=fluence_dim)
    return Config(f"small-{seed}", vol, density, synthetic_spectrum(1000 + seed), synthetic_kernel(1000 + seed), [b],
                  [np.ones(fluence_dim * fluence_dim, dtype=np.uint8)])
