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


def direction_to_varian_iec(direction):
    """(gantry, couch) in rad of a source->isocenter direction (dockerScripts/HeadNeck/scripts.py:47-62)."""
    eps = 1e-4
    cos_g = float(direction[1])
    sin_g = math.sqrt(max(0.0, 1.0 - cos_g * cos_g))
    gantry = math.acos(max(-1.0, min(1.0, cos_g)))
    if sin_g < eps:
        minus_couch = 0.0
    else:
        sin_mc = direction[2] / sin_g
        cos_mc = max(-1.0, min(1.0, -direction[0] / sin_g))
        minus_couch = math.acos(cos_mc)
        if sin_mc < 0:
            minus_couch = -minus_couch
    return gantry, -minus_couch


def fibonacci_sphere(n: int) -> np.ndarray:
    """n unit vectors spread evenly over the sphere (golden-angle spiral)."""
    i = np.arange(n) + 0.5
    z = 1.0 - 2.0 * i / n
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = math.pi * (3.0 - math.sqrt(5.0)) * i
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)


def body_density(dim, seed: int, sigma: float = 4.0):
    """Gaussian-smoothed seeded noise mapped to [0.2, 1.8] g/cc inside an ellipsoidal body, 0.0012
    outside.  Returns (density (Z, Y, X) float32, body mask uint8)."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    noise = gaussian_filter(rng.standard_normal(dim[::-1]).astype(np.float32), sigma=sigma)
    z, y, x = np.meshgrid(*(np.arange(d, dtype=np.float32) + 0.5 for d in dim[::-1]), indexing="ij")
    cx, cy, cz = (d / 2.0 for d in dim)
    body = ((x - cx) / (0.45 * dim[0])) ** 2 + ((y - cy) / (0.40 * dim[1])) ** 2 + \
        ((z - cz) / (0.47 * dim[2])) ** 2 <= 1.0
    inside = noise[body]
    lo, hi = float(inside.min()), float(inside.max())
    rho = np.full(dim[::-1], 0.0012, dtype=np.float32)
    rho[body] = 0.2 + 1.6 * (inside - lo) / (hi - lo)
    return rho, body.astype(np.uint8)


def ptv_segments(dim, center, radii, n_seg: int = 4):
    """An ellipsoidal PTV (voxel units) split into n_seg slabs along x: list of uint8 masks."""
    z, y, x = np.meshgrid(*(np.arange(d, dtype=np.float32) + 0.5 for d in dim[::-1]), indexing="ij")
    ell = ((x - center[0]) / radii[0]) ** 2 + ((y - center[1]) / radii[1]) ** 2 + \
        ((z - center[2]) / radii[2]) ** 2 <= 1.0
    edges = np.linspace(center[0] - radii[0], center[0] + radii[0], n_seg + 1)
    segs = []
    for k in range(n_seg):
        segs.append((ell & (x >= edges[k]) & (x < edges[k + 1] + (k == n_seg - 1))).astype(np.uint8))
    return segs, ell.astype(np.uint8)


def fourpi_config(name, dim, seed, n_beams, ptv_radii, fluence_dim=20, with_flags=True) -> Config:
    """4pi plan: body phantom, PTV in 4 segments with one isocenter each (segment centroids), n_beams
    Fibonacci-sphere beams assigned round-robin to segments, beamlets kept where they intersect their
    segment (beamletFlagInit, IMRTBeamBundle.cu:130-316)."""
    from . import host

    vol = make_volume(dim)
    density, body = body_density(dim, seed)
    center = [d / 2.0 for d in dim]
    segs, ptv = ptv_segments(dim, center, ptv_radii)
    isos = [host.ptv_isocenter(m, vol) for m in segs]
    bundles, flags, seg_of = [], [], []
    for k, d in enumerate(fibonacci_sphere(n_beams)):
        gantry, couch = direction_to_varian_iec(d)
        deg = (round(math.degrees(gantry), 4), round(math.degrees(couch), 4), 0.0)  # beamlist text precision
        s = k % len(segs)
        b = make_bundle(isos[s], deg, fluence_dim=fluence_dim)
        bundles.append(b)
        seg_of.append(s)
        flags.append(host.beamlet_flag_init(b, segs[s], vol) if with_flags else None)
    cfg = Config(name, vol, density, synthetic_spectrum(seed), synthetic_kernel(seed), bundles, flags, ptv=ptv)
    cfg.notes = {"segments": segs, "segment_of_beam": seg_of, "isocenters": [list(map(float, i)) for i in isos],
                 "body": body}
    return cfg


def config3(n_beams: int = 180) -> Config:
    """Batched 4pi: 200x200x160 head-and-neck-sized body, 4 isocenters, 180 beams, 20x20 fluence."""
    return fourpi_config("batched-4pi-200x200x160", (200, 200, 160), 1003, n_beams, (22.0, 15.0, 13.0))


def config5_candidates(n_beams: int = 1160) -> Config:
    """Serving: 220x220x150 pancreas-sized body with 1160 candidate 4pi beams."""
    return fourpi_config("serving-4pi-220x220x150", (220, 220, 150), 1005, n_beams, (26.0, 18.0, 15.0))
