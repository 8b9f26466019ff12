"""Pins of the CPU oracle against known answers, closed forms and brute force.

Each test names the reference file the known answer or formula comes from.  These run on the
CPU only; they check the oracle itself, not the HIP library.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as orc
from fastdose_amd import synth


def f3(*v):
    return (C.c_float * 3)(*v)


def vec(a):
    return np.array([a[0], a[1], a[2]], dtype=np.float64)


# ---------------------------------------------------------------- rotations (geom.cu:58-111)

def test_rotation_known_answers():
    out = (C.c_float * 3)()
    # geom.cu:64-70: (3,0,0) about z by pi/3 -> (1.5, 2.5981, 0)
    orc.lib().orc_rotate_axis(f3(3, 0, 0), f3(0, 0, 1), np.float32(math.pi / 3), out)
    assert np.allclose(vec(out), [1.5, 3 * math.sqrt(3) / 2, 0.0], atol=1e-5)
    # geom.cu:72-79: about (1,1,1)/sqrt(3) by 2 pi/3 -> (0, 3, 0)
    r = np.float32(math.sqrt(1.0 / 3))
    orc.lib().orc_rotate_axis(f3(3, 0, 0), f3(r, r, r), np.float32(2 * math.pi / 3), out)
    assert np.allclose(vec(out), [0.0, 3.0, 0.0], atol=1e-5)


def test_rotation_inverse_roundtrip_and_matrix():
    rng = np.random.default_rng(3)
    out = (C.c_float * 3)()
    back = (C.c_float * 3)()
    M = (C.c_double * 9)()
    Mi = (C.c_double * 9)()
    for _ in range(50):
        v = rng.uniform(-1, 1, 3).astype(np.float32)
        th, ph, co = (np.float32(x) for x in rng.uniform(-math.pi, math.pi, 3))
        orc.lib().orc_inverse_rotate_beam(f3(*v), th, ph, co, out)
        orc.lib().orc_rotate_beam(out, th, ph, co, back)
        assert np.allclose(vec(back), v, atol=2e-6)  # geom.cu:99-110 identity
        # FP64 matrices compose the same rotations: M(BEV->PVCS) applied to v == FP32 chain
        orc.lib().orc_beam_matrix64(float(th), float(ph), float(co), 1, Mi)
        orc.lib().orc_beam_matrix64(float(th), float(ph), float(co), 0, M)
        Mi_np = np.array(Mi[:]).reshape(3, 3)
        M_np = np.array(M[:]).reshape(3, 3)
        assert np.allclose(Mi_np @ v.astype(np.float64), vec(out), atol=2e-6)
        assert np.allclose(M_np @ Mi_np, np.eye(3), atol=1e-12)


def test_angle_vector_theta_zero():
    # angle2Vector(0, phi) = (0, 0, 1) (geom.cu:113-124); in grid units it stays along z
    bm = orc.Beam()
    bm.beamlet_size[0] = bm.beamlet_size[1] = 0.125
    bm.long_spacing = 0.25
    d = (C.c_float * 3)()
    inv = (C.c_float * 3)()
    orc.lib().orc_cone_direction(0.0, 1.0, C.byref(bm), d, inv)
    assert abs(d[0]) < 1e-7 and abs(d[1]) < 1e-7 and d[2] == 1.0


# ---------------------------------------------------------------- calc_range (beam.cu:134-221)

def test_calc_range_axis_aligned_cube():
    # gantry-0 beam through a bbox: lim_min = SAD - (iso.y - y0), lim_max = SAD + (y1 - iso.y),
    # the closed form singleBeamBEV.cu:173-175 also states
    vol = orc.cast(orc.Volume, synth.make_volume((40, 40, 40), bbox_start=(10, 8, 10), bbox_dim=(20, 24, 20)))
    bm = orc.Beam()
    for i, x in enumerate((5.0, 5.0, 5.0)):
        bm.isocenter[i] = x
    bm.beamlet_size[0] = bm.beamlet_size[1] = 0.125
    bm.fmap_size[0] = bm.fmap_size[1] = 16
    bm.sad = 100.0
    bm.long_spacing = 0.25
    assert orc.calc_range(bm, vol) == 0
    assert abs(bm.lim_min - (100.0 - (5.0 - 2.0))) < 1e-4
    assert abs(bm.lim_max - (100.0 + (8.0 - 5.0))) < 1e-4
    assert bm.long_dim == math.ceil((bm.lim_max - bm.lim_min) / 0.25)
    assert bm.long_dim in (24, 25)
    assert np.allclose(vec(bm.source), [5.0, -95.0, 5.0], atol=1e-5)


def test_calc_range_miss():
    vol = orc.cast(orc.Volume, synth.make_volume((20, 20, 20)))
    bm = orc.Beam()
    for i, x in enumerate((200.0, 2.5, 2.5)):  # isocenter far outside, beam along y misses
        bm.isocenter[i] = x
    bm.beamlet_size[0] = bm.beamlet_size[1] = 0.125
    bm.fmap_size[0] = bm.fmap_size[1] = 16
    bm.sad = 100.0
    bm.long_spacing = 0.25
    assert orc.calc_range(bm, vol) == 1


# ---------------------------------------------------------------- lattice helpers (DoseComp.cu:76-164)

def test_line_seg_closed_forms():
    # geom.cu:239-264: both points on the iso plane -> factor 1; both on the axis -> |dz| ls
    bm = orc.Beam()
    bm.beamlet_size[0] = 0.125
    bm.beamlet_size[1] = 0.13
    bm.fmap_size[0] = bm.fmap_size[1] = 16
    bm.sad = 100.0
    bm.lim_min = 90.0
    bm.long_spacing = 0.25
    zn = (100.0 - 90.0) / 0.25
    s = (C.c_float * 3)(3.0, 5.0, zn)
    e = (C.c_float * 3)(11.0, 2.0, zn)
    d = orc.lib().orc_line_seg(s, e, C.byref(bm))
    assert abs(d - math.hypot(8 * 0.125, 3 * float(np.float32(0.13)))) < 1e-12
    s = (C.c_float * 3)(8.0, 8.0, 3.0)
    e = (C.c_float * 3)(8.0, 8.0, 11.5)
    assert abs(orc.lib().orc_line_seg(s, e, C.byref(bm)) - 8.5 * 0.25) < 1e-12


def test_next_point_properties():
    rng = np.random.default_rng(11)
    out = (C.c_float * 3)()
    for _ in range(2000):
        d = rng.normal(size=3)
        d[2] = abs(d[2]) + 0.05
        d = (d / np.linalg.norm(d)).astype(np.float32)
        inv = (np.float32(1) / d).astype(np.float32)
        cur = rng.uniform(-30, 30, 3).astype(np.float32)
        cur[2] = np.float32(np.round(cur[2])) if rng.random() < 0.3 else cur[2]
        flag = orc.lib().orc_next_point(f3(*cur), f3(*d), f3(*inv), out)
        nxt = vec(out)
        step = nxt - cur.astype(np.float64)
        k = int(np.argmax(np.abs(d)))
        t = step[k] / d[k]
        assert t > 0
        # moved along the direction (FP32 positions: absolute error ~ ulp(30))
        assert np.allclose(step, t * d.astype(np.float64), rtol=0, atol=2e-5)
        if flag:
            assert nxt[2] == round(nxt[2])  # slice ends are snapped
        else:
            # landed on an x or y voxel boundary
            assert min(abs(nxt[0] - round(nxt[0])), abs(nxt[1] - round(nxt[1]))) < 1e-3


def test_wrap_index_is_periodic_floor():
    # d_calcCoords (DoseComp.cu:147-164): fmod + shift to [0, 16), floored; away from FP32
    # rounding edges this is floor(m) mod 16
    rng = np.random.default_rng(5)
    for m in np.concatenate([rng.uniform(-3000, 3000, 4000), np.arange(-64, 64, 0.5)]).astype(np.float32):
        r = math.fmod(float(m), 16.0)
        r = r + 16.0 if r < 0 else r
        assert orc.lib().orc_wrap_index(m, 16.0) == int(math.floor(r)) % 16
    # the rounding edge: just below a multiple of 16 stays at 15 (fmodf + 16 would give 16)
    assert orc.lib().orc_wrap_index(np.float32(-1e-7), 16.0) == 15
    assert orc.lib().orc_wrap_index(np.float32(-32.0), 16.0) == 0


# ---------------------------------------------------------------- table readers (spectrum.cu, kernel.cu)

def test_spectrum_fixture(golden):
    s = orc.Spectrum()
    assert orc.lib().orc_read_spectrum(str(golden / "spec_6mv.spec").encode(), C.byref(s)) == 0
    rows = [l.split() for l in (golden / "spec_6mv.spec").read_text().splitlines() if l.strip()]
    assert s.n == len(rows) == 14
    w = np.array([float(r[1]) for r in rows])
    assert np.allclose(np.array(s.fluence[: s.n]), w / w.sum(), rtol=1e-6)
    assert abs(sum(s.fluence[: s.n]) - 1.0) < 1e-6
    assert np.allclose(np.array(s.energy[: s.n]), [float(r[0]) for r in rows])
    assert np.allclose(np.array(s.mu[: s.n]), [float(r[2]) for r in rows])
    assert np.allclose(np.array(s.mu_en[: s.n]), [float(r[3]) for r in rows])


def test_kernel_fixture(golden):
    k = orc.Kernel()
    assert orc.lib().orc_read_kernel(str(golden / "kernel_exp_6mv.txt").encode(), 8, C.byref(k)) == 0
    rows = [l.split() for l in (golden / "kernel_exp_6mv.txt").read_text().splitlines()[1:] if l.strip()]
    assert k.n_theta == 6 and k.n_phi == 8
    for i, r in enumerate(rows):
        assert k.theta_mid[i] == np.float32((np.float32(r[0]) + np.float32(r[1])) / np.float32(2))
        assert k.A[i] == np.float32(r[2]) and k.b[i] == np.float32(r[5])
    # kernel.cu:49 phi_i = (i + 0.5) * 2 pi / nPhi; 4 forward cones (theta_mid <= pi/2)
    assert np.allclose(np.array(k.phi[:8]), (np.arange(8) + 0.5) * 2 * math.pi / 8, atol=1e-6)
    assert sum(1 for i in range(6) if k.theta_mid[i] <= math.pi / 2) == 4


# ---------------------------------------------------------------- Terma closed form (TermaComp.cu:179-253)

def _padded_water_beam(density_value=1.0):
    vol = orc.cast(orc.Volume, synth.make_volume((40, 40, 40), bbox_start=(10, 10, 10), bbox_dim=(20, 20, 20)))
    dens = np.full((40, 40, 40), density_value, dtype=np.float32)
    bm = orc.Beam()
    for i in range(3):
        bm.isocenter[i] = 5.0
    bm.beamlet_size[0] = bm.beamlet_size[1] = 0.125
    bm.fmap_size[0] = bm.fmap_size[1] = 16
    bm.sad = 100.0
    bm.long_spacing = 0.25
    assert orc.calc_range(bm, vol) == 0
    return vol, dens, bm


@pytest.mark.parametrize("rho", [1.0, 0.0])
def test_terma_closed_form(rho):
    vol, dens, bm = _padded_water_beam(rho)
    spec = orc.cast(orc.Spectrum, synth.synthetic_spectrum(1001))
    fl = np.random.default_rng(0).uniform(0.5, 1.5, 256).astype(np.float32)
    T, R = orc.terma(bm, fl, vol, dens, spec)
    L = bm.long_dim
    n = spec.n
    E = np.array(spec.energy[:n], np.float64)
    mu = np.array(spec.mu[:n], np.float64)
    w = np.array(spec.fluence[:n], np.float64) * E * mu
    sad, ls = float(bm.sad), float(bm.long_spacing)
    dt = ls / (sad * 4)
    t0 = float(bm.lim_min) / sad
    bs = float(bm.beamlet_size[0])
    for p in (0, 17, 119, 136, 255):
        ix, iy = p % 16, p // 16
        direction = np.array([(ix + 0.5 - 8) * bs, sad, (iy + 0.5 - 8) * bs])
        step = np.linalg.norm(direction) * dt
        k = np.arange(1, 4 * L + 1, dtype=np.float64)
        t = t0 + k * dt
        rpl = rho * k * step
        terms = (w[None, :] * np.exp(-mu[None, :] * rpl[:, None])).sum(1) / t ** 2
        want = fl[p] * terms.reshape(L, 4).mean(1)
        assert np.allclose(T[:, p], want, rtol=1e-12, atol=0)
        assert np.allclose(R[:, p], rho, rtol=0, atol=1e-12)


# ---------------------------------------------------------------- CCCS (DoseComp.cu:217-392, derivation.md)

def _near_axis_kernel(n_phi=8, theta=1e-3):
    k = orc.Kernel()
    k.n_theta = 2
    k.n_phi = n_phi
    k.theta_begin[0], k.theta_end[0] = 0.0, np.float32(2 * theta)
    k.theta_mid[0] = np.float32(theta)
    k.theta_begin[1], k.theta_end[1] = 2.0, 2.2   # backward cone: skipped (DoseComp.cu:265)
    k.theta_mid[1] = 2.1
    k.A[0], k.a[0], k.B[0], k.b[0] = 1.7, 3.1, 0.02, 0.07
    k.A[1], k.a[1], k.B[1], k.b[1] = 5.0, 5.0, 5.0, 5.0
    for i in range(n_phi):
        k.phi[i] = np.float32((i + 0.5) * 2 * math.pi / n_phi)
    return k


@pytest.mark.parametrize("exact_g", [False, True])
def test_cccs_one_dimensional_closed_form(exact_g):
    """Near-axis cone: each ray stays in its own pixel, one segment per slice, so the CCCS of
    every pixel is the 1-D problem of derivation.md; X is checked against its closed sum
    X_i = sum_{j<i} T_j e^{-a sum_{j<k<i} p_k} (1 - e^{-a p_j}) rather than the recursion."""
    rng = np.random.default_rng(1)
    L = 40
    bm = orc.Beam()
    bm.beamlet_size[0], bm.beamlet_size[1] = 0.125, 0.12
    bm.fmap_size[0] = bm.fmap_size[1] = 16
    bm.sad, bm.lim_min, bm.long_spacing, bm.long_dim = 100.0, 93.0, 0.25, L
    T = rng.uniform(0.0, 1.0, (L, 256))
    R = rng.uniform(0.2, 1.8, (L, 256))
    kern = _near_axis_kernel()
    D, nseg = orc.cccs(bm, T, R, kern, orc.EXACT_G if exact_g else 0)
    assert nseg == 8 * L * 256  # one segment per slice per azimuth
    A, a, B, b = (float(kern.A[0]), float(kern.a[0]), float(kern.B[0]), float(kern.b[0]))
    bs = np.array([float(bm.beamlet_size[0]), float(bm.beamlet_size[1])])
    sad, lm, ls = float(bm.sad), float(bm.lim_min), float(bm.long_spacing)
    want = np.zeros((L, 256))
    for p in range(256):
        ix, iy = p % 16, p // 16
        total = np.zeros(L)
        last_p = None
        for f in range(8):
            th, ph = float(np.float32(1e-3)), float(kern.phi[f])
            u = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
            g = u / np.array([bs[0], bs[1], ls])
            g /= np.linalg.norm(g)
            z = np.arange(L + 1, dtype=np.float64)
            x = 0.5 + ix + g[0] / g[2] * z
            y = 0.5 + iy + g[1] / g[2] * z
            fz = (z * ls + lm) / sad
            phys = np.stack([(x - 8) * bs[0] * fz, (y - 8) * bs[1] * fz, z * ls], 1)
            seg = np.linalg.norm(np.diff(phys, axis=0), axis=1)
            pp = seg * R[:, p]
            contrib = np.zeros(L)
            for (K, k_) in ((A, a), (B, b)):
                cum = np.concatenate([[0.0], np.cumsum(pp)])
                X = np.zeros(L)
                for i in range(L):
                    j = np.arange(i)
                    X[i] = np.sum(T[j, p] * np.exp(-k_ * (cum[i] - cum[j + 1])) * (1 - np.exp(-k_ * pp[j])))
                if exact_g:
                    gi = (1 - np.exp(-k_ * pp)) / (k_ * pp)
                    contrib += K / k_ * ((1 - gi) * T[:, p] + gi * X) * pp
                else:
                    contrib += K / k_ * X * pp
            total += contrib
            last_p = pp
        want[:, p] = total / (last_p + float(np.float32(1e-4)))
    assert np.allclose(D, want, rtol=2e-6, atol=1e-12)


def test_cccs_linear_in_terma_and_zero_density():
    cfg = synth.small_config(seed=3, dim=(16, 18, 14), fluence_dim=1)
    beams = [orc.cast(orc.Beam, x) for x in _beamlets(cfg)]
    bm = beams[0]
    kern = orc.cast(orc.Kernel, cfg.kernel)
    rng = np.random.default_rng(2)
    T1 = rng.uniform(0, 1, (bm.long_dim, 256))
    T2 = rng.uniform(0, 1, (bm.long_dim, 256))
    R = rng.uniform(0.2, 1.8, (bm.long_dim, 256))
    D1, _ = orc.cccs(bm, T1, R, kern)
    D2, _ = orc.cccs(bm, T2, R, kern)
    D12, _ = orc.cccs(bm, 2.0 * T1 + T2, R, kern)
    assert np.allclose(D12, 2.0 * D1 + D2, rtol=1e-10, atol=1e-14)
    assert (D1 >= 0).all() and D1.max() > 0
    D0, _ = orc.cccs(bm, T1, np.zeros_like(R), kern)
    assert not D0.any()
    assert orc.cccs_count(bm, kern) == orc.cccs(bm, T1, R, kern)[1]


def _beamlets(cfg):
    import fastdose_amd as fd

    out = []
    for b, f in zip(cfg.bundles, cfg.flags):
        beams, _ = fd.beamlets_init(b, f, cfg.volume)
        out.extend(beams)
    return out


# ---------------------------------------------------------------- resampler (IMRTgeom.cu:370-516)

def _rot_fp64(v, th, ph, co):
    """rotateBeamAtOriginRHS (geom.cu:36-44) in FP64 by successive Rodrigues turns."""
    def rod(p, r, t):
        r = np.asarray(r, dtype=np.float64)
        return (1 - math.cos(t)) * np.dot(p, r) * r + math.cos(t) * p + math.sin(t) * np.cross(r, p)
    axis = np.array([math.sin(-ph), 0.0, math.cos(-ph)])
    return rod(rod(v, axis, -th), [0.0, 1.0, 0.0], ph + co)


def _rot_fp64_many(v, th, ph, co):
    def rod(p, r, t):
        r = np.asarray(r, dtype=np.float64)
        return ((1 - math.cos(t)) * (p @ r))[:, None] * r + math.cos(t) * p + math.sin(t) * np.cross(r, p)
    axis = np.array([math.sin(-ph), 0.0, math.cos(-ph)])
    return rod(rod(v, axis, -th), [0.0, 1.0, 0.0], ph + co)


def test_resample_brute_force_fp64():
    cfg = synth.small_config(seed=5, dim=(20, 22, 18), fluence_dim=1, angles=(40.0, 25.0, 15.0))
    bm = orc.cast(orc.Beam, _beamlets(cfg)[0])
    vol = orc.cast(orc.Volume, cfg.volume)
    rng = np.random.default_rng(9)
    L = bm.long_dim
    dose = rng.uniform(0.0, 1.0, (L, 256))
    cols, vals, _ = orc.resample(bm, dose, vol, extent=2.0)
    X, Y, Z = (int(vol.dim[i]) for i in range(3))
    vs = float(vol.voxel_size[0])
    th, ph, co = (float(bm.angles[i]) for i in range(3))
    src = vec(bm.source)
    o = (np.arange(5) + 0.5) / 5
    gx, gy, gz = np.meshgrid(np.arange(X), np.arange(Y), np.arange(Z), indexing="ij")
    vox = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1)                        # (V, 3)
    off = np.stack(np.meshgrid(o, o, o, indexing="ij"), -1).reshape(-1, 3)          # (125, 3)
    q = (vox[:, None, :] + off[None, :, :]).reshape(-1, 3) * vs - src
    b = _rot_fp64_many(q, th, ph, co)
    d = (b[:, 1] - float(bm.lim_min)) / float(bm.long_spacing)
    u = b[:, 0] / (float(bm.beamlet_size[0]) * b[:, 1] / float(bm.sad)) + 8
    w = b[:, 2] / (float(bm.beamlet_size[1]) * b[:, 1] / float(bm.sad)) + 8
    ok = (d >= 0) & (d < L) & (u >= 0) & (u < 16) & (w >= 0) & (w < 16)
    sample = np.zeros(len(q))
    sample[ok] = dose[d[ok].astype(int), w[ok].astype(int) * 16 + u[ok].astype(int)]
    per_vox = sample.reshape(-1, 125).sum(1) / 125
    brute = np.zeros(X * Y * Z)
    brute[vox[:, 0] + X * (vox[:, 1] + Y * vox[:, 2])] = per_vox
    got = np.zeros(X * Y * Z)
    got[cols] = vals
    assert np.all(np.diff(cols) > 0)  # ascending columns
    diff = np.abs(got - brute)
    # FP32 vs FP64 sample decisions may differ only for samples within FP32 rounding (~5e-5
    # pixel) of a BEV voxel boundary: a few samples per thousand voxels, one sample each
    assert (diff > 1e-9).sum() <= 0.02 * max(1, (brute > 0).sum())
    assert diff.max() <= 3.0 / 125
    assert (brute > 0).sum() > 100


def test_resample_constant_field_and_extent():
    cfg = synth.small_config(seed=6, dim=(24, 24, 24), fluence_dim=1, angles=(0.0, 0.0, 0.0))
    bm = orc.cast(orc.Beam, _beamlets(cfg)[0])
    vol = orc.cast(orc.Volume, cfg.volume)
    c = 0.37
    dose = np.full((bm.long_dim, 256), c)
    cols, vals, _ = orc.resample(bm, dose, vol, extent=2.0)
    k = np.rint(vals / c * 125)
    assert np.allclose(vals, k * c / 125, rtol=1e-12)    # multiples of c/125
    assert np.isclose(vals.max(), c, rtol=1e-12)          # fully covered voxels
    assert (k >= 1).all() and (k <= 125).all()
    # a small extent removes every voxel whose centre is farther than extent + half diagonal
    cols2, vals2, _ = orc.resample(bm, dose, vol, extent=0.2)
    assert set(cols2.tolist()) <= set(cols.tolist()) and len(cols2) < len(cols)
    X, Y = int(vol.dim[0]), int(vol.dim[1])
    vs = float(vol.voxel_size[0])
    for col in cols2:
        x, y, z = col % X, (col // X) % Y, col // (X * Y)
        ctr = (np.array([x, y, z]) + 0.5) * vs - vec(bm.source)
        b = _rot_fp64(ctr, float(bm.angles[0]), float(bm.angles[1]), float(bm.angles[2]))
        assert math.hypot(b[0], b[2]) <= 0.2 + 0.5 * math.sqrt(3) * vs + 1e-4


# ---------------------------------------------------------------- CSR (IMRTDebug.cu:403-474, 773-870)

def test_dense_to_csr_known_matrices():
    dense = np.array([[0, 0, 1, 0, 2, 0], [2, 0, 0, 1, 0, 0], [0, 0, 1, 2, 0, 1], [1, 0, 1, 0, 0, 0],
                      [1, 1, 0, 0, 0, 1]], dtype=np.float32)
    off, col, val = orc.dense_to_csr(dense)
    assert off.tolist() == [0, 2, 4, 7, 9, 12]
    assert col.tolist() == [2, 4, 0, 3, 2, 3, 5, 0, 2, 0, 1, 5]
    assert val.tolist() == [1, 2, 2, 1, 1, 2, 1, 1, 1, 1, 1, 1]
    dense4 = np.array([[0, 0, 1, 2], [1, 0, 0, 1], [0, 0, 1, 0], [2, 0, 0, 0]], dtype=np.float32)
    off, col, val = orc.dense_to_csr(dense4)
    assert off.tolist() == [0, 2, 4, 5, 6]
    assert col.tolist() == [2, 3, 0, 3, 2, 0]
    assert val.tolist() == [1.0, 2.0, 1.0, 1.0, 1.0, 2.0]


# ---------------------------------------------------------------- trilinear density (TermaComp.cu:222-233)

def _bev_to_pvcs64(th, ph, co):
    """inverseRotateBeamAtOriginRHS (geom.cu:47-56) as an FP64 matrix: the transpose of the forward
    rotation, whose columns are the images of the unit vectors under _rot_fp64."""
    F = np.stack([_rot_fp64(e, th, ph, co) for e in np.eye(3)], 1)
    return F.T


def _ramp_1d(c, n, alpha, beta):
    """1-D linear texture (texel i centred at i + 0.5, border 0) of the field g(i) = alpha + beta (i + 0.5),
    i in [0, n), sampled at c: the line itself between the first and last texel centres, falling linearly
    to 0 over the half texel beyond them."""
    c = np.asarray(c, dtype=np.float64)
    g0, gn = alpha + beta * 0.5, alpha + beta * (n - 0.5)
    out = np.where((c > 0.5) & (c < n - 0.5), alpha + beta * c, 0.0)
    out = np.where((c > -0.5) & (c <= 0.5), (c + 0.5) * g0, out)
    out = np.where((c >= n - 0.5) & (c < n + 0.5), (n + 0.5 - c) * gn, out)
    return out


def test_terma_density_ramp_closed_form():
    """A separable density fx(x) fy(y) fz(z) of three linear ramps: trilinear filtering reproduces it
    exactly between texel centres and falls to 0 over the half texel at each border (CUDA border
    addressing), so DensityBEV is the mean of four closed-form samples and the radiological path their
    running sum.  Anisotropic voxels and distinct ramps catch an axis mix-up; the beam enters and leaves
    through the border half-texels and one side of its field runs along the x = 0 face."""
    vol = orc.Volume()
    dims = (24, 20, 22)
    vs = tuple(float(np.float32(v)) for v in (0.25, 0.3, 0.2))   # as the FP32 volume holds them
    for i in range(3):
        vol.voxel_size[i] = vs[i]
        vol.dim[i] = dims[i]
        vol.bbox_start[i] = 0
        vol.bbox_dim[i] = dims[i]
    ramps = ((0.6, 0.05), (1.3, -0.04), (0.4, 0.07))       # (alpha, beta) per axis x, y, z
    g = [ramps[d][0] + ramps[d][1] * (np.arange(dims[d]) + 0.5) for d in range(3)]
    dens = (g[2][:, None, None] * g[1][None, :, None] * g[0][None, None, :]).astype(np.float32)
    # the field is stored in FP32: use the stored values' own ramps (exact in FP64 up to their rounding)
    assert np.allclose(dens, g[2][:, None, None] * g[1][None, :, None] * g[0][None, None, :], rtol=1e-6)
    bm = orc.Beam()
    iso = (0.9, 3.0, 2.2)
    for i in range(3):
        bm.isocenter[i] = iso[i]
    bm.beamlet_size[0] = bm.beamlet_size[1] = 0.125
    bm.fmap_size[0] = bm.fmap_size[1] = 16
    bm.sad = 100.0
    bm.long_spacing = 0.25
    bm.angles[0], bm.angles[1], bm.angles[2] = 0.35, 0.6, 0.2
    assert orc.calc_range(bm, vol) == 0
    spec = orc.cast(orc.Spectrum, synth.synthetic_spectrum(1009))
    fl = np.random.default_rng(4).uniform(0.5, 1.5, 256).astype(np.float32)
    T, R = orc.terma(bm, fl, vol, dens, spec)
    L = bm.long_dim
    Mi = _bev_to_pvcs64(float(bm.angles[0]), float(bm.angles[1]), float(bm.angles[2]))
    sad, ls, lm = float(bm.sad), float(bm.long_spacing), float(bm.lim_min)
    src = vec(bm.source)
    dt = ls / (sad * 4)
    t = lm / sad + np.arange(1, 4 * L + 1) * dt
    n = spec.n
    mu = np.array(spec.mu[:n], np.float64)
    w = np.array(spec.fluence[:n], np.float64) * np.array(spec.energy[:n], np.float64) * mu
    border = 0
    for p in range(256):
        ix, iy = p % 16, p // 16
        pv = Mi @ np.array([(ix + 0.5 - 8) * 0.125, sad, (iy + 0.5 - 8) * 0.125])
        c = (src[None, :] + pv[None, :] * t[:, None]) / np.array(vs)[None, :]
        f = [_ramp_1d(c[:, d], dims[d], *ramps[d]) for d in range(3)]
        rho = f[0] * f[1] * f[2]
        inner = np.all((c > 0.5) & (c < np.array(dims) - 0.5), axis=1)
        border += int(((rho > 0) & ~inner).sum())
        assert np.allclose(R[:, p], rho.reshape(L, 4).mean(1), rtol=2e-6, atol=1e-9)
        rpl = np.cumsum(rho) * np.linalg.norm(pv) * dt
        terms = (w[None, :] * np.exp(-mu[None, :] * rpl[:, None])).sum(1) / t ** 2
        assert np.allclose(T[:, p], fl[p] * terms.reshape(L, 4).mean(1), rtol=2e-6, atol=0)
    assert border > 500  # the border half-texels are exercised


# ---------------------------------------------------------------- CCCS per-ray transcription (DoseComp.cu:217-392)

def _cccs_per_ray(bm, T, R, kern):
    """DoseComp.cu:265-390 restated ray by ray in FP64, as each reference thread runs it: its own
    coordinates (shared lattice point + (idx_x, idx_y)), the d_calcCoords fmod wrap of its midpoint,
    d_calcLineSeg of its own endpoints, and its own floor(c / 16) edge test.  The shared lattice point
    (d_nextPoint) and cone direction come from the oracle's primitives, which the next-point and
    angle tests above pin; g == 1 (DoseComp.cu:167-175).  Returns the dose and the number of steps
    whose midpoint lies within 1e-6 of a voxel boundary (where an FP32 midpoint may round across it)."""
    fx, fy = int(bm.fmap_size[0]), int(bm.fmap_size[1])
    npix = fx * fy
    L = int(bm.long_dim)
    ix = np.arange(npix) % fx
    iy = np.arange(npix) // fx
    bs = (float(bm.beamlet_size[0]), float(bm.beamlet_size[1]))
    ls, lm, sad = float(bm.long_spacing), float(bm.lim_min), float(bm.sad)
    eps_f32 = float(np.float32(1e-4))

    def phys(x, y, z):
        f = (z * ls + lm) / sad
        return np.stack([(x - fx * 0.5) * bs[0] * f, (y - fy * 0.5) * bs[1] * f, np.broadcast_to(z * ls, x.shape)], 1)

    dose = np.zeros((L, npix))
    near = 0
    d = (C.c_float * 3)()
    inv = (C.c_float * 3)()
    out = (C.c_float * 3)()
    for t in range(kern.n_theta):
        th = kern.theta_mid[t]
        if th > math.pi / 2:
            continue
        A, a, B, b = (float(kern.A[t]), float(kern.a[t]), float(kern.B[t]), float(kern.b[t]))
        nphi = kern.n_phi
        XA = np.zeros((nphi, npix))
        XB = np.zeros((nphi, npix))
        dirs, base = [], []
        for f in range(nphi):
            orc.lib().orc_cone_direction(th, kern.phi[f], C.byref(bm), d, inv)
            dirs.append((f3(*d), f3(*inv)))
            base.append(np.array([0.5, 0.5, 0.0], dtype=np.float32))
        while True:
            s = int(math.floor(float(base[0][2]) + 0.5))
            if s >= L:
                break
            Ts, Rs = T[s], R[s]
            SD = np.zeros(npix)
            cumu = None
            for f in range(nphi):
                cumu = np.zeros(npix)
                while True:
                    flag = orc.lib().orc_next_point(f3(*base[f]), dirs[f][0], dirs[f][1], out)
                    nb = np.array([out[0], out[1], out[2]], dtype=np.float32)
                    b0 = [float(base[f][0]) + ix, float(base[f][1]) + iy]
                    n0 = [float(nb[0]) + ix, float(nb[1]) + iy]
                    mx, my = (b0[0] + n0[0]) * 0.5, (b0[1] + n0[1]) * 0.5
                    for m_ in (mx, my):
                        fr = m_ - np.floor(m_)
                        near += int(np.count_nonzero((fr < 1e-6) | (fr > 1 - 1e-6)))
                    wx = np.fmod(mx, fx)
                    wx = np.where(wx < 0, wx + fx, wx)
                    wy = np.fmod(my, fy)
                    wy = np.where(wy < 0, wy + fy, wy)
                    m = np.floor(wx).astype(int) + fx * np.floor(wy).astype(int)
                    seg = np.linalg.norm(phys(n0[0], n0[1], float(nb[2])) - phys(b0[0], b0[1], float(base[f][2])),
                                         axis=1)
                    p_i = seg * Rs[m]
                    np.add.at(SD, m, (A / a * XA[f] + B / b * XB[f]) * p_i)
                    np.add.at(cumu, m, p_i)
                    ea, eb = np.exp(-a * p_i), np.exp(-b * p_i)
                    xa = Ts[m] * (1 - ea) + ea * XA[f]
                    xb = Ts[m] * (1 - eb) + eb * XB[f]
                    crossed = (np.floor(b0[0] / fx) != np.floor(n0[0] / fx)) | \
                              (np.floor(b0[1] / fy) != np.floor(n0[1] / fy))
                    xa[crossed] = 0.0
                    xb[crossed] = 0.0
                    XA[f], XB[f] = xa, xb
                    base[f] = nb
                    if flag:
                        break
            dose[s] += SD / (cumu + eps_f32)
    return dose, near


@pytest.mark.parametrize("L,n_phi,seed", [(6, 8, 1), (9, 5, 2), (7, 12, 3)])
def test_cccs_per_ray_transcription(L, n_phi, seed):
    """Every forward cone of the configs' kernel, including the near-horizontal one (tan theta = 5:
    ~14 segments per slice, many rays crossing the BEV edge each slice), on a beamlet with unequal
    lateral spacings: the oracle's shared-lattice CCCS (one wrapped shift per step, edge resets from
    integer floors) equals the per-ray transcription of the reference kernel."""
    rng = np.random.default_rng(seed)
    bm = orc.Beam()
    bm.beamlet_size[0], bm.beamlet_size[1] = 0.125, 0.11
    bm.fmap_size[0] = bm.fmap_size[1] = 16
    bm.sad, bm.lim_min, bm.long_spacing, bm.long_dim = 100.0, 93.0, 0.25, L
    kern = orc.cast(orc.Kernel, synth.synthetic_kernel(1000 + seed, n_phi=n_phi))
    T = rng.uniform(0.0, 1.0, (L, 256))
    R = rng.uniform(0.2, 1.8, (L, 256))
    D, nseg = orc.cccs(bm, T, R, kern)
    want, near = _cccs_per_ray(bm, T, R, kern)
    assert near == 0  # no midpoint within FP32 reach of a voxel boundary: the index maps must agree
    assert nseg > 40 * L * 256 * n_phi / 8
    assert np.allclose(D, want, rtol=1e-11, atol=1e-14)
    # the edge resets matter here: without them the dose differs
    assert D.max() > 0


# ---------------------------------------------------------------- PTV flags (IMRTBeamBundle.cu:269-316)

def _flags_brute(bundle, mask, vol, delta=0.0):
    """d_beamletFlagInit in FP64: 3 x 3 lateral samples per depth step i < long_dim of a beam spanning
    the bundle's fluence map, at source + R^-1 (x, sad, z) SXD / sad; a beamlet is on when the linear
    texture of the mask is positive at a sample, i.e. some texel within one texel pitch (per axis) of
    the sample is set.  delta widens (> 0) or narrows (< 0) that pitch, to find beamlets whose flag
    hangs on rounding."""
    bm = orc.Beam()
    for i in range(3):
        bm.isocenter[i] = bundle.isocenter[i]
        bm.angles[i] = bundle.angles[i]
    bm.beamlet_size[0], bm.beamlet_size[1] = bundle.beamlet_size[0], bundle.beamlet_size[1]
    bm.fmap_size[0], bm.fmap_size[1] = bundle.fluence_dim[0], bundle.fluence_dim[1]
    bm.sad = bundle.sad
    bm.long_spacing = bundle.long_spacing
    assert orc.calc_range(bm, vol) == 0
    fx, fy = int(bundle.fluence_dim[0]), int(bundle.fluence_dim[1])
    Mi = _bev_to_pvcs64(float(bm.angles[0]), float(bm.angles[1]), float(bm.angles[2]))
    src = vec(bm.source)
    vs = np.array([float(vol.voxel_size[i]) for i in range(3)])
    dims = np.array([int(vol.dim[i]) for i in range(3)])
    sad = float(bm.sad)
    sxd = float(bm.lim_min) + np.arange(int(bm.long_dim)) * float(bm.long_spacing)
    off = (np.arange(3) + 0.5) / 3
    bx = np.arange(fx * fy) % fx
    by = np.arange(fx * fy) // fx
    X = (bx[:, None] + off[None, :] - fx * 0.5) * float(bm.beamlet_size[0])       # (nb, 3)
    Z = (by[:, None] + off[None, :] - fy * 0.5) * float(bm.beamlet_size[1])
    disp = np.zeros((fx * fy, 3, 3, 3))
    disp[:, :, :, 0] = X[:, :, None]
    disp[:, :, :, 1] = sad
    disp[:, :, :, 2] = Z[:, None, :]
    pv = disp @ Mi.T                                                                # (nb, j, k, 3)
    c = (src + pv[:, :, :, None, :] * (sxd / sad)[None, None, None, :, None]) / vs  # (nb, j, k, i, 3)
    c = c.reshape(fx * fy, -1, 3)
    pad = np.zeros(tuple(dims[::-1] + 6), dtype=bool)
    pad[3:-3, 3:-3, 3:-3] = mask.reshape(dims[::-1]) > 0
    t0 = np.floor(c - 0.5).astype(int)
    on = np.zeros(c.shape[:2], dtype=bool)
    for ox in (-1, 0, 1, 2):
        for oy in (-1, 0, 1, 2):
            for oz in (-1, 0, 1, 2):
                t = t0 + np.array([ox, oy, oz])
                ok = np.all(np.abs(c - (t + 0.5)) < 1.0 + delta, axis=2)
                t = np.clip(t, -3, dims + 2) + 3
                on |= ok & pad[t[:, :, 2], t[:, :, 1], t[:, :, 0]]
    return on.any(1).astype(np.uint8), bm


def _flag_bundle(iso, angles, fd=10, bs=0.5):
    b = orc.Bundle()
    for i in range(3):
        b.isocenter[i] = iso[i]
        b.angles[i] = angles[i]
    b.beamlet_size[0] = b.beamlet_size[1] = bs
    b.fluence_dim[0] = b.fluence_dim[1] = fd
    b.sad = 100.0
    b.long_spacing = 0.25
    b.sub_dim[0] = b.sub_dim[1] = 16
    b.sub_on[0] = b.sub_on[1] = 4
    return b


def test_beamlet_flags_box_closed_form():
    """Gantry 0: sample (bx, j, by, k) at depth SXD sits at iso + (x_j, SXD - SAD, z_k) with
    x_j = (bx + (j + .5)/3 - 5) bs SXD/SAD.  A box PTV of voxels [x0, x1) x [y0, y1) x [z0, z1) has a
    positive linear texture on the open box (x0 - .5, x1 + .5) ... in voxel units, so a beamlet is on
    exactly when, over the depths whose y lies in the box, one of its x samples and one of its z
    samples sweep into the box -- a rectangle of beamlets."""
    vol = orc.cast(orc.Volume, synth.make_volume((40, 40, 40)))
    mask = np.zeros((40, 40, 40), dtype=np.uint8)
    box = ((14, 26), (15, 25), (16, 24))   # x, y, z voxel ranges
    mask[box[2][0]:box[2][1], box[1][0]:box[1][1], box[0][0]:box[0][1]] = 1
    bundle = _flag_bundle((5.0, 5.0, 5.0), (0.0, 0.0, 0.0))
    brute, bm = _flags_brute(bundle, mask, vol)
    vs = 0.25
    sad = 100.0
    src_y = float(bm.source[1])
    sxd = float(bm.lim_min) + np.arange(int(bm.long_dim)) * 0.25
    yv = (src_y + sxd) / vs
    ins = (yv > box[1][0] - 0.5) & (yv < box[1][1] + 0.5)
    f_lo, f_hi = sxd[ins].min() / sad, sxd[ins].max() / sad
    want = np.zeros(100, dtype=np.uint8)
    for b in range(100):
        bx, by = b % 10, b // 10
        okx = okz = False
        for j in range(3):
            o = (bx + (j + 0.5) / 3 - 5) * 0.5
            lo, hi = sorted((o * f_lo, o * f_hi))
            okx |= lo + 5.0 < (box[0][1] + 0.5) * vs and hi + 5.0 > (box[0][0] - 0.5) * vs
            o = (by + (j + 0.5) / 3 - 5) * 0.5
            lo, hi = sorted((o * f_lo, o * f_hi))
            okz |= lo + 5.0 < (box[2][1] + 0.5) * vs and hi + 5.0 > (box[2][0] - 0.5) * vs
        want[b] = okx and okz
    assert want.reshape(10, 10)[2:8, 1:9].all() and want.sum() == 48   # bx in [1, 8], by in [2, 7]
    assert (brute == want).all()
    assert (_flags_brute(bundle, mask, vol, delta=0.02)[0] == want).all()   # no flag near a rounding edge
    assert (_flags_brute(bundle, mask, vol, delta=-0.02)[0] == want).all()
    assert (orc.beamlet_flags(bundle, mask, vol) == want).all()


def test_beamlet_flags_depth_range_end():
    """Depth samples are SXD = lim_min + i * long_spacing for i < long_dim (IMRTBeamBundle.cu:289-291):
    with the bounding box shorter than the volume, a PTV slab beyond the box is flagged exactly when
    its texture support starts before the last sample.  The slab is placed one texel either side of
    that sample (with margins far above FP32 rounding), so a depth grid shifted by one step flips it."""
    vol = orc.cast(orc.Volume, synth.make_volume((40, 48, 40), bbox_dim=(40, 32, 40)))
    bundle = _flag_bundle((5.0, 4.0, 5.0), (0.0, 0.0, 0.0))
    _, bm = _flags_brute(bundle, np.zeros((40, 48, 40), dtype=np.uint8), vol)
    L = int(bm.long_dim)
    c_last = (float(bm.source[1]) + float(bm.lim_min) + (L - 1) * 0.25) / 0.25
    s_far = math.floor(c_last + 0.5) + 1            # support (s_far - 0.5, ...) starts past the last sample
    assert 0.01 < s_far - 0.5 - c_last < 0.99
    for start, want in ((s_far, 0), (s_far - 1, 1)):
        mask = np.zeros((40, 48, 40), dtype=np.uint8)
        mask[:, start:start + 4, :] = 1
        brute, _ = _flags_brute(bundle, mask, vol)
        assert (brute == want).all()
        assert (orc.beamlet_flags(bundle, mask, vol) == want).all()


@pytest.mark.parametrize("angles", [(0.7, -0.4, 0.3), (2.2, 1.1, -0.6), (1.5708, 0.0, 0.0)])
def test_beamlet_flags_brute_force(angles):
    """Rotated bundles through an ellipsoidal PTV: the oracle's FP32 flags equal the FP64 brute force on
    every beamlet whose flag does not hang on a sample within 1e-3 texel of a texel-weight edge."""
    vol = orc.cast(orc.Volume, synth.make_volume((36, 40, 32)))
    z, y, x = np.meshgrid(*(np.arange(n) + 0.5 for n in (32, 40, 36)), indexing="ij")
    mask = (((x - 17) / 7.0) ** 2 + ((y - 21) / 9.0) ** 2 + ((z - 15) / 6.0) ** 2 <= 1.0).astype(np.uint8)
    bundle = _flag_bundle((4.4, 5.1, 3.9), angles, fd=12, bs=0.5)
    exact, _ = _flags_brute(bundle, mask, vol)
    wide, _ = _flags_brute(bundle, mask, vol, delta=1e-3)
    narrow, _ = _flags_brute(bundle, mask, vol, delta=-1e-3)
    sure = wide == narrow
    got = orc.beamlet_flags(bundle, mask, vol)
    assert sure.sum() >= 140 and 10 < exact.sum() < 130
    assert (got[sure] == exact[sure]).all()


# ---------------------------------------------------------------- beamletInit invariants (IMRTBeamBundle.cu:341-569)

def _n(v):
    return v / np.linalg.norm(v)


def _angle(u, v):
    return math.acos(max(-1.0, min(1.0, float(np.dot(_n(u), _n(v))))))


def _beamlet_geometry(bundle, ix, iy):
    """The beamlet's pyramid in the bundle's BEV, FP64: the corner rays (reference order), the two
    diagonal bisectors, and for each the in-plane distance of the nearest corner (IMRTBeamBundle.cu:
    359-421 define the canonical pyramid by these)."""
    bs = float(bundle.beamlet_size[0]), float(bundle.beamlet_size[1])
    hx = bundle.fluence_dim[0] * bs[0] * 0.5
    hy = bundle.fluence_dim[1] * bs[1] * 0.5
    sad = float(bundle.sad)
    v = [np.array([(ix + dx) * bs[0] - hx, sad, (iy + dy) * bs[1] - hy]) for dx, dy in ((0, 0), (1, 0), (1, 1), (0, 1))]
    v = [_n(a) for a in v]
    out = {}
    for case, (i, j), others in ((1, (0, 2), (0, 1, 3)), (2, (1, 3), (0, 1, 2))):
        c = _n(v[i] + v[j])
        proj = [a / np.dot(a, c) - c for a in v]
        out[case] = (c, proj, min(np.linalg.norm(proj[k]) for k in others))
    return v, out


def test_beamlet_init_canonical_pyramid():
    """For random bundles and all their beamlets, the canonical beamlet reproduces the geometry of the
    beamlet's pyramid (checked in FP64 from the definitions, not from the oracle's arithmetic):
      * its axis R'^-1 (0, 1, 0) is the diagonal bisector of the beamlet's corner rays (the case with the
        larger nearest-corner distance), so it aims at the beamlet's centre to within the bisector's
        offset;
      * its source iso' + R'^-1 (0, -sad', 0) is the bundle's source (sad' = SAD / axis.y);
      * its BEV -z axis is the in-plane direction toward corner 0 (case 1: the reference's v0 + v0, Q4)
        or toward the midpoint of corners 0 and 1 (case 2);
      * its field (sub_on x beamlet_size') has the sides of the rectangle whose half-diagonals reach the
        nearest corner distance toward corners 0 and 1; in case 2 the field's corner rays are that
        rectangle's corners;
      * the sub-fluence is the centred sub_on x sub_on block."""
    rng = np.random.default_rng(17)
    vol = orc.cast(orc.Volume, synth.make_volume((80, 80, 80)))
    seen = {1: 0, 2: 0, "mirror": 0, "tolerance": 0}
    for trial in range(6):
        ang = rng.uniform([-3.0, -3.0, -1.5], [3.0, 3.0, 1.5]).astype(np.float32)
        bundle = _flag_bundle((10.0, 10.0, 10.0), tuple(float(a) for a in ang), fd=16, bs=0.5)
        Mb = _bev_to_pvcs64(*(float(a) for a in bundle.angles))
        S = vec(bundle.isocenter) + Mb @ np.array([0.0, -100.0, 0.0])
        for ix in range(0, 16, 3):
            for iy in range(0, 16, 3):
                rc, out, flu = orc.beamlet_init(bundle, ix, iy, vol)
                assert rc == 0
                v, cases = _beamlet_geometry(bundle, ix, iy)
                case = 1 if cases[1][2] > cases[2][2] else 2
                c, proj, dist = cases[case]
                Mc = _bev_to_pvcs64(*(float(a) for a in out.angles))
                axis = Mc @ np.array([0.0, 1.0, 0.0])
                # (theta, phi): the first of (t, p), (t, pi - p), (-t, p), (-t, pi - p) whose axis aligns
                # with the bisector to cos >= 1 - 1e-4 (IMRTBeamBundle.cu:457-499)
                d = Mb @ c
                t = math.acos(d[1])
                p = math.asin(d[2] / math.sqrt(d[2] ** 2 + d[0] ** 2 + 1e-8))
                cand = [(t, p), (t, math.pi - p), (-t, p), (-t, math.pi - p)]
                align = [float(np.dot(_bev_to_pvcs64(a, b, 0.0) @ np.array([0.0, 1.0, 0.0]), d)) for a, b in cand]
                first = next(k for k in range(4) if align[k] >= 1 - 1e-4)
                if all(abs(x - (1 - 1e-4)) > 2e-6 for x in align[:first + 1]):
                    assert abs(float(out.angles[0]) - cand[first][0]) < 1e-4
                    assert abs(math.remainder(float(out.angles[1]) - cand[first][1], 2 * math.pi)) < 1e-4
                if _angle(axis, d) > 2e-5:
                    # accepted within the reference's tolerance, not exactly aligned: the axis (and
                    # with it the source) is off by up to acos(1 - 1e-4)
                    assert _angle(axis, d) < math.acos(1 - 1e-4) + 1e-5
                    seen["tolerance"] += 1
                    continue
                centre = Mb @ _n(v[0] + v[1] + v[2] + v[3])
                assert _angle(axis, centre) < 1e-5
                sad_c = float(out.sad)
                assert abs(sad_c - 100.0 / c[1]) < 2e-5 * sad_c
                Sc = vec(out.isocenter) + Mc @ np.array([0.0, -sad_c, 0.0])
                assert np.linalg.norm(Sc - S) < 2.5e-3   # the axis tolerance (2e-5 rad) times ~100 cm
                p0, p1 = _n(proj[0]), _n(proj[1])
                codir = Mc @ np.array([0.0, 0.0, -1.0])
                want = p0 if case == 1 else _n(p0 + p1)
                if abs(cases[1][2] - cases[2][2]) > 1e-6 * cases[1][2]:
                    err = _angle(codir, Mb @ want)
                    if err > 2e-3:
                        # the reference keeps +coll whenever it aligns to cos >= 1 - 1e-3
                        # (IMRTBeamBundle.cu:519-533), so a small collimator angle may keep the
                        # wrong sign: then the mirror (-coll) is the aligned one
                        Mm = _bev_to_pvcs64(float(out.angles[0]), float(out.angles[1]), -float(out.angles[2]))
                        assert _angle(Mm @ np.array([0.0, 0.0, -1.0]), Mb @ want) < 2e-3
                        assert err < math.acos(1 - 1e-3) + 1e-4
                        seen["mirror"] += 1
                    seen[case] += 1
                q0, q1 = c + dist * p0, c + dist * p1
                fw, fh = float(out.beamlet_size[0]) * 4, float(out.beamlet_size[1]) * 4
                assert abs(fw - np.linalg.norm(q0 - q1) * sad_c) < 1e-5 * fw
                assert abs(fh - np.linalg.norm(q0 + q1 - 2 * c) * sad_c) < 1e-5 * fh
                if case == 2 and abs(cases[1][2] - cases[2][2]) > 1e-6 * cases[1][2]:
                    corners = [Mc @ np.array([sx * fw / 2, sad_c, sz * fh / 2]) for sx in (-1, 1) for sz in (-1, 1)]
                    rect = [Mb @ (c + s * dist * p) for p in (p0, p1) for s in (-1, 1)]
                    for k in corners:
                        assert min(_angle(k, r) for r in rect) < 2e-5
                on = np.zeros((16, 16), dtype=np.float32)
                on[6:10, 6:10] = 1
                assert (flu.reshape(16, 16) == on).all()
                assert out.fmap_size[0] == 16 and out.long_spacing == np.float32(0.25)
    assert seen[1] >= 10 and seen[2] >= 10, seen
