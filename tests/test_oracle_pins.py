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
