"""Host geometry of the library against the oracle: bit-exact (CPU only, no GPU calls).

calc_range (beam.cu:134-221), beamletInit (IMRTBeamBundle.cu:341-569), beamlet/PTV flags
(IMRTBeamBundle.cu:130-316), PTV isocenter (IMRTBeamBundle.cu:36-65), table readers
(spectrum.cu:14-60, kernel.cu:13-72).
"""
import ctypes as C

import numpy as np
import pytest

import fastdose_amd as fd
import oracle as orc
from fastdose_amd import synth


def same_bytes(a, b):
    return bytes(a) == bytes(b)


def random_beam(rng):
    b = fd.Beam()
    for i in range(3):
        b.isocenter[i] = rng.uniform(4.0, 12.0)
        b.angles[i] = np.float32(rng.uniform(-np.pi, np.pi))
    b.beamlet_size[0] = b.beamlet_size[1] = rng.choice([0.125, 0.25, 0.5])
    b.fmap_size[0] = b.fmap_size[1] = 16
    b.sad = 100.0
    b.long_spacing = 0.25
    return b


def test_calc_range_bit_exact():
    rng = np.random.default_rng(42)
    vol = synth.make_volume((64, 60, 56), bbox_start=(4, 3, 2), bbox_dim=(50, 52, 48))
    ovol = orc.cast(orc.Volume, vol)
    hits = 0
    for _ in range(300):
        b = random_beam(rng)
        ob = orc.cast(orc.Beam, b)
        rc_o = orc.calc_range(ob, ovol)
        try:
            fd.calc_range(b, vol)
            rc_p = 0
        except fd.FastDoseError as e:
            assert e.code == -3
            rc_p = 1
        assert rc_o == rc_p
        if rc_p == 0:
            hits += 1
            assert same_bytes(b, ob)
    assert hits > 250


@pytest.mark.parametrize("make", [synth.config1, synth.config2])
def test_beamlet_init_bit_exact(make):
    cfg = make()
    bundle = cfg.bundles[0]
    beams, flu = fd.beamlets_init(bundle, cfg.flags[0], cfg.volume)
    assert len(beams) == int(cfg.flags[0].sum())
    ob = orc.cast(orc.Bundle, bundle)
    ov = orc.cast(orc.Volume, cfg.volume)
    fx = bundle.fluence_dim[0]
    k = 0
    for j in range(bundle.fluence_dim[1]):
        for i in range(fx):
            if not cfg.flags[0][i + j * fx]:
                continue
            rc, obm, ofl = orc.beamlet_init(ob, i, j, ov)
            assert rc == 0
            assert same_bytes(beams[k], obm), (i, j)
            assert np.array_equal(flu[k], ofl)
            k += 1
    # dose-mode sub-fluence: centre 4 x 4 of 16 x 16 is on (IMRTBeamBundle.cu:555-564)
    f = flu[0].reshape(16, 16)
    assert f[6:10, 6:10].all() and f.sum() == 16


def test_beamlet_init_random_bundles():
    rng = np.random.default_rng(7)
    vol = synth.make_volume((80, 80, 80))
    ov = orc.cast(orc.Volume, vol)
    for _ in range(20):
        angles = rng.uniform(-180, 180, 3)
        bundle = synth.make_bundle([10.0, 10.0, 10.0], angles, fluence_dim=20)
        ob = orc.cast(orc.Bundle, bundle)
        for (i, j) in [(0, 0), (19, 19), (3, 17), (10, 9), (12, 4)]:
            rc, obm, _ = orc.beamlet_init(ob, i, j, ov)
            try:
                pb, _ = fd.beamlet_init(bundle, i, j, vol)
                rc_p = 0
            except fd.FastDoseError:
                rc_p = 1
            assert rc == rc_p
            if rc == 0:
                assert same_bytes(pb, obm)


def ellipsoid_mask(dim, center, radii):
    z, y, x = np.meshgrid(np.arange(dim[2]), np.arange(dim[1]), np.arange(dim[0]), indexing="ij")
    r = ((x + 0.5 - center[0]) / radii[0]) ** 2 + ((y + 0.5 - center[1]) / radii[1]) ** 2 + \
        ((z + 0.5 - center[2]) / radii[2]) ** 2
    return (r <= 1.0).astype(np.uint8)


def test_ptv_flags_and_isocenter_bit_exact():
    dim = (48, 44, 40)
    vol = synth.make_volume(dim)
    ov = orc.cast(orc.Volume, vol)
    mask = ellipsoid_mask(dim, (26, 20, 18), (7, 5, 6))
    iso = fd.ptv_isocenter(mask, vol)
    oiso = (C.c_float * 3)()
    assert orc.lib().orc_ptv_isocenter(mask.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(ov), oiso) == 0
    assert np.array_equal(iso, np.array(oiso[:], dtype=np.float32))
    rng = np.random.default_rng(3)
    for _ in range(6):
        angles = rng.uniform(-180, 180, 3)
        bundle = synth.make_bundle(iso, angles, fluence_dim=20)
        flags = fd.beamlet_flag_init(bundle, mask, vol)
        oflags = orc.beamlet_flags(orc.cast(orc.Bundle, bundle), mask, ov)
        assert np.array_equal(flags, oflags)
        assert 0 < flags.sum() < 400


def test_table_readers_match_oracle(golden):
    s = fd.read_spectrum(golden / "spec_6mv.spec")
    os_ = orc.Spectrum()
    orc.lib().orc_read_spectrum(str(golden / "spec_6mv.spec").encode(), C.byref(os_))
    assert same_bytes(s, os_)
    k = fd.read_kernel(golden / "kernel_exp_6mv.txt", 8)
    ok = orc.Kernel()
    orc.lib().orc_read_kernel(str(golden / "kernel_exp_6mv.txt").encode(), 8, C.byref(ok))
    assert same_bytes(k, ok)


def test_reader_errors(tmp_path):
    with pytest.raises(fd.FastDoseError) as e:
        fd.read_spectrum(tmp_path / "missing.spec")
    assert e.value.code == -4
    bad = tmp_path / "k.txt"
    bad.write_text("header\n" + "0 0.1 1 1 1 1\n" * 25)
    with pytest.raises(fd.FastDoseError):
        fd.read_kernel(bad, 8)
    with pytest.raises(fd.FastDoseError):
        fd.read_kernel(bad, 17)


def test_beamlet_init_argument_errors():
    cfg = synth.config1()
    with pytest.raises(fd.FastDoseError):
        fd.beamlet_init(cfg.bundles[0], 8, 0, cfg.volume)
