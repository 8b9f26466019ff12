"""HIP path vs the FP64 oracle, through the C-ABI (needs an MI355X).

Bar: BEV geometry and CSR column sets bit-exact; Terma, CCCS dose and resampled dose within 0.5%
of each beamlet's maximum (|D_gpu - D_fp64| <= 0.005 max_b D_fp64).
"""
import numpy as np
import pytest

import fastdose_amd as fd
import oracle as orc
from fastdose_amd import synth

from gpu_util import beamlets_of, dense_diff, gamma_pass_rate, gpu_stages, log, oracle_beamlet, row

pytestmark = pytest.mark.gpu

TOL = 0.005  # per-beamlet max-dose tolerance


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def check_against_oracle(cfg, beams, flu, g, idx, flags=0, stage_isolated=True):
    worst = {"terma": 0.0, "dose": 0.0, "csr": 0.0, "gamma_fail": 0.0}
    for i in idx:
        b = beams[i]
        L = b.long_dim
        ref = oracle_beamlet(cfg, b, flu[i], flags)
        T, R, D = g["terma"][i, :L], g["dens"][i, :L], g["dose"][i, :L]
        et = np.abs(T - ref["terma"]).max() / ref["terma"].max()
        er = np.abs(R - ref["dens"]).max() / max(ref["dens"].max(), 1e-30)
        ed = np.abs(D - ref["dose"]).max() / ref["dose"].max()
        cols, vals = row(g, i)
        assert np.all(np.diff(cols.astype(np.int64)) > 0), "columns not ascending"
        ec, sym = dense_diff(cols, vals, ref["cols"], ref["vals"])
        ec /= max(ref["vals"].max(), 1e-30)
        worst["terma"] = max(worst["terma"], et)
        worst["dose"] = max(worst["dose"], ed)
        worst["csr"] = max(worst["csr"], ec)
        assert et < 1e-4 and er < 1e-4, (i, et, er)
        assert ed <= TOL, (i, ed)
        assert ec <= TOL, (i, ec)
        gp = gamma_pass_rate(cfg.volume, ref["cols"], ref["vals"], cols, vals, dd=0.01, dta=0.1)
        worst["gamma_fail"] = max(worst["gamma_fail"], 1.0 - gp)
        assert gp >= 0.99, (i, gp)  # gamma 1 %/1 mm
        # index sets: identical except for entries that are structurally ~0 on one side
        if len(sym):
            u = np.union1d(cols, ref["cols"])
            tiny = 1e-6 * ref["vals"].max()
            vg = dict(zip(cols.tolist(), vals.tolist()))
            vr = dict(zip(ref["cols"].tolist(), ref["vals"].tolist()))
            for c in sym.tolist():
                assert max(vg.get(c, 0.0), vr.get(c, 0.0)) <= tiny, (i, c)
            assert len(sym) <= 1e-4 * len(u) + 2
        if stage_isolated:
            # the resampler on the GPU's own BEV dose: decisions and index sets bit-exact
            c2, v2, _ = orc.resample(orc.cast(orc.Beam, b), D.astype(np.float64), orc.cast(orc.Volume, cfg.volume),
                                     cfg.extent)
            assert np.array_equal(cols.astype(np.int64), c2), i
            # same samples, same BEV pixels; the GPU sums up to 125 of them in FP32 (relative error
            # <= 124 * 2^-24 < 1e-5), the oracle in FP64
            rel = np.abs(vals - v2) / np.maximum(np.abs(v2), 1e-30)
            assert np.all((rel <= 1e-5) | (np.abs(vals - v2) <= 1e-7 * v2.max())), (i, float(rel.max()))
    log(f"worst rel errors {worst}")
    return worst


def test_small_hetero(torch_cuda):
    cfg = synth.small_config(seed=7, dim=(24, 28, 20), fluence_dim=2)
    beams, flu = beamlets_of(cfg)
    log("small_hetero ctx")
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel) as ctx:
        g = gpu_stages(ctx, beams, flu, cfg.extent)
    check_against_oracle(cfg, beams, flu, g, range(len(beams)))


def test_small_exact_g(torch_cuda):
    cfg = synth.small_config(seed=8, dim=(22, 22, 26), fluence_dim=2, angles=(-50.0, 70.0, 30.0))
    beams, flu = beamlets_of(cfg)
    log("small_exact_g ctx")
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel, options=fd.FD_OPT_EXACT_G) as ctx:
        g = gpu_stages(ctx, beams, flu, cfg.extent)
    check_against_oracle(cfg, beams, flu, g, range(len(beams)), flags=orc.EXACT_G)


@pytest.mark.parametrize("n_phi", [2, 5, 16, 17, 24])
def test_azimuth_counts(torch_cuda, n_phi):
    """Workgroup shapes other than 8 azimuths: 2 (no slab prefetch), 5 (two slab values per
    thread), 16 (the kernel whose wavefronts march themselves, one azimuth each) and 17 and 24 (two
    azimuths per self-marching wavefront, the last one alone for 17; the reference accepts up to 24,
    DoseComp.cu:26-30)."""
    cfg = synth.small_config(seed=11, dim=(22, 24, 22), fluence_dim=2, angles=(40.0, -30.0, 15.0))
    cfg.kernel = synth.synthetic_kernel(1011, n_phi=n_phi)
    beams, flu = beamlets_of(cfg)
    log(f"azimuths {n_phi} ctx")
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel) as ctx:
        g = gpu_stages(ctx, beams, flu, cfg.extent)
    check_against_oracle(cfg, beams, flu, g, range(len(beams)))


def test_config1_all_beamlets(torch_cuda):
    cfg = synth.config1()
    beams, flu = beamlets_of(cfg)
    assert len(beams) == 64
    log("config1 ctx")
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel) as ctx:
        g = gpu_stages(ctx, beams, flu, cfg.extent)
    check_against_oracle(cfg, beams, flu, g, range(0, 64, 3))


def test_config2_sampled_beamlets(torch_cuda):
    cfg = synth.config2()
    beams, flu = beamlets_of(cfg)
    assert len(beams) == 256
    log("config2 ctx")
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel) as ctx:
        g = gpu_stages(ctx, beams, flu, cfg.extent)
    # size-independent properties on every row
    off = g["offsets"]
    assert off[0] == 0 and np.all(np.diff(off) > 0)
    check_against_oracle(cfg, beams, flu, g, [0, 17, 37, 60, 85, 120, 135, 150, 178, 200, 230, 255],
                         stage_isolated=True)


def test_segment_count_matches_oracle(torch_cuda):
    cfg = synth.config1()
    beams, flu = beamlets_of(cfg)
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel) as ctx:
        nseg = ctx.cccs_segments(beams)
    k = orc.cast(orc.Kernel, cfg.kernel)
    for i in (0, 9, 63):
        assert nseg[i] == orc.cccs_count(orc.cast(orc.Beam, beams[i]), k)


def test_deterministic_and_batch_independent(torch_cuda):
    """Each beamlet's result is the same alone and inside a batch, bit for bit."""
    cfg = synth.small_config(seed=9, dim=(26, 24, 22), fluence_dim=3)
    beams, flu = beamlets_of(cfg)
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel) as ctx:
        g1 = gpu_stages(ctx, beams, flu, cfg.extent)
        g2 = gpu_stages(ctx, beams, flu, cfg.extent)
        for k in (0, 4, len(beams) - 1):
            alone = gpu_stages(ctx, fd.beam_array([beams[k]]), flu[k:k + 1], cfg.extent)
            L = beams[k].long_dim
            assert np.array_equal(alone["dose"][0, :L], g1["dose"][k, :L])
            c1, v1 = row(g1, k)
            c2, v2 = row(alone, 0)
            assert np.array_equal(c1, c2) and np.array_equal(v1, v2)
    for key in ("terma", "dose", "cols", "vals", "offsets"):
        assert np.array_equal(g1[key], g2[key]), key


def test_stage_descriptors_follow_the_beams(torch_cuda):
    """The stage calls reuse the device's beamlet descriptors only when they are the ones passed:
    stages of beams A, of beams B (same count), a batched run (which replaces the descriptors) and
    A again give A's first result bit for bit, and B equals B in a fresh context."""
    cfg = synth.small_config(seed=31, dim=(24, 26, 22), fluence_dim=2)
    iso = cfg.bundles[0].isocenter
    cfg.bundles.append(synth.make_bundle(iso, (130.0, -20.0, 10.0), fluence_dim=2))
    cfg.flags.append(np.ones(4, dtype=np.uint8))
    a, fa = fd.beamlets_init(cfg.bundles[0], cfg.flags[0], cfg.volume)
    b, fb = fd.beamlets_init(cfg.bundles[1], cfg.flags[1], cfg.volume)
    assert len(a) == len(b)
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel) as ctx:
        ga = gpu_stages(ctx, a, fa, cfg.extent)
        gb = gpu_stages(ctx, b, fb, cfg.extent)
        fd.dose_mat_construction(ctx, cfg.bundles, cfg.flags, cfg.extent, batched=True)
        ga2 = gpu_stages(ctx, a, fa, cfg.extent)
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel) as ctx:
        gb_fresh = gpu_stages(ctx, b, fb, cfg.extent)
    for key in ("terma", "dose", "cols", "vals", "offsets"):
        assert np.array_equal(ga[key], ga2[key]), key
        assert np.array_equal(gb[key], gb_fresh[key]), key


def test_dense_to_csr_known_and_random(torch_cuda):
    torch = torch_cuda
    dense = np.array([[0, 0, 1, 0, 2, 0], [2, 0, 0, 1, 0, 0], [0, 0, 1, 2, 0, 1], [1, 0, 1, 0, 0, 0],
                      [1, 1, 0, 0, 0, 1]], dtype=np.float32)
    off, col, val = fd.dense_to_csr(torch.from_numpy(dense).cuda())
    assert off.cpu().tolist() == [0, 2, 4, 7, 9, 12]
    assert col.cpu().tolist() == [2, 4, 0, 3, 2, 3, 5, 0, 2, 0, 1, 5]
    assert val.cpu().tolist() == [1, 2, 2, 1, 1, 2, 1, 1, 1, 1, 1, 1]
    dense4 = np.array([[0, 0, 1, 2], [1, 0, 0, 1], [0, 0, 1, 0], [2, 0, 0, 0]], dtype=np.float32)
    off, col, val = fd.dense_to_csr(torch.from_numpy(dense4).cuda())
    assert off.cpu().tolist() == [0, 2, 4, 5, 6]
    assert col.cpu().tolist() == [2, 3, 0, 3, 2, 0]
    rng = np.random.default_rng(0)
    for rows, cols in [(1, 1), (3, 1000), (37, 5003), (0, 10), (5, 0)]:
        d = rng.uniform(size=(rows, cols)).astype(np.float32)
        d[d < 0.7] = 0.0
        off, col, val = fd.dense_to_csr(torch.from_numpy(d).cuda())
        o2, c2, v2 = orc.dense_to_csr(d) if rows and cols else (np.zeros(rows + 1, np.uint64), [], [])
        assert np.array_equal(off.cpu().numpy().astype(np.uint64), np.asarray(o2, dtype=np.uint64))
        assert np.array_equal(col.cpu().numpy(), np.asarray(c2, dtype=np.int64))
        assert np.array_equal(val.cpu().numpy(), np.asarray(v2, dtype=np.float32))


def test_dosemat_files_roundtrip(torch_cuda, tmp_path):
    """DoseMatConstruction -> tofile -> readDose.m semantics -> fromfile."""
    cfg = synth.small_config(seed=10, dim=(24, 24, 24), fluence_dim=2)
    b2 = synth.make_bundle(cfg.bundles[0].isocenter, (120.0, -30.0, 0.0), fluence_dim=2)
    cfg.bundles.append(b2)
    cfg.flags.append(np.array([1, 0, 1, 1], dtype=np.uint8))
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel) as ctx:
        dm = fd.dose_mat_construction(ctx, cfg.bundles, cfg.flags, cfg.extent)
        nm, rows, nnz = dm.info()
        assert nm == 2 and rows == 4 + 3
        folder = tmp_path / "doseMatFolder"
        dm.tofile(folder)
        fd.write_fluence_map(tmp_path / "fluenceMap.bin", cfg.flags)
        mats = fd.read_dosemat_folder(folder)
        for k in range(2):
            off, col, val = dm.matrix(k)
            assert np.array_equal(mats[k][0], off)
            assert np.array_equal(mats[k][1], col)
            assert np.array_equal(mats[k][2], val)
        # the file set of MatCSREnsemble::tofile
        for name in ("NonZeroElements", "CumuNonZeroElements", "numRowsPerMat", "CumuNumRowsPerMat",
                     "OffsetBufferIdx", "offsetsBuffer", "columnsBuffer", "valuesBuffer"):
            assert (folder / f"{name}.bin").exists()
        assert np.fromfile(folder / "OffsetBufferIdx.bin", dtype=np.uint64).tolist() == [0, 5]
        assert np.fromfile(tmp_path / "fluenceMap.bin", dtype=np.uint8).tolist() == [1, 1, 1, 1, 1, 0, 1, 1]
        dm2 = fd.DoseMat(dm.n_cols)
        dm2.fromfile(folder)
        assert dm2.info() == dm.info()
        for k in range(2):
            a, b = dm.matrix(k), dm2.matrix(k)
            assert all(np.array_equal(x, y) for x, y in zip(a, b))
        # a truncated folder is refused on the host and the loaded ensemble stays as it was
        bad = tmp_path / "truncated"
        bad.mkdir()
        for f in folder.iterdir():
            (bad / f.name).write_bytes(f.read_bytes())
        (bad / "columnsBuffer.bin").write_bytes((folder / "columnsBuffer.bin").read_bytes()[:-8])
        with pytest.raises(fd.FastDoseError):
            dm2.fromfile(bad)
        assert dm2.info() == dm.info()
        assert all(np.array_equal(x, y) for x, y in zip(dm.matrix(1), dm2.matrix(1)))
        # doseMerge.py: two runs' folders merged read back as their matrices in order
        dm_b = fd.dose_mat_construction(ctx, cfg.bundles[1:], cfg.flags[1:], cfg.extent)
        folder_b = tmp_path / "second" / "doseMatFolder"
        folder_b.parent.mkdir()
        dm_b.tofile(folder_b)
        fd.write_fluence_map(folder / "fluenceMap.bin", cfg.flags)
        fd.write_fluence_map(folder_b / "fluenceMap.bin", cfg.flags[1:])
        merged = tmp_path / "merged"
        fd.merge_dosemat_folders([folder, folder_b], merged)
        dm3 = fd.DoseMat(dm.n_cols)
        dm3.fromfile(merged)
        assert dm3.info()[0] == 3
        for k, (src, j) in enumerate(((dm, 0), (dm, 1), (dm_b, 0))):
            assert all(np.array_equal(x, y) for x, y in zip(src.matrix(j), dm3.matrix(k))), k
        # the dose matrix rows equal the split-stage rows
        beams, flu = fd.beamlets_init(cfg.bundles[0], cfg.flags[0], cfg.volume)
        g = gpu_stages(ctx, beams, flu, cfg.extent)
        off, col, val = dm.matrix(0)
        assert np.array_equal(off.astype(np.int64), g["offsets"])
        assert np.array_equal(col.astype(np.int64), g["cols"]) and np.array_equal(val, g["vals"])


def test_batched_beams_match_per_beam(torch_cuda):
    """fd_dosemat_add_beams: beams batched into chunks give each beam's per-beam matrix, bit for bit."""
    cfg = synth.small_config(seed=12, dim=(28, 26, 24), fluence_dim=3)
    iso = cfg.bundles[0].isocenter
    for ang in ((200.0, 10.0, 5.0), (75.0, -40.0, 0.0), (310.0, 80.0, 30.0)):
        cfg.bundles.append(synth.make_bundle(iso, ang, fluence_dim=3))
        cfg.flags.append(np.ones(9, dtype=np.uint8))
    cfg.flags[2] = np.zeros(9, dtype=np.uint8)  # a beam with no active beamlet
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel) as ctx:
        ref = fd.dose_mat_construction(ctx, cfg.bundles, cfg.flags, cfg.extent)
        for chunk in (0, 7, 1):
            log(f"batched chunk={chunk}")
            dm = fd.dose_mat_construction(ctx, cfg.bundles, cfg.flags, cfg.extent, batched=True, max_chunk=chunk)
            assert dm.info() == ref.info()
            for k in range(len(cfg.bundles)):
                a, b = ref.matrix(k), dm.matrix(k)
                assert all(np.array_equal(x, y) for x, y in zip(a, b)), (chunk, k)
        assert ref.matrix(2)[0].tolist() == [0]


def test_reciprocal_variants(torch_cuda):
    """The resampler with the significand correction forced on (FD_OPT_IEEE_DIV) gives the same
    matrix, bit for bit, as the default (which skips it where the device check passed), and both
    pass the oracle checks."""
    cfg = synth.small_config(seed=23, dim=(24, 26, 22), fluence_dim=2, angles=(-35.0, 25.0, 60.0))
    beams, flu = beamlets_of(cfg)
    out = []
    for opt in (0, fd.FD_OPT_IEEE_DIV):
        log(f"reciprocal variant options={opt}")
        with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel, options=opt) as ctx:
            out.append(gpu_stages(ctx, beams, flu, cfg.extent))
    for i in range(len(beams)):
        c0, v0 = row(out[0], i)
        c1, v1 = row(out[1], i)
        assert np.array_equal(c0, c1) and np.array_equal(v0, v1), i
    check_against_oracle(cfg, beams, flu, out[1], range(len(beams)))


@pytest.mark.parametrize("segments", [1, 2, 4, 8, 16])
def test_terma_segments(torch_cuda, segments):
    """Terma with 1, 2, 4, 8 or 16 ray segments (a tuning parameter) stays within the bar."""
    cfg = synth.small_config(seed=21, dim=(26, 24, 30), fluence_dim=2, angles=(15.0, 50.0, -20.0))
    beams, flu = beamlets_of(cfg)
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel) as ctx:
        ctx.set_tuning(8192, 0, segments)
        assert ctx.tuning["terma_segments"] == segments
        g = gpu_stages(ctx, beams, flu, cfg.extent)
    check_against_oracle(cfg, beams, flu, g, range(len(beams)))


def test_autotune(torch_cuda):
    """fd_autotune picks a valid setting, keeps it in the context, and batched construction with it
    gives each beam's per-beam matrix (same context, same setting) bit for bit."""
    cfg = synth.small_config(seed=22, dim=(28, 26, 24), fluence_dim=3)
    iso = cfg.bundles[0].isocenter
    for ang in ((200.0, 10.0, 5.0), (75.0, -40.0, 0.0), (310.0, 80.0, 30.0), (120.0, 20.0, 60.0)):
        cfg.bundles.append(synth.make_bundle(iso, ang, fluence_dim=3))
        cfg.flags.append(np.ones(9, dtype=np.uint8))
    beamlets, fls, counts = [], [], []
    for b, f in zip(cfg.bundles, cfg.flags):
        bl, fl = fd.beamlets_init(b, f, cfg.volume)
        beamlets.extend(bl)
        fls.append(fl)
        counts.append(len(bl))
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel) as ctx:
        with pytest.raises(fd.FastDoseError):
            ctx.set_tuning(0, 0, 4)
        with pytest.raises(fd.FastDoseError):
            ctx.set_tuning(1024, 0, 3)
        t = ctx.autotune(beamlets, np.concatenate(fls, 0), counts, cfg.extent)
        log(f"autotune {t}")
        assert t["batch_beamlets"] in (1024, 2048, 4096, 8192, 16384) and t["graph"] in (0, 1)
        assert t["terma_segments"] in (4, 8, 16) and t["beamlets_per_s"] > 0
        assert ctx.tuning == t
        ref = fd.dose_mat_construction(ctx, cfg.bundles, cfg.flags, cfg.extent)
        dm = fd.dose_mat_construction(ctx, cfg.bundles, cfg.flags, cfg.extent, batched=True, max_chunk=0)
        assert dm.info() == ref.info()
        for k in range(len(cfg.bundles)):
            assert all(np.array_equal(x, y) for x, y in zip(ref.matrix(k), dm.matrix(k))), k


def test_device_beamlet_flags_match_host(torch_cuda):
    """fd_beamlet_flags (device, many bundles) == the oracle's flags (orc_beamlet_flags, pinned by its
    closed-form and brute-force tests) == fd_beamlet_flags_host, bit for bit."""
    torch = torch_cuda
    cfg = synth.fourpi_config("flags-4pi", (60, 56, 48), 1011, 24, (7.0, 5.0, 4.0), fluence_dim=12,
                              with_flags=False)
    segs = cfg.notes["segments"]
    seg_of = cfg.notes["segment_of_beam"]
    ovol = orc.cast(orc.Volume, cfg.volume)
    on = 0
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel) as ctx:
        for s_idx, mask in enumerate(segs):
            ks = [k for k in range(len(cfg.bundles)) if seg_of[k] == s_idx]
            bundles = [cfg.bundles[k] for k in ks]
            mask_d = torch.from_numpy(np.ascontiguousarray(mask, dtype=np.uint8)).cuda()
            dev = ctx.beamlet_flags(mask_d, bundles)
            for b, f in zip(bundles, dev):
                assert np.array_equal(orc.beamlet_flags(orc.cast(orc.Bundle, b), mask, ovol), f)
                assert np.array_equal(fd.beamlet_flag_init(b, mask, cfg.volume), f)
                on += int(f.sum())
        assert on > 0
        # a bundle far outside the volume: all flags 0
        far = synth.make_bundle(np.array([500.0, 500.0, 500.0], np.float32), (0.0, 0.0, 0.0), fluence_dim=4)
        mask_d = torch.from_numpy(np.ascontiguousarray(segs[0], dtype=np.uint8)).cuda()
        assert not ctx.beamlet_flags(mask_d, [far])[0].any()


def test_rpl_fp64_rule(torch_cuda):
    """FD_OPT_RPL_FP64 uses FP64 only on rays longer than 256 voxels (none in this volume, so the
    output is the FP32 one bit for bit); FD_OPT_RPL_FP64_ALL uses it everywhere and stays within the
    FP32 tolerance of the oracle."""
    cfg = synth.small_config(seed=14, dim=(24, 24, 24), fluence_dim=2)
    beams, flu = beamlets_of(cfg)
    outs = {}
    for opt in (0, fd.FD_OPT_RPL_FP64, fd.FD_OPT_RPL_FP64_ALL):
        with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel, options=opt) as ctx:
            outs[opt] = gpu_stages(ctx, beams, flu, cfg.extent)
    assert np.array_equal(outs[0]["terma"], outs[fd.FD_OPT_RPL_FP64]["terma"])
    g = outs[fd.FD_OPT_RPL_FP64_ALL]
    for i in range(len(beams)):
        L = beams[i].long_dim
        ref = oracle_beamlet(cfg, beams[i], flu[i])
        et = np.abs(g["terma"][i, :L] - ref["terma"]).max() / ref["terma"].max()
        assert et < 1e-4, (i, et)


def test_graph_replays_match_eager(torch_cuda):
    """FD_OPT_GRAPH: each batch captured into a HIP graph and replayed (instantiated once, then
    updated in place); every replay equals the eager launches bit for bit, with eager work and
    synchronisations between the replays."""
    torch = torch_cuda
    cfg = synth.small_config(seed=18, dim=(28, 26, 24), fluence_dim=3)
    iso = cfg.bundles[0].isocenter
    for ang in ((200.0, 10.0, 5.0), (75.0, -40.0, 0.0)):
        cfg.bundles.append(synth.make_bundle(iso, ang, fluence_dim=3))
        cfg.flags.append(np.ones(9, dtype=np.uint8))
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel) as ctx:
        ref = fd.dose_mat_construction(ctx, cfg.bundles, cfg.flags, cfg.extent)
        ctx.set_options(fd.FD_OPT_GRAPH)
        for rep in range(4):
            log(f"graph replay {rep}")
            chunk = (0, 7, 4, 0)[rep]
            dm = fd.dose_mat_construction(ctx, cfg.bundles, cfg.flags, cfg.extent, batched=True, max_chunk=chunk)
            x = torch.ones(1 << 20, device="cuda")  # eager work between replays
            assert float(x.sum()) == float(1 << 20)
            torch.cuda.synchronize()
            for k in range(len(cfg.bundles)):
                assert all(np.array_equal(a, b) for a, b in zip(ref.matrix(k), dm.matrix(k))), (rep, k)
        with fd.Server(ctx, n_slots=2, max_batch_beamlets=10, options=fd.FD_OPT_GRAPH, extent=cfg.extent) as srv:
            beams, flu, counts = fd.bundle_beamlets(cfg.bundles, cfg.flags, cfg.volume)
            dm = srv.fetch(srv.submit(beams, flu, counts))
        for k in range(len(cfg.bundles)):
            assert all(np.array_equal(a, b) for a, b in zip(ref.matrix(k), dm.matrix(k))), ("serve", k)


def test_empty_and_dark_inputs(torch_cuda):
    """No beamlets: every stage is a no-op and the CSR has one zero offset.  A beamlet with zero
    fluence: zero Terma, zero dose, an empty CSR row between two ordinary ones (exact zeros are
    dropped, as cusparseDenseToSparse does), and the neighbours equal their results alone."""
    torch = torch_cuda
    cfg = synth.small_config(seed=13, dim=(20, 22, 18), fluence_dim=2)
    beams, flu = beamlets_of(cfg)
    assert len(beams) >= 3
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel) as ctx:
        T = torch.zeros(256, device="cuda")
        ctx.terma([], torch.zeros(256, device="cuda"), T, T, 256)
        ctx.cccs([], T, T, T, 256)
        off, cols, vals = ctx.resample_csr([], T, 256, cfg.extent)
        torch.cuda.synchronize()
        assert off.cpu().tolist() == [0]
        dark = flu.copy()
        dark[1] = 0.0
        g = gpu_stages(ctx, beams, dark, cfg.extent)
        ref = gpu_stages(ctx, beams, flu, cfg.extent)
    L = beams[1].long_dim
    assert not g["terma"][1, :L].any() and not g["dose"][1, :L].any()
    c1, v1 = row(g, 1)
    assert len(c1) == 0
    for k in (0, 2):
        ca, va = row(g, k)
        cb, vb = row(ref, k)
        assert np.array_equal(ca, cb) and np.array_equal(va, vb), k


def test_unaligned_stride(torch_cuda):
    """A stride that is not a multiple of four floats (the cone sum then takes its scalar form)
    gives the same BEV volumes, bit for bit, as the aligned one."""
    torch = torch_cuda
    cfg = synth.small_config(seed=17, dim=(20, 20, 22), fluence_dim=2)
    beams, flu = beamlets_of(cfg)
    n = len(beams)
    Lmax = max(b.long_dim for b in beams)
    out = []
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel) as ctx:
        for stride in (Lmax * 256, Lmax * 256 + 3):
            flu_d = torch.from_numpy(np.ascontiguousarray(flu, dtype=np.float32)).cuda()
            T = torch.zeros(n * stride, device="cuda")
            R = torch.zeros_like(T)
            D = torch.zeros_like(T)
            ctx.terma(beams, flu_d, T, R, stride)
            ctx.cccs(beams, T, R, D, stride)
            torch.cuda.synchronize()
            d = D.cpu().numpy()
            out.append([d[i * stride:i * stride + beams[i].long_dim * 256] for i in range(n)])
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_error_paths(torch_cuda):
    torch = torch_cuda
    cfg = synth.small_config(seed=11, dim=(16, 16, 16), fluence_dim=1)
    beams, flu = beamlets_of(cfg)
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel) as ctx:
        stride = beams[0].long_dim * 256
        T = torch.zeros(stride, device="cuda")
        with pytest.raises(fd.FastDoseError):
            ctx.terma(beams, torch.from_numpy(flu).cuda(), T, T, stride - 1)  # stride too small
        bad = fd.Beam()
        bad.fmap_size[0] = bad.fmap_size[1] = 8
        bad.long_dim = 4
        with pytest.raises(fd.FastDoseError) as e:
            ctx.cccs([bad], T, T, T, stride)
        assert e.value.code == -6
        # subFluenceDim 8 (IMRTArgs.cpp --subFluenceDim): the host geometry builds the beamlet, the
        # whole-path call rejects it before any kernel runs
        b8 = synth.make_bundle(cfg.bundles[0].isocenter, (30.0, 20.0, 10.0), fluence_dim=1, sub_dim=8, sub_on=2)
        bb, _ = fd.beamlets_init(b8, np.ones(1, np.uint8), cfg.volume)
        assert len(bb) == 1 and bb[0].fmap_size[0] == 8
        dm = fd.DoseMat(16 ** 3)
        with pytest.raises(fd.FastDoseError) as e:
            dm.add_beam(ctx, bb, np.zeros((1, 256), np.float32), cfg.extent)
        assert e.value.code == -6 and dm.info()[0] == 0
        dm.close()
