"""The batched, precision and serving workloads at their full sizes (needs an MI355X).

- Batched 4pi: 200 x 200 x 160 body, 4 isocenters, 180 Fibonacci-sphere beams, 20 x 20 fluence,
  beamlets flagged on the device: every beam's matrix from the batched call (several chunks, and
  again with HIP-graph replay) equals the plain one-beam-per-call matrix bit for bit, and sampled
  beamlets (the longest among them) match the FP64 oracle.
- Precision: the 128^3 slab beam with the 1e-3 relative CSR threshold, BF16 / FP16 values with a
  per-beamlet FP32 scale, and 256 fluence vectors through D^T; the FP64 radiological path on every
  ray, and the 256-voxel rule on a volume whose rays are longer than that.
- Serving: 220 x 220 x 150 body with 1160 candidate beams; requests alone and in a mix are identical
  bit for bit, equal to the plain path, and sampled beamlets match the oracle.

Bar: per-beamlet |D - D_fp64| <= 0.5 % of the beamlet's maximum with gamma 1 %/1 mm >= 99 %, bitwise
equality wherever two paths of this library compute the same beam.
"""
import numpy as np
import pytest

import fastdose_amd as fd
import oracle as orc
from fastdose_amd import synth, workload

from gpu_util import beamlets_of, check_row, gpu_stages, log, oracle_many, same_matrix

pytestmark = pytest.mark.gpu

TOL = 0.005


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _sample_beamlets(store, n, seed, include_longest=True):
    rng = np.random.default_rng(seed)
    idx = sorted(set(rng.choice(store.n, size=n, replace=False).tolist()))
    if include_longest:
        L = np.array([store.beams[g].long_dim for g in range(store.n)])
        idx = sorted(set(idx) | {int(L.argmax())})
    return idx


def _check_sampled(cfg, store, idx, matrix_of_beam, tag):
    """Rows of sampled global beamlets (matrix_of_beam(k) -> host CSR of beam k) against the oracle."""
    items = [(store.beamlet(g), store.flu[g]) for g in idx]
    refs = oracle_many(cfg, items)
    worst = 0.0
    cache = {}
    for g, ref in zip(idx, refs):
        k, r = store.beam_of(g)
        if k not in cache:
            cache[k] = matrix_of_beam(k)
        off, col, val = cache[k]
        a, b = int(off[r]), int(off[r + 1])
        worst = max(worst, check_row(cfg.volume, col[a:b], val[a:b], ref, TOL, f"{tag} beamlet {g} (beam {k})"))
    log(f"{tag}: {len(idx)} sampled beamlets (L up to {max(store.beams[g].long_dim for g in idx)}), "
        f"worst |D - D_fp64| / max = {worst:.2e}")
    return worst


def test_config3_batched_4pi_full_size(torch_cuda):
    torch = torch_cuda
    log("config3: building the 200x200x160 body and 180 beams")
    cfg = synth.fourpi_config("batched-4pi-200x200x160", (200, 200, 160), 1003, 180, (22.0, 15.0, 13.0),
                              with_flags=False)
    nvox = 200 * 200 * 160
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel) as ctx:
        flags = workload.device_flags(ctx, cfg, torch)
        # device flags bit-exact with the oracle's FP32 flag test on a spread of bundles
        segs, seg_of = cfg.notes["segments"], cfg.notes["segment_of_beam"]
        for k in (0, 45, 91, 136, 179):
            want = orc.beamlet_flags(orc.cast(orc.Bundle, cfg.bundles[k]), segs[seg_of[k]],
                                     orc.cast(orc.Volume, cfg.volume))
            assert np.array_equal(flags[k], want), k
        store = workload.BeamStore(cfg, flags)
        assert 30 <= store.counts.min() and store.counts.max() <= 400
        log(f"config3: {store.n} beamlets, {store.counts.min()}..{store.counts.max()} per beam")
        arr, flu, cnt = store.request(range(180))
        plain = fd.DoseMat(nvox)
        for k in range(180):
            a, f, _ = store.request([k])
            plain.add_beam(ctx, a, f, cfg.extent)
        torch.cuda.synchronize()
        log("config3: plain done; batched in chunks of 8192 beamlets")
        batched = fd.DoseMat(nvox)
        batched.add_beams(ctx, arr, flu, cnt, 8192, cfg.extent)
        ctx.set_options(fd.FD_OPT_GRAPH)
        log("config3: batched with graph replay, chunks of 4096")
        graphed = fd.DoseMat(nvox)
        graphed.add_beams(ctx, arr, flu, cnt, 4096, cfg.extent)
        ctx.set_options(0)
        torch.cuda.synchronize()
        assert plain.info() == batched.info() == graphed.info()
        for k in range(180):
            p = plain.matrix(k)
            assert same_matrix(p, batched.matrix(k)), k
            assert same_matrix(p, graphed.matrix(k)), k
        log(f"config3: 180 matrices bitwise equal (plain / batched / graph), nnz {plain.info()[2]}")
        _check_sampled(cfg, store, _sample_beamlets(store, 6, 3), plain.matrix, "config3")
        for dm in (plain, batched, graphed):
            dm.close()


def _csr_of(refs, nvox, thr=0.0):
    """scipy CSR (FP64) of oracle rows; thr > 0 keeps entries >= thr x the row's maximum."""
    import scipy.sparse as sp

    indptr, cols, vals = [0], [], []
    for ref in refs:
        c, v = ref["cols"], ref["vals"]
        if thr > 0 and len(v):
            keep = v >= thr * v.max()
            c, v = c[keep], v[keep]
        cols.append(c)
        vals.append(v)
        indptr.append(indptr[-1] + len(c))
    return sp.csr_matrix((np.concatenate(vals), np.concatenate(cols), np.array(indptr)), shape=(len(refs), nvox))


def _decode16(q, vtype):
    if vtype == fd.FD_VAL_BF16:
        return (q.astype(np.uint32) << 16).view(np.float32)
    return q.view(np.float16).astype(np.float32)


def test_config4_precision_full_size(torch_cuda):
    torch = torch_cuda
    cfg = synth.config2()
    beams, flu = beamlets_of(cfg)
    n = len(beams)
    assert n == 256
    log("config4: oracle rows of all 256 slab beamlets")
    refs = oracle_many(cfg, [(beams[i], flu[i]) for i in range(n)])
    nvox = 128 ** 3
    D_full, D_thr = _csr_of(refs, nvox), _csr_of(refs, nvox, thr=1e-3)
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel) as ctx:
        ctx.set_threshold(1e-3)
        dm = fd.dose_mat_construction(ctx, cfg.bundles, cfg.flags, cfg.extent)
        off, col, val = dm.matrix(0)
        # per-beamlet max-dose error of the thresholded FP32, BF16 and FP16 columns vs FP64 (every row)
        worst = {}
        for name, vt in (("f32", fd.FD_VAL_F32), ("bf16", fd.FD_VAL_BF16), ("fp16", fd.FD_VAL_FP16)):
            if vt == fd.FD_VAL_F32:
                v = val
            else:
                dm.pack(vt)
                scale, q = dm.packed()
                rows_of = np.repeat(np.arange(n), np.diff(off.astype(np.int64)))
                v = _decode16(q, vt) * scale[rows_of]
            e = 0.0
            for i in range(n):
                a, b = int(off[i]), int(off[i + 1])
                ref = refs[i]
                u = np.union1d(col[a:b].astype(np.int64), ref["cols"])
                x = np.zeros(len(u))
                y = np.zeros(len(u))
                x[np.searchsorted(u, col[a:b].astype(np.int64))] = v[a:b]
                y[np.searchsorted(u, ref["cols"])] = ref["vals"]
                e = max(e, float(np.abs(x - y).max() / y.max()))
            worst[name] = e
            assert e <= TOL, (name, e)
        log(f"config4: per-beamlet max-dose error vs FP64 (threshold 1e-3): {worst}")
        # 256 seeded U[0,1] fluence vectors through D^T, FP32 accumulation on the device
        gen = torch.Generator(device="cuda").manual_seed(4)
        X = torch.rand((n, 256), dtype=torch.float32, device="cuda", generator=gen)
        Xh = X.cpu().numpy().astype(np.float64)
        sel = np.arange(0, 256, 16)
        Yfull = (D_full.T @ Xh[:, sel])
        Ythr = (D_thr.T @ Xh[:, sel])
        ymax = float(np.abs(Yfull).max())
        thr_cost = float(np.abs(Ythr - Yfull).max()) / ymax
        errs = {}
        for name, vt in (("f32", fd.FD_VAL_F32), ("bf16", fd.FD_VAL_BF16), ("fp16", fd.FD_VAL_FP16)):
            with fd.Spmm(dm, vt) as p:
                Y = p.run(X)
                torch.cuda.synchronize()
                Yg = Y[:, torch.from_numpy(sel).cuda()].cpu().numpy().astype(np.float64)
            # against the same thresholded product through the oracle's matrix: what the value type
            # costs; against the unthresholded oracle product: what the threshold costs on top
            e_thr = float(np.abs(Yg - Ythr).max()) / ymax
            e_full = float(np.abs(Yg - Yfull).max()) / ymax
            errs[name] = (e_thr, e_full)
            assert e_thr <= (1e-3 if vt == fd.FD_VAL_F32 else 4e-3), (name, e_thr)
            assert e_full <= thr_cost + e_thr + 1e-6, (name, e_full)
        log(f"config4: D^T X errors (vs oracle thresholded, vs oracle full): {errs}; "
            f"the 1e-3 threshold alone costs {thr_cost:.2e} of max")
        dm.close()
        # FP64 radiological path on every ray: Terma against the oracle on sampled beamlets
        ctx.set_threshold(0.0)
        ctx.set_options(fd.FD_OPT_RPL_FP64_ALL)
        idx = [0, 40, 100, 170, 255]
        sub = fd.beam_array([beams[i] for i in idx])
        g = gpu_stages(ctx, sub, flu[idx], cfg.extent)
        for j, i in enumerate(idx):
            L = beams[i].long_dim
            et = np.abs(g["terma"][j, :L] - refs[i]["terma"]).max() / refs[i]["terma"].max()
            ed = np.abs(g["dose"][j, :L] - refs[i]["dose"]).max() / refs[i]["dose"].max()
            assert et < 1e-4 and ed <= TOL, (i, et, ed)


def test_rpl_fp64_rule_on_long_rays(torch_cuda):
    """A 48 x 300 x 48 volume: a beam along y has rays of ~300 slices (> 256 voxels), one along x of ~48.
    FD_OPT_RPL_FP64 must take the FP64 path on the long beam only: identical to FD_OPT_RPL_FP64_ALL there,
    identical to the FP32 default on the short one; both within 1e-4 of the oracle's Terma."""
    rng = np.random.default_rng(77)
    dim = (48, 300, 48)
    vol = synth.make_volume(dim)
    density = rng.uniform(0.2, 1.8, size=dim[::-1]).astype(np.float32)
    iso = [np.float32(0.25 * d * 0.5) for d in dim]
    long_b = synth.make_bundle(iso, (0.0, 0.0, 0.0), fluence_dim=2)
    short_b = synth.make_bundle(iso, (90.0, 0.0, 0.0), fluence_dim=2)
    spec, kern = synth.synthetic_spectrum(1077), synth.synthetic_kernel(1077)
    bl, fl = fd.beamlets_init(long_b, np.ones(4, np.uint8), vol)
    bs, fs = fd.beamlets_init(short_b, np.ones(4, np.uint8), vol)
    assert min(b.long_dim for b in bl) * 0.25 > 256 * 0.25, [b.long_dim for b in bl]
    assert max(b.long_dim for b in bs) * 0.25 < 256 * 0.25
    out = {}
    for opt in (0, fd.FD_OPT_RPL_FP64, fd.FD_OPT_RPL_FP64_ALL):
        with fd.Context(vol, density, spec, kern, options=opt) as ctx:
            out[opt] = (gpu_stages(ctx, bl, fl), gpu_stages(ctx, bs, fs))
    g_long = {o: out[o][0]["terma"] for o in out}
    g_short = {o: out[o][1]["terma"] for o in out}
    assert np.array_equal(g_long[fd.FD_OPT_RPL_FP64], g_long[fd.FD_OPT_RPL_FP64_ALL])
    assert not np.array_equal(g_long[fd.FD_OPT_RPL_FP64], g_long[0])
    assert np.array_equal(g_short[fd.FD_OPT_RPL_FP64], g_short[0])
    ovol, ospec = orc.cast(orc.Volume, vol), orc.cast(orc.Spectrum, spec)
    for i in range(len(bl)):
        L = bl[i].long_dim
        T_ref, _ = orc.terma(orc.cast(orc.Beam, bl[i]), fl[i], ovol, density, ospec)
        for o in (0, fd.FD_OPT_RPL_FP64):
            et = np.abs(g_long[o][i, :L] - T_ref).max() / T_ref.max()
            assert et < 1e-4, (o, i, et)
    log(f"rpl64 rule: long beam L={bl[0].long_dim} takes the FP64 path, short beam L={bs[0].long_dim} does not")


def test_config5_serving_full_size(torch_cuda):
    torch = torch_cuda
    log("config5: building the 220x220x150 body and 1160 candidate beams")
    cfg = synth.fourpi_config("serving-4pi-220x220x150", (220, 220, 150), 1005, 1160, (26.0, 18.0, 15.0),
                              with_flags=False)
    nvox = 220 * 220 * 150
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel) as ctx:
        flags = workload.device_flags(ctx, cfg, torch)
        store = workload.BeamStore(cfg, flags)
        log(f"config5: {store.n} candidate beamlets, {store.counts.min()}..{store.counts.max()} per beam")
        rng = np.random.default_rng(1005)
        reqs = [list(rng.choice(1160, size=int(s), replace=False)) for s in (7, 20, 1, 13, 4)]
        A = reqs[0]
        plain = {}
        for k in sorted(set(k for ks in reqs for k in ks)):
            dm = fd.DoseMat(nvox)
            a, f, _ = store.request([k])
            dm.add_beam(ctx, a, f, cfg.extent)
            plain[k] = dm.matrix(0)
            dm.close()
        for slots, batch, budget in ((3, 4096, 0), (2, 1024, 1)):
            log(f"config5: server slots={slots} batch={batch} budget={budget}")
            with fd.Server(ctx, n_slots=slots, max_batch_beamlets=batch, hbm_budget=budget,
                           extent=cfg.extent) as srv:
                alone = srv.fetch(srv.submit(*store.request(A)))
                ids = [srv.submit(*store.request(ks)) for ks in reqs[1:3] + [A] + reqs[3:]]
                mats = [srv.fetch(i) for i in ids]
                st = srv.stats()
            mixed = mats[2]
            for j, k in enumerate(A):
                assert same_matrix(alone.matrix(j), mixed.matrix(j)), (slots, k)
                assert same_matrix(alone.matrix(j), plain[k]), (slots, k)
            for ks, dm in zip(reqs[1:3] + [A] + reqs[3:], mats):
                for j, k in enumerate(ks):
                    assert same_matrix(dm.matrix(j), plain[k]), (slots, k)
                dm.close()
            alone.close()
            if budget == 1:
                assert st["peak_inflight_batches"] == 1
            log(f"config5: stats {st}")
        # sampled beamlets of the requests against the oracle (the longest of them included)
        beams_in = sorted(set(k for ks in reqs for k in ks))
        cand = [int(store.first[k]) + r for k in beams_in for r in range(int(store.counts[k]))]
        L = np.array([store.beams[g].long_dim for g in cand])
        rs = np.random.default_rng(5)
        idx = sorted(set(rs.choice(cand, size=5, replace=False).tolist()) | {cand[int(L.argmax())]})
        _check_sampled(cfg, store, idx, lambda k: plain[k], "config5")
