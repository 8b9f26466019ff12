"""Serving loop (fd_serve_*): continuous batching over a resident context (needs an MI355X).

Each beam's matrix must be bitwise identical whether its request runs alone or in a mix of
concurrent requests, and identical to the plain one-beam-at-a-time path."""
import numpy as np
import pytest

import fastdose_amd as fd
from fastdose_amd import synth

from gpu_util import log

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _case():
    cfg = synth.small_config(seed=13, dim=(30, 28, 26), fluence_dim=4)
    rng = np.random.default_rng(5)
    iso0 = np.asarray(cfg.bundles[0].isocenter, dtype=np.float32)
    bundles, flags = [], []
    for k in range(8):
        ang = (float(rng.uniform(0, 360)), float(rng.uniform(-60, 60)), float(rng.uniform(0, 90)))
        iso = iso0 + rng.uniform(-0.5, 0.5, 3).astype(np.float32)
        bundles.append(synth.make_bundle(iso, ang, fluence_dim=4))
        f = (rng.uniform(size=16) < 0.7).astype(np.uint8)
        f[5] = 1
        flags.append(f)
    return cfg, bundles, flags


def _request(per_beam, ks):
    beams, flus, counts = [], [], []
    for k in ks:
        b, f = per_beam[k]
        beams.extend(b)
        flus.append(f)
        counts.append(len(b))
    return fd.beam_array(beams), np.concatenate(flus, 0), counts


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def test_serve_alone_and_in_mix(torch_cuda):
    cfg, bundles, flags = _case()
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel) as ctx:
        ref = fd.dose_mat_construction(ctx, bundles, flags, cfg.extent)
        per_beam = [fd.beamlets_init(b, f, cfg.volume) for b, f in zip(bundles, flags)]
        mixes = ([0], [1, 4], [2, 3], [5, 6, 7], [3], list(range(8)), [6, 2])
        for slots, batch, budget in ((3, 20, 0), (2, 64, 1), (1, 4096, 0)):
            log(f"serve slots={slots} batch={batch} budget={budget}")
            with fd.Server(ctx, n_slots=slots, max_batch_beamlets=batch, hbm_budget=budget,
                           extent=cfg.extent) as srv:
                rid = srv.submit(*_request(per_beam, [2, 3]))
                alone = srv.fetch(rid)
                srv.reset_stats()
                ids = [srv.submit(*_request(per_beam, ks)) for ks in mixes]
                for i in ids:
                    srv.wait(i, 120.0)
                    assert srv.poll(i) == 1
                mats = [srv.fetch(i) for i in ids]
                st = srv.stats()
            assert _same(alone.matrix(0), mats[2].matrix(0)) and _same(alone.matrix(1), mats[2].matrix(1))
            for ks, dm in zip(mixes, mats):
                assert dm.info()[0] == len(ks)
                for j, k in enumerate(ks):
                    assert _same(dm.matrix(j), ref.matrix(k)), (slots, batch, ks, k)
            nb = sum(len(ks) for ks in mixes)
            assert st["beams_done"] == nb and st["requests_done"] == len(mixes)
            assert st["p99_ms"] >= st["p50_ms"] > 0
            if budget == 1:
                assert st["peak_inflight_batches"] == 1  # admission by footprint serialises
            log(f"stats {st}")


def test_serve_errors(torch_cuda):
    cfg, bundles, flags = _case()
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel) as ctx:
        with pytest.raises(fd.FastDoseError):
            fd.Server(ctx, n_slots=99)
        with fd.Server(ctx, n_slots=1) as srv:
            with pytest.raises(fd.FastDoseError):
                srv.poll(12345)
            b, f = fd.beamlets_init(bundles[0], flags[0], cfg.volume)
            with pytest.raises(ValueError):
                srv.submit(b, f, [len(b) + 1])
            bad = fd.beam_array(list(b))
            bad[0].long_dim = 0
            with pytest.raises(fd.FastDoseError):
                srv.submit(bad, f, [len(b)])
            rid = srv.submit(b, f, [len(b)])
            dm = srv.fetch(rid)
            assert dm.info()[0] == 1
