"""CCCS march tables (needs an MI355X): the batched paths walk the lattice once per march geometry
(BEV pitch, longitudinal spacing) and cone, and every beamlet of that geometry reads the walk from
the table instead of marching in its own workgroup.  The results must be those of the march
wavefront bit for bit (the arithmetic is the same, split between the two kernels), which ties the
table path to the oracle checks of the march path.
"""
import numpy as np
import pytest

import fastdose_amd as fd
from fastdose_amd import synth
from gpu_util import log, same_matrix

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def four_beams(seed, n_phi, fluence_dim=3):
    """Four beams of one isocentre: beamlet k of every beam has the same march geometry."""
    cfg = synth.small_config(seed=seed, dim=(28, 26, 24), fluence_dim=fluence_dim)
    cfg.kernel = synth.synthetic_kernel(1000 + seed, n_phi=n_phi)
    iso = cfg.bundles[0].isocenter
    for ang in ((200.0, 10.0, 5.0), (75.0, -40.0, 0.0), (310.0, 80.0, 30.0)):
        cfg.bundles.append(synth.make_bundle(iso, ang, fluence_dim=fluence_dim))
        cfg.flags.append(np.ones(fluence_dim * fluence_dim, dtype=np.uint8))
    return cfg


def batched(cfg, options, chunk=0):
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel, options=options) as ctx:
        return fd.dose_mat_construction(ctx, cfg.bundles, cfg.flags, cfg.extent, batched=True, max_chunk=chunk)


@pytest.mark.parametrize("n_phi", [2, 5, 8, 10])
def test_tables_match_march_wavefront(torch_cuda, n_phi):
    cfg = four_beams(40 + n_phi, n_phi)
    for g in (0, fd.FD_OPT_EXACT_G):
        log(f"tables n_phi={n_phi} exact_g={bool(g)}")
        ref = batched(cfg, g | fd.FD_OPT_NO_TABLES)
        for chunk in (0, 13):
            tab = batched(cfg, g | fd.FD_OPT_TABLES, chunk)
            assert tab.info() == ref.info()
            for k in range(len(cfg.bundles)):
                assert same_matrix(ref.matrix(k), tab.matrix(k)), (n_phi, g, chunk, k)
        assert ref.info()[2] > 0


def test_tables_unsupported_beyond_ten_azimuths(torch_cuda):
    cfg = four_beams(61, 12)
    with pytest.raises(fd.FastDoseError) as e:
        batched(cfg, fd.FD_OPT_TABLES)
    assert e.value.code == fd.FD_ERR_UNSUPPORTED
    batched(cfg, 0)  # the default path marches in every workgroup there


def test_tables_under_graph_replay_and_serving(torch_cuda):
    """Tables built inside captured graphs (replayed with eager work between) and by the serving
    slots give the march wavefront's matrices."""
    torch = torch_cuda
    cfg = four_beams(52, 8)
    ref = batched(cfg, fd.FD_OPT_NO_TABLES)
    with fd.Context(cfg.volume, cfg.density, cfg.spectrum, cfg.kernel, options=fd.FD_OPT_TABLES) as ctx:
        ctx.set_options(fd.FD_OPT_TABLES | fd.FD_OPT_GRAPH)
        for rep in range(3):
            log(f"tables graph replay {rep}")
            dm = fd.dose_mat_construction(ctx, cfg.bundles, cfg.flags, cfg.extent, batched=True,
                                          max_chunk=(0, 20, 0)[rep])
            x = torch.ones(1 << 20, device="cuda")
            assert float(x.sum()) == float(1 << 20)
            torch.cuda.synchronize()
            for k in range(len(cfg.bundles)):
                assert same_matrix(ref.matrix(k), dm.matrix(k)), (rep, k)
        log("tables serving")
        with fd.Server(ctx, n_slots=2, max_batch_beamlets=16, options=fd.FD_OPT_TABLES, extent=cfg.extent) as srv:
            beams, flu, counts = fd.bundle_beamlets(cfg.bundles, cfg.flags, cfg.volume)
            dm = srv.fetch(srv.submit(beams, flu, counts))
        for k in range(len(cfg.bundles)):
            assert same_matrix(ref.matrix(k), dm.matrix(k)), ("serve", k)
