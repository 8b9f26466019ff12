"""The gamma-index helper the GPU parity tests use (1 %/1 mm), checked on hand-made columns."""
import numpy as np

from fastdose_amd import synth
from gpu_util import gamma_pass_rate


def _column(vol, f):
    X, Y, Z = (int(d) for d in vol.dim)
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    v = f(x, y, z).astype(np.float32).ravel()
    cols = np.nonzero(v)[0]
    return cols, v[cols]


def test_gamma_identity_scale_and_shift():
    vol = synth.make_volume((12, 10, 8), voxel=0.25)
    c, v = _column(vol, lambda x, y, z: np.exp(-((x - 6.0) ** 2 + (y - 5.0) ** 2) / 4.0) * (z < 6))
    assert gamma_pass_rate(vol, c, v, c, v) == 1.0
    assert gamma_pass_rate(vol, c, v, c, v * 1.009) == 1.0          # within 1 % of max everywhere
    assert gamma_pass_rate(vol, c, v, c, v * 1.05) < 1.0            # 5 % at the flat peak fails
    assert gamma_pass_rate(vol, c, v, c, v * 1.2) < 0.95
    # a one-voxel (2.5 mm > 1 mm) lateral shift of a steep profile fails
    c2, v2 = _column(vol, lambda x, y, z: np.exp(-((x - 7.0) ** 2 + (y - 5.0) ** 2) / 4.0) * (z < 6))
    assert gamma_pass_rate(vol, c, v, c2, v2) < 0.99
    # with a 3 mm distance criterion the same shift passes at the peak region
    assert gamma_pass_rate(vol, c, v, c2, v2, dd=0.01, dta=0.3) > gamma_pass_rate(vol, c, v, c2, v2)


def test_gamma_sub_voxel_distance():
    """A steep penumbra-like edge (12.5 % of max per mm) on a 2.5 mm grid: shifted by 0.5 mm it fails
    the 1 % dose criterion near the edge but passes gamma 1 %/1 mm through the distance criterion,
    which needs the sub-voxel search; shifted by 2 mm it fails."""
    vol = synth.make_volume((40, 6, 6), voxel=0.25)

    def edge(shift):
        return _column(vol, lambda x, y, z: 0.02 + 0.5 + 0.5 * np.tanh(((x + 0.5) * 0.25 - 5.0 - shift) / 0.4)
                       + 0 * y + 0 * z)

    c, v = edge(0.0)
    for s in (0.05, -0.05, 0.06):
        cs, vs = edge(s)
        assert np.abs(vs - v).max() > 0.05                            # the dose criterion alone fails
        assert gamma_pass_rate(vol, c, v, cs, vs) >= 0.99, s
    for s in (0.2, -0.2):
        cs, vs = edge(s)
        assert gamma_pass_rate(vol, c, v, cs, vs) < 0.95, s
    # a dose error of 2 % of max in the flat region fails whatever the distance
    c2, v2 = edge(0.0)
    v2 = v2 + 0.02 * (v2 > 0.9)
    assert gamma_pass_rate(vol, c, v, c2, v2) < 0.9
