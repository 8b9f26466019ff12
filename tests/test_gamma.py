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
    assert gamma_pass_rate(vol, c, v, c, v * 1.05) < 0.99           # 5 % near the peak fails
    # a one-voxel (2.5 mm > 1 mm) lateral shift of a steep profile fails
    c2, v2 = _column(vol, lambda x, y, z: np.exp(-((x - 7.0) ** 2 + (y - 5.0) ** 2) / 4.0) * (z < 6))
    assert gamma_pass_rate(vol, c, v, c2, v2) < 0.99
    # with a 3 mm distance criterion the same shift passes at the peak region
    assert gamma_pass_rate(vol, c, v, c2, v2, dd=0.01, dta=0.3) > gamma_pass_rate(vol, c, v, c2, v2)
