"""fastdose_imrt: the command-line drop-in for `IMRT --mode 0` (IMRTOpt/main.cu:33-84)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import fastdose_amd as fd
from fastdose_amd import synth

ROOT = Path(__file__).resolve().parent.parent
CLI = ROOT / "fastdose_amd" / "fastdose_imrt"


def run(args, **kw):
    return subprocess.run([str(CLI), *args], capture_output=True, text=True, timeout=300, **kw)


def test_help_and_flag_errors():
    r = run(["--help"])
    assert r.returncode == 0
    for name in ("phantomDim", "primaryROI", "subFluenceOn", "EstNonZeroElementsPerMat", "outputFolder"):
        assert f"--{name}" in r.stdout
    r = run(["--SAD", "100"])
    assert r.returncode == 1 and "required" in r.stderr
    r = run(["--bogus", "1"])
    assert r.returncode == 1 and "unknown option" in r.stderr


def write_case(tmp: Path, golden: Path):
    dim = (32, 30, 28)
    rng = np.random.default_rng(4)
    density = rng.uniform(0.5, 1.5, dim[::-1]).astype(np.float32)
    z, y, x = np.meshgrid(np.arange(dim[2]), np.arange(dim[1]), np.arange(dim[0]), indexing="ij")
    ptv = (((x - 15.5) / 4) ** 2 + ((y - 14.5) / 3) ** 2 + ((z - 13.5) / 3) ** 2 <= 1).astype(np.uint8)
    body = np.zeros(dim[::-1], np.uint8)
    body[2:-2, 2:-2, 2:-2] = 1
    (tmp / "masks").mkdir()
    ptv.tofile(tmp / "masks" / "PTV.bin")
    body.tofile(tmp / "masks" / "BODY.bin")
    density.tofile(tmp / "density.raw")
    (tmp / "beamlist.txt").write_text("0 0 0\n90 0 0\n30 40 10\n")
    args = ["--phantomDim", *map(str, dim), "--voxelSize", "0.25", "0.25", "0.25", "--SAD", "100",
            "--density", str(tmp / "density.raw"), "--structures", "BODY", "PTV", "--masks", str(tmp / "masks"),
            "--primaryROI", "PTV", "--bboxROI", "BODY", "--beamlist", str(tmp / "beamlist.txt"), "--mode", "0",
            "--deviceIdx", "0", "--spectrum", str(golden / "spec_6mv.spec"), "--kernel",
            str(golden / "kernel_exp_6mv.txt"), "--fluenceDim", "6", "--outputFolder", str(tmp / "out")]
    return dim, density, ptv, body, args


def test_missing_mask_file(tmp_path, golden):
    _, _, _, _, args = write_case(tmp_path, golden)
    (tmp_path / "masks" / "PTV.bin").unlink()
    r = run(args)
    assert r.returncode == 1 and "PTV.bin" in r.stderr


@pytest.mark.gpu
def test_cli_matches_library(tmp_path, golden):
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    dim, density, ptv, body, args = write_case(tmp_path, golden)
    r = run(args)
    assert r.returncode == 0, r.stderr + r.stdout
    folder = tmp_path / "out" / "doseMatFolder"
    mats = fd.read_dosemat_folder(folder)
    assert len(mats) == 3
    # one beam at a time gives the same files
    args1 = list(args)
    args1[args1.index("--outputFolder") + 1] = str(tmp_path / "out1")
    r = run([*args1, "--beamletBatch", "0"])
    assert r.returncode == 0, r.stderr + r.stdout
    for name in ("offsetsBuffer", "columnsBuffer", "valuesBuffer", "NonZeroElements", "numRowsPerMat"):
        a = (folder / f"{name}.bin").read_bytes()
        b = (tmp_path / "out1" / "doseMatFolder" / f"{name}.bin").read_bytes()
        assert a == b, name
    # same pipeline through the Python host API
    zz, yy, xx = np.nonzero(body)
    vol = synth.make_volume(dim, bbox_start=(xx.min(), yy.min(), zz.min()),
                            bbox_dim=(xx.max() - xx.min() + 1, yy.max() - yy.min() + 1, zz.max() - zz.min() + 1))
    iso = fd.ptv_isocenter(ptv, vol)
    bundles = [synth.make_bundle(iso, ang, fluence_dim=6) for ang in ((0, 0, 0), (90, 0, 0), (30, 40, 10))]
    flags = [fd.beamlet_flag_init(b, ptv, vol) for b in bundles]
    fmap = np.fromfile(folder / "fluenceMap.bin", dtype=np.uint8).reshape(3, 36)
    assert np.array_equal(fmap, np.stack(flags))
    spec = fd.read_spectrum(golden / "spec_6mv.spec")
    kern = fd.read_kernel(golden / "kernel_exp_6mv.txt", 8)
    with fd.Context(vol, density, spec, kern) as ctx:
        dm = fd.dose_mat_construction(ctx, bundles, flags, 2.0)
        for k in range(3):
            off, col, val = dm.matrix(k)
            assert np.array_equal(mats[k][0], off)
            assert np.array_equal(mats[k][1], col)
            assert np.array_equal(mats[k][2], val)
