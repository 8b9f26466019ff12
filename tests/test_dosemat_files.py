"""doseMatFolder files on the host (no GPU): the checks fd_dosemat_fromfile makes before it touches
the device, and the folder merge of scripts/doseMerge.py:9-63.  Folders are written here with numpy
in the layout MatCSREnsemble::tofile (IMRTSparseMatEns.cu:168-239) produces."""
import ctypes as C
import shutil

import numpy as np
import pytest

import fastdose_amd as fd
from fastdose_amd import _lib

U64 = np.uint64


def _write_folder(path, mats, fluence=None):
    """mats: list of (offsets, columns, values) per beam, offsets local (start at 0)."""
    path.mkdir(parents=True, exist_ok=True)
    nnz = np.array([len(m[1]) for m in mats], U64)
    rows = np.array([len(m[0]) - 1 for m in mats], U64)
    oidx = np.concatenate([[0], np.cumsum(rows + 1)[:-1]]).astype(U64) if mats else np.zeros(0, U64)
    nnz.tofile(path / "NonZeroElements.bin")
    np.cumsum(nnz).astype(U64).tofile(path / "CumuNonZeroElements.bin")
    rows.tofile(path / "numRowsPerMat.bin")
    np.cumsum(rows).astype(U64).tofile(path / "CumuNumRowsPerMat.bin")
    oidx.tofile(path / "OffsetBufferIdx.bin")
    np.concatenate([m[0] for m in mats] or [np.zeros(0)]).astype(U64).tofile(path / "offsetsBuffer.bin")
    np.concatenate([m[1] for m in mats] or [np.zeros(0)]).astype(U64).tofile(path / "columnsBuffer.bin")
    np.concatenate([m[2] for m in mats] or [np.zeros(0)]).astype(np.float32).tofile(path / "valuesBuffer.bin")
    if fluence is not None:
        np.asarray(fluence, np.uint8).tofile(path / "fluenceMap.bin")


def _random_mats(rng, n_beams, n_cols):
    mats = []
    for _ in range(n_beams):
        rows = int(rng.integers(0, 6))
        counts = rng.integers(0, 9, rows)
        off = np.concatenate([[0], np.cumsum(counts)]).astype(U64)
        cols = np.concatenate([np.sort(rng.choice(n_cols, c, replace=False)) for c in counts] or
                              [np.zeros(0)]).astype(U64)
        vals = rng.uniform(0.1, 2.0, int(off[-1])).astype(np.float32)
        mats.append((off, cols, vals))
    return mats


def test_merge_is_the_concatenation(tmp_path):
    rng = np.random.default_rng(1)
    folders, all_mats, maps = [], [], []
    for k, nb in enumerate((3, 1, 4)):
        mats = _random_mats(rng, nb, 500)
        flu = rng.integers(0, 2, nb * 9).astype(np.uint8)
        _write_folder(tmp_path / f"doseMat{k}" / "doseMatFolder", mats, flu)
        folders.append(tmp_path / f"doseMat{k}" / "doseMatFolder")
        all_mats.extend(mats)
        maps.append(flu)
    out = tmp_path / "doseMatMerge" / "doseMatFolder"
    out.parent.mkdir()
    fd.merge_dosemat_folders(folders, out)
    # doseMerge.py:46-62: each of these files is the byte concatenation of the inputs'
    for name in ("NonZeroElements", "numRowsPerMat", "offsetsBuffer", "columnsBuffer", "valuesBuffer", "fluenceMap"):
        want = b"".join((f / f"{name}.bin").read_bytes() for f in folders)
        assert (out / f"{name}.bin").read_bytes() == want, name
    # and the index files the script leaves out describe the merged set
    rows = np.fromfile(out / "numRowsPerMat.bin", U64)
    nnz = np.fromfile(out / "NonZeroElements.bin", U64)
    assert np.array_equal(np.fromfile(out / "CumuNumRowsPerMat.bin", U64), np.cumsum(rows))
    assert np.array_equal(np.fromfile(out / "CumuNonZeroElements.bin", U64), np.cumsum(nnz))
    assert np.array_equal(np.fromfile(out / "OffsetBufferIdx.bin", U64), np.concatenate([[0], np.cumsum(rows + 1)[:-1]]))
    # readDose.m semantics over the merged folder give every input matrix, in order
    got = fd.read_dosemat_folder(out)
    assert len(got) == len(all_mats)
    for a, b in zip(got, all_mats):
        assert all(np.array_equal(x, y) for x, y in zip(a, b))


def _corruptions(tmp_path, good):
    """Folders that each break one rule fd_dosemat_fromfile checks."""
    cases = {}

    def variant(name, edit):
        d = tmp_path / name
        shutil.copytree(good, d)
        edit(d)
        cases[name] = d

    def set_u64(d, fname, i, v):
        a = np.fromfile(d / fname, U64)
        a[i] = v
        a.tofile(d / fname)

    variant("column_beyond_volume", lambda d: set_u64(d, "columnsBuffer.bin", 3, 10 ** 6))
    variant("truncated_values", lambda d: (d / "valuesBuffer.bin").write_bytes((d / "valuesBuffer.bin").read_bytes()[:-4]))
    variant("truncated_offsets", lambda d: (d / "offsetsBuffer.bin").write_bytes((d / "offsetsBuffer.bin").read_bytes()[:-8]))
    variant("offset_index_past_end", lambda d: set_u64(d, "OffsetBufferIdx.bin", 1, 10 ** 4))
    variant("nonzeros_too_many", lambda d: set_u64(d, "NonZeroElements.bin", 0, 10 ** 5))
    variant("rows_too_many", lambda d: set_u64(d, "numRowsPerMat.bin", 0, 10 ** 5))
    variant("offsets_decrease", lambda d: set_u64(d, "offsetsBuffer.bin", 2, 0))
    variant("offsets_not_from_zero", lambda d: set_u64(d, "offsetsBuffer.bin", 0, 1))
    variant("table_sizes_differ", lambda d: (d / "numRowsPerMat.bin").write_bytes((d / "numRowsPerMat.bin").read_bytes()[:-8]))
    variant("missing_file", lambda d: (d / "columnsBuffer.bin").unlink())
    return cases


def test_fromfile_rejects_bad_folders_before_any_device_work(tmp_path):
    """Every corrupted or mismatched folder gives FD_ERR_IO and leaves the ensemble untouched; the
    checks come before any device allocation or copy, which is why this runs without a GPU."""
    rng = np.random.default_rng(2)
    mats = _random_mats(rng, 3, 500)
    while sum(len(m[1]) for m in mats) < 8 or any(len(m[0]) < 3 for m in mats[:2]):
        mats = _random_mats(rng, 3, 500)
    good = tmp_path / "good"
    _write_folder(good, mats)
    dm = fd.DoseMat(500)
    for name, folder in _corruptions(tmp_path, good).items():
        with pytest.raises(fd.FastDoseError) as e:
            dm.fromfile(folder)
        assert e.value.code == _lib.FD_ERR_IO, name
        assert dm.info() == (0, 0, 0), name
    # the good folder read with too few voxels: a column beyond n_cols
    small = fd.DoseMat(int(max(int(m[1].max()) for m in mats if len(m[1]))))
    with pytest.raises(fd.FastDoseError) as e:
        small.fromfile(good)
    assert e.value.code == _lib.FD_ERR_IO and b"beyond" in fd.load().fd_last_error()


def test_merge_checks_inputs_and_writes_nothing_on_error(tmp_path):
    rng = np.random.default_rng(3)
    good = tmp_path / "good"
    _write_folder(good, _random_mats(rng, 2, 100), np.ones(18, np.uint8))
    bad = tmp_path / "bad"
    _write_folder(bad, _random_mats(rng, 2, 100), np.ones(18, np.uint8))
    (bad / "valuesBuffer.bin").write_bytes(b"\0\0")
    out = tmp_path / "merged"
    with pytest.raises(fd.FastDoseError) as e:
        fd.merge_dosemat_folders([good, bad], out)
    assert e.value.code == _lib.FD_ERR_IO and not out.exists()
    nomap = tmp_path / "nomap"
    _write_folder(nomap, _random_mats(rng, 1, 100))
    with pytest.raises(fd.FastDoseError):
        fd.merge_dosemat_folders([good, nomap], out)   # fluenceMap.bin in one input only
    assert not out.exists()
    with pytest.raises(fd.FastDoseError) as e:
        fd.merge_dosemat_folders([], out)
    assert e.value.code == _lib.FD_ERR_ARG
