"""fd_dense_to_csr_noalloc (no allocation by the library) against fd_dense_to_csr, through the C-ABI.

Bar: bit-exact offsets, columns and values (integer and index work), on two calls in a row with the
second smaller than the first (the shared device scratch keeps the first call's contents), plus the
two-phase capacity rule and the row limit, which the host rejects before any kernel runs.
"""
import ctypes as C

import numpy as np
import pytest

import fastdose_amd as fd
from fastdose_amd import _lib

pytestmark = pytest.mark.gpu

MAX_ROWS = 1 << 20  # FD_NOALLOC_MAX_ROWS


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _csr_call(name, torch, dense, capacity=None):
    """Two-phase call of fd_dense_to_csr(_noalloc) on a (rows, cols) float32 device tensor."""
    rows, cols = dense.shape
    dev = dense.device
    offsets = torch.full((rows + 1,), -7, dtype=torch.int64, device=dev)
    csr = _lib.CSR(0, 0, 0, offsets.data_ptr(), None, None)
    fn = getattr(fd.load(), name)
    rc = fn(C.c_void_p(dense.data_ptr()), rows, cols, C.byref(csr), None)
    nnz = int(csr.nnz)
    if nnz == 0:
        assert rc == _lib.FD_OK, fd.load().fd_last_error()
        torch.cuda.synchronize()
        return offsets.cpu().numpy(), np.zeros(0, np.int64), np.zeros(0, np.float32)
    assert rc == _lib.FD_ERR_CAPACITY, (name, rc)
    cap = nnz if capacity is None else capacity
    c = torch.full((max(cap, 1),), -1, dtype=torch.int64, device=dev)
    v = torch.full((max(cap, 1),), -1.0, dtype=torch.float32, device=dev)
    csr = _lib.CSR(0, 0, cap, offsets.data_ptr(), c.data_ptr(), v.data_ptr())
    rc = fn(C.c_void_p(dense.data_ptr()), rows, cols, C.byref(csr), None)
    if cap < nnz:
        assert rc == _lib.FD_ERR_CAPACITY and int(csr.nnz) == nnz
        torch.cuda.synchronize()
        return None
    assert rc == _lib.FD_OK, fd.load().fd_last_error()
    torch.cuda.synchronize()
    return offsets.cpu().numpy(), c[:nnz].cpu().numpy(), v[:nnz].cpu().numpy()


def _same(a, b):
    for x, y in zip(a, b):
        assert x.dtype == y.dtype and np.array_equal(x, y)


def _random_dense(rng, rows, cols, zero_frac):
    d = rng.uniform(0.1, 1.0, size=(rows, cols)).astype(np.float32)
    d[rng.uniform(size=(rows, cols)) < zero_frac] = 0.0
    return d


def test_noalloc_matches_allocating_entry_point(torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(7)
    # two calls in a row, the second smaller: the scratch still holds the first call's counts and
    # offsets, which must not leak into the second result
    for rows, cols, zf in [(3000, 2000, 0.6), (257, 999, 0.9), (1, 1, 0.0), (5, 0, 0.0), (0, 4, 0.0)]:
        print(f"[no-alloc] rows {rows} cols {cols}", flush=True)
        d = torch.from_numpy(_random_dense(rng, rows, cols, zf)).cuda()
        ref = _csr_call("fd_dense_to_csr", torch, d)
        got = _csr_call("fd_dense_to_csr_noalloc", torch, d)
        _same(ref, got)
        assert got[0][-1] == len(got[1])


def test_noalloc_known_matrix_and_capacity(torch_cuda):
    torch = torch_cuda
    # the 5 x 6 matrix of IMRTDebug.cu:405-409
    dense = np.array([[0, 0, 1, 0, 2, 0], [2, 0, 0, 1, 0, 0], [0, 0, 1, 2, 0, 1], [1, 0, 1, 0, 0, 0],
                      [1, 1, 0, 0, 0, 1]], dtype=np.float32)
    d = torch.from_numpy(dense).cuda()
    off, col, val = _csr_call("fd_dense_to_csr_noalloc", torch, d)
    assert off.tolist() == [0, 2, 4, 7, 9, 12]
    assert col.tolist() == [2, 4, 0, 3, 2, 3, 5, 0, 2, 0, 1, 5]
    assert val.tolist() == [1, 2, 2, 1, 1, 2, 1, 1, 1, 1, 1, 1]
    # capacity one short: FD_ERR_CAPACITY with nnz reported, then a full call still matches
    assert _csr_call("fd_dense_to_csr_noalloc", torch, d, capacity=11) is None
    off2, col2, val2 = _csr_call("fd_dense_to_csr_noalloc", torch, d)
    assert off2.tolist() == off.tolist() and col2.tolist() == col.tolist()


def test_noalloc_rejects_rows_beyond_the_limit(torch_cuda):
    torch = torch_cuda
    lib = fd.load()
    # rejected on the host before any kernel runs (no device memory is touched: dense is 1 x 1)
    one = torch.ones((1, 1), dtype=torch.float32, device="cuda:0")
    offsets = torch.zeros(2, dtype=torch.int64, device="cuda:0")
    csr = _lib.CSR(0, 0, 0, offsets.data_ptr(), None, None)
    rc = lib.fd_dense_to_csr_noalloc(C.c_void_p(one.data_ptr()), MAX_ROWS + 1, 1, C.byref(csr), None)
    assert rc == _lib.FD_ERR_ARG and b"FD_NOALLOC_MAX_ROWS" in lib.fd_last_error()
    rc = lib.fd_dense_to_csr_noalloc(C.c_void_p(one.data_ptr()), 1, 1, None, None)
    assert rc == _lib.FD_ERR_ARG


def test_noalloc_at_the_row_limit(torch_cuda):
    torch = torch_cuda
    # FD_NOALLOC_MAX_ROWS rows of 3 columns (12 MiB of input): the whole scratch in use
    rng = np.random.default_rng(11)
    d = torch.from_numpy(_random_dense(rng, MAX_ROWS, 3, 0.5)).cuda()
    print(f"[no-alloc] rows {MAX_ROWS} cols 3", flush=True)
    _same(_csr_call("fd_dense_to_csr", torch, d), _csr_call("fd_dense_to_csr_noalloc", torch, d))
    small = torch.from_numpy(_random_dense(rng, 100, 50, 0.5)).cuda()
    _same(_csr_call("fd_dense_to_csr", torch, small), _csr_call("fd_dense_to_csr_noalloc", torch, small))
