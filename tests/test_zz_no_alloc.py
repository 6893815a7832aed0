"""The allocation-free CSR SpMV (bmfp_csr_spmv_noalloc) through the C ABI.

It runs the merge-path kernels of bmfp_spmv(A, BMFP_CSR_MERGE, ...) over
caller-owned device arrays with the scratch in the library's fixed device
array, so its result must equal the handle-based merge SpMV bit for bit; both
are also checked against the oracle's CSR product (FP64: per row
|y - y_oracle| <= 1e-12 (|A||x|)_i).  Calls in a row share the scratch, so a
larger call followed by a smaller one checks that nothing of the first leaks
into the second.
"""
import ctypes

import numpy as np
import pytest

import block_matrix_format_performance_amd as bm
import oracle

F64_TOL = 1e-12
U32 = 2.0 ** -24


def csr_case(lengths, n_cols, seed):
    """CSR with the given row lengths (sorted random columns, values
    uniform[-1,1]) and x uniform[0,1]."""
    rng = np.random.default_rng(seed)
    lengths = np.asarray(lengths, np.int64)
    rp = np.concatenate([[0], np.cumsum(lengths)])
    cols = np.concatenate([np.sort(rng.choice(n_cols, ln, replace=ln > n_cols)) if ln else np.zeros(0, np.int64)
                           for ln in lengths]).astype(np.int32)
    return oracle.CSR(len(lengths), n_cols, rp, cols, rng.uniform(-1, 1, int(rp[-1]))), rng.uniform(0, 1, n_cols)


def noalloc_spmv(c, x, dtype):
    """y = A x by bmfp_csr_spmv_noalloc on device copies of c's arrays."""
    vt = np.float64 if dtype == bm.F64 else np.float32
    rp = bm.DeviceArray.from_host(c.row_ptr)
    cols = bm.DeviceArray.from_host(c.columns if c.nnz else np.zeros(1, np.int32))
    vals = bm.DeviceArray.from_host(c.values.astype(vt) if c.nnz else np.zeros(1, vt))
    dx = bm.DeviceArray.from_host(x.astype(vt))
    dy = bm.DeviceArray(c.n_rows, vt)
    dy.upload(np.full(c.n_rows, np.nan, vt))  # every row must be written
    bm.check(bm.lib().bmfp_csr_spmv_noalloc(dtype, c.n_rows, c.n_cols, c.nnz, rp.ptr, cols.ptr, vals.ptr,
                                            dx.ptr, dy.ptr, None))
    bm.synchronize()
    return dy.download()


def handle_merge_spmv(c, x, dtype):
    A = bm.Matrix.csr(c.row_ptr, c.columns, c.values, c.n_rows, c.n_cols, dtype=dtype)
    y = bm.spmv_host(A, x, bm.CSR_MERGE)
    A.destroy()
    return y


def check_against_oracle(c, x, y, dtype):
    if dtype == bm.F64:
        want, tol = oracle.csr_spmv(c, x), F64_TOL * oracle.csr_absmv(c, x)
    else:
        c32 = oracle.CSR(c.n_rows, c.n_cols, c.row_ptr, c.columns, c.values.astype(np.float32).astype(np.float64))
        x32 = x.astype(np.float32).astype(np.float64)
        want, tol = oracle.csr_spmv(c32, x32), (np.diff(c.row_ptr) + 2) * U32 * oracle.csr_absmv(c32, x32)
    bad = np.nonzero(np.abs(y.astype(np.float64) - want) > tol)[0]
    assert bad.size == 0, bad[:5]


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [bm.F64, bm.F32])
def test_noalloc_equals_handle_merge_on_two_calls(device, dtype):
    rng = np.random.default_rng(7)
    # first call: ~100k path items over ~200 tiles, a row over 40 tiles,
    # empty runs; second call: a smaller matrix, so the scratch still holds
    # the first call's carries and boundaries beyond its own
    big = np.concatenate([rng.integers(0, 40, 3000), [20000], np.zeros(700, np.int64), rng.integers(0, 9, 2000)])
    small = np.concatenate([rng.integers(0, 30, 400), [1500], np.zeros(90, np.int64), [3]])
    for k, lengths in enumerate((big, small)):
        print(f"[gpu-case] noalloc dtype={dtype} call={k} rows={len(lengths)} nnz={int(np.sum(lengths))}",
              flush=True)
        c, x = csr_case(lengths, 5000, seed=11 + k)
        y = noalloc_spmv(c, x, dtype)
        want = handle_merge_spmv(c, x, dtype)
        assert np.array_equal(y, want), np.nonzero(y != want)[0][:5]
        check_against_oracle(c, x, y, dtype)


@pytest.mark.gpu
def test_noalloc_stencil_expansion_and_empty_rows(device):
    m = oracle.gen_stencil9(20, 17, 4, 1, 3)
    c = oracle.bcsr_to_csr(m)
    x = oracle.rng_unit_vec(3, 4, c.n_cols)
    y = noalloc_spmv(c, x, bm.F64)
    assert np.array_equal(y, handle_merge_spmv(c, x, bm.F64))
    check_against_oracle(c, x, y, bm.F64)
    e = oracle.CSR(1000, 10, np.zeros(1001, np.int32), np.zeros(0, np.int32), np.zeros(0))
    assert np.array_equal(noalloc_spmv(e, np.ones(10), bm.F64), np.zeros(1000))


def test_noalloc_rejects_sizes_beyond_the_scratch():
    # refused by the size check, before any device call: the pointers are
    # never dereferenced
    L = bm.lib()
    fake = ctypes.c_void_p(256)
    st = L.bmfp_csr_spmv_noalloc(bm.F64, 1 << 27, 1 << 27, (1 << 28) - (1 << 27) + 1, fake, fake, fake, fake,
                                 fake, None)
    assert st == bm.ERR_INVALID
    assert b"allocation-free limit" in L.bmfp_last_error()
    assert L.bmfp_csr_spmv_noalloc(bm.F16, 10, 10, 10, fake, fake, fake, fake, fake, None) == bm.ERR_UNSUPPORTED
    assert L.bmfp_csr_spmv_noalloc(bm.F64, 10, 10, 10, None, fake, fake, fake, fake, None) == bm.ERR_INVALID
