"""Pins the CPU oracle (oracle/) before anything is checked against it.

Each check is chosen so that a plausible slip in the restatement (wrong index,
dropped term, transposed operand, wrong sign) fails: golden vectors from the
reference's own host code compiled from its sources (tests/golden/ref_host.json),
closed forms, brute-force dense products and LAPACK solves on small systems.
"""
import json
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_host.json")


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def dense_of_bcsr(m, layout=0, values=None):
    v = m.values if values is None else values
    bs = m.bs
    A = np.zeros((m.n_rows * bs, m.n_cols * bs))
    for br in range(m.n_rows):
        for b in range(m.row_ptr[br], m.row_ptr[br + 1]):
            blk = v[b * bs * bs:(b + 1) * bs * bs].reshape(bs, bs)
            if layout == 1:
                blk = blk.T
            A[br * bs:(br + 1) * bs, m.columns[b] * bs:(m.columns[b] + 1) * bs] += blk
    return A


def dense_of_csr(c):
    A = np.zeros((c.n_rows, c.n_cols))
    for r in range(c.n_rows):
        for e in range(c.row_ptr[r], c.row_ptr[r + 1]):
            A[r, c.columns[e]] += c.values[e]
    return A


# ---- RNG ----------------------------------------------------------------------
def test_rng_known_answers():
    # (seed 0, stream 0) keys to 0, so indices 0 and 1 are the first two outputs
    # of the published SplitMix64 generator started from state 0.
    L = oracle.lib()
    assert L.orc_rng_u64(0, 0, 0) == 0xE220A8397B1DCDAF
    assert L.orc_rng_u64(0, 0, 1) == 0x6E789E6AA1B965F4
    u = oracle.rng_unit_vec(3, 4, 20000)
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01
    s = oracle.lib().orc_rng_sym(3, 4, 7)
    assert s == 2.0 * oracle.lib().orc_rng_unit(3, 4, 7) - 1.0
    # the vector fills are the scalar functions element by element
    idx = [0, 1, 7, 19999]
    assert [u[i] for i in idx] == [L.orc_rng_unit(3, 4, i) for i in idx]
    v = oracle.rng_sym_vec(3, 4, 20000)
    assert [v[i] for i in idx] == [L.orc_rng_sym(3, 4, i) for i in idx]


# ---- generators and conversions vs the reference's compiled host code ---------
def test_gen_n_diag_matches_reference(golden):
    for case in golden["matrices"]:
        m = oracle.gen_n_diag_bcsr(case["n_rows"], case["blocks_per_row"], case["bs"])
        assert m.nnzb == case["nnzb"]
        np.testing.assert_array_equal(m.row_ptr, case["bcsr_row_ptr"])
        np.testing.assert_array_equal(m.columns, case["bcsr_columns"])
        assert np.array_equal(m.values, np.array(case["bcsr_values"]))  # bit-exact


def test_bcsr_to_csr_matches_reference(golden):
    for case in golden["matrices"]:
        m = oracle.gen_n_diag_bcsr(case["n_rows"], case["blocks_per_row"], case["bs"])
        c = oracle.bcsr_to_csr(m)
        np.testing.assert_array_equal(c.row_ptr, case["csr_row_ptr"])
        np.testing.assert_array_equal(c.columns, case["csr_columns"])
        assert np.array_equal(c.values, np.array(case["csr_values"]))


def test_transpose_matches_reference_and_is_involution(golden):
    for case in golden["matrices"]:
        m = oracle.gen_n_diag_bcsr(case["n_rows"], case["blocks_per_row"], case["bs"])
        t = oracle.transpose_blocks(m)
        assert np.array_equal(t, np.array(case["bcsr_values_transposed"]))
        m2 = oracle.BCSR(m.n_rows, m.n_cols, m.bs, m.row_ptr, m.columns, t)
        assert np.array_equal(oracle.transpose_blocks(m2), m.values)


def test_block_lookup_matches_reference(golden):
    # get_block_data_by_column (lower_bound over the row's columns)
    for case in golden["matrices"]:
        m = oracle.gen_n_diag_bcsr(case["n_rows"], case["blocks_per_row"], case["bs"])
        bpr, bs = case["blocks_per_row"], case["bs"]
        got = []
        for row in range(m.n_rows):
            col = m.columns[m.row_ptr[row] + bpr // 2]
            k = int(np.searchsorted(m.columns[m.row_ptr[row]:m.row_ptr[row + 1]], col))
            got.append(m.values[(m.row_ptr[row] + k) * bs * bs])
        assert np.array_equal(np.array(got), np.array(case["block_lookup_mid"]))


def test_csr_bcsr_round_trip():
    for bs in (1, 2, 3, 5):
        m = oracle.gen_stencil9(5, 4, bs, 0, 11)
        c = oracle.bcsr_to_csr(m)
        m2 = oracle.csr_to_bcsr(c, bs)
        np.testing.assert_array_equal(m2.row_ptr, m.row_ptr)
        np.testing.assert_array_equal(m2.columns, m.columns)
        assert np.array_equal(m2.values, m.values)
        c2 = oracle.bcsr_to_csr(m2)
        assert np.array_equal(c2.values, c.values)
        np.testing.assert_array_equal(c2.columns, c.columns)


def test_csr_to_bcsr_fills_partial_blocks_with_zeros():
    # one nonzero at (1, 2) of a 4x4 matrix, bs = 2 -> one block (0, 1)
    c = oracle.CSR(4, 4, [0, 0, 1, 1, 1], [2], [7.0])
    m = oracle.csr_to_bcsr(c, 2)
    np.testing.assert_array_equal(m.row_ptr, [0, 1, 1])
    np.testing.assert_array_equal(m.columns, [1])
    np.testing.assert_array_equal(m.values, [0.0, 0.0, 7.0, 0.0])
    np.testing.assert_array_equal(dense_of_bcsr(m), dense_of_csr(c))


# ---- SpMV ---------------------------------------------------------------------
@pytest.mark.parametrize("bs", [1, 2, 3, 4, 6])
def test_csr_spmv_closed_form_x_ones(bs):
    # gen_n_diag_bcsr with x = 1: scalar row rho of a block row sums
    # (bs*col + bs^2*rho + bs(bs-1)/2)/N over its blocks.
    N, bpr = 11, 3
    m = oracle.gen_n_diag_bcsr(N, bpr, bs)
    c = oracle.bcsr_to_csr(m)
    y = oracle.csr_spmv(c, np.ones(c.n_cols))
    for br in range(N):
        cols = m.columns[m.row_ptr[br]:m.row_ptr[br + 1]]
        for rho in range(bs):
            want = sum((bs * col + bs * bs * rho + bs * (bs - 1) / 2.0) / N for col in cols)
            assert abs(y[br * bs + rho] - want) <= 1e-13 * max(1.0, abs(want))


def test_spmv_brute_force_dense():
    rng = np.random.default_rng(5)
    for bs in (1, 2, 3, 4):
        m = oracle.gen_stencil9(4, 3, bs, 0, 2)
        x = rng.uniform(-1, 1, m.n_cols * bs)
        A = dense_of_bcsr(m)
        want = A @ x
        y = oracle.bcsr_spmv(m, x)
        scale = np.abs(A) @ np.abs(x)
        assert np.all(np.abs(y - want) <= 1e-13 * scale)
        # column-major values give the same product
        t = oracle.transpose_blocks(m)
        np.testing.assert_array_equal(oracle.bcsr_spmv(m, x, layout=1, values=t), y)
        # CSR expansion: same summation order -> bit-identical
        c = oracle.bcsr_to_csr(m)
        np.testing.assert_array_equal(oracle.csr_spmv(c, x), y)
        np.testing.assert_allclose(oracle.bcsr_absmv(m, x), scale, rtol=1e-13)


def test_spmv_transposed_operand_is_detected():
    # using column-major indexing on row-major values must change the result
    m = oracle.gen_stencil9(3, 3, 3, 0, 4)
    x = oracle.rng_unit_vec(0, 4, m.n_cols * 3)
    assert not np.allclose(oracle.bcsr_spmv(m, x, layout=1), oracle.bcsr_spmv(m, x, layout=0))


# ---- stencil generator algebra -------------------------------------------------
@pytest.mark.parametrize("nx,ny", [(1, 1), (1, 5), (4, 1), (5, 4), (7, 7)])
def test_stencil_structure(nx, ny):
    m = oracle.gen_stencil9(nx, ny, 2, 0, 0)
    assert m.nnzb == (3 * nx - 2) * (3 * ny - 2)
    for i in range(nx * ny):
        iy, ix = divmod(i, nx)
        want = sorted(jy * nx + jx for jy in range(iy - 1, iy + 2) for jx in range(ix - 1, ix + 2)
                      if 0 <= jy < ny and 0 <= jx < nx)
        np.testing.assert_array_equal(m.columns[m.row_ptr[i]:m.row_ptr[i + 1]], want)


def test_stencil_kind0_diagonal_dominance():
    m = oracle.gen_stencil9(6, 5, 3, 0, 0)
    A = dense_of_bcsr(m)
    off = np.abs(A).sum(axis=1) - np.abs(np.diag(A))
    d = np.diag(A)
    # a_rr = u + (offrow + 1) with u in [-1, 1)
    assert np.all(d - (off + 1.0) >= -1.0) and np.all(d - (off + 1.0) < 1.0)
    assert np.all(d > off)
    assert not np.allclose(A, A.T)          # nonsymmetric
    # off-diagonal entries uniform in [-1, 1)
    offd = A[~np.eye(A.shape[0], dtype=bool)]
    assert offd.min() >= -1.0 and offd.max() < 1.0


def test_stencil_kind1_convective_algebra():
    bs = 4
    m = oracle.gen_stencil9(5, 4, bs, 1, 0)
    for br in range(m.n_rows):
        for b in range(m.row_ptr[br], m.row_ptr[br + 1]):
            blk = m.values[b * bs * bs:(b + 1) * bs * bs].reshape(bs, bs)
            if m.columns[b] == br:
                assert np.count_nonzero(blk - np.diag(np.diag(blk))) == 0
                continue
            K = -(blk + blk.T) / 2.0              # C cancels: K = -(A + A^T)/2
            C = -(blk - blk.T) / 0.6              # K cancels
            assert np.all(K >= -1e-15) and np.all(K <= 1.0 + 1e-15)
            assert np.allclose(np.diag(C), 0.0)
            assert np.all(np.abs(C) <= 1.0 + 1e-12)
    A = dense_of_bcsr(m)
    off = np.abs(A).sum(axis=1) - np.abs(np.diag(A))
    np.testing.assert_allclose(np.diag(A), 1.02 * off + 1.0, rtol=1e-14)


# ---- BiCGStab ------------------------------------------------------------------
def test_bicgstab_matches_lapack():
    m = oracle.gen_stencil9(6, 6, 2, 0, 3)
    c = oracle.bcsr_to_csr(m)
    A = dense_of_csr(c)
    b = oracle.rng_sym_vec(1, 5, c.n_rows)
    x, info = oracle.bicgstab(c, b, 1e-12, 500)
    assert info.status == 0
    xs = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-10
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-11


def test_bicgstab_manufactured_ones_and_jacobi():
    m = oracle.gen_stencil9(8, 6, 3, 1, 0)
    c = oracle.bcsr_to_csr(m)
    b = oracle.csr_spmv(c, np.ones(c.n_rows))
    for jac in (False, True):
        x, info = oracle.bicgstab(c, b, 1e-10, 2000, use_jacobi=jac)
        assert info.status == 0, (jac, info.status, info.iterations)
        assert np.max(np.abs(x - 1.0)) < 1e-7
        assert info.rel_residual < 1e-10


def test_bicgstab_identity_converges_in_one_iteration():
    n = 10
    c = oracle.CSR(n, n, np.arange(n + 1), np.arange(n), np.ones(n))
    b = oracle.rng_sym_vec(0, 6, n)
    x, info = oracle.bicgstab(c, b, 1e-12, 10)
    assert info.status == 0 and info.iterations == 1
    np.testing.assert_array_equal(x, b)


def test_bicgstab_breakdown_reported():
    # rotation: (r^, A p) = 0 on the first step
    c = oracle.CSR(2, 2, [0, 1, 2], [1, 0], [1.0, -1.0])
    x, info = oracle.bicgstab(c, np.array([1.0, 0.0]), 1e-12, 10)
    assert info.status == 2
    assert np.all(np.isfinite(x))


def test_bicgstab_max_iterations():
    m = oracle.gen_stencil9(10, 10, 2, 1, 0)
    c = oracle.bcsr_to_csr(m)
    b = oracle.csr_spmv(c, np.ones(c.n_rows))
    x, info = oracle.bicgstab(c, b, 1e-14, 3)
    assert info.status == 1 and info.iterations == 3


def test_bicgstab_odd_iteration_returns_current_iterate():
    # stops after an odd count: the returned x must be the current iterate
    m = oracle.gen_stencil9(6, 6, 2, 0, 3)
    c = oracle.bcsr_to_csr(m)
    b = oracle.rng_sym_vec(1, 5, c.n_rows)
    for maxit in (1, 2, 3):
        x, info = oracle.bicgstab(c, b, 1e-30, maxit)
        r = b - oracle.csr_spmv(c, x)
        assert abs(np.linalg.norm(r) / np.linalg.norm(b) - info.rel_residual) < 1e-6


# ---- measurement arithmetic ----------------------------------------------------
def test_compare_metric_matches_reference_text(golden):
    r1 = np.array([1.0, 2.0, 3.0, 4.0])
    bad = np.array([1.0, 2.5, 3.0, 4.0])
    err = oracle.compare_metric(bad, r1)
    assert err == pytest.approx(0.25 / 30.0, rel=1e-15)
    # the reference prints "ERROR: <err with 6 significant digits>" above 1e-9
    assert golden["compare"]["bad"].startswith("ERROR: %g" % err)
    assert golden["compare"]["same"] == "" and golden["compare"]["tiny"] == ""
    assert oracle.compare_metric(np.array([1.0, 2.0 + 1e-6, 3.0, 4.0]), r1) < 1e-9
