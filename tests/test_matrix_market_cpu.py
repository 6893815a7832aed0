"""Matrix Market I/O and BCSR block-size detection (bmfp_mm_read, bmfp_mm_write,
bmfp_block_size_detect), CPU only.

The reader is pinned against scipy.io (an independent Matrix Market
implementation) on every supported flavour; the writer's text is pinned
byte for byte against the reference's own write_mm in test_cpp_host.py.
Block-size detection must recover the block size of matrices built from
known blocks."""
import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

import block_matrix_format_performance_amd as bm
import oracle


def scipy_csr(path):
    a = scipy.io.mmread(path)
    c = sp.csr_matrix(a)
    c.sum_duplicates()
    c.sort_indices()
    return c


def assert_same_csr(mine, c):
    rp, cols, vals, n, m = mine
    assert (n, m) == c.shape
    np.testing.assert_array_equal(rp, c.indptr)
    np.testing.assert_array_equal(cols, c.indices)
    assert np.array_equal(vals, c.data.astype(np.float64))


def random_sparse(n, m, density, seed, integer=False):
    rng = np.random.default_rng(seed)
    a = sp.random(n, m, density=density, format="coo", random_state=seed,
                  data_rvs=(lambda k: rng.integers(-9, 10, k).astype(float)) if integer else None)
    return a


@pytest.mark.parametrize("symmetry", ["general", "symmetric", "skew-symmetric"])
@pytest.mark.parametrize("field", ["real", "integer", "pattern"])
def test_reader_matches_scipy(tmp_path, symmetry, field):
    if symmetry == "skew-symmetric" and field == "pattern":
        pytest.skip("pattern matrices cannot be skew-symmetric")
    a = random_sparse(60, 60, 0.08, seed=len(symmetry) + len(field), integer=field == "integer")
    if symmetry == "symmetric":
        a = (a + a.T).tocoo()
    elif symmetry == "skew-symmetric":
        a = (a - a.T).tocoo()
    path = tmp_path / "m.mtx"
    scipy.io.mmwrite(str(path), a, field=field, symmetry=symmetry)
    assert_same_csr(bm.read_mm(path), scipy_csr(path))


def test_reader_rectangular_and_dense_array(tmp_path):
    a = random_sparse(37, 91, 0.1, seed=3)
    path = tmp_path / "r.mtx"
    scipy.io.mmwrite(str(path), a)
    assert_same_csr(bm.read_mm(path), scipy_csr(path))
    d = a.toarray()
    scipy.io.mmwrite(str(tmp_path / "d.mtx"), d)              # "array" format
    assert_same_csr(bm.read_mm(tmp_path / "d.mtx"), sp.csr_matrix(d))


def test_reader_comments_duplicates_and_order(tmp_path):
    path = tmp_path / "h.mtx"
    path.write_text("%%MatrixMarket MATRIX Coordinate Integer General\n"
                    "% a comment\n"
                    "\n"
                    "3 4 6\n"
                    "3 4 7\n"
                    "1 2 5\n"
                    "   \n"
                    "1 2 -2\n"
                    "2 1 1\n"
                    "% another comment\n"
                    "1 1 3\n"
                    "3 1 4\n")
    rp, cols, vals, n, m = bm.read_mm(path)
    assert (n, m) == (3, 4)
    np.testing.assert_array_equal(rp, [0, 2, 3, 5])
    np.testing.assert_array_equal(cols, [0, 1, 0, 0, 3])
    np.testing.assert_array_equal(vals, [3.0, 3.0, 1.0, 4.0, 7.0])  # (1,2) summed: 5 - 2


def test_empty_and_all_zero_matrices(tmp_path):
    path = tmp_path / "e.mtx"
    path.write_text("%%MatrixMarket matrix coordinate real general\n5 3 0\n")
    rp, cols, vals, n, m = bm.read_mm(path)
    assert (n, m, len(cols)) == (5, 3, 0) and np.all(rp == 0)


@pytest.mark.parametrize("text,status", [
    ("5 5 1\n1 1 1.0\n", bm.ERR_INVALID),                                          # no banner
    ("%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n", bm.ERR_UNSUPPORTED),
    ("%%MatrixMarket matrix coordinate real hermitian\n1 1 1\n1 1 1\n", bm.ERR_UNSUPPORTED),
    ("%%MatrixMarket vector coordinate real general\n1 1 1\n1 1 1\n", bm.ERR_UNSUPPORTED),
    ("%%MatrixMarket matrix coordinate real general\n3 3 2\n1 1 1.0\n", bm.ERR_INVALID),   # truncated
    ("%%MatrixMarket matrix coordinate real general\n3 3 1\n4 1 1.0\n", bm.ERR_INVALID),   # row 4 of 3
    ("%%MatrixMarket matrix coordinate real general\n3 3 1\n1 0 1.0\n", bm.ERR_INVALID),   # column 0
    ("%%MatrixMarket matrix coordinate real general\n3 3 1\n1 1 x\n", bm.ERR_INVALID),
    ("%%MatrixMarket matrix coordinate real symmetric\n3 4 1\n1 1 1.0\n", bm.ERR_INVALID),
    ("%%MatrixMarket matrix array real symmetric\n2 2\n1\n2\n3\n", bm.ERR_UNSUPPORTED),
    # size lines declaring far more entries than the file holds: rejected
    # without reserving memory for the declared count
    ("%%MatrixMarket matrix coordinate real general\n3 3 2000000000\n1 1 1.0\n", bm.ERR_INVALID),
    ("%%MatrixMarket matrix coordinate real symmetric\n3 3 9000000000000\n1 1 1.0\n", bm.ERR_INVALID),
])
def test_reader_errors(tmp_path, text, status):
    path = tmp_path / "bad.mtx"
    path.write_text(text)
    with pytest.raises(bm.BmfpError) as e:
        bm.read_mm(path)
    assert e.value.status == status


def test_missing_file():
    with pytest.raises(bm.BmfpError) as e:
        bm.read_mm("/nonexistent/dir/m.mtx")
    assert e.value.status == bm.ERR_INVALID


def test_writer_round_trips_through_scipy(tmp_path):
    a = sp.csr_matrix(random_sparse(50, 70, 0.1, seed=9))
    a.data = np.random.default_rng(1).standard_normal(a.nnz) * 1e-7   # awkward decimal expansions
    path = tmp_path / "w.mtx"
    bm.write_mm(path, a.indptr, a.indices, a.data, 50, 70)
    back = scipy_csr(path)
    np.testing.assert_array_equal(back.indptr, a.indptr)
    np.testing.assert_array_equal(back.indices, a.indices)
    assert np.array_equal(back.data, a.data)                     # %.20g is exact for doubles
    assert_same_csr(bm.read_mm(path), a)


@pytest.mark.parametrize("bs", [1, 2, 3, 4, 6, 8])
def test_block_size_detection_recovers_stencil_blocks(bs):
    m = oracle.gen_stencil9(9, 8, bs, 0, bs)
    c = oracle.bcsr_to_csr(m)
    got_bs, nnzb, fill = bm.detect_block_size(c.row_ptr, c.columns, c.n_rows, c.n_cols)
    assert got_bs == bs and nnzb == m.nnzb and fill == 1.0


def test_block_size_detection_on_a_random_pattern_brute_force():
    # sparse random pattern: blocking only adds explicit zeros, bs = 1 wins;
    # nnzb of each candidate checked against a brute-force count
    a = sp.csr_matrix(random_sparse(96, 96, 0.03, seed=5))
    a.sort_indices()
    for cand in (1, 2, 3, 4, 6, 8):
        r, c = a.nonzero()
        want = len(set(zip(r // cand, c // cand)))
        got_bs, nnzb, fill = bm.detect_block_size(a.indptr, a.indices, 96, 96, candidates=[cand])
        assert (got_bs, nnzb) == (cand, want)
        assert fill == pytest.approx(want * cand * cand / a.nnz)
    assert bm.detect_block_size(a.indptr, a.indices, 96, 96)[0] == 1


def test_block_size_detection_skips_non_dividing_sizes():
    m = oracle.gen_stencil9(5, 5, 3, 0, 1)
    c = oracle.bcsr_to_csr(m)          # 75 x 75: only 1 and 3 of the defaults divide it
    assert bm.detect_block_size(c.row_ptr, c.columns, 75, 75)[0] == 3
    with pytest.raises(bm.BmfpError):
        bm.detect_block_size(c.row_ptr, c.columns, 75, 75, candidates=[2, 4])


def test_reader_survives_mutated_files(tmp_path):
    # seeded mutations of valid files (deleted, duplicated and truncated lines,
    # garbage tokens): every read either fails with a status or returns a
    # well-formed CSR (row pointers from 0 to nnz, nondecreasing, columns in
    # range and strictly ascending within a row)
    rng = np.random.default_rng(2024)
    junk = ["-1", "0", "2147483647", "99999999999", "1e309", "nan", "x", " ", "%", "3 3", "1 1 1 1", "\n"]
    path = tmp_path / "f.mtx"
    parsed = 0
    for case in range(600):
        field = ["real", "integer", "pattern"][case % 3]
        sym = ["general", "symmetric", "skew-symmetric"][(case // 3) % 3]
        if sym == "skew-symmetric" and field == "pattern":
            sym = "general"
        n = int(rng.integers(1, 9))
        lines = [f"%%MatrixMarket matrix coordinate {field} {sym}", f"{n} {n} 6"]
        for _ in range(6):
            i, j = sorted(rng.integers(1, n + 1, 2).tolist(), reverse=sym != "general")
            lines.append(f"{i} {j}" + ("" if field == "pattern" else f" {rng.uniform(-5, 5):.6g}"))
        text = "\n".join(lines) + "\n"
        for _ in range(int(rng.integers(0, 4))):
            p = int(rng.integers(0, len(text)))
            op = int(rng.integers(0, 4))
            if op == 0:
                text = text[:p] + text[p + int(rng.integers(1, 6)):]
            elif op == 1:
                text = text[:p] + junk[int(rng.integers(0, len(junk)))] + text[p:]
            elif op == 2:
                text = text[:p]
            else:
                q = text.find("\n", p)
                text = text if q < 0 else text[:q + 1] + text[p:q + 1] + text[q + 1:]
        path.write_text(text)
        try:
            rp, cols, vals, nr, nc = bm.read_mm(path)
        except bm.BmfpError as e:
            assert e.status in (bm.ERR_INVALID, bm.ERR_UNSUPPORTED)
            continue
        parsed += 1
        assert rp[0] == 0 and rp[-1] == len(cols) == len(vals) and len(rp) == nr + 1
        assert np.all(np.diff(rp) >= 0)
        assert np.all((cols >= 0) & (cols < nc))
        for r in range(nr):
            assert np.all(np.diff(cols[rp[r]:rp[r + 1]]) > 0)
    assert parsed > 100
