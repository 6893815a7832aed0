"""GPU parity of every SpMV kernel family against the CPU oracle.

Bar (FP64): per row |y_gpu - y_oracle| <= 1e-12 * (|A||x|)_i.
FP32 / FP16+scale storage: the oracle runs on the device's own rounded values
(downloaded) and x rounded to FP32, so only accumulation error remains:
|y_gpu - y_oracle| <= (row_len + 2) * 2^-24 * (|A||x|)_i.
"""
import numpy as np
import pytest

import block_matrix_format_performance_amd as bm
import oracle

pytestmark = pytest.mark.gpu

F64_TOL = 1e-12
U32 = 2.0 ** -24


@pytest.fixture(autouse=True)
def _name_the_case(request, device):
    print(f"\n[gpu-case] {request.node.nodeid}", flush=True)
    yield


def stencil_case(nx, ny, bs, kind=0, seed=0):
    m = oracle.gen_stencil9(nx, ny, bs, kind, seed)
    x = oracle.rng_unit_vec(seed, 4, m.n_cols * bs)
    return m, x


def ragged_case(n_rows, n_cols, bs, max_len, seed, empty_every=5):
    """Random BCSR with ragged rows (0..max_len blocks), sorted columns."""
    rng = np.random.default_rng(seed)
    rp = [0]
    cols = []
    for r in range(n_rows):
        ln = 0 if (empty_every and r % empty_every == 3) else int(rng.integers(1, max_len + 1))
        ln = min(ln, n_cols)
        cols.extend(sorted(rng.choice(n_cols, ln, replace=False).tolist()))
        rp.append(len(cols))
    nnzb = len(cols)
    vals = rng.uniform(-1, 1, nnzb * bs * bs)
    m = oracle.BCSR(n_rows, n_cols, bs, np.array(rp), np.array(cols, dtype=np.int32), vals)
    x = rng.uniform(0, 1, n_cols * bs)
    return m, x


def row_lengths(m):
    return np.repeat(np.diff(m.row_ptr) * m.bs, m.bs)


def check_bcsr(m, x, layout, variant, dtype=bm.F64):
    vals = m.values if layout == bm.ROW_MAJOR else oracle.transpose_blocks(m)
    A = bm.Matrix.bcsr(m.row_ptr, m.columns, vals, m.n_rows, m.n_cols, m.bs, dtype=dtype, layout=layout)
    y = bm.spmv_host(A, x, variant)
    if dtype == bm.F64:
        want = oracle.bcsr_spmv(m, x, layout=layout, values=vals)
        scale = oracle.bcsr_absmv(m, x, layout=layout, values=vals)
        tol = F64_TOL * scale
    else:
        dv = A.download_values()
        x32 = x.astype(np.float32).astype(np.float64)
        want = oracle.bcsr_spmv(m, x32, layout=layout, values=dv)
        scale = oracle.bcsr_absmv(m, x32, layout=layout, values=dv)
        tol = (row_lengths(m) + 2) * U32 * scale
    err = np.abs(y.astype(np.float64) - want)
    bad = np.nonzero(err > tol)[0]
    assert bad.size == 0, (f"variant {variant} layout {layout} bs {m.bs} dtype {dtype}: "
                           f"{bad.size} rows off, first {bad[:5]}, err {err[bad[:5]]}, tol {tol[bad[:5]]}")
    return A, y


VARIANTS_ROW = [bm.BCSR_THREAD_PER_ROW, bm.BCSR_GROUP_PER_ROW, bm.BCSR_SHUFFLE_X]
VARIANTS_COL = VARIANTS_ROW + [bm.BCSR_BLOCK_ROW_GROUP]


@pytest.mark.parametrize("bs", bm.PRECOMPILED_BS)
@pytest.mark.parametrize("layout", [bm.ROW_MAJOR, bm.COL_MAJOR])
def test_bcsr_f64_stencil_all_variants(bs, layout):
    m, x = stencil_case(13, 11, bs, kind=0, seed=bs)
    for v in (VARIANTS_ROW if layout == bm.ROW_MAJOR else VARIANTS_COL):
        check_bcsr(m, x, layout, v)
    check_bcsr(m, x, layout, bm.SPMV_AUTO)


# rows of 0-40 blocks: the dense lane-group shapes take full batches of U
# blocks, a batch and a tail, several batches, and empty rows (group sums of
# 6 and 16 lanes included)
@pytest.mark.parametrize("bs", [1, 2, 3, 4, 6, 8, 12, 16, 24, 32])
@pytest.mark.parametrize("layout", [bm.ROW_MAJOR, bm.COL_MAJOR])
def test_bcsr_f64_ragged_empty_rows(bs, layout):
    m, x = ragged_case(97, 83, bs, max_len=40, seed=100 + bs)
    for v in (VARIANTS_ROW if layout == bm.ROW_MAJOR else VARIANTS_COL):
        check_bcsr(m, x, layout, v)


@pytest.mark.parametrize("dtype", [bm.F64, bm.F32])
@pytest.mark.parametrize("bs", [1, 2, 3])
@pytest.mark.parametrize("layout", [bm.ROW_MAJOR, bm.COL_MAJOR])
def test_bcsr_small_blocks_short_and_long_row_runs(bs, layout, dtype):
    # wavefronts of 64/bs block rows: runs of short rows (at most 9 blocks
    # each) alternate with runs holding a few long rows, so some wavefronts
    # stage their blocks in LDS and others read them directly
    rng = np.random.default_rng(40 + bs)
    n_rows, n_cols = 700, 650
    rp, cols = [0], []
    for r in range(n_rows):
        long_run = (r // (64 // bs)) % 3 == 1
        ln = int(rng.integers(30, 120)) if long_run and r % 7 == 0 else int(rng.integers(0, 10))
        cols.extend(sorted(rng.choice(n_cols, ln, replace=False).tolist()))
        rp.append(len(cols))
    m = oracle.BCSR(n_rows, n_cols, bs, np.array(rp), np.array(cols, dtype=np.int32),
                    rng.uniform(-1, 1, len(cols) * bs * bs))
    x = rng.uniform(0, 1, n_cols * bs)
    check_bcsr(m, x, layout, bm.BCSR_GROUP_PER_ROW, dtype=dtype)
    check_bcsr(m, x, layout, bm.SPMV_AUTO, dtype=dtype)


@pytest.mark.parametrize("bs", [1, 4, 16])
def test_bcsr_f64_xcd_sized_grid(bs):
    # big enough that the launch uses the XCD tile order (>= 64 workgroups)
    m, x = stencil_case(72, 70, bs, kind=1, seed=3)
    for layout in (bm.ROW_MAJOR, bm.COL_MAJOR):
        for v in (VARIANTS_ROW if layout == bm.ROW_MAJOR else VARIANTS_COL):
            check_bcsr(m, x, layout, v)


# grids past 8 XCD bands (a partial last round of bands included), so the
# lane-group family takes its banded tile order: the staged kernel at bs 1
# and 2 (8192 / 4096 block rows per band), the dense kernels at bs 8 and 12
# (~8192 scalar rows per band)
@pytest.mark.parametrize("bs,side", [(1, 270), (2, 190), (8, 95), (12, 76)])
def test_bcsr_f64_banded_grid(bs, side):
    m, x = stencil_case(side, side, bs, kind=1, seed=5)
    for layout in (bm.ROW_MAJOR, bm.COL_MAJOR):
        check_bcsr(m, x, layout, bm.BCSR_GROUP_PER_ROW)
        check_bcsr(m, x, layout, bm.SPMV_AUTO)


@pytest.mark.parametrize("bs", [5, 7, 40])
def test_bcsr_runtime_block_size(bs):
    m, x = stencil_case(6, 5, bs, kind=0, seed=1)
    for layout in (bm.ROW_MAJOR, bm.COL_MAJOR):
        check_bcsr(m, x, layout, bm.BCSR_THREAD_PER_ROW)
        check_bcsr(m, x, layout, bm.SPMV_AUTO)
        A = bm.Matrix.bcsr(m.row_ptr, m.columns, m.values, m.n_rows, m.n_cols, bs, layout=layout)
        with pytest.raises(bm.BmfpError) as e:
            bm.spmv_host(A, x, bm.BCSR_GROUP_PER_ROW)
        assert e.value.status == bm.ERR_UNSUPPORTED


@pytest.mark.parametrize("bs", [1, 2, 3, 4, 8, 12, 32])
@pytest.mark.parametrize("dtype", [bm.F32, bm.F16])
def test_bcsr_low_precision_storage(bs, dtype):
    m, x = stencil_case(9, 8, bs, kind=0, seed=2)
    for layout in (bm.ROW_MAJOR, bm.COL_MAJOR):
        for v in (VARIANTS_ROW if layout == bm.ROW_MAJOR else VARIANTS_COL):
            check_bcsr(m, x, layout, v, dtype=dtype)


def test_bcsr_identity_blocks_give_x():
    bs, n = 4, 50
    rp = np.arange(n + 1)
    cols = np.arange(n)
    vals = np.tile(np.eye(bs).ravel(), n)
    x = oracle.rng_sym_vec(0, 4, n * bs)
    for layout in (bm.ROW_MAJOR, bm.COL_MAJOR):
        A = bm.Matrix.bcsr(rp, cols, vals, n, n, bs, layout=layout)
        for v in (VARIANTS_ROW if layout == bm.ROW_MAJOR else VARIANTS_COL):
            np.testing.assert_array_equal(bm.spmv_host(A, x, v), x)


def test_bcsr_single_block_is_dense_gemv():
    bs = 32
    rng = np.random.default_rng(1)
    blk = rng.uniform(-1, 1, (bs, bs))
    x = rng.uniform(-1, 1, bs)
    want = blk @ x
    for layout, vals in ((bm.ROW_MAJOR, blk.ravel()), (bm.COL_MAJOR, blk.T.ravel())):
        A = bm.Matrix.bcsr([0, 1], [0], vals, 1, 1, bs, layout=layout)
        for v in (VARIANTS_ROW if layout == bm.ROW_MAJOR else VARIANTS_COL):
            y = bm.spmv_host(A, x, v)
            assert np.all(np.abs(y - want) <= 1e-13 * (np.abs(blk) @ np.abs(x)))


def test_bcsr_empty_matrix_is_noop():
    A = bm.Matrix.bcsr([0], [], [], 0, 0, 4)
    dx = bm.DeviceArray(1)
    dy = bm.DeviceArray(1)
    A.spmv(dx, dy, bm.BCSR_THREAD_PER_ROW)
    bm.synchronize()


def test_bcsr_results_are_bitwise_reproducible():
    m, x = stencil_case(40, 33, 6, kind=1, seed=9)
    for layout in (bm.ROW_MAJOR, bm.COL_MAJOR):
        for v in (VARIANTS_ROW if layout == bm.ROW_MAJOR else VARIANTS_COL):
            _, y1 = check_bcsr(m, x, layout, v)
            _, y2 = check_bcsr(m, x, layout, v)
            assert np.array_equal(y1, y2)


def test_bcsr_library_baseline_rows():
    m, x = stencil_case(20, 17, 4, kind=0, seed=4)
    for layout in (bm.ROW_MAJOR, bm.COL_MAJOR):
        check_bcsr(m, x, layout, bm.BCSR_LIBRARY)


def test_invalid_structures_rejected():
    with pytest.raises(bm.BmfpError) as e:
        bm.Matrix.bcsr([0, 2], [0, 5], np.zeros(8), 1, 2, 2)   # column out of range
    assert e.value.status == bm.ERR_INVALID
    with pytest.raises(bm.BmfpError):
        bm.Matrix.bcsr([0, 2, 1], [0, 1], np.zeros(8), 2, 2, 2)  # row_ptr decreases
    m, x = stencil_case(4, 4, 2)
    A = bm.Matrix.bcsr(m.row_ptr, m.columns, m.values, m.n_rows, m.n_cols, 2)
    with pytest.raises(bm.BmfpError):
        bm.spmv_host(A, x, bm.CSR_SCALAR)


# ---- CSR ----------------------------------------------------------------------
def check_csr(c, x, variant, dtype=bm.F64):
    A = bm.Matrix.csr(c.row_ptr, c.columns, c.values, c.n_rows, c.n_cols, dtype=dtype)
    y = bm.spmv_host(A, x, variant).astype(np.float64)
    if dtype == bm.F64:
        want = oracle.csr_spmv(c, x)
        tol = F64_TOL * oracle.csr_absmv(c, x)
    else:
        v32 = c.values.astype(np.float32).astype(np.float64)
        c32 = oracle.CSR(c.n_rows, c.n_cols, c.row_ptr, c.columns, v32)
        x32 = x.astype(np.float32).astype(np.float64)
        want = oracle.csr_spmv(c32, x32)
        tol = (np.diff(c.row_ptr) + 2) * U32 * oracle.csr_absmv(c32, x32)
    bad = np.nonzero(np.abs(y - want) > tol)[0]
    assert bad.size == 0, (variant, bad[:5])


@pytest.mark.parametrize("bs", [1, 2, 4, 8, 32])
def test_csr_stencil_expansion(bs):
    m, x = stencil_case(12, 10, bs, kind=0, seed=bs)
    c = oracle.bcsr_to_csr(m)
    for v in (bm.CSR_SCALAR, bm.CSR_VECTOR, bm.SPMV_AUTO):
        check_csr(c, x, v)
    check_csr(c, x, bm.CSR_SCALAR, dtype=bm.F32)
    check_csr(c, x, bm.CSR_VECTOR, dtype=bm.F32)


def test_csr_ragged_rows_every_group_width():
    # mean row lengths from ~1 to ~300 cover lane groups of 2 .. 64
    for max_len in (2, 8, 24, 60, 130, 300):
        m, x = ragged_case(150, 400, 1, max_len=max_len, seed=max_len)
        c = oracle.bcsr_to_csr(m)
        for v in (bm.CSR_SCALAR, bm.CSR_VECTOR):
            check_csr(c, x, v)


def csr_from_lengths(lengths, n_cols, seed):
    """CSR with the given row lengths, random sorted columns (with repeats
    allowed across rows), values uniform[-1,1]; x uniform[0,1]."""
    rng = np.random.default_rng(seed)
    lengths = np.asarray(lengths, np.int64)
    rp = np.concatenate([[0], np.cumsum(lengths)])
    cols = np.concatenate([np.sort(rng.choice(n_cols, ln, replace=ln > n_cols)) if ln else np.zeros(0, np.int64)
                           for ln in lengths]).astype(np.int32) if lengths.sum() else np.zeros(0, np.int32)
    vals = rng.uniform(-1, 1, int(rp[-1]))
    x = rng.uniform(0, 1, n_cols)
    return oracle.CSR(len(lengths), n_cols, rp, cols, vals), x


@pytest.mark.parametrize("dtype", [bm.F64, bm.F32])
@pytest.mark.parametrize("base_mean,long_len,n_long", [(6, 100, 11), (20, 200, 10), (100, 800, 4)])
def test_csr_staged_and_direct_wavefronts(base_mean, long_len, n_long, dtype):
    # short rows keep most wavefronts within the LDS staging capacity; a run
    # of long rows from row 640 (a wavefront boundary for every group width)
    # pushes one wavefront beyond it, and that one reads global memory
    # directly.  The longest row stays under 8x the mean + 64, where staging
    # is still chosen.
    rng = np.random.default_rng(base_mean)
    n = 2000
    lengths = rng.integers(0, 2 * base_mean + 1, n)
    lengths[640:640 + n_long] = long_len
    assert lengths.max() <= 8 * lengths.mean() + 64
    c, x = csr_from_lengths(lengths, 3000, seed=base_mean)
    for v in (bm.CSR_SCALAR, bm.CSR_VECTOR):
        check_csr(c, x, v, dtype=dtype)


@pytest.mark.parametrize("bs", [1, 2, 4, 8, 32])
def test_csr_merge_stencil_expansion(bs):
    m, x = stencil_case(12, 10, bs, kind=1, seed=bs + 20)
    c = oracle.bcsr_to_csr(m)
    check_csr(c, x, bm.CSR_MERGE)
    check_csr(c, x, bm.CSR_MERGE, dtype=bm.F32)


MERGE_PATTERNS = {
    # a merge tile is 512 path items (row ends + nonzeros)
    "one_row_over_200_tiles": [100000],
    "long_rows_between_empty_runs": [0] * 700 + [3000] + [0] * 1300 + [1] + [2500, 7] + [0] * 600,
    "all_rows_empty": [0] * 3000,
    "power_law": None,
    "uniform_short": None,
    "tile_boundary_rows": [511, 1, 510, 2, 512, 0, 1023, 1, 1, 511],
}


@pytest.mark.parametrize("name", sorted(MERGE_PATTERNS))
@pytest.mark.parametrize("dtype", [bm.F64, bm.F32])
def test_csr_merge_irregular_rows(name, dtype):
    rng = np.random.default_rng(len(name))
    lengths = MERGE_PATTERNS[name]
    if name == "power_law":
        lengths = np.minimum((rng.pareto(1.1, 6000) * 3).astype(np.int64), 40000)
    elif name == "uniform_short":
        lengths = rng.integers(0, 6, 20000)
    c, x = csr_from_lengths(lengths, 5000, seed=len(name) + 1)
    check_csr(c, x, bm.CSR_MERGE, dtype=dtype)


def test_kernel_symbol_names_the_launched_kernel():
    # bmfp_spmv_kernel_symbol resolves the family the same way bmfp_spmv does
    m, x = stencil_case(9, 7, 4, kind=0, seed=4)
    R = bm.Matrix.bcsr(m.row_ptr, m.columns, m.values, m.n_rows, m.n_cols, 4)
    C = bm.Matrix.bcsr(m.row_ptr, m.columns, oracle.transpose_blocks(m), m.n_rows, m.n_cols, 4,
                       layout=bm.COL_MAJOR)
    assert "k_bcsr_dgr" in R.kernel_symbol(bm.BCSR_GROUP_PER_ROW)
    assert "k_bcsr_dgc" in C.kernel_symbol(bm.BCSR_GROUP_PER_ROW)
    assert "k_bcsr_shfx" in C.kernel_symbol(bm.BCSR_SHUFFLE_X)
    assert "k_bcsr_tpr" in R.kernel_symbol(bm.BCSR_THREAD_PER_ROW)
    assert R.kernel_symbol(bm.SPMV_AUTO) == R.kernel_symbol(R.auto_variant())
    assert R.kernel_symbol(bm.BCSR_GROUP_PER_ROW) != C.kernel_symbol(bm.BCSR_GROUP_PER_ROW)
    c = oracle.bcsr_to_csr(m)
    A = bm.Matrix.csr(c.row_ptr, c.columns, c.values, c.n_rows, c.n_cols)
    sym = A.kernel_symbol(bm.CSR_MERGE)
    assert "k_csr_merge" in sym and "bounds" not in sym and "fixup" not in sym
    with pytest.raises(bm.BmfpError) as e:
        R.kernel_symbol(bm.BCSR_LIBRARY)
    assert e.value.status == bm.ERR_UNSUPPORTED
    with pytest.raises(bm.BmfpError):
        R.kernel_symbol(bm.CSR_MERGE)             # a CSR family on a BCSR matrix
    # the query launched nothing: y keeps its contents
    dx = bm.DeviceArray.from_host(x)
    dy = bm.DeviceArray.from_host(np.full(R.y_len, 7.0))
    R.kernel_symbol(bm.BCSR_GROUP_PER_ROW)
    bm.synchronize()
    assert np.all(dy.download() == 7.0)
    del dx


def test_bcsr_auto_choice_small_blocks():
    # bs 1-3: LDS-staged lane groups; bs 1 only while rows average <= 9 blocks;
    # FP64 column-major bs 4-24: column lane groups, bs 32: shuffle-shared x
    for bs, layout, want in [(1, bm.ROW_MAJOR, bm.BCSR_GROUP_PER_ROW), (2, bm.COL_MAJOR, bm.BCSR_GROUP_PER_ROW),
                             (3, bm.ROW_MAJOR, bm.BCSR_GROUP_PER_ROW), (4, bm.ROW_MAJOR, bm.BCSR_GROUP_PER_ROW),
                             (4, bm.COL_MAJOR, bm.BCSR_GROUP_PER_ROW), (32, bm.COL_MAJOR, bm.BCSR_SHUFFLE_X)]:
        m, x = stencil_case(9, 7, bs, kind=0, seed=bs)
        vals = m.values if layout == bm.ROW_MAJOR else oracle.transpose_blocks(m)
        A = bm.Matrix.bcsr(m.row_ptr, m.columns, vals, m.n_rows, m.n_cols, bs, layout=layout)
        assert A.auto_variant() == want, (bs, layout)
    # column lane groups are FP64 only: FP32 column-major shares x by shuffles
    m, x = stencil_case(9, 7, 4, kind=0, seed=4)
    A = bm.Matrix.bcsr(m.row_ptr, m.columns, oracle.transpose_blocks(m), m.n_rows, m.n_cols, 4,
                       dtype=bm.F32, layout=bm.COL_MAJOR)
    assert A.auto_variant() == bm.BCSR_SHUFFLE_X
    m, x = stencil_case(9, 7, 1, kind=0, seed=1)
    A = bm.Matrix.bcsr(m.row_ptr, m.columns, m.values, m.n_rows, m.n_cols, 1, dtype=bm.F32)
    assert A.auto_variant() == bm.BCSR_GROUP_PER_ROW       # 4-byte blocks are staged too
    A = bm.Matrix.bcsr(m.row_ptr, m.columns, m.values, m.n_rows, m.n_cols, 1, dtype=bm.F16)
    assert A.auto_variant() == bm.BCSR_THREAD_PER_ROW      # 2-byte blocks are not
    m, x = ragged_case(200, 300, 1, max_len=40, seed=5, empty_every=0)   # mean ~20 blocks per row
    A = bm.Matrix.bcsr(m.row_ptr, m.columns, m.values, m.n_rows, m.n_cols, 1)
    assert A.auto_variant() == bm.BCSR_THREAD_PER_ROW
    check_bcsr(m, x, bm.ROW_MAJOR, bm.SPMV_AUTO)


def test_csr_auto_choice_follows_row_lengths():
    cases = [(bm.CSR_SCALAR, [9] * 500), (bm.CSR_VECTOR, [300] * 50), (bm.CSR_VECTOR, [20] * 100),
             (bm.CSR_MERGE, [40] * 50 + [10] * 50), (bm.CSR_SCALAR, list(range(1, 16)) * 30),
             (bm.CSR_VECTOR, list(range(0, 6)) * 80), (bm.CSR_MERGE, list(range(20, 61)) * 10),
             (bm.CSR_SCALAR, [4] * 400)]
    for want, lengths in cases:
        c, x = csr_from_lengths(lengths, 4000, seed=want)
        A = bm.Matrix.csr(c.row_ptr, c.columns, c.values, c.n_rows, c.n_cols)
        assert A.auto_variant() == want
        check_csr(c, x, bm.SPMV_AUTO)
    c, x = csr_from_lengths([2] * 2000 + [5000], 6000, seed=9)  # one row dwarfs the mean
    assert bm.Matrix.csr(c.row_ptr, c.columns, c.values, c.n_rows, c.n_cols).auto_variant() == bm.CSR_MERGE


def test_csr_merge_is_bitwise_reproducible():
    lengths = np.minimum((np.random.default_rng(3).pareto(1.1, 20000) * 4).astype(np.int64), 60000)
    c, x = csr_from_lengths(lengths, 8000, seed=4)
    A = bm.Matrix.csr(c.row_ptr, c.columns, c.values, c.n_rows, c.n_cols)
    y1 = bm.spmv_host(A, x, bm.CSR_MERGE)
    y2 = bm.spmv_host(A, x, bm.CSR_MERGE)
    assert np.array_equal(y1.view(np.uint64), y2.view(np.uint64))


@pytest.mark.parametrize("bs,layout", [(1, bm.ROW_MAJOR), (2, bm.ROW_MAJOR), (3, bm.COL_MAJOR), (5, bm.ROW_MAJOR),
                                       (8, bm.COL_MAJOR)])
def test_device_csr_expansion_is_bit_exact(bs, layout):
    m, x = ragged_case(40, 33, bs, max_len=7, seed=bs)
    vals = m.values if layout == bm.ROW_MAJOR else oracle.transpose_blocks(m)
    A = bm.Matrix.bcsr(m.row_ptr, m.columns, vals, m.n_rows, m.n_cols, bs, layout=layout)
    c = oracle.bcsr_to_csr(m)
    C = A.to_csr()
    assert (C.is_bcsr, C.n_rows, C.n_cols, C.nnzb) == (0, c.n_rows, c.n_cols, len(c.columns))
    rp, cols = C.download_structure()
    np.testing.assert_array_equal(rp, c.row_ptr)
    np.testing.assert_array_equal(cols, c.columns)
    assert np.array_equal(C.download_values().view(np.uint64), c.values.view(np.uint64))
    assert np.array_equal(A.to_csr(bm.F32).download_values(), c.values.astype(np.float32).astype(np.float64))
    for v in (bm.SPMV_AUTO, bm.CSR_MERGE):
        check_csr(c, x, v)


def test_matrix_market_file_through_every_path(tmp_path):
    # a file's CSR through AUTO and merge, and its detected blocks through BCSR AUTO
    m, x = stencil_case(14, 11, 3, kind=1, seed=31)
    c = oracle.bcsr_to_csr(m)
    path = tmp_path / "s3.mtx"
    bm.write_mm(path, c.row_ptr, c.columns, c.values, c.n_rows, c.n_cols)
    rp, cols, vals, n, ncol = bm.read_mm(path)
    r = oracle.CSR(n, ncol, rp, cols, vals)
    for v in (bm.SPMV_AUTO, bm.CSR_MERGE):
        check_csr(r, x, v)
    bs, nnzb, fill = bm.detect_block_size(rp, cols, n, ncol)
    assert (bs, nnzb, fill) == (3, m.nnzb, 1.0)
    b = oracle.csr_to_bcsr(r, bs)
    check_bcsr(b, x, bm.ROW_MAJOR, bm.SPMV_AUTO)


def test_fill_matches_host_rng():
    n = 1000
    d = bm.DeviceArray(n)
    d.fill(0, seed=5, stream_id=4)
    assert np.array_equal(d.download(), oracle.rng_unit_vec(5, 4, n))
    d.fill(1, seed=5, stream_id=6)
    assert np.array_equal(d.download(), oracle.rng_sym_vec(5, 6, n))
    d.fill(2, value=2.5)
    assert np.all(d.download() == 2.5)


@pytest.mark.parametrize("bs,kind,layout,dtype", [(1, 0, bm.ROW_MAJOR, bm.F64), (3, 1, bm.ROW_MAJOR, bm.F64),
                                                  (4, 0, bm.COL_MAJOR, bm.F64), (6, 1, bm.ROW_MAJOR, bm.F32),
                                                  (8, 0, bm.ROW_MAJOR, bm.F16), (1, 1, bm.ROW_MAJOR, bm.F64),
                                                  (2, 0, bm.ROW_MAJOR, bm.F64), (2, 1, bm.COL_MAJOR, bm.F64),
                                                  (3, 0, bm.ROW_MAJOR, bm.F64), (4, 1, bm.ROW_MAJOR, bm.F64),
                                                  (6, 0, bm.ROW_MAJOR, bm.F64), (8, 1, bm.ROW_MAJOR, bm.F64),
                                                  (5, 0, bm.ROW_MAJOR, bm.F64), (12, 1, bm.ROW_MAJOR, bm.F64)])
def test_device_stencil_matches_host_generator(bs, kind, layout, dtype):
    # bmfp_stencil_create generates on the device; the matrix must equal the
    # one uploaded from the oracle's generator, bit for bit
    nx, ny = 13, 7
    D = bm.Matrix.stencil(nx, ny, bs, kind, seed=9, dtype=dtype, layout=layout)
    m = oracle.gen_stencil9(nx, ny, bs, kind, 9)
    H = bm.Matrix.bcsr(m.row_ptr, m.columns, m.values, m.n_rows, m.n_cols, bs, dtype=dtype, layout=bm.ROW_MAJOR)
    if layout != bm.ROW_MAJOR:
        H = H.convert(dtype, layout)
    assert (D.n_rows, D.nnzb, D.bs, D.layout) == (H.n_rows, H.nnzb, H.bs, H.layout)
    assert np.array_equal(D.download_values(), H.download_values())
    x = oracle.rng_unit_vec(2, 4, m.n_cols * bs)
    assert np.array_equal(bm.spmv_host(D, x), bm.spmv_host(H, x))



def test_stencil_regenerate_in_place():
    # bmfp_stencil_generate rewrites a resident stencil matrix for another
    # kind / seed: the values equal a fresh bmfp_stencil_create and the host
    # generator's, bitwise
    A = bm.Matrix.stencil(13, 9, 4, 0, seed=1)
    A.regenerate(1, 77)
    bm.synchronize()
    m = oracle.gen_stencil9(13, 9, 4, 1, 77)
    assert np.array_equal(A.download_values(), m.values)
    with pytest.raises(bm.BmfpError):
        A.convert(bm.F32).regenerate(0, 1)   # only F64 row-major stencil handles

@pytest.mark.parametrize("bs", [5, 7, 9, 10, 13, 20, 40, 70])
@pytest.mark.parametrize("dtype", [bm.F64, bm.F32])
def test_bcsr_jit_kernel(bs, dtype):
    # hipRTC-specialised kernel (the AUTO choice for these block sizes)
    m, x = stencil_case(5, 4, bs, kind=1, seed=bs)
    for layout in (bm.ROW_MAJOR, bm.COL_MAJOR):
        A, y = check_bcsr(m, x, layout, bm.BCSR_JIT, dtype=dtype)
        assert A.auto_variant() == bm.BCSR_JIT
        y2 = bm.spmv_host(A, x, bm.SPMV_AUTO)
        assert np.array_equal(y, y2)               # cached module, same kernel


@pytest.mark.parametrize("bs", [1, 4, 12])
def test_bcsr_jit_on_precompiled_sizes_and_ragged_rows(bs):
    m, x = ragged_case(61, 57, bs, max_len=30, seed=7 + bs)
    for layout in (bm.ROW_MAJOR, bm.COL_MAJOR):
        check_bcsr(m, x, layout, bm.BCSR_JIT)


def test_bcsr_jit_unsupported_for_f16():
    m, x = stencil_case(4, 4, 5, kind=0, seed=1)
    A = bm.Matrix.bcsr(m.row_ptr, m.columns, m.values, m.n_rows, m.n_cols, 5, dtype=bm.F16)
    assert A.auto_variant() == bm.BCSR_THREAD_PER_ROW
    with pytest.raises(bm.BmfpError) as e:
        bm.spmv_host(A, x, bm.BCSR_JIT)
    assert e.value.status == bm.ERR_UNSUPPORTED


@pytest.mark.parametrize("n16", [1, 255, 256 * 8 * 8192 + 37])
def test_stream_yardsticks(n16):
    # the copy kernel's unrolled loop and its tail both land every chunk; the
    # read kernel reads and leaves its sink alone (its fold of these values
    # is not the sentinel); misaligned sizes are rejected
    L = bm.lib()
    src = bm.DeviceArray.from_host(np.random.default_rng(n16).uniform(-1, 1, 2 * n16))
    dst = bm.DeviceArray(2 * n16)
    dst.zero()
    bm.check(L.bmfp_stream_copy(dst.ptr, src.ptr, 16 * n16, None))
    bm.synchronize()
    assert np.array_equal(dst.download(), src.download())
    sink = bm.DeviceArray(2)
    sink.upload(np.array([7.0, 7.0]))
    bm.check(L.bmfp_stream_read(src.ptr, 16 * n16, sink.ptr, None))
    bm.synchronize()
    assert np.array_equal(sink.download(), [7.0, 7.0])
    assert L.bmfp_stream_read(src.ptr, 16 * n16 + 8, sink.ptr, None) == bm.ERR_INVALID
    assert L.bmfp_stream_copy(dst.ptr, src.ptr, 8, None) == bm.ERR_INVALID


# Values written by a device kernel moments before (the stencil generator,
# regenerating the matrix in place under new seeds) must be what every family
# reads, the families with non-temporal value loads included; small grids
# keep the fresh values inside the caches.
@pytest.mark.parametrize("bs,layout", [(8, bm.ROW_MAJOR), (12, bm.ROW_MAJOR), (16, bm.ROW_MAJOR),
                                       (32, bm.ROW_MAJOR), (8, bm.COL_MAJOR), (4, bm.ROW_MAJOR)])
def test_spmv_reads_freshly_generated_values(bs, layout):
    nx, ny = 24, 20
    A = bm.Matrix.stencil(nx, ny, bs, 0, seed=0)
    x = oracle.rng_unit_vec(5, 4, nx * ny * bs)
    dx = bm.DeviceArray.from_host(x)
    dy = bm.DeviceArray(nx * ny * bs)
    for seed in (11, 12, 13, 14):
        A.regenerate(0, seed)
        M = A if layout == bm.ROW_MAJOR else A.convert(bm.F64, bm.COL_MAJOR)
        for v in (bm.BCSR_GROUP_PER_ROW, bm.SPMV_AUTO, bm.BCSR_THREAD_PER_ROW):
            M.spmv(dx, dy, v)
            y = dy.download()
            m = oracle.gen_stencil9(nx, ny, bs, 0, seed)
            want = oracle.bcsr_spmv(m, x)
            tol = F64_TOL * oracle.bcsr_absmv(m, x)
            bad = np.nonzero(np.abs(y - want) > tol)[0]
            assert bad.size == 0, (bs, layout, seed, v, bad[:5])
        if M is not A:
            M.destroy()
    A.destroy()
