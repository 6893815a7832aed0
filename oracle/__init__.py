"""Python (ctypes + numpy) view of the CPU oracle in oracle.c.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, and there only as the checker / the timed CPU
baseline.  The product library never imports, links or calls it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class SolveInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("status", ctypes.c_int),
                ("rel_residual", ctypes.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        u64, c_int, c_dbl = ctypes.c_uint64, ctypes.c_int, ctypes.c_double
        L.orc_rng_u64.restype = u64
        L.orc_rng_u64.argtypes = [u64, u64, u64]
        L.orc_rng_unit.restype = c_dbl
        L.orc_rng_unit.argtypes = [u64, u64, u64]
        L.orc_rng_sym.restype = c_dbl
        L.orc_rng_sym.argtypes = [u64, u64, u64]
        L.orc_rng_fill.argtypes = [u64, u64, c_int, ctypes.c_int64, _f64p]
        L.orc_gen_n_diag_bcsr.argtypes = [c_int, c_int, c_int, _i32p, _i32p, _f64p]
        L.orc_stencil_nnzb.restype = ctypes.c_int64
        L.orc_stencil_nnzb.argtypes = [c_int, c_int]
        L.orc_gen_stencil9.argtypes = [c_int, c_int, c_int, c_int, u64, _i32p, _i32p, _f64p]
        L.orc_transpose_blocks.argtypes = [c_int, c_int, _i32p, _f64p, _f64p]
        L.orc_bcsr_to_csr.argtypes = [c_int, c_int, _i32p, _i32p, _f64p, _i32p, _i32p, _f64p]
        L.orc_csr_to_bcsr_count.restype = c_int
        L.orc_csr_to_bcsr_count.argtypes = [c_int, c_int, _i32p, _i32p]
        L.orc_csr_to_bcsr.restype = c_int
        L.orc_csr_to_bcsr.argtypes = [c_int, c_int, _i32p, _i32p, _f64p, _i32p, _i32p, _f64p,
                                      ctypes.POINTER(c_int)]
        L.orc_csr_spmv.argtypes = [c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.orc_csr_absmv.argtypes = [c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.orc_bcsr_spmv.argtypes = [c_int, c_int, c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.orc_bcsr_absmv.argtypes = [c_int, c_int, c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.orc_bicgstab.argtypes = [c_int, _i32p, _i32p, _f64p, _f64p, c_dbl, c_int, c_int, _f64p,
                                   ctypes.POINTER(SolveInfo)]
        L.orc_compare_metric.restype = c_dbl
        L.orc_compare_metric.argtypes = [c_int, _f64p, _f64p]
        _lib = L
    return _lib


class BCSR:
    """Host BCSR matrix (int32 indices, float64 values, row-major blocks)."""

    def __init__(self, n_rows, n_cols, bs, row_ptr, columns, values):
        self.n_rows, self.n_cols, self.bs = int(n_rows), int(n_cols), int(bs)
        self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int32)
        self.columns = np.ascontiguousarray(columns, dtype=np.int32)
        self.values = np.ascontiguousarray(values, dtype=np.float64)

    @property
    def nnzb(self):
        return int(self.row_ptr[-1])


class CSR:
    def __init__(self, n_rows, n_cols, row_ptr, columns, values):
        self.n_rows, self.n_cols = int(n_rows), int(n_cols)
        self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int32)
        self.columns = np.ascontiguousarray(columns, dtype=np.int32)
        self.values = np.ascontiguousarray(values, dtype=np.float64)

    @property
    def nnz(self):
        return int(self.row_ptr[-1])


def rng_unit_vec(seed, stream, n):
    out = np.empty(int(n), np.float64)
    lib().orc_rng_fill(seed, stream, 0, int(n), out)
    return out


def rng_sym_vec(seed, stream, n):
    out = np.empty(int(n), np.float64)
    lib().orc_rng_fill(seed, stream, 1, int(n), out)
    return out


def gen_n_diag_bcsr(n_rows, blocks_per_row, bs):
    nnzb = n_rows * blocks_per_row
    rp = np.zeros(n_rows + 1, np.int32)
    cols = np.zeros(nnzb, np.int32)
    vals = np.zeros(nnzb * bs * bs, np.float64)
    lib().orc_gen_n_diag_bcsr(n_rows, blocks_per_row, bs, rp, cols, vals)
    return BCSR(n_rows, n_rows, bs, rp, cols, vals)


def gen_stencil9(nx, ny, bs, kind, seed):
    L = lib()
    nnzb = int(L.orc_stencil_nnzb(nx, ny))
    n = nx * ny
    rp = np.zeros(n + 1, np.int32)
    cols = np.zeros(nnzb, np.int32)
    vals = np.zeros(nnzb * bs * bs, np.float64)
    L.orc_gen_stencil9(nx, ny, bs, kind, seed, rp, cols, vals)
    return BCSR(n, n, bs, rp, cols, vals)


def transpose_blocks(m):
    out = np.empty_like(m.values)
    lib().orc_transpose_blocks(m.n_rows, m.bs, m.row_ptr, m.values, out)
    return out


def bcsr_to_csr(m):
    n = m.n_rows * m.bs
    nnz = m.nnzb * m.bs * m.bs
    rp = np.zeros(n + 1, np.int32)
    cols = np.zeros(nnz, np.int32)
    vals = np.zeros(nnz, np.float64)
    lib().orc_bcsr_to_csr(m.n_rows, m.bs, m.row_ptr, m.columns, m.values, rp, cols, vals)
    return CSR(n, m.n_cols * m.bs, rp, cols, vals)


def csr_to_bcsr(c, bs):
    L = lib()
    nnzb = L.orc_csr_to_bcsr_count(c.n_rows, bs, c.row_ptr, c.columns)
    if nnzb < 0:
        raise ValueError("n_rows is not a multiple of bs")
    nb = c.n_rows // bs
    rp = np.zeros(nb + 1, np.int32)
    cols = np.zeros(max(nnzb, 1), np.int32)
    vals = np.zeros(max(nnzb, 1) * bs * bs, np.float64)
    out = ctypes.c_int(0)
    rc = L.orc_csr_to_bcsr(c.n_rows, bs, c.row_ptr, c.columns, c.values, rp, cols, vals,
                           ctypes.byref(out))
    if rc != 0:
        raise ValueError("csr_to_bcsr failed")
    return BCSR(nb, c.n_cols // bs, bs, rp, cols[:nnzb], vals[:nnzb * bs * bs])


def csr_spmv(c, x):
    y = np.zeros(c.n_rows, np.float64)
    lib().orc_csr_spmv(c.n_rows, c.row_ptr, c.columns, c.values, np.ascontiguousarray(x, np.float64), y)
    return y


def csr_absmv(c, x):
    y = np.zeros(c.n_rows, np.float64)
    lib().orc_csr_absmv(c.n_rows, c.row_ptr, c.columns, c.values, np.ascontiguousarray(x, np.float64), y)
    return y


def bcsr_spmv(m, x, layout=0, values=None):
    v = m.values if values is None else np.ascontiguousarray(values, np.float64)
    y = np.zeros(m.n_rows * m.bs, np.float64)
    lib().orc_bcsr_spmv(m.n_rows, m.bs, layout, m.row_ptr, m.columns, v,
                        np.ascontiguousarray(x, np.float64), y)
    return y


def bcsr_absmv(m, x, layout=0, values=None):
    v = m.values if values is None else np.ascontiguousarray(values, np.float64)
    y = np.zeros(m.n_rows * m.bs, np.float64)
    lib().orc_bcsr_absmv(m.n_rows, m.bs, layout, m.row_ptr, m.columns, v,
                         np.ascontiguousarray(x, np.float64), y)
    return y


def bicgstab(c, b, eps, max_iterations, use_jacobi=False):
    x = np.zeros(c.n_rows, np.float64)
    info = SolveInfo()
    lib().orc_bicgstab(c.n_rows, c.row_ptr, c.columns, c.values, np.ascontiguousarray(b, np.float64),
                       eps, max_iterations, 1 if use_jacobi else 0, x, ctypes.byref(info))
    return x, info


def compare_metric(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return lib().orc_compare_metric(len(a), a, b)
