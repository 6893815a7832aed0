"""Python binding of the gfx950 block-sparse SpMV / BiCGStab library (libbmfp.so).

The product is the C ABI declared in include/bmfp.h, implemented by hand-written
HIP kernels for gfx950.  This module is a thin ctypes layer over it for tests and
the benchmark script; every compute call goes through libbmfp.so, and importing
anything that needs the library fails loudly if it has not been built.
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "libbmfp.so")
CPP_LIB_PATH = os.path.join(PKG_DIR, "libbmfp_cpp.so")
MAIN_PATH = os.path.join(PKG_DIR, "bmfp_main")
HEADER_PATH = os.path.join(REPO_DIR, "include", "bmfp.h")

# ---- constants mirrored from include/bmfp.h ---------------------------------
OK = 0
ERR_INVALID = -1
ERR_UNSUPPORTED = -2
ERR_NO_MEMORY = -3
ERR_BREAKDOWN = -4
ERR_NOT_CONVERGED = -5
ERR_HIP = -6
ERR_NONFINITE = -7
ERR_BUDGET = -8
ERR_NO_DEVICE = -9

F64, F32, F16 = 0, 1, 2
ROW_MAJOR, COL_MAJOR = 0, 1

SPMV_AUTO = 0
CSR_SCALAR = 1
CSR_VECTOR = 2
BCSR_THREAD_PER_ROW = 3
BCSR_GROUP_PER_ROW = 4
BCSR_SHUFFLE_X = 5
BCSR_BLOCK_ROW_GROUP = 6
BCSR_LIBRARY = 7
BCSR_JIT = 8
CSR_MERGE = 9

BCSR_VARIANTS = (BCSR_THREAD_PER_ROW, BCSR_GROUP_PER_ROW, BCSR_SHUFFLE_X, BCSR_BLOCK_ROW_GROUP)
PRECOMPILED_BS = (1, 2, 3, 4, 6, 8, 12, 16, 24, 32)

_VEC_NP = {F64: np.float64, F32: np.float32, F16: np.float32}


class BmfpError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"bmfp status {status}: {message}")
        self.status = status


_lib = None


def lib():
    """Loads libbmfp.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                              "or `make -C block_matrix_format_performance_amd`")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    c_int, c_dbl, vp, sz = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t
    ip = ctypes.POINTER(ctypes.c_int)
    dp = ctypes.POINTER(ctypes.c_double)
    L.bmfp_last_error.restype = ctypes.c_char_p
    L.bmfp_status_string.restype = ctypes.c_char_p
    L.bmfp_status_string.argtypes = [c_int]
    L.bmfp_version.restype = c_int
    L.bmfp_init.argtypes = [c_int]
    L.bmfp_device_count.argtypes = [ip]
    L.bmfp_device_info.argtypes = [ctypes.c_char_p, sz, ctypes.POINTER(sz), ip, ip]
    L.bmfp_malloc.argtypes = [ctypes.POINTER(vp), sz]
    L.bmfp_free.argtypes = [vp]
    L.bmfp_memcpy_h2d.argtypes = [vp, vp, sz]
    L.bmfp_memcpy_d2h.argtypes = [vp, vp, sz]
    L.bmfp_memcpy_d2d.argtypes = [vp, vp, sz]
    L.bmfp_memset.argtypes = [vp, c_int, sz]
    L.bmfp_stream_create.argtypes = [ctypes.POINTER(vp)]
    L.bmfp_stream_destroy.argtypes = [vp]
    L.bmfp_stream_synchronize.argtypes = [vp]
    L.bmfp_event_create.argtypes = [ctypes.POINTER(vp)]
    L.bmfp_event_destroy.argtypes = [vp]
    L.bmfp_event_record.argtypes = [vp, vp]
    L.bmfp_event_elapsed.argtypes = [vp, vp, dp]
    L.bmfp_stream_copy.argtypes = [vp, vp, sz, vp]
    L.bmfp_stream_read.argtypes = [vp, sz, vp, vp]
    L.bmfp_fill.argtypes = [vp, c_int, sz, c_int, ctypes.c_uint64, ctypes.c_uint64, c_dbl, vp]
    L.bmfp_gen_n_diag_bcsr.argtypes = [c_int, c_int, c_int, vp, vp, vp]
    L.bmfp_stencil_nnzb.restype = ctypes.c_int64
    L.bmfp_stencil_nnzb.argtypes = [c_int, c_int]
    L.bmfp_gen_stencil9.argtypes = [c_int, c_int, c_int, c_int, ctypes.c_uint64, vp, vp, vp]
    L.bmfp_mm_read.argtypes = [ctypes.c_char_p, ip, ip, ip, vp, vp, vp]
    L.bmfp_mm_write.argtypes = [ctypes.c_char_p, c_int, c_int, vp, vp, vp]
    L.bmfp_block_size_detect.argtypes = [c_int, c_int, vp, vp, c_int, vp, c_int, ip, ip, dp]
    L.bmfp_bcsr_create.argtypes = [ctypes.POINTER(vp), c_int, c_int, c_int, c_int, c_int, c_int, vp, vp, vp]
    L.bmfp_csr_create.argtypes = [ctypes.POINTER(vp), c_int, c_int, c_int, c_int, vp, vp, vp]
    L.bmfp_stencil_create.argtypes = [ctypes.POINTER(vp), c_int, c_int, c_int, c_int, ctypes.c_uint64, c_int, c_int]
    L.bmfp_matrix_convert.argtypes = [vp, c_int, c_int, ctypes.POINTER(vp), vp]
    L.bmfp_matrix_destroy.argtypes = [vp]
    L.bmfp_matrix_shape.argtypes = [vp, ip, ip, ip, ip, ip, ip, ip]
    L.bmfp_spmv_traffic.argtypes = [vp, dp, dp]
    L.bmfp_matrix_download_values.argtypes = [vp, vp]
    L.bmfp_spmv.argtypes = [vp, c_int, vp, vp, vp]
    L.bmfp_spmv_timed.argtypes = [vp, c_int, vp, vp, c_int, c_int, vp, dp]
    L.bmfp_spmv_auto_variant.argtypes = [vp, ip]
    L.bmfp_spmv_kernel_symbol.argtypes = [vp, c_int, ctypes.c_char_p, sz]
    L.bmfp_jit_build.argtypes = [c_int, c_int, c_int, ctypes.POINTER(sz), dp]
    L.bmfp_csr_spmv_noalloc.argtypes = [c_int, c_int, c_int, c_int, vp, vp, vp, vp, vp, vp]
    L.bmfp_stencil_generate.argtypes = [vp, c_int, ctypes.c_uint64, vp]
    _declare_solver(L)


def _declare_solver(L):
    """Solver / refinement / serving entry points (present once built)."""
    c_int, c_dbl, vp = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    L.bmfp_bicgstab_create.argtypes = [ctypes.POINTER(vp), vp, c_int]
    L.bmfp_bicgstab_solve.argtypes = [vp, vp, vp, c_dbl, c_int, vp, ctypes.POINTER(SolveInfo)]
    L.bmfp_bicgstab_destroy.argtypes = [vp]
    L.bmfp_bicgstab_workspace_bytes.argtypes = [vp, c_int, ctypes.POINTER(ctypes.c_size_t)]
    L.bmfp_refine_create.argtypes = [ctypes.POINTER(vp), vp, c_int, c_int]
    L.bmfp_refine_solve.argtypes = [vp, vp, vp, c_dbl, c_dbl, c_int, c_int, vp, ctypes.POINTER(SolveInfo)]
    L.bmfp_refine_destroy.argtypes = [vp]
    L.bmfp_refine_copy_seconds.argtypes = [vp, c_int, ctypes.POINTER(c_dbl)]


class SolveInfo(ctypes.Structure):
    """Mirror of bmfp_solve_info (include/bmfp.h)."""
    _fields_ = [("iterations", ctypes.c_int), ("status", ctypes.c_int),
                ("rel_residual", ctypes.c_double), ("true_rel_residual", ctypes.c_double),
                ("seconds", ctypes.c_double), ("outer_iterations", ctypes.c_int),
                ("inner_iterations", ctypes.c_int), ("inner_dtype", ctypes.c_int),
                ("escalations", ctypes.c_int)]


def check(status):
    if status != OK:
        raise BmfpError(status, lib().bmfp_last_error().decode(errors="replace"))
    return status


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# ---- device plumbing -----------------------------------------------------------
def device_count():
    n = ctypes.c_int(0)
    check(lib().bmfp_device_count(ctypes.byref(n)))
    return n.value


def init(device=0):
    check(lib().bmfp_init(device))


def device_info():
    name = ctypes.create_string_buffer(256)
    mem = ctypes.c_size_t(0)
    cus = ctypes.c_int(0)
    arch = ctypes.c_int(0)
    check(lib().bmfp_device_info(name, 256, ctypes.byref(mem), ctypes.byref(cus), ctypes.byref(arch)))
    return {"name": name.value.decode(), "total_mem": mem.value, "compute_units": cus.value,
            "arch": arch.value}


def synchronize():
    check(lib().bmfp_device_synchronize())


class DeviceArray:
    """Owning device buffer of `n` elements of numpy dtype `dtype`."""

    def __init__(self, n, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(lib().bmfp_malloc(ctypes.byref(p), max(self.nbytes, 1)))
        self.ptr = p

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        check(lib().bmfp_memcpy_h2d(self.ptr, _ptr(a), self.nbytes))

    def download(self):
        out = np.empty(self.n, self.dtype)
        check(lib().bmfp_memcpy_d2h(_ptr(out), self.ptr, self.nbytes))
        return out

    def fill(self, kind, seed=0, stream_id=0, value=0.0, stream=None):
        dt = F64 if self.dtype == np.float64 else F32
        check(lib().bmfp_fill(self.ptr, dt, self.n, kind, seed, stream_id, value, stream))

    def zero(self):
        check(lib().bmfp_memset(self.ptr, 0, self.nbytes))

    def view(self, start, n):
        """Non-owning view of elements [start, start+n) (keeps this buffer alive)."""
        if start < 0 or n < 0 or start + n > self.n:
            raise IndexError(f"view [{start}, {start + n}) outside [0, {self.n})")
        return DeviceView(self, start, n)

    def free(self):
        if self.ptr is not None and self.ptr.value:
            lib().bmfp_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceView:
    """A contiguous slice of a DeviceArray; usable wherever a DeviceArray is read or written."""

    def __init__(self, owner, start, n):
        self.owner = owner
        self.n = int(n)
        self.dtype = owner.dtype
        self.nbytes = self.n * self.dtype.itemsize
        self.ptr = ctypes.c_void_p(owner.ptr.value + int(start) * self.dtype.itemsize)

    def download(self):
        out = np.empty(self.n, self.dtype)
        check(lib().bmfp_memcpy_d2h(_ptr(out), self.ptr, self.nbytes))
        return out


class Stream:
    def __init__(self):
        s = ctypes.c_void_p()
        check(lib().bmfp_stream_create(ctypes.byref(s)))
        self.handle = s

    def synchronize(self):
        check(lib().bmfp_stream_synchronize(self.handle))

    def destroy(self):
        if self.handle is not None:
            lib().bmfp_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Event:
    """hipEvent_t recorded on a stream (None = null stream)."""

    def __init__(self):
        e = ctypes.c_void_p()
        check(lib().bmfp_event_create(ctypes.byref(e)))
        self.handle = e

    def record(self, stream=None):
        check(lib().bmfp_event_record(self.handle, stream))

    def elapsed_since(self, start):
        t = ctypes.c_double(0)
        check(lib().bmfp_event_elapsed(start.handle, self.handle, ctypes.byref(t)))
        return t.value

    def __del__(self):
        try:
            lib().bmfp_event_destroy(self.handle)
        except Exception:
            pass


# ---- host generators (product code, libbmfp.so) ----------------------------------
def gen_stencil9(nx, ny, bs, kind=0, seed=0):
    """9-point block stencil (row-major blocks); returns (row_ptr, columns, values)."""
    L = lib()
    nnzb = int(L.bmfp_stencil_nnzb(nx, ny))
    n = nx * ny
    rp = np.zeros(n + 1, np.int32)
    cols = np.zeros(nnzb, np.int32)
    vals = np.zeros(nnzb * bs * bs, np.float64)
    check(L.bmfp_gen_stencil9(nx, ny, bs, kind, seed, _ptr(rp), _ptr(cols), _ptr(vals)))
    return rp, cols, vals


def read_mm(path):
    """Matrix Market file -> CSR (row_ptr, columns, values, n_rows, n_cols): columns
    ascending within a row, duplicates summed, symmetric halves mirrored."""
    L = lib()
    p = os.fsencode(path)
    n, m, nnz = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    check(L.bmfp_mm_read(p, ctypes.byref(n), ctypes.byref(m), ctypes.byref(nnz), None, None, None))
    rp = np.zeros(n.value + 1, np.int32)
    cols = np.zeros(nnz.value, np.int32)
    vals = np.zeros(nnz.value, np.float64)
    check(L.bmfp_mm_read(p, ctypes.byref(n), ctypes.byref(m), ctypes.byref(nnz), _ptr(rp), _ptr(cols), _ptr(vals)))
    return rp, cols, vals, n.value, m.value


def write_mm(path, row_ptr, columns, values, n_rows, n_cols):
    """CSR -> "coordinate real general" Matrix Market text (csr_matrix_class::write_mm)."""
    rp = np.ascontiguousarray(row_ptr, np.int32)
    cols = np.ascontiguousarray(columns, np.int32)
    vals = np.ascontiguousarray(values, np.float64)
    check(lib().bmfp_mm_write(os.fsencode(path), n_rows, n_cols, _ptr(rp), _ptr(cols), _ptr(vals)))


BLOCK_SIZE_CANDIDATES = (1, 2, 3, 4, 6, 8, 12, 16, 24, 32)


def detect_block_size(row_ptr, columns, n_rows, n_cols, value_bytes=8, candidates=BLOCK_SIZE_CANDIDATES):
    """(bs, nnzb, fill): the candidate block size whose BCSR SpMV moves the fewest
    compulsory bytes for this CSR pattern."""
    rp = np.ascontiguousarray(row_ptr, np.int32)
    cols = np.ascontiguousarray(columns, np.int32)
    cand = np.ascontiguousarray(candidates, np.int32)
    bs, nnzb, fill = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_double(0)
    check(lib().bmfp_block_size_detect(n_rows, n_cols, _ptr(rp), _ptr(cols), value_bytes, _ptr(cand), len(cand),
                                       ctypes.byref(bs), ctypes.byref(nnzb), ctypes.byref(fill)))
    return bs.value, nnzb.value, fill.value


def gen_n_diag_bcsr(n_rows, blocks_per_row, bs):
    nnzb = n_rows * blocks_per_row
    rp = np.zeros(n_rows + 1, np.int32)
    cols = np.zeros(nnzb, np.int32)
    vals = np.zeros(nnzb * bs * bs, np.float64)
    check(lib().bmfp_gen_n_diag_bcsr(n_rows, blocks_per_row, bs, _ptr(rp), _ptr(cols), _ptr(vals)))
    return rp, cols, vals


# ---- matrices ---------------------------------------------------------------------
class Matrix:
    """Device-resident BCSR or CSR matrix (a bmfp_matrix handle)."""

    def __init__(self, handle, keep=None):
        self.handle = handle
        self._keep = keep
        vals = [ctypes.c_int(0) for _ in range(7)]
        check(lib().bmfp_matrix_shape(handle, *[ctypes.byref(v) for v in vals]))
        (self.is_bcsr, self.dtype, self.layout, self.n_rows, self.n_cols,
         self.bs, self.nnzb) = [v.value for v in vals]

    @classmethod
    def bcsr(cls, row_ptr, columns, values, n_rows, n_cols, bs, dtype=F64, layout=ROW_MAJOR):
        rp = np.ascontiguousarray(row_ptr, np.int32)
        cols = np.ascontiguousarray(columns, np.int32)
        vals = np.ascontiguousarray(values, np.float64)
        h = ctypes.c_void_p()
        check(lib().bmfp_bcsr_create(ctypes.byref(h), dtype, layout, n_rows, n_cols, bs,
                                     int(rp[-1]), _ptr(rp), _ptr(cols), _ptr(vals)))
        return cls(h)

    @classmethod
    def stencil(cls, nx, ny, bs, kind=0, seed=0, dtype=F64, layout=ROW_MAJOR):
        """9-point block stencil generated in device memory (same values as gen_stencil9)."""
        h = ctypes.c_void_p()
        check(lib().bmfp_stencil_create(ctypes.byref(h), nx, ny, bs, kind, seed, dtype, layout))
        return cls(h)

    @classmethod
    def csr(cls, row_ptr, columns, values, n_rows, n_cols, dtype=F64):
        rp = np.ascontiguousarray(row_ptr, np.int32)
        cols = np.ascontiguousarray(columns, np.int32)
        vals = np.ascontiguousarray(values, np.float64)
        h = ctypes.c_void_p()
        check(lib().bmfp_csr_create(ctypes.byref(h), dtype, n_rows, n_cols, int(rp[-1]),
                                    _ptr(rp), _ptr(cols), _ptr(vals)))
        return cls(h)

    def regenerate(self, kind, seed, stream=None):
        """bmfp_stencil_generate: new values in place (a stencil-made F64 row-major matrix)."""
        check(lib().bmfp_stencil_generate(self.handle, kind, seed, stream))

    def convert(self, dtype, layout=None, stream=None):
        h = ctypes.c_void_p()
        check(lib().bmfp_matrix_convert(self.handle, dtype, self.layout if layout is None else layout,
                                        ctypes.byref(h), stream))
        return Matrix(h)

    @property
    def vec_dtype(self):
        return _VEC_NP[self.dtype]

    @property
    def x_len(self):
        return self.n_cols * self.bs

    @property
    def y_len(self):
        return self.n_rows * self.bs

    def spmv(self, x, y, variant=SPMV_AUTO, stream=None):
        check(lib().bmfp_spmv(self.handle, variant, x.ptr, y.ptr, stream))

    def spmv_timed(self, x, y, variant=SPMV_AUTO, warmup=20, reps=200, stream=None):
        t = ctypes.c_double(0)
        check(lib().bmfp_spmv_timed(self.handle, variant, x.ptr, y.ptr, warmup, reps, stream,
                                    ctypes.byref(t)))
        return t.value

    def traffic(self):
        b, f = ctypes.c_double(0), ctypes.c_double(0)
        check(lib().bmfp_spmv_traffic(self.handle, ctypes.byref(b), ctypes.byref(f)))
        return b.value, f.value

    def auto_variant(self):
        v = ctypes.c_int(0)
        check(lib().bmfp_spmv_auto_variant(self.handle, ctypes.byref(v)))
        return v.value

    def kernel_symbol(self, variant=SPMV_AUTO):
        """Mangled symbol of the kernel spmv(x, y, variant) launches."""
        buf = ctypes.create_string_buffer(1024)
        check(lib().bmfp_spmv_kernel_symbol(self.handle, variant, buf, len(buf)))
        return buf.value.decode()

    def to_csr(self, dtype=F64, stream=None):
        """CSR expansion in device memory (csr_matrix_class(bcsr) entry order)."""
        h = ctypes.c_void_p()
        check(lib().bmfp_matrix_to_csr(self.handle, dtype, ctypes.byref(h), stream))
        return Matrix(h)

    def download_structure(self):
        """(row_ptr, columns) as int32 arrays."""
        rp = np.zeros(self.n_rows + 1, np.int32)
        cols = np.zeros(self.nnzb, np.int32)
        check(lib().bmfp_matrix_download_structure(self.handle, _ptr(rp), _ptr(cols)))
        return rp, cols

    def download_values(self):
        out = np.empty(self.nnzb * self.bs * self.bs, np.float64)
        check(lib().bmfp_matrix_download_values(self.handle, _ptr(out)))
        return out

    def destroy(self):
        if self.handle is not None:
            lib().bmfp_matrix_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def jit_build(bs, dtype=F64, layout=ROW_MAJOR):
    """Compiles (hipRTC, gfx950) the BCSR_JIT kernel for (bs, dtype, layout): (code bytes, seconds)."""
    nb, sec = ctypes.c_size_t(0), ctypes.c_double(0)
    check(lib().bmfp_jit_build(bs, dtype, layout, ctypes.byref(nb), ctypes.byref(sec)))
    return nb.value, sec.value


def spmv_host(A, x, variant=SPMV_AUTO):
    """Uploads x, runs one SpMV through the library, returns y on the host."""
    dx = DeviceArray.from_host(np.ascontiguousarray(x, A.vec_dtype))
    dy = DeviceArray(A.y_len, A.vec_dtype)
    A.spmv(dx, dy, variant)
    synchronize()
    return dy.download()


# ---- solvers ------------------------------------------------------------------------
SOLVE_STATUSES = (OK, ERR_NOT_CONVERGED, ERR_BREAKDOWN, ERR_NONFINITE)


class BiCGStab:
    """bmfp_bicgstab_* handle: BiCGStab on a device-resident BCSR matrix."""

    def __init__(self, A, use_jacobi=False):
        self.A = A
        h = ctypes.c_void_p()
        check(lib().bmfp_bicgstab_create(ctypes.byref(h), A.handle, 1 if use_jacobi else 0))
        self.handle = h

    def solve(self, b, x, eps, max_iterations, stream=None):
        """b, x: DeviceArray.  Returns SolveInfo; raises only for bad arguments."""
        info = SolveInfo()
        st = lib().bmfp_bicgstab_solve(self.handle, b.ptr, x.ptr, eps, max_iterations, stream,
                                       ctypes.byref(info))
        if st not in SOLVE_STATUSES:
            check(st)
        return info

    def solve_host(self, b, eps=1e-10, max_iterations=10000):
        vt = self.A.vec_dtype
        db = DeviceArray.from_host(np.ascontiguousarray(b, vt))
        dx = DeviceArray(self.A.y_len, vt)
        info = self.solve(db, dx, eps, max_iterations)
        return dx.download(), info

    def destroy(self):
        if self.handle is not None:
            lib().bmfp_bicgstab_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Refine:
    """bmfp_refine_*: FP64 iterative refinement around an FP32 / FP16 BiCGStab."""

    def __init__(self, A64, inner_dtype=F32, use_jacobi=False):
        self.A = A64
        h = ctypes.c_void_p()
        check(lib().bmfp_refine_create(ctypes.byref(h), A64.handle, inner_dtype, 1 if use_jacobi else 0))
        self.handle = h
        self.inner_dtype = inner_dtype

    def solve(self, b, x, eps=1e-10, inner_tol=1e-5, max_outer=50, max_inner=10000, stream=None):
        info = SolveInfo()
        st = lib().bmfp_refine_solve(self.handle, b.ptr, x.ptr, eps, inner_tol, max_outer, max_inner, stream,
                                     ctypes.byref(info))
        if st not in SOLVE_STATUSES:
            check(st)
        return info

    def copy_seconds(self, inner_dtype=None):
        """Device time of the handle's copy of A in inner_dtype (default: the requested one)."""
        t = ctypes.c_double(0)
        check(lib().bmfp_refine_copy_seconds(self.handle, self.inner_dtype if inner_dtype is None else inner_dtype,
                                             ctypes.byref(t)))
        return t.value

    def solve_host(self, b, eps=1e-10, inner_tol=1e-5, max_outer=50, max_inner=10000):
        db = DeviceArray.from_host(np.ascontiguousarray(b, np.float64))
        dx = DeviceArray(self.A.y_len, np.float64)
        info = self.solve(db, dx, eps, inner_tol, max_outer, max_inner)
        return dx.download(), info

    def destroy(self):
        if self.handle is not None:
            lib().bmfp_refine_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


# ---- serving ------------------------------------------------------------------------
class ServerConfig(ctypes.Structure):
    _fields_ = [("memory_budget_bytes", ctypes.c_size_t), ("max_systems", ctypes.c_int),
                ("iterations_per_round", ctypes.c_int), ("slot_groups", ctypes.c_int),
                ("inner_tol", ctypes.c_double), ("max_outer", ctypes.c_int), ("max_inner", ctypes.c_int),
                ("chunks_per_tile", ctypes.c_int), ("use_jacobi", ctypes.c_int),
                ("inner_f16_min_bs", ctypes.c_int), ("inner_tol_f16", ctypes.c_double)]


class Request(ctypes.Structure):
    _fields_ = [("n_rows", ctypes.c_int), ("bs", ctypes.c_int), ("nnzb", ctypes.c_int),
                ("row_ptr", ctypes.c_void_p), ("columns", ctypes.c_void_p), ("values", ctypes.c_void_p),
                ("stencil_nx", ctypes.c_int), ("stencil_ny", ctypes.c_int), ("stencil_kind", ctypes.c_int),
                ("stencil_seed", ctypes.c_uint64), ("nrhs", ctypes.c_int), ("b", ctypes.c_void_p),
                ("rhs_seed", ctypes.c_uint64), ("tol", ctypes.c_double)]


class RequestResult(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("status", ctypes.c_int), ("nrhs", ctypes.c_int), ("n", ctypes.c_int64),
                ("outer_iterations", ctypes.c_int), ("inner_iterations", ctypes.c_int),
                ("rel_residual", ctypes.c_double), ("submit_time", ctypes.c_double),
                ("admit_time", ctypes.c_double), ("done_time", ctypes.c_double)]


def _declare_server(L):
    vp = ctypes.c_void_p
    ip = ctypes.POINTER(ctypes.c_int)
    L.bmfp_server_default_config.argtypes = [ctypes.POINTER(ServerConfig)]
    L.bmfp_server_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(ServerConfig)]
    L.bmfp_server_destroy.argtypes = [vp]
    L.bmfp_server_request_bytes.argtypes = [ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_size_t)]
    L.bmfp_server_submit.argtypes = [vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]
    L.bmfp_server_step.argtypes = [vp, ip, ip]
    L.bmfp_server_poll.argtypes = [vp, ctypes.POINTER(RequestResult), vp, ctypes.c_size_t, ip]
    L.bmfp_server_peek.argtypes = [vp, ctypes.POINTER(RequestResult), ip]
    L.bmfp_server_schedule.argtypes = [vp, ip, ip, ip]
    L.bmfp_server_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t), ip, ip,
                                    ctypes.POINTER(ctypes.c_double)]


class Server:
    """bmfp_server_*: continuous-batching BiCGStab request loop."""

    def __init__(self, **overrides):
        L = lib()
        if not getattr(L, "_server_declared", False):
            _declare_server(L)
            L._server_declared = True
        cfg = ServerConfig()
        check(L.bmfp_server_default_config(ctypes.byref(cfg)))
        for k, v in overrides.items():
            setattr(cfg, k, v)
        h = ctypes.c_void_p()
        check(L.bmfp_server_create(ctypes.byref(h), ctypes.byref(cfg)))
        self.handle = h
        self._keep = {}

    @staticmethod
    def stencil_request(nx, ny, bs, kind=0, seed=0, nrhs=1, rhs_seed=0, tol=1e-10, b=None):
        q = Request()
        q.bs, q.stencil_nx, q.stencil_ny, q.stencil_kind, q.stencil_seed = bs, nx, ny, kind, seed
        q.nrhs, q.rhs_seed, q.tol = nrhs, rhs_seed, tol
        keep = []
        if b is not None:
            bb = np.ascontiguousarray(b, np.float64)
            q.b = bb.ctypes.data
            keep.append(bb)
        return q, keep

    @staticmethod
    def bcsr_request(row_ptr, columns, values, n_rows, bs, nrhs=1, b=None, rhs_seed=0, tol=1e-10):
        rp = np.ascontiguousarray(row_ptr, np.int32)
        cols = np.ascontiguousarray(columns, np.int32)
        vals = np.ascontiguousarray(values, np.float64)
        q = Request()
        q.n_rows, q.bs, q.nnzb = n_rows, bs, int(rp[-1])
        q.row_ptr, q.columns, q.values = rp.ctypes.data, cols.ctypes.data, vals.ctypes.data
        q.nrhs, q.rhs_seed, q.tol = nrhs, rhs_seed, tol
        keep = [rp, cols, vals]
        if b is not None:
            bb = np.ascontiguousarray(b, np.float64)
            q.b = bb.ctypes.data
            keep.append(bb)
        return q, keep

    def submit(self, req_keep):
        q, keep = req_keep
        rid = ctypes.c_uint64(0)
        check(lib().bmfp_server_submit(self.handle, ctypes.byref(q), ctypes.byref(rid)))
        return rid.value

    @staticmethod
    def request_bytes(req_keep):
        nb = ctypes.c_size_t(0)
        check(lib().bmfp_server_request_bytes(ctypes.byref(req_keep[0]), ctypes.byref(nb)))
        return nb.value

    def step(self):
        a, q = ctypes.c_int(0), ctypes.c_int(0)
        check(lib().bmfp_server_step(self.handle, ctypes.byref(a), ctypes.byref(q)))
        return a.value, q.value

    def poll(self, want_x=True):
        """Returns (RequestResult, x[nrhs, n] or None) for one finished request, or None."""
        res = RequestResult()
        got = ctypes.c_int(0)
        check(lib().bmfp_server_peek(self.handle, ctypes.byref(res), ctypes.byref(got)))
        if not got.value:
            return None
        x = np.empty(res.nrhs * res.n if want_x else 1, np.float64)
        check(lib().bmfp_server_poll(self.handle, ctypes.byref(res), x.ctypes.data if want_x else None,
                                     x.size, ctypes.byref(got)))
        return res, (x.reshape(res.nrhs, res.n) if want_x else None)

    def schedule(self):
        """(slot_groups, xcd_run, {bs: chunks per tile}) from the configuration or the autotuner.

        slot_groups > 0 is a persistent grid of that many workgroups; -1 means one
        wavefront per tile, the grid following the tile table each round."""
        g, seg = ctypes.c_int(0), ctypes.c_int(0)
        per = (ctypes.c_int * 9)()
        check(lib().bmfp_server_schedule(self.handle, ctypes.byref(g), ctypes.byref(seg), per))
        return g.value, seg.value, {bs: per[bs] for bs in range(9) if per[bs]}

    def stats(self):
        used, cap = ctypes.c_size_t(0), ctypes.c_size_t(0)
        act, q = ctypes.c_int(0), ctypes.c_int(0)
        busy = ctypes.c_double(0)
        check(lib().bmfp_server_stats(self.handle, ctypes.byref(used), ctypes.byref(cap), ctypes.byref(act),
                                      ctypes.byref(q), ctypes.byref(busy)))
        return {"used": used.value, "capacity": cap.value, "active": act.value, "queued": q.value,
                "busy_seconds": busy.value}

    def drain(self, max_steps=1000000):
        """Steps until nothing is active or queued; returns {id: (result, x)}."""
        out = {}
        for _ in range(max_steps):
            a, q = self.step()
            while True:
                r = self.poll()
                if r is None:
                    break
                out[r[0].id] = r
            if a == 0 and q == 0:
                break
        return out

    def destroy(self):
        if self.handle is not None:
            lib().bmfp_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
