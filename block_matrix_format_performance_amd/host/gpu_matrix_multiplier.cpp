// gpu_matrix_multiplier.cpp -- the reference's SpMV entry points
// (gpu/gpu_matrix_multiplier.h:13-27) over libbmfp's C ABI.
#include "bmfp/gpu_matrix_multiplier.h"

#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "bmfp.h"

namespace {

void check(int st, const char *what)
{
  if (st != BMFP_OK) throw std::runtime_error(std::string(what) + ": " + bmfp_last_error());
}

// The device is initialised once per process (device 0; the reference
// hard-codes device 1, main.cpp:388).
void ensure_init()
{
  static int st = bmfp_init(0);
  check(st, "bmfp_init");
}

struct DevBuf {
  void *p = nullptr;
  explicit DevBuf(size_t bytes) { check(bmfp_malloc(&p, bytes), "bmfp_malloc"); }
  ~DevBuf() { bmfp_free(p); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
};

struct Handle {
  bmfp_matrix m = nullptr;
  ~Handle() { bmfp_matrix_destroy(m); }
};

template <class T> constexpr int dtype_of() { return sizeof(T) == 8 ? BMFP_F64 : BMFP_F32; }

template <class I> std::vector<int> to_int(const I *a, size_t n)
{
  std::vector<int> out(n);
  for (size_t i = 0; i < n; i++) out[i] = (int) a[i];
  return out;
}

template <class T> std::vector<double> to_double(const T *a, size_t n)
{
  return std::vector<double>(a, a + n);
}

constexpr int WARMUP = 1;
constexpr int REPS = 10;

// Times `variant` on A (mean of REPS launches) with x uploaded from x_host
// (or x = 1 when x_host is null), checks against reference_y.
template <class T>
measurement_class time_variant(bmfp_matrix A, int variant, const char *name, const T *x_host, size_t x_len,
                               size_t y_len, const T *reference_y, T *y_out)
{
  DevBuf dx(x_len * sizeof(T)), dy(y_len * sizeof(T));
  if (x_host)
    check(bmfp_memcpy_h2d(dx.p, x_host, x_len * sizeof(T)), "upload x");
  else
    check(bmfp_fill(dx.p, dtype_of<T>(), x_len, 2, 0, 0, 1.0, nullptr), "fill x");
  double t = 0;
  const int st = bmfp_spmv_timed(A, variant, dx.p, dy.p, WARMUP, REPS, nullptr, &t);
  if (st == BMFP_ERR_UNSUPPORTED) {
    std::cerr << "skip " << name << ": " << bmfp_last_error() << std::endl;
    return measurement_class(name, std::numeric_limits<double>::quiet_NaN(), 0, 0);
  }
  check(st, name);
  std::vector<T> y(y_len);
  check(bmfp_memcpy_d2h(y.data(), dy.p, y_len * sizeof(T)), "download y");
  if (reference_y) compare_results(y_len, reference_y, y.data());
  if (y_out) std::copy(y.begin(), y.end(), y_out);
  double bytes = 0, flops = 0;
  check(bmfp_spmv_traffic(A, &bytes, &flops), "traffic");
  return measurement_class(name, t, bytes, flops);
}

template <class T, class I>
void csr_handle(Handle &h, const csr_matrix_class<T, I> &m)
{
  ensure_init();
  const auto rp = to_int(m.row_ptr.get(), (size_t) m.n_rows + 1);
  const auto cols = to_int(m.columns.get(), (size_t) m.nnz);
  const auto vals = to_double(m.values.get(), (size_t) m.nnz);
  check(bmfp_csr_create(&h.m, dtype_of<T>(), (int) m.n_rows, (int) m.n_cols, (int) m.nnz, rp.data(), cols.data(),
                        vals.data()),
        "bmfp_csr_create");
}

template <class T, class I>
void bcsr_handle(Handle &h, const bcsr_matrix_class<T, I> &m, const T *values, int layout)
{
  ensure_init();
  const auto rp = to_int(m.row_ptr.get(), (size_t) m.n_rows + 1);
  const auto cols = to_int(m.columns.get(), (size_t) m.nnzb);
  const auto vals = to_double(values, (size_t) m.nnzb * m.bs * m.bs);
  check(bmfp_bcsr_create(&h.m, dtype_of<T>(), layout, (int) m.n_rows, (int) m.n_cols, (int) m.bs, (int) m.nnzb,
                         rp.data(), cols.data(), vals.data()),
        "bmfp_bcsr_create");
}

struct Entry {
  const char *name;
  int layout;
  int variant;
};

// gpu_matrix_multiplier.cu:804-1317 order.  The library rows use rocSPARSE
// bsrmv where the reference used cuSPARSE; the "template", "fastdiv" and
// "po2" rows of the reference are specialisations of one kernel family here
// (every kernel is compiled per block size) and time the same family.
const Entry BCSR_ENTRIES[] = {
    {"GPU BCSR (row major, thread per row)", BMFP_ROW_MAJOR, BMFP_BCSR_THREAD_PER_ROW},
    {"GPU BCSR (row major, warp per row)", BMFP_ROW_MAJOR, BMFP_BCSR_GROUP_PER_ROW},
    {"GPU BSR (rocSPARSE, row major)", BMFP_ROW_MAJOR, BMFP_BCSR_LIBRARY},
    {"GPU BSR (rocSPARSE, column major)", BMFP_COL_MAJOR, BMFP_BCSR_LIBRARY},
    {"GPU BCSR (column major, thread per row)", BMFP_COL_MAJOR, BMFP_BCSR_THREAD_PER_ROW},
    {"GPU BCSR (column major, thread per row, template)", BMFP_COL_MAJOR, BMFP_BCSR_THREAD_PER_ROW},
    {"GPU BCSR (column major, thread per row, coal x)", BMFP_COL_MAJOR, BMFP_BCSR_SHUFFLE_X},
    {"GPU BCSR (column major, thread per row, coal x, template)", BMFP_COL_MAJOR, BMFP_BCSR_SHUFFLE_X},
    {"GPU BCSR (column major, column-by-column)", BMFP_COL_MAJOR, BMFP_BCSR_BLOCK_ROW_GROUP},
    {"GPU BCSR (column major, column-by-column, int fastdiv)", BMFP_COL_MAJOR, BMFP_BCSR_BLOCK_ROW_GROUP},
    {"GPU BCSR (column major, column-by-column, int po2)", BMFP_COL_MAJOR, BMFP_BCSR_BLOCK_ROW_GROUP},
    {"GPU BCSR (column major, column-by-column, template)", BMFP_COL_MAJOR, BMFP_BCSR_BLOCK_ROW_GROUP},
};

}  // namespace

const std::vector<std::string> &gpu_bcsr_spmv_names()
{
  static const std::vector<std::string> names = [] {
    std::vector<std::string> v;
    for (const Entry &e : BCSR_ENTRIES) v.push_back(e.name);
    return v;
  }();
  return names;
}

template <typename T, typename I>
measurement_class gpu_csr_spmv(const csr_matrix_class<T, I> &m, const T *x, const T *reference_y, T *y_out)
{
  Handle h;
  csr_handle(h, m);
  return time_variant<T>(h.m, BMFP_CSR_SCALAR, "GPU CSR", x, (size_t) m.n_cols, (size_t) m.n_rows, reference_y, y_out);
}

template <typename T, typename I>
measurement_class gpu_csr_spmv(const csr_matrix_class<T, I> &m, const T *reference_y)
{
  return gpu_csr_spmv<T, I>(m, nullptr, reference_y, nullptr);
}

template <typename T, typename I>
measurement_class gpu_csr_vector_spmv(const csr_matrix_class<T, I> &m, const T *x, const T *reference_y, T *y_out)
{
  Handle h;
  csr_handle(h, m);
  return time_variant<T>(h.m, BMFP_CSR_VECTOR, "GPU CSR-Vector", x, (size_t) m.n_cols, (size_t) m.n_rows,
                         reference_y, y_out);
}

template <typename T, typename I>
measurement_class gpu_csr_vector_spmv(const csr_matrix_class<T, I> &m, const T *reference_y)
{
  return gpu_csr_vector_spmv<T, I>(m, nullptr, reference_y, nullptr);
}

template <typename T, typename I>
measurement_class gpu_csr_merge_spmv(const csr_matrix_class<T, I> &m, const T *x, const T *reference_y, T *y_out)
{
  Handle h;
  csr_handle(h, m);
  return time_variant<T>(h.m, BMFP_CSR_MERGE, "GPU Merge SpMV", x, (size_t) m.n_cols, (size_t) m.n_rows,
                         reference_y, y_out);
}

template <typename T, typename I>
std::vector<measurement_class> gpu_bcsr_spmv(bcsr_matrix_class<T, I> &m, const T *transposed, const T *x,
                                             const T *reference_y)
{
  Handle row, col;
  bcsr_handle(row, m, m.values.get(), BMFP_ROW_MAJOR);
  bcsr_handle(col, m, transposed, BMFP_COL_MAJOR);
  const size_t xn = (size_t) m.n_cols * m.bs, yn = (size_t) m.n_rows * m.bs;
  std::vector<measurement_class> out;
  for (const Entry &e : BCSR_ENTRIES)
    out.push_back(time_variant<T>(e.layout == BMFP_ROW_MAJOR ? row.m : col.m, e.variant, e.name, x, xn, yn,
                                  reference_y, nullptr));
  return out;
}

template <typename T, typename I>
std::vector<measurement_class> gpu_bcsr_spmv(bcsr_matrix_class<T, I> &m, const T *transposed, const T *reference_y)
{
  return gpu_bcsr_spmv<T, I>(m, transposed, nullptr, reference_y);
}

template <typename T, typename I>
measurement_class gpu_bcsr_jit_spmv(bcsr_matrix_class<T, I> &m, const T *transposed, const T *x, const T *reference_y)
{
  Handle col;
  bcsr_handle(col, m, transposed, BMFP_COL_MAJOR);
  const size_t xn = (size_t) m.n_cols * m.bs, yn = (size_t) m.n_rows * m.bs;
  DevBuf dx(xn * sizeof(T)), dy(yn * sizeof(T));
  if (x)
    check(bmfp_memcpy_h2d(dx.p, x, xn * sizeof(T)), "upload x");
  else
    check(bmfp_fill(dx.p, dtype_of<T>(), xn, 2, 0, 0, 1.0, nullptr), "fill x");
  // compile (hipRTC) and load, untimed, then one timed launch
  check(bmfp_spmv(col.m, BMFP_BCSR_JIT, dx.p, dy.p, nullptr), "jit compile");
  double t = 0;
  check(bmfp_spmv_timed(col.m, BMFP_BCSR_JIT, dx.p, dy.p, 0, 1, nullptr, &t), "jit");
  std::vector<T> y(yn);
  check(bmfp_memcpy_d2h(y.data(), dy.p, yn * sizeof(T)), "download y");
  if (reference_y) compare_results(yn, reference_y, y.data());
  double bytes = 0, flops = 0;
  check(bmfp_spmv_traffic(col.m, &bytes, &flops), "traffic");
  return measurement_class("jit", t, bytes, flops);
}

// Instantiations (reference: gpu_matrix_multiplier.cu:1333-1341).
#define BMFP_INSTANTIATE(T, I)                                                                                  \
  template measurement_class gpu_csr_spmv<T, I>(const csr_matrix_class<T, I> &, const T *);                   \
  template measurement_class gpu_csr_spmv<T, I>(const csr_matrix_class<T, I> &, const T *, const T *, T *);   \
  template measurement_class gpu_csr_vector_spmv<T, I>(const csr_matrix_class<T, I> &, const T *);            \
  template measurement_class gpu_csr_vector_spmv<T, I>(const csr_matrix_class<T, I> &, const T *, const T *,  \
                                                       T *);                                                    \
  template measurement_class gpu_csr_merge_spmv<T, I>(const csr_matrix_class<T, I> &, const T *, const T *, T *); \
  template std::vector<measurement_class> gpu_bcsr_spmv<T, I>(bcsr_matrix_class<T, I> &, const T *, const T *); \
  template std::vector<measurement_class> gpu_bcsr_spmv<T, I>(bcsr_matrix_class<T, I> &, const T *, const T *,  \
                                                              const T *);                                        \
  template measurement_class gpu_bcsr_jit_spmv<T, I>(bcsr_matrix_class<T, I> &, const T *, const T *, const T *);
BMFP_INSTANTIATE(float, int)
BMFP_INSTANTIATE(double, int)
#undef BMFP_INSTANTIATE

// Version tag of the C++ drop-in layer (lets tests check the library loads).
extern "C" int bmfp_cpp_abi_version(void) { return BMFP_VERSION; }
