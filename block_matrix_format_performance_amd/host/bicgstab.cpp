// bicgstab.cpp -- gpu_bicgstab<T, C> (reference: gpu/bicgstab.h:12-47,
// gpu/bicgstab.cu:23-531) over libbmfp's fused BiCGStab.
#include "bmfp/bicgstab.h"

#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

void check(int st, const char *what)
{
  if (st != BMFP_OK) throw std::runtime_error(std::string(what) + ": " + bmfp_last_error());
}

void ensure_init()
{
  static int st = bmfp_init(0);
  check(st, "bmfp_init");
}

bool quiet()
{
  const char *e = std::getenv("BMFP_QUIET");
  return e && *e && *e != '0';
}

template <class T> constexpr int dtype_of() { return sizeof(T) == 8 ? BMFP_F64 : BMFP_F32; }

template <class C> std::vector<int> to_int(const C *a, size_t n)
{
  std::vector<int> out(n);
  for (size_t i = 0; i < n; i++) out[i] = (int) a[i];
  return out;
}

struct Handle {
  bmfp_matrix m = nullptr;
  ~Handle() { bmfp_matrix_destroy(m); }
};

}  // namespace

template <class T, class C>
void gpu_bicgstab<T, C>::reserve(size_t rows, size_t nnz_values, size_t bs)
{
  ensure_init();
  size_t total = 0;
  check(bmfp_device_info(nullptr, 0, &total, nullptr, nullptr), "bmfp_device_info");
  // bicgstab.cu:30-45: vectors + matrix against 80% of device memory.
  const size_t nblocks = nnz_values / (bs * bs);
  const size_t required = rows * sizeof(T) * (use_precond ? 12 : 11)      // 9-10 solver vectors + b + x
                          + nnz_values * sizeof(T) + nblocks * 4 + (rows / bs + 1) * 4;
  if ((double) required > 0.8 * (double) total)
    throw std::runtime_error("Error! There is not enough memory on gpu (" + std::to_string(required / 1024 / 1024) +
                             "MB)!");
  if (!quiet()) std::cout << "Allocate " << (double) required / 1024 / 1024 << "MB on GPU" << std::endl;
  check(bmfp_malloc_host(reinterpret_cast<void **>(&h_x), rows * sizeof(T)), "bmfp_malloc_host");
  check(bmfp_malloc(&d_b, rows * sizeof(T)), "bmfp_malloc");
  check(bmfp_malloc(&d_x, rows * sizeof(T)), "bmfp_malloc");
}

template <class T, class C>
gpu_bicgstab<T, C>::gpu_bicgstab(const csr_matrix_class<T, C> &A, bool precond)
  : n_rows((unsigned) A.n_rows), use_precond(precond)
{
  reserve((size_t) A.n_rows, (size_t) A.nnz, 1);
}

template <class T, class C>
gpu_bicgstab<T, C>::gpu_bicgstab(const bcsr_matrix_class<T, C> &A, bool precond)
  : n_rows((unsigned) (A.n_rows * A.bs)), use_precond(precond)
{
  reserve((size_t) A.n_rows * A.bs, (size_t) A.nnzb * A.bs * A.bs, (size_t) A.bs);
}

template <class T, class C>
gpu_bicgstab<T, C>::~gpu_bicgstab()
{
  bmfp_free_host(h_x);
  bmfp_free(d_b);
  bmfp_free(d_x);
}

template <class T, class C>
T *gpu_bicgstab<T, C>::run(bmfp_matrix m, const T *b, T epsilon, unsigned int max_iterations)
{
  check(bmfp_memcpy_h2d(d_b, b, (size_t) n_rows * sizeof(T)), "upload b");
  bmfp_solver s = nullptr;
  check(bmfp_bicgstab_create(&s, m, use_precond ? 1 : 0), "bmfp_bicgstab_create");
  const int st = bmfp_bicgstab_solve(s, d_b, d_x, (double) epsilon, (int) max_iterations, nullptr, &info);
  bmfp_bicgstab_destroy(s);
  if (st != BMFP_OK && st != BMFP_ERR_NOT_CONVERGED && st != BMFP_ERR_BREAKDOWN && st != BMFP_ERR_NONFINITE)
    check(st, "bmfp_bicgstab_solve");
  check(bmfp_memcpy_d2h(h_x, d_x, (size_t) n_rows * sizeof(T)), "download x");
  if (!quiet())
    std::cout << "i: " << info.iterations << "; rhs norm: " << info.rel_residual << "; status: "
              << bmfp_status_string(info.status) << "\n\nCalculation complete in " << info.seconds << "s"
              << std::endl;
  return h_x;
}

template <class T, class C>
T *gpu_bicgstab<T, C>::solve(const csr_matrix_class<T, C> &A, const T *b, T epsilon, unsigned int max_iterations)
{
  if ((unsigned) A.n_rows != n_rows) throw std::invalid_argument("gpu_bicgstab::solve: size mismatch");
  Handle h;
  const auto rp = to_int(A.row_ptr.get(), (size_t) A.n_rows + 1);
  const auto cols = to_int(A.columns.get(), (size_t) A.nnz);
  const std::vector<double> vals(A.values.get(), A.values.get() + A.nnz);
  check(bmfp_bcsr_create(&h.m, dtype_of<T>(), BMFP_ROW_MAJOR, (int) A.n_rows, (int) A.n_cols, 1, (int) A.nnz,
                         rp.data(), cols.data(), vals.data()),
        "bmfp_bcsr_create");
  return run(h.m, b, epsilon, max_iterations);
}

template <class T, class C>
T *gpu_bicgstab<T, C>::solve(const bcsr_matrix_class<T, C> &A, const T *b, T epsilon, unsigned int max_iterations)
{
  if ((unsigned) (A.n_rows * A.bs) != n_rows) throw std::invalid_argument("gpu_bicgstab::solve: size mismatch");
  Handle h;
  const auto rp = to_int(A.row_ptr.get(), (size_t) A.n_rows + 1);
  const auto cols = to_int(A.columns.get(), (size_t) A.nnzb);
  const size_t nv = (size_t) A.nnzb * A.bs * A.bs;
  const std::vector<double> vals(A.values.get(), A.values.get() + nv);
  check(bmfp_bcsr_create(&h.m, dtype_of<T>(), BMFP_ROW_MAJOR, (int) A.n_rows, (int) A.n_cols, (int) A.bs,
                         (int) A.nnzb, rp.data(), cols.data(), vals.data()),
        "bmfp_bcsr_create");
  return run(h.m, b, epsilon, max_iterations);
}

// bicgstab.cu:528-531
template class gpu_bicgstab<float, int>;
template class gpu_bicgstab<double, int>;
template class gpu_bicgstab<float, unsigned int>;
template class gpu_bicgstab<double, unsigned int>;
