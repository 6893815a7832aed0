// bicgstab.h -- drop-in for the reference's gpu/bicgstab.h:12-47: the
// gpu_bicgstab<T, C> class keeps its constructor and solve() signatures; the
// solve runs libbmfp's fused device-resident BiCGStab (include/bmfp.h).
//
// Behaviour kept: the constructor throws std::runtime_error when the device
// memory the solver needs exceeds 80% of the device's memory
// (bicgstab.cu:30-45); solve() uploads A and b, starts from x0 = 0, iterates
// until the recurrence residual ||r||/||b|| < epsilon or max_iterations, and
// returns a pointer to pinned host memory owned by the solver.
// Behaviour changed: the returned x is the current iterate (the reference
// returns buffer 0, a stale iterate after an odd number of iterations);
// breakdowns stop the solve with the last valid iterate instead of NaNs;
// last_info() reports iterations, status and residuals.
// Extension: a constructor and solve() taking a BCSR matrix, which keeps its
// blocks (the CSR form runs as 1x1 blocks).
#ifndef BMFP_BICGSTAB_H
#define BMFP_BICGSTAB_H

#include <cstddef>
#include <memory>

#include "bmfp.h"
#include "matrix_converters.h"

template <class T, class C = std::size_t>
class gpu_bicgstab
{
public:
  gpu_bicgstab() = delete;
  explicit gpu_bicgstab(const csr_matrix_class<T, C> &A_b, bool use_precond);
  explicit gpu_bicgstab(const bcsr_matrix_class<T, C> &A_b, bool use_precond);
  ~gpu_bicgstab();
  gpu_bicgstab(const gpu_bicgstab &) = delete;
  gpu_bicgstab &operator=(const gpu_bicgstab &) = delete;

  T *solve(const csr_matrix_class<T, C> &A, const T *b, T epsilon, unsigned int max_iterations);
  T *solve(const bcsr_matrix_class<T, C> &A, const T *b, T epsilon, unsigned int max_iterations);

  const bmfp_solve_info &last_info() const { return info; }

private:
  T *run(bmfp_matrix m, const T *b, T epsilon, unsigned int max_iterations);
  void reserve(size_t rows, size_t nnz_values, size_t bs);

  unsigned int n_rows = 0;
  const bool use_precond = false;
  T *h_x = nullptr;           // pinned host result
  void *d_b = nullptr;        // device right-hand side
  void *d_x = nullptr;        // device solution
  bmfp_solve_info info{};
};

#endif
