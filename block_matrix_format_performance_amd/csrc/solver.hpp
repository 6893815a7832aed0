// solver.hpp -- BiCGStab solver handle shared by the plain solver, the
// mixed-precision refinement loop and the serving loop.
#pragma once

#include "matrix.hpp"

namespace bmfp {
struct SolverState;
struct RefineState;
void refine_destroy(RefineState *R);
}

struct bmfp_solver_t {
  bmfp_matrix_t *A = nullptr;   // borrowed
  int use_jacobi = 0;
  int np = 1;                   // persistent workgroups of every solver kernel
  int split = 0;                // 1: five-kernel split schedule, 0: three fused kernels
  int64_t n = 0;                // scalar rows
  bmfp::DeviceBuffer vecs;      // x r r^ p0 p1 v0 v1 s t [pinv]
  bmfp::DeviceBuffer parts;     // partial sums
  bmfp::DeviceBuffer state;     // SolverState
  bmfp::SolverState *h_state = nullptr;  // pinned mirror of the device state
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_sync = nullptr;
  hipGraphExec_t graphs[4] = {};  // 1, 2, 4 and 8 iterations
  int (*ops_init)(const bmfp_solver_t &, const void *, double, int, hipStream_t) = nullptr;
  int (*ops_iterate)(const bmfp_solver_t &, hipStream_t) = nullptr;
  int (*ops_jacobi)(const bmfp_solver_t &, hipStream_t) = nullptr;
  int (*spmv_groups)(const bmfp_solver_t &) = nullptr;  // split schedule's SpMV grid
  bmfp::RefineState *refine = nullptr;  // set for bmfp_refine_* handles only
  ~bmfp_solver_t();
};

namespace bmfp {
size_t solver_workspace_bytes(const bmfp_matrix_t &A, int use_jacobi);
// Block rows per pass of the solver's pass-sliced engine for (dtype, bs), 0 if
// it has none (BMFP_F32 / BMFP_F16, bs 4, 6, 8).
int sliced_pass_rows(int dtype, int bs);
int solver_create(bmfp_solver_t &S, bmfp_matrix_t *A, int use_jacobi);
// first_batch > 0: iterations launched before the first read of the solver
// state (a caller's prediction, e.g. the previous solve's count + 1); 0: 8.
// timed = false: no closing event synchronisation and no info->seconds (the
// solution copy is still ordered on the solver's stream).
int solver_run(bmfp_solver_t &S, const void *b, void *x, double eps, int max_iterations, hipStream_t user,
               bmfp_solve_info *info, bool true_residual = true, int first_batch = 0, bool timed = true);
}  // namespace bmfp
