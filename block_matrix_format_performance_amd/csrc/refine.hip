// refine.hip -- mixed-precision iterative refinement (bmfp_refine_*): FP64
// outer loop on the FP64 matrix, FP32-vector BiCGStab inner solves on an FP32
// or FP16+scale copy of the matrix, escalating FP16 -> FP32 on stagnation.
#include <algorithm>
#include <cmath>
#include <vector>

#include "matrix.hpp"
#include "solver.hpp"
#include "solver_kernels.hpp"

namespace bmfp {

namespace {

// Refinement step 0 (x = 0): r = b, so no matrix read; r32 = (float) b and
// partial sums of (b, b) in both slots, as the residual kernel writes them.
__global__ __launch_bounds__(256) void k_refine_first(const double *__restrict__ b, int64_t n,
                                                      float *__restrict__ r32, double *__restrict__ part)
{
  __shared__ double lds[2][4];
  double acc[2] = {0.0, 0.0};
  for (int64_t i = (int64_t) blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t) gridDim.x * 256) {
    const double bi = b[i];
    r32[i] = (float) bi;
    acc[0] += bi * bi;
  }
  acc[1] = acc[0];
  write_partials<2>(part, acc, lds);
}

// Refinement step k >= 1, first half: x_new = x_old + d.
__global__ __launch_bounds__(256) void k_refine_update(int64_t n, const double *__restrict__ x_old,
                                                       const float *__restrict__ d, double *__restrict__ x_new)
{
  const int64_t i = (int64_t) blockIdx.x * 256 + threadIdx.x;
  if (i < n) x_new[i] = x_old[i] + (double) d[i];
}

// Outer decision of the refinement loop, taken on the device: reduces the
// residual kernel's partial sums (fixed order), forms ||r_k|| / ||b|| and
// decides converged / non-finite / escalate FP16 -> FP32 (the residual
// shrank by less than 2x in a step, from step 2) / continue.  The host reads
// this record and acts on `decision`.
enum : int { REFINE_CONTINUE = 0, REFINE_CONVERGED = 1, REFINE_NONFINITE = 2, REFINE_ESCALATE = 3 };
struct RefineCtl {
  double norm_b;    // ||b||, from step 0
  double rel;       // ||r_k|| / ||b||
  double prev_rel;  // the previous step's rel
  double eps;       // outer tolerance
  int decision;
  int pad[3];
};

__global__ __launch_bounds__(256) void k_refine_decide(const double *__restrict__ part, int np, int k, int fp16,
                                                       RefineCtl *__restrict__ c)
{
  __shared__ double lds[2][4];
  double red[2];
  reduce_partials<2>(part, np, red, lds);
  if (threadIdx.x == 0) {
    if (k == 0) {
      c->norm_b = sqrt(red[1]);
      c->prev_rel = INFINITY;
    }
    const double rel = c->norm_b > 0 ? sqrt(red[0]) / c->norm_b : 0.0;
    int d = REFINE_CONTINUE;
    if (!isfinite(rel)) d = REFINE_NONFINITE;
    else if (rel < c->eps) d = REFINE_CONVERGED;
    else if (fp16 && k >= 2 && rel > 0.5 * c->prev_rel) d = REFINE_ESCALATE;
    c->rel = rel;
    c->prev_rel = rel;
    c->decision = d;
  }
}

}  // namespace

struct RefineState {
  bmfp_matrix_t *A = nullptr;                // FP64 matrix (borrowed)
  int use_jacobi = 0;
  int dtype = BMFP_F32;                      // current inner storage type
  int dtype0 = BMFP_F32;                     // requested inner storage type
  std::unique_ptr<bmfp_matrix_t> low[3];     // indexed by dtype (F32, F16 used)
  std::unique_ptr<bmfp_solver_t> inner[3];
  DeviceBuffer x;                            // two FP64 iterates (ping-pong)
  DeviceBuffer r32, d32;                     // FP32 residual / correction
  DeviceBuffer part;                         // residual partial sums
  DeviceBuffer ctl;                          // RefineCtl: the outer decision
  RefineCtl *h_ctl = nullptr;                // pinned mirror
  int np = 1;                                // workgroups of the step-0 pass
  int resid_groups = 1;                      // workgroups of the residual SpMV
  DeviceBuffer partW;                        // their partial sums
  int64_t n = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_sync = nullptr;
  int (*resid)(const RefineState &, const double *b, int xin, hipStream_t) = nullptr;
  double copy_seconds[3] = {};               // device time of each inner copy (by dtype)
  ~RefineState()
  {
    if (stream) (void) hipStreamSynchronize(stream);
    if (h_ctl) (void) hipHostFree(h_ctl);
    if (ev0) (void) hipEventDestroy(ev0);
    if (ev1) (void) hipEventDestroy(ev1);
    if (ev_sync) (void) hipEventDestroy(ev_sync);
    if (stream) (void) hipStreamDestroy(stream);
  }
};

namespace {

// The FP64 residual: x_new = x_old + d as a vector pass, then the SpMV of
// the stored x_new with the split schedule's engine (on the 2048x1024 bs 4
// system 30 + 500 us against 578 us forming x_old + d at every gathered
// index), its per-workgroup partials summed by REDUCE_GROUPS workgroups.
template <int BS, int LAYOUT>
struct Resid {
  using Engine = SplitEngine<double, double, BS, LAYOUT, false>;
  static int groups(const bmfp_matrix_t &A)
  {
    if constexpr (one_tile<Engine>::value)
      return (int) std::max<int64_t>(1, Engine::groups1(A.n_rows));
    const int64_t t = Engine::threads(A.n_rows);
    return (int) std::max<int64_t>(1, std::min<int64_t>((t + 255) / 256, SOLVER_MAX_GROUPS));
  }
  static int run(const RefineState &R, const double *b, int xin, hipStream_t s)
  {
    MatArgs<double, double> M;
    M.n_brows = R.A->n_rows;
    M.rp = R.A->row_ptr->as<int>();
    M.cols = R.A->columns->as<int>();
    M.vals = R.A->values.as<double>();
    M.scale = nullptr;
    const double *xo = R.x.as<double>() + (int64_t) xin * R.n;
    double *xn = R.x.as<double>() + (int64_t) (1 - xin) * R.n;
    k_refine_update<<<(unsigned) ((R.n + 255) / 256), 256, 0, s>>>(R.n, xo, R.d32.as<float>(), xn);
    BMFP_LAUNCHED();
    k_refine_resid_split<Engine><<<R.resid_groups, 256, 0, s>>>(M, xn, b, R.r32.as<float>(),
                                                                  R.partW.as<double>());
    BMFP_LAUNCHED();
    k_bicg_reduce<2><<<REDUCE_GROUPS, 256, 0, s>>>(R.partW.as<double>(), R.resid_groups, R.part.as<double>(),
                                                   nullptr);
    BMFP_LAUNCHED();
    return BMFP_OK;
  }
};

template <int LAYOUT>
bool pick(RefineState &R)
{
#define BMFP_RESID_BS(B)                         \
  case B:                                        \
    R.resid = Resid<B, LAYOUT>::run;             \
    R.resid_groups = Resid<B, LAYOUT>::groups(*R.A); \
    return true;
  switch (R.A->bs) {
    BMFP_RESID_BS(1)
    BMFP_RESID_BS(2)
    BMFP_RESID_BS(3)
    BMFP_RESID_BS(4)
    BMFP_RESID_BS(6)
    BMFP_RESID_BS(8)
    BMFP_RESID_BS(12)
    BMFP_RESID_BS(16)
    BMFP_RESID_BS(24)
    BMFP_RESID_BS(32)
  default:
    return false;
  }
#undef BMFP_RESID_BS
}

// Creates the inner matrix copy and solver for `dtype` on first use.
// The inner copy of the matrix in `dtype`: pass-sliced for the split
// schedule's engine where one exists (row-major bs 4, 6, 8 without Jacobi,
// at least 32 MB of inner values, padding within a quarter of the blocks),
// else BCSR order.
int ensure_level(RefineState &R, int dtype)
{
  if (R.inner[dtype]) return BMFP_OK;
  const bmfp_matrix_t &A = *R.A;
  const int rpw = sliced_pass_rows(dtype, A.bs);
  const int J = (A.max_row_blocks + 2) / 3 * 3;
  const int64_t nsb = ((int64_t) A.n_rows + std::max(rpw, 1) - 1) / std::max(rpw, 1) * rpw * J;
  const bool slice = rpw > 0 && A.layout == BMFP_ROW_MAJOR && !R.use_jacobi && J > 0 &&
                     (double) A.n_vals() * dtype_size(dtype) >= 32e6 && nsb <= (int64_t) A.nnzb + A.nnzb / 4;
  if (slice) {
    std::unique_ptr<bmfp_matrix_t> m(new bmfp_matrix_t);
    m->is_bcsr = 1;
    m->dtype = dtype;
    m->layout = A.layout;
    m->n_rows = A.n_rows;
    m->n_cols = A.n_cols;
    m->bs = A.bs;
    m->nnzb = A.nnzb;
    m->max_row_blocks = A.max_row_blocks;
    m->mean_row_blocks = A.mean_row_blocks;
    m->device = A.device;
    m->row_ptr = A.row_ptr;
    m->columns = A.columns;
    BMFP_TRY(make_sliced_copy(A, dtype, rpw, J, *m, R.stream, R.ev0, R.ev1));
    R.low[dtype].reset(m.release());
  } else {
    // (the span includes the copy's allocation)
    BMFP_HIP(hipEventRecord(R.ev0, R.stream));
    bmfp_matrix out = nullptr;
    BMFP_TRY(bmfp_matrix_convert(R.A, dtype, R.A->layout, &out, R.stream));
    R.low[dtype].reset(out);
    BMFP_HIP(hipEventRecord(R.ev1, R.stream));
  }
  BMFP_HIP(hipEventSynchronize(R.ev1));
  float ms = 0.f;
  BMFP_HIP(hipEventElapsedTime(&ms, R.ev0, R.ev1));
  R.copy_seconds[dtype] = ms * 1e-3;
  BMFP_HIP(hipStreamSynchronize(R.stream));
  std::unique_ptr<bmfp_solver_t> S(new bmfp_solver_t);
  BMFP_TRY(solver_create(*S, R.low[dtype].get(), R.use_jacobi));
  R.inner[dtype] = std::move(S);
  return BMFP_OK;
}

}  // namespace

int refine_create(bmfp_solver_t &H, bmfp_matrix_t *A, int inner_dtype, int use_jacobi)
{
  if (!A->is_bcsr || A->dtype != BMFP_F64) return fail(BMFP_ERR_INVALID, "refinement needs an F64 BCSR matrix");
  if (A->n_rows != A->n_cols) return fail(BMFP_ERR_INVALID, "refinement needs a square matrix");
  if (inner_dtype != BMFP_F32 && inner_dtype != BMFP_F16)
    return fail(BMFP_ERR_INVALID, "inner_dtype must be BMFP_F32 or BMFP_F16");
  std::unique_ptr<RefineState> R(new RefineState);
  R->A = A;
  R->use_jacobi = use_jacobi ? 1 : 0;
  R->dtype = R->dtype0 = inner_dtype;
  R->n = A->y_len();
  const bool ok = A->layout == BMFP_ROW_MAJOR ? pick<BMFP_ROW_MAJOR>(*R) : pick<BMFP_COL_MAJOR>(*R);
  if (!ok) return fail(BMFP_ERR_UNSUPPORTED, "no refinement kernels for bs %d", A->bs);
  const size_t n = (size_t) std::max<int64_t>(R->n, 1);
  BMFP_TRY(R->x.alloc(2 * n * sizeof(double)));
  BMFP_TRY(R->r32.alloc(n * sizeof(float)));
  BMFP_TRY(R->d32.alloc(n * sizeof(float)));
  R->np = (int) std::max<int64_t>(1, std::min<int64_t>((R->n + 255) / 256, SOLVER_MAX_GROUPS));
  BMFP_TRY(R->part.alloc(2 * SOLVER_MAX_GROUPS * sizeof(double)));
  BMFP_TRY(R->partW.alloc(2 * (size_t) R->resid_groups * sizeof(double)));
  BMFP_TRY(R->ctl.alloc(sizeof(RefineCtl)));
  BMFP_HIP(hipHostMalloc(reinterpret_cast<void **>(&R->h_ctl), sizeof(RefineCtl), hipHostMallocDefault));
  BMFP_HIP(hipStreamCreateWithFlags(&R->stream, hipStreamNonBlocking));
  BMFP_HIP(hipEventCreate(&R->ev0));
  BMFP_HIP(hipEventCreate(&R->ev1));
  BMFP_HIP(hipEventCreateWithFlags(&R->ev_sync, hipEventDisableTiming));
  BMFP_TRY(ensure_level(*R, inner_dtype));
  H.refine = R.release();
  return BMFP_OK;
}

void refine_destroy(RefineState *R) { delete R; }

int refine_run(RefineState &R, const double *b, double *x, double eps, double inner_tol, int max_outer,
               int max_inner, hipStream_t user, bmfp_solve_info *info)
{
  if (!(eps > 0.0) || !(inner_tol > 0.0) || max_outer < 1 || max_inner < 1)
    return fail(BMFP_ERR_INVALID, "refinement: eps, inner_tol > 0 and max_outer, max_inner >= 1 required");
  bmfp_solve_info out{};
  out.inner_dtype = R.dtype0;
  if (R.n == 0) {
    if (info) *info = out;
    return BMFP_OK;
  }
  if (!b || !x) return fail(BMFP_ERR_INVALID, "NULL vector");
  hipStream_t s = R.stream;
  BMFP_HIP(hipEventRecord(R.ev_sync, user));
  BMFP_HIP(hipStreamWaitEvent(s, R.ev_sync, 0));
  BMFP_HIP(hipEventRecord(R.ev0, s));
  // each solve starts from the requested inner precision (an escalation
  // lasts for the rest of its own solve only)
  R.dtype = R.dtype0;
  out.inner_dtype = R.dtype;
  BMFP_HIP(hipMemsetAsync(R.x.ptr, 0, (size_t) R.n * sizeof(double), s));
  *R.h_ctl = RefineCtl{};
  R.h_ctl->eps = eps;
  BMFP_HIP(hipMemcpyAsync(R.ctl.ptr, R.h_ctl, sizeof(RefineCtl), hipMemcpyHostToDevice, s));
  int xin = 0;
  int status = BMFP_ERR_NOT_CONVERGED;
  int hint = 0;  // launch batch of the next inner solve: the previous one's iterations + 1
  for (int k = 0; k <= max_outer; k++) {
    if (k == 0) {
      // x_0 = 0: r = b without reading the matrix
      k_refine_first<<<R.np, 256, 0, s>>>(b, R.n, R.r32.as<float>(), R.part.as<double>());
      BMFP_LAUNCHED();
    } else {
      // x_k = x_{k-1} + d, r = b - A x_k (FP64); x_k lands in buffer 1 - xin
      BMFP_TRY(R.resid(R, b, xin, s));
      xin = 1 - xin;
    }
    k_refine_decide<<<1, 256, 0, s>>>(R.part.as<double>(), k == 0 ? R.np : REDUCE_GROUPS, k,
                                      R.dtype == BMFP_F16 ? 1 : 0, R.ctl.as<RefineCtl>());
    BMFP_LAUNCHED();
    BMFP_HIP(hipMemcpyAsync(R.h_ctl, R.ctl.ptr, sizeof(RefineCtl), hipMemcpyDeviceToHost, s));
    BMFP_HIP(hipStreamSynchronize(s));
    const RefineCtl &c = *R.h_ctl;
    out.rel_residual = out.true_rel_residual = c.rel;
    out.outer_iterations = k;
    if (c.decision == REFINE_NONFINITE) {
      status = BMFP_ERR_NONFINITE;
      break;
    }
    if (c.decision == REFINE_CONVERGED) {
      status = BMFP_OK;
      break;
    }
    if (k == max_outer) break;
    // FP16 inner solves that stopped paying off escalate to FP32
    if (c.decision == REFINE_ESCALATE) {
      BMFP_TRY(ensure_level(R, BMFP_F32));
      R.dtype = BMFP_F32;
      out.escalations++;
    }
    bmfp_solve_info in{};
    int st = solver_run(*R.inner[R.dtype], R.r32.ptr, R.d32.ptr, inner_tol, max_inner, s, &in, false, hint, false);
    out.inner_iterations += in.iterations;
    if ((st == BMFP_ERR_BREAKDOWN || st == BMFP_ERR_NONFINITE) && R.dtype == BMFP_F16) {
      BMFP_TRY(ensure_level(R, BMFP_F32));
      R.dtype = BMFP_F32;
      out.escalations++;
      st = solver_run(*R.inner[R.dtype], R.r32.ptr, R.d32.ptr, inner_tol, max_inner, s, &in, false, 0, false);
      out.inner_iterations += in.iterations;
    }
    hint = in.iterations + 1;
    if (st == BMFP_ERR_NONFINITE) {
      status = st;
      break;
    }
    if (st != BMFP_OK && st != BMFP_ERR_NOT_CONVERGED && st != BMFP_ERR_BREAKDOWN) return st;
    // a breakdown still leaves the last valid inner iterate in d
  }
  BMFP_HIP(hipMemcpyAsync(x, R.x.as<double>() + (int64_t) xin * R.n, (size_t) R.n * sizeof(double),
                          hipMemcpyDeviceToDevice, s));
  BMFP_HIP(hipEventRecord(R.ev1, s));
  BMFP_HIP(hipEventSynchronize(R.ev1));
  float ms = 0.f;
  BMFP_HIP(hipEventElapsedTime(&ms, R.ev0, R.ev1));
  out.seconds = ms * 1e-3;
  out.iterations = out.inner_iterations;
  out.status = status;
  out.inner_dtype = R.dtype;
  if (info) *info = out;
  return status;
}

}  // namespace bmfp

using namespace bmfp;

extern "C" int bmfp_refine_create(bmfp_solver *out, bmfp_matrix A64, int inner_dtype, int use_jacobi)
{
  return bmfp::guarded("bmfp_refine_create", [&]() -> int {
    if (!out || !A64) return fail(BMFP_ERR_INVALID, "NULL argument");
    *out = nullptr;
    std::unique_ptr<bmfp_solver_t> H(new bmfp_solver_t);
    BMFP_TRY(refine_create(*H, A64, inner_dtype, use_jacobi));
    *out = H.release();
    return BMFP_OK;
  });
}

extern "C" int bmfp_refine_solve(bmfp_solver s, const double *b, double *x, double eps, double inner_tol,
                                 int max_outer, int max_inner, bmfp_stream stream, bmfp_solve_info *info)
{
  return bmfp::guarded("bmfp_refine_solve", [&]() -> int {
    if (!s || !s->refine) return fail(BMFP_ERR_INVALID, "not a refinement handle");
    const int st = refine_run(*s->refine, b, x, eps, inner_tol, max_outer, max_inner,
                              static_cast<hipStream_t>(stream), info);
    if (st == BMFP_ERR_NOT_CONVERGED) set_error("refinement: max_outer (%d) reached", max_outer);
    return st;
  });
}

extern "C" int bmfp_refine_copy_seconds(bmfp_solver s, int inner_dtype, double *seconds)
{
  return bmfp::guarded("bmfp_refine_copy_seconds", [&]() -> int {
    if (!s || !s->refine || !seconds) return fail(BMFP_ERR_INVALID, "not a refinement handle");
    if (inner_dtype != BMFP_F32 && inner_dtype != BMFP_F16) return fail(BMFP_ERR_INVALID, "bad inner_dtype");
    *seconds = s->refine->copy_seconds[inner_dtype];
    return BMFP_OK;
  });
}

extern "C" int bmfp_refine_destroy(bmfp_solver s)
{
  return bmfp::guarded("bmfp_refine_destroy", [&]() -> int {
    delete s;
    return BMFP_OK;
  });
}
