// bicgstab.hip -- BiCGStab solver handle (bmfp_bicgstab_*): resident Krylov
// vectors, device-side scalars, fused kernels replayed from a HIP graph.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "matrix.hpp"
#include "solver.hpp"
#include "solver_kernels.hpp"

namespace bmfp {

namespace {

constexpr int GRAPH_ITERS = 8;  // iterations captured per graph launch

template <class VT, class AT, bool SCALED>
MatArgs<VT, AT> mat_args(const bmfp_matrix_t &A)
{
  MatArgs<VT, AT> m;
  m.n_brows = A.n_rows;
  m.rp = A.row_ptr->as<int>();
  m.cols = A.columns->as<int>();
  m.vals = A.values.as<VT>();
  m.scale = SCALED ? A.scale.as<float>() : nullptr;
  m.cols_sl = A.sl_j > 0 ? A.cols_sl.as<int>() : nullptr;
  m.sl_j = A.sl_j;
  return m;
}

template <class AT>
SolverVecs<AT> vec_args(const bmfp_solver_t &S, const void *b)
{
  SolverVecs<AT> V;
  AT *base = S.vecs.as<AT>();
  const int64_t n = S.n;
  V.b = static_cast<const AT *>(b);
  V.x = base + 0 * n;
  V.r = base + 1 * n;
  V.rh = base + 2 * n;
  V.p[0] = base + 3 * n;
  V.p[1] = base + 4 * n;
  V.v[0] = base + 5 * n;
  V.v[1] = base + 6 * n;
  V.s = base + 7 * n;
  V.t = base + 8 * n;
  V.pinv = S.use_jacobi ? base + 9 * n : nullptr;
  double *parts = S.parts.as<double>();
  V.partA = parts;
  V.partB = parts + SOLVER_MAX_GROUPS;
  V.partC = parts + 3 * SOLVER_MAX_GROUPS;
  V.partW = S.split ? parts + 5 * SOLVER_MAX_GROUPS : nullptr;
  V.st = S.state.as<SolverState>();
  V.n = n;
  V.npa = V.npb = S.split ? REDUCE_GROUPS : S.np;
  return V;
}

// Block sizes with a pass-sliced split engine: FP32 / FP16 row-major 4, 6, 8
// (the refinement's inner copies).
template <class VT, int BS, int LAYOUT> constexpr bool sliceable()
{
  return sizeof(VT) < 8 && LAYOUT == BMFP_ROW_MAJOR && (BS == 4 || BS == 6 || BS == 8);
}

template <class VT, class AT, bool SCALED, int BS, int LAYOUT, bool SL = false>
struct Ops {
  using Engine = SolverEngine<VT, AT, BS, LAYOUT, SCALED>;
  using Split = typename std::conditional<SL, SlicedRowsEngine<VT, AT, BS, SCALED>,
                                          SplitEngine<VT, AT, BS, LAYOUT, SCALED>>::type;

  static int groups(const bmfp_matrix_t &A, bool split)
  {
    const int64_t t = split ? Split::threads(A.n_rows) : Engine::threads(A.n_rows);
    return (int) std::max<int64_t>(1, std::min<int64_t>((t + 255) / 256, SOLVER_MAX_GROUPS));
  }
  // workgroups of the split schedule's SpMV launches
  static int spmv_groups(const bmfp_solver_t &S)
  {
    if constexpr (one_tile<Split>::value) return (int) std::max<int64_t>(1, Split::groups1(S.A->n_rows));
    return S.np;
  }

  static int init(const bmfp_solver_t &S, const void *b, double eps, int max_iter, hipStream_t s)
  {
    k_bicg_init<AT><<<S.np, 256, 0, s>>>(vec_args<AT>(S, b), eps, max_iter);
    BMFP_LAUNCHED();
    return BMFP_OK;
  }

  static int iterate(const bmfp_solver_t &S, hipStream_t s)
  {
    const MatArgs<VT, AT> M = mat_args<VT, AT, SCALED>(*S.A);
    const SolverVecs<AT> V = vec_args<AT>(S, nullptr);
    if (S.split) {
      const int gw = spmv_groups(S);
      k_bicg_sp<AT><<<S.np, 256, 0, s>>>(V);
      k_bicg_sv<VT, AT, Split><<<gw, 256, 0, s>>>(M, V);
      k_bicg_reduce<1><<<REDUCE_GROUPS, 256, 0, s>>>(V.partW, gw, V.partA, V.st);
      k_bicg_ss<AT><<<S.np, 256, 0, s>>>(V);
      k_bicg_st<VT, AT, Split><<<gw, 256, 0, s>>>(M, V);
      k_bicg_reduce<2><<<REDUCE_GROUPS, 256, 0, s>>>(V.partW, gw, V.partB, V.st);
      k_bicg_sx<AT><<<S.np, 256, 0, s>>>(V);
    } else {
      k_bicg_a<VT, AT, Engine><<<S.np, 256, 0, s>>>(M, V);
      k_bicg_b<VT, AT, Engine><<<S.np, 256, 0, s>>>(M, V);
      k_bicg_c<AT><<<S.np, 256, 0, s>>>(V);
    }
    BMFP_LAUNCHED();
    return BMFP_OK;
  }

  static int jacobi(const bmfp_solver_t &S, hipStream_t s)
  {
    const MatArgs<VT, AT> M = mat_args<VT, AT, SCALED>(*S.A);
    const unsigned g = (unsigned) ((S.n + 255) / 256);
    k_jacobi<VT, AT, LAYOUT, SCALED><<<g, 256, 0, s>>>(M, S.A->bs, S.vecs.as<AT>() + 9 * S.n);
    BMFP_LAUNCHED();
    return BMFP_OK;
  }
};

template <class VT, class AT, bool SCALED, int BS, int LAYOUT, bool SL>
void assign_ops(bmfp_solver_t &S)
{
  using O = Ops<VT, AT, SCALED, BS, LAYOUT, SL>;
  S.ops_init = O::init;
  S.ops_iterate = O::iterate;
  S.ops_jacobi = O::jacobi;
  S.np = O::groups(*S.A, S.split);
  S.spmv_groups = O::spmv_groups;
}

template <class VT, class AT, bool SCALED, int BS, int LAYOUT>
bool pick_ops(bmfp_solver_t &S)
{
  if (S.A->sl_j > 0) {
    if constexpr (sliceable<VT, BS, LAYOUT>()) {
      assign_ops<VT, AT, SCALED, BS, LAYOUT, true>(S);
      return true;
    }
    return false;
  }
  assign_ops<VT, AT, SCALED, BS, LAYOUT, false>(S);
  return true;
}

template <class VT, class AT, bool SCALED, int LAYOUT>
bool pick_bs(bmfp_solver_t &S)
{
#define BMFP_SOLVER_BS(B) \
  case B: return pick_ops<VT, AT, SCALED, B, LAYOUT>(S);
  switch (S.A->bs) {
    BMFP_SOLVER_BS(1)
    BMFP_SOLVER_BS(2)
    BMFP_SOLVER_BS(3)
    BMFP_SOLVER_BS(4)
    BMFP_SOLVER_BS(6)
    BMFP_SOLVER_BS(8)
    BMFP_SOLVER_BS(12)
    BMFP_SOLVER_BS(16)
    BMFP_SOLVER_BS(24)
    BMFP_SOLVER_BS(32)
  default:
    return false;
  }
#undef BMFP_SOLVER_BS
}

template <class VT, class AT, bool SCALED>
bool pick_layout(bmfp_solver_t &S)
{
  return S.A->layout == BMFP_ROW_MAJOR ? pick_bs<VT, AT, SCALED, BMFP_ROW_MAJOR>(S)
                                       : pick_bs<VT, AT, SCALED, BMFP_COL_MAJOR>(S);
}

// ||b - y||^2 partial sums (one per workgroup) for the true residual.
template <class AT>
__global__ __launch_bounds__(256) void k_diff_norm2(const AT *__restrict__ b, const AT *__restrict__ y, int64_t n,
                                                    double *__restrict__ part)
{
  __shared__ double lds[2][4];
  double v[2] = {0.0, 0.0};
  for (int64_t i = (int64_t) blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t) gridDim.x * 256) {
    const double d = (double) b[i] - (double) y[i];
    v[0] += d * d;
    v[1] += (double) b[i] * (double) b[i];
  }
  write_partials<2>(part, v, lds);
}

// Schedule: the split schedule moves the matrix at the standalone SpMV rate
// but runs seven kernels per iteration against three, so it pays from about
// 32 MB of stored values on (plain FP64 solves of the serving stream's
// 9-point systems: up to 13% faster above, up to 27% slower below, where a
// solve is bound by launches).  BMFP_SOLVER_FUSED=1 forces the fused
// schedule (measurement switch).
int split_schedule(const bmfp_matrix_t &A)
{
  static const int forced = [] {
    const char *e = std::getenv("BMFP_SOLVER_FUSED");
    return e ? std::atoi(e) : 0;
  }();
  constexpr double SPLIT_MIN_VALUE_BYTES = 32e6;
  return !forced && (double) A.n_vals() * dtype_size(A.dtype) >= SPLIT_MIN_VALUE_BYTES;
}

}  // namespace

int sliced_pass_rows(int dtype, int bs)
{
  auto pick = [&](auto vt) -> int {
    using VT = decltype(vt);
    switch (bs) {
    case 4: return sliced_rpw<VT, 4>();
    case 6: return sliced_rpw<VT, 6>();
    case 8: return sliced_rpw<VT, 8>();
    default: return 0;
    }
  };
  if (dtype == BMFP_F32) return pick(float{});
  if (dtype == BMFP_F16) return pick(__half{});
  return 0;
}

size_t solver_workspace_bytes(const bmfp_matrix_t &A, int use_jacobi)
{
  const int64_t n = A.y_len();
  // partial sums: 5 * SOLVER_MAX_GROUPS, plus 2 per workgroup of the split
  // schedule's SpMVs (at most one workgroup per 16 rows)
  return (size_t) n * vec_size(A.dtype) * (use_jacobi ? 10 : 9) + sizeof(double) * 5 * SOLVER_MAX_GROUPS +
         sizeof(double) * 2 * (size_t) ((n + 15) / 16 + SOLVER_MAX_GROUPS) + sizeof(SolverState) + 256;
}

int solver_create(bmfp_solver_t &S, bmfp_matrix_t *A, int use_jacobi)
{
  S.A = A;
  S.use_jacobi = use_jacobi ? 1 : 0;
  S.n = A->y_len();
  // a pass-sliced copy runs on the split schedule only, without Jacobi
  S.split = split_schedule(*A) || A->sl_j > 0;
  if (A->n_rows != A->n_cols) return fail(BMFP_ERR_INVALID, "BiCGStab needs a square matrix");
  if (!A->is_bcsr) return fail(BMFP_ERR_UNSUPPORTED, "BiCGStab runs on BCSR handles (CSR: use bs = 1)");
  if (A->sl_j > 0 && use_jacobi) return fail(BMFP_ERR_UNSUPPORTED, "no Jacobi preconditioning on a pass-sliced copy");
  bool ok = false;
  switch (A->dtype) {
  case BMFP_F64: ok = pick_layout<double, double, false>(S); break;
  case BMFP_F32: ok = pick_layout<float, float, false>(S); break;
  case BMFP_F16: ok = pick_layout<__half, float, true>(S); break;
  }
  if (!ok) return fail(BMFP_ERR_UNSUPPORTED, "no BiCGStab kernels for bs %d", A->bs);
  const size_t vb = vec_size(A->dtype);
  BMFP_TRY(S.vecs.alloc((size_t) std::max<int64_t>(S.n, 1) * vb * (S.use_jacobi ? 10 : 9)));
  // fused: partA, partB, partC; split adds the per-workgroup partials of its SpMVs
  const size_t wparts = S.split ? (size_t) 2 * std::max(S.spmv_groups(S), (int) ((S.n + 15) / 16)) : 0;
  BMFP_TRY(S.parts.alloc(sizeof(double) * (5 * SOLVER_MAX_GROUPS + wparts)));
  BMFP_TRY(S.state.alloc(sizeof(SolverState)));
  BMFP_HIP(hipHostMalloc(reinterpret_cast<void **>(&S.h_state), sizeof(SolverState), hipHostMallocDefault));
  BMFP_HIP(hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking));
  BMFP_HIP(hipEventCreate(&S.ev0));
  BMFP_HIP(hipEventCreate(&S.ev1));
  BMFP_HIP(hipEventCreateWithFlags(&S.ev_sync, hipEventDisableTiming));
  if (S.use_jacobi && S.n) {
    BMFP_TRY(S.ops_jacobi(S, S.stream));
    BMFP_HIP(hipStreamSynchronize(S.stream));
  }
  return BMFP_OK;
}

// Captures graphs[k] = 2^k iterations (k = 0..3).
static int build_graphs(bmfp_solver_t &S)
{
  for (int k = 0; k < 4; k++) {
    if (S.graphs[k]) continue;
    hipGraph_t g = nullptr;
    BMFP_HIP(hipStreamBeginCapture(S.stream, hipStreamCaptureModeThreadLocal));
    int st = BMFP_OK;
    for (int i = 0; i < (1 << k) && st == BMFP_OK; i++) st = S.ops_iterate(S, S.stream);
    hipError_t e = hipStreamEndCapture(S.stream, &g);
    if (st != BMFP_OK) {
      if (g) (void) hipGraphDestroy(g);
      return st;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
    e = hipGraphInstantiate(&S.graphs[k], g, nullptr, nullptr, 0);
    (void) hipGraphDestroy(g);
    if (e != hipSuccess) {
      S.graphs[k] = nullptr;
      return hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
    }
  }
  return BMFP_OK;
}

// Launches exactly `iters` iterations: whole 8-iteration graphs, then the
// binary digits of the remainder.
static int launch_iterations(bmfp_solver_t &S, int iters, hipStream_t s)
{
  for (; iters >= GRAPH_ITERS; iters -= GRAPH_ITERS) BMFP_HIP(hipGraphLaunch(S.graphs[3], s));
  for (int k = 2; k >= 0; k--)
    if (iters & (1 << k)) BMFP_HIP(hipGraphLaunch(S.graphs[k], s));
  return BMFP_OK;
}

int solver_run(bmfp_solver_t &S, const void *b, void *x, double eps, int max_iterations, hipStream_t user,
               bmfp_solve_info *info, bool true_residual, int first_batch, bool timed)
{
  if (!(eps > 0.0) || max_iterations < 0) return fail(BMFP_ERR_INVALID, "eps must be > 0, max_iterations >= 0");
  if (S.n == 0) {
    if (info) *info = bmfp_solve_info{};
    return BMFP_OK;
  }
  if (!b || !x) return fail(BMFP_ERR_INVALID, "NULL vector");
  BMFP_TRY(build_graphs(S));
  hipStream_t s = S.stream;
  // order after the caller's prior work on its stream
  BMFP_HIP(hipEventRecord(S.ev_sync, user));
  BMFP_HIP(hipStreamWaitEvent(s, S.ev_sync, 0));
  BMFP_HIP(hipEventRecord(S.ev0, s));
  BMFP_TRY(S.ops_init(S, b, eps, max_iterations, s));
  // Iterations run in batches between reads of the solver state.  Kernels
  // after the stop are no-ops but each still costs a launch, so after the
  // first batch (8) each batch is the number of iterations the residual's
  // decrease so far predicts are left, plus the one whose first kernel sees
  // the stop (doubling while the residual has not decreased), launched as
  // graphs of 8, 4, 2 and 1 iterations.
  int want = first_batch > 0 ? std::min(first_batch, 8 * GRAPH_ITERS) : GRAPH_ITERS;
  want = std::max(1, std::min(want, max_iterations + 1));
  for (;;) {
    BMFP_TRY(launch_iterations(S, want, s));
    BMFP_HIP(hipMemcpyAsync(S.h_state, S.state.ptr, sizeof(SolverState), hipMemcpyDeviceToHost, s));
    BMFP_HIP(hipStreamSynchronize(s));
    const SolverState &st = *S.h_state;
    if (st.done) break;
    int next = std::min(want * 2, 8 * GRAPH_ITERS);
    const double r0 = std::sqrt(st.norm_b2), rk = std::sqrt(st.rr), goal = st.eps * r0;
    if (st.iter > 0 && rk > 0 && rk < r0 && goal > 0) {
      const double per_iter = std::log(rk / r0) / st.iter;  // < 0
      const double left = std::log(goal / rk) / per_iter;
      next = (int) std::min<double>(8 * GRAPH_ITERS, std::max(0.0, std::ceil(left)) + 1);
    }
    // never past the iteration whose first kernel reports max_iterations
    want = std::max(1, std::min(next, st.max_iter - st.iter + 1));
  }
  const size_t vb = vec_size(S.A->dtype);
  BMFP_HIP(hipMemcpyAsync(x, S.vecs.ptr, (size_t) S.n * vb, hipMemcpyDeviceToDevice, s));
  // the caller's later work on its stream reads x: order it after the copy
  // (an untimed solve returns without waiting for it)
  BMFP_HIP(hipEventRecord(S.ev_sync, s));
  BMFP_HIP(hipStreamWaitEvent(user, S.ev_sync, 0));
  float ms = 0.f;
  if (timed) {
    BMFP_HIP(hipEventRecord(S.ev1, s));
    BMFP_HIP(hipEventSynchronize(S.ev1));
    BMFP_HIP(hipEventElapsedTime(&ms, S.ev0, S.ev1));
  }
  const SolverState &st = *S.h_state;
  bmfp_solve_info out{};
  out.iterations = st.iter;
  out.status = st.status;
  out.rel_residual = st.norm_b2 > 0 ? std::sqrt(st.rr) / std::sqrt(st.norm_b2) : 0.0;
  out.seconds = ms * 1e-3;
  // true residual ||b - A x|| / ||b|| (after the timed region; BCSR-order values only)
  if (true_residual && S.A->sl_j == 0) {
    void *y = S.vecs.as<char>() + (size_t) 8 * S.n * vb;  // t buffer
    BMFP_TRY(launch_bcsr(*S.A, auto_variant(*S.A), x, y, s));
    const int ng = std::min<int64_t>((S.n + 255) / 256, SOLVER_MAX_GROUPS);
    double *part = S.parts.as<double>();
    if (S.A->dtype == BMFP_F64)
      k_diff_norm2<double><<<ng, 256, 0, s>>>(static_cast<const double *>(b), static_cast<const double *>(y), S.n, part);
    else
      k_diff_norm2<float><<<ng, 256, 0, s>>>(static_cast<const float *>(b), static_cast<const float *>(y), S.n, part);
    BMFP_LAUNCHED();
    std::vector<double> h((size_t) 2 * ng);
    BMFP_HIP(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    BMFP_HIP(hipStreamSynchronize(s));
    double rr = 0, bb = 0;
    for (int g = 0; g < ng; g++) {
      rr += h[2 * g];
      bb += h[2 * g + 1];
    }
    out.true_rel_residual = bb > 0 ? std::sqrt(rr) / std::sqrt(bb) : 0.0;
  }
  if (info) *info = out;
  return out.status;
}

}  // namespace bmfp

using namespace bmfp;

bmfp_solver_t::~bmfp_solver_t()
{
  // A solve that returned early on an error can leave graph launches and the
  // state copy into h_state queued: let them finish before the graphs, the
  // pinned state and the stream go.
  if (stream) (void) hipStreamSynchronize(stream);
  if (refine) refine_destroy(refine);
  for (hipGraphExec_t g : graphs)
    if (g) (void) hipGraphExecDestroy(g);
  if (h_state) (void) hipHostFree(h_state);
  if (ev0) (void) hipEventDestroy(ev0);
  if (ev1) (void) hipEventDestroy(ev1);
  if (ev_sync) (void) hipEventDestroy(ev_sync);
  if (stream) (void) hipStreamDestroy(stream);
}

extern "C" int bmfp_bicgstab_workspace_bytes(bmfp_matrix A, int use_jacobi, size_t *bytes)
{
  return bmfp::guarded("bmfp_bicgstab_workspace_bytes", [&]() -> int {
    if (!A || !bytes) return fail(BMFP_ERR_INVALID, "NULL argument");
    *bytes = solver_workspace_bytes(*A, use_jacobi);
    return BMFP_OK;
  });
}

extern "C" int bmfp_bicgstab_create(bmfp_solver *out, bmfp_matrix A, int use_jacobi)
{
  return bmfp::guarded("bmfp_bicgstab_create", [&]() -> int {
    if (!out || !A) return fail(BMFP_ERR_INVALID, "NULL argument");
    *out = nullptr;
    std::unique_ptr<bmfp_solver_t> S(new bmfp_solver_t);
    BMFP_TRY(solver_create(*S, A, use_jacobi));
    *out = S.release();
    return BMFP_OK;
  });
}

extern "C" int bmfp_bicgstab_solve(bmfp_solver s, const void *b, void *x, double eps, int max_iterations,
                                   bmfp_stream stream, bmfp_solve_info *info)
{
  return bmfp::guarded("bmfp_bicgstab_solve", [&]() -> int {
    if (!s || s->refine) return fail(BMFP_ERR_INVALID, "not a BiCGStab handle");
    const int st = solver_run(*s, b, x, eps, max_iterations, static_cast<hipStream_t>(stream), info);
    if (st == BMFP_ERR_NOT_CONVERGED) set_error("BiCGStab: max_iterations (%d) reached", max_iterations);
    else if (st == BMFP_ERR_BREAKDOWN) set_error("BiCGStab breakdown");
    else if (st == BMFP_ERR_NONFINITE) set_error("BiCGStab: non-finite value in the iteration");
    return st;
  });
}

extern "C" int bmfp_bicgstab_destroy(bmfp_solver s)
{
  return bmfp::guarded("bmfp_bicgstab_destroy", [&]() -> int {
    delete s;
    return BMFP_OK;
  });
}
