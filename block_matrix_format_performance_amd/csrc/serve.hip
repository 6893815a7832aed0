// serve.hip -- continuous-batching request loop (bmfp_server_*).
//
// Requests (a BCSR FP64 system with 1..4 right-hand sides) queue on the host,
// are admitted while their footprint fits the device memory budget, and are
// solved by FP64 iterative refinement around FP32 BiCGStab.  Every admitted
// (request, right-hand side) pair is one "system"; all active systems advance
// together through a per-system state machine on the device:
//
//   RESID   (slot 1)  x' = x + d, r = b - A64 x', r32 = r      fin1: DONE | INIT
//   INIT    (slot 2)  d = 0, ri = rh = r32, p = v = 0           fin2: INNER_A
//   INNER_A (slot 1)  p = ri + beta (p - omega v), v = A32 p    fin1: alpha
//   INNER_B (slot 2)  s = ri - alpha v, t = A32 s               fin2: omega
//   INNER_C (slot 3)  ri = s - omega t, d += alpha p + omega s  fin3: INNER_A | RESID
//
// One batched iteration is six kernels (slot1 fin1 slot2 fin2 slot3 fin3),
// captured several times over in one HIP graph that is replayed every round.
// Slot kernels walk a tile table (a tile = 4 wavefronts x 64/bs block rows of
// one system) with a persistent grid; a tile works only if its system's phase
// matches the slot.  Each tile writes its own partial sums and the fin
// kernels reduce each system's tiles in tile order, so a system's arithmetic
// is the same whatever else is in the batch (bitwise identical results).
// New requests are set up on a separate stream while rounds run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <vector>

#include "matrix.hpp"
#include "solver_kernels.hpp"

namespace bmfp {

enum : int {
  SYS_EMPTY = 0,
  SYS_RESID = 1,
  SYS_INIT = 2,
  SYS_INNER_A = 3,
  SYS_INNER_B = 4,
  SYS_INNER_C = 5,
  SYS_DONE = 6
};

struct SysDesc {
  int bs, n_brows, tile0, n_tiles;
  int64_t n;
  const int *rp;
  const int *cols;
  const double *v64;
  const float *v32;
  const double *b;
  double *x[2];
  float *r32, *d, *ri, *rh, *p[2], *v[2], *s, *t;
  double tol, inner_tol;
  int max_outer, max_inner;
};

struct SysState {
  int phase, status, outer, inner_it, inner_total, pad;
  double norm_b2, rel, rho, alpha, omega, beta, tt, inner_rr0;
};

struct TileDesc {
  int sys, br0;
};

__host__ __device__ constexpr int serve_tile_brows(int bs) { return 4 * (64 / bs); }

namespace {

// y = (A op)_row for the rows of one tile: lane (w, lbr, r) owns scalar row
// (br0 + w*RPW + lbr)*BS + r; operand elements are shared by shuffles.
template <class VT, int BS, class OP, class EPI>
__device__ __forceinline__ void tile_spmv(const SysDesc &D, const VT *__restrict__ vals, int br0, const OP &op,
                                          EPI &&epi)
{
  constexpr int RPW = 64 / BS;
  using AT = decltype(op((int64_t) 0));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int lbr = lane / BS, r = lane - lbr * BS, base_lane = lbr * BS;
  const int br = br0 + w * RPW + lbr;
  const bool active = lane < RPW * BS && br < D.n_brows;
  int b = 0, b1 = 0;
  if (active) {
    b = D.rp[br];
    b1 = D.rp[br + 1];
  }
  AT acc = 0;
  for (; b + 3 <= b1; b += 3) {
    int col[3];
    VT v[3][BS];
    AT xe[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) col[k] = D.cols[b + k];
#pragma unroll
    for (int k = 0; k < 3; ++k) load_block_row<VT, BS, BMFP_ROW_MAJOR>(vals + (int64_t) (b + k) * (BS * BS), r, v[k]);
#pragma unroll
    for (int k = 0; k < 3; ++k) xe[k] = op((int64_t) col[k] * BS + r);
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int c = 0; c < BS; ++c) acc = fmadd((AT) v[k][c], __shfl(xe[k], base_lane + c), acc);
  }
  for (; b < b1; ++b) {
    VT v[BS];
    load_block_row<VT, BS, BMFP_ROW_MAJOR>(vals + (int64_t) b * (BS * BS), r, v);
    const AT xe = op((int64_t) D.cols[b] * BS + r);
#pragma unroll
    for (int c = 0; c < BS; ++c) acc = fmadd((AT) v[c], __shfl(xe, base_lane + c), acc);
  }
  if (active) epi((int64_t) br * BS + r, acc);
}

#define BMFP_SERVE_BS_SWITCH(bs, CALL) \
  switch (bs) {                        \
  case 1: { constexpr int BS = 1; CALL; } break; \
  case 2: { constexpr int BS = 2; CALL; } break; \
  case 3: { constexpr int BS = 3; CALL; } break; \
  case 4: { constexpr int BS = 4; CALL; } break; \
  case 6: { constexpr int BS = 6; CALL; } break; \
  case 8: { constexpr int BS = 8; CALL; } break; \
  default: break;                      \
  }

__device__ __forceinline__ int tile_phase(const SysState *ss, int sys)
{
  __shared__ int s_phase;
  __syncthreads();
  if (threadIdx.x == 0) s_phase = ss[sys].phase;
  __syncthreads();
  return s_phase;
}

__device__ __forceinline__ void tile_partials(double *__restrict__ part, int tile, double (&v)[2], double (*lds)[4])
{
  block_sum<2>(v, lds);
  if (threadIdx.x == 0) {
    part[(int64_t) tile * 2] = v[0];
    part[(int64_t) tile * 2 + 1] = v[1];
  }
}

// ---- slot 1: RESID (FP64 residual) and INNER_A (v = A32 p) -------------------
__global__ __launch_bounds__(256) void k_slot1(const SysDesc *__restrict__ sd, const SysState *__restrict__ ss,
                                               const TileDesc *__restrict__ tiles, const int *__restrict__ ntp,
                                               double *__restrict__ part)
{
  __shared__ double lds[2][4];
  const int nt = *ntp;
  for (int tile = blockIdx.x; tile < nt; tile += gridDim.x) {
    const TileDesc T = tiles[tile];
    const int phase = tile_phase(ss, T.sys);
    if (phase != SYS_RESID && phase != SYS_INNER_A) continue;
    const SysDesc &D = sd[T.sys];
    const SysState &S = ss[T.sys];
    double acc[2] = {0.0, 0.0};
    if (phase == SYS_RESID) {
      const int xp = S.outer & 1;
      const OperandXD op{D.x[xp], D.d};
      double *xn = D.x[1 - xp];
      BMFP_SERVE_BS_SWITCH(D.bs, (tile_spmv<double, BS>(D, D.v64, T.br0, op, [&](int64_t row, double y) {
                             xn[row] = op.raw(row);
                             const double bi = D.b[row];
                             const double ri = bi - y;
                             D.r32[row] = (float) ri;
                             acc[0] += ri * ri;
                             acc[1] += bi * bi;
                           })));
    } else {
      const int it = S.inner_it, pv = (it - 1) & 1, cu = it & 1;
      const OperandP<float> op{D.ri, D.p[pv], D.v[pv], nullptr, (float) S.beta, (float) (it == 0 ? 1.0 : S.omega)};
      float *pc = D.p[cu];
      float *vc = D.v[cu];
      BMFP_SERVE_BS_SWITCH(D.bs, (tile_spmv<float, BS>(D, D.v32, T.br0, op, [&](int64_t row, float y) {
                             pc[row] = op.raw(row);
                             vc[row] = y;
                             acc[0] += (double) D.rh[row] * (double) y;
                           })));
    }
    tile_partials(part, tile, acc, lds);
  }
}

// ---- slot 2: INIT (inner start) and INNER_B (t = A32 s) ----------------------
__global__ __launch_bounds__(256) void k_slot2(const SysDesc *__restrict__ sd, const SysState *__restrict__ ss,
                                               const TileDesc *__restrict__ tiles, const int *__restrict__ ntp,
                                               double *__restrict__ part)
{
  __shared__ double lds[2][4];
  const int nt = *ntp;
  for (int tile = blockIdx.x; tile < nt; tile += gridDim.x) {
    const TileDesc T = tiles[tile];
    const int phase = tile_phase(ss, T.sys);
    if (phase != SYS_INIT && phase != SYS_INNER_B) continue;
    const SysDesc &D = sd[T.sys];
    const SysState &S = ss[T.sys];
    double acc[2] = {0.0, 0.0};
    if (phase == SYS_INIT) {
      const int rows = serve_tile_brows(D.bs) * D.bs;
      const int64_t row = (int64_t) T.br0 * D.bs + threadIdx.x;
      if (threadIdx.x < rows && row < D.n) {
        const float r0 = D.r32[row];
        D.d[row] = 0.f;
        D.ri[row] = r0;
        D.rh[row] = r0;
        D.p[0][row] = D.p[1][row] = 0.f;
        D.v[0][row] = D.v[1][row] = 0.f;
        acc[0] = (double) r0 * (double) r0;
        acc[1] = acc[0];
      }
    } else {
      const int cu = S.inner_it & 1;
      const OperandS<float> op{D.ri, D.v[cu], nullptr, (float) S.alpha};
      BMFP_SERVE_BS_SWITCH(D.bs, (tile_spmv<float, BS>(D, D.v32, T.br0, op, [&](int64_t row, float y) {
                             const float s = op.raw(row);
                             D.s[row] = s;
                             D.t[row] = y;
                             acc[0] += (double) y * (double) s;
                             acc[1] += (double) y * (double) y;
                           })));
    }
    tile_partials(part, tile, acc, lds);
  }
}

// ---- slot 3: INNER_C (ri = s - omega t, d += alpha p + omega s) --------------
__global__ __launch_bounds__(256) void k_slot3(const SysDesc *__restrict__ sd, const SysState *__restrict__ ss,
                                               const TileDesc *__restrict__ tiles, const int *__restrict__ ntp,
                                               double *__restrict__ part)
{
  __shared__ double lds[2][4];
  const int nt = *ntp;
  for (int tile = blockIdx.x; tile < nt; tile += gridDim.x) {
    const TileDesc T = tiles[tile];
    const int phase = tile_phase(ss, T.sys);
    if (phase != SYS_INNER_C) continue;
    const SysDesc &D = sd[T.sys];
    const SysState &S = ss[T.sys];
    double acc[2] = {0.0, 0.0};
    const int rows = serve_tile_brows(D.bs) * D.bs;
    const int64_t row = (int64_t) T.br0 * D.bs + threadIdx.x;
    if (threadIdx.x < rows && row < D.n) {
      const float a = (float) S.alpha, w = (float) S.omega;
      const float s = D.s[row];
      const float h = fmadd(a, D.p[S.inner_it & 1][row], D.d[row]);
      const float rn = fmadd(-w, D.t[row], s);
      D.d[row] = fmadd(w, s, h);
      D.ri[row] = rn;
      acc[0] = (double) D.rh[row] * (double) rn;
      acc[1] = (double) rn * (double) rn;
    }
    tile_partials(part, tile, acc, lds);
  }
}

// ---- finalize: one wavefront per system, fixed-order reduction ----------------
template <int SLOT>
__global__ __launch_bounds__(256) void k_fin(const SysDesc *__restrict__ sd, SysState *__restrict__ ss,
                                             const int *__restrict__ list, const int *__restrict__ nlp,
                                             const double *__restrict__ part)
{
  const int wave = (int) ((blockIdx.x * 256 + threadIdx.x) >> 6), lane = threadIdx.x & 63;
  if (wave >= *nlp) return;
  const int sys = list[wave];
  SysState &S = ss[sys];
  const int phase = S.phase;
  const bool mine = SLOT == 1 ? (phase == SYS_RESID || phase == SYS_INNER_A)
                  : SLOT == 2 ? (phase == SYS_INIT || phase == SYS_INNER_B)
                              : phase == SYS_INNER_C;
  if (!mine) return;
  const SysDesc &D = sd[sys];
  double a0 = 0.0, a1 = 0.0;
  for (int k = lane; k < D.n_tiles; k += 64) {
    a0 += part[(int64_t) (D.tile0 + k) * 2];
    a1 += part[(int64_t) (D.tile0 + k) * 2 + 1];
  }
  a0 = wave_sum(a0);
  a1 = wave_sum(a1);
  if (lane != 0) return;
  if (phase == SYS_RESID) {
    // a0 = ||r||^2, a1 = ||b||^2 for x' = x + d (now in x[1 - outer%2])
    if (S.outer == 0) S.norm_b2 = a1;
    S.outer += 1;
    // b = 0 gives x = 0 exactly; a non-finite ||b|| or ||r|| gives a NaN rel
    const double rel = S.norm_b2 == 0.0 ? 0.0 : sqrt(a0) / sqrt(S.norm_b2);
    S.rel = rel;
    if (!isfinite(rel)) { S.status = BMFP_ERR_NONFINITE; S.phase = SYS_DONE; }
    else if (rel < D.tol) { S.status = BMFP_OK; S.phase = SYS_DONE; }
    else if (S.outer > D.max_outer) { S.status = BMFP_ERR_NOT_CONVERGED; S.phase = SYS_DONE; }
    else S.phase = SYS_INIT;
  } else if (phase == SYS_INIT) {
    // a0 = (rh, ri) = rho_0, a1 = ||ri||^2
    S.inner_rr0 = a1;
    S.inner_it = 0;
    S.rho = a0;
    S.alpha = 1.0;
    S.omega = 1.0;
    S.beta = a0;  // rho_0 / 1 * 1 / 1
    S.phase = (a0 == 0.0 || !isfinite(a0)) ? SYS_RESID : SYS_INNER_A;
  } else if (phase == SYS_INNER_A) {
    // a0 = (rh, v)
    if (a0 == 0.0 || !isfinite(a0)) {
      S.inner_total += S.inner_it;
      S.phase = SYS_RESID;  // inner breakdown: d keeps the last valid iterate
    } else {
      S.alpha = S.rho / a0;
      S.phase = SYS_INNER_B;
    }
  } else if (phase == SYS_INNER_B) {
    // a0 = (t, s), a1 = (t, t)
    S.tt = a1;
    S.omega = a1 == 0.0 ? 0.0 : a0 / a1;
    S.phase = SYS_INNER_C;
  } else {
    // a0 = (rh, ri), a1 = ||ri||^2 after the update
    S.inner_it += 1;
    const double rel_in = S.inner_rr0 > 0.0 ? sqrt(a1) / sqrt(S.inner_rr0) : 0.0;
    const bool stop = !(rel_in >= D.inner_tol) || S.inner_it >= D.max_inner || S.tt == 0.0 || S.omega == 0.0 ||
                      a0 == 0.0 || !isfinite(a0);
    if (stop) {
      S.inner_total += S.inner_it;
      S.phase = SYS_RESID;
    } else {
      S.beta = a0 / S.rho * S.alpha / S.omega;
      S.rho = a0;
      S.phase = SYS_INNER_A;
    }
  }
}

// ---- admission-time device work --------------------------------------------------
// 9-point stencil values, one thread per scalar row; identical operations (and
// order of the row sums) to the host generator, with explicit rounding so no
// multiply-add is fused.
__global__ __launch_bounds__(256) void k_gen_stencil_rows(int nx, int ny, int bs, int kind, uint64_t seed,
                                                          const int *__restrict__ rp, int *__restrict__ cols,
                                                          double *__restrict__ vals)
{
  const int64_t t = (int64_t) blockIdx.x * 256 + threadIdx.x;
  const int64_t n = (int64_t) nx * ny * bs;
  if (t >= n) return;
  const int64_t i = t / bs;
  const int r = (int) (t - i * bs);
  const int iy = (int) (i / nx), ix = (int) (i - (int64_t) iy * nx);
  const uint64_t ka = rng_key(seed, kind == 0 ? STREAM_DD : STREAM_K), kc = rng_key(seed, STREAM_C);
  const int64_t bs2 = (int64_t) bs * bs;
  int64_t nb = rp[i];
  int64_t diag = -1;
  double sum = 0.0;
  for (int s = 0; s < 9; s++) {
    const int jy = iy + s / 3 - 1, jx = ix + s % 3 - 1;
    if (jy < 0 || jy >= ny || jx < 0 || jx >= nx) continue;
    if (r == 0) cols[nb] = jy * nx + jx;
    const uint64_t base = ((uint64_t) i * 9u + (uint64_t) s) * (uint64_t) bs2;
    double *row = vals + nb * bs2 + (int64_t) r * bs;
    if (s == 4) diag = nb;
    for (int c = 0; c < bs; c++) {
      double v;
      if (kind == 0) {
        v = rng_sym(ka, base + (uint64_t) r * bs + c);
      } else if (s == 4) {
        v = 0.0;
      } else {
        const int lo = r < c ? r : c, hi = r < c ? c : r;
        const double k = rng_unit(ka, base + (uint64_t) lo * bs + hi);
        double cc = 0.0;
        if (r < c) cc = rng_sym(kc, base + (uint64_t) r * bs + c);
        else if (r > c) cc = -rng_sym(kc, base + (uint64_t) c * bs + r);
        v = -__dadd_rn(k, __dmul_rn(0.3, cc));
      }
      row[c] = v;
      if (!(s == 4 && c == r)) sum = __dadd_rn(sum, fabs(v));
    }
    nb++;
  }
  double *dg = vals + diag * bs2 + (int64_t) r * bs + r;
  if (kind == 0) *dg = __dadd_rn(*dg, __dadd_rn(sum, 1.0));
  else *dg = __dadd_rn(__dmul_rn(1.02, sum), 1.0);
}

__global__ __launch_bounds__(256) void k_to_f32(const double *__restrict__ a, float *__restrict__ b, int64_t n)
{
  for (int64_t i = (int64_t) blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t) gridDim.x * 256) b[i] = (float) a[i];
}

// b[k] for right-hand side k of a request: uniform[-1,1] keyed by (seed, k*n + i).
__global__ __launch_bounds__(256) void k_rhs(double *__restrict__ b, int64_t n, int k, uint64_t seed)
{
  const uint64_t key = rng_key(seed, STREAM_RHS);
  for (int64_t i = (int64_t) blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t) gridDim.x * 256)
    b[i] = rng_sym(key, (uint64_t) k * (uint64_t) n + (uint64_t) i);
}

}  // namespace
}  // namespace bmfp
#include "serve_host.inc"
