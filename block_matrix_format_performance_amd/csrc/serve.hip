// serve.hip -- continuous-batching request loop (bmfp_server_*).
//
// A request is an FP64 BCSR system with 1..4 right-hand sides, solved by FP64
// iterative refinement around FP32 BiCGStab.  Admitted requests advance
// together through a per-request state machine on the device; the right-hand
// sides of one request move in lockstep so that every block row of its
// matrix is read once per SpMV for all of them:
//
//   RESID   (slot 0)  x' = x + d, r = b - A64 x', r32 = r              fin0: DONE | INNER_A | RESID
//   INNER_A (slot 1)  p = ri + beta (p - omega v), v = A32 p            fin1: INNER_B | RESID
//   INNER_B (slot 2)  s = ri - alpha v, t = A32 s                       fin2: INNER_C
//   INNER_C (slot 3)  ri = s - omega t, d += alpha p + omega s          fin3: INNER_A | RESID
//
// The inner solve starts from ri = r32 with the shadow residual r32 itself
// (it stays constant until the next RESID), p = v = d = 0: iteration 0 reads
// r32 in place of ri and skips the zero terms, so no initialisation pass is
// needed.  With x = 0, the first RESID is r = b and reads no matrix.
//
// Per right-hand side k three bit masks select the work: `live` (not yet
// converged or failed) for RESID, `act` (still inside the current inner
// solve) for the rest, and `dval` (d holds this refinement step's correction;
// otherwise it counts as 0).  A right-hand side whose inner solve stops is
// frozen until the others stop too, so its arithmetic is exactly that of
// solving it alone.
//
// One batched iteration is eight kernels (slot0 fin0 slot1 fin1 slot2 fin2
// slot3 fin3), captured iterations_per_round times in one HIP graph.  Slot
// kernels give every wavefront its own tiles from a tile table (no workgroup
// barriers); a tile is `passes` consecutive chunks of one request, a chunk is
// 4 wavefront passes of 64/bs block rows.  Partial sums are kept per chunk and
// the fin kernels (one wavefront per request) reduce a request's chunks in
// chunk order, so the arithmetic of a request depends neither on the rest of
// the batch nor on the tile schedule: results are bitwise identical alone, in
// any mix, and under any schedule the autotuner picks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <vector>

#include "matrix.hpp"
#include "solver_kernels.hpp"
#include "stencil_kernels.hpp"

namespace bmfp {

constexpr int SERVE_MAX_RHS = 4;
constexpr int SERVE_PART = 2 * SERVE_MAX_RHS;  // partial sums per chunk
constexpr int SERVE_CHUNK_PASSES = 4;          // wavefront passes per chunk

enum : int { REQ_EMPTY = 0, REQ_RESID = 1, REQ_INNER_A = 2, REQ_INNER_B = 3, REQ_INNER_C = 4, REQ_DONE = 5 };

// Device view of one request.  Vectors interleave the nrhs right-hand sides:
// element i of right-hand side k is at i*nrhs + k.
struct ReqDesc {
  int bs, n_brows, nrhs, chunk0, n_chunks, pad;
  int64_t n;  // scalar rows of one right-hand side
  const int *rp;
  const int *cols;
  const double *v64;
  const float *v32;
  const float *dinv;  // Jacobi: 1 / diagonal entry per scalar row (FP32); null without
  const double *b;
  double *x[2];
  float *r32, *d, *ri, *p[2], *v[2], *s, *t;
  double tol, inner_tol;
  int max_outer, max_inner;
};

struct RhsState {
  int status, outer, inner_it, inner_total;
  double norm_b2, rel, rho, alpha, omega, beta, tt, inner_rr0;
};

// `outer` counts the refinement steps of the live right-hand sides and
// `inner_it` the iterations of the current inner solve of the active ones
// (each set moves in lockstep, so one counter serves all its members).
struct ReqState {
  int phase;
  unsigned live, act, dval;
  int outer, inner_it, pad[2];
  RhsState k[SERVE_MAX_RHS];
};

struct TileDesc {
  int req, chunk, passes;
};

__host__ __device__ constexpr int serve_chunk_brows(int bs) { return SERVE_CHUNK_PASSES * (64 / bs); }

namespace {

// Slot and finalize kernel SLOT serve phase REQ_RESID + SLOT.
template <int SLOT> __device__ __forceinline__ bool slot_owns(int phase) { return phase == REQ_RESID + SLOT; }

using Acc = double[SERVE_MAX_RHS][2];

// Pointers read from the request tables are generic; accesses through these
// global-address-space views compile to global (not flat) loads and stores.
template <class T> using gptr = __attribute__((address_space(1))) T *;
template <class T> __device__ __forceinline__ gptr<T> glb(T *p) { return (gptr<T>) p; }

// Element `off` of an array: the byte offset fits 32 bits (guaranteed at
// submission), so the access is a uniform base plus a 32-bit lane offset.
template <class T> __device__ __forceinline__ T ld(gptr<const T> base, uint32_t off)
{
  return *(gptr<const T>) ((gptr<const char>) base + off * (uint32_t) sizeof(T));
}

// The NR interleaved right-hand-side values at element offset `off`
// (off = i*NR, so the group is NR*sizeof(T)-aligned), in the widest loads and
// stores that alignment allows.
template <class T, int NR> __device__ __forceinline__ void ldv(gptr<const T> base, uint32_t off, T (&o)[NR])
{
  gptr<const char> p = (gptr<const char>) base + off * (uint32_t) sizeof(T);
  constexpr int BYTES = (int) sizeof(T) * NR;
  if constexpr (BYTES == 8 || BYTES == 16 || BYTES == 32) {
    constexpr int PER = (BYTES >= 16 ? 16 : BYTES) / (int) sizeof(T);
    typedef T vec_t __attribute__((ext_vector_type(PER)));
#pragma unroll
    for (int c = 0; c < NR; c += PER) {
      const vec_t t = *(gptr<const vec_t>) (p + c * (int) sizeof(T));
#pragma unroll
      for (int e = 0; e < PER; ++e) o[c + e] = t[e];
    }
  } else {
#pragma unroll
    for (int c = 0; c < NR; ++c) o[c] = *(gptr<const T>) (p + c * (int) sizeof(T));
  }
}
template <class T, int NR> __device__ __forceinline__ void stv(gptr<T> base, uint32_t off, const T (&v)[NR])
{
  gptr<char> p = (gptr<char>) base + off * (uint32_t) sizeof(T);
  constexpr int BYTES = (int) sizeof(T) * NR;
  if constexpr (BYTES == 8 || BYTES == 16 || BYTES == 32) {
    constexpr int PER = (BYTES >= 16 ? 16 : BYTES) / (int) sizeof(T);
    typedef T vec_t __attribute__((ext_vector_type(PER)));
#pragma unroll
    for (int c = 0; c < NR; c += PER) {
      vec_t t;
#pragma unroll
      for (int e = 0; e < PER; ++e) t[e] = v[c + e];
      *(gptr<vec_t>) (p + c * (int) sizeof(T)) = t;
    }
  } else {
#pragma unroll
    for (int c = 0; c < NR; ++c) *(gptr<T>) (p + c * (int) sizeof(T)) = v[c];
  }
}

// N consecutive values at p (row r of a row-major block), in the widest loads
// the alignment of a block row allows (up to 16 bytes).
template <class T, int N> __device__ __forceinline__ void load_row(gptr<const T> p, T (&v)[N])
{
  constexpr int ALIGN = lowbit((int) sizeof(T) * N) < 16 ? lowbit((int) sizeof(T) * N) : 16;
  constexpr int PER = ALIGN / (int) sizeof(T);
  if constexpr (PER <= 1) {
#pragma unroll
    for (int c = 0; c < N; ++c) v[c] = p[c];
  } else {
    typedef T vec_t __attribute__((ext_vector_type(PER)));
#pragma unroll
    for (int c = 0; c < N; c += PER) {
      const vec_t t = *(gptr<const vec_t>) (p + c);
#pragma unroll
      for (int e = 0; e < PER; ++e) v[c + e] = t[e];
    }
  }
}

template <int CTRL> __device__ __forceinline__ double dpp_mov(double v)
{
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double lane_value(double v, int lane)
{
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                          __builtin_amdgcn_readlane(__double2loint(v), lane));
}

// Sum over the 64 lanes of a wavefront in a fixed order of additions: DPP
// butterflies inside each 16-lane row, then the four row sums in row order.
// The result is uniform.
__device__ __forceinline__ double wave_total(double v)
{
  v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_mov<0x141>(v);  // row_half_mirror
  v += dpp_mov<0x140>(v);  // row_mirror
  return ((lane_value(v, 0) + lane_value(v, 16)) + lane_value(v, 32)) + lane_value(v, 48);
}

// Multi-right-hand-side SpMV over one chunk.  Lane (lbr, r) owns scalar row
// i = (br0 + pass*RPW + lbr)*BS + r; each block row is loaded once and applied
// to the gathered operands of all NR right-hand sides (operand elements shared
// across the block row's lanes by shuffles).  All lanes of a block row run
// the same trip counts, so every exchange reads an active lane.  Vectors
// interleave the right-hand sides: element i of k sits at i*NR + k.
// Without use_matrix the product is 0 (no block is read).
// gather(off, x[NR]) forms the operand at element offset off; own(off, m)
// loads what the epilogue needs of the owned row before the block loop;
// epi(off, y[NR], m) finishes the row.  Every right-hand side is computed;
// the epilogues keep the values the state machine must not change.
// Blocks whose loads are issued together: as many as keep the batch's
// registers (block row values + G gathered vectors of NR operands) within
// about 40, up to a whole 9-block stencil row.
template <class VT, class AT, int BS, int NR, int G> constexpr int spmv_batch()
{
  constexpr int regs = BS * (int) sizeof(VT) / 4 + G * NR * (int) sizeof(AT) / 4;
  return 40 / regs > 9 ? 9 : (40 / regs < 1 ? 1 : 40 / regs);
}

template <class VT, class AT, int BS, int NR, int G, class OWNT, class GATHER, class OWN, class EPI>
__device__ __forceinline__ void chunk_spmv(const ReqDesc &D, gptr<const VT> vals, int br0, bool use_matrix,
                                           const GATHER &gather, const OWN &own, const EPI &epi)
{
  constexpr int RPW = 64 / BS;
  constexpr int U = spmv_batch<VT, AT, BS, NR, G>();
  const int lane = threadIdx.x & 63;
  const int lbr = lane / BS, r = lane - lbr * BS, base_lane = lbr * BS;
  const gptr<const int> rp = glb(D.rp), cols = glb(D.cols);
  // a pass's row pointers are loaded while the previous pass computes
  auto row_bounds = [&](int ps, int &lo, int &hi) {
    const int brq = br0 + ps * RPW + lbr;
    lo = hi = 0;
    if (use_matrix && ps < SERVE_CHUNK_PASSES && lane < RPW * BS && brq < D.n_brows) {
      lo = ld(rp, (uint32_t) brq);
      hi = ld(rp, (uint32_t) brq + 1u);
    }
  };
  int next_b, next_b1;
  row_bounds(0, next_b, next_b1);
#pragma unroll 1
  for (int ps = 0; ps < SERVE_CHUNK_PASSES; ++ps) {
    const int br = br0 + ps * RPW + lbr;
    const bool on = lane < RPW * BS && br < D.n_brows;
    const uint32_t off = (uint32_t) (br * BS + r) * NR;
    int b = next_b, b1 = next_b1;
    row_bounds(ps + 1, next_b, next_b1);
    OWNT mine;
    if (on) own(off, mine);
    AT acc[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) acc[k] = 0;
    for (; b + U <= b1; b += U) {
      uint32_t xo[U];
      VT v[U][BS];
#pragma unroll
      for (int u = 0; u < U; ++u) xo[u] = (uint32_t) (ld(cols, (uint32_t) (b + u)) * BS + r) * NR;
#pragma unroll
      for (int u = 0; u < U; ++u) load_row<VT, BS>(vals + ((int64_t) (b + u) * BS + r) * BS, v[u]);
      AT xe[U][NR];
#pragma unroll
      for (int u = 0; u < U; ++u) gather(xo[u], xe[u]);
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int k = 0; k < NR; ++k)
#pragma unroll
          for (int c = 0; c < BS; ++c) acc[k] = fmadd((AT) v[u][c], group_bcast<BS>(xe[u][k], c, base_lane), acc[k]);
    }
    for (; U > 1 && b < b1; ++b) {
      VT v[BS];
      load_row<VT, BS>(vals + ((int64_t) b * BS + r) * BS, v);
      AT xe[NR];
      gather((uint32_t) (ld(cols, (uint32_t) b) * BS + r) * NR, xe);
#pragma unroll
      for (int k = 0; k < NR; ++k)
#pragma unroll
        for (int c = 0; c < BS; ++c) acc[k] = fmadd((AT) v[c], group_bcast<BS>(xe[k], c, base_lane), acc[k]);
    }
    if (on) epi(off, acc, mine);
  }
}

// Elementwise pass over the scalar rows of one chunk: f(off) per row.
template <int BS, int NR, class F> __device__ __forceinline__ void chunk_rows(const ReqDesc &D, int br0, const F &f)
{
  constexpr int RPW = 64 / BS;
  const int lane = threadIdx.x & 63;
#pragma unroll 1
  for (int ps = 0; ps < SERVE_CHUNK_PASSES; ++ps) {
    const int64_t row = (int64_t) (br0 + ps * RPW) * BS + lane;
    if (lane < RPW * BS && row < D.n) f((uint32_t) row * NR);
  }
}

template <int NR> struct ResidOwn {
  double xr[NR], b[NR], old[NR];
};

// x' = x + d (d counted as 0 where dval is clear), r = b - A64 x', r32 = r;
// partials ||r||^2 (FP64) and ||r32||^2.  With x = 0 (step 0) the matrix is
// not read: r = b.
template <int BS, int NR>
__device__ __forceinline__ void resid_chunk(const ReqDesc &D, const ReqState &S, int br0, Acc &acc)
{
  const int par = S.outer & 1;
  const unsigned live = S.live, dval = S.dval;
  const bool first = S.outer == 0;
  const gptr<const double> xs = glb((const double *) D.x[par]), b = glb(D.b);
  const gptr<double> xn = glb(D.x[1 - par]);
  const gptr<const float> d = glb((const float *) D.d);
  const gptr<float> r32 = glb(D.r32);
  auto gather = [&](uint32_t o, double (&x)[NR]) {
    if (first) {
#pragma unroll
      for (int k = 0; k < NR; ++k) x[k] = 0.0;
      return;
    }
    ldv<double, NR>(xs, o, x);
    if (dval) {
      float dd[NR];
      ldv<float, NR>(d, o, dd);
#pragma unroll
      for (int k = 0; k < NR; ++k)
        if ((dval >> k) & 1u) x[k] += (double) dd[k];
    }
  };
  auto own = [&](uint32_t o, ResidOwn<NR> &m) {
    gather(o, m.xr);
    ldv<double, NR>(b, o, m.b);
    if (live != (1u << NR) - 1u) ldv<double, NR>((gptr<const double>) xn, o, m.old);
  };
  chunk_spmv<double, double, BS, NR, 2, ResidOwn<NR>>(
      D, glb(D.v64), br0, !first, gather, own, [&](uint32_t o, const double (&y)[NR], const ResidOwn<NR> &m) {
        double xw[NR];
        float rw[NR];
#pragma unroll
        for (int k = 0; k < NR; ++k) {
          // a finished right-hand side keeps its iterate
          xw[k] = ((live >> k) & 1u) ? m.xr[k] : m.old[k];
          const double ri = m.b[k] - y[k];
          rw[k] = (float) ri;
          acc[k][0] += ri * ri;
          acc[k][1] += (double) rw[k] * (double) rw[k];
        }
        stv<double, NR>(xn, o, xw);
        stv<float, NR>(r32, o, rw);
      });
}

template <int NR> struct AOwn {
  float p[NR], rh[NR];
};

// p = ri + beta (p_prev - omega v_prev) (p = r32 at iteration 0), v = A32 p;
// partial (r32, v).
template <int BS, int NR>
__device__ __forceinline__ void inner_a_chunk(const ReqDesc &D, const ReqState &S, int br0, Acc &acc)
{
  const int it = S.inner_it;
  float beta[NR], omega[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    beta[k] = (float) S.k[k].beta;
    omega[k] = (float) S.k[k].omega;
  }
  const gptr<const float> pp = glb((const float *) D.p[(it - 1) & 1]), vp = glb((const float *) D.v[(it - 1) & 1]);
  const gptr<float> pc = glb(D.p[it & 1]), vc = glb(D.v[it & 1]);
  const gptr<const float> ri = glb((const float *) D.ri), rh = glb((const float *) D.r32);
  const gptr<const float> dinv = glb(D.dinv);
  const bool jac = D.dinv != nullptr;
  auto form_p = [&](uint32_t o, float (&x)[NR]) {
    if (it == 0) {
      ldv<float, NR>(rh, o, x);
      return;
    }
    float rr[NR], p0[NR], v0[NR];
    ldv<float, NR>(ri, o, rr);
    ldv<float, NR>(pp, o, p0);
    ldv<float, NR>(vp, o, v0);
#pragma unroll
    for (int k = 0; k < NR; ++k) x[k] = p_update(rr[k], p0[k], v0[k], beta[k], omega[k]);
  };
  // the SpMV operand: p, or D^-1 p with Jacobi (right preconditioning)
  auto gather = [&](uint32_t o, float (&x)[NR]) {
    form_p(o, x);
    if (jac) {
      const float m = ld(dinv, o / NR);
#pragma unroll
      for (int k = 0; k < NR; ++k) x[k] *= m;
    }
  };
  auto own = [&](uint32_t o, AOwn<NR> &m) {
    form_p(o, m.p);
    if (it == 0) {
#pragma unroll
      for (int k = 0; k < NR; ++k) m.rh[k] = m.p[k];
    } else {
      ldv<float, NR>(rh, o, m.rh);
    }
  };
  chunk_spmv<float, float, BS, NR, 3, AOwn<NR>>(D, glb(D.v32), br0, true, gather, own,
                                                [&](uint32_t o, const float (&y)[NR], const AOwn<NR> &m) {
                                                  stv<float, NR>(pc, o, m.p);
                                                  stv<float, NR>(vc, o, y);
#pragma unroll
                                                  for (int k = 0; k < NR; ++k)
                                                    acc[k][0] += (double) m.rh[k] * (double) y[k];
                                                });
}

template <int NR> struct BOwn {
  float s[NR];
};

// s = ri - alpha v (ri = r32 at iteration 0), t = A32 s; partials (t, s), (t, t).
template <int BS, int NR>
__device__ __forceinline__ void inner_b_chunk(const ReqDesc &D, const ReqState &S, int br0, Acc &acc)
{
  float alpha[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) alpha[k] = (float) S.k[k].alpha;
  const int it = S.inner_it;
  const gptr<const float> vcur = glb((const float *) D.v[it & 1]);
  const gptr<const float> ri = glb((const float *) (it == 0 ? D.r32 : D.ri));
  const gptr<float> s_out = glb(D.s), t_out = glb(D.t);
  const gptr<const float> dinv = glb(D.dinv);
  const bool jac = D.dinv != nullptr;
  auto form_s = [&](uint32_t o, float (&x)[NR]) {
    float rr[NR], vv[NR];
    ldv<float, NR>(ri, o, rr);
    ldv<float, NR>(vcur, o, vv);
#pragma unroll
    for (int k = 0; k < NR; ++k) x[k] = s_update(rr[k], vv[k], alpha[k]);
  };
  // the SpMV operand: s, or D^-1 s with Jacobi
  auto gather = [&](uint32_t o, float (&x)[NR]) {
    form_s(o, x);
    if (jac) {
      const float m = ld(dinv, o / NR);
#pragma unroll
      for (int k = 0; k < NR; ++k) x[k] *= m;
    }
  };
  chunk_spmv<float, float, BS, NR, 2, BOwn<NR>>(D, glb(D.v32), br0, true, gather,
                                                [&](uint32_t o, BOwn<NR> &m) { form_s(o, m.s); },
                                                [&](uint32_t o, const float (&y)[NR], const BOwn<NR> &m) {
                                                  stv<float, NR>(s_out, o, m.s);
                                                  stv<float, NR>(t_out, o, y);
#pragma unroll
                                                  for (int k = 0; k < NR; ++k) {
                                                    acc[k][0] += (double) y[k] * (double) m.s[k];
                                                    acc[k][1] += (double) y[k] * (double) y[k];
                                                  }
                                                });
}

// ri = s - omega t, d += alpha p + omega s, or d += D^-1 (alpha p + omega s)
// with Jacobi (d = 0 before iteration 0; a live right-hand side outside the
// inner solve gets d = 0 at iteration 0 and keeps d afterwards); partials
// (r32, ri), (ri, ri).
template <int BS, int NR>
__device__ __forceinline__ void inner_c_chunk(const ReqDesc &D, const ReqState &S, int br0, Acc &acc)
{
  const unsigned act = S.act;
  const int it = S.inner_it;
  float alpha[NR], omega[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    alpha[k] = (float) S.k[k].alpha;
    omega[k] = (float) S.k[k].omega;
  }
  const gptr<const float> pcur = glb((const float *) D.p[it & 1]), s_in = glb((const float *) D.s),
                          t_in = glb((const float *) D.t), rh = glb((const float *) D.r32);
  const gptr<float> d = glb(D.d), ri = glb(D.ri);
  const gptr<const float> dinv = glb(D.dinv);
  const bool jac = D.dinv != nullptr;
  chunk_rows<BS, NR>(D, br0, [&](uint32_t o) {
    float s[NR], p[NR], dd[NR], t[NR], h[NR], rn[NR];
    const float jm = jac ? ld(dinv, o / NR) : 1.f;
    ldv<float, NR>(s_in, o, s);
    ldv<float, NR>(pcur, o, p);
    if (it == 0) {
#pragma unroll
      for (int k = 0; k < NR; ++k) dd[k] = 0.f;
    } else {
      ldv<float, NR>((gptr<const float>) d, o, dd);
    }
    ldv<float, NR>(t_in, o, t);
    ldv<float, NR>(rh, o, h);
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      rn[k] = fmadd(-omega[k], t[k], s[k]);
      if ((act >> k) & 1u)
        dd[k] = jac ? fmadd(jm, fmadd(omega[k], s[k], alpha[k] * p[k]), dd[k])
                    : fmadd(omega[k], s[k], fmadd(alpha[k], p[k], dd[k]));
      acc[k][0] += (double) h[k] * (double) rn[k];
      acc[k][1] += (double) rn[k] * (double) rn[k];
    }
    stv<float, NR>(d, o, dd);
    stv<float, NR>(ri, o, rn);
  });
}

#define BMFP_SERVE_BS_SWITCH(bs, CALL)           \
  switch (bs) {                                  \
  case 1: { constexpr int BS = 1; CALL; } break; \
  case 2: { constexpr int BS = 2; CALL; } break; \
  case 3: { constexpr int BS = 3; CALL; } break; \
  case 4: { constexpr int BS = 4; CALL; } break; \
  case 6: { constexpr int BS = 6; CALL; } break; \
  case 8: { constexpr int BS = 8; CALL; } break; \
  default: break;                                \
  }
#define BMFP_SERVE_NR_SWITCH(nr, CALL)           \
  switch (nr) {                                  \
  case 1: { constexpr int NR = 1; CALL; } break; \
  case 2: { constexpr int NR = 2; CALL; } break; \
  case 3: { constexpr int NR = 3; CALL; } break; \
  case 4: { constexpr int NR = 4; CALL; } break; \
  default: break;                                \
  }

template <int SLOT, int BS, int NR>
__device__ __forceinline__ void run_tile(const ReqDesc &D, const ReqState &S, int phase, const TileDesc &T,
                                         gptr<double> part)
{
  const int lane = threadIdx.x & 63;
#pragma unroll 1
  for (int c = 0; c < T.passes; ++c) {
    const int chunk = T.chunk + c;
    const int br0 = chunk * serve_chunk_brows(BS);
    Acc acc = {};
    if constexpr (SLOT == 0) {
      resid_chunk<BS, NR>(D, S, br0, acc);
    } else if constexpr (SLOT == 1) {
      inner_a_chunk<BS, NR>(D, S, br0, acc);
    } else if constexpr (SLOT == 2) {
      inner_b_chunk<BS, NR>(D, S, br0, acc);
    } else {
      inner_c_chunk<BS, NR>(D, S, br0, acc);
    }
    // per-chunk partial sums (0 for absent right-hand sides), lane j < 8 writes value j
    double mine = 0.0;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const double s0 = wave_total(acc[k][0]), s1 = wave_total(acc[k][1]);
      mine = lane == 2 * k ? s0 : (lane == 2 * k + 1 ? s1 : mine);
    }
    if (lane < SERVE_PART) *(part + (int64_t) (D.chunk0 + chunk) * SERVE_PART + lane) = mine;
  }
}

// ---- slot kernels ----------------------------------------------------------------
// Two launch shapes, results identical (partial sums are per chunk):
//  * persistent = 0: one wavefront per tile, the grid sized to the tile table
//    of the round (the hardware overlaps starting and finishing wavefronts);
//  * persistent = 1: a fixed grid whose wavefronts each walk the table.
// With xcd_run = K > 0 the table is dealt to the 8 XCDs in runs of K
// consecutive tiles (run j to XCD j mod 8; workgroups are dispatched to the
// XCDs round-robin), so neighbouring chunks of a request, which gather the
// same operand rows, share one L2 while every XCD still sees tiles from the
// whole table.  With 0 the tiles go to wavefronts in table order.  Grids are
// multiples of 8 workgroups when K > 0.
template <int SLOT>
__device__ __forceinline__ void slot_tile(const ReqDesc *__restrict__ rd, const ReqState *__restrict__ rs,
                                          const TileDesc *__restrict__ tiles, double *__restrict__ part, int ti)
{
  const TileDesc T = tiles[ti];
  const ReqState &S = rs[T.req];
  const int phase = S.phase;
  if (!slot_owns<SLOT>(phase)) return;
  const ReqDesc &D = rd[T.req];
  BMFP_SERVE_BS_SWITCH(D.bs, BMFP_SERVE_NR_SWITCH(D.nrhs, (run_tile<SLOT, BS, NR>(D, S, phase, T, glb(part)))));
}

template <int SLOT>
__global__ __launch_bounds__(256) void k_slot(const ReqDesc *__restrict__ rd, const ReqState *__restrict__ rs,
                                              const TileDesc *__restrict__ tiles, const int *__restrict__ ntp,
                                              double *__restrict__ part, int xcd_run, int persistent)
{
  const int nt = *ntp;
  const int wib = (int) (threadIdx.x >> 6);
  if (!persistent) {
    int64_t ti = (int64_t) blockIdx.x * 4 + wib;
    if (xcd_run > 0) {
      const int64_t l = (int64_t) (blockIdx.x >> 3) * 4 + wib;
      ti = (l / xcd_run * 8 + (blockIdx.x & 7)) * xcd_run + l % xcd_run;
    }
    if (ti < nt) slot_tile<SLOT>(rd, rs, tiles, part, __builtin_amdgcn_readfirstlane((int) ti));
    return;
  }
  if (xcd_run > 0) {
    const int xcd = (int) blockIdx.x & 7;
    const int lw = __builtin_amdgcn_readfirstlane((int) (blockIdx.x >> 3) * 4 + wib);
    const int nwx = (int) (gridDim.x >> 3) * 4;
    for (int j = lw;; j += nwx) {
      const int64_t ti64 = ((int64_t) (j / xcd_run) * 8 + xcd) * xcd_run + j % xcd_run;
      if (ti64 >= nt) break;
      slot_tile<SLOT>(rd, rs, tiles, part, (int) ti64);
    }
    return;
  }
  const int wave = __builtin_amdgcn_readfirstlane((int) blockIdx.x * 4 + wib);
  const int nw = (int) gridDim.x * 4;
  for (int ti = wave; ti < nt; ti += nw) slot_tile<SLOT>(rd, rs, tiles, part, ti);
}

// ---- finalize: one wavefront per request, lane k owns right-hand side k --------
template <int SLOT>
__global__ __launch_bounds__(256) void k_fin(const ReqDesc *__restrict__ rd, ReqState *__restrict__ rs,
                                             const int *__restrict__ list, const int *__restrict__ nlp,
                                             const double *__restrict__ part)
{
  const int w = (int) ((blockIdx.x * 256 + threadIdx.x) >> 6), lane = threadIdx.x & 63;
  if (w >= *nlp) return;
  const int q = list[w];
  ReqState &S = rs[q];
  const int phase = S.phase;
  if (!slot_owns<SLOT>(phase)) return;
  const ReqDesc &D = rd[q];
  double a[SERVE_PART];
#pragma unroll
  for (int j = 0; j < SERVE_PART; ++j) a[j] = 0.0;
  for (int c = lane; c < D.n_chunks; c += 64) {
    const double *pc = part + (int64_t) (D.chunk0 + c) * SERVE_PART;
#pragma unroll
    for (int j = 0; j < SERVE_PART; ++j) a[j] += pc[j];
  }
#pragma unroll
  for (int j = 0; j < SERVE_PART; ++j) a[j] = wave_sum(a[j]);
  double a0 = 0.0, a1 = 0.0;
#pragma unroll
  for (int k = 0; k < SERVE_MAX_RHS; ++k)
    if (lane == k) {
      a0 = a[2 * k];
      a1 = a[2 * k + 1];
    }
  const unsigned live = S.live, act = S.act;
  const int outer = S.outer, it = S.inner_it;
  const int k = lane;
  bool keep = false, inner = false;
  if (k < D.nrhs) {
    RhsState &R = S.k[k];
    if (SLOT == 0) {
      if ((live >> k) & 1u) {
        // a0 = ||r||^2 (FP64), a1 = ||r32||^2 for x' = x + d (now in x[1 - outer%2])
        if (outer == 0) R.norm_b2 = a0;  // x = 0: r = b
        R.outer = outer + 1;
        // b = 0 gives x = 0 exactly; a non-finite ||b|| or ||r|| gives a NaN rel
        const double rel = R.norm_b2 == 0.0 ? 0.0 : sqrt(a0) / sqrt(R.norm_b2);
        R.rel = rel;
        if (!isfinite(rel)) R.status = BMFP_ERR_NONFINITE;
        else if (rel < D.tol) R.status = BMFP_OK;
        else if (R.outer > D.max_outer) R.status = BMFP_ERR_NOT_CONVERGED;
        else keep = true;
        if (keep) {
          // the inner solve starts: ri = rh = r32, rho_0 = (rh, ri)
          R.inner_rr0 = a1;
          R.inner_it = 0;
          R.rho = a1;
          R.alpha = 1.0;
          R.omega = 1.0;
          R.beta = a1;
          inner = !(a1 == 0.0 || !isfinite(a1));
        }
      }
    } else if ((act >> k) & 1u) {
      if (SLOT == 1) {
        // a0 = (rh, v); a breakdown ends the inner solve, d keeps the last iterate
        if (a0 == 0.0 || !isfinite(a0)) {
          R.inner_total += it;
        } else {
          R.alpha = R.rho / a0;
          keep = true;
        }
      } else if (SLOT == 2) {
        // a0 = (t, s), a1 = (t, t)
        R.tt = a1;
        R.omega = a1 == 0.0 ? 0.0 : a0 / a1;
        keep = true;
      } else {
        // a0 = (rh, ri), a1 = ||ri||^2 after the update
        R.inner_it = it + 1;
        const double rel_in = R.inner_rr0 > 0.0 ? sqrt(a1) / sqrt(R.inner_rr0) : 0.0;
        const bool stop = !(rel_in >= D.inner_tol) || it + 1 >= D.max_inner || R.tt == 0.0 || R.omega == 0.0 ||
                          a0 == 0.0 || !isfinite(a0);
        if (stop) {
          R.inner_total += it + 1;
        } else {
          R.beta = a0 / R.rho * R.alpha / R.omega;
          R.rho = a0;
          keep = true;
        }
      }
    }
  }
  const unsigned kept = (unsigned) (__ballot(keep) & 0xFull);
  const unsigned inner_set = (unsigned) (__ballot(inner) & 0xFull);
  if (lane == 0) {
    if (SLOT == 0) {
      S.outer = outer + 1;
      S.inner_it = 0;
      S.live = kept;
      S.act = inner_set;
      S.dval = 0u;
      S.phase = kept == 0u ? REQ_DONE : (inner_set ? REQ_INNER_A : REQ_RESID);
    } else if (SLOT == 1) {
      S.act = kept;
      S.phase = kept ? REQ_INNER_B : REQ_RESID;
    } else if (SLOT == 2) {
      S.phase = REQ_INNER_C;
    } else {
      // d now holds this step's correction for every live right-hand side
      S.inner_it = it + 1;
      S.dval = live;
      S.act = kept;
      S.phase = kept ? REQ_INNER_A : REQ_RESID;
    }
  }
}

// ---- admission-time device work (the stencil generator is in stencil_kernels.hpp)
// Stencil block row pointers in closed form: the nodes before (iy, ix) hold
// F(iy, ny) * Sx + f(iy, ny) * F(ix, nx) blocks, with f(i, m) = 3 - [i = 0] -
// [i = m-1] neighbours per axis, F(i, m) = sum of f over the first i indices
// and Sx = F(nx, nx).
__device__ __forceinline__ int64_t stencil_prefix(int i, int m) { return 3 * (int64_t) i - (i >= 1) - (i >= m); }

__global__ __launch_bounds__(256) void k_stencil_rp(int nx, int ny, int *__restrict__ rp)
{
  const int64_t nodes = (int64_t) nx * ny;
  const int64_t sx = stencil_prefix(nx, nx);
  for (int64_t i = (int64_t) blockIdx.x * 256 + threadIdx.x; i <= nodes; i += (int64_t) gridDim.x * 256) {
    int64_t v;
    if (i == nodes) {
      v = stencil_prefix(ny, ny) * sx;
    } else {
      const int iy = (int) (i / nx), ix = (int) (i - (int64_t) iy * nx);
      const int64_t fy = 3 - (iy == 0) - (iy == ny - 1);
      v = stencil_prefix(iy, ny) * sx + fy * stencil_prefix(ix, nx);
    }
    rp[i] = (int) v;
  }
}

__global__ __launch_bounds__(256) void k_to_f32(const double *__restrict__ a, float *__restrict__ b, int64_t n)
{
  for (int64_t i = (int64_t) blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t) gridDim.x * 256) b[i] = (float) a[i];
}

// Jacobi preconditioner of a request: 1 / a_ii per scalar row from the FP64
// row-major blocks, 1 where the diagonal block is absent or |a_ii| < 1e-20
// (the rule of the BiCGStab solver's k_jacobi), stored in FP32.
__global__ __launch_bounds__(256) void k_serve_dinv(int n_brows, int bs, const int *__restrict__ rp,
                                                    const int *__restrict__ cols, const double *__restrict__ v64,
                                                    float *__restrict__ dinv)
{
  const int64_t n = (int64_t) n_brows * bs;
  for (int64_t row = (int64_t) blockIdx.x * 256 + threadIdx.x; row < n; row += (int64_t) gridDim.x * 256) {
    const int br = (int) (row / bs), r = (int) (row - (int64_t) br * bs);
    double a = 0.0;
    bool found = false;
    for (int b = rp[br]; b < rp[br + 1]; ++b)
      if (cols[b] == br) {
        a = v64[((int64_t) b * bs + r) * bs + r];
        found = true;
        break;
      }
    dinv[row] = (float) ((!found || fabs(a) < 1e-20) ? 1.0 : 1.0 / a);
  }
}

// b of a request with nr right-hand sides, interleaved (element i of k at
// i*nr + k): uniform[-1,1] keyed by (seed, k*n + i).
__global__ __launch_bounds__(256) void k_rhs(double *__restrict__ b, int64_t n, int nr, uint64_t seed)
{
  const uint64_t key = rng_key(seed, STREAM_RHS);
  for (int64_t t = (int64_t) blockIdx.x * 256 + threadIdx.x; t < n * nr; t += (int64_t) gridDim.x * 256) {
    const int64_t i = t / nr, k = t - i * nr;
    b[t] = rng_sym(key, (uint64_t) (k * n + i));
  }
}

// Solutions in right-hand-side-major order: out[k*n + i] = x_par(k)[i*nr + k],
// where bit k of par_mask names the iterate buffer holding right-hand side k.
__global__ __launch_bounds__(256) void k_gather_solution(const double *__restrict__ x0, const double *__restrict__ x1,
                                                         unsigned par_mask, double *__restrict__ out, int64_t n,
                                                         int nr)
{
  for (int64_t t = (int64_t) blockIdx.x * 256 + threadIdx.x; t < n * nr; t += (int64_t) gridDim.x * 256) {
    const int64_t k = t / n, i = t - k * n;
    out[t] = (((par_mask >> k) & 1u) ? x1 : x0)[i * nr + k];
  }
}

// Initial state of a request: every right-hand side live, refinement step 0.
// `phase`, `act` and the scalars allow forcing a phase (schedule tuning).
__global__ void k_req_init(ReqState *__restrict__ st, int nrhs, int phase)
{
  if (threadIdx.x != 0) return;
  ReqState S{};
  S.phase = phase;
  S.live = S.act = (1u << nrhs) - 1u;
  S.dval = 0u;
  S.inner_it = phase == REQ_RESID ? 0 : 1;
  for (int k = 0; k < SERVE_MAX_RHS; ++k) {
    S.k[k] = RhsState{};
    S.k[k].status = BMFP_OK;
    S.k[k].alpha = S.k[k].omega = S.k[k].beta = 0.5;
    S.k[k].rho = 1.0;
    S.k[k].inner_it = phase == REQ_RESID ? 0 : 1;
  }
  *st = S;
}

}  // namespace
}  // namespace bmfp
#include "serve_host.inc"
