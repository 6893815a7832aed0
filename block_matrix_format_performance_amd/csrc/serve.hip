// serve.hip -- continuous-batching request loop (bmfp_server_*).
//
// A request is an FP64 BCSR system with 1..4 right-hand sides, solved by FP64
// iterative refinement around FP32 BiCGStab.  Admitted requests advance
// together through a per-request state machine on the device; the right-hand
// sides of one request move in lockstep so that every block row of its
// matrix is read once per SpMV for all of them.  Each phase is one or two
// slot kernels (an elementwise pass that stores the SpMV's operand, then the
// SpMV, which gathers that one vector) and a finalize kernel:
//
//   RESID    slot X  x' = x + d                                          (elementwise)
//            slot R  r = b - A64 x', r32 = r                             fin0: DONE | INNER_A | RESID
//   INNER_A  slot P  p = ri + beta (p - omega v)                         (elementwise)
//            slot A  v = A32 p                                           fin1: INNER_B | RESID
//   INNER_B  slot S  s = ri - alpha v                                    (elementwise)
//            slot B  t = A32 s                                           fin2: INNER_C
//   INNER_C  slot C  ri = s - omega t, d += alpha p + omega s            fin3: INNER_A | RESID
//
// Forming p, s or x' at every gathered index instead (two or three vector
// loads per gathered element) held these SpMVs at ~3.8 TB/s of compulsory
// traffic; the stored operand costs one more vector pass per phase.
//
// The inner solve starts from ri = r32 with the shadow residual r32 itself
// (it stays constant until the next RESID), p = v = d = 0: iteration 0 reads
// r32 in place of ri and p and skips the zero terms, so no initialisation
// pass is needed.  With x = 0, the first RESID is r = b and reads no matrix.
//
// Per right-hand side k three bit masks select the work: `live` (not yet
// converged or failed) for RESID, `act` (still inside the current inner
// solve) for the rest, and `dval` (d holds this refinement step's correction;
// otherwise it counts as 0).  A right-hand side whose inner solve stops is
// frozen until the others stop too, so its arithmetic is exactly that of
// solving it alone.
//
// One batched iteration is eleven kernels (X R fin0 P A fin1 S B fin2 C fin3),
// captured iterations_per_round times in one HIP graph.  Slot kernels give
// every wavefront its own tiles from a tile table (no workgroup barriers); a
// tile is `passes` consecutive chunks of one request, a chunk is 4 wavefront
// passes of 64/bs block rows.  Partial sums are kept per chunk and the fin
// kernels (one wavefront per request) reduce a request's chunks in chunk
// order, so the arithmetic of a request depends neither on the rest of the
// batch nor on the tile schedule: results are bitwise identical alone, in any
// mix, and under any schedule the autotuner picks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <type_traits>
#include <vector>

#include "matrix.hpp"
#include "solver_kernels.hpp"
#include "stencil_kernels.hpp"

namespace bmfp {

constexpr int SERVE_MAX_RHS = 4;
constexpr int SERVE_PART = 2 * SERVE_MAX_RHS;  // partial sums per chunk
constexpr int SERVE_CHUNK_PASSES = 4;          // wavefront passes per chunk
constexpr int SERVE_F16_MIN_BS = 4;            // smallest block size with FP16 inner matrices

enum : int { REQ_EMPTY = 0, REQ_RESID = 1, REQ_INNER_A = 2, REQ_INNER_B = 3, REQ_INNER_C = 4, REQ_DONE = 5 };

// Device view of one request.  Vectors interleave the nrhs right-hand sides:
// element i of right-hand side k is at i*nrhs + k.
struct ReqDesc {
  int bs, n_brows, nrhs, chunk0, n_chunks;
  int sl_j;   // > 0: the inner matrix is pass-sliced with sl_j block slots per row; 0: BCSR order
  int sl64;   // 1: v64 is pass-sliced like the inner matrix (same slots and cols_sl); 0: BCSR order
  int64_t n;  // scalar rows of one right-hand side
  const int *rp;
  const int *cols;
  const int *cols_sl;   // block columns of the pass-sliced inner matrix (sl_j > 0)
  const double *v64;
  const float *v32;     // FP32 inner matrix, or null
  const __half *v16;    // FP16 inner matrix (one FP32 scale per block row in sc16), or null
  const float *sc16;
  const float *dinv;  // Jacobi: 1 / diagonal entry per scalar row (FP32); null without
  const double *b;
  double *x[2];
  float *r32, *d, *ri, *p[2], *v[2], *s, *t;
  double tol, inner_tol;
  int max_outer, max_inner;
};

struct RhsState {
  int status, outer, inner_it, inner_total;
  double norm_b2, rel, rho, alpha, omega, beta, tt, inner_rr0, inner_tol;
};

// A refinement step whose outer residual rel is already close to the outer
// tolerance tol stops its inner solve at SERVE_LAST_STEP_MARGIN * tol / rel
// when that is looser than the inner tolerance: the next residual is then
// about (that + the inner precision's own error) * rel, below tol, and the
// step reads the inner matrix for fewer iterations.
constexpr double SERVE_LAST_STEP_MARGIN = 0.6;

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

// The tile table lists the tiles of each served block size together.  The
// counts array holds the table size, the finalize list size, then per block
// size (1, 2, 3, 4, 6, 8, in that order) the first tile and the tile count.
constexpr int SERVE_NBS = 6;
__host__ __device__ constexpr int serve_bs_index(int bs)
{
  return bs <= 4 ? bs - 1 : bs == 6 ? 4 : 5;
}
constexpr int SERVE_COUNTS = 2 + 2 * SERVE_NBS;

__host__ __device__ constexpr int serve_chunk_brows(int bs) { return SERVE_CHUNK_PASSES * (64 / bs); }

namespace {

// Slot kernels, in launch order within a batched iteration, and the phase
// each one serves; finalize kernel F serves phase REQ_RESID + F.
enum : int { SLOT_X = 0, SLOT_R = 1, SLOT_P = 2, SLOT_A = 3, SLOT_S = 4, SLOT_B = 5, SLOT_C = 6, SERVE_SLOTS = 7 };
__host__ __device__ constexpr int slot_phase(int slot)
{
  return slot <= SLOT_R ? REQ_RESID : slot <= SLOT_A ? REQ_INNER_A : slot <= SLOT_B ? REQ_INNER_B : REQ_INNER_C;
}
// slots that leave per-chunk partial sums for their phase's finalize kernel
__host__ __device__ constexpr bool slot_reduces(int slot)
{
  return slot == SLOT_R || slot == SLOT_A || slot == SLOT_B || slot == SLOT_C;
}
template <int F> __device__ __forceinline__ bool fin_owns(int phase) { return phase == REQ_RESID + F; }

using Acc = double[SERVE_MAX_RHS][2];

// Pointers read from the request tables are generic; accesses through these
// global-address-space views compile to global (not flat) loads and stores.
template <class T> using gptr = __attribute__((address_space(1))) T *;
template <class T> __device__ __forceinline__ gptr<T> glb(T *p) { return (gptr<T>) p; }

// Element `off` of an array: the byte offset fits 32 bits (submission
// refuses requests of 2^30 or more blocks or 2^27 or more rows times
// right-hand sides), so the access is a uniform base plus a 32-bit lane offset.
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
// the alignment of a block row allows (up to 16 bytes), copied out of
// 32-bit lanes (so 2-byte elements load as whole words too).
template <class T, int N> __device__ __forceinline__ void load_row(gptr<const T> p, T (&v)[N])
{
  constexpr int BYTES = (int) sizeof(T) * N;
  constexpr int CH = lowbit(BYTES) < 16 ? lowbit(BYTES) : 16;
  const gptr<const char> src = (gptr<const char>) p;
  char *dst = reinterpret_cast<char *>(v);
  if constexpr (CH >= 4) {
    typedef unsigned int w_t __attribute__((ext_vector_type(CH / 4)));
#pragma unroll
    for (int c = 0; c < BYTES; c += CH) {
      const w_t t = *(gptr<const w_t>) (src + c);
      __builtin_memcpy(dst + c, &t, CH);
    }
  } else {
#pragma unroll
    for (int c = 0; c < N; ++c) v[c] = p[c];
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
// gather(off, x[NR]) loads the operand at element offset off; own(off, m)
// loads what the epilogue needs of the owned row before the block loop;
// epi(off, y[NR], m) finishes the row.  Every right-hand side is computed;
// the epilogues keep the values the state machine must not change.
// Blocks whose loads are issued together: as many as keep the batch's
// registers (block row values + the NR gathered operands) within about 40,
// up to a whole 9-block stencil row.
template <class VT, class AT, int BS, int NR> constexpr int spmv_batch()
{
  // FP16 rows are sized as FP32 ones: they are converted before use
  constexpr int regs = BS * ((int) sizeof(VT) < 4 ? 4 : (int) sizeof(VT)) / 4 + NR * (int) sizeof(AT) / 4;
  return 40 / regs > 9 ? 9 : (40 / regs < 1 ? 1 : 40 / regs);
}

template <class VT, class AT, int BS, int NR, class OWNT, int UB = 0, class GATHER, class OWN, class EPI>
__device__ __forceinline__ void chunk_spmv(const ReqDesc &D, gptr<const VT> vals, int br0, bool use_matrix,
                                           const GATHER &gather, const OWN &own, const EPI &epi,
                                           gptr<const float> scale = nullptr)
{
  constexpr int RPW = 64 / BS;
  constexpr int U = UB > 0 ? UB : spmv_batch<VT, AT, BS, NR>();
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
    // FP16 values: the block row's scale multiplies the finished sums
    const AT sc = (scale && on) ? (AT) ld(scale, (uint32_t) br) : (AT) 1;
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
    if (scale)
#pragma unroll
      for (int k = 0; k < NR; ++k) acc[k] *= sc;
    if (on) epi(off, acc, mine);
  }
}

// Pass-sliced inner matrices.  A pass is the RPW = 64/BS consecutive block
// rows one wavefront covers; the inner (FP32 or FP16) copy stores pass P's
// blocks slot by slot: block j of the pass's row lbr is sliced block
// (P*J + j)*RPW + lbr, J = the longest row rounded up to a multiple of 3, and
// rows shorter than J are padded with zero blocks (column: the row's last
// block column, 0 for an empty row).  Row r of that block then starts at
// element ((P*J + j)*RPW*BS + lane)*BS for lane = lbr*BS + r: every load of a
// wavefront reads one contiguous range (RPW*BS*BS values per slot) instead of
// RPW separate blocks, and every lane of the wavefront runs the same J
// iterations.  Products and their order are those of the BCSR path (a zero
// block adds 0 * x).
// Without use_matrix the product is 0 (no block is read).
template <class VT, class AT, int BS, int NR, class OWNT, class GATHER, class OWN, class EPI>
__device__ __forceinline__ void chunk_spmv_sl(const ReqDesc &D, gptr<const VT> vals, int br0, bool use_matrix,
                                              const GATHER &gather, const OWN &own, const EPI &epi,
                                              gptr<const float> scale = nullptr)
{
  constexpr int RPW = 64 / BS, LPW = RPW * BS, U = 3;
  const int lane = threadIdx.x & 63;
  const int ln = lane < LPW ? lane : LPW - 1;  // spare lanes repeat the last row's loads
  const int lbr = ln / BS, r = ln - lbr * BS, base_lane = lbr * BS;
  const gptr<const int> cs = glb(D.cols_sl);
  const int J = use_matrix ? D.sl_j : 0;  // a multiple of U
  // passes of this chunk that hold rows (uniform)
  const int passes = min(SERVE_CHUNK_PASSES, (D.n_brows - br0 + RPW - 1) / RPW);
  // operand offsets of slots q0..q0+U-1 of the chunk's slot sequence (pass
  // by pass): loaded one batch ahead, so each batch waits on one round trip
  auto load_xo = [&](int q0, uint32_t (&xo)[U]) {
    const int ps = q0 / J, j0 = q0 - ps * J;
    const int64_t sb = ((int64_t) (br0 / RPW + ps) * J + j0) * RPW + lbr;
#pragma unroll
    for (int u = 0; u < U; ++u) xo[u] = (uint32_t) (ld(cs, (uint32_t) (sb + u * RPW)) * BS + r) * NR;
  };
  uint32_t xo[U];
  if (passes > 0 && J > 0) load_xo(0, xo);
#pragma unroll 1
  for (int ps = 0; ps < passes; ++ps) {
    const int br = br0 + ps * RPW + lbr;
    const bool on = lane < LPW && br < D.n_brows;
    const uint32_t off = (uint32_t) (br * BS + r) * NR;
    OWNT mine;
    if (on) own(off, mine);
    const AT sc = (scale && on) ? (AT) ld(scale, (uint32_t) br) : (AT) 1;
    AT acc[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) acc[k] = 0;
    const int64_t s0 = (int64_t) (br0 / RPW + ps) * J * RPW;  // the pass's first sliced block
    for (int j = 0; j < J; j += U) {
      VT v[U][BS];
#pragma unroll
      for (int u = 0; u < U; ++u) load_row<VT, BS>(vals + ((s0 + (int64_t) (j + u) * RPW) * BS + ln) * BS, v[u]);
      AT xe[U][NR];
#pragma unroll
      for (int u = 0; u < U; ++u) gather(xo[u], xe[u]);
      const int qn = ps * J + j + U;  // the next batch, in this pass or the next
      if (qn < passes * J) load_xo(qn, xo);
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int k = 0; k < NR; ++k)
#pragma unroll
          for (int c = 0; c < BS; ++c) acc[k] = fmadd((AT) v[u][c], group_bcast<BS>(xe[u][k], c, base_lane), acc[k]);
    }
    if (scale)
#pragma unroll
      for (int k = 0; k < NR; ++k) acc[k] *= sc;
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

// x' = x + d (d counted as 0 where dval is clear; a finished right-hand side
// keeps its iterate; x' = 0 at step 0).
template <int BS, int NR>
__device__ __forceinline__ void resid_x_chunk(const ReqDesc &D, const ReqState &S, int br0)
{
  const int par = S.outer & 1;
  const unsigned live = S.live, dval = S.dval;
  const bool first = S.outer == 0;
  const gptr<const double> xs = glb((const double *) D.x[par]);
  const gptr<double> xn = glb(D.x[1 - par]);
  const gptr<const float> d = glb((const float *) D.d);
  chunk_rows<BS, NR>(D, br0, [&](uint32_t o) {
    double xw[NR];
    if (first) {
#pragma unroll
      for (int k = 0; k < NR; ++k) xw[k] = 0.0;
    } else {
      double xr[NR], old[NR];
      ldv<double, NR>(xs, o, xr);
      if (dval) {
        float dd[NR];
        ldv<float, NR>(d, o, dd);
#pragma unroll
        for (int k = 0; k < NR; ++k)
          if ((dval >> k) & 1u) xr[k] += (double) dd[k];
      }
      if (live != (1u << NR) - 1u) ldv<double, NR>((gptr<const double>) xn, o, old);
#pragma unroll
      for (int k = 0; k < NR; ++k) xw[k] = ((live >> k) & 1u) ? xr[k] : old[k];
    }
    stv<double, NR>(xn, o, xw);
  });
}

template <int NR> struct ResidOwn {
  double b[NR];
};

// r = b - A64 x', r32 = r; partials ||r||^2 (FP64) and ||r32||^2.  With x = 0
// (step 0) the matrix is not read: r = b.  A pass-sliced FP64 copy (stencil
// requests) is read slot by slot; its products and their order are those of
// the BCSR path, which reads one block at a time (batched BCSR loads beside
// the sliced path cost the kernel a wave per SIMD at bs 6).
template <int BS, int NR>
__device__ __forceinline__ void resid_chunk(const ReqDesc &D, const ReqState &S, int br0, Acc &acc)
{
  const int par = S.outer & 1;
  const bool first = S.outer == 0;
  const gptr<const double> xn = glb((const double *) D.x[1 - par]), b = glb(D.b);
  const gptr<float> r32 = glb(D.r32);
  auto gather = [&](uint32_t o, double (&x)[NR]) { ldv<double, NR>(xn, o, x); };
  auto own = [&](uint32_t o, ResidOwn<NR> &m) { ldv<double, NR>(b, o, m.b); };
  auto epi = [&](uint32_t o, const double (&y)[NR], const ResidOwn<NR> &m) {
    float rw[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const double ri = m.b[k] - y[k];
      rw[k] = (float) ri;
      acc[k][0] += ri * ri;
      acc[k][1] += (double) rw[k] * (double) rw[k];
    }
    stv<float, NR>(r32, o, rw);
  };
  if (D.sl64) chunk_spmv_sl<double, double, BS, NR, ResidOwn<NR>>(D, glb(D.v64), br0, !first, gather, own, epi);
  else chunk_spmv<double, double, BS, NR, ResidOwn<NR>, 1>(D, glb(D.v64), br0, !first, gather, own, epi);
}

// p = ri + beta (p - omega v) for iterations >= 1 (iteration 0 uses p = r32,
// read in place by the other phases).
template <int BS, int NR>
__device__ __forceinline__ void inner_p_chunk(const ReqDesc &D, const ReqState &S, int br0)
{
  if (S.inner_it == 0) return;
  float beta[NR], omega[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    beta[k] = (float) S.k[k].beta;
    omega[k] = (float) S.k[k].omega;
  }
  const gptr<float> p = glb(D.p[0]);
  // the previous p: r32 after iteration 0, which stores none
  const gptr<const float> pp = S.inner_it == 1 ? glb((const float *) D.r32) : (gptr<const float>) p;
  const gptr<const float> v = glb((const float *) D.v[0]), ri = glb((const float *) D.ri);
  chunk_rows<BS, NR>(D, br0, [&](uint32_t o) {
    float rr[NR], p0[NR], v0[NR], pn[NR];
    ldv<float, NR>(ri, o, rr);
    ldv<float, NR>(pp, o, p0);
    ldv<float, NR>(v, o, v0);
#pragma unroll
    for (int k = 0; k < NR; ++k) pn[k] = p_update(rr[k], p0[k], v0[k], beta[k], omega[k]);
    stv<float, NR>(p, o, pn);
  });
}

template <int NR> struct RhOwn {
  float rh[NR];
};

// v = A32 q, q = p (r32 at iteration 0), or D^-1 p with Jacobi (right
// preconditioning, scaled at the gathered index); partial (r32, v).
template <int BS, int NR>
__device__ __forceinline__ void inner_a_chunk(const ReqDesc &D, const ReqState &S, int br0, Acc &acc)
{
  const gptr<const float> rh = glb((const float *) D.r32);
  const gptr<const float> q = S.inner_it == 0 ? rh : glb((const float *) D.p[0]);
  const gptr<float> v = glb(D.v[0]);
  const gptr<const float> dinv = glb(D.dinv);
  const bool jac = D.dinv != nullptr;
  auto gather = [&](uint32_t o, float (&x)[NR]) {
    ldv<float, NR>(q, o, x);
    if (jac) {
      const float m = ld(dinv, o / NR);
#pragma unroll
      for (int k = 0; k < NR; ++k) x[k] *= m;
    }
  };
  auto own = [&](uint32_t o, RhOwn<NR> &m) { ldv<float, NR>(rh, o, m.rh); };
  auto epi = [&](uint32_t o, const float (&y)[NR], const RhOwn<NR> &m) {
    stv<float, NR>(v, o, y);
#pragma unroll
    for (int k = 0; k < NR; ++k) acc[k][0] += (double) m.rh[k] * (double) y[k];
  };
  if constexpr (BS >= SERVE_F16_MIN_BS)
    if (D.v16) {
      if (D.sl_j) chunk_spmv_sl<__half, float, BS, NR, RhOwn<NR>>(D, glb(D.v16), br0, true, gather, own, epi, glb(D.sc16));
      else chunk_spmv<__half, float, BS, NR, RhOwn<NR>, 1>(D, glb(D.v16), br0, true, gather, own, epi, glb(D.sc16));
      return;
    }
  if (D.sl_j) chunk_spmv_sl<float, float, BS, NR, RhOwn<NR>>(D, glb(D.v32), br0, true, gather, own, epi);
  else chunk_spmv<float, float, BS, NR, RhOwn<NR>, 1>(D, glb(D.v32), br0, true, gather, own, epi);
}

// s = ri - alpha v (ri = r32 at iteration 0).
template <int BS, int NR>
__device__ __forceinline__ void inner_s_chunk(const ReqDesc &D, const ReqState &S, int br0)
{
  float alpha[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) alpha[k] = (float) S.k[k].alpha;
  const gptr<const float> v = glb((const float *) D.v[0]);
  const gptr<const float> ri = glb((const float *) (S.inner_it == 0 ? D.r32 : D.ri));
  const gptr<float> s_out = glb(D.s);
  chunk_rows<BS, NR>(D, br0, [&](uint32_t o) {
    float rr[NR], vv[NR], sn[NR];
    ldv<float, NR>(ri, o, rr);
    ldv<float, NR>(v, o, vv);
#pragma unroll
    for (int k = 0; k < NR; ++k) sn[k] = s_update(rr[k], vv[k], alpha[k]);
    stv<float, NR>(s_out, o, sn);
  });
}

template <int NR> struct SOwn {
  float s[NR];
};

// t = A32 z, z = s or D^-1 s with Jacobi; partials (t, s), (t, t).
template <int BS, int NR>
__device__ __forceinline__ void inner_b_chunk(const ReqDesc &D, const ReqState &S, int br0, Acc &acc)
{
  (void) S;
  const gptr<const float> s_in = glb((const float *) D.s);
  const gptr<float> t_out = glb(D.t);
  const gptr<const float> dinv = glb(D.dinv);
  const bool jac = D.dinv != nullptr;
  auto gather = [&](uint32_t o, float (&x)[NR]) {
    ldv<float, NR>(s_in, o, x);
    if (jac) {
      const float m = ld(dinv, o / NR);
#pragma unroll
      for (int k = 0; k < NR; ++k) x[k] *= m;
    }
  };
  auto own = [&](uint32_t o, SOwn<NR> &m) { ldv<float, NR>(s_in, o, m.s); };
  auto epi = [&](uint32_t o, const float (&y)[NR], const SOwn<NR> &m) {
    stv<float, NR>(t_out, o, y);
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      acc[k][0] += (double) y[k] * (double) m.s[k];
      acc[k][1] += (double) y[k] * (double) y[k];
    }
  };
  if constexpr (BS >= SERVE_F16_MIN_BS)
    if (D.v16) {
      if (D.sl_j) chunk_spmv_sl<__half, float, BS, NR, SOwn<NR>>(D, glb(D.v16), br0, true, gather, own, epi, glb(D.sc16));
      else chunk_spmv<__half, float, BS, NR, SOwn<NR>, 1>(D, glb(D.v16), br0, true, gather, own, epi, glb(D.sc16));
      return;
    }
  if (D.sl_j) chunk_spmv_sl<float, float, BS, NR, SOwn<NR>>(D, glb(D.v32), br0, true, gather, own, epi);
  else chunk_spmv<float, float, BS, NR, SOwn<NR>, 1>(D, glb(D.v32), br0, true, gather, own, epi);
}

// ri = s - omega t, d += alpha p + omega s, or d += D^-1 (alpha p + omega s)
// with Jacobi (p = r32 and d = 0 before at iteration 0; a live right-hand
// side outside the inner solve gets d = 0 at iteration 0 and keeps d
// afterwards); partials (r32, ri), (ri, ri).
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
  const gptr<const float> rh = glb((const float *) D.r32);
  const gptr<const float> pcur = it == 0 ? rh : glb((const float *) D.p[0]), s_in = glb((const float *) D.s),
                          t_in = glb((const float *) D.t);
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
    if constexpr (SLOT == SLOT_X) {
      resid_x_chunk<BS, NR>(D, S, br0);
    } else if constexpr (SLOT == SLOT_R) {
      resid_chunk<BS, NR>(D, S, br0, acc);
    } else if constexpr (SLOT == SLOT_P) {
      inner_p_chunk<BS, NR>(D, S, br0);
    } else if constexpr (SLOT == SLOT_A) {
      inner_a_chunk<BS, NR>(D, S, br0, acc);
    } else if constexpr (SLOT == SLOT_S) {
      inner_s_chunk<BS, NR>(D, S, br0);
    } else if constexpr (SLOT == SLOT_B) {
      inner_b_chunk<BS, NR>(D, S, br0, acc);
    } else {
      inner_c_chunk<BS, NR>(D, S, br0, acc);
    }
    if constexpr (!slot_reduces(SLOT)) continue;
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
// The SpMV slots (R, A, B) run as one kernel per block size (BSK > 0) over
// that block size's range of the tile table: a kernel holding every block
// size's paths needs the registers of their union (about 1.4x the largest
// one), which would hold all of them to 2-3 waves per SIMD.  The elementwise
// slots (BSK = 0) walk the whole table.
template <int SLOT> constexpr bool slot_per_bs() { return SLOT == SLOT_R || SLOT == SLOT_A || SLOT == SLOT_B; }

template <int SLOT, int BSK>
__device__ __forceinline__ void slot_tile(const ReqDesc *__restrict__ rd, const ReqState *__restrict__ rs,
                                          const TileDesc *__restrict__ tiles, double *__restrict__ part, int ti)
{
  const TileDesc T = tiles[ti];
  const ReqState &S = rs[T.req];
  const int phase = S.phase;
  if (phase != slot_phase(SLOT)) return;
  const ReqDesc &D = rd[T.req];
  if constexpr (BSK > 0) {
    constexpr int BS = BSK;
    BMFP_SERVE_NR_SWITCH(D.nrhs, (run_tile<SLOT, BS, NR>(D, S, phase, T, glb(part))));
  } else {
    BMFP_SERVE_BS_SWITCH(D.bs, BMFP_SERVE_NR_SWITCH(D.nrhs, (run_tile<SLOT, BS, NR>(D, S, phase, T, glb(part)))));
  }
}

template <int SLOT, int BSK>
__device__ __forceinline__ void slot_body(const ReqDesc *__restrict__ rd, const ReqState *__restrict__ rs,
                                          const TileDesc *__restrict__ tiles, const int *__restrict__ counts,
                                          double *__restrict__ part, int xcd_run, int persistent)
{
  int t0 = 0, nt = counts[0];
  if constexpr (BSK > 0) {
    t0 = counts[2 + 2 * serve_bs_index(BSK)];
    nt = counts[3 + 2 * serve_bs_index(BSK)];
  }
  const int wib = (int) (threadIdx.x >> 6);
  if (!persistent) {
    int64_t ti = (int64_t) blockIdx.x * 4 + wib;
    if (xcd_run > 0) {
      const int64_t l = (int64_t) (blockIdx.x >> 3) * 4 + wib;
      ti = (l / xcd_run * 8 + (blockIdx.x & 7)) * xcd_run + l % xcd_run;
    }
    if (ti < nt) slot_tile<SLOT, BSK>(rd, rs, tiles, part, __builtin_amdgcn_readfirstlane(t0 + (int) ti));
    return;
  }
  if (xcd_run > 0) {
    const int xcd = (int) blockIdx.x & 7;
    const int lw = __builtin_amdgcn_readfirstlane((int) (blockIdx.x >> 3) * 4 + wib);
    const int nwx = (int) (gridDim.x >> 3) * 4;
    for (int j = lw;; j += nwx) {
      const int64_t ti64 = ((int64_t) (j / xcd_run) * 8 + xcd) * xcd_run + j % xcd_run;
      if (ti64 >= nt) break;
      slot_tile<SLOT, BSK>(rd, rs, tiles, part, t0 + (int) ti64);
    }
    return;
  }
  const int wave = __builtin_amdgcn_readfirstlane((int) blockIdx.x * 4 + wib);
  const int nw = (int) gridDim.x * 4;
  for (int ti = wave; ti < nt; ti += nw) slot_tile<SLOT, BSK>(rd, rs, tiles, part, t0 + ti);
}

template <int SLOT, int BSK = 0>
__global__ __launch_bounds__(256) void k_slot(const ReqDesc *__restrict__ rd, const ReqState *__restrict__ rs,
                                              const TileDesc *__restrict__ tiles, const int *__restrict__ counts,
                                              double *__restrict__ part, int xcd_run, int persistent)
{
  slot_body<SLOT, BSK>(rd, rs, tiles, counts, part, xcd_run, persistent);
}

// ---- finalize: one workgroup per request; after the reduction lane k of the
// first wavefront owns right-hand side k.  SLOT here is the phase index (0
// RESID, 1 INNER_A, 2 INNER_B, 3 INNER_C).
template <int SLOT>
__global__ __launch_bounds__(256) void k_fin(const ReqDesc *__restrict__ rd, ReqState *__restrict__ rs,
                                             const int *__restrict__ list, const int *__restrict__ nlp,
                                             const double *__restrict__ part)
{
  __shared__ double red[4][SERVE_PART];
  const int w = (int) blockIdx.x, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (w >= *nlp) return;
  const int q = list[w];
  ReqState &S = rs[q];
  const int phase = S.phase;
  if (!fin_owns<SLOT>(phase)) return;
  const ReqDesc &D = rd[q];
  // the request's chunk partials: thread t adds chunks t, t + 256, ... (8
  // doubles each, 16-byte loads), then the waves' sums are added in order
  double a[SERVE_PART];
#pragma unroll
  for (int j = 0; j < SERVE_PART; ++j) a[j] = 0.0;
  for (int c = threadIdx.x; c < D.n_chunks; c += 256) {
    const double2 *pc = reinterpret_cast<const double2 *>(part + (int64_t) (D.chunk0 + c) * SERVE_PART);
#pragma unroll
    for (int j = 0; j < SERVE_PART / 2; ++j) {
      const double2 t = pc[j];
      a[2 * j] += t.x;
      a[2 * j + 1] += t.y;
    }
  }
#pragma unroll
  for (int j = 0; j < SERVE_PART; ++j) a[j] = wave_sum(a[j]);
  if (lane == 0)
#pragma unroll
    for (int j = 0; j < SERVE_PART; ++j) red[wv][j] = a[j];
  __syncthreads();
  if (wv != 0) return;
#pragma unroll
  for (int j = 0; j < SERVE_PART; ++j) a[j] = (red[0][j] + red[1][j]) + (red[2][j] + red[3][j]);
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
          R.inner_tol = fmax(D.inner_tol, rel > 0.0 ? SERVE_LAST_STEP_MARGIN * D.tol / rel : 0.0);
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
        const bool stop = !(rel_in >= R.inner_tol) || it + 1 >= D.max_inner || R.tt == 0.0 || R.omega == 0.0 ||
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

// FP16 inner matrices scale each block row by one FP32 factor: sc = the row's
// max |entry| rounded to FP32 (1 for an all-zero row), values stored as
// FP16(FP32(v / sc)) and multiplied by sc after the row's sums.
//
// The scale of every block row: one wavefront per row, one read of its
// values (up to 9 per lane held in registers, longer rows in further passes).
__global__ __launch_bounds__(256) void k_serve_rowmax(int n_brows, int bs, const int *__restrict__ rp,
                                                      const double *__restrict__ v64, float *__restrict__ sc)
{
  const int64_t br = ((int64_t) blockIdx.x * 256 + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (br >= n_brows) return;
  const int64_t bs2 = (int64_t) bs * bs, e0 = (int64_t) rp[br] * bs2, e1 = (int64_t) rp[br + 1] * bs2;
  double m = 0.0;
  for (int64_t e = e0; e < e1; e += 9 * 64) {
    double v[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) v[k] = e + k * 64 + lane < e1 ? v64[e + k * 64 + lane] : 0.0;
#pragma unroll
    for (int k = 0; k < 9; ++k) m = fmax(m, fabs(v[k]));
  }
  for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if (lane == 0) sc[br] = m == 0.0 ? 1.0f : (float) m;
}

template <class OT> __device__ __forceinline__ OT to_inner(double v, double inv);
template <> __device__ __forceinline__ float to_inner<float>(double v, double) { return (float) v; }
template <> __device__ __forceinline__ __half to_inner<__half>(double v, double inv)
{
  return __float2half_rn((float) (v * inv));
}

// FP16 inner matrix in BCSR order (the fallback for matrices whose rows are
// too uneven to slice); sc from k_serve_rowmax.  One thread per value.
__global__ __launch_bounds__(256) void k_serve_f16(int n_brows, int bs, const int *__restrict__ rp,
                                                   const double *__restrict__ v64, const float *__restrict__ sc,
                                                   __half *__restrict__ v16)
{
  const int64_t bs2 = (int64_t) bs * bs, nv = (int64_t) rp[n_brows] * bs2;
  for (int64_t e = (int64_t) blockIdx.x * 256 + threadIdx.x; e < nv; e += (int64_t) gridDim.x * 256) {
    // the block row of value e: binary search of the row pointers
    const int b = (int) (e / bs2);
    int lo = 0, hi = n_brows;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (rp[mid] <= b) lo = mid;
      else hi = mid;
    }
    v16[e] = to_inner<__half>(v64[e], 1.0 / (double) sc[lo]);
  }
}

// Pass-sliced inner copy of a request's FP64 BCSR matrix (layout: see
// chunk_spmv_sl), FP32, or FP16 with the scales in sc.  One thread per row of
// a sliced block (grid-stride; 32-bit index arithmetic, the sliced values
// staying below 2^31): consecutive threads store consecutive rows, each row
// read from its block in the FP64 copy, zero in padding blocks; the thread of
// row 0 also writes the block's column (padding: the row's last block column,
// 0 for an empty row).
template <class OT, int BS>
__global__ __launch_bounds__(256) void k_serve_slice(int n_brows, int sl_j, const int *__restrict__ rp,
                                                     const int *__restrict__ cols, const double *__restrict__ v64,
                                                     const float *__restrict__ sc, OT *__restrict__ out,
                                                     int *__restrict__ cols_sl)
{
  constexpr int RPW = 64 / BS, BS2 = BS * BS;
  const int n_rows = ((n_brows + RPW - 1) / RPW * RPW * sl_j) * BS;  // rows of all sliced blocks
  for (int t = blockIdx.x * 256 + threadIdx.x; t < n_rows; t += gridDim.x * 256) {
    // sliced block sb = (P*J + j)*RPW + lbr holds block j of block row P*RPW + lbr
    const int sb = t / BS, r = t - sb * BS;
    const int pj = sb / RPW, lbr = sb - pj * RPW, pp = pj / sl_j, j = pj - pp * sl_j, br = pp * RPW + lbr;
    int b0 = 0, len = 0;
    if (br < n_brows) {
      b0 = rp[br];
      len = rp[br + 1] - b0;
    }
    if (r == 0) cols_sl[sb] = j < len ? cols[b0 + j] : len > 0 ? cols[b0 + len - 1] : 0;
    OT o[BS];
    if (j < len) {
      const double inv = sizeof(OT) == 2 ? 1.0 / (double) sc[br] : 1.0;
      double v[BS];
      load_n<double, BS, lowbit(8 * BS)>(v64 + ((int64_t) (b0 + j) * BS2 + r * BS), v);
#pragma unroll
      for (int c = 0; c < BS; ++c) o[c] = to_inner<OT>(v[c], inv);
    } else {
#pragma unroll
      for (int c = 0; c < BS; ++c) o[c] = (OT) 0.0f;
    }
    store_words<OT, BS>(out + (int64_t) t * BS, o);
  }
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

// The same from a pass-sliced FP64 copy: the row's slots in order, the first
// whose column is the row's own (padding slots follow the real blocks).
__global__ __launch_bounds__(256) void k_serve_dinv_sl(int n_brows, int bs, int sl_j, const int *__restrict__ cols_sl,
                                                       const double *__restrict__ v64, float *__restrict__ dinv)
{
  const int rpw = 64 / bs;
  const int64_t n = (int64_t) n_brows * bs;
  for (int64_t row = (int64_t) blockIdx.x * 256 + threadIdx.x; row < n; row += (int64_t) gridDim.x * 256) {
    const int br = (int) (row / bs), r = (int) (row - (int64_t) br * bs);
    const int64_t s0 = (int64_t) (br / rpw) * sl_j * rpw + br % rpw;
    double a = 0.0;
    bool found = false;
    for (int j = 0; j < sl_j; ++j) {
      const int64_t sb = s0 + (int64_t) j * rpw;
      if (cols_sl[sb] == br) {
        a = v64[(sb * bs + r) * bs + r];
        found = true;
        break;
      }
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
