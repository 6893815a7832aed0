// solver_kernels.hpp -- fused BiCGStab kernels for gfx950.
//
// One BiCGStab iteration (bicgstab.cu:426-513 without per-step host syncs) is
// three kernels:
//   A  v = A q,  q = P (r + beta (p_prev - omega_prev v_prev))   (p formed on
//      the fly at every gathered index and written for owned rows), partial
//      sums of (r^, v)
//   B  t = A z,  z = P (r - alpha v)  (s written for owned rows), partial sums
//      of (t, s) and (t, t)
//   C  r = s - omega t,  x = (x + alpha P p) + omega P s, partial sums of
//      (r^, r) and (r, r)
// P is the point-Jacobi inverse diagonal (bicgstab.cu:201-238) or identity.
//
// Scalars never leave the device.  Every kernel runs a persistent grid of the
// same NP workgroups; each workgroup first reduces the previous kernel's NP
// partial sums itself, in one fixed order, so all workgroups derive
// bit-identical alpha/beta/omega without atomics or grid barriers.  The
// iteration counter and the stop flag travel through SolverState fields that
// only earlier kernels write.
#pragma once

#include <type_traits>

#include "spmv_kernels.hpp"

namespace bmfp {

constexpr int SOLVER_MAX_GROUPS = 2048;

struct SolverState {
  double rho[2], alpha[2], omega[2], tt[2];  // by iteration parity
  double norm_b2;      // ||b||^2
  double rr;           // ||r||^2 after the last completed iteration
  double eps;          // relative tolerance on ||r||/||b||
  int iter;            // completed iterations (written by C)
  int iter_b, iter_c;  // hand-off copies (written by A, B)
  int done;            // 1 once the solve has stopped
  int status;          // BMFP_OK / _NOT_CONVERGED / _BREAKDOWN / _NONFINITE
  int max_iter;
  int pad;
};

template <class AT>
struct SolverVecs {
  const AT *__restrict__ b;
  AT *__restrict__ x;
  AT *__restrict__ r;
  AT *__restrict__ rh;
  AT *__restrict__ p[2];
  AT *__restrict__ v[2];
  AT *__restrict__ s;
  AT *__restrict__ t;
  const AT *__restrict__ pinv;   // Jacobi inverse diagonal, or null
  double *__restrict__ partA;    // [npa]
  double *__restrict__ partB;    // [npb][2]
  double *__restrict__ partC;    // [NP][2]
  double *__restrict__ partW;    // split schedule: per-workgroup partials of its SpMVs
  SolverState *st;
  int64_t n;                     // scalar rows
  int npa, npb;                  // partial sums in partA / partB (NP fused, 1 split)
};

// ---- fixed-order reductions ----------------------------------------------------
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Sum of one value per thread over a 256-thread workgroup; result valid in
// every thread.  Order fixed by thread index.
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double (*lds)[4])
{
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    v[k] = wave_sum(v[k]);
    if (lane == 0) lds[k][w] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = (lds[k][0] + lds[k][1]) + (lds[k][2] + lds[k][3]);
  __syncthreads();
}

// Reduces K interleaved partial sums part[g*K + k], g < np, in a fixed order.
template <int K>
__device__ __forceinline__ void reduce_partials(const double *__restrict__ part, int np, double (&out)[K],
                                                double (*lds)[4])
{
#pragma unroll
  for (int k = 0; k < K; ++k) out[k] = 0.0;
  for (int g = threadIdx.x; g < np; g += 256)
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] += part[g * K + k];
  block_sum<K>(out, lds);
}

template <int K>
__device__ __forceinline__ void write_partials(double *__restrict__ part, double (&v)[K], double (*lds)[4])
{
  block_sum<K>(v, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) part[blockIdx.x * K + k] = v[k];
}

// Reads the stop flag once per workgroup: WG 0 of the running kernel may set
// it concurrently, and every thread of a workgroup must take the same branch.
__device__ __forceinline__ bool solver_stopped(const SolverState *st)
{
  __shared__ int s_done;
  if (threadIdx.x == 0) s_done = st->done;
  __syncthreads();
  return s_done != 0;
}

// ---- operands formed on the fly ---------------------------------------------------
// The same explicit fused multiply-adds are used where an owned row's value is
// stored and where neighbours gather it, so both see identical bits.
template <class AT> __device__ __forceinline__ AT p_update(AT r, AT p, AT v, AT beta, AT omega)
{
  return fmadd(beta, fmadd(-omega, v, p), r);
}
template <class AT> __device__ __forceinline__ AT s_update(AT r, AT v, AT alpha) { return fmadd(-alpha, v, r); }

// Kernel A operand: q_i = P_i (r_i + beta (p_i - omega v_i)).  `raw` returns
// the unpreconditioned p_i that is stored for owned rows.
template <class AT>
struct OperandP {
  const AT *__restrict__ r;
  const AT *__restrict__ p;
  const AT *__restrict__ v;
  const AT *__restrict__ pinv;
  AT beta, omega;
  __device__ __forceinline__ AT raw(int64_t i) const { return p_update(r[i], p[i], v[i], beta, omega); }
  __device__ __forceinline__ AT operator()(int64_t i) const
  {
    const AT q = raw(i);
    return pinv ? pinv[i] * q : q;
  }
  template <int N, int ALIGN>
  __device__ __forceinline__ void load(int64_t i, AT (&out)[N]) const
  {
    AT rr[N], pp[N], vv[N];
    load_n<AT, N, ALIGN>(r + i, rr);
    load_n<AT, N, ALIGN>(p + i, pp);
    load_n<AT, N, ALIGN>(v + i, vv);
#pragma unroll
    for (int c = 0; c < N; ++c) out[c] = p_update(rr[c], pp[c], vv[c], beta, omega);
    if (pinv) {
      AT d[N];
      load_n<AT, N, ALIGN>(pinv + i, d);
#pragma unroll
      for (int c = 0; c < N; ++c) out[c] = d[c] * out[c];
    }
  }
};

// Kernel B operand: z_i = P_i (r_i - alpha v_i); raw = s_i.
template <class AT>
struct OperandS {
  const AT *__restrict__ r;
  const AT *__restrict__ v;
  const AT *__restrict__ pinv;
  AT alpha;
  __device__ __forceinline__ AT raw(int64_t i) const { return s_update(r[i], v[i], alpha); }
  __device__ __forceinline__ AT operator()(int64_t i) const
  {
    const AT q = raw(i);
    return pinv ? pinv[i] * q : q;
  }
  template <int N, int ALIGN>
  __device__ __forceinline__ void load(int64_t i, AT (&out)[N]) const
  {
    AT rr[N], vv[N];
    load_n<AT, N, ALIGN>(r + i, rr);
    load_n<AT, N, ALIGN>(v + i, vv);
#pragma unroll
    for (int c = 0; c < N; ++c) out[c] = s_update(rr[c], vv[c], alpha);
    if (pinv) {
      AT d[N];
      load_n<AT, N, ALIGN>(pinv + i, d);
#pragma unroll
      for (int c = 0; c < N; ++c) out[c] = d[c] * out[c];
    }
  }
};

// Plain vector operand (split schedule): the SpMV reads x as stored.
template <class AT>
struct OperandVec {
  const AT *__restrict__ x;
  __device__ __forceinline__ AT raw(int64_t i) const { return x[i]; }
  __device__ __forceinline__ AT operator()(int64_t i) const { return x[i]; }
  template <int N, int ALIGN>
  __device__ __forceinline__ void load(int64_t i, AT (&out)[N]) const
  {
    load_n<AT, N, ALIGN>(x + i, out);
  }
};

// ---- fused SpMV body -------------------------------------------------------------------
// Lane group of G lanes per scalar row (G = 1: thread per row); lane g owns
// columns [g*VEC, (g+1)*VEC) of every block.  Returns, for lane 0 of the
// group, y_row = sum_b A_(row,b) op(x_b) and the owned operand op.raw(row).
template <class VT, class AT, int BS, int G, int VEC, int LAYOUT, bool SCALED, class OP>
__device__ __forceinline__ AT fused_row(const int *__restrict__ rp, const int *__restrict__ cols,
                                        const VT *__restrict__ vals, const float *__restrict__ scale,
                                        const OP &op, int br, int r, int g)
{
  constexpr int GA = BS / VEC;
  constexpr int CC = VEC % 8 == 0 ? 8 : VEC;  // columns per step (bounds registers)
  constexpr int VALIGN =
      cmin(cmin(lowbit((int) sizeof(VT) * BS), lowbit((int) sizeof(VT) * VEC)), lowbit((int) sizeof(VT) * CC));
  constexpr int XALIGN =
      cmin(cmin(lowbit((int) sizeof(AT) * BS), lowbit((int) sizeof(AT) * VEC)), lowbit((int) sizeof(AT) * CC));
  AT acc = 0;
  if (g < GA) {
    const int b0 = rp[br], b1 = rp[br + 1];
#pragma unroll 2
    for (int b = b0; b < b1; ++b) {
      const int64_t x0 = (int64_t) cols[b] * BS + g * VEC;
      const VT *blk = vals + (int64_t) b * (BS * BS);
#pragma unroll
      for (int c0 = 0; c0 < VEC; c0 += CC) {
        AT xv[CC];
        op.template load<CC, XALIGN>(x0 + c0, xv);
        if constexpr (LAYOUT == BMFP_ROW_MAJOR) {
          VT v[CC];
          load_n<VT, CC, VALIGN>(blk + r * BS + g * VEC + c0, v);
#pragma unroll
          for (int c = 0; c < CC; ++c) acc = fmadd(cvt<AT>(v[c]), xv[c], acc);
        } else {
#pragma unroll
          for (int c = 0; c < CC; ++c) acc = fmadd(cvt<AT>(blk[(g * VEC + c0 + c) * BS + r]), xv[c], acc);
        }
      }
    }
  }
#pragma unroll
  for (int off = G / 2; off >= 1; off /= 2) acc += __shfl_xor(acc, off, G);
  if constexpr (SCALED) acc *= (AT) scale[br];
  return acc;
}

// ---- kernels ---------------------------------------------------------------------------
template <class VT, class AT>
struct MatArgs {
  int n_brows;
  const int *__restrict__ rp;
  const int *__restrict__ cols;
  const VT *__restrict__ vals;
  const float *__restrict__ scale;
  const int *__restrict__ cols_sl = nullptr;  // pass-sliced matrices: block columns
  int sl_j = 0;                                // pass-sliced matrices: block slots per row
};

// Row engines: walk every scalar row of the matrix in a persistent grid and
// call epi(row, y) once per row with y = (A op)_row.
//
// GroupEngine: G lanes per row, lane g owning a column slice of every block
// (row-major blocks with bs >= 12: whole 128-byte lines per block row).
template <class VT, class AT, int BS, int G, int VEC, int LAYOUT, bool SCALED>
struct GroupEngine {
  static int64_t threads(int64_t n_brows) { return n_brows * BS * G; }
  template <class OP, class EPI>
  __device__ __forceinline__ static void run(const MatArgs<VT, AT> &M, const OP &op, EPI &&epi)
  {
    const int64_t threads = (int64_t) M.n_brows * BS * G;
    for (int64_t base = (int64_t) blockIdx.x * 256; base < threads; base += (int64_t) gridDim.x * 256) {
      const int64_t tt = base + threadIdx.x;
      if (tt < threads) {
        const int g = (int) (tt % G);
        const int64_t row = tt / G;
        const int br = (int) (row / BS), r = (int) (row % BS);
        const AT y = fused_row<VT, AT, BS, G, VEC, LAYOUT, SCALED>(M.rp, M.cols, M.vals, M.scale, op, br, r, g);
        if (g == 0) epi(row, y);
      }
    }
  }
};

// Lane c of each group of BS consecutive lanes, broadcast to the group: no
// exchange for BS = 1, DPP quad permutations (VALU) for BS = 2 and 4, a
// bit-mode swizzle for BS = 8, a general permute otherwise.
template <int BS> __device__ __forceinline__ int group_bcast_i32(int v, int c, int base_lane)
{
  if constexpr (BS == 1) {
    return v;
  } else if constexpr (BS == 2) {
    return c == 0 ? __builtin_amdgcn_mov_dpp(v, 0xA0, 0xF, 0xF, false) : __builtin_amdgcn_mov_dpp(v, 0xF5, 0xF, 0xF, false);
  } else if constexpr (BS == 4) {
    switch (c) {
    case 0: return __builtin_amdgcn_mov_dpp(v, 0x00, 0xF, 0xF, false);
    case 1: return __builtin_amdgcn_mov_dpp(v, 0x55, 0xF, 0xF, false);
    case 2: return __builtin_amdgcn_mov_dpp(v, 0xAA, 0xF, 0xF, false);
    default: return __builtin_amdgcn_mov_dpp(v, 0xFF, 0xF, 0xF, false);
    }
  } else if constexpr (BS == 8) {
    switch (c) {  // lane (i & 0x18) | c within each 32-lane half
    case 0: return __builtin_amdgcn_ds_swizzle(v, 0x18 | (0 << 5));
    case 1: return __builtin_amdgcn_ds_swizzle(v, 0x18 | (1 << 5));
    case 2: return __builtin_amdgcn_ds_swizzle(v, 0x18 | (2 << 5));
    case 3: return __builtin_amdgcn_ds_swizzle(v, 0x18 | (3 << 5));
    case 4: return __builtin_amdgcn_ds_swizzle(v, 0x18 | (4 << 5));
    case 5: return __builtin_amdgcn_ds_swizzle(v, 0x18 | (5 << 5));
    case 6: return __builtin_amdgcn_ds_swizzle(v, 0x18 | (6 << 5));
    default: return __builtin_amdgcn_ds_swizzle(v, 0x18 | (7 << 5));
    }
  } else {
    return __shfl(v, base_lane + c);
  }
}

template <int BS> __device__ __forceinline__ float group_bcast(float v, int c, int base_lane)
{
  return __int_as_float(group_bcast_i32<BS>(__float_as_int(v), c, base_lane));
}
template <int BS> __device__ __forceinline__ double group_bcast(double v, int c, int base_lane)
{
  return __hiloint2double(group_bcast_i32<BS>(__double2hiint(v), c, base_lane),
                          group_bcast_i32<BS>(__double2loint(v), c, base_lane));
}

// ShflEngine: one lane per scalar row, 64/BS block rows per wavefront.  For
// each block, lane r forms only operand element r of the block's x segment
// (the operands here cost 2-4 loads each) and reads the other BS-1 from its
// block-row neighbours by group broadcasts; loads of U blocks are issued
// together.
template <class VT, class AT, int BS, int LAYOUT, bool SCALED>
struct ShflEngine {
  static constexpr int U = BS <= 2 ? 9 : BS <= 8 ? 3 : 1;
  static int64_t threads(int64_t n_brows) { return (n_brows + 64 / BS - 1) / (64 / BS) * 64; }
  template <class OP, class EPI>
  __device__ __forceinline__ static void run(const MatArgs<VT, AT> &M, const OP &op, EPI &&epi)
  {
    constexpr int RPW = 64 / BS;
    const int lane = threadIdx.x & 63;
    const int lbr = lane / BS, r = lane - lbr * BS, base_lane = lbr * BS;
    const int64_t n_tiles = ((int64_t) M.n_brows + RPW - 1) / RPW;
    const int64_t waves = (int64_t) gridDim.x * 4;
    // a tile's row pointers are loaded while the wavefront's previous tile
    // computes: one dependent global latency less per tile
    auto row_bounds = [&](int64_t t, int &lo, int &hi) {
      const int64_t row = t * RPW + lbr;
      lo = hi = 0;
      if (t < n_tiles && lane < RPW * BS && row < M.n_brows) {
        lo = M.rp[row];
        hi = M.rp[row + 1];
      }
    };
    int64_t wt = ((int64_t) blockIdx.x * 256 + threadIdx.x) >> 6;
    int next_b, next_b1;
    row_bounds(wt, next_b, next_b1);
    for (; wt < n_tiles; wt += waves) {
      const int64_t br64 = wt * RPW + lbr;
      const bool active = lane < RPW * BS && br64 < M.n_brows;
      const int br = (int) br64;
      int b = next_b, b1 = next_b1;
      row_bounds(wt + waves, next_b, next_b1);
      // all lanes of a block row share b..b1, so every shuffle source below
      // is active in the same iteration
      AT acc = 0;
      for (; b + U <= b1; b += U) {
        int col[U];
#pragma unroll
        for (int k = 0; k < U; ++k) col[k] = M.cols[b + k];
        VT v[U][BS];
#pragma unroll
        for (int k = 0; k < U; ++k) load_block_row<VT, BS, LAYOUT>(M.vals + (int64_t) (b + k) * (BS * BS), r, v[k]);
        AT xe[U];
#pragma unroll
        for (int k = 0; k < U; ++k) xe[k] = op((int64_t) col[k] * BS + r);
#pragma unroll
        for (int k = 0; k < U; ++k)
#pragma unroll
          for (int c = 0; c < BS; ++c) acc = fmadd(cvt<AT>(v[k][c]), group_bcast<BS>(xe[k], c, base_lane), acc);
      }
      for (; b < b1; ++b) {
        VT v[BS];
        load_block_row<VT, BS, LAYOUT>(M.vals + (int64_t) b * (BS * BS), r, v);
        const AT xe = op((int64_t) M.cols[b] * BS + r);
#pragma unroll
        for (int c = 0; c < BS; ++c) acc = fmadd(cvt<AT>(v[c]), group_bcast<BS>(xe, c, base_lane), acc);
      }
      if (active) {
        if constexpr (SCALED) acc *= (AT) M.scale[br];
        epi((int64_t) br * BS + r, acc);
      }
    }
  }
};

// ShflEngine with each wavefront's blocks staged in LDS first.  The values
// and column indices of a tile's RPW block rows are one contiguous range,
// copied with 16-byte loads at consecutive addresses across the lanes
// (StagedRange); the per-lane loop is ShflEngine's, reading blocks and
// indices from LDS, so every product and sum is the same, in the same order
// (results bitwise equal to ShflEngine's).  Used for FP32 blocks of up to 4
// columns, whose lane-per-row loads are narrow (16 bytes per block row): the
// FP32 refinement ran 2-3% faster.  FP16 blocks gained nothing (their
// kernels wait on the operand gathers) and lost occupancy, so they keep
// ShflEngine.  A tile over CAP blocks (rows of more than 9 blocks on average)
// reads global memory as ShflEngine does.
template <class VT, class AT, int BS, int LAYOUT, bool SCALED>
struct StagedShflEngine {
  using Base = ShflEngine<VT, AT, BS, LAYOUT, SCALED>;
  // ShflEngine's batch: whole 9-block rows (one gather latency per row) need
  // 156 VGPRs in FP32 and ran 1.3x slower
  static constexpr int U = Base::U, RPW = 64 / BS, CAP = 9 * RPW;
  static constexpr int BB = BS * BS * (int) sizeof(VT);  // bytes per block
  static_assert(BB % 4 == 0, "staging copies whole 4-byte words");
  static int64_t threads(int64_t n_brows) { return Base::threads(n_brows); }

  template <class OP, class COLF, class BLKF>
  __device__ __forceinline__ static AT block_row(int b, int b1, int r, int base_lane, const OP &op, COLF col_at,
                                                 BLKF blk_at)
  {
    AT acc = 0;
    for (; b + U <= b1; b += U) {
      int col[U];
#pragma unroll
      for (int k = 0; k < U; ++k) col[k] = col_at(b + k);
      VT v[U][BS];
#pragma unroll
      for (int k = 0; k < U; ++k) load_block_row<VT, BS, LAYOUT>(blk_at(b + k), r, v[k]);
      AT xe[U];
#pragma unroll
      for (int k = 0; k < U; ++k) xe[k] = op((int64_t) col[k] * BS + r);
#pragma unroll
      for (int k = 0; k < U; ++k)
#pragma unroll
        for (int c = 0; c < BS; ++c) acc = fmadd(cvt<AT>(v[k][c]), group_bcast<BS>(xe[k], c, base_lane), acc);
    }
    for (; b < b1; ++b) {
      VT v[BS];
      load_block_row<VT, BS, LAYOUT>(blk_at(b), r, v);
      const AT xe = op((int64_t) col_at(b) * BS + r);
#pragma unroll
      for (int c = 0; c < BS; ++c) acc = fmadd(cvt<AT>(v[c]), group_bcast<BS>(xe, c, base_lane), acc);
    }
    return acc;
  }

  template <class OP, class EPI>
  __device__ __forceinline__ static void run(const MatArgs<VT, AT> &M, const OP &op, EPI &&epi)
  {
    __shared__ __align__(16) char vals_s[4][staged_bytes(CAP * BB)];
    __shared__ __align__(16) char cols_s[4][staged_bytes(CAP * 4)];
    const int lane = threadIdx.x & 63, wib = threadIdx.x >> 6;
    const int lbr = lane / BS, r = lane - lbr * BS, base_lane = lbr * BS;
    const int64_t n_tiles = ((int64_t) M.n_brows + RPW - 1) / RPW;
    const int64_t waves = (int64_t) gridDim.x * 4;
    for (int64_t wt = ((int64_t) blockIdx.x * 256 + threadIdx.x) >> 6; wt < n_tiles; wt += waves) {
      const int r0 = (int) (wt * RPW), r1 = r0 + RPW < M.n_brows ? r0 + RPW : M.n_brows;
      const int B0 = M.rp[r0], B1 = M.rp[r1];
      const int br = r0 + lbr;
      // all lanes of a block row share b..b1, so every shuffle source is
      // active in the same iteration
      const bool active = lane < RPW * BS && br < r1;
      int b = 0, b1 = 0;
      if (active) {
        b = M.rp[br];
        b1 = M.rp[br + 1];
      }
      AT acc;
      if (B1 - B0 <= CAP) {
        StagedRange<CAP * BB> sv;
        StagedRange<CAP * 4> sc;
        sv.load(M.vals, (int64_t) B0 * BB, (int64_t) B1 * BB, lane);
        sc.load(M.cols, (int64_t) B0 * 4, (int64_t) B1 * 4, lane);
        wave_lds_sync();  // the previous tile's reads of the slice come first
        sv.store(vals_s[wib], lane);
        sc.store(cols_s[wib], lane);
        wave_lds_sync();
        const VT *vs = reinterpret_cast<const VT *>(vals_s[wib] + sv.shift);
        const int *cs = reinterpret_cast<const int *>(cols_s[wib] + sc.shift);
        acc = block_row(
            b, b1, r, base_lane, op, [&](int k) { return cs[k - B0]; },
            [&](int k) { return vs + (int64_t) (k - B0) * (BS * BS); });
      } else {
        acc = block_row(
            b, b1, r, base_lane, op, [&](int k) { return M.cols[k]; },
            [&](int k) { return M.vals + (int64_t) k * (BS * BS); });
      }
      if (active) {
        if constexpr (SCALED) acc *= (AT) M.scale[br];
        epi((int64_t) br * BS + r, acc);
      }
    }
  }
};

// DenseEngine: the lane mapping of k_bcsr_dgr (row-major blocks).  G =
// BS/VEC lanes per scalar row, lane g owning columns [g*VEC, (g+1)*VEC) of
// every block, U blocks' loads issued together, the group's partial sums
// combined by the same fixed shuffle tree.  The split schedule launches it
// like the standalone kernel, one wavefront per tile of 64/G rows (tiles()
// wavefronts, 4 per workgroup): a persistent loop around the same body held
// 102 instead of 68 VGPRs at bs 4 and ran the SpMV 1.4x slower.
template <class VT, class AT, int BS, int VEC, int U, bool SCALED>
struct DenseEngine {
  static constexpr int G = BS / VEC, RPW = 64 / G;
  static_assert(BS % VEC == 0 && G <= 64, "bad group shape");
  static int64_t threads(int64_t n_brows) { return (n_brows * BS + RPW - 1) / RPW * 64; }
  // workgroups of a one-tile-per-wavefront launch
  static int64_t groups1(int64_t n_brows) { return ((n_brows * BS + RPW - 1) / RPW + 3) / 4; }
  template <class OP, class EPI>
  __device__ __forceinline__ static void tile(const MatArgs<VT, AT> &M, const OP &op, EPI &&epi, int64_t wt,
                                              const int *stop = nullptr)
  {
    constexpr int ALIGN = cmin(lowbit((int) sizeof(VT) * BS), lowbit((int) sizeof(VT) * VEC));
    constexpr int XALIGN = cmin(lowbit((int) sizeof(AT) * BS), lowbit((int) sizeof(AT) * VEC));
    const int lane = threadIdx.x & 63;
    const int lr = lane / G, g = lane - lr * G;
    const int64_t rows = (int64_t) M.n_brows * BS;
    const int64_t row = wt * RPW + lr;
    const bool on = lane < RPW * G && row < rows;
    const int64_t br = on ? row / BS : 0;
    const int r = (int) (row - br * BS);
    int b = 0, b1 = 0;
    AT sc = 1, pre = 0;  // the row's scale and epilogue operand, loaded with its bounds
    if (on) {
      b = M.rp[br];
      b1 = M.rp[br + 1];
      if constexpr (SCALED) sc = (AT) M.scale[br];
      if (g == 0) pre = epi.pre(row);
    }
    if (stop) {
      // a stopped solve's tile returns here; the flag is read by a vector
      // load (relaxed, device scope: no scalar-cache wait in front of the
      // bounds) issued beside the bounds, and an empty asm that reads all
      // three keeps the compiler from sinking the loads past the test
      const int dn = __hip_atomic_load(stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      asm volatile("" ::"v"(b), "v"(b1), "v"(dn));
      if (__builtin_amdgcn_readfirstlane(dn)) return;
    }
    AT acc = 0;
    for (; b + U <= b1; b += U) {
      int col[U];
      VT v[U][VEC];
      AT xv[U][VEC];
#pragma unroll
      for (int k = 0; k < U; ++k) col[k] = M.cols[b + k];
#pragma unroll
      for (int k = 0; k < U; ++k) load_n<VT, VEC, ALIGN>(M.vals + ((int64_t) (b + k) * BS + r) * BS + g * VEC, v[k]);
#pragma unroll
      for (int k = 0; k < U; ++k) op.template load<VEC, XALIGN>((int64_t) col[k] * BS + g * VEC, xv[k]);
#pragma unroll
      for (int k = 0; k < U; ++k)
#pragma unroll
        for (int c = 0; c < VEC; ++c) acc = fmadd(cvt<AT>(v[k][c]), xv[k][c], acc);
    }
    for (; U > 1 && b < b1; ++b) {
      VT v[VEC];
      AT xv[VEC];
      load_n<VT, VEC, ALIGN>(M.vals + ((int64_t) b * BS + r) * BS + g * VEC, v);
      op.template load<VEC, XALIGN>((int64_t) M.cols[b] * BS + g * VEC, xv);
#pragma unroll
      for (int c = 0; c < VEC; ++c) acc = fmadd(cvt<AT>(v[c]), xv[c], acc);
    }
    acc = group_tree_sum<G>(acc, lane, lr, g);
    if (on && g == 0) {
      if constexpr (SCALED) acc *= sc;
      epi(row, acc, pre);
    }
  }
  // one tile per wavefront (grid of groups1 workgroups); false when *stop
  // was set (the tile did nothing)
  template <class OP, class EPI>
  __device__ __forceinline__ static bool run1(const MatArgs<VT, AT> &M, const OP &op, EPI &&epi,
                                              const int *stop = nullptr)
  {
    tile(M, op, epi, ((int64_t) blockIdx.x * 256 + threadIdx.x) >> 6, stop);
    return !(stop && __builtin_amdgcn_readfirstlane(*stop));
  }
  template <class OP, class EPI>
  __device__ __forceinline__ static void run(const MatArgs<VT, AT> &M, const OP &op, EPI &&epi)
  {
    const int64_t n_tiles = ((int64_t) M.n_brows * BS + RPW - 1) / RPW;
    const int64_t waves = (int64_t) gridDim.x * 4;
    for (int64_t wt = ((int64_t) blockIdx.x * 256 + threadIdx.x) >> 6; wt < n_tiles; wt += waves) tile(M, op, epi, wt);
  }
};

// Rows per thread for RowsEngine: as many consecutive rows of a row-major
// block as fill one 16-byte load (1 when a single row already does).
template <class VT, int BS> constexpr int rows_per_thread()
{
  int r = 1;
  while (BS % (2 * r) == 0 && 2 * r * BS * (int) sizeof(VT) <= 16) r *= 2;
  return r;
}

// RowsEngine: one thread per R consecutive scalar rows of a block row (row-
// major blocks), R = RR or rows_per_thread: its R block rows are one
// contiguous load (16 bytes for FP16 bs 4 and FP32 bs 2, where a single row
// is 8), the x segment one more; U blocks' loads issued together; no
// cross-lane sums.
template <class VT, class AT, int BS, int U, bool SCALED, int RR = 0>
struct RowsEngine {
  static constexpr int R = RR > 0 && BS % RR == 0 ? RR : rows_per_thread<VT, BS>();
  static constexpr int TPB = BS / R;  // threads per block row
  static int64_t threads(int64_t n_brows) { return n_brows * TPB; }
  static int64_t groups1(int64_t n_brows) { return (n_brows * TPB + 255) / 256; }
  template <class OP, class EPI>
  __device__ __forceinline__ static void item(const MatArgs<VT, AT> &M, const OP &op, EPI &&epi, int64_t t,
                                              const int *stop = nullptr)
  {
    constexpr int CC = BS % 8 == 0 ? 8 : BS;  // columns per step (bounds registers)
    constexpr int VALIGN = cmin(lowbit((int) sizeof(VT) * BS * R), 16);
    constexpr int XALIGN = cmin(lowbit((int) sizeof(AT) * CC), lowbit((int) sizeof(AT) * BS));
    const int br = (int) (t / TPB);
    const int r0 = (int) (t - (int64_t) br * TPB) * R;
    const int b0 = M.rp[br], b1 = M.rp[br + 1];
    // the scale and the epilogue operands load with the bounds, not after the row
    const AT sc = SCALED ? (AT) M.scale[br] : (AT) 1;
    AT pre[R], acc[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
      pre[i] = epi.pre((int64_t) br * BS + r0 + i);
      acc[i] = 0;
    }
    if (stop) {  // the stop test, as in DenseEngine::tile
      const int dn = __hip_atomic_load(stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      asm volatile("" ::"v"(b0), "v"(b1), "v"(dn));
      if (__builtin_amdgcn_readfirstlane(dn)) return;
    }
    int b = b0;
    if constexpr (U > 1) {
      for (; b + U <= b1; b += U) {
        int col[U];
        VT v[U][R * BS];
        AT xv[U][BS];
#pragma unroll
        for (int k = 0; k < U; ++k) col[k] = M.cols[b + k];
#pragma unroll
        for (int k = 0; k < U; ++k) load_n<VT, R * BS, VALIGN>(M.vals + ((int64_t) (b + k) * BS + r0) * BS, v[k]);
#pragma unroll
        for (int k = 0; k < U; ++k) op.template load<BS, XALIGN>((int64_t) col[k] * BS, xv[k]);
#pragma unroll
        for (int k = 0; k < U; ++k)
#pragma unroll
          for (int i = 0; i < R; ++i)
#pragma unroll
            for (int c = 0; c < BS; ++c) acc[i] = fmadd(cvt<AT>(v[k][i * BS + c]), xv[k][c], acc[i]);
      }
    }
#pragma unroll 3
    for (; b < b1; ++b) {
      const int64_t x0 = (int64_t) M.cols[b] * BS;
      const VT *blk = M.vals + ((int64_t) b * BS + r0) * BS;
#pragma unroll
      for (int c0 = 0; c0 < BS; c0 += CC) {
        AT xv[CC];
        op.template load<CC, XALIGN>(x0 + c0, xv);
#pragma unroll
        for (int i = 0; i < R; ++i) {
          VT v[CC];
          load_n<VT, CC, cmin(VALIGN, lowbit((int) sizeof(VT) * CC))>(blk + i * BS + c0, v);
#pragma unroll
          for (int c = 0; c < CC; ++c) acc[i] = fmadd(cvt<AT>(v[c]), xv[c], acc[i]);
        }
      }
    }
#pragma unroll
    for (int i = 0; i < R; ++i) {
      if constexpr (SCALED) acc[i] *= sc;
      epi((int64_t) br * BS + r0 + i, acc[i], pre[i]);
    }
  }
  template <class OP, class EPI>
  __device__ __forceinline__ static bool run1(const MatArgs<VT, AT> &M, const OP &op, EPI &&epi,
                                              const int *stop = nullptr)
  {
    const int64_t t = (int64_t) blockIdx.x * 256 + threadIdx.x;
    if (t < (int64_t) M.n_brows * TPB) item(M, op, epi, t, stop);
    return !(stop && __builtin_amdgcn_readfirstlane(*stop));
  }
  template <class OP, class EPI>
  __device__ __forceinline__ static void run(const MatArgs<VT, AT> &M, const OP &op, EPI &&epi)
  {
    const int64_t total = (int64_t) M.n_brows * TPB;
    for (int64_t t = (int64_t) blockIdx.x * 256 + threadIdx.x; t < total; t += (int64_t) gridDim.x * 256)
      item(M, op, epi, t);
  }
};

// Rows per thread and block rows per wavefront of SlicedRowsEngine: one
// 16-byte row load per block (FP16 bs 4: two 8-byte rows).  Measured on the
// refinement of the 2048x1024 bs 4 system (same iterations and bits): FP32
// one row per thread (16 block rows per pass) 18.2 ms against two
// (RowsEngine's choice) 18.4 and four 18.7; FP16 two rows 15.2 ms against
// one 16.9 and four 16.8.
template <class VT, int BS> constexpr int sliced_rows() { return rows_per_thread<VT, BS>(); }
template <class VT, int BS> constexpr int sliced_rpw() { return 64 / (BS / sliced_rows<VT, BS>()); }

// SlicedRowsEngine: RowsEngine over a pass-sliced copy (MatArgs::cols_sl,
// sl_j).  One wavefront per pass of RPW block rows; lane (lbr, r0) reads R
// rows of the pass's slot-j block of row lbr, so each load instruction of the
// wavefront covers one contiguous range (RPW*BS*BS values per slot); every
// lane runs the same sl_j slots in batches of 3 (sl_j is a multiple of 3;
// padding blocks are zero).  Products and their order are those of RowsEngine.
// FP16 values load non-temporal (streamed once per SpMV; x and the vectors
// keep the caches): the FP16 refinement of the 2048x1024 bs 4 system 16.6 ->
// 15.2 ms, while the FP32 one slowed 18.4 -> 20.0 ms and keeps cached loads.
template <class VT, class AT, int BS, bool SCALED>
struct SlicedRowsEngine {
  static constexpr int R = sliced_rows<VT, BS>(), TPB = BS / R, RPW = 64 / TPB, LPW = RPW * TPB, U = 3;
  static int64_t threads(int64_t n_brows) { return (n_brows + RPW - 1) / RPW * 64; }
  static int64_t groups1(int64_t n_brows) { return (threads(n_brows) + 255) / 256; }
  template <class OP, class EPI>
  __device__ __forceinline__ static bool run1(const MatArgs<VT, AT> &M, const OP &op, EPI &&epi,
                                              const int *stop = nullptr)
  {
    constexpr int VALIGN = cmin(lowbit((int) sizeof(VT) * BS * R), 16);
    constexpr int CC = BS % 8 == 0 ? 8 : BS;
    constexpr int XALIGN = cmin(lowbit((int) sizeof(AT) * CC), lowbit((int) sizeof(AT) * BS));
    const int64_t pass = ((int64_t) blockIdx.x * 256 + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int ln = lane < LPW ? lane : LPW - 1;  // spare lanes repeat the last row's loads
    const int lbr = ln / TPB, r0 = (ln - lbr * TPB) * R;
    const int64_t br = pass * RPW + lbr;
    const bool on = lane < LPW && br < M.n_brows;
    const int J = M.sl_j;
    const AT sc = (SCALED && on) ? (AT) M.scale[br] : (AT) 1;
    AT pre[R], acc[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
      pre[i] = on ? epi.pre(br * BS + r0 + i) : (AT) 0;
      acc[i] = 0;
    }
    if (stop) {  // the stop test, as in RowsEngine::item
      const int dn = __hip_atomic_load(stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      asm volatile("" ::"v"(dn));
      if (__builtin_amdgcn_readfirstlane(dn)) return false;
    }
    if (pass * RPW < M.n_brows) {  // uniform: the wavefront's pass holds rows
      const int64_t s0 = pass * J * RPW;
      for (int j = 0; j < J; j += U) {
        int col[U];
        VT v[U][R * BS];
        AT xv[U][BS];
#pragma unroll
        for (int k = 0; k < U; ++k) col[k] = M.cols_sl[s0 + (int64_t) (j + k) * RPW + lbr];
#pragma unroll
        for (int k = 0; k < U; ++k)
          load_n<VT, R * BS, VALIGN, (sizeof(VT) == 2)>(M.vals + ((s0 + (int64_t) (j + k) * RPW + lbr) * BS + r0) * BS,
                                                      v[k]);
#pragma unroll
        for (int k = 0; k < U; ++k) op.template load<BS, XALIGN>((int64_t) col[k] * BS, xv[k]);
#pragma unroll
        for (int k = 0; k < U; ++k)
#pragma unroll
          for (int i = 0; i < R; ++i)
#pragma unroll
            for (int c = 0; c < BS; ++c) acc[i] = fmadd(cvt<AT>(v[k][i * BS + c]), xv[k][c], acc[i]);
      }
    }
    if (on) {
#pragma unroll
      for (int i = 0; i < R; ++i) {
        if constexpr (SCALED) acc[i] *= sc;
        epi(br * BS + r0 + i, acc[i], pre[i]);
      }
    }
    return true;
  }
};

// Columns per lane for GroupEngine: one 16-byte load, two for FP64 bs >= 12.
template <class VT> constexpr int engine_vec(int bs)
{
  constexpr int W = 16 / (int) sizeof(VT);
  int v = 1;
  while (v < W && bs % (2 * v) == 0) v *= 2;
  if (sizeof(VT) == 8 && bs >= 12 && bs % 4 == 0) v = 4;
  return v;
}

// Engine of the fused SpMVs: shuffle-shared operands (one operand element per
// lane per block) except for row-major blocks of 12 and more columns, where a
// lane row of bs values is too long and column slices stream better; with
// the blocks staged in LDS for FP32 blocks of up to 4 columns (at most 9.2 KB
// of LDS per wavefront).
template <class VT, class AT, int BS, int LAYOUT, bool SCALED>
using SolverEngine = typename std::conditional<
    (LAYOUT == BMFP_ROW_MAJOR && BS >= 12),
    GroupEngine<VT, AT, BS, pow2ceil(BS / engine_vec<VT>(BS)), engine_vec<VT>(BS), LAYOUT, SCALED>,
    typename std::conditional<(sizeof(VT) == 4 && BS <= 4),
                              StagedShflEngine<VT, AT, BS, LAYOUT, SCALED>,
                              ShflEngine<VT, AT, BS, LAYOUT, SCALED>>::type>::type;

// Engine of the split schedule's SpMVs (a stored operand), row-major blocks:
// thread per row for FP64 bs 4 (on the 2048x1024 bs 4 system 497 / 478 us
// per SpMV against 546 / 534 us with the dense lane groups), the dense lane
// groups of the standalone FP64 kernel for 6+ columns, 16-byte row groups
// for FP32 / FP16 blocks of 4-8 columns (FP16 bs 4: batches of 9 blocks,
// 168 / 156 us against 178 / 168 with 3; FP32 bs 4: two rows, 32 bytes, per
// thread in batches of 3 blocks, the FP32 refinement of the 2048x1024 bs 4
// system 19.6 ms against 19.9 with one row in batches of 9, 21.4 with two
// rows in batches of 9 and 24.2 with four rows; staging a wavefront's 32
// block rows in LDS first, so that each load instruction covers whole cache
// lines, ran no faster in FP32 or FP16); the fused schedule's engines
// otherwise.
template <class VT, int BS> constexpr int split_vec() { return dense_shape<VT, BS>().vec > 0 ? dense_shape<VT, BS>().vec : 1; }
template <class VT, int BS> constexpr int split_u() { return dense_shape<VT, BS>().u > 0 ? dense_shape<VT, BS>().u : 1; }
template <class VT, class AT, int BS, int LAYOUT, bool SCALED>
using SplitEngine = typename std::conditional<
    (LAYOUT == BMFP_ROW_MAJOR && sizeof(VT) == 8 && BS == 4), RowsEngine<VT, AT, BS, 1, SCALED>,
    typename std::conditional<
        (LAYOUT == BMFP_ROW_MAJOR && sizeof(VT) == 8 && BS >= 6 && dense_shape<VT, BS>().vec > 0),
        DenseEngine<VT, AT, BS, split_vec<VT, BS>(), split_u<VT, BS>(), SCALED>,
        typename std::conditional<(LAYOUT == BMFP_ROW_MAJOR && sizeof(VT) < 8 && BS >= 4 && BS <= 8),
                                  RowsEngine<VT, AT, BS, (sizeof(VT) == 2 && BS <= 4 ? 9 : 3), SCALED,
                                             (sizeof(VT) == 4 && BS == 4 ? 2 : 0)>,
                                  SolverEngine<VT, AT, BS, LAYOUT, SCALED>>::type>::type>::type;

// Init: x = 0, r = r^ = b, p = v = 0 (both parities); partC = ((b,b), (b,b)).
template <class AT>
__global__ __launch_bounds__(256) void k_bicg_init(SolverVecs<AT> V, double eps, int max_iter)
{
  __shared__ double lds[2][4];
  double part[2] = {0.0, 0.0};
  for (int64_t i = (int64_t) blockIdx.x * 256 + threadIdx.x; i < V.n; i += (int64_t) gridDim.x * 256) {
    const AT bi = V.b[i];
    V.x[i] = 0;
    V.r[i] = bi;
    V.rh[i] = bi;
    V.p[0][i] = 0;
    V.p[1][i] = 0;
    V.v[0][i] = 0;
    V.v[1][i] = 0;
    part[0] += (double) bi * (double) bi;
  }
  part[1] = part[0];
  write_partials<2>(V.partC, part, lds);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    SolverState *st = V.st;
    for (int k = 0; k < 2; ++k) st->rho[k] = st->alpha[k] = st->omega[k] = st->tt[k] = 1.0;
    st->norm_b2 = 0.0;
    st->rr = 0.0;
    st->eps = eps;
    st->iter = st->iter_b = st->iter_c = 0;
    st->done = 0;
    st->status = BMFP_OK;
    st->max_iter = max_iter;
  }
}

// Head of kernel A (and of the split schedule's p kernel): reduces the
// previous iteration's (r^, r), (r, r), applies the stopping tests in a fixed
// order and, when the iteration goes on, returns beta and omega_prev.  WG 0
// records the outcome in the state.  Uniform over the workgroup.
template <class AT>
__device__ __forceinline__ bool bicg_head_a(const SolverVecs<AT> &V, double (*lds)[4], int &it, double &beta,
                                            double &omega_p)
{
  SolverState *st = V.st;
  it = st->iter;
  double red[2];
  reduce_partials<2>(V.partC, gridDim.x, red, lds);
  const double rho_c = red[0], rr = red[1];
  const double nb2 = it == 0 ? rr : st->norm_b2;
  int stop = 1;  // 1: continue; otherwise the final status
  if (it == 0 && rr == 0.0) stop = BMFP_OK;                         // b = 0: x = 0 is exact
  else if (!isfinite(rr) || !isfinite(rho_c)) stop = BMFP_ERR_NONFINITE;
  else if (it > 0 && sqrt(rr) / sqrt(nb2) < st->eps) stop = BMFP_OK;
  else if (it > 0 && (st->tt[(it - 1) & 1] == 0.0 || st->omega[(it - 1) & 1] == 0.0)) stop = BMFP_ERR_BREAKDOWN;
  else if (it >= st->max_iter) stop = BMFP_ERR_NOT_CONVERGED;
  else if (rho_c == 0.0) stop = BMFP_ERR_BREAKDOWN;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->rr = rr;
    if (it == 0) st->norm_b2 = rr;
    if (stop != 1) {
      st->status = stop;
      st->done = 1;
    } else {
      st->rho[it & 1] = rho_c;
      st->iter_b = it;
    }
  }
  if (stop != 1) return false;
  const int pv = (it - 1) & 1;
  const double rho_p = it == 0 ? 1.0 : st->rho[pv];
  const double alpha_p = it == 0 ? 1.0 : st->alpha[pv];
  omega_p = it == 0 ? 1.0 : st->omega[pv];
  beta = rho_c / rho_p * alpha_p / omega_p;
  return true;
}

// Head of kernel B (and of the split schedule's s kernel): reduces (r^, v)
// and returns alpha = rho / (r^, v), or stops on a zero or non-finite (r^, v).
template <class AT>
__device__ __forceinline__ bool bicg_head_b(const SolverVecs<AT> &V, double (*lds)[4], int &it, double &alpha)
{
  SolverState *st = V.st;
  it = st->iter_b;
  double red[1];
  reduce_partials<1>(V.partA, V.npa, red, lds);
  const double rhv = red[0];
  const double rho_c = st->rho[it & 1];
  const int stop = !isfinite(rhv) ? BMFP_ERR_NONFINITE : rhv == 0.0 ? BMFP_ERR_BREAKDOWN : 1;
  alpha = rho_c / rhv;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (stop != 1) {
      st->status = stop;
      st->done = 1;
    } else {
      st->alpha[it & 1] = alpha;
      st->iter_c = it;
    }
  }
  return stop == 1;
}

// Head of kernel C: reduces (t, s), (t, t) into omega.
template <class AT>
__device__ __forceinline__ void bicg_head_c(const SolverVecs<AT> &V, double (*lds)[4], int &it, double &alpha,
                                            double &omega, double &tt)
{
  SolverState *st = V.st;
  it = st->iter_c;
  double red[2];
  reduce_partials<2>(V.partB, V.npb, red, lds);
  const double ts = red[0];
  tt = red[1];
  omega = tt == 0.0 ? 0.0 : ts / tt;
  alpha = st->alpha[it & 1];
}

template <class AT>
__device__ __forceinline__ void bicg_tail_c(const SolverVecs<AT> &V, int it, double omega, double tt)
{
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    SolverState *st = V.st;
    st->omega[it & 1] = omega;
    st->tt[it & 1] = tt;
    st->iter = it + 1;
  }
}

template <class VT, class AT, class ENGINE>
__global__ __launch_bounds__(256) void k_bicg_a(MatArgs<VT, AT> M, SolverVecs<AT> V)
{
  __shared__ double lds[2][4];
  if (solver_stopped(V.st)) return;
  int it;
  double beta, omega_p;
  if (!bicg_head_a(V, lds, it, beta, omega_p)) return;
  const int pv = (it - 1) & 1, cu = it & 1;
  OperandP<AT> op{V.r, V.p[pv], V.v[pv], V.pinv, (AT) beta, (AT) omega_p};
  double part[1] = {0.0};
  AT *__restrict__ pc = V.p[cu];
  AT *__restrict__ vc = V.v[cu];
  ENGINE::run(M, op, [&](int64_t row, AT y) {
    pc[row] = op.raw(row);
    vc[row] = y;
    part[0] += (double) V.rh[row] * (double) y;
  });
  write_partials<1>(V.partA, part, lds);
}

template <class VT, class AT, class ENGINE>
__global__ __launch_bounds__(256) void k_bicg_b(MatArgs<VT, AT> M, SolverVecs<AT> V)
{
  __shared__ double lds[2][4];
  if (solver_stopped(V.st)) return;
  int it;
  double alpha;
  if (!bicg_head_b(V, lds, it, alpha)) return;
  OperandS<AT> op{V.r, V.v[it & 1], V.pinv, (AT) alpha};
  double part[2] = {0.0, 0.0};
  ENGINE::run(M, op, [&](int64_t row, AT y) {
    const AT s = op.raw(row);
    V.s[row] = s;
    V.t[row] = y;
    part[0] += (double) y * (double) s;
    part[1] += (double) y * (double) y;
  });
  write_partials<2>(V.partB, part, lds);
}

template <class AT>
__global__ __launch_bounds__(256) void k_bicg_c(SolverVecs<AT> V)
{
  __shared__ double lds[2][4];
  if (solver_stopped(V.st)) return;
  int it;
  double alpha, omega, tt;
  bicg_head_c(V, lds, it, alpha, omega, tt);
  const AT a = (AT) alpha, w = (AT) omega;
  const AT *__restrict__ p = V.p[it & 1];
  double part[2] = {0.0, 0.0};
  for (int64_t i = (int64_t) blockIdx.x * 256 + threadIdx.x; i < V.n; i += (int64_t) gridDim.x * 256) {
    const AT si = V.s[i];
    const AT q = V.pinv ? V.pinv[i] * p[i] : p[i];
    const AT z = V.pinv ? V.pinv[i] * si : si;
    const AT h = fmadd(a, q, V.x[i]);
    const AT rn = fmadd(-w, V.t[i], si);
    V.x[i] = fmadd(w, z, h);
    V.r[i] = rn;
    part[0] += (double) V.rh[i] * (double) rn;
    part[1] += (double) rn * (double) rn;
  }
  write_partials<2>(V.partC, part, lds);
  bicg_tail_c(V, it, omega, tt);
}

// ---- split schedule ------------------------------------------------------------------
// The same iteration as kernels A, B, C, in five kernels whose SpMVs read a
// stored operand instead of forming it at every gathered index:
//   P   p = r + beta (p - omega v)  [q = P^-1 p]            (elementwise)
//   SV  v = A q, partial (r^, v) per workgroup              (SpMV)
//   R   (r^, v) summed in a fixed order                     (16 workgroups)
//   S   s = r - alpha v  [z = P^-1 s]                       (elementwise)
//   ST  t = A z, partial (t, s), (t, t) per workgroup       (SpMV)
//   R   (t, s), (t, t) summed in a fixed order              (16 workgroups)
//   X   r = s - omega t, x = (x + alpha q) + omega z, partial (r^, r), (r, r)
// Same formulas and the same fused multiply-adds per element as A, B, C; the
// SpMV runs on the dense engines of the standalone kernels, so it moves its
// bytes at the standalone rate (forming the operand at each gather cost the
// fused kernels 30-60% per SpMV against about 8% for the extra passes).
// Buffers: p = p[0], q = p[1], v = v[0], z = v[1] (q, z with Jacobi only;
// without it the SpMVs read p and s directly).
template <class AT>
__global__ __launch_bounds__(256) void k_bicg_sp(SolverVecs<AT> V)
{
  __shared__ double lds[2][4];
  if (solver_stopped(V.st)) return;
  int it;
  double beta, omega_p;
  if (!bicg_head_a(V, lds, it, beta, omega_p)) return;
  const AT bt = (AT) beta, om = (AT) omega_p;
  AT *__restrict__ p = V.p[0];
  AT *__restrict__ q = V.p[1];
  const AT *__restrict__ v = V.v[0];
  for (int64_t i = (int64_t) blockIdx.x * 256 + threadIdx.x; i < V.n; i += (int64_t) gridDim.x * 256) {
    const AT pn = p_update(V.r[i], p[i], v[i], bt, om);
    p[i] = pn;
    if (V.pinv) q[i] = V.pinv[i] * pn;
  }
}

// Engines with a one-tile-per-wavefront launch (groups1) run the split
// schedule's SpMVs in that shape; others in their persistent grid.
template <class E, class = void> struct one_tile : std::false_type {};
template <class E> struct one_tile<E, decltype((void) &E::groups1)> : std::true_type {};

// The stop flag: a one-tile launch has tens of thousands of workgroups, so
// each wavefront tests the flag itself (a scalar load, no workgroup barrier),
// after issuing its row bounds (the test's latency then overlaps them), and a
// stopped solve's launches end at once; persistent engines check it once per
// workgroup.  The flag is written only by earlier kernels, so every
// wavefront of a workgroup takes the same branch.
template <class ENGINE, class VT, class AT, class OP, class EPI>
__device__ __forceinline__ bool split_run(const MatArgs<VT, AT> &M, const OP &op, EPI &&epi, const SolverState *st)
{
  if constexpr (one_tile<ENGINE>::value) {
    return ENGINE::run1(M, op, epi, &st->done);
  } else {
    if (solver_stopped(st)) return false;
    ENGINE::run(M, op, epi);
  }
  return true;
}

// Split SpMV epilogues: store y and accumulate its dot products.  The
// one-tile engines call pre(row) with the row bounds and pass the value on;
// the persistent engines call the two-argument form.  v and t are stored
// non-temporal (the SpMV's L2 then keeps its operand and matrix lines): on
// the 2048x1024 bs 4 system the FP64 solve 37.3 -> 36.9 ms, the FP16
// refinement 15.1 -> 14.7 ms, FP32 unchanged, same iterations and bits.
template <class AT>
struct EpiV {  // v = y; (r^, v)
  AT *__restrict__ v;
  const AT *__restrict__ rh;
  double part[1] = {0.0};
  __device__ __forceinline__ AT pre(int64_t row) const { return rh[row]; }
  __device__ __forceinline__ void operator()(int64_t row, AT y, AT rhv)
  {
    __builtin_nontemporal_store(y, &v[row]);
    part[0] += (double) rhv * (double) y;
  }
  __device__ __forceinline__ void operator()(int64_t row, AT y) { (*this)(row, y, rh[row]); }
};
template <class AT>
struct EpiT {  // t = y; (t, s), (t, t)
  AT *__restrict__ t;
  const AT *__restrict__ s;
  double part[2] = {0.0, 0.0};
  __device__ __forceinline__ AT pre(int64_t row) const { return s[row]; }
  __device__ __forceinline__ void operator()(int64_t row, AT y, AT si)
  {
    __builtin_nontemporal_store(y, &t[row]);
    part[0] += (double) y * (double) si;
    part[1] += (double) y * (double) y;
  }
  __device__ __forceinline__ void operator()(int64_t row, AT y) { (*this)(row, y, s[row]); }
};

template <class VT, class AT, class ENGINE>
__global__ __launch_bounds__(256) void k_bicg_sv(MatArgs<VT, AT> M, SolverVecs<AT> V)
{
  __shared__ double lds[2][4];
  const OperandVec<AT> op{V.pinv ? V.p[1] : V.p[0]};
  EpiV<AT> epi{V.v[0], V.rh};
  if (split_run<ENGINE>(M, op, epi, V.st)) write_partials<1>(V.partW, epi.part, lds);
}

// Sums the `count` per-workgroup partials of a split SpMV (K values each; st
// may be null: no stop flag):
// workgroup g of the REDUCE_GROUPS adds the contiguous slice g of them in a
// fixed order and writes out[g*K + k]; the next kernel's head adds those
// REDUCE_GROUPS values in order (V.npa / V.npb).  One workgroup alone took
// 8-10 us for 33k partials.
constexpr int REDUCE_GROUPS = 16;
template <int K>
__global__ __launch_bounds__(256) void k_bicg_reduce(const double *__restrict__ in, int count,
                                                     double *__restrict__ out, const SolverState *st)
{
  __shared__ double lds[2][4];
  if (st && st->done) return;  // uniform: written only by earlier kernels
  const int per = (count + REDUCE_GROUPS - 1) / REDUCE_GROUPS;
  const int g0 = blockIdx.x * per, g1 = g0 + per < count ? g0 + per : count;
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = 0.0;
#pragma unroll 4
  for (int g = g0 + threadIdx.x; g < g1; g += 256)
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] += in[(int64_t) g * K + k];
  write_partials<K>(out, v, lds);
}

template <class AT>
__global__ __launch_bounds__(256) void k_bicg_ss(SolverVecs<AT> V)
{
  __shared__ double lds[2][4];
  if (solver_stopped(V.st)) return;
  int it;
  double alpha;
  if (!bicg_head_b(V, lds, it, alpha)) return;
  const AT a = (AT) alpha;
  const AT *__restrict__ v = V.v[0];
  AT *__restrict__ z = V.v[1];
  for (int64_t i = (int64_t) blockIdx.x * 256 + threadIdx.x; i < V.n; i += (int64_t) gridDim.x * 256) {
    const AT si = s_update(V.r[i], v[i], a);
    V.s[i] = si;
    if (V.pinv) z[i] = V.pinv[i] * si;
  }
}

template <class VT, class AT, class ENGINE>
__global__ __launch_bounds__(256) void k_bicg_st(MatArgs<VT, AT> M, SolverVecs<AT> V)
{
  __shared__ double lds[2][4];
  const OperandVec<AT> op{V.pinv ? V.v[1] : V.s};
  EpiT<AT> epi{V.t, V.s};
  if (split_run<ENGINE>(M, op, epi, V.st)) write_partials<2>(V.partW, epi.part, lds);
}

template <class AT>
__global__ __launch_bounds__(256) void k_bicg_sx(SolverVecs<AT> V)
{
  __shared__ double lds[2][4];
  if (solver_stopped(V.st)) return;
  int it;
  double alpha, omega, tt;
  bicg_head_c(V, lds, it, alpha, omega, tt);
  const AT a = (AT) alpha, w = (AT) omega;
  const AT *__restrict__ q = V.pinv ? V.p[1] : V.p[0];
  const AT *__restrict__ z = V.pinv ? V.v[1] : V.s;
  double part[2] = {0.0, 0.0};
  for (int64_t i = (int64_t) blockIdx.x * 256 + threadIdx.x; i < V.n; i += (int64_t) gridDim.x * 256) {
    const AT si = V.s[i];
    const AT h = fmadd(a, q[i], V.x[i]);
    const AT rn = fmadd(-w, V.t[i], si);
    V.x[i] = fmadd(w, z[i], h);
    V.r[i] = rn;
    part[0] += (double) V.rh[i] * (double) rn;
    part[1] += (double) rn * (double) rn;
  }
  write_partials<2>(V.partC, part, lds);
  bicg_tail_c(V, it, omega, tt);
}

// Point-Jacobi inverse diagonal from BCSR diagonal blocks (bicgstab.cu:201-225:
// 1/a_ii, or 1 when |a_ii| < 1e-20 or the diagonal block is absent).
template <class VT, class AT, int LAYOUT, bool SCALED>
__global__ __launch_bounds__(256) void k_jacobi(MatArgs<VT, AT> M, int bs, AT *__restrict__ pinv)
{
  const int64_t row = (int64_t) blockIdx.x * 256 + threadIdx.x;
  if (row >= (int64_t) M.n_brows * bs) return;
  const int br = (int) (row / bs), r = (int) (row % bs);
  double d = 0.0;
  bool found = false;
  for (int b = M.rp[br]; b < M.rp[br + 1]; ++b)
    if (M.cols[b] == br) {
      const VT v = M.vals[(int64_t) b * bs * bs + (int64_t) r * bs + r];
      d = (double) cvt<AT>(v);
      if (SCALED) d *= (double) M.scale[br];
      found = true;
      break;
    }
  pinv[row] = (AT) ((!found || fabs(d) < 1e-20) ? 1.0 : 1.0 / d);
}

}  // namespace bmfp

namespace bmfp {

// Refinement residual of a stored iterate, for the split SpMV engines:
// r = b - A x in FP64, r32 = (float) r and the partial sums ((r, r), (b, b))
// per workgroup (one-tile engines) or per persistent workgroup.
struct EpiR {
  const double *__restrict__ b;
  float *__restrict__ r32;
  double part[2] = {0.0, 0.0};
  __device__ __forceinline__ double pre(int64_t row) const { return b[row]; }
  __device__ __forceinline__ void operator()(int64_t row, double y, double bi)
  {
    const double ri = bi - y;
    r32[row] = (float) ri;
    part[0] += ri * ri;
    part[1] += bi * bi;
  }
  __device__ __forceinline__ void operator()(int64_t row, double y) { (*this)(row, y, b[row]); }
};

template <class ENGINE>
__global__ __launch_bounds__(256) void k_refine_resid_split(MatArgs<double, double> M, const double *__restrict__ x,
                                                            const double *__restrict__ b, float *__restrict__ r32,
                                                            double *__restrict__ part)
{
  __shared__ double lds[2][4];
  const OperandVec<double> op{x};
  EpiR epi{b, r32};
  if constexpr (one_tile<ENGINE>::value)
    ENGINE::run1(M, op, epi);
  else
    ENGINE::run(M, op, epi);
  write_partials<2>(part, epi.part, lds);
}

}  // namespace bmfp
