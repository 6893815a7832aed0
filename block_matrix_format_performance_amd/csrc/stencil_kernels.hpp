// stencil_kernels.hpp -- the 9-point block stencil generated in device memory.
//
// Produces exactly the values of the host generator (generate.cpp, built with
// -ffp-contract=off): the same counter-based random draws, the same order of
// the diagonal row sums, and floating-point contraction switched off in the
// kernel (HIP's __dadd_rn / __dmul_rn are plain + and *, which the device
// compiler would otherwise fuse into a multiply-add).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "common.hpp"
#include "spmv_kernels.hpp"

namespace bmfp {

inline int64_t stencil_block_count(int nx, int ny) { return (3 * (int64_t) nx - 2) * (3 * (int64_t) ny - 2); }

// Block row pointers of the stencil: the 3x3 window of each node clipped at
// the grid edges, nodes numbered iy*nx + ix.
inline void stencil_row_ptr(int nx, int ny, std::vector<int> &rp)
{
  rp.resize((size_t) nx * ny + 1);
  int64_t nb = 0;
  rp[0] = 0;
  for (int iy = 0; iy < ny; iy++)
    for (int ix = 0; ix < nx; ix++) {
      nb += (3 - (iy == 0) - (iy == ny - 1)) * (3 - (ix == 0) - (ix == nx - 1));
      rp[(size_t) iy * nx + ix + 1] = (int) nb;
    }
}

namespace {

// One thread per scalar row t = node*bs + r: writes row r of each of the
// node's blocks (row-major blocks) and, for r == 0, the block columns; with
// vals32 also the FP32 rounding of every value.  Grids up to 2^31 scalar rows
// (32-bit index arithmetic).
__global__ __launch_bounds__(256) void k_gen_stencil_rows(int nx, int ny, int bs, int kind, uint64_t seed,
                                                          const int *__restrict__ rp, int *__restrict__ cols,
                                                          double *__restrict__ vals, float *__restrict__ vals32)
{
#pragma clang fp contract(off)
  const int64_t t64 = (int64_t) blockIdx.x * 256 + threadIdx.x;
  if (t64 >= (int64_t) nx * ny * bs) return;
  const int t = (int) t64;
  const int i = t / bs;
  const int r = t - i * bs;
  const int iy = i / nx, ix = i - iy * nx;
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
    const int64_t at = nb * bs2 + (int64_t) r * bs;
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
        v = -(k + 0.3 * cc);
      }
      if (!(s == 4 && c == r)) {
        sum += fabs(v);
        vals[at + c] = v;
        if (vals32) vals32[at + c] = (float) v;
      } else if (kind == 0) {
        vals[at + c] = v;  // the random part of the diagonal entry; the shift is added below
      }
    }
    nb++;
  }
  const int64_t dg = diag * bs2 + (int64_t) r * bs + r;
  const double dv = kind == 0 ? vals[dg] + (sum + 1.0) : 1.02 * sum + 1.0;
  vals[dg] = dv;
  if (vals32) vals32[dg] = (float) dv;
}

// N values at p, whose byte address is a multiple of the lowest set bit of
// N*sizeof(T), in stores of up to 16 bytes.
template <class T, int N> __device__ __forceinline__ void store_row(T *p, const T (&v)[N])
{
  constexpr int BYTES = N * (int) sizeof(T);
  constexpr int LOW = BYTES & -BYTES;
  constexpr int PER = (LOW < 16 ? LOW : 16) / (int) sizeof(T);
  if constexpr (PER <= 1) {
#pragma unroll
    for (int c = 0; c < N; ++c) p[c] = v[c];
  } else {
    typedef T vec_t __attribute__((ext_vector_type(PER)));
#pragma unroll
    for (int c = 0; c < N; c += PER) {
      vec_t t;
#pragma unroll
      for (int e = 0; e < PER; ++e) t[e] = v[c + e];
      *reinterpret_cast<vec_t *>(p + c) = t;
    }
  }
}

// Copies in the pass-sliced order of the request loop's matrices: block j of
// node i is sliced block ((i / RPW) * j_slots + j) * RPW + i % RPW (RPW =
// 64 / BS nodes per pass), rows are padded to j_slots blocks with zero blocks
// whose column is the row's last one, and the nodes past the grid in the
// last pass get zero blocks and column 0.  vals: FP32 values, vals64: FP64
// values, cols: the sliced block columns (written when either copy is).
// vals16 (bs 4, 6 and 8 only, beside vals64): FP16 values, each FP64 value v
// of block row i stored as half(float(v * (1 / scale_i))) with scale_i the row's
// largest |entry| rounded to FP32 (1 for an all-zero row).
struct SlicedOut {
  float *vals = nullptr;
  int *cols = nullptr;
  int j_slots = 0;
  double *vals64 = nullptr;
  __half *vals16 = nullptr;
  __host__ __device__ bool on() const { return vals != nullptr || vals64 != nullptr; }
};

// Row of BS FP16 values v[c] * inv at p (p a multiple of 2*BS bytes), in
// 4- to 16-byte stores.
template <int BS>
__device__ __forceinline__ void put_half_row(__half *p, const double (&v)[BS], double inv)
{
  __half h[BS];
#pragma unroll
  for (int c = 0; c < BS; c++) h[c] = __float2half_rn((float) (v[c] * inv));
  store_words<__half, BS>(p, h);
}

// Row r (of BS values) of sliced slot `at` (the element offset of the row) in
// the FP64 and FP32 sliced copies that are present.
template <int BS>
__device__ __forceinline__ void put_sliced_row(const SlicedOut &sl, int64_t at, const double (&v)[BS])
{
  if (sl.vals64) store_row<double, BS>(sl.vals64 + at, v);
  if (sl.vals) {
    float f[BS];
#pragma unroll
    for (int c = 0; c < BS; c++) f[c] = (float) v[c];
    store_row<float, BS>(sl.vals + at, f);
  }
}

// k_gen_stencil_rows for a compile-time block size: the same draws, the same
// arithmetic and the same order of the diagonal row sum, with each block row
// held in registers and written by up-to-16-byte stores (the generic kernel
// stores one value at a time).  The diagonal block's row is written once, at
// the end, with its diagonal entry in place.
template <int BS>
__global__ __launch_bounds__(256) void k_gen_stencil_rows_t(int nx, int ny, int kind, uint64_t seed,
                                                            const int *__restrict__ rp, int *__restrict__ cols,
                                                            double *__restrict__ vals, float *__restrict__ vals32,
                                                            SlicedOut sl = SlicedOut{})
{
#pragma clang fp contract(off)
  constexpr int RPW = 64 / BS;
  const int64_t n_nodes = (int64_t) nx * ny;
  const int64_t t64 = (int64_t) blockIdx.x * 256 + threadIdx.x;
  const int64_t n_pad = sl.on() ? (n_nodes + RPW - 1) / RPW * RPW : n_nodes;
  if (t64 >= n_pad * BS) return;
  const int t = (int) t64;
  const int i = t / BS;
  const int r = t - i * BS;
  // the row of slot j of node i in the sliced copies
  auto sl_row = [&](int j) { return (((int64_t) (i / RPW) * sl.j_slots + j) * RPW + i % RPW) * (BS * BS) + r * BS; };
  auto sl_col = [&](int j) { return sl.cols + ((int64_t) (i / RPW) * sl.j_slots + j) * RPW + i % RPW; };
  if (i >= n_nodes) {  // a node past the grid in the last pass: zero blocks
    const double z[BS] = {};
    for (int j = 0; j < sl.j_slots; j++) {
      put_sliced_row<BS>(sl, sl_row(j), z);
      if (r == 0) *sl_col(j) = 0;
    }
    return;
  }
  const int iy = i / nx, ix = i - iy * nx;
  const uint64_t ka = rng_key(seed, kind == 0 ? STREAM_DD : STREAM_K), kc = rng_key(seed, STREAM_C);
  constexpr int64_t bs2 = (int64_t) BS * BS;
  const int64_t nb0 = rp[i];
  int64_t nb = nb0;
  int64_t diag_at = 0;
  int diag_j = 0, last_col = 0;
  double dg[BS];
  double sum = 0.0;
  for (int s = 0; s < 9; s++) {
    const int jy = iy + s / 3 - 1, jx = ix + s % 3 - 1;
    if (jy < 0 || jy >= ny || jx < 0 || jx >= nx) continue;
    last_col = jy * nx + jx;
    if (r == 0) {
      cols[nb] = last_col;
      if (sl.on()) *sl_col((int) (nb - nb0)) = last_col;
    }
    const uint64_t base = ((uint64_t) i * 9u + (uint64_t) s) * (uint64_t) bs2;
    const int64_t at = nb * bs2 + (int64_t) r * BS;
    double v[BS];
#pragma unroll
    for (int c = 0; c < BS; c++) {
      if (kind == 0) {
        v[c] = rng_sym(ka, base + (uint64_t) r * BS + c);
      } else if (s == 4) {
        v[c] = 0.0;
      } else {
        const int lo = r < c ? r : c, hi = r < c ? c : r;
        const double k = rng_unit(ka, base + (uint64_t) lo * BS + hi);
        double cc = 0.0;
        if (r < c) cc = rng_sym(kc, base + (uint64_t) r * BS + c);
        else if (r > c) cc = -rng_sym(kc, base + (uint64_t) c * BS + r);
        v[c] = -(k + 0.3 * cc);
      }
      if (!(s == 4 && c == r)) sum += fabs(v[c]);
    }
    if (s == 4) {
      diag_at = at;
      diag_j = (int) (nb - nb0);
#pragma unroll
      for (int c = 0; c < BS; c++) dg[c] = v[c];
    } else {
      if (vals) store_row<double, BS>(vals + at, v);
      if (vals32) {
        float f[BS];
#pragma unroll
        for (int c = 0; c < BS; c++) f[c] = (float) v[c];
        store_row<float, BS>(vals32 + at, f);
      }
      if (sl.on()) put_sliced_row<BS>(sl, sl_row((int) (nb - nb0)), v);
    }
    nb++;
  }
#pragma unroll
  for (int c = 0; c < BS; c++)
    if (c == r) dg[c] = kind == 0 ? dg[c] + (sum + 1.0) : 1.02 * sum + 1.0;
  if (vals) store_row<double, BS>(vals + diag_at, dg);
  if (vals32) {
    float f[BS];
#pragma unroll
    for (int c = 0; c < BS; c++) f[c] = (float) dg[c];
    store_row<float, BS>(vals32 + diag_at, f);
  }
  if (sl.on()) {
    put_sliced_row<BS>(sl, sl_row(diag_j), dg);
    const double z[BS] = {};  // padding slots
    for (int j = (int) (nb - nb0); j < sl.j_slots; j++) {
      put_sliced_row<BS>(sl, sl_row(j), z);
      if (r == 0) *sl_col(j) = last_col;
    }
  }
}

// Stores the wavefront's blocks staged in LDS (NPW blocks of BS*BS values,
// node q's at q*BS*BS): lanes take 16-byte chunks in memory order, block q
// going to block index nb of the lane holding node q's row 0 (skipped where
// that node has no such block); the FP32 copy in 8-byte chunks.
template <int BS>
__device__ __forceinline__ void store_staged_blocks(const double *blk, int nb, bool present, double *vals,
                                                    float *vals32, int lane)
{
  typedef double d2 __attribute__((ext_vector_type(2)));
  typedef float f2 __attribute__((ext_vector_type(2)));
  constexpr int NPW = 64 / BS, BS2 = BS * BS, CPB = BS2 / 2, NCH = (NPW * CPB + 63) / 64;
#pragma unroll
  for (int k = 0; k < NCH; ++k) {
    const int ch = lane + 64 * k;
    const int qq = ch / CPB < NPW ? ch / CPB : NPW - 1;
    const int nbq = __shfl(nb, qq * BS);
    const int pq = __shfl((int) present, qq * BS);
    if (ch < NPW * CPB && pq) {
      const int off = (ch - qq * CPB) * 2;
      const d2 v = *reinterpret_cast<const d2 *>(blk + qq * BS2 + off);
      const int64_t at = (int64_t) nbq * BS2 + off;
      *reinterpret_cast<d2 *>(vals + at) = v;
      if (vals32) *reinterpret_cast<f2 *>(vals32 + at) = f2{(float) v.x, (float) v.y};
    }
  }
}

// k_gen_stencil_rows_t with the stores coalesced, for bs 4, 6 and 8; with
// `scale`, also each block row's largest |entry| rounded to FP32 (1 for an
// all-zero row), the scale of an FP16 copy, and with sl.vals16 that copy,
// pass-sliced.  There a
// lane's 16-byte stores of its block row still leave every store instruction
// spread over one cache line per row.  A wavefront holds NPW = 64/BS
// consecutive nodes; for each neighbour s the lanes put their block rows in
// the wavefront's LDS slice, then store the NPW blocks (each contiguous in
// memory) chunk by chunk in memory order.  The diagonal block goes last, its
// diagonal entries in place.  Same draws, arithmetic and row-sum order as
// k_gen_stencil_rows_t; no workgroup barrier (each wavefront uses its own
// slice).
template <int BS>
__global__ __launch_bounds__(256) void k_gen_stencil_rows_c(int nx, int ny, int kind, uint64_t seed,
                                                            const int *__restrict__ rp, int *__restrict__ cols,
                                                            double *__restrict__ vals, float *__restrict__ vals32,
                                                            float *__restrict__ scale, SlicedOut sl = SlicedOut{})
{
#pragma clang fp contract(off)
  constexpr int NPW = 64 / BS, BS2 = BS * BS;
  static_assert(BS2 % 2 == 0, "blocks of whole 16-byte chunks");
  __shared__ __align__(16) double blk_s[4][NPW * BS2];
  const int lane = threadIdx.x & 63, wib = threadIdx.x >> 6;
  const int64_t n_nodes = (int64_t) nx * ny;
  const int64_t i0 = ((int64_t) blockIdx.x * 4 + wib) * NPW;
  if (i0 >= n_nodes) return;  // the whole wavefront leaves; no barrier follows
  const int q = lane / BS, r = lane - q * BS;
  const bool on = q < NPW && i0 + q < n_nodes;
  const int i = (int) (on ? i0 + q : i0);
  const int iy = i / nx, ix = i - iy * nx;
  const uint64_t ka = rng_key(seed, kind == 0 ? STREAM_DD : STREAM_K), kc = rng_key(seed, STREAM_C);
  double *blk = blk_s[wib];
  const int nb0 = rp[i];
  int nb = nb0, diag_nb = 0, last_col = 0;
  // the wavefront is one pass of the sliced copies: lane (q, r) writes row r
  // of its node's blocks there (nodes past the grid: zero blocks).  Without
  // the BCSR FP64 output (vals null) no block is staged in LDS.
  const bool sl_on = sl.on() && q < NPW;
  auto sl_row = [&](int j) { return (((int64_t) (i0 / NPW) * sl.j_slots + j) * NPW + q) * BS2 + r * BS; };
  auto sl_col = [&](int j) { return sl.cols + ((int64_t) (i0 / NPW) * sl.j_slots + j) * NPW + q; };
  // row r of the node's block toward neighbour s, before the diagonal shift
  auto draw = [&](int s, double (&v)[BS]) {
    const uint64_t base = ((uint64_t) i * 9u + (uint64_t) s) * (uint64_t) BS2;
#pragma unroll
    for (int c = 0; c < BS; c++) {
      if (kind == 0) {
        v[c] = rng_sym(ka, base + (uint64_t) r * BS + c);
      } else if (s == 4) {
        v[c] = 0.0;
      } else {
        const int lo = r < c ? r : c, hi = r < c ? c : r;
        const double k = rng_unit(ka, base + (uint64_t) lo * BS + hi);
        double cc = 0.0;
        if (r < c) cc = rng_sym(kc, base + (uint64_t) r * BS + c);
        else if (r > c) cc = -rng_sym(kc, base + (uint64_t) c * BS + r);
        v[c] = -(k + 0.3 * cc);
      }
    }
  };
  double dg[BS];
  double sum = 0.0, mx = 0.0;  // mx: largest |stored entry| of the lane's row
  for (int s = 0; s < 9; s++) {
    const int jy = iy + s / 3 - 1, jx = ix + s % 3 - 1;
    const bool present = on && jy >= 0 && jy < ny && jx >= 0 && jx < nx;
    double v[BS];
#pragma unroll
    for (int c = 0; c < BS; c++) v[c] = 0.0;
    if (present) {
      last_col = jy * nx + jx;
      if (r == 0) {
        cols[nb] = last_col;
        if (sl_on) *sl_col(nb - nb0) = last_col;
      }
      draw(s, v);
#pragma unroll
      for (int c = 0; c < BS; c++)
        if (!(s == 4 && c == r)) sum += fabs(v[c]);
    }
    if (s == 4) {
      diag_nb = nb;
#pragma unroll
      for (int c = 0; c < BS; c++) dg[c] = v[c];
    } else {
      if (present) {
#pragma unroll
        for (int c = 0; c < BS; c++) {
          if (vals) blk[q * BS2 + r * BS + c] = v[c];
          mx = fmax(mx, fabs(v[c]));
        }
        if (sl_on) put_sliced_row<BS>(sl, sl_row(nb - nb0), v);
      }
      if (vals) {  // (uniform)
        wave_lds_sync();
        store_staged_blocks<BS>(blk, nb, present, vals, vals32, lane);
        wave_lds_sync();  // the stores' LDS reads come before the next block's writes
      }
    }
    if (present) nb++;
  }
  if (on) {
#pragma unroll
    for (int c = 0; c < BS; c++) {
      if (c == r) dg[c] = kind == 0 ? dg[c] + (sum + 1.0) : 1.02 * sum + 1.0;
      if (vals) blk[q * BS2 + r * BS + c] = dg[c];
      mx = fmax(mx, fabs(dg[c]));
    }
  }
  if (vals) {
    wave_lds_sync();
    store_staged_blocks<BS>(blk, diag_nb, on, vals, vals32, lane);
  }
  if (sl_on) {
    if (on) put_sliced_row<BS>(sl, sl_row(diag_nb - nb0), dg);
    // padding slots (all of them for a node past the grid)
    const double z[BS] = {};
    for (int j = on ? nb - nb0 : 0; j < sl.j_slots; j++) {
      put_sliced_row<BS>(sl, sl_row(j), z);
      if (r == 0) *sl_col(j) = on ? last_col : 0;
    }
  }
  if (scale || sl.vals16) {  // (uniform)
    // the node's BS rows are lanes q*BS .. q*BS + BS - 1
    double m = 0.0;
#pragma unroll
    for (int k = 0; k < BS; k++) m = fmax(m, __shfl(mx, (q * BS + k) & 63));
    const float sc = m == 0.0 ? 1.0f : (float) m;
    if (scale && on && r == 0) scale[i] = sc;
    if (sl.vals16 && q < NPW) {
      // the FP16 copy needs the whole block row's scale first: the
      // off-diagonal rows are drawn again (the same draws), the diagonal row
      // is dg, then the padding slots (all of them past the grid) are zeros
      const double inv = 1.0 / (double) sc;
      int j = 0;
      for (int s = 0; s < 9; s++) {
        const int jy = iy + s / 3 - 1, jx = ix + s % 3 - 1;
        if (!(on && jy >= 0 && jy < ny && jx >= 0 && jx < nx)) continue;
        double v[BS];
        if (s == 4) {
#pragma unroll
          for (int c = 0; c < BS; c++) v[c] = dg[c];
        } else {
          draw(s, v);
        }
        put_half_row<BS>(sl.vals16 + sl_row(j), v, inv);
        j++;
      }
      const double z[BS] = {};
      for (; j < sl.j_slots; j++) put_half_row<BS>(sl.vals16 + sl_row(j), z, inv);
    }
  }
}

// The generator for any block size: the register-row kernel for the block
// sizes the request loop serves (its BCSR-order FP64 stores coalesced through
// LDS for 4, 6 and 8), the generic one otherwise.  `vals` null: no BCSR-order
// values.  `scale` (FP16 block-row scales) is written for bs 4, 6 and 8 only;
// `sl` (pass-sliced FP64 and / or FP32 copies) for bs 1, 2, 3, 4, 6, 8, its
// FP16 copy for bs 4, 6 and 8.
inline void launch_gen_stencil(int nx, int ny, int bs, int kind, uint64_t seed, const int *rp, int *cols,
                               double *vals, float *vals32, hipStream_t s, float *scale = nullptr,
                               SlicedOut sl = SlicedOut{})
{
  // threads of the register-row kernels: the nodes of whole passes with a sliced copy
  const int64_t nodes = sl.on() ? ((int64_t) nx * ny + 64 / bs - 1) / (64 / bs) * (64 / bs) : (int64_t) nx * ny;
  const unsigned g = (unsigned) ((nodes * bs + 255) / 256);
  auto waves4 = [&](int npw) { return (unsigned) (((int64_t) nx * ny + 4 * npw - 1) / (4 * npw)); };
  switch (bs) {
  case 1: k_gen_stencil_rows_t<1><<<g, 256, 0, s>>>(nx, ny, kind, seed, rp, cols, vals, vals32, sl); break;
  case 2: k_gen_stencil_rows_t<2><<<g, 256, 0, s>>>(nx, ny, kind, seed, rp, cols, vals, vals32, sl); break;
  case 3: k_gen_stencil_rows_t<3><<<g, 256, 0, s>>>(nx, ny, kind, seed, rp, cols, vals, vals32, sl); break;
  case 4: k_gen_stencil_rows_c<4><<<waves4(16), 256, 0, s>>>(nx, ny, kind, seed, rp, cols, vals, vals32, scale, sl); break;
  case 6: k_gen_stencil_rows_c<6><<<waves4(10), 256, 0, s>>>(nx, ny, kind, seed, rp, cols, vals, vals32, scale, sl); break;
  case 8: k_gen_stencil_rows_c<8><<<waves4(8), 256, 0, s>>>(nx, ny, kind, seed, rp, cols, vals, vals32, scale, sl); break;
  default: k_gen_stencil_rows<<<g, 256, 0, s>>>(nx, ny, bs, kind, seed, rp, cols, vals, vals32); break;
  }
}

}  // namespace
}  // namespace bmfp
