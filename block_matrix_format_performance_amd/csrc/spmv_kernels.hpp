// spmv_kernels.hpp -- gfx950 SpMV kernels (BCSR and CSR), templated on the
// value storage type VT, the vector/accumulator type AT and the block size BS.
//
// All kernels are HBM-bandwidth bound (0.14-0.25 flop/byte in FP64), so the
// design goals are: every value byte read exactly once with wide (up to
// 16-byte) loads, x re-reads served from L2/MALL, no atomics (results are
// bitwise reproducible run to run), 64-lane wavefront reductions done with
// cross-lane shuffles in a fixed order.
//
// Mapping of the reference's CUDA kernels (gpu_matrix_multiplier.cu) onto the
// families here:
//   k_bcsr_tpr   thread per scalar row       K4 :275-304, K8 :341-370, K9 :372-400
//   k_bcsr_gpr   lane group per scalar row   K5 :306-339 (32 fixed lanes there)
//   k_bcsr_stage lane group per block row,   the same family for bs 1-3
//                wavefront's blocks in LDS,
//                blocks split over its lanes
//   k_bcsr_split as k_bcsr_stage without     the same family (FP16 bs 3:
//                the LDS copy                blocks not of whole 4-byte words)
//   k_bcsr_shfx  thread per row, x segment   K10/K11 :402-469 (there: shared memory,
//                shared by lane shuffles      wrong for bs<32; here: per block row)
//   k_bcsr_bgr   lane group per block row    K12-K16 :485-673 (there: wrong for
//                (column-major streaming)    bs in {3,5,6,9,10}; here: exact)
//   k_csr_scalar thread per row              K2 :65-86
//   k_csr_vector lane group per row          K3 :165-196
//   k_csr_stage  thread / lane group per     K2 and K3 where a wavefront's
//                row, wavefront's entries    rows fill its LDS slice
//                in LDS
//   k_csr_merge  merge-path tiles            (no kernel in the tree; the
//                                             "GPU Merge SpMV" rows of
//                                             results/bs_range.json)
#pragma once

#include "common.hpp"

namespace bmfp {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// ---- element conversions --------------------------------------------------
template <class AT, class VT> __device__ __forceinline__ AT cvt(VT v) { return (AT) v; }
template <> __device__ __forceinline__ float cvt<float, __half>(__half v) { return __half2float(v); }

__device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// ---- wide loads -------------------------------------------------------------
constexpr int lowbit(int v) { return v & -v; }
constexpr int cmin(int a, int b) { return a < b ? a : b; }
constexpr int pow2ceil(int v) { int p = 1; while (p < v) p *= 2; return p; }

// Loads N consecutive VT from p, whose byte address is known to be a
// multiple of ALIGN, in the widest chunks (<= 16 bytes) that alignment and
// size allow; NT marks 8- and 16-byte chunks non-temporal (streamed data).
template <class VT, int N, int ALIGN, bool NT = false>
__device__ __forceinline__ void load_n(const VT *__restrict__ p, VT (&out)[N])
{
  constexpr int BYTES = N * (int) sizeof(VT);
  constexpr int CH = cmin(cmin(ALIGN, 16), lowbit(BYTES));
  const char *src = reinterpret_cast<const char *>(p);
  char *dst = reinterpret_cast<char *>(out);
  if constexpr (CH == 16) {
#pragma unroll
    for (int i = 0; i < BYTES / 16; ++i) {
      const u32x4 *q = reinterpret_cast<const u32x4 *>(src + 16 * i);
      u32x4 v = NT ? __builtin_nontemporal_load(q) : *q;
      __builtin_memcpy(dst + 16 * i, &v, 16);
    }
  } else if constexpr (CH == 8) {
#pragma unroll
    for (int i = 0; i < BYTES / 8; ++i) {
      const u32x2 *q = reinterpret_cast<const u32x2 *>(src + 8 * i);
      u32x2 v = NT ? __builtin_nontemporal_load(q) : *q;
      __builtin_memcpy(dst + 8 * i, &v, 8);
    }
  } else {
#pragma unroll
    for (int i = 0; i < N; ++i) out[i] = p[i];
  }
}

// Stores the N values of v at p in the widest chunks (up to 16 bytes) the
// row's size allows (p is aligned to the lowest set bit of N*sizeof(T)).
template <class T, int N> __device__ __forceinline__ void store_words(T *p, const T (&v)[N])
{
  constexpr int BYTES = N * (int) sizeof(T);
  constexpr int CH = lowbit(BYTES) < 16 ? lowbit(BYTES) : 16;
  char *dst = reinterpret_cast<char *>(p);
  const char *src = reinterpret_cast<const char *>(v);
  if constexpr (CH >= 4) {
    typedef unsigned int w_t __attribute__((ext_vector_type(CH / 4)));
#pragma unroll
    for (int c = 0; c < BYTES; c += CH) {
      w_t t;
      __builtin_memcpy(&t, src + c, CH);
      *reinterpret_cast<w_t *>(dst + c) = t;
    }
  } else {
#pragma unroll
    for (int c = 0; c < N; ++c) p[c] = v[c];
  }
}

// Maps the hardware workgroup id so that each XCD (workgroups are dealt to
// the 8 XCDs round robin) processes one contiguous run of tiles: neighbouring
// block rows then share x segments, column indices and row pointers in the
// same L2.  `per` = launched groups / 8 (launch a multiple of 8); 0 = off.
__device__ __forceinline__ int64_t tile_id(int per)
{
  const int64_t i = blockIdx.x;
  if (per == 0) return i;
  return (i & 7) * (int64_t) per + (i >> 3);
}
// Banded order: XCD j (workgroup ids j, j + 8, ...) takes bands j, j + 8, ...
// of `band` consecutive tiles, so the block rows of a band, and the x
// segments their neighbouring rows share, stay in one L2 without the eight
// XCDs streaming eight distant regions (launch a multiple of 8 * band
// groups).  band = 0: tile_id(per).
__device__ __forceinline__ int64_t tile_id(int per, int band)
{
  if (band == 0) return tile_id(per);
  const int64_t l = (int64_t) (blockIdx.x >> 3), xcd = blockIdx.x & 7;
  return (l / band * 8 + xcd) * band + l % band;
}

// FP64 SpMV kernels whose y stores go out non-temporal (y is written once and
// not read again by the kernel, so its lines need not take L2 space from x):
// measured in paired sweeps on the HBM-resident stencils, % of 8 TB/s, staged
// bs 1 72.1 -> 76.2 and bs 2 74.6 -> 76.9, dense lane groups bs 4 73.6 ->
// 74.6, bs 8 78.7 -> 79.3, bs 16 84.0 -> 86.1; bs 3, 6, 12, 24 and 32 were
// 0.3-2.2 points slower and keep ordinary stores.
template <class VT, int BS> constexpr bool y_nontemporal()
{
  return sizeof(VT) == 8 && (BS == 1 || BS == 2 || BS == 4 || BS == 8 || BS == 16);
}
template <bool NT, class T> __device__ __forceinline__ void store_y(T *p, T v)
{
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// ---- shared launch arguments --------------------------------------------------
template <class VT, class AT>
struct BcsrArgs {
  int n_brows;
  int xcd_per;
  int band = 0;  // tiles per XCD band (tile_id(per, band)); 0: none
  const int *__restrict__ rp;
  const int *__restrict__ cols;
  const VT *__restrict__ vals;
  const float *__restrict__ scale;   // per block row, F16 storage only
  const AT *__restrict__ x;
  AT *__restrict__ y;
};

// Chunk of columns processed per step in the thread-per-row kernels: keeps
// the register footprint bounded for large BS.
template <int BS> constexpr int col_chunk() { return BS % 8 == 0 ? 8 : BS; }

// Blocks whose loads a thread issues together before using them (thread-per-
// row kernels).  Small blocks carry few bytes per block, so a whole 9-point
// row is fetched at once to keep enough bytes in flight; 0 = plain loop.
template <int BS> constexpr int tpr_batch() { return BS <= 2 ? 9 : BS == 3 ? 3 : 0; }

// Row r of one block (BS values in the thread's layout).
template <class VT, int BS, int LAYOUT>
__device__ __forceinline__ void load_block_row(const VT *__restrict__ blk, int r, VT (&v)[BS])
{
  if constexpr (LAYOUT == BMFP_ROW_MAJOR) {
    load_n<VT, BS, lowbit((int) sizeof(VT) * BS)>(blk + r * BS, v);
  } else {
#pragma unroll
    for (int c = 0; c < BS; ++c) v[c] = blk[c * BS + r];
  }
}

// Sum over blocks [b0, b1) of row r of each block times its x segment.  Whole
// batches of U blocks issue all their loads (indices, values, x) before any
// is used; the remaining < U blocks go one at a time.  Unpredicated batches
// matter: per-block guards inside a batch serialise the loads.
template <class VT, class AT, int BS, int LAYOUT, int U>
__device__ __forceinline__ AT tpr_batched(const int *__restrict__ cols, const VT *__restrict__ vals,
                                          const AT *__restrict__ x, int b0, int b1, int r)
{
  constexpr int XALIGN = lowbit((int) sizeof(AT) * BS);
  AT acc = 0;
  int b = b0;
  for (; b + U <= b1; b += U) {
    int col[U];
#pragma unroll
    for (int k = 0; k < U; ++k) col[k] = cols[b + k];
    VT v[U][BS];
#pragma unroll
    for (int k = 0; k < U; ++k) load_block_row<VT, BS, LAYOUT>(vals + (int64_t) (b + k) * (BS * BS), r, v[k]);
    AT xv[U][BS];
#pragma unroll
    for (int k = 0; k < U; ++k) load_n<AT, BS, XALIGN>(x + (int64_t) col[k] * BS, xv[k]);
#pragma unroll
    for (int k = 0; k < U; ++k)
#pragma unroll
      for (int c = 0; c < BS; ++c) acc = fmadd(cvt<AT>(v[k][c]), xv[k][c], acc);
  }
  for (; b < b1; ++b) {
    VT v[BS];
    AT xv[BS];
    load_block_row<VT, BS, LAYOUT>(vals + (int64_t) b * (BS * BS), r, v);
    load_n<AT, BS, XALIGN>(x + (int64_t) cols[b] * BS, xv);
#pragma unroll
    for (int c = 0; c < BS; ++c) acc = fmadd(cvt<AT>(v[c]), xv[c], acc);
  }
  return acc;
}

// ============================================================================
// Thread per scalar row.  Row-major: the thread's row of each block is BS
// contiguous values, loaded in 16-byte pieces; the BS threads of a block row
// together read the whole block contiguously.  Column-major: for each column
// c the BS threads of a block row read one contiguous column.
// ============================================================================
template <class VT, class AT, int BS, int LAYOUT, bool SCALED>
__global__ __launch_bounds__(256) void k_bcsr_tpr(BcsrArgs<VT, AT> a)
{
  const int64_t t = tile_id(a.xcd_per) * 256 + threadIdx.x;
  const int64_t br64 = t / BS;
  if (br64 >= a.n_brows) return;
  const int br = (int) br64;
  const int r = (int) (t - br64 * BS);
  const int b0 = a.rp[br], b1 = a.rp[br + 1];
  AT acc = 0;
  if constexpr (tpr_batch<BS>() > 0) {
    acc = tpr_batched<VT, AT, BS, LAYOUT, tpr_batch<BS>()>(a.cols, a.vals, a.x, b0, b1, r);
  } else {
    constexpr int CC = col_chunk<BS>();
    constexpr int XALIGN = lowbit((int) sizeof(AT) * BS);
    constexpr int VALIGN = lowbit((int) sizeof(VT) * BS);
#pragma unroll 3
    for (int b = b0; b < b1; ++b) {
      const AT *xs = a.x + (int64_t) a.cols[b] * BS;
      const VT *blk = a.vals + (int64_t) b * (BS * BS);
#pragma unroll
      for (int c0 = 0; c0 < BS; c0 += CC) {
        AT xv[CC];
        load_n<AT, CC, cmin(XALIGN, lowbit((int) sizeof(AT) * CC))>(xs + c0, xv);
        if constexpr (LAYOUT == BMFP_ROW_MAJOR) {
          VT v[CC];
          load_n<VT, CC, cmin(VALIGN, lowbit((int) sizeof(VT) * CC))>(blk + r * BS + c0, v);
#pragma unroll
          for (int c = 0; c < CC; ++c) acc = fmadd(cvt<AT>(v[c]), xv[c], acc);
        } else {
#pragma unroll
          for (int c = 0; c < CC; ++c) acc = fmadd(cvt<AT>(blk[(c0 + c) * BS + r]), xv[c], acc);
        }
      }
    }
  }
  if constexpr (SCALED) acc *= (AT) a.scale[br];
  a.y[t] = acc;
}

// ============================================================================
// Lane group per scalar row.  G = pow2ceil(BS/VEC) lanes; lane g < BS/VEC
// owns columns [g*VEC, g*VEC+VEC) of every block in the row.  Partial sums
// are combined by a butterfly over the G lanes.
// ============================================================================
template <class VT, class AT, int BS, int VEC, int LAYOUT, bool SCALED>
__global__ __launch_bounds__(256) void k_bcsr_gpr(BcsrArgs<VT, AT> a)
{
  constexpr int GA = BS / VEC;
  constexpr int G = pow2ceil(GA);
  static_assert(BS % VEC == 0 && G <= 64, "bad group shape");
  const int64_t t = tile_id(a.xcd_per) * 256 + threadIdx.x;
  const int g = (int) (t % G);
  const int64_t row = t / G;
  const int64_t br64 = row / BS;
  if (br64 >= a.n_brows) return;  // whole groups leave together
  const int br = (int) br64;
  const int r = (int) (row - br64 * BS);
  const int b0 = a.rp[br], b1 = a.rp[br + 1];
  constexpr int XALIGN = cmin(lowbit((int) sizeof(AT) * BS), lowbit((int) sizeof(AT) * VEC));
  constexpr int VALIGN = cmin(lowbit((int) sizeof(VT) * BS), lowbit((int) sizeof(VT) * VEC));
  AT acc = 0;
  if (g < GA) {
#pragma unroll 3
    for (int b = b0; b < b1; ++b) {
      const AT *xs = a.x + (int64_t) a.cols[b] * BS + g * VEC;
      const VT *blk = a.vals + (int64_t) b * (BS * BS);
      AT xv[VEC];
      load_n<AT, VEC, XALIGN>(xs, xv);
      if constexpr (LAYOUT == BMFP_ROW_MAJOR) {
        VT v[VEC];
        load_n<VT, VEC, VALIGN>(blk + r * BS + g * VEC, v);
#pragma unroll
        for (int c = 0; c < VEC; ++c) acc = fmadd(cvt<AT>(v[c]), xv[c], acc);
      } else {
#pragma unroll
        for (int c = 0; c < VEC; ++c) acc = fmadd(cvt<AT>(blk[(g * VEC + c) * BS + r]), xv[c], acc);
      }
    }
  }
#pragma unroll
  for (int off = G / 2; off >= 1; off /= 2) acc += __shfl_xor(acc, off, G);
  if (g == 0) {
    if constexpr (SCALED) acc *= (AT) a.scale[br];
    a.y[row] = acc;
  }
}

// ============================================================================
// Dense lane group per scalar row (row-major blocks): G = BS/VEC lanes, so
// every lane of a group works (64/G rows per wavefront; the 64 mod G spare
// lanes idle).  Lane g owns columns [g*VEC, g*VEC+VEC) of every block of its
// row; loads of U blocks are issued together; partial sums are combined by a
// fixed shuffle tree over the group.  Shapes per block size (FP64) from
// timing on the HBM-resident stencils, dense_shape(): VEC = 2 with U = 9 (a
// whole 9-block stencil row in one batch) is the fastest everywhere but at
// bs 6 (U = 3) and bs 24 (VEC = 4, U = 1).  Non-temporal value loads (the
// values are streamed once; x and the indices keep the caches) raise the
// 9-block shapes at bs 8-32 (paired runs, % of 8 TB/s: bs 8 73.6 -> 76.0 /
// 78.1, bs 16 76.0 / 79.1 -> 81.3 / 83.5, bs 32 78.2 / 80.0 -> 84.9 / 87.0)
// and lower bs 4, 6 and 24, which keep cached loads.
// ============================================================================
struct DenseShape {
  int vec, u;
  bool nt = false;  // values loaded non-temporal
};
template <class VT, int BS> constexpr DenseShape dense_shape()
{
  if constexpr (sizeof(VT) != 8) return {0, 0};
  switch (BS) {
  case 4: return {2, 9};
  case 6: return {2, 3};
  case 8: return {2, 9, true};
  case 12: return {2, 9, true};
  case 16: return {2, 9, true};
  case 24: return {4, 1};
  case 32: return {2, 9, true};
  default: return {0, 0};
  }
}

// Sum over the G consecutive lanes of group lr (lane = lr*G + g), left in
// lane g = 0: lane g += lane g + s for s = 1, 2, 4, ... (all lanes take part;
// the spare lanes past the last whole group read only each other).
template <int G, class AT>
__device__ __forceinline__ AT group_tree_sum(AT acc, int lane, int lr, int g)
{
  const int end = lr * G + G < 64 ? lr * G + G : 64;
#pragma unroll
  for (int s2 = 1; s2 < G; s2 *= 2) {
    const int src = lane + s2 < end ? lane + s2 : lane;
    const AT o = __shfl(acc, src);
    if (g % (2 * s2) == 0 && g + s2 < G) acc += o;
  }
  return acc;
}

template <class VT, class AT, int BS, int VEC, int U, bool SCALED, bool NT = false>
__global__ __launch_bounds__(256) void k_bcsr_dgr(BcsrArgs<VT, AT> a)
{
  constexpr int G = BS / VEC;
  constexpr int RPW = 64 / G;
  static_assert(BS % VEC == 0 && G <= 64, "bad group shape");
  const int lane = threadIdx.x & 63;
  const int64_t wave = (tile_id(a.xcd_per, a.band) * 256 + threadIdx.x) >> 6;
  const int lr = lane / G, g = lane - lr * G;
  const int64_t row = wave * RPW + lr;
  const int64_t rows = (int64_t) a.n_brows * BS;
  const bool on = lane < RPW * G && row < rows;
  const int64_t br = on ? row / BS : 0;
  const int r = (int) (row - br * BS);
  int b = 0, b1 = 0;
  if (on) {
    b = a.rp[br];
    b1 = a.rp[br + 1];
  }
  constexpr int ALIGN = cmin(lowbit((int) sizeof(VT) * BS), lowbit((int) sizeof(VT) * VEC));
  constexpr int XALIGN = cmin(lowbit((int) sizeof(AT) * BS), lowbit((int) sizeof(AT) * VEC));
  AT acc = 0;
  for (; b + U <= b1; b += U) {
    int col[U];
    VT v[U][VEC];
    AT xv[U][VEC];
#pragma unroll
    for (int k = 0; k < U; ++k) col[k] = a.cols[b + k];
#pragma unroll
    for (int k = 0; k < U; ++k) load_n<VT, VEC, ALIGN, NT>(a.vals + ((int64_t) (b + k) * BS + r) * BS + g * VEC, v[k]);
#pragma unroll
    for (int k = 0; k < U; ++k) load_n<AT, VEC, XALIGN>(a.x + (int64_t) col[k] * BS + g * VEC, xv[k]);
#pragma unroll
    for (int k = 0; k < U; ++k)
#pragma unroll
      for (int c = 0; c < VEC; ++c) acc = fmadd(cvt<AT>(v[k][c]), xv[k][c], acc);
  }
  for (; U > 1 && b < b1; ++b) {
    VT v[VEC];
    AT xv[VEC];
    load_n<VT, VEC, ALIGN, NT>(a.vals + ((int64_t) b * BS + r) * BS + g * VEC, v);
    load_n<AT, VEC, XALIGN>(a.x + (int64_t) a.cols[b] * BS + g * VEC, xv);
#pragma unroll
    for (int c = 0; c < VEC; ++c) acc = fmadd(cvt<AT>(v[c]), xv[c], acc);
  }
  acc = group_tree_sum<G>(acc, lane, lr, g);
  if (on && g == 0) {
    if constexpr (SCALED) acc *= (AT) a.scale[br];
    store_y<y_nontemporal<VT, BS>()>(&a.y[row], acc);
  }
}

// ============================================================================
// Dense lane group per block row for column-major blocks, the counterpart of
// k_bcsr_dgr: column c of a block is BS contiguous values, and lane l = c*H +
// h of a block row's group of L = BS*H lanes (H = BS/VEC) owns its rows
// [h*VEC, h*VEC + VEC), so the group reads every block as one contiguous range
// and each lane needs a single x element per block.  Loads of U blocks are
// issued together; each lane keeps VEC row partials over the block row; the
// partials of a row set are summed over the BS columns by a fixed shuffle tree
// (stride H), and lane h (c = 0) stores its VEC rows.  64/L block rows per
// wavefront.
// ============================================================================
template <class VT, class AT, int BS, int VEC, int U, bool SCALED, bool NT = false>
__global__ __launch_bounds__(256) void k_bcsr_dgc(BcsrArgs<VT, AT> a)
{
  constexpr int H = BS / VEC, L = BS * H, BPW = 64 / L;
  static_assert(BS % VEC == 0 && L <= 64, "bad group shape");
  constexpr int ALIGN = cmin(lowbit((int) sizeof(VT) * BS), lowbit((int) sizeof(VT) * VEC));
  const int lane = threadIdx.x & 63;
  const int64_t wave = (tile_id(a.xcd_per, a.band) * 256 + threadIdx.x) >> 6;
  const int q = lane / L, l = lane - q * L;
  const int c = l / H, h = l - c * H;
  const int64_t br = wave * BPW + q;
  const bool on = lane < BPW * L && br < a.n_brows;
  int b = 0, b1 = 0;
  if (on) {
    b = a.rp[br];
    b1 = a.rp[br + 1];
  }
  const int off = c * BS + h * VEC;  // the lane's first element within a block
  AT acc[VEC];
#pragma unroll
  for (int i = 0; i < VEC; ++i) acc[i] = 0;
  for (; b + U <= b1; b += U) {
    int col[U];
    VT v[U][VEC];
    AT xv[U];
#pragma unroll
    for (int k = 0; k < U; ++k) col[k] = a.cols[b + k];
#pragma unroll
    for (int k = 0; k < U; ++k) load_n<VT, VEC, ALIGN, NT>(a.vals + (int64_t) (b + k) * (BS * BS) + off, v[k]);
#pragma unroll
    for (int k = 0; k < U; ++k) xv[k] = a.x[(int64_t) col[k] * BS + c];
#pragma unroll
    for (int k = 0; k < U; ++k)
#pragma unroll
      for (int i = 0; i < VEC; ++i) acc[i] = fmadd(cvt<AT>(v[k][i]), xv[k], acc[i]);
  }
  for (; U > 1 && b < b1; ++b) {
    VT v[VEC];
    load_n<VT, VEC, ALIGN, NT>(a.vals + (int64_t) b * (BS * BS) + off, v);
    const AT xe = a.x[(int64_t) a.cols[b] * BS + c];
#pragma unroll
    for (int i = 0; i < VEC; ++i) acc[i] = fmadd(cvt<AT>(v[i]), xe, acc[i]);
  }
  // column c += column c + s, s = 1, 2, 4, ... (fixed order; lanes of other
  // groups and spare lanes read only themselves)
#pragma unroll
  for (int s2 = 1; s2 < BS; s2 *= 2) {
    const bool take = c % (2 * s2) == 0 && c + s2 < BS;
    const int src = (c + s2 < BS && lane + s2 * H < 64) ? lane + s2 * H : lane;
#pragma unroll
    for (int i = 0; i < VEC; ++i) {
      const AT o = __shfl(acc[i], src);
      if (take) acc[i] += o;
    }
  }
  if (on && c == 0) {
    AT sc = 1;
    if constexpr (SCALED) sc = (AT) a.scale[br];
#pragma unroll
    for (int i = 0; i < VEC; ++i) store_y<y_nontemporal<VT, BS>()>(&a.y[br * BS + h * VEC + i], SCALED ? acc[i] * sc : acc[i]);
  }
}

// ============================================================================
// Lane group per block row with the blocks split across the group (bs 2, 3):
// the BS lanes of a block row each take whole blocks b0+h, b0+h+BS, ... in
// predicated batches of P (all index, value and x loads of a batch issued
// together) and accumulate all BS output rows; lane r then sums the group's
// partials of row r in BS shuffle rounds (round k: lane h offers its partial
// of row (h - k) mod BS, lane r reads lane (r + k) mod BS), a fixed order.
// For a 9-block row a lane issues about 9/BS x (block + index + x segment)
// requests instead of 9 x (block row + index + x segment).
// ============================================================================
template <int BS> constexpr int split_batch() { return BS == 1 ? 9 : BS == 2 ? 5 : 3; }

template <class VT, class AT, int BS, int LAYOUT, bool SCALED>
__global__ __launch_bounds__(256) void k_bcsr_split(BcsrArgs<VT, AT> a)
{
  constexpr int P = split_batch<BS>(), RPW = 64 / BS;
  constexpr int VALIGN = lowbit((int) sizeof(VT) * BS * BS), XALIGN = lowbit((int) sizeof(AT) * BS);
  const int lane = threadIdx.x & 63;
  const int64_t wave = (tile_id(a.xcd_per) * 256 + threadIdx.x) >> 6;
  const int q = lane / BS, h = lane - q * BS, base = q * BS;
  const int64_t br64 = wave * RPW + q;
  const bool on = lane < RPW * BS && br64 < a.n_brows;
  const int br = on ? (int) br64 : 0;
  int b = 0, b1 = 0;
  if (on) {
    b = a.rp[br] + h;
    b1 = a.rp[br + 1];
  }
  AT part[BS];
#pragma unroll
  for (int r = 0; r < BS; ++r) part[r] = 0;
  for (; b < b1; b += BS * P) {
    int c[P];
    VT v[P][BS * BS];
    AT xv[P][BS];
#pragma unroll
    for (int u = 0; u < P; ++u)
      if (b + BS * u < b1) c[u] = a.cols[b + BS * u];
#pragma unroll
    for (int u = 0; u < P; ++u)
      if (b + BS * u < b1) load_n<VT, BS * BS, VALIGN>(a.vals + (int64_t) (b + BS * u) * (BS * BS), v[u]);
#pragma unroll
    for (int u = 0; u < P; ++u)
      if (b + BS * u < b1) load_n<AT, BS, XALIGN>(a.x + (int64_t) c[u] * BS, xv[u]);
#pragma unroll
    for (int u = 0; u < P; ++u)
      if (b + BS * u < b1) {
#pragma unroll
        for (int r = 0; r < BS; ++r)
#pragma unroll
          for (int cc = 0; cc < BS; ++cc)
            part[r] = fmadd(cvt<AT>(v[u][LAYOUT == BMFP_ROW_MAJOR ? r * BS + cc : cc * BS + r]), xv[u][cc], part[r]);
      }
  }
  AT acc = 0;
#pragma unroll
  for (int k = 0; k < BS; ++k) {
    const int si = (h - k + BS) % BS;
    AT send = part[0];
#pragma unroll
    for (int r = 1; r < BS; ++r) send = si == r ? part[r] : send;
    acc += __shfl(send, lane < RPW * BS ? base + (h + k) % BS : lane);
  }
  if (on) {
    if constexpr (SCALED) acc *= (AT) a.scale[br];
    a.y[(int64_t) br * BS + h] = acc;
  }
}

// ============================================================================
// k_bcsr_split with the wavefront's blocks staged in LDS first.  The values
// and column indices of the wavefront's RPW block rows are one contiguous
// range; it is copied to LDS with 16-byte loads at consecutive addresses
// across the lanes (one load instruction moves 1 KB; StagedRange), then each lane reads
// its whole blocks from LDS and gathers x from global memory as in
// k_bcsr_split.  A wavefront whose range exceeds CAP blocks (rows longer than
// 9 blocks on average) reads global memory directly.  No workgroup barrier:
// each wavefront stages and reads only its own LDS slice.  For bs = 1 this is
// thread per row (RPW = 64, P = 9: a whole 9-point row in one batch).
// ============================================================================
template <int BS> constexpr int stage_cap() { return 9 * (64 / BS); }

// LDS operations of one wavefront complete in issue order, so a compiler
// barrier is all that lanes need to read what other lanes of the same
// wavefront wrote.
__device__ __forceinline__ void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Copy of bytes [s, e) of a buffer whose base is 16-byte aligned into LDS,
// in two halves so that several ranges' loads can be in flight together:
// load() issues 16-byte loads at consecutive addresses across the lanes from
// s rounded down to 16 bytes (up to 12 bytes before s come along), the
// partial last chunk 4 bytes at a time (s and e are multiples of 4,
// e - s <= CAPB); store() writes them to dst; shift is s's offset in dst.
// LDS bytes that hold a staged range of up to capb bytes: the start rounded
// down to 16 bytes adds up to 12, and the copy writes whole 16-byte chunks.
constexpr int staged_bytes(int capb) { return (capb + 12 + 15) / 16 * 16; }

template <int CAPB>
struct StagedRange {
  static constexpr int NCH = (staged_bytes(CAPB) / 16 + 63) / 64;
  u32x4 t[NCH];
  int n = 0, shift = 0;
  __device__ __forceinline__ void load(const void *base_v, int64_t s, int64_t e, int lane)
  {
    const char *base = static_cast<const char *>(base_v);
    const int64_t s_al = s & ~(int64_t) 15;
    n = (int) (e - s_al);
    shift = (int) (s - s_al);
#pragma unroll
    for (int m = 0; m < NCH; ++m) {
      const int k = 16 * (lane + 64 * m);
      if (k < n) {
        if (k + 16 <= n) {
          t[m] = *reinterpret_cast<const u32x4 *>(base + s_al + k);
        } else {
          const unsigned *w = reinterpret_cast<const unsigned *>(base + s_al + k);
          t[m] = u32x4{w[0], n - k > 4 ? w[1] : 0u, n - k > 8 ? w[2] : 0u, 0u};
        }
      }
    }
  }
  __device__ __forceinline__ void store(char *dst, int lane) const
  {
#pragma unroll
    for (int m = 0; m < NCH; ++m) {
      const int k = 16 * (lane + 64 * m);
      if (k < n) *reinterpret_cast<u32x4 *>(dst + k) = t[m];
    }
  }
};

// Accumulates blocks b, b + BS, ... < b1 of one block row into part[] in
// predicated batches of P: col_at(k) and block_at(k, v) fetch block k's column
// and values.  X_FIRST issues the x gathers before the block loads (blocks in
// LDS: the gathers are the long-latency loads); otherwise the global block
// loads go out first, before the gathers wait on their column indices.
template <class VT, class AT, int BS, int LAYOUT, int P, bool X_FIRST, class ColF, class BlkF>
__device__ __forceinline__ void split_accumulate(int b, int b1, const AT *__restrict__ x, AT (&part)[BS],
                                                 ColF col_at, BlkF block_at)
{
  constexpr int XALIGN = lowbit((int) sizeof(AT) * BS);
  for (; b < b1; b += BS * P) {
    int c[P];
    VT v[P][BS * BS];
    AT xv[P][BS];
#pragma unroll
    for (int u = 0; u < P; ++u)
      if (b + BS * u < b1) c[u] = col_at(b + BS * u);
    if constexpr (!X_FIRST) {
#pragma unroll
      for (int u = 0; u < P; ++u)
        if (b + BS * u < b1) block_at(b + BS * u, v[u]);
    }
#pragma unroll
    for (int u = 0; u < P; ++u)
      if (b + BS * u < b1) load_n<AT, BS, XALIGN>(x + (int64_t) c[u] * BS, xv[u]);
    if constexpr (X_FIRST) {
#pragma unroll
      for (int u = 0; u < P; ++u)
        if (b + BS * u < b1) block_at(b + BS * u, v[u]);
    }
#pragma unroll
    for (int u = 0; u < P; ++u)
      if (b + BS * u < b1) {
#pragma unroll
        for (int r = 0; r < BS; ++r)
#pragma unroll
          for (int cc = 0; cc < BS; ++cc)
            part[r] = fmadd(cvt<AT>(v[u][LAYOUT == BMFP_ROW_MAJOR ? r * BS + cc : cc * BS + r]), xv[u][cc], part[r]);
      }
  }
}

template <class VT, class AT, int BS, int LAYOUT, bool SCALED, int WPG>
__global__ __launch_bounds__(64 * WPG) void k_bcsr_stage(BcsrArgs<VT, AT> a)
{
  constexpr int P = split_batch<BS>(), RPW = 64 / BS, CAP = stage_cap<BS>();
  constexpr int BB = BS * BS * (int) sizeof(VT);  // bytes per block
  static_assert(BB % 4 == 0, "staging copies whole 4-byte words");
  constexpr int VALIGN = lowbit(BB);
  __shared__ __align__(16) char vals_s[WPG][staged_bytes(CAP * BB)];
  __shared__ __align__(16) char cols_s[WPG][staged_bytes(CAP * 4)];
  const int lane = threadIdx.x & 63, wib = threadIdx.x >> 6;
  const int64_t r0_64 = (tile_id(a.xcd_per, a.band) * WPG + wib) * RPW;
  if (r0_64 >= a.n_brows) return;  // the whole wavefront leaves; no barrier follows
  const int r0 = (int) r0_64, r1 = r0 + RPW < a.n_brows ? r0 + RPW : a.n_brows;
  const int B0 = a.rp[r0], B1 = a.rp[r1];
  const int q = lane / BS, h = lane - q * BS, base = q * BS;
  const bool on = lane < RPW * BS && r0 + q < r1;
  const int br = on ? r0 + q : r0;
  int b = 0, b1 = 0;
  if (on) {
    b = a.rp[br] + h;
    b1 = a.rp[br + 1];
  }
  AT part[BS];
#pragma unroll
  for (int r = 0; r < BS; ++r) part[r] = 0;
  if (B1 - B0 <= CAP) {
    StagedRange<CAP * BB> sv;
    StagedRange<CAP * 4> sc;
    sv.load(a.vals, (int64_t) B0 * BB, (int64_t) B1 * BB, lane);
    sc.load(a.cols, (int64_t) B0 * 4, (int64_t) B1 * 4, lane);
    sv.store(vals_s[wib], lane);
    sc.store(cols_s[wib], lane);
    wave_lds_sync();
    const char *vs = vals_s[wib] + sv.shift;
    const int *cs = reinterpret_cast<const int *>(cols_s[wib] + sc.shift);
    split_accumulate<VT, AT, BS, LAYOUT, P, true>(
        b, b1, a.x, part, [&](int k) { return cs[k - B0]; },
        [&](int k, VT(&v)[BS * BS]) {
          load_n<VT, BS * BS, VALIGN>(reinterpret_cast<const VT *>(vs + (k - B0) * BB), v);
        });
  } else {
    split_accumulate<VT, AT, BS, LAYOUT, P, false>(
        b, b1, a.x, part, [&](int k) { return a.cols[k]; },
        [&](int k, VT(&v)[BS * BS]) { load_n<VT, BS * BS, VALIGN>(a.vals + (int64_t) k * (BS * BS), v); });
  }
  AT acc = 0;
#pragma unroll
  for (int k = 0; k < BS; ++k) {
    const int si = (h - k + BS) % BS;
    AT send = part[0];
#pragma unroll
    for (int r = 1; r < BS; ++r) send = si == r ? part[r] : send;
    acc += __shfl(send, lane < RPW * BS ? base + (h + k) % BS : lane);
  }
  if (on) {
    if constexpr (SCALED) acc *= (AT) a.scale[br];
    store_y<y_nontemporal<VT, BS>()>(&a.y[(int64_t) br * BS + h], acc);
  }
}

// ============================================================================
// Thread per scalar row with the x segment shared by shuffles: the BS lanes
// of a block row each load one element of x[col*BS .. col*BS+BS) (one
// coalesced load instead of BS broadcast loads per lane) and read the others
// from their neighbours' registers.  RPW = 64/BS block rows per wavefront;
// the 64 - RPW*BS spare lanes idle.
// ============================================================================
template <class VT, class AT, int BS, int LAYOUT, bool SCALED>
__global__ __launch_bounds__(256) void k_bcsr_shfx(BcsrArgs<VT, AT> a)
{
  constexpr int RPW = 64 / BS;
  const int lane = threadIdx.x & 63;
  const int64_t wave = (tile_id(a.xcd_per) * 256 + threadIdx.x) >> 6;
  const int lbr = lane / BS;
  const int r = lane - lbr * BS;
  const int64_t br64 = wave * RPW + lbr;
  if (lane >= RPW * BS || br64 >= a.n_brows) return;  // whole block rows leave together
  const int br = (int) br64;
  const int b0 = a.rp[br], b1 = a.rp[br + 1];
  const int base = lbr * BS;
  constexpr int VALIGN = lowbit((int) sizeof(VT) * BS);
  constexpr int CC = col_chunk<BS>();
  AT acc = 0;
  for (int b = b0; b < b1; ++b) {
    const AT xr = a.x[(int64_t) a.cols[b] * BS + r];
    const VT *blk = a.vals + (int64_t) b * (BS * BS);
#pragma unroll
    for (int c0 = 0; c0 < BS; c0 += CC) {
      if constexpr (LAYOUT == BMFP_ROW_MAJOR) {
        VT v[CC];
        load_n<VT, CC, cmin(VALIGN, lowbit((int) sizeof(VT) * CC))>(blk + r * BS + c0, v);
#pragma unroll
        for (int c = 0; c < CC; ++c) acc = fmadd(cvt<AT>(v[c]), __shfl(xr, base + c0 + c), acc);
      } else {
#pragma unroll
        for (int c = 0; c < CC; ++c)
          acc = fmadd(cvt<AT>(blk[(c0 + c) * BS + r]), __shfl(xr, base + c0 + c), acc);
      }
    }
  }
  if constexpr (SCALED) acc *= (AT) a.scale[br];
  a.y[(int64_t) br * BS + r] = acc;
}

// ============================================================================
// Lane group per block row, column-major blocks.  The block row's values are
// one contiguous run of nblk*BS*BS elements; GL lanes stream it in strides of
// L = (GL/BS)*BS elements (one coalesced load per lane per step).  Because L
// is a multiple of BS, lane j always meets row r = j % BS; the M = GL/BS
// lanes of each row are then summed by a fixed-order shuffle tree.  This is
// the reference's "column-by-column" kernel with the lane accounting made
// exact for every BS (the reference folds duplicate columns in for BS that
// do not divide 32).
// ============================================================================
template <class VT, class AT, int BS, int GL, bool SCALED>
__global__ __launch_bounds__(256) void k_bcsr_bgr(BcsrArgs<VT, AT> a)
{
  static_assert(GL == 16 || GL == 32 || GL == 64, "GL");
  static_assert(BS <= GL, "BS > GL");
  constexpr int M = GL / BS;
  constexpr int L = M * BS;
  constexpr int BS2 = BS * BS;
  const int64_t t = tile_id(a.xcd_per) * 256 + threadIdx.x;
  const int j = (int) (t % GL);
  const int64_t br64 = t / GL;
  if (br64 >= a.n_brows) return;  // whole groups leave together
  const int br = (int) br64;
  const int b0 = a.rp[br], b1 = a.rp[br + 1];
  const int nelem = (b1 - b0) * BS2;
  const VT *vals = a.vals + (int64_t) b0 * BS2;
  const int *cols = a.cols + b0;
  AT acc = 0;
  if (j < L) {
#pragma unroll 4
    for (int e = j; e < nelem; e += L) {
      const int k = e / BS2;
      const int c = (e - k * BS2) / BS;
      acc = fmadd(cvt<AT>(vals[e]), a.x[(int64_t) cols[k] * BS + c], acc);
    }
  }
  // lanes r, r+BS, ..., r+(M-1)*BS hold row r's partial sums
  const int q = j / BS;
#pragma unroll
  for (int s = 1; s < M; s *= 2) {
    const AT o = __shfl_down(acc, s * BS, GL);
    if ((q % (2 * s)) == 0 && q + s < M) acc += o;
  }
  if (j < BS) {
    if constexpr (SCALED) acc *= (AT) a.scale[br];
    a.y[(int64_t) br * BS + j] = acc;
  }
}

// ============================================================================
// Runtime block size (any bs), thread per scalar row; used for block sizes
// without a precompiled instantiation.
// ============================================================================
template <class VT, class AT, int LAYOUT, bool SCALED>
__global__ __launch_bounds__(256) void k_bcsr_tpr_rt(BcsrArgs<VT, AT> a, int bs)
{
  const int64_t t = tile_id(a.xcd_per) * 256 + threadIdx.x;
  const int64_t br64 = t / bs;
  if (br64 >= a.n_brows) return;
  const int br = (int) br64;
  const int r = (int) (t - br64 * bs);
  const int b0 = a.rp[br], b1 = a.rp[br + 1];
  const int64_t bs2 = (int64_t) bs * bs;
  AT acc = 0;
  for (int b = b0; b < b1; ++b) {
    const AT *xs = a.x + (int64_t) a.cols[b] * bs;
    const VT *blk = a.vals + (int64_t) b * bs2;
    for (int c = 0; c < bs; ++c) {
      const VT v = LAYOUT == BMFP_ROW_MAJOR ? blk[r * bs + c] : blk[c * bs + r];
      acc = fmadd(cvt<AT>(v), xs[c], acc);
    }
  }
  if constexpr (SCALED) acc *= (AT) a.scale[br];
  a.y[t] = acc;
}

// ============================================================================
// CSR.
// ============================================================================
template <class VT, class AT>
struct CsrArgs {
  int n_rows;
  int xcd_per;
  const int *__restrict__ rp;
  const int *__restrict__ cols;
  const VT *__restrict__ vals;
  const AT *__restrict__ x;
  AT *__restrict__ y;
};

// Sum over e = e0, e0 + stride, ... < e1 of vals[e] * x[cols[e]], in that
// order.  Batches of U entries issue their index and value loads together,
// then their x gathers, so a lane waits on one gather latency per batch
// instead of one per entry.  TAIL_ONLY: full batches run unpredicated (guards
// inside a batch serialise its loads) and only the last, partial batch is
// predicated -- best for a thread walking a whole row; otherwise every batch
// is predicated -- best for lane groups, whose lanes see a few entries each.
template <class VT, class AT, int U, bool TAIL_ONLY>
__device__ __forceinline__ AT csr_dot(const CsrArgs<VT, AT> &a, int e, int e1, int stride)
{
  AT acc = 0;
  int c[U];
  VT v[U];
  AT xv[U];
  if constexpr (TAIL_ONLY) {
    for (; e + (U - 1) * stride < e1; e += U * stride) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        c[u] = a.cols[e + u * stride];
        v[u] = a.vals[e + u * stride];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) xv[u] = a.x[c[u]];
#pragma unroll
      for (int u = 0; u < U; ++u) acc = fmadd(cvt<AT>(v[u]), xv[u], acc);
    }
  }
  for (; e < e1; e += U * stride) {
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (e + u * stride < e1) {
        c[u] = a.cols[e + u * stride];
        v[u] = a.vals[e + u * stride];
      }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (e + u * stride < e1) xv[u] = a.x[c[u]];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (e + u * stride < e1) acc = fmadd(cvt<AT>(v[u]), xv[u], acc);
  }
  return acc;
}

template <class VT, class AT>
__global__ __launch_bounds__(256) void k_csr_scalar(CsrArgs<VT, AT> a)
{
  const int64_t row = tile_id(a.xcd_per) * 256 + threadIdx.x;
  if (row >= a.n_rows) return;
  a.y[row] = csr_dot<VT, AT, 8, true>(a, a.rp[row], a.rp[row + 1], 1);
}

template <class VT, class AT, int G>
__global__ __launch_bounds__(256) void k_csr_vector(CsrArgs<VT, AT> a)
{
  const int64_t t = tile_id(a.xcd_per) * 256 + threadIdx.x;
  const int g = (int) (t % G);
  const int64_t row = t / G;
  if (row >= a.n_rows) return;  // whole groups leave together
  AT acc = csr_dot<VT, AT, 4, false>(a, a.rp[row] + g, a.rp[row + 1], G);
#pragma unroll
  for (int off = G / 2; off >= 1; off /= 2) acc += __shfl_xor(acc, off, G);
  if (g == 0) a.y[row] = acc;
}

// ============================================================================
// CSR rows staged per wavefront (the CSR counterpart of k_bcsr_stage): RPW =
// 64/G consecutive rows, G lanes per row.  The wavefront's nonzeros are one
// contiguous range of values and of column indices; both are copied to the
// wavefront's LDS slice with 16-byte loads at consecutive addresses, then
// lane g of a row sums its entries g, g + G, ... from LDS (U x gathers from
// global memory in flight) and the group adds its lanes' sums by xor
// shuffles (a fixed order).  Ranges over CSR_STAGE_CAP nonzeros read global
// memory directly (csr_dot).  No workgroup barrier.
// ============================================================================
constexpr int CSR_STAGE_CAP = 640;

template <class VT, class AT, int G, int WPG>
__global__ __launch_bounds__(64 * WPG) void k_csr_stage(CsrArgs<VT, AT> a)
{
  // wavefronts hold ~576 nonzeros, ~9 per lane: a lane's entries in one batch
  constexpr int RPW = 64 / G, CAP = CSR_STAGE_CAP, U = 9;
  constexpr int VCAP = CAP * (int) sizeof(VT), CCAP = CAP * 4;
  __shared__ __align__(16) char vals_s[WPG][staged_bytes(VCAP)];
  __shared__ __align__(16) char cols_s[WPG][staged_bytes(CCAP)];
  const int lane = threadIdx.x & 63, wib = threadIdx.x >> 6;
  const int64_t r0_64 = (tile_id(a.xcd_per) * WPG + wib) * RPW;
  if (r0_64 >= a.n_rows) return;  // the whole wavefront leaves; no barrier follows
  const int r0 = (int) r0_64, r1 = r0 + RPW < a.n_rows ? r0 + RPW : a.n_rows;
  const int J0 = a.rp[r0], J1 = a.rp[r1];
  const int q = lane / G, g = lane - q * G, row = r0 + q;
  const bool on = row < r1;
  int e = 0, e1 = 0;
  if (on) {
    e = a.rp[row] + g;
    e1 = a.rp[row + 1];
  }
  AT acc = 0;
  if (J1 - J0 <= CAP) {
    StagedRange<VCAP> sv;
    StagedRange<CCAP> sc;
    sv.load(a.vals, (int64_t) J0 * (int) sizeof(VT), (int64_t) J1 * (int) sizeof(VT), lane);
    sc.load(a.cols, (int64_t) J0 * 4, (int64_t) J1 * 4, lane);
    sv.store(vals_s[wib], lane);
    sc.store(cols_s[wib], lane);
    wave_lds_sync();
    const VT *vs = reinterpret_cast<const VT *>(vals_s[wib] + sv.shift);
    const int *cs = reinterpret_cast<const int *>(cols_s[wib] + sc.shift);
    for (; e < e1; e += U * G) {
      int c[U];
      AT xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (e + u * G < e1) c[u] = cs[e + u * G - J0];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (e + u * G < e1) xv[u] = a.x[c[u]];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (e + u * G < e1) acc = fmadd(cvt<AT>(vs[e + u * G - J0]), xv[u], acc);
    }
  } else {
    acc = csr_dot<VT, AT, U, G == 1>(a, e, e1, G);
  }
#pragma unroll
  for (int off = G / 2; off >= 1; off /= 2) acc += __shfl_xor(acc, off, G);
  if (on && g == 0) a.y[row] = acc;
}

// ============================================================================
// Merge-path CSR SpMV: load-balanced for any row-length distribution.
// The merge of the row ends (row i ends before nonzero j iff rp[i+1] <= j)
// with the nonzeros 0..nnz-1 is a path of n + nnz items, cut into tiles of
// MERGE_IPT items per lane, 64 lanes per wavefront.  k_csr_merge_bounds
// finds where every tile starts (one thread per tile boundary, a binary
// search over row_ptr); each wavefront of k_csr_merge stages its tile's
// products vals[j] * x[cols[j]] and row ends in LDS with coalesced loads, and
// each lane walks its MERGE_IPT items from there.  A row that crosses lanes
// is summed by a segmented scan over the lanes (keyed by the row each lane
// ends in); a row that crosses wavefronts leaves a per-wavefront carry that
// k_csr_merge_fixup adds in wavefront order.  Every row is written once by
// the lane that reaches its end, so y needs no clearing, and the order of
// additions is fixed: results are bitwise reproducible.
// ============================================================================
constexpr int MERGE_IPT = 8;
constexpr int MERGE_TILE = 64 * MERGE_IPT;

// Rows consumed at merge-path diagonal d: the first x in [max(0, d-nnz),
// min(d, n)] whose row end comes after nonzero d-x-1.
__device__ __forceinline__ int merge_search(int64_t d, int n, int nnz, const int *__restrict__ rp)
{
  int64_t lo = d - nnz > 0 ? d - nnz : 0, hi = d < n ? d : n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t) rp[mid + 1] <= d - mid - 1) lo = mid + 1;
    else hi = mid;
  }
  return (int) lo;
}

// Row coordinate of every tile boundary t = 0..n_waves (diagonal t * MERGE_TILE,
// clamped to the end of the path).
template <int TILE>
__global__ __launch_bounds__(256) void k_csr_merge_bounds(int n, int nnz, int64_t n_waves, const int *__restrict__ rp,
                                                          int *__restrict__ bounds)
{
  const int64_t t = (int64_t) blockIdx.x * 256 + threadIdx.x;
  if (t > n_waves) return;
  const int64_t total = (int64_t) n + nnz;
  bounds[t] = merge_search(t * TILE < total ? t * TILE : total, n, nnz, rp);
}

template <class VT, class AT>
__global__ __launch_bounds__(256) void k_csr_merge(CsrArgs<VT, AT> a, int nnz, int64_t n_waves,
                                                   const int *__restrict__ bounds, int *__restrict__ carry_row,
                                                   AT *__restrict__ carry_val)
{
  // products padded by one slot per 8: lanes walk runs of ~8 consecutive
  // products, which would otherwise fall on the same LDS banks
  __shared__ AT prod_s[4][MERGE_TILE + MERGE_TILE / 8];
  __shared__ int ends_s[4][MERGE_TILE + 1];
  const int lane = threadIdx.x & 63, wib = threadIdx.x >> 6;
  const int n = a.n_rows;
  const int64_t total = (int64_t) n + nnz;
  const int64_t w = (int64_t) blockIdx.x * 4 + wib;
  // wavefronts past the path take part in the barrier with an empty tile
  const int64_t d0 = w * MERGE_TILE < total ? w * MERGE_TILE : total;
  const int64_t d1 = d0 + MERGE_TILE < total ? d0 + MERGE_TILE : total;
  const int i0 = w < n_waves ? bounds[w] : n, i1 = w < n_waves ? bounds[w + 1] : n;
  const int j0 = (int) (d0 - i0), j1 = (int) (d1 - i1);
  {
    // all index and value loads first, then all x gathers, then the products:
    // no load waits on another except each gather on its own index
    const int cnt = j1 - j0, nends = i1 - i0 + 1;
    int ci[MERGE_IPT];
    VT v[MERGE_IPT];
    AT xv[MERGE_IPT];
#pragma unroll
    for (int m = 0; m < MERGE_IPT; ++m) {
      const int k = lane + 64 * m;
      if (k < cnt) {
        ci[m] = a.cols[j0 + k];
        v[m] = a.vals[j0 + k];
      }
    }
#pragma unroll
    for (int m = 0; m < MERGE_IPT + 1; ++m) {
      const int k = lane + 64 * m;
      if (k < nends) ends_s[wib][k] = i0 + k < n ? a.rp[i0 + k + 1] : nnz;
    }
#pragma unroll
    for (int m = 0; m < MERGE_IPT; ++m)
      if (lane + 64 * m < cnt) xv[m] = a.x[ci[m]];
#pragma unroll
    for (int m = 0; m < MERGE_IPT; ++m) {
      const int k = lane + 64 * m;
      if (k < cnt) prod_s[wib][k + (k >> 3)] = cvt<AT>(v[m]) * xv[m];
    }
  }
  __syncthreads();
  if (w >= n_waves) return;
  // this lane's items [dl, de) start at (i, j): search the staged row ends
  const int64_t dl = d0 + lane * MERGE_IPT < d1 ? d0 + lane * MERGE_IPT : d1;
  const int64_t de = dl + MERGE_IPT < d1 ? dl + MERGE_IPT : d1;
  int lo = (int) (dl - j1 > i0 ? dl - j1 : i0), hi = (int) (dl - j0 < i1 ? dl - j0 : i1);
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if ((int64_t) ends_s[wib][mid - i0] <= dl - mid - 1) lo = mid + 1;
    else hi = mid;
  }
  int i = lo, j = (int) (dl - lo);
  AT acc = 0, first_part = 0;
  int first_row = -1;
  for (int64_t t = dl; t < de; ++t) {
    if (j < ends_s[wib][i - i0]) {
      acc += prod_s[wib][(j - j0) + ((j - j0) >> 3)];
      ++j;
    } else {
      if (first_row < 0) {
        first_row = i;
        first_part = acc;
      } else {
        a.y[i] = acc;  // the whole row lies in this lane's items
      }
      ++i;
      acc = 0;
    }
  }
  // segmented inclusive scan of the lanes' open-row sums, keyed by the row
  // each lane ends in (keys ascend with the lane, so equal keys are adjacent)
  AT sum = acc;
#pragma unroll
  for (int off = 1; off < 64; off *= 2) {
    const AT o = __shfl_up(sum, off);
    const int oi = __shfl_up(i, off);
    if (lane >= off && oi == i) sum += o;
  }
  const AT prev_sum = __shfl_up(sum, 1);
  const int prev_row = __shfl_up(i, 1);
  if (first_row >= 0) a.y[first_row] = (lane > 0 && prev_row == first_row ? prev_sum : (AT) 0) + first_part;
  if (lane == 63) {
    carry_row[w] = i;
    carry_val[w] = sum;
  }
}

// Adds the wavefront carries: each run of wavefronts ending in the same row
// is summed in wavefront order by the thread of its first wavefront.
template <class AT>
__global__ __launch_bounds__(256) void k_csr_merge_fixup(int64_t n_waves, int n, const int *__restrict__ carry_row,
                                                         const AT *__restrict__ carry_val, AT *__restrict__ y)
{
  const int64_t w = (int64_t) blockIdx.x * 256 + threadIdx.x;
  if (w >= n_waves) return;
  const int r = carry_row[w];
  if (r >= n || (w > 0 && carry_row[w - 1] == r)) return;
  AT s = 0;
  for (int64_t k = w; k < n_waves && carry_row[k] == r; ++k) s += carry_val[k];
  y[r] += s;
}

}  // namespace bmfp
