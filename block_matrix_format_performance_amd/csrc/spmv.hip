// spmv.hip -- instantiation tables and launchers for the SpMV kernels.
#include <rocsparse/rocsparse.h>

#include <cstdlib>

#include "matrix.hpp"
#include "spmv_kernels.hpp"

static_assert(bmfp::MERGE_TILE == bmfp::MERGE_TILE_ITEMS, "merge tile and carry sizing disagree");

namespace bmfp {

namespace {

// XCD-contiguous tile order (tile_id).  Off by default: measured on the
// 9-point stencil sweep it is neutral to slightly slower than the hardware's
// round-robin order; on the LDS-staged bs 1-3 kernels it brings the L2-miss
// traffic down to the compulsory bytes (x read once per XCD) yet runs 7-9%
// slower, the MALL already serving the x re-reads.  BMFP_XCD_REMAP=1
// switches it on.
int xcd_remap_enabled()
{
  static int v = [] {
    const char *e = std::getenv("BMFP_XCD_REMAP");
    return e ? std::atoi(e) : 0;
  }();
  return v;
}

struct Grid {
  unsigned groups;
  int per;
};

// Kernel-symbol queries (spmv_kernel_fn) run a launcher with tl_probe set: the
// launch of the path's main kernel records its host stub and nothing runs.
thread_local bool tl_probe = false;
thread_local const void *tl_probe_fn = nullptr;

template <class... P, class... A>
void launch(void (*k)(P...), unsigned groups, unsigned threads, hipStream_t s, A... a)
{
  if (tl_probe) {
    tl_probe_fn = reinterpret_cast<const void *>(k);
    return;
  }
  k<<<groups, threads, 0, s>>>(a...);
}

// A helper kernel of a multi-kernel path (merge-path bounds and fix-up): a
// probe skips it without recording it.
template <class... P, class... A>
void launch_aux(void (*k)(P...), unsigned groups, unsigned threads, hipStream_t s, A... a)
{
  if (tl_probe) return;
  k<<<groups, threads, 0, s>>>(a...);
}

Grid make_grid(int64_t threads)
{
  int64_t g = (threads + 255) / 256;
  if (g < 1) g = 1;
  Grid out{(unsigned) g, 0};
  if (xcd_remap_enabled() && g >= 64) {
    const int64_t g8 = (g + 7) / 8 * 8;
    out.groups = (unsigned) g8;
    out.per = (int) (g8 / 8);
  }
  return out;
}

// Columns per lane in the lane-group-per-row kernel: one 16-byte load, or
// two for FP64 blocks of 12 and more columns (fewer lanes per row, more bytes
// in flight per lane: +3-4 points of HBM peak measured at bs 12-32).
template <class VT> constexpr int gpr_vec(int bs)
{
  constexpr int W = 16 / (int) sizeof(VT);  // elements per 16-byte load
  int v = 1;
  while (v < W && bs % (2 * v) == 0) v *= 2;
  if (sizeof(VT) == 8 && bs >= 12 && bs % 4 == 0) v = 4;
  return v;
}
constexpr int bgr_lanes(int bs) { return bs == 1 ? 16 : bs <= 3 ? 32 : 64; }

template <class VT, class AT, bool SCALED>
BcsrArgs<VT, AT> bcsr_args(const bmfp_matrix_t &A, const void *x, void *y, int per)
{
  BcsrArgs<VT, AT> a;
  a.n_brows = A.n_rows;
  a.xcd_per = per;
  a.rp = A.row_ptr->as<int>();
  a.cols = A.columns->as<int>();
  a.vals = A.values.as<VT>();
  a.scale = SCALED ? A.scale.as<float>() : nullptr;
  a.x = static_cast<const AT *>(x);
  a.y = static_cast<AT *>(y);
  return a;
}

// Wavefronts per workgroup of k_bcsr_stage (LDS per wavefront: 6.9 KB at bs 1,
// 10.4 KB at bs 2, 14.4 KB at bs 3 in FP64), the fastest of 1, 2 and 4 on the
// HBM-resident 9-point stencils.
template <int BS> constexpr int stage_wpg() { return BS == 1 ? 4 : 2; }
// XCD bands of k_bcsr_stage at bs 1 and 2: 128 wavefronts (8192 block rows
// at bs 1, 4096 at bs 2).  On the HBM-resident 9-point stencils (a band ~1.2
// grid rows) they cut the fetched bytes by 4.8% at bs 1 and 5.7% at bs 2,
// at -0.4 and +0.9 points of HBM peak in paired sweeps (bs 3: -0.9, so it
// keeps the hardware order); bands of 4 grid rows and more fetch less still
// but run 1.5-3 points slower.
template <int BS> constexpr int stage_band_waves() { return BS <= 2 ? 128 : 0; }

template <class VT, class AT, int BS, int LAYOUT, bool SCALED>
void launch_stage(const bmfp_matrix_t &A, const void *x, void *y, hipStream_t s)
{
  constexpr int RPW = 64 / BS, WPG = stage_wpg<BS>(), BAND = stage_band_waves<BS>() / WPG;
  const int64_t waves = std::max<int64_t>((A.n_rows + RPW - 1) / RPW, 1);
  const int64_t groups = (waves + WPG - 1) / WPG;  // one group per WPG wavefronts
  if (BAND > 0 && groups >= 8 * BAND && !xcd_remap_enabled()) {
    // whole rounds of 8 bands; groups past the last block row exit at once
    const int64_t gb = (groups + 8 * BAND - 1) / (8 * BAND) * (8 * BAND);
    BcsrArgs<VT, AT> a = bcsr_args<VT, AT, SCALED>(A, x, y, 0);
    a.band = BAND;
    launch(k_bcsr_stage<VT, AT, BS, LAYOUT, SCALED, WPG>, (unsigned) gb, 64 * WPG, s, a);
    return;
  }
  const Grid g = make_grid(groups * 256);
  launch(k_bcsr_stage<VT, AT, BS, LAYOUT, SCALED, WPG>, g.groups, 64 * WPG, s,
         bcsr_args<VT, AT, SCALED>(A, x, y, g.per));
}

// XCD bands of the dense lane-group kernels (tile_id(per, band)): about 8192
// scalar rows per band, at bs 8 and 12, where paired sweeps on the
// HBM-resident stencils measured row-major bs 12 74.4 -> 76.8% of HBM peak
// and bs 8 +0.3 points in both layouts; at bs 4 (-0.6), 6 (0) and 16-32
// (-0.3 to -0.9) the hardware order is kept.
template <int BS> constexpr bool dense_banded() { return BS == 8 || BS == 12; }

// Launches k with `groups` workgroups of 256 lanes covering rows_per_group
// scalar rows each, dealt to XCDs in bands when dense_banded<BS>().
template <class VT, class AT, bool SCALED, int BS, class K>
void launch_dense(K k, const bmfp_matrix_t &A, const void *x, void *y, hipStream_t s, int64_t groups,
                  int64_t rows_per_group)
{
  const int band = (int) std::max<int64_t>(1, 8192 / rows_per_group);
  if (dense_banded<BS>() && groups >= 8 * band && !xcd_remap_enabled()) {
    // whole rounds of 8 bands; waves past the last row exit at once
    const int64_t gb = (groups + 8 * band - 1) / (8 * band) * (8 * band);
    BcsrArgs<VT, AT> a = bcsr_args<VT, AT, SCALED>(A, x, y, 0);
    a.band = band;
    launch(k, (unsigned) gb, 256, s, a);
    return;
  }
  const Grid g = make_grid(groups * 256);
  launch(k, g.groups, 256, s, bcsr_args<VT, AT, SCALED>(A, x, y, g.per));
}

template <class VT, class AT, int BS, int VEC, int U, bool SCALED, bool NT>
void launch_dgc(const bmfp_matrix_t &A, const void *x, void *y, hipStream_t s)
{
  constexpr int BPW = 64 / (BS * (BS / VEC));
  const int64_t waves = (A.n_rows + BPW - 1) / BPW;
  launch_dense<VT, AT, SCALED, BS>(k_bcsr_dgc<VT, AT, BS, VEC, U, SCALED, NT>, A, x, y, s, (waves + 3) / 4,
                                   4 * BPW * BS);
}

// Column-major FP64 lane-group shapes (values per lane, blocks in flight per
// lane), the fastest measured on HBM-resident 9-point stencil matrices
// (~4.4e8 stored values, interleaved timings, % of 8 TB/s): bs 4 (2, 4) 74.1
// (2, 3: 73.5), bs 6 (3, 4) 71.8-71.9 (3, 3: 71.4; 2, 9: 69.0; one x element
// and three 8-byte loads per lane, 5 block rows in 60 lanes of a wavefront),
// bs 8 (2, 9) 78.1, bs 12 (6, 2) 73.5 (4, 3: 72.0), bs 16 (4, 3) 75.2,
// bs 24 (12, 1) 75.3.  At bs 32 every shape stays near 50 (64 lanes hold one
// block row, so a wavefront has one x segment per block), and the per-row
// lane-slice kernel is kept.  bs 8 loads its values non-temporal (73.3 ->
// 76.1 / 78.4 in paired runs); at the other sizes that was slower (bs 6
// -6 to -26, bs 12 -7, bs 16 -5 to -30, bs 24 -40 points).
template <int BS>
constexpr DenseShape dgc_shape()
{
  if constexpr (BS == 4) return {2, 4};
  else if constexpr (BS == 6) return {3, 4};
  else if constexpr (BS == 8) return {2, 9, true};
  else if constexpr (BS == 12) return {6, 2};
  else if constexpr (BS == 16) return {4, 3};
  else if constexpr (BS == 24) return {12, 1};
  else return {0, 0};
}

template <class VT, class AT, bool SCALED, int BS, int LAYOUT>
int launch_fixed(const bmfp_matrix_t &A, int variant, const void *x, void *y, hipStream_t s)
{
  const int64_t nb = A.n_rows;
  switch (variant) {
  case BMFP_BCSR_THREAD_PER_ROW: {
    Grid g = make_grid(nb * BS);
    launch(k_bcsr_tpr<VT, AT, BS, LAYOUT, SCALED>, g.groups, 256, s, bcsr_args<VT, AT, SCALED>(A, x, y, g.per));
    break;
  }
  case BMFP_BCSR_GROUP_PER_ROW: {
    constexpr DenseShape D = dense_shape<VT, BS>();
    if constexpr (BS <= 3 && (BS * BS * (int) sizeof(VT)) % 4 == 0) {
      launch_stage<VT, AT, BS, LAYOUT, SCALED>(A, x, y, s);
    } else if constexpr (BS == 2 || BS == 3) {
      constexpr int RPW = 64 / BS;
      Grid g = make_grid((nb + RPW - 1) / RPW * 64);
      launch(k_bcsr_split<VT, AT, BS, LAYOUT, SCALED>, g.groups, 256, s, bcsr_args<VT, AT, SCALED>(A, x, y, g.per));
    } else if constexpr (LAYOUT == BMFP_ROW_MAJOR && D.vec > 0) {
      constexpr int RPW = 64 / (BS / D.vec);
      const int64_t waves = (nb * BS + RPW - 1) / RPW;
      launch_dense<VT, AT, SCALED, BS>(k_bcsr_dgr<VT, AT, BS, D.vec, D.u, SCALED, D.nt>, A, x, y, s,
                                       (waves + 3) / 4, 4 * RPW);
    } else if constexpr (LAYOUT == BMFP_COL_MAJOR && sizeof(VT) == 8 && dgc_shape<BS>().vec > 0) {
      constexpr DenseShape G = dgc_shape<BS>();
      launch_dgc<VT, AT, BS, G.vec, G.u, SCALED, G.nt>(A, x, y, s);
    } else {
      constexpr int VEC = gpr_vec<VT>(BS);
      constexpr int G = pow2ceil(BS / VEC);
      Grid g = make_grid(nb * BS * G);
      launch(k_bcsr_gpr<VT, AT, BS, VEC, LAYOUT, SCALED>, g.groups, 256, s, bcsr_args<VT, AT, SCALED>(A, x, y, g.per));
    }
    break;
  }
  case BMFP_BCSR_SHUFFLE_X: {
    constexpr int RPW = 64 / BS;
    Grid g = make_grid((nb + RPW - 1) / RPW * 64);
    launch(k_bcsr_shfx<VT, AT, BS, LAYOUT, SCALED>, g.groups, 256, s, bcsr_args<VT, AT, SCALED>(A, x, y, g.per));
    break;
  }
  case BMFP_BCSR_BLOCK_ROW_GROUP: {
    if constexpr (LAYOUT != BMFP_COL_MAJOR)
      return fail(BMFP_ERR_UNSUPPORTED, "BMFP_BCSR_BLOCK_ROW_GROUP needs column-major blocks");
    else {
      constexpr int GL = bgr_lanes(BS);
      Grid g = make_grid(nb * GL);
      launch(k_bcsr_bgr<VT, AT, BS, GL, SCALED>, g.groups, 256, s, bcsr_args<VT, AT, SCALED>(A, x, y, g.per));
    }
    break;
  }
  default:
    return fail(BMFP_ERR_INVALID, "variant %d is not a BCSR kernel family", variant);
  }
  BMFP_LAUNCHED();
  return BMFP_OK;
}

template <class VT, class AT, bool SCALED, int LAYOUT>
int launch_runtime_bs(const bmfp_matrix_t &A, int variant, const void *x, void *y, hipStream_t s)
{
  if (variant != BMFP_BCSR_THREAD_PER_ROW)
    return fail(BMFP_ERR_UNSUPPORTED, "block size %d has no precompiled kernel for variant %d", A.bs, variant);
  Grid g = make_grid((int64_t) A.n_rows * A.bs);
  launch(k_bcsr_tpr_rt<VT, AT, LAYOUT, SCALED>, g.groups, 256, s, bcsr_args<VT, AT, SCALED>(A, x, y, g.per), A.bs);
  BMFP_LAUNCHED();
  return BMFP_OK;
}

template <class VT, class AT, bool SCALED, int LAYOUT>
int launch_layout(const bmfp_matrix_t &A, int variant, const void *x, void *y, hipStream_t s)
{
#define BMFP_BS_CASE(B) \
  case B: return launch_fixed<VT, AT, SCALED, B, LAYOUT>(A, variant, x, y, s);
  switch (A.bs) {
    BMFP_BS_CASE(1)
    BMFP_BS_CASE(2)
    BMFP_BS_CASE(3)
    BMFP_BS_CASE(4)
    BMFP_BS_CASE(6)
    BMFP_BS_CASE(8)
    BMFP_BS_CASE(12)
    BMFP_BS_CASE(16)
    BMFP_BS_CASE(24)
    BMFP_BS_CASE(32)
  default:
    return launch_runtime_bs<VT, AT, SCALED, LAYOUT>(A, variant, x, y, s);
  }
#undef BMFP_BS_CASE
}

template <class VT, class AT, bool SCALED>
int launch_typed(const bmfp_matrix_t &A, int variant, const void *x, void *y, hipStream_t s)
{
  return A.layout == BMFP_ROW_MAJOR ? launch_layout<VT, AT, SCALED, BMFP_ROW_MAJOR>(A, variant, x, y, s)
                                    : launch_layout<VT, AT, SCALED, BMFP_COL_MAJOR>(A, variant, x, y, s);
}

bool precompiled_bs(int bs)
{
  switch (bs) {
  case 1: case 2: case 3: case 4: case 6: case 8: case 12: case 16: case 24: case 32: return true;
  default: return false;
  }
}

}  // namespace

bool bcsr_supported(int variant, int dtype, int layout, int bs)
{
  if (bs < 1) return false;
  if (variant == BMFP_BCSR_LIBRARY || variant == BMFP_BCSR_JIT) return dtype != BMFP_F16;
  if (!precompiled_bs(bs)) return variant == BMFP_BCSR_THREAD_PER_ROW;
  if (variant == BMFP_BCSR_BLOCK_ROW_GROUP) return layout == BMFP_COL_MAJOR;
  return variant == BMFP_BCSR_THREAD_PER_ROW || variant == BMFP_BCSR_GROUP_PER_ROW ||
         variant == BMFP_BCSR_SHUFFLE_X;
}

int auto_variant(const bmfp_matrix_t &A)
{
  if (!A.is_bcsr) {
    // measured on ~1.4e8-nonzero matrices (equal, ragged and power-law rows,
    // and the CSR expansions of the 9-point block stencils): merge-path tiles
    // whenever one row dwarfs the mean (a lane group would serialise it) and
    // for ragged mid-length rows, whose lengths leave lanes of a group idle;
    // LDS-staged lane groups for near-equal mid-length rows; lane groups for
    // long rows and for short ragged rows; thread per row for short rows of
    // (nearly) equal length and for short rows that fill a staged wavefront
    const double mean = A.mean_row_blocks, mx = A.max_row_blocks;
    if (mx > 8.0 * mean + 64.0) return BMFP_CSR_MERGE;
    if (mean >= 128.0) return BMFP_CSR_VECTOR;
    if (mean < 12.0)
      return mx <= mean + 2.0 || csr_stage_group(mean, mx, 1) == 1 ? BMFP_CSR_SCALAR : BMFP_CSR_VECTOR;
    return mx <= 1.25 * mean + 2.0 ? BMFP_CSR_VECTOR : BMFP_CSR_MERGE;
  }
  // Fastest family per (layout, bs) measured on HBM-resident 9-point stencil
  // matrices (FP64, ~4.4e8 stored values): for bs 1-3 the wavefront's blocks
  // staged in LDS and split over each block row's lanes (bs 1: FP64 and FP32,
  // while rows average at most 9 blocks, the staging capacity; otherwise
  // batched thread per row); 16-byte lane slices (row-major) from bs = 4 up;
  // column-major: the column lane groups (FP64, bs 4-24, 69-77% of 8 TB/s
  // against 64-73% for shuffle-shared x segments), shuffle-shared x segments
  // otherwise.
  // block sizes without a precompiled instantiation: a kernel specialised by
  // hipRTC (F64 / F32), the runtime-bs thread-per-row kernel for F16
  if (!precompiled_bs(A.bs)) return A.dtype == BMFP_F16 ? BMFP_BCSR_THREAD_PER_ROW : BMFP_BCSR_JIT;
  if (A.bs == 1)  // staged for FP64 and FP32 (whole 4-byte words)
    return A.dtype != BMFP_F16 && A.mean_row_blocks <= 9.0 ? BMFP_BCSR_GROUP_PER_ROW : BMFP_BCSR_THREAD_PER_ROW;
  if (A.bs <= 3) return BMFP_BCSR_GROUP_PER_ROW;
  if (A.layout == BMFP_ROW_MAJOR) return BMFP_BCSR_GROUP_PER_ROW;
  return A.dtype == BMFP_F64 && A.bs <= 24 ? BMFP_BCSR_GROUP_PER_ROW : BMFP_BCSR_SHUFFLE_X;
}

int launch_bcsr(const bmfp_matrix_t &A, int variant, const void *x, void *y, hipStream_t s)
{
  if (A.n_rows == 0) return BMFP_OK;
  if (!bcsr_supported(variant, A.dtype, A.layout, A.bs))
    return fail(BMFP_ERR_UNSUPPORTED, "BCSR variant %d unsupported for dtype %d layout %d bs %d",
                variant, A.dtype, A.layout, A.bs);
  if (variant == BMFP_BCSR_JIT) return launch_bcsr_jit(A, x, y, s);
  switch (A.dtype) {
  case BMFP_F64: return launch_typed<double, double, false>(A, variant, x, y, s);
  case BMFP_F32: return launch_typed<float, float, false>(A, variant, x, y, s);
  case BMFP_F16: return launch_typed<__half, float, true>(A, variant, x, y, s);
  default: return fail(BMFP_ERR_INVALID, "bad dtype %d", A.dtype);
  }
}

// Lanes per row for the CSR lane-group kernel: the power of two nearest half
// the mean row length, clamped to [2, 64].
int csr_group_for(double mean_len)
{
  int g = 2;
  while (g < 64 && 2.0 * g < mean_len) g *= 2;
  return g;
}

// Lanes per row of k_csr_stage: the smallest power of two (>= lo) whose
// wavefront of 64/G rows holds at most 576 nonzeros on average (9/10 of the
// staging capacity, room for rows longer than the mean).  0 where staging
// was measured not to pay: under 288 nonzeros per wavefront (rows of 4 with
// one lane each: 4.7 vs 5.3 TB/s), 32+ lanes per row (rows of 256: 5.0 vs
// 5.6 TB/s for the lane-group kernel), or a longest row over 8x the mean +
// 64 (its wavefront would walk it with the narrow group: power-law rows 26
// vs 95 GB/s).
int csr_stage_group(double mean_len, double max_len, int lo)
{
  constexpr double FILL = 0.9 * CSR_STAGE_CAP;
  if (max_len > 8.0 * mean_len + 64.0) return 0;
  int g = lo;
  while (g < 64 && (64.0 / g) * mean_len > FILL) g *= 2;
  const double fill = (64.0 / g) * mean_len;
  return fill <= FILL && fill >= 0.45 * CSR_STAGE_CAP && g <= 16 ? g : 0;
}

template <class VT, class AT, int G>
void launch_csr_stage(const bmfp_matrix_t &A, CsrArgs<VT, AT> a, hipStream_t s)
{
  constexpr int WPG = 2, RPW = 64 / G;
  const int64_t waves = std::max<int64_t>((A.n_rows + RPW - 1) / RPW, 1);
  const Grid g = make_grid((waves + WPG - 1) / WPG * 256);  // one group per WPG wavefronts
  a.xcd_per = g.per;
  launch(k_csr_stage<VT, AT, G, WPG>, g.groups, 64 * WPG, s, a);
}

template <class VT, class AT>
int launch_csr_typed(const bmfp_matrix_t &A, int variant, const void *x, void *y, hipStream_t s)
{
  CsrArgs<VT, AT> a;
  a.n_rows = A.n_rows;
  a.rp = A.row_ptr->as<int>();
  a.cols = A.columns->as<int>();
  a.vals = A.values.as<VT>();
  a.x = static_cast<const AT *>(x);
  a.y = static_cast<AT *>(y);
  if (variant == BMFP_CSR_SCALAR) {
    if (csr_stage_group(A.mean_row_blocks, A.max_row_blocks, 1) == 1) {
      launch_csr_stage<VT, AT, 1>(A, a, s);
    } else {
      Grid g = make_grid(A.n_rows);
      a.xcd_per = g.per;
      launch(k_csr_scalar<VT, AT>, g.groups, 256, s, a);
    }
  } else if (variant == BMFP_CSR_VECTOR && csr_stage_group(A.mean_row_blocks, A.max_row_blocks, 2) > 0) {
    switch (csr_stage_group(A.mean_row_blocks, A.max_row_blocks, 2)) {
    case 2: launch_csr_stage<VT, AT, 2>(A, a, s); break;
    case 4: launch_csr_stage<VT, AT, 4>(A, a, s); break;
    case 8: launch_csr_stage<VT, AT, 8>(A, a, s); break;
    default: launch_csr_stage<VT, AT, 16>(A, a, s); break;
    }
  } else if (variant == BMFP_CSR_VECTOR) {
    const int G = A.csr_group;
    Grid g = make_grid((int64_t) A.n_rows * G);
    a.xcd_per = g.per;
    switch (G) {
    case 2: launch(k_csr_vector<VT, AT, 2>, g.groups, 256, s, a); break;
    case 4: launch(k_csr_vector<VT, AT, 4>, g.groups, 256, s, a); break;
    case 8: launch(k_csr_vector<VT, AT, 8>, g.groups, 256, s, a); break;
    case 16: launch(k_csr_vector<VT, AT, 16>, g.groups, 256, s, a); break;
    case 32: launch(k_csr_vector<VT, AT, 32>, g.groups, 256, s, a); break;
    default: launch(k_csr_vector<VT, AT, 64>, g.groups, 256, s, a); break;
    }
  } else if (variant == BMFP_CSR_MERGE) {
    const int64_t waves = merge_waves(A.n_rows, A.nnzb);
    if (A.merge_carry.bytes < merge_carry_bytes(A.n_rows, A.nnzb))
      return fail(BMFP_ERR_INVALID, "matrix has no merge carry buffer");
    int *carry_row = A.merge_carry.as<int>();
    AT *carry_val = reinterpret_cast<AT *>(A.merge_carry.as<char>() + (size_t) waves * 8);
    int *bounds = reinterpret_cast<int *>(A.merge_carry.as<char>() + (size_t) waves * 16);
    launch_aux(k_csr_merge_bounds<MERGE_TILE>, (unsigned) ((waves + 1 + 255) / 256), 256, s, A.n_rows, A.nnzb,
               waves, a.rp, bounds);
    BMFP_LAUNCHED();
    launch(k_csr_merge<VT, AT>, (unsigned) ((waves + 3) / 4), 256, s, a, A.nnzb, waves, bounds, carry_row, carry_val);
    BMFP_LAUNCHED();
    launch_aux(k_csr_merge_fixup<AT>, (unsigned) ((waves + 255) / 256), 256, s, waves, A.n_rows, carry_row,
               carry_val, a.y);
  } else {
    return fail(BMFP_ERR_INVALID, "variant %d is not a CSR kernel family", variant);
  }
  BMFP_LAUNCHED();
  return BMFP_OK;
}

int launch_csr(const bmfp_matrix_t &A, int variant, const void *x, void *y, hipStream_t s)
{
  if (A.n_rows == 0) return BMFP_OK;
  switch (A.dtype) {
  case BMFP_F64: return launch_csr_typed<double, double>(A, variant, x, y, s);
  case BMFP_F32: return launch_csr_typed<float, float>(A, variant, x, y, s);
  default: return fail(BMFP_ERR_UNSUPPORTED, "CSR SpMV supports F64 and F32 storage");
  }
}

int spmv_kernel_fn(const bmfp_matrix_t &A, int variant, const void **fn)
{
  if (variant == BMFP_SPMV_AUTO) variant = auto_variant(A);
  if (variant == BMFP_BCSR_LIBRARY || variant == BMFP_BCSR_JIT)
    return fail(BMFP_ERR_UNSUPPORTED, "variant %d runs library or run-time compiled code", variant);
  const bool csr_family = variant == BMFP_CSR_SCALAR || variant == BMFP_CSR_VECTOR || variant == BMFP_CSR_MERGE;
  if (A.is_bcsr == csr_family) return fail(BMFP_ERR_INVALID, "variant %d does not match the matrix format", variant);
  tl_probe = true;
  tl_probe_fn = nullptr;
  const int st = A.is_bcsr ? launch_bcsr(A, variant, nullptr, nullptr, nullptr)
                           : launch_csr(A, variant, nullptr, nullptr, nullptr);
  tl_probe = false;
  *fn = tl_probe_fn;
  if (st == BMFP_OK && !*fn) return fail(BMFP_ERR_INVALID, "the matrix has no rows: no kernel runs");
  return st;
}

// ---- rocSPARSE bsrmv: the library baseline row (reference: cuSPARSE) -------
#define BMFP_RS(call)                                                                  \
  do {                                                                                 \
    rocsparse_status st_ = (call);                                                     \
    if (st_ != rocsparse_status_success)                                               \
      return fail(BMFP_ERR_HIP, "%s failed: rocsparse status %d", #call, (int) st_);   \
  } while (0)

int launch_library(bmfp_matrix_t &A, const void *x, void *y, hipStream_t s)
{
  if (!A.is_bcsr || A.dtype == BMFP_F16)
    return fail(BMFP_ERR_UNSUPPORTED, "library bsrmv needs an F64/F32 BCSR matrix");
  if (A.n_rows == 0) return BMFP_OK;
  rocsparse_handle h = static_cast<rocsparse_handle>(A.lib_handle);
  if (!h) {
    BMFP_RS(rocsparse_create_handle(&h));
    A.lib_handle = h;
    rocsparse_mat_descr d;
    BMFP_RS(rocsparse_create_mat_descr(&d));
    A.lib_descr = d;
  }
  BMFP_RS(rocsparse_set_stream(h, s));
  rocsparse_mat_descr d = static_cast<rocsparse_mat_descr>(A.lib_descr);
  const rocsparse_direction dir =
      A.layout == BMFP_ROW_MAJOR ? rocsparse_direction_row : rocsparse_direction_column;
  if (A.dtype == BMFP_F64) {
    const double one = 1.0, zero = 0.0;
    BMFP_RS(rocsparse_dbsrmv(h, dir, rocsparse_operation_none, A.n_rows, A.n_cols, A.nnzb, &one, d,
                             A.values.as<double>(), A.row_ptr->as<int>(), A.columns->as<int>(), A.bs,
                             nullptr, static_cast<const double *>(x), &zero, static_cast<double *>(y)));
  } else {
    const float one = 1.0f, zero = 0.0f;
    BMFP_RS(rocsparse_sbsrmv(h, dir, rocsparse_operation_none, A.n_rows, A.n_cols, A.nnzb, &one, d,
                             A.values.as<float>(), A.row_ptr->as<int>(), A.columns->as<int>(), A.bs,
                             nullptr, static_cast<const float *>(x), &zero, static_cast<float *>(y)));
  }
  return BMFP_OK;
}

void destroy_library(bmfp_matrix_t &A)
{
  if (A.lib_descr) rocsparse_destroy_mat_descr(static_cast<rocsparse_mat_descr>(A.lib_descr));
  if (A.lib_handle) rocsparse_destroy_handle(static_cast<rocsparse_handle>(A.lib_handle));
  A.lib_descr = nullptr;
  A.lib_handle = nullptr;
}

}  // namespace bmfp
