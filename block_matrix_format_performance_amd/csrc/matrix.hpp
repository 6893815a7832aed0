// matrix.hpp -- device-resident sparse matrix handle (bmfp_matrix).
#pragma once

#include "common.hpp"

struct bmfp_matrix_t {
  int is_bcsr = 1;
  int dtype = BMFP_F64;        // storage type of values
  int layout = BMFP_ROW_MAJOR; // block layout (BCSR only)
  int n_rows = 0;              // block rows (BCSR) or rows (CSR)
  int n_cols = 0;              // block columns (BCSR) or columns (CSR)
  int bs = 1;                  // 1 for CSR
  int nnzb = 0;                // nonzero blocks (BCSR) or nonzeros (CSR)
  int device = 0;
  int max_row_blocks = 0;      // longest (block) row
  double mean_row_blocks = 0;  // mean (block) row length
  int csr_group = 0;           // lanes per row chosen for BMFP_CSR_VECTOR
  int stencil_nx = 0, stencil_ny = 0;  // grid of a bmfp_stencil_create matrix (0: other)

  std::shared_ptr<bmfp::DeviceBuffer> row_ptr;  // int[n_rows+1]  (shared by converted copies)
  std::shared_ptr<bmfp::DeviceBuffer> columns;  // int[nnzb]
  bmfp::DeviceBuffer values;                    // dtype[nnzb*bs*bs]
  bmfp::DeviceBuffer scale;                     // float[n_rows] (BMFP_F16 only)
  bmfp::DeviceBuffer merge_carry;               // per-wavefront carries of BMFP_CSR_MERGE (CSR only)
  // Pass-sliced values (the solver's inner copies only): block j of block row
  // P*sl_rpw + l at ((P*sl_j + j)*sl_rpw + l), rows padded with zero blocks to
  // sl_j; its block columns in cols_sl.  sl_j = 0: values in BCSR order.
  int sl_j = 0, sl_rpw = 0;
  bmfp::DeviceBuffer cols_sl;

  // rocSPARSE state for BMFP_BCSR_LIBRARY, created on first use.
  void *lib_handle = nullptr;
  void *lib_descr = nullptr;
  ~bmfp_matrix_t();

  int64_t n_vals() const { return (int64_t) nnzb * bs * bs; }
  int64_t x_len() const { return (int64_t) n_cols * bs; }
  int64_t y_len() const { return (int64_t) n_rows * bs; }
};

namespace bmfp {
// Merge-path CSR SpMV: wavefront tiles of 512 path items (row ends +
// nonzeros) and the scratch they need: per wavefront a carry row and value
// (8-byte slots each), then the n_waves + 1 tile boundaries (int).
constexpr int64_t MERGE_TILE_ITEMS = 512;
inline int64_t merge_waves(int64_t n_rows, int64_t nnz) { return (n_rows + nnz + MERGE_TILE_ITEMS - 1) / MERGE_TILE_ITEMS; }
inline size_t merge_carry_bytes(int64_t n_rows, int64_t nnz)
{
  return (size_t) merge_waves(n_rows, nnz) * 16 + (size_t) (merge_waves(n_rows, nnz) + 1) * 4;
}
int launch_bcsr(const bmfp_matrix_t &A, int variant, const void *x, void *y, hipStream_t s);
// The hipRTC-specialised kernel for A's block size (F64 / F32).
int launch_bcsr_jit(const bmfp_matrix_t &A, const void *x, void *y, hipStream_t s);
int launch_csr(const bmfp_matrix_t &A, int variant, const void *x, void *y, hipStream_t s);
// Host stub of the kernel launch_bcsr / launch_csr would run for `variant`
// (the tile kernel of the merge path); nothing is launched.
int spmv_kernel_fn(const bmfp_matrix_t &A, int variant, const void **fn);
int launch_library(bmfp_matrix_t &A, const void *x, void *y, hipStream_t s);
void destroy_library(bmfp_matrix_t &A);
int auto_variant(const bmfp_matrix_t &A);
int csr_group_for(double mean_len);
int csr_stage_group(double mean_len, double max_len, int lo);
bool bcsr_supported(int variant, int dtype, int layout, int bs);
// A pass-sliced FP32 / FP16 copy of the F64 row-major BCSR matrix src (layout
// above, rpw block rows per pass, j_slots slots per row) into m, whose shape
// fields the caller has filled; FP16 rows scaled by one FP32 factor each.
// e0 / e1 (optional) are recorded on s around the copy's kernels, after its
// allocations.
int make_sliced_copy(const bmfp_matrix_t &src, int dtype, int rpw, int j_slots, bmfp_matrix_t &m, hipStream_t s,
                     hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// Device address of the allocation-free entry point's scratch on `device`
// (the current device), resolved once per device.
int noalloc_setup(int device, void **out);
}  // namespace bmfp
