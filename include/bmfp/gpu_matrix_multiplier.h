// gpu_matrix_multiplier.h -- drop-in for the reference's
// gpu/gpu_matrix_multiplier.h:13-27: the three SpMV entry points keep their
// names, signatures and measurement names; underneath they call libbmfp's
// gfx950 kernels through the C ABI (include/bmfp.h).
//
// Behaviour kept: each call uploads the matrix, runs y = A x with x = 1
// (gpu_matrix_multiplier.cu:122), downloads y, reports a mismatch against
// reference_y with compare_results and returns timings.
// Behaviour changed (reference defects not reproduced):
//   - timings are the mean of 10 launches after one warm-up, recorded with
//     HIP events on the work stream (the reference times one cold launch);
//   - every measurement carries the compulsory bytes and flops of one SpMV
//     (the reference records 0 bytes for GPU rows);
//   - every BCSR family is exact for every block size (the reference's
//     "coal x" and "column-by-column" kernels are wrong for many bs and its
//     switch statements launch nothing for bs outside {1,2,3,4,8,16,32}).
// Extensions: overloads taking an explicit x, and a variant returning y.
#ifndef BMFP_GPU_MATRIX_MULTIPLIER_H
#define BMFP_GPU_MATRIX_MULTIPLIER_H

#include <vector>

#include "matrix_converters.h"
#include "measurement_class.h"

// "GPU CSR": thread per row (gpu_matrix_multiplier.cu:88-163).
template <typename data_type, typename index_type>
measurement_class gpu_csr_spmv(const csr_matrix_class<data_type, index_type> &matrix, const data_type *reference_y);

// "GPU CSR-Vector": lane group per row (gpu_matrix_multiplier.cu:198-273).
template <typename data_type, typename index_type>
measurement_class gpu_csr_vector_spmv(const csr_matrix_class<data_type, index_type> &matrix,
                                      const data_type *reference_y);

// All BCSR families, in the reference's order and names
// (gpu_matrix_multiplier.cu:736-1331); transpose_matrix_data holds the
// column-major copy of the values (bcsr_matrix_class::transpose_blocks).
template <typename data_type, typename index_type>
std::vector<measurement_class> gpu_bcsr_spmv(bcsr_matrix_class<data_type, index_type> &matrix,
                                             const data_type *transpose_matrix_data, const data_type *reference_y);

// ---- extensions ---------------------------------------------------------------
// Same as above with an explicit x (length n_cols); y_out (length n_rows)
// receives the last result when non-null.
template <typename data_type, typename index_type>
measurement_class gpu_csr_spmv(const csr_matrix_class<data_type, index_type> &matrix, const data_type *x,
                               const data_type *reference_y, data_type *y_out);

template <typename data_type, typename index_type>
measurement_class gpu_csr_vector_spmv(const csr_matrix_class<data_type, index_type> &matrix, const data_type *x,
                                      const data_type *reference_y, data_type *y_out);

template <typename data_type, typename index_type>
std::vector<measurement_class> gpu_bcsr_spmv(bcsr_matrix_class<data_type, index_type> &matrix,
                                             const data_type *transpose_matrix_data, const data_type *x,
                                             const data_type *reference_y);

// "GPU Merge SpMV": merge-path tiles, equal work per wavefront whatever the
// row lengths (the name of the rows in the reference's results/bs_range.json,
// whose kernel is not in the tree).
template <typename data_type, typename index_type>
measurement_class gpu_csr_merge_spmv(const csr_matrix_class<data_type, index_type> &matrix, const data_type *x,
                                     const data_type *reference_y, data_type *y_out);

// Measurement names of gpu_bcsr_spmv, in order.
const std::vector<std::string> &gpu_bcsr_spmv_names();

// "jit": the reference's run-time-specialised BCSR kernel (NVRTC, inline in
// main.cpp:170-294) as the hipRTC-compiled kernel for the matrix's bs on the
// column-major values; compiled first, then one timed launch, as there.
// data_type float or double (the reference's is float only).
template <typename data_type, typename index_type>
measurement_class gpu_bcsr_jit_spmv(bcsr_matrix_class<data_type, index_type> &matrix,
                                    const data_type *transpose_matrix_data, const data_type *x,
                                    const data_type *reference_y);

#endif
