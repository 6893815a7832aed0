/*
 * oracle.h -- CPU restatement of the reference's SpMV / BiCGStab path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links this code;
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it,
 * and only as the checker (or as the timed CPU baseline), never as the thing
 * that is measured on the GPU or shipped.
 *
 * Every function cites the reference file:line it restates.  Indices are
 * int32 (the reference's index_type = int), values double unless stated.
 */
#ifndef BMFP_ORACLE_H
#define BMFP_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- counter-based synthetic-input RNG (no reference counterpart) ------- */
uint64_t orc_rng_u64(uint64_t seed, uint64_t stream, uint64_t index);
double   orc_rng_unit(uint64_t seed, uint64_t stream, uint64_t index);  /* [0,1)  */
double   orc_rng_sym(uint64_t seed, uint64_t stream, uint64_t index);   /* [-1,1) */
/* out[i] = orc_rng_unit (kind 0) or orc_rng_sym (kind 1) of (seed, stream, i) */
void     orc_rng_fill(uint64_t seed, uint64_t stream, int kind, int64_t n, double *out);

/* ---- generators ---------------------------------------------------------- */
/* gen_n_diag_bcsr, matrix_converters.h:157-196 (double values). */
void orc_gen_n_diag_bcsr(int n_rows, int blocks_per_row, int bs,
                         int *row_ptr, int *columns, double *values);

/* 9-point block stencil on an nx x ny node grid.  kind 0: diagonally
 * dominant nonsymmetric (uniform[-1,1] blocks, diagonal shifted by the
 * absolute row sum + 1).  kind 1: convective (off-diagonal -(K+0.3C), K
 * symmetric uniform[0,1], C skew uniform[-1,1]; diagonal 1.02*rowsum + 1).
 * nnzb = stencil_nnzb(nx, ny).  Values row-major inside each block. */
int64_t orc_stencil_nnzb(int nx, int ny);
void orc_gen_stencil9(int nx, int ny, int bs, int kind, uint64_t seed,
                      int *row_ptr, int *columns, double *values);

/* ---- format conversions -------------------------------------------------- */
/* transpose_blocks, matrix_converters.h:32-49 */
void orc_transpose_blocks(int n_rows, int bs, const int *row_ptr,
                          const double *values, double *new_values);
/* csr_matrix_class(const bcsr_matrix_class&), matrix_converters.h:88-117 */
void orc_bcsr_to_csr(int n_rows, int bs, const int *brow_ptr, const int *bcols,
                     const double *bvalues, int *row_ptr, int *cols, double *values);
/* inverse of the above (no reference counterpart).  Returns 0 on success,
 * -1 if the CSR pattern is not made of whole bs x bs blocks. */
int orc_csr_to_bcsr(int n, int bs, const int *row_ptr, const int *cols,
                    const double *values, int *brow_ptr, int *bcols,
                    double *bvalues, int *nnzb_out);
int orc_csr_to_bcsr_count(int n, int bs, const int *row_ptr, const int *cols);

/* ---- SpMV ---------------------------------------------------------------- */
/* cpu_csr_spmv_single_thread_naive, main.cpp:36-45, with x as an input
 * (the reference overwrites x with 1.0 at main.cpp:27). */
void orc_csr_spmv(int n_rows, const int *row_ptr, const int *cols,
                  const double *values, const double *x, double *y);
/* |A||x| per row, the scale of the per-row parity tolerance. */
void orc_csr_absmv(int n_rows, const int *row_ptr, const int *cols,
                   const double *values, const double *x, double *y);
/* BCSR SpMV in scalar CSR order (layout 0 = row-major blocks, 1 = column
 * major, indexing gpu_matrix_multiplier.cu:299 / :365). */
void orc_bcsr_spmv(int n_rows, int bs, int layout, const int *row_ptr,
                   const int *cols, const double *values, const double *x,
                   double *y);
void orc_bcsr_absmv(int n_rows, int bs, int layout, const int *row_ptr,
                    const int *cols, const double *values, const double *x,
                    double *y);

/* ---- BiCGStab (bicgstab.cu:388-513; B3/B4 Jacobi at :201-238) ------------ */
typedef struct {
  int iterations;       /* iterations performed                          */
  int status;           /* 0 converged, 1 max_iterations, 2 breakdown     */
  double rel_residual;  /* recurrence ||r||/||b|| at exit                 */
} orc_solve_info;

/* Solves A x = b for a CSR matrix.  x0 = 0.  Returns the final iterate in x
 * (the reference returns buffer 0, bicgstab.cu:515, a stale iterate after an
 * odd iteration count; this returns the current one).  Breakdown guards:
 * (t,t) == 0 means s == 0 (A nonsingular): x = h, r = s, converged; a zero
 * (r^,v) or rho is reported as breakdown instead of producing NaN. */
void orc_bicgstab(int n, const int *row_ptr, const int *cols,
                  const double *values, const double *b, double eps,
                  int max_iterations, int use_jacobi, double *x,
                  orc_solve_info *info);

/* ---- measurement arithmetic (measurement_class.h:29-49, .cpp:7-16) ------- */
/* compare_results metric, measurement_class.h:60-70: sum (a-b)^2 / sum b^2 */
double orc_compare_metric(int n, const double *a, const double *b);

#ifdef __cplusplus
}
#endif
#endif
