/*
 * bmfp.h -- C ABI of the MI355X (gfx950) block-sparse SpMV / BiCGStab library.
 *
 * The reference exposes C++ templates with CUDA inside them; this header is the
 * plain-C boundary underneath the drop-in C++ headers in include/bmfp/.  Each
 * entry point names the reference interface it replaces.
 *
 * Conventions
 *   - Every function returns an int status: BMFP_OK (0) or a negative code.
 *     No exception crosses the ABI and nothing calls exit().
 *     bmfp_last_error() returns the calling thread's last error text.
 *   - Device pointers are caller-owned unless a function says otherwise.
 *   - Calls taking a bmfp_stream are ordered on that stream (NULL = the
 *     device's null stream) and return without synchronising unless noted.
 *   - Indices are int32 (the reference's index_type = int); n_rows, nnz and
 *     nnzb*bs*bs must fit in int32.
 */
#ifndef BMFP_H
#define BMFP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BMFP_VERSION 100

/* ---- status codes ------------------------------------------------------- */
enum {
  BMFP_OK = 0,
  BMFP_ERR_INVALID = -1,        /* bad argument / shape                      */
  BMFP_ERR_UNSUPPORTED = -2,    /* e.g. a block size with no kernel          */
  BMFP_ERR_NO_MEMORY = -3,      /* device (or host) allocation failed        */
  BMFP_ERR_BREAKDOWN = -4,      /* BiCGStab breakdown ((r^,v)=0, rho=0, ...) */
  BMFP_ERR_NOT_CONVERGED = -5,  /* max_iterations reached                    */
  BMFP_ERR_HIP = -6,            /* HIP runtime error                         */
  BMFP_ERR_NONFINITE = -7,      /* NaN/Inf in inputs or iterates             */
  BMFP_ERR_BUDGET = -8,         /* memory-budget admission refused           */
  BMFP_ERR_NO_DEVICE = -9       /* no usable gfx950 device                   */
};

/* ---- element types and layouts ----------------------------------------- */
enum { BMFP_F64 = 0, BMFP_F32 = 1, BMFP_F16 = 2 };
enum { BMFP_ROW_MAJOR = 0, BMFP_COL_MAJOR = 1 };

/* ---- SpMV kernel families (reference kernel names in brackets) ---------- */
enum {
  BMFP_SPMV_AUTO = 0,            /* best family for (format, layout, bs)             */
  BMFP_CSR_SCALAR = 1,           /* [csr_spmv_kernel] gpu_matrix_multiplier.cu:65-86  */
  BMFP_CSR_VECTOR = 2,           /* [csr_spmv_vector_kernel] :165-196; lane groups    */
                                 /*  of 2..64 lanes chosen from the mean row length   */
  BMFP_BCSR_THREAD_PER_ROW = 3,  /* [..._thread_per_row_{row,column}_major_matrix     */
                                 /*  (_template)] :275-304, :341-400                  */
  BMFP_BCSR_GROUP_PER_ROW = 4,   /* [..._warp_per_row_row_major_matrix] :306-339      */
  BMFP_BCSR_SHUFFLE_X = 5,       /* [..._coal_x(_template)] :402-469, x segments      */
                                 /*  shared across a block row by lane shuffles       */
  BMFP_BCSR_BLOCK_ROW_GROUP = 6, /* [..._column_by_column(_fastdiv,_po2,_template)]   */
                                 /*  :485-673, lane group per block row               */
  BMFP_BCSR_LIBRARY = 7,         /* [cusparse?bsrmv] :676-734 -> rocSPARSE bsrmv      */
  BMFP_BCSR_JIT = 8,             /* [bcsr_jit, NVRTC] main.cpp:170-294 -> a kernel    */
                                 /*  compiled by hipRTC for A's bs (F64 / F32, either */
                                 /*  layout); the AUTO choice for bs outside the      */
                                 /*  precompiled {1,2,3,4,6,8,12,16,24,32}            */
  BMFP_CSR_MERGE = 9             /* ["GPU Merge SpMV" rows of results/bs_range.json;  */
                                 /*  no kernel in the tree] merge-path tiles: equal   */
                                 /*  work per wavefront for any row lengths.  Uses    */
                                 /*  the matrix's carry buffer: merge SpMVs on one    */
                                 /*  handle must not overlap in time                  */
};

typedef struct bmfp_matrix_t *bmfp_matrix;
typedef struct bmfp_solver_t *bmfp_solver;
typedef void *bmfp_stream;     /* a hipStream_t */

/* ---- diagnostics --------------------------------------------------------- */
const char *bmfp_last_error(void);
const char *bmfp_status_string(int status);
int bmfp_version(void);

/* ---- device, memory, streams -------------------------------------------
 * Replace the raw cudaSetDevice / cudaMalloc / cudaMemcpy / cudaFree /
 * cudaDeviceSynchronize calls made inline by main.cpp:388 and
 * gpu_matrix_multiplier.cu:88-163 (and every other wrapper). */
int bmfp_init(int device);
int bmfp_device_count(int *count);
int bmfp_device_info(char *name, size_t name_len, size_t *total_mem_bytes,
                     int *compute_units, int *gcn_arch_major_minor);
int bmfp_malloc(void **ptr, size_t bytes);
int bmfp_free(void *ptr);
/* Pinned (page-locked) host memory, as the reference's cudaMallocHost for
 * the solver result (bicgstab.cu:62). */
int bmfp_malloc_host(void **ptr, size_t bytes);
int bmfp_free_host(void *ptr);
int bmfp_memcpy_h2d(void *dst, const void *src, size_t bytes);
int bmfp_memcpy_d2h(void *dst, const void *src, size_t bytes);
int bmfp_memcpy_d2d(void *dst, const void *src, size_t bytes);
int bmfp_memset(void *dst, int value, size_t bytes);
int bmfp_stream_create(bmfp_stream *stream);
int bmfp_stream_destroy(bmfp_stream stream);
int bmfp_stream_synchronize(bmfp_stream stream);
int bmfp_device_synchronize(void);
/* Timing events on a stream (hipEvent_t underneath).  bmfp_event_elapsed
 * waits for `stop` and returns the seconds between the two records. */
typedef void *bmfp_event;
int bmfp_event_create(bmfp_event *event);
int bmfp_event_destroy(bmfp_event event);
int bmfp_event_record(bmfp_event event, bmfp_stream stream);
int bmfp_event_elapsed(bmfp_event start, bmfp_event stop, double *seconds);
/* Device-to-device copy by a plain 16-byte-per-lane streaming kernel (not
 * the DMA engines): the measured-bandwidth yardstick for the SpMV kernels.
 * bytes must be a multiple of 16 and both pointers 16-byte aligned. */
int bmfp_stream_copy(void *dst, const void *src, size_t bytes, bmfp_stream stream);
/* Reads `bytes` of device memory once with the same streaming loads and
 * writes nothing but (practically never) one word to `sink` (device, 4-byte
 * aligned): the read-only bandwidth yardstick.  No counterpart in the
 * reference; it states what an HBM-read-bound kernel can reach on the GPU. */
int bmfp_stream_read(const void *src, size_t bytes, void *sink, bmfp_stream stream);
/* Deterministic device fill with the library's counter-based RNG:
 * kind 0 uniform[0,1), kind 1 uniform[-1,1), kind 2 constant `value`.
 * Element i of a vector filled with (seed, stream_id) is identical to the
 * host generator's element i.  dtype is BMFP_F64 or BMFP_F32. */
int bmfp_fill(void *dst, int dtype, size_t n, int kind, uint64_t seed,
              uint64_t stream_id, double value, bmfp_stream stream);

/* ---- host-side synthetic matrices ---------------------------------------
 * gen_n_diag_bcsr: matrix_converters.h:157-196 (values double, row-major).
 * Arrays: row_ptr[n_rows+1], columns[n_rows*blocks_per_row],
 * values[n_rows*blocks_per_row*bs*bs]. */
int bmfp_gen_n_diag_bcsr(int n_rows, int blocks_per_row, int bs,
                         int *row_ptr, int *columns, double *values);
/* 9-point block stencil on an nx*ny node grid (no reference counterpart:
 * the synthetic workloads' generator).  kind 0: diagonally dominant,
 * nonsymmetric; kind 1: convective -(K+0.3C) off-diagonals.  Values are
 * row-major inside blocks.  nnzb = bmfp_stencil_nnzb(nx, ny). */
int64_t bmfp_stencil_nnzb(int nx, int ny);
int bmfp_gen_stencil9(int nx, int ny, int bs, int kind, uint64_t seed,
                      int *row_ptr, int *columns, double *values);

/* ---- device-resident matrices -------------------------------------------
 * Replace the per-call cudaMalloc + cudaMemcpy of matrix arrays inside
 * gpu_csr_spmv / gpu_csr_vector_spmv / gpu_bcsr_spmv
 * (gpu_matrix_multiplier.cu:88-163, :198-273, :736-770) and the uploads in
 * gpu_bicgstab::solve (bicgstab.cu:403-405).  The handle owns its device
 * copies; host arrays can be freed after the call returns.
 *
 * bmfp_bcsr_create: `values` are host doubles in `layout`; `dtype` is the
 * device storage type.  BMFP_F16 stores each block row scaled by one FP32
 * factor (max |entry| of the block row maps to 1.0). */
int bmfp_bcsr_create(bmfp_matrix *out, int dtype, int layout, int n_rows,
                     int n_cols, int bs, int nnzb, const int *row_ptr,
                     const int *columns, const double *values);
int bmfp_csr_create(bmfp_matrix *out, int dtype, int n_rows, int n_cols,
                    int nnz, const int *row_ptr, const int *columns,
                    const double *values);
/* The 9-point block stencil of bmfp_gen_stencil9 generated directly in
 * device memory (bit-identical values), stored as dtype / layout.  No
 * reference counterpart: it builds the multi-GB synthetic workloads without
 * a host round trip. */
int bmfp_stencil_create(bmfp_matrix *out, int nx, int ny, int bs, int kind,
                        uint64_t seed, int dtype, int layout);
/* Regenerates, in place and ordered on `stream`, the column indices and
 * values of an F64 row-major matrix made by bmfp_stencil_create (same grid
 * and block size) for another kind / seed: a request's matrix built on the
 * device without an allocation (no reference counterpart). */
int bmfp_stencil_generate(bmfp_matrix m, int kind, uint64_t seed, bmfp_stream stream);
/* Device-side copy of a BCSR matrix into another storage type / layout
 * (indices shared by reference count).  Used for mixed precision. */
int bmfp_matrix_convert(bmfp_matrix src, int dtype, int layout,
                        bmfp_matrix *out, bmfp_stream stream);
int bmfp_matrix_destroy(bmfp_matrix m);
/* Shape query: any pointer may be NULL.  For CSR, bs = 1 and nnzb = nnz. */
int bmfp_matrix_shape(bmfp_matrix m, int *is_bcsr, int *dtype, int *layout,
                      int *n_rows, int *n_cols, int *bs, int *nnzb);
/* Compulsory HBM bytes and flops of one SpMV on this matrix: values once,
 * indices once, x once, y once (never the reference's CPU model, which
 * counts x per nonzero). */
int bmfp_spmv_traffic(bmfp_matrix m, double *bytes, double *flops);
/* ---- Matrix Market I/O and block-size detection (host only) ---------------
 * Replace common/mmio.{h,c} and csr_matrix_class::write_mm
 * (matrix_converters.h:119-144).  bmfp_mm_read parses coordinate (real,
 * integer, pattern; general, symmetric, skew-symmetric) and array (real,
 * integer; general) files into CSR: columns ascending within a row, duplicate
 * entries summed, the mirrored half of symmetric storage added, zeros of
 * array files dropped.  Call it with row_ptr == NULL for the sizes, then with
 * arrays of n_rows+1, nnz and nnz elements (the sizes as returned). */
int bmfp_mm_read(const char *path, int *n_rows, int *n_cols, int *nnz,
                 int *row_ptr, int *columns, double *values);
/* "coordinate real general", 1-based "%d %d %10.20g" entries in CSR order:
 * the text write_mm produces through mmio. */
int bmfp_mm_write(const char *path, int n_rows, int n_cols, const int *row_ptr,
                  const int *columns, const double *values);
/* Block size for BCSR storage of a CSR pattern (no reference counterpart):
 * of the candidates dividing both dimensions, the one whose BCSR SpMV with
 * value_bytes-wide values moves the fewest compulsory bytes (values, block
 * columns, block row pointers, x, y).  Also the block count and the fill
 * (stored values / nonzeros) of the choice. */
int bmfp_block_size_detect(int n_rows, int n_cols, const int *row_ptr,
                           const int *columns, int value_bytes,
                           const int *candidates, int n_candidates, int *bs,
                           int *nnzb, double *fill);

/* Download values (as doubles, in the handle's layout, unscaled). */
int bmfp_matrix_download_values(bmfp_matrix m, double *values);
/* Download row pointers (n_rows+1) and (block) column indices (nnzb). */
int bmfp_matrix_download_structure(bmfp_matrix m, int *row_ptr, int *columns);
/* CSR expansion of an F64 BCSR matrix in device memory, into F64 or F32
 * values: replaces csr_matrix_class(const bcsr_matrix_class&)
 * (matrix_converters.h:88-117) with the same entry order, bit-exact for F64. */
int bmfp_matrix_to_csr(bmfp_matrix src, int dtype, bmfp_matrix *out,
                       bmfp_stream stream);

/* ---- SpMV ----------------------------------------------------------------
 * y = A x.  x, y are device vectors of n_cols*bs / n_rows*bs elements in the
 * vector type of A (F64 for F64 matrices, F32 for F32 and F16 matrices).
 * variant: BMFP_SPMV_AUTO or a family valid for A's format
 * (CSR: BMFP_CSR_*; BCSR: BMFP_BCSR_*).  Replaces the kernel launches of
 * gpu_csr_spmv (:132-139), gpu_csr_vector_spmv (:242-249) and the 12 BCSR
 * variants of gpu_bcsr_spmv (:784-1305). */
int bmfp_spmv(bmfp_matrix A, int variant, const void *x, void *y,
              bmfp_stream stream);
/* `warmup` untimed, then `reps` SpMVs between two hipEvents recorded on
 * `stream`; *avg_seconds = elapsed / reps.  Replaces the single-launch
 * cudaEvent timing in every reference wrapper (e.g. :125-149). */
int bmfp_spmv_timed(bmfp_matrix A, int variant, const void *x, void *y,
                    int warmup, int reps, bmfp_stream stream,
                    double *avg_seconds);
/* Compiles with hipRTC (no device needed) the kernel BMFP_BCSR_JIT uses for
 * (bs, dtype, layout) on gfx950; reports its code size and compile time.
 * The SpMV entry points compile and cache the kernel themselves on first
 * use; this call checks the toolchain and measures the compile cost. */
int bmfp_jit_build(int bs, int dtype, int layout, size_t *code_bytes, double *seconds);
/* The family BMFP_SPMV_AUTO resolves to for A. */
int bmfp_spmv_auto_variant(bmfp_matrix A, int *variant);
/* Writes into buf (len bytes) the code-object symbol of the kernel that
 * bmfp_spmv(A, variant, ...) launches, mangled as `rocprofv3 -M` prints it
 * (for the merge path, its tile kernel); launches nothing.  No counterpart
 * in the reference: measurement support, so that counter profiles can be
 * matched to the kernel a run picked.  BMFP_ERR_UNSUPPORTED for the library
 * and run-time compiled families; BMFP_ERR_INVALID when A has no rows. */
int bmfp_spmv_kernel_symbol(bmfp_matrix A, int variant, char *buf, size_t len);

/* ---- BiCGStab ------------------------------------------------------------
 * Replaces gpu_bicgstab<T,C> (bicgstab.h:12-47, bicgstab.cu:23-526).  The
 * solver keeps its Krylov vectors and scalars resident; each iteration is
 * three fused kernels (two SpMVs with the vector updates formed on the fly,
 * one update kernel) whose dot products are reduced on the device in a fixed
 * order, so results are bitwise reproducible.  Convergence test as in the
 * reference: recurrence ||r||/||b|| < eps (bicgstab.cu:512).  Breakdowns
 * ((r^,v) = 0, rho = 0, omega = 0) return BMFP_ERR_BREAKDOWN with x the last
 * valid iterate instead of NaNs; t = 0 (s = 0, exact solution) converges.
 * The vector type follows the matrix: F64 matrices solve in FP64, F32 and F16
 * (per-block-row scaled) matrices in FP32 vectors with FP64 reductions. */
typedef struct {
  int iterations;           /* BiCGStab iterations (total inner ones for refinement)  */
  int status;               /* BMFP_OK, _NOT_CONVERGED, _BREAKDOWN or _NONFINITE       */
  double rel_residual;      /* recurrence ||r||/||b|| at exit                          */
  double true_rel_residual; /* ||b - A x||/||b|| recomputed in FP64 after the solve    */
  double seconds;           /* device time of the solve (hipEvents on its stream)      */
  int outer_iterations;     /* refinement steps (0 for plain BiCGStab)                 */
  int inner_iterations;     /* inner BiCGStab iterations summed over refinement steps  */
  int inner_dtype;          /* refinement: storage type of the last inner solve        */
  int escalations;          /* refinement: precision escalations (FP16 -> FP32)        */
} bmfp_solve_info;

/* use_jacobi: point-Jacobi preconditioning (bicgstab.cu:201-238). */
int bmfp_bicgstab_create(bmfp_solver *out, bmfp_matrix A, int use_jacobi);
/* b, x: device vectors of A's vector type; x is output only (x0 = 0, as
 * bicgstab_init_kernel, bicgstab.cu:257-284).  Synchronous: returns when the
 * solve has stopped.  Returns info->status (BMFP_OK when converged). */
int bmfp_bicgstab_solve(bmfp_solver s, const void *b, void *x, double eps,
                        int max_iterations, bmfp_stream stream,
                        bmfp_solve_info *info);
int bmfp_bicgstab_destroy(bmfp_solver s);
/* ---- mixed-precision iterative refinement (no reference counterpart) ----
 * Solves A x = b for an F64 BCSR matrix A:
 *   x_0 = 0;  r_k = b - A x_k   (FP64 matrix and vectors, fused SpMV kernel)
 *   stop when ||r_k||/||b|| < eps  (true residual, FP64)
 *   solve A_low d_k = r_k by BiCGStab to inner_tol (relative), FP32 vectors,
 *   A_low an FP32 copy of A (inner_dtype BMFP_F32) or FP16 values with one
 *   FP32 scale per block row (BMFP_F16);  x_{k+1} = x_k + d_k  (FP64).
 * If an FP16 inner solve breaks down, or the outer residual shrinks by less
 * than 2x in a step, the inner matrix escalates to FP32 (info->escalations).
 * b, x are device FP64 vectors.  Synchronous; returns info->status. */
int bmfp_refine_create(bmfp_solver *out, bmfp_matrix A64, int inner_dtype,
                       int use_jacobi);
int bmfp_refine_solve(bmfp_solver s, const double *b, double *x, double eps,
                      double inner_tol, int max_outer, int max_inner,
                      bmfp_stream stream, bmfp_solve_info *info);
/* Device time, in seconds, of the copy of A in inner_dtype (BMFP_F32 or
 * BMFP_F16) the handle made (0 if it has none): the one-time cost to add to
 * its solves.  No reference counterpart: measurement support. */
int bmfp_refine_copy_seconds(bmfp_solver s, int inner_dtype, double *seconds);
int bmfp_refine_destroy(bmfp_solver s);

/* ---- serving: continuous-batching request loop (no reference counterpart) --
 * Requests -- an FP64 BCSR system with 1..4 right-hand sides -- queue on the
 * host and are admitted first-come first-served while their device footprint
 * (FP64 matrix + FP32 copy, or FP16 copy and block-row scales, + 12 vectors
 * per right-hand side) fits the memory budget.  Admitted requests are set up on their own stream while earlier
 * ones compute, then join the batch: each request is solved by FP64
 * iterative refinement around FP32 BiCGStab, its right-hand sides in lockstep
 * (one read of the matrix per SpMV for all of them), all requests advancing
 * together through batched kernels replayed from a HIP graph
 * (iterations_per_round batched iterations per bmfp_server_step).  A
 * request's arithmetic is independent of the batch it shares and of the tile
 * schedule: results are bitwise identical alone and in any mix.  Block sizes
 * 1,2,3,4,6,8. */
typedef struct bmfp_server_t *bmfp_server;

typedef struct {
  size_t memory_budget_bytes; /* device bytes for admitted requests (capped at 90% free) */
  int max_systems;            /* resident requests (active + finished, not yet polled) */
  int iterations_per_round;   /* batched iterations per graph replay                   */
  int slot_groups;            /* slot-kernel grid: > 0 persistent workgroups walking the
                                 tile table; -1 one wavefront per tile (grid sized to
                                 the tables each round); 0: tuned                     */
  double inner_tol;           /* FP32 BiCGStab relative tolerance per refinement step;
                                 a step whose outer residual rel is near the request's
                                 tol stops at max(inner_tol, 0.6 tol / rel) instead   */
  int max_outer;              /* refinement steps per right-hand side                  */
  int max_inner;              /* inner iterations per refinement step                  */
  int chunks_per_tile;        /* tile height in 256-row chunks; 0: tuned per bs        */
  int use_jacobi;             /* 1: point-Jacobi (1/diagonal, right) preconditioning of
                                 the inner solves, as gpu_bicgstab(A, true); 0: none   */
  int inner_f16_min_bs;       /* requests with bs >= this (4, 6 or 8) keep their inner
                                 matrix in FP16 with one FP32 scale per block row
                                 instead of FP32 (default 4); 0: FP32 for every
                                 request                                              */
  double inner_tol_f16;       /* inner tolerance of FP16 requests; 0: 1e-3 for bs 8,
                                 1e-4 below                                            */
} bmfp_server_config;

typedef struct {
  /* explicit system (used when stencil_nx == 0): row-major blocks, FP64 */
  int n_rows, bs, nnzb;
  const int *row_ptr;
  const int *columns;
  const double *values;
  /* synthetic 9-point block stencil generated on the device, bit-identical
   * to bmfp_gen_stencil9(stencil_nx, stencil_ny, bs, kind, seed) */
  int stencil_nx, stencil_ny, stencil_kind;
  uint64_t stencil_seed;
  int nrhs;
  const double *b;   /* host, nrhs * n_rows*bs doubles; NULL: uniform[-1,1]   */
  uint64_t rhs_seed; /* (key rhs_seed, element k*n + i) for right-hand side k  */
  double tol;        /* relative residual ||b - A x|| / ||b|| (FP64)           */
} bmfp_request;

typedef struct {
  uint64_t id;
  int status;            /* worst status over the right-hand sides               */
  int nrhs;
  int64_t n;             /* rows per right-hand side                             */
  int outer_iterations;  /* max over right-hand sides                            */
  int inner_iterations;  /* max over right-hand sides (summed over outer steps)  */
  double rel_residual;   /* max FP64 true relative residual                      */
  double submit_time, admit_time, done_time;  /* host steady-clock seconds       */
} bmfp_request_result;

int bmfp_server_default_config(bmfp_server_config *cfg);
int bmfp_server_create(bmfp_server *out, const bmfp_server_config *cfg);
int bmfp_server_destroy(bmfp_server s);
/* Device bytes a request needs (the admission footprint). */
int bmfp_server_request_bytes(const bmfp_request *req, size_t *bytes);
/* Queues a request (copies explicit host arrays).  BMFP_ERR_BUDGET when it
 * could never fit the budget. */
int bmfp_server_submit(bmfp_server s, const bmfp_request *req, uint64_t *id);
/* Waits for the round of batched iterations the previous call launched and
 * retires the requests it finished (bmfp_server_poll then returns them),
 * admits what fits, launches the next round and returns while it runs, so
 * the caller's work until the next call overlaps it.  Reports the active
 * systems (those in the launched round) and the waiting requests; 0 and 0
 * when nothing is left to run.  Destroying the server waits for the round. */
int bmfp_server_step(bmfp_server s, int *n_active, int *n_queued);
/* Pops one finished request (FIFO by completion): result and, if x is not
 * NULL, its solutions (nrhs * n doubles, right-hand side major).  A finished
 * request keeps its device memory until it is polled. */
int bmfp_server_poll(bmfp_server s, bmfp_request_result *res, double *x,
                     size_t x_capacity, int *got);
/* The result bmfp_server_poll would return next, without removing it. */
int bmfp_server_peek(bmfp_server s, bmfp_request_result *res, int *got);
/* The tile schedule in use (from the configuration or the autotuner run by
 * bmfp_server_create): slot-kernel grid (as slot_groups in the configuration,
 * -1 = one wavefront per tile), tiles per run dealt to one XCD (0: table
 * order) and chunks per tile, the latter indexed by block size (9 entries, 0
 * for unsupported sizes). */
int bmfp_server_schedule(bmfp_server s, int *slot_groups, int *xcd_run,
                         int *chunks_per_tile_by_bs);
int bmfp_server_stats(bmfp_server s, size_t *used_bytes, size_t *budget_bytes,
                      int *active_systems, int *queued, double *busy_seconds);

/* Device bytes the solver for A holds beyond A itself (the reference's
 * ctor memory estimate, bicgstab.cu:30-45). */
int bmfp_bicgstab_workspace_bytes(bmfp_matrix A, int use_jacobi, size_t *bytes);

/* ---- allocation-free CSR SpMV ----------------------------------------------
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  y = A x by the merge-path kernels of BMFP_CSR_MERGE,
 * bitwise equal to bmfp_spmv(A, BMFP_CSR_MERGE, ...) on a handle holding the
 * same arrays, but over caller-owned device arrays (row_ptr[n_rows+1],
 * columns[nnz], values[nnz], x[n_cols], y[n_rows]; values, x and y of dtype
 * BMFP_F64 or BMFP_F32) and with the kernels' scratch (per-wavefront carries
 * and tile boundaries) in one fixed device array of BMFP_NOALLOC_SCRATCH_BYTES
 * that the library holds on each device it runs on until the process ends.
 * Calls no allocating routine.  n_rows + nnz above BMFP_NOALLOC_MAX_ITEMS is
 * refused with BMFP_ERR_INVALID.  Every call on a device shares that device's
 * scratch: calls that overlap in time on one device (two streams, two host
 * threads) race on it and must not be issued.  Stream-ordered. */
#define BMFP_NOALLOC_MAX_ITEMS (1 << 28)
#define BMFP_NOALLOC_SCRATCH_BYTES ((size_t) 8388864)
int bmfp_csr_spmv_noalloc(int dtype, int n_rows, int n_cols, int nnz, const int *row_ptr,
                          const int *columns, const void *values, const void *x, void *y,
                          bmfp_stream stream);

#ifdef __cplusplus
}
#endif
#endif
