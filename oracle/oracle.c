/*
 * oracle.c -- CPU restatement of the reference's SpMV / BiCGStab path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Built with -ffp-contract=off so
 * that every a*b+c is a rounded multiply followed by a rounded add, the way
 * the reference's host code compiles with g++ -O3 on x86-64 (no FMA).
 *
 * Pinning (what checks each function, beyond restating its formula):
 *   orc_gen_n_diag_bcsr  golden vectors from the compiled reference header
 *                        (tests/golden/ref_host.json, oracle/_ref)
 *   orc_bcsr_to_csr      golden vectors from the compiled reference header
 *   orc_transpose_blocks golden vectors + involution
 *   orc_csr_to_bcsr      round trip with orc_bcsr_to_csr (bit-exact)
 *   orc_csr_spmv         closed form of y for x = 1 on gen_n_diag_bcsr
 *                        matrices; brute-force dense A*x on tiny inputs
 *   orc_bcsr_spmv        bit-identical to orc_csr_spmv on the CSR expansion
 *   orc_gen_stencil9     structure closed forms (nnzb = (3nx-2)(3ny-2)),
 *                        diagonal dominance, K symmetry / C skew identities
 *   orc_bicgstab         numpy.linalg.solve on small systems, x* = 1 for
 *                        b = A*1, A = I one-step convergence, iteration counts
 *   orc_compare_metric   hand-computed cases and the golden compare_results
 *                        output of the compiled reference header
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* RNG: SplitMix64 finaliser applied to a (seed, stream)-keyed counter.      */
/* ------------------------------------------------------------------------ */
#define ORC_GOLDEN 0x9E3779B97F4A7C15ULL

static uint64_t orc_mix(uint64_t z)
{
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

uint64_t orc_rng_u64(uint64_t seed, uint64_t stream, uint64_t index)
{
  const uint64_t key = orc_mix(seed * ORC_GOLDEN + stream);
  return orc_mix(key + (index + 1) * ORC_GOLDEN);
}

double orc_rng_unit(uint64_t seed, uint64_t stream, uint64_t index)
{
  return (double) (orc_rng_u64(seed, stream, index) >> 11) * 0x1.0p-53;
}

double orc_rng_sym(uint64_t seed, uint64_t stream, uint64_t index)
{
  return 2.0 * orc_rng_unit(seed, stream, index) - 1.0;
}

void orc_rng_fill(uint64_t seed, uint64_t stream, int kind, int64_t n, double *out)
{
  /* element i = orc_rng_unit (kind 0) or orc_rng_sym (kind 1) of index i */
  for (int64_t i = 0; i < n; i++)
    out[i] = kind == 0 ? orc_rng_unit(seed, stream, (uint64_t) i) : orc_rng_sym(seed, stream, (uint64_t) i);
}

/* ------------------------------------------------------------------------ */
/* gen_n_diag_bcsr: matrix_converters.h:157-196.                             */
/* ------------------------------------------------------------------------ */
void orc_gen_n_diag_bcsr(int n_rows, int blocks_per_row, int bs,
                         int *row_ptr, int *columns, double *values)
{
  int block_id = 0;
  for (int row = 0; row < n_rows; row++) {
    row_ptr[row] = row * blocks_per_row;
    int first_column;                     /* :176-183 */
    if (row < blocks_per_row / 2)
      first_column = 0;
    else if (row > n_rows - blocks_per_row)
      first_column = n_rows - blocks_per_row;
    else
      first_column = row - blocks_per_row / 2;
    for (int e = 0; e < blocks_per_row; e++) {
      const int col = first_column + e;
      double *blk = values + (size_t) block_id * bs * bs;
      for (unsigned i = 0; i < (unsigned) (bs * bs); i++)   /* :188-189 */
        blk[i] = ((double) col + (double) i) / (double) n_rows;
      columns[block_id++] = col;
    }
  }
  row_ptr[n_rows] = n_rows * blocks_per_row;
}

/* ------------------------------------------------------------------------ */
/* 9-point block stencil generator (synthetic inputs; no reference code).    */
/* ------------------------------------------------------------------------ */
enum { ORC_STREAM_DD = 1, ORC_STREAM_K = 2, ORC_STREAM_C = 3 };

int64_t orc_stencil_nnzb(int nx, int ny)
{
  return (int64_t) (3 * (int64_t) nx - 2) * (3 * (int64_t) ny - 2);
}

/* Random entry (r, c) of the block at stencil slot s of block row i. */
static double orc_stencil_entry(int kind, uint64_t seed, int64_t i, int s,
                                int bs, int r, int c)
{
  const uint64_t base = ((uint64_t) i * 9u + (uint64_t) s) * (uint64_t) bs * (uint64_t) bs;
  if (kind == 0)
    return orc_rng_sym(seed, ORC_STREAM_DD, base + (uint64_t) r * bs + c);
  /* kind 1: -(K + 0.3 C), K symmetric uniform[0,1], C skew uniform[-1,1] */
  const int lo = r < c ? r : c, hi = r < c ? c : r;
  const double k = orc_rng_unit(seed, ORC_STREAM_K, base + (uint64_t) lo * bs + hi);
  double cc = 0.0;
  if (r < c)
    cc = orc_rng_sym(seed, ORC_STREAM_C, base + (uint64_t) r * bs + c);
  else if (r > c)
    cc = -orc_rng_sym(seed, ORC_STREAM_C, base + (uint64_t) c * bs + r);
  const double t = 0.3 * cc;
  return -(k + t);
}

void orc_gen_stencil9(int nx, int ny, int bs, int kind, uint64_t seed,
                      int *row_ptr, int *columns, double *values)
{
  const int64_t bs2 = (int64_t) bs * bs;
  int64_t nb = 0;
  row_ptr[0] = 0;
  for (int iy = 0; iy < ny; iy++)
    for (int ix = 0; ix < nx; ix++) {
      const int64_t i = (int64_t) iy * nx + ix;
      const int64_t row_first = nb;
      int diag_block = -1;
      /* pass 1: pattern and raw entries, slots in ascending column order */
      for (int s = 0; s < 9; s++) {
        const int dy = s / 3 - 1, dx = s % 3 - 1;
        const int jy = iy + dy, jx = ix + dx;
        if (jy < 0 || jy >= ny || jx < 0 || jx >= nx)
          continue;
        columns[nb] = jy * nx + jx;
        double *blk = values + nb * bs2;
        if (s == 4) {
          diag_block = (int) (nb - row_first);
          for (int r = 0; r < bs; r++)
            for (int c = 0; c < bs; c++)
              blk[r * bs + c] = kind == 0 ? orc_stencil_entry(0, seed, i, s, bs, r, c) : 0.0;
        } else {
          for (int r = 0; r < bs; r++)
            for (int c = 0; c < bs; c++)
              blk[r * bs + c] = orc_stencil_entry(kind, seed, i, s, bs, r, c);
        }
        nb++;
      }
      /* pass 2: diagonal shift from the absolute off-diagonal row sum */
      double *dblk = values + (row_first + diag_block) * bs2;
      for (int r = 0; r < bs; r++) {
        double sum = 0.0;
        for (int64_t b = row_first; b < nb; b++)
          for (int c = 0; c < bs; c++) {
            if (b == row_first + diag_block && c == r)
              continue;
            sum += fabs(values[b * bs2 + r * bs + c]);
          }
        if (kind == 0) {
          const double shift = sum + 1.0;
          dblk[r * bs + r] = dblk[r * bs + r] + shift;
        } else {
          const double scaled = 1.02 * sum;
          dblk[r * bs + r] = scaled + 1.0;
        }
      }
      row_ptr[i + 1] = (int) nb;
    }
}

/* ------------------------------------------------------------------------ */
/* Conversions.                                                              */
/* ------------------------------------------------------------------------ */
void orc_transpose_blocks(int n_rows, int bs, const int *row_ptr,
                          const double *values, double *new_values)
{
  /* matrix_converters.h:32-49: new[j*bs+i] = old[i*bs+j] per block */
  const size_t bs2 = (size_t) bs * bs;
  for (int row = 0; row < n_rows; row++)
    for (int b = row_ptr[row]; b < row_ptr[row + 1]; b++) {
      const double *o = values + bs2 * b;
      double *n = new_values + bs2 * b;
      for (int i = 0; i < bs; i++)
        for (int j = 0; j < bs; j++)
          n[j * bs + i] = o[i * bs + j];
    }
}

void orc_bcsr_to_csr(int n_rows, int bs, const int *brow_ptr, const int *bcols,
                     const double *bvalues, int *row_ptr, int *cols, double *values)
{
  /* matrix_converters.h:88-117 */
  size_t off = 0;
  for (int br = 0; br < n_rows; br++)
    for (int r = 0; r < bs; r++) {
      row_ptr[br * bs + r] = (int) off;
      for (int b = brow_ptr[br]; b < brow_ptr[br + 1]; b++)
        for (int c = 0; c < bs; c++) {
          cols[off] = bcols[b] * bs + c;
          values[off] = bvalues[(size_t) b * bs * bs + r * bs + c];
          off++;
        }
    }
  row_ptr[n_rows * bs] = (int) off;
}

static int orc_cmp_int(const void *a, const void *b)
{
  const int x = *(const int *) a, y = *(const int *) b;
  return (x > y) - (x < y);
}

/* Distinct block columns of block row br, sorted, written to out; returns count. */
static int orc_block_cols_of_row(int br, int bs, const int *row_ptr,
                                 const int *cols, int *scratch, int *out)
{
  int m = 0;
  for (int r = 0; r < bs; r++)
    for (int e = row_ptr[br * bs + r]; e < row_ptr[br * bs + r + 1]; e++)
      scratch[m++] = cols[e] / bs;
  qsort(scratch, (size_t) m, sizeof(int), orc_cmp_int);
  int u = 0;
  for (int k = 0; k < m; k++)
    if (u == 0 || out[u - 1] != scratch[k])
      out[u++] = scratch[k];
  return u;
}

static int orc_max_block_row_nnz(int nb, int bs, const int *row_ptr)
{
  int mx = 0;
  for (int br = 0; br < nb; br++) {
    const int m = row_ptr[(br + 1) * bs] - row_ptr[br * bs];
    if (m > mx) mx = m;
  }
  return mx;
}

int orc_csr_to_bcsr_count(int n, int bs, const int *row_ptr, const int *cols)
{
  if (bs <= 0 || n % bs) return -1;
  const int nb = n / bs;
  const int mx = orc_max_block_row_nnz(nb, bs, row_ptr);
  int *scratch = (int *) malloc(sizeof(int) * (size_t) (mx + 1));
  int *out = (int *) malloc(sizeof(int) * (size_t) (mx + 1));
  int total = 0;
  for (int br = 0; br < nb; br++)
    total += orc_block_cols_of_row(br, bs, row_ptr, cols, scratch, out);
  free(scratch);
  free(out);
  return total;
}

int orc_csr_to_bcsr(int n, int bs, const int *row_ptr, const int *cols,
                    const double *values, int *brow_ptr, int *bcols,
                    double *bvalues, int *nnzb_out)
{
  if (bs <= 0 || n % bs) return -1;
  const int nb = n / bs;
  const size_t bs2 = (size_t) bs * bs;
  const int mx = orc_max_block_row_nnz(nb, bs, row_ptr);
  int *scratch = (int *) malloc(sizeof(int) * (size_t) (mx + 1));
  int nnzb = 0;
  brow_ptr[0] = 0;
  for (int br = 0; br < nb; br++) {
    const int u = orc_block_cols_of_row(br, bs, row_ptr, cols, scratch, bcols + nnzb);
    memset(bvalues + bs2 * nnzb, 0, sizeof(double) * bs2 * (size_t) u);
    for (int r = 0; r < bs; r++)
      for (int e = row_ptr[br * bs + r]; e < row_ptr[br * bs + r + 1]; e++) {
        const int bc = cols[e] / bs;
        int lo = 0, hi = u;           /* lower_bound, as matrix_converters.h:61-70 */
        while (lo < hi) {
          const int mid = (lo + hi) / 2;
          if (bcols[nnzb + mid] < bc) lo = mid + 1; else hi = mid;
        }
        bvalues[bs2 * (size_t) (nnzb + lo) + (size_t) r * bs + cols[e] % bs] = values[e];
      }
    nnzb += u;
    brow_ptr[br + 1] = nnzb;
  }
  free(scratch);
  if (nnzb_out) *nnzb_out = nnzb;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* SpMV.                                                                     */
/* ------------------------------------------------------------------------ */
void orc_csr_spmv(int n_rows, const int *row_ptr, const int *cols,
                  const double *values, const double *x, double *y)
{
  /* main.cpp:36-45 */
  for (int row = 0; row < n_rows; row++) {
    double dot = 0;
    for (int e = row_ptr[row]; e < row_ptr[row + 1]; e++)
      dot += values[e] * x[cols[e]];
    y[row] = dot;
  }
}

void orc_csr_absmv(int n_rows, const int *row_ptr, const int *cols,
                   const double *values, const double *x, double *y)
{
  for (int row = 0; row < n_rows; row++) {
    double dot = 0;
    for (int e = row_ptr[row]; e < row_ptr[row + 1]; e++)
      dot += fabs(values[e]) * fabs(x[cols[e]]);
    y[row] = dot;
  }
}

static double orc_bval(const double *values, int bs, int layout, size_t b, int r, int c)
{
  /* gpu_matrix_multiplier.cu:299 (row major) and :365 (column major) */
  const size_t bs2 = (size_t) bs * bs;
  return layout == 0 ? values[b * bs2 + (size_t) r * bs + c]
                     : values[b * bs2 + (size_t) c * bs + r];
}

void orc_bcsr_spmv(int n_rows, int bs, int layout, const int *row_ptr,
                   const int *cols, const double *values, const double *x,
                   double *y)
{
  for (int br = 0; br < n_rows; br++)
    for (int r = 0; r < bs; r++) {
      double dot = 0;
      for (int b = row_ptr[br]; b < row_ptr[br + 1]; b++)
        for (int c = 0; c < bs; c++)
          dot += orc_bval(values, bs, layout, (size_t) b, r, c) * x[(size_t) cols[b] * bs + c];
      y[(size_t) br * bs + r] = dot;
    }
}

void orc_bcsr_absmv(int n_rows, int bs, int layout, const int *row_ptr,
                    const int *cols, const double *values, const double *x,
                    double *y)
{
  for (int br = 0; br < n_rows; br++)
    for (int r = 0; r < bs; r++) {
      double dot = 0;
      for (int b = row_ptr[br]; b < row_ptr[br + 1]; b++)
        for (int c = 0; c < bs; c++)
          dot += fabs(orc_bval(values, bs, layout, (size_t) b, r, c)) * fabs(x[(size_t) cols[b] * bs + c]);
      y[(size_t) br * bs + r] = dot;
    }
}

/* ------------------------------------------------------------------------ */
/* BiCGStab.                                                                 */
/* ------------------------------------------------------------------------ */
static double orc_dot(int n, const double *a, const double *b)
{
  double s = 0;                 /* dot_product, bicgstab.cu:156-174 */
  for (int i = 0; i < n; i++)
    s += a[i] * b[i];
  return s;
}

void orc_bicgstab(int n, const int *row_ptr, const int *cols,
                  const double *values, const double *b, double eps,
                  int max_iterations, int use_jacobi, double *x_out,
                  orc_solve_info *info)
{
  const size_t sz = sizeof(double) * (size_t) n;
  double *x[2], *r[2], *v[2], *p[2];
  for (int k = 0; k < 2; k++) {
    x[k] = (double *) calloc((size_t) n, sizeof(double));
    r[k] = (double *) calloc((size_t) n, sizeof(double));
    v[k] = (double *) calloc((size_t) n, sizeof(double));
    p[k] = (double *) calloc((size_t) n, sizeof(double));
  }
  double *h = (double *) calloc((size_t) n, sizeof(double));
  double *s = (double *) calloc((size_t) n, sizeof(double));
  double *t = (double *) calloc((size_t) n, sizeof(double));
  double *rh = (double *) malloc(sz);
  double *P = NULL, *q = NULL, *z = NULL;

  if (use_jacobi) {             /* calculate_jacobi_preconditioner, :201-225 */
    P = (double *) malloc(sz);
    q = (double *) malloc(sz);
    z = (double *) malloc(sz);
    for (int i = 0; i < n; i++) {
      P[i] = 1.0;
      for (int e = row_ptr[i]; e < row_ptr[i + 1]; e++)
        if (cols[e] == i) {
          P[i] = fabs(values[e]) < 1e-20 ? 1.0 : 1.0 / values[e];
          break;
        }
    }
  }

  const double norm_b = sqrt(orc_dot(n, b, b));           /* :408-412 */
  memcpy(r[0], b, sz);                                    /* init, :257-284 */
  memcpy(rh, b, sz);
  double omega[2] = {1.0, 1.0}, rho[2] = {1.0, 1.0}, alpha = 1.0;

  int it = 0, status = 1, cur = 0;
  double rel = norm_b > 0 ? 1.0 : 0.0;
  if (norm_b == 0.0) {          /* b = 0: x = 0 is exact */
    status = 0;
  } else {
    while (it < max_iterations) {                         /* :426-513 */
      const int pv = it % 2;
      it++;
      cur = it % 2;
      const double rhoc = orc_dot(n, rh, r[pv]);          /* 1 */
      if (rhoc == 0.0) { status = 2; cur = pv; it--; break; }
      const double beta = rhoc / rho[pv] * alpha / omega[pv];     /* 2 */
      for (int i = 0; i < n; i++)                          /* 3 */
        p[cur][i] = r[pv][i] + beta * (p[pv][i] - omega[pv] * v[pv][i]);
      const double *qq = p[cur];
      if (P) { for (int i = 0; i < n; i++) q[i] = P[i] * p[cur][i]; qq = q; }
      orc_csr_spmv(n, row_ptr, cols, values, qq, v[cur]);  /* 4 */
      const double rhv = orc_dot(n, rh, v[cur]);           /* 5 */
      if (rhv == 0.0) { status = 2; cur = pv; it--; break; }
      alpha = rhoc / rhv;
      for (int i = 0; i < n; i++) {                        /* 6, 7 */
        h[i] = x[pv][i] + alpha * qq[i];
        s[i] = r[pv][i] - alpha * v[cur][i];
      }
      const double *zz = s;
      if (P) { for (int i = 0; i < n; i++) z[i] = P[i] * s[i]; zz = z; }
      orc_csr_spmv(n, row_ptr, cols, values, zz, t);       /* 8 */
      const double ts = orc_dot(n, t, s), tt = orc_dot(n, t, t);   /* 9 */
      double omegac;
      if (tt == 0.0) {
        /* t = A z = 0 => s = 0: h is the exact solution of this step */
        omegac = 0.0;
      } else {
        omegac = ts / tt;
      }
      rho[cur] = rhoc;
      omega[cur] = omegac;
      for (int i = 0; i < n; i++) {                        /* 10, 11 */
        r[cur][i] = s[i] - omegac * t[i];
        x[cur][i] = h[i] + omegac * zz[i];
      }
      rel = sqrt(orc_dot(n, r[cur], r[cur])) / norm_b;     /* 12 */
      if (rel < eps) { status = 0; break; }
      if (tt == 0.0 || omegac == 0.0) { status = 2; break; }
    }
  }

  memcpy(x_out, x[cur], sz);
  if (info) { info->iterations = it; info->status = status; info->rel_residual = rel; }

  for (int k = 0; k < 2; k++) { free(x[k]); free(r[k]); free(v[k]); free(p[k]); }
  free(h); free(s); free(t); free(rh); free(P); free(q); free(z);
}

double orc_compare_metric(int n, const double *a, const double *b)
{
  double num = 0.0, den = 0.0;      /* measurement_class.h:63-70 */
  for (int i = 0; i < n; i++) {
    num += (a[i] - b[i]) * (a[i] - b[i]);
    den += b[i] * b[i];
  }
  return num / den;
}
