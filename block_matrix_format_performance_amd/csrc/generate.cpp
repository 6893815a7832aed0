// generate.cpp -- host generators for synthetic BCSR matrices.  Compiled with
// -ffp-contract=off: every entry is a fixed sequence of rounded IEEE
// operations, so the same (grid, bs, kind, seed) gives the same bits on any
// host.
#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "common.hpp"

using namespace bmfp;

extern "C" int bmfp_gen_n_diag_bcsr(int n_rows, int blocks_per_row, int bs, int *row_ptr, int *columns,
                                    double *values)
{
  return bmfp::guarded("bmfp_gen_n_diag_bcsr", [&]() -> int {
    // matrix_converters.h:157-196
    if (n_rows < 1 || bs < 1 || blocks_per_row < 1 || blocks_per_row > n_rows)
      return fail(BMFP_ERR_INVALID, "gen_n_diag_bcsr: bad shape");
    if (!row_ptr || !columns || !values) return fail(BMFP_ERR_INVALID, "gen_n_diag_bcsr: NULL array");
    const size_t bs2 = (size_t) bs * bs;
    size_t block = 0;
    for (int row = 0; row < n_rows; row++) {
      row_ptr[row] = row * blocks_per_row;
      int first = row - blocks_per_row / 2;
      if (row < blocks_per_row / 2) first = 0;
      else if (row > n_rows - blocks_per_row) first = n_rows - blocks_per_row;
      for (int k = 0; k < blocks_per_row; k++, block++) {
        const int col = first + k;
        columns[block] = col;
        double *blk = values + block * bs2;
        for (size_t i = 0; i < bs2; i++) blk[i] = ((double) col + (double) i) / (double) n_rows;
      }
    }
    row_ptr[n_rows] = n_rows * blocks_per_row;
    return BMFP_OK;
  });
}

extern "C" int64_t bmfp_stencil_nnzb(int nx, int ny)
{
  if (nx < 1 || ny < 1) return 0;
  return (3 * (int64_t) nx - 2) * (3 * (int64_t) ny - 2);
}

namespace {

// Entry (r, c) of the block at stencil slot s (0..8, ascending column order)
// of block row i.  kind 0: uniform[-1,1].  kind 1: -(K + 0.3*C) with K
// symmetric uniform[0,1] and C skew uniform[-1,1] within the block.
inline double stencil_entry(int kind, uint64_t key_a, uint64_t key_b, int64_t i, int s, int bs, int r, int c)
{
  const uint64_t base = ((uint64_t) i * 9u + (uint64_t) s) * (uint64_t) bs * (uint64_t) bs;
  if (kind == 0) return rng_sym(key_a, base + (uint64_t) r * bs + c);
  const int lo = std::min(r, c), hi = std::max(r, c);
  const double k = rng_unit(key_a, base + (uint64_t) lo * bs + hi);
  double cc = 0.0;
  if (r < c) cc = rng_sym(key_b, base + (uint64_t) r * bs + c);
  else if (r > c) cc = -rng_sym(key_b, base + (uint64_t) c * bs + r);
  const double t = 0.3 * cc;
  return -(k + t);
}

void stencil_rows(int nx, int ny, int bs, int kind, uint64_t seed, int iy0, int iy1, const int *row_ptr,
                  int *columns, double *values)
{
  const uint64_t key_a = rng_key(seed, kind == 0 ? STREAM_DD : STREAM_K);
  const uint64_t key_b = rng_key(seed, STREAM_C);
  const int64_t bs2 = (int64_t) bs * bs;
  for (int iy = iy0; iy < iy1; iy++)
    for (int ix = 0; ix < nx; ix++) {
      const int64_t i = (int64_t) iy * nx + ix;
      int64_t nb = row_ptr[i];
      const int64_t first = nb;
      int64_t diag = -1;
      for (int s = 0; s < 9; s++) {
        const int jy = iy + s / 3 - 1, jx = ix + s % 3 - 1;
        if (jy < 0 || jy >= ny || jx < 0 || jx >= nx) continue;
        columns[nb] = jy * nx + jx;
        double *blk = values + nb * bs2;
        if (s == 4) {
          diag = nb;
          for (int r = 0; r < bs; r++)
            for (int c = 0; c < bs; c++)
              blk[r * bs + c] = kind == 0 ? stencil_entry(0, key_a, key_b, i, s, bs, r, c) : 0.0;
        } else {
          for (int r = 0; r < bs; r++)
            for (int c = 0; c < bs; c++) blk[r * bs + c] = stencil_entry(kind, key_a, key_b, i, s, bs, r, c);
        }
        nb++;
      }
      double *dblk = values + diag * bs2;
      for (int r = 0; r < bs; r++) {
        double sum = 0.0;
        for (int64_t b = first; b < nb; b++)
          for (int c = 0; c < bs; c++) {
            if (b == diag && c == r) continue;
            sum += std::fabs(values[b * bs2 + r * bs + c]);
          }
        if (kind == 0) {
          const double shift = sum + 1.0;
          dblk[r * bs + r] = dblk[r * bs + r] + shift;
        } else {
          const double scaled = 1.02 * sum;
          dblk[r * bs + r] = scaled + 1.0;
        }
      }
    }
}

}  // namespace

extern "C" int bmfp_gen_stencil9(int nx, int ny, int bs, int kind, uint64_t seed, int *row_ptr, int *columns,
                                 double *values)
{
  return bmfp::guarded("bmfp_gen_stencil9", [&]() -> int {
    if (nx < 1 || ny < 1 || bs < 1 || (kind != 0 && kind != 1))
      return fail(BMFP_ERR_INVALID, "gen_stencil9: bad arguments (nx %d ny %d bs %d kind %d)", nx, ny, bs, kind);
    if (!row_ptr || !columns || !values) return fail(BMFP_ERR_INVALID, "gen_stencil9: NULL array");
    if ((int64_t) nx * ny > INT32_MAX || bmfp_stencil_nnzb(nx, ny) * bs * bs > INT32_MAX)
      return fail(BMFP_ERR_INVALID, "gen_stencil9: matrix too large for int32 indices");
    // row pointers: count of in-grid neighbours (3x3 window clipped at edges)
    int64_t nb = 0;
    row_ptr[0] = 0;
    for (int iy = 0; iy < ny; iy++) {
      const int wy = 3 - (iy == 0) - (iy == ny - 1);
      for (int ix = 0; ix < nx; ix++) {
        const int wx = 3 - (ix == 0) - (ix == nx - 1);
        nb += wy * wx;
        row_ptr[(int64_t) iy * nx + ix + 1] = (int) nb;
      }
    }
    unsigned nt = std::max(1u, std::min(std::thread::hardware_concurrency(), 16u));
    if ((int64_t) nb * bs * bs < (1 << 20)) nt = 1;
    nt = std::min<unsigned>(nt, (unsigned) ny);
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; t++) {
      const int y0 = (int) ((int64_t) ny * t / nt), y1 = (int) ((int64_t) ny * (t + 1) / nt);
      pool.emplace_back(stencil_rows, nx, ny, bs, kind, seed, y0, y1, row_ptr, columns, values);
    }
    for (auto &th : pool) th.join();
    return BMFP_OK;
  });
}
