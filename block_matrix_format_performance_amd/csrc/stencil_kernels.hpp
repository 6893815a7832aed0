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
// node's blocks (row-major blocks) and, for r == 0, the block columns.
__global__ __launch_bounds__(256) void k_gen_stencil_rows(int nx, int ny, int bs, int kind, uint64_t seed,
                                                          const int *__restrict__ rp, int *__restrict__ cols,
                                                          double *__restrict__ vals)
{
#pragma clang fp contract(off)
  const int64_t t = (int64_t) blockIdx.x * 256 + threadIdx.x;
  const int64_t n = (int64_t) nx * ny * bs;
  if (t >= n) return;
  const int64_t i = t / bs;
  const int r = (int) (t - i * bs);
  const int iy = (int) (i / nx), ix = (int) (i - (int64_t) iy * nx);
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
    double *row = vals + nb * bs2 + (int64_t) r * bs;
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
      row[c] = v;
      if (!(s == 4 && c == r)) sum += fabs(v);
    }
    nb++;
  }
  double *dg = vals + diag * bs2 + (int64_t) r * bs + r;
  if (kind == 0) *dg = *dg + (sum + 1.0);
  else *dg = 1.02 * sum + 1.0;
}

}  // namespace
}  // namespace bmfp
