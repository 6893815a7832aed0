// noalloc.hip -- merge-path CSR SpMV over caller-owned device arrays that
// allocates nothing (bmfp_csr_spmv_noalloc).
//
// bmfp_spmv(A, BMFP_CSR_MERGE, ...) keeps its scratch -- the per-wavefront
// carries and the tile boundaries of the merge path -- in a buffer allocated
// with the matrix handle.  Callers that cannot allocate device memory (code
// under graph capture, serving loops, harnesses that forbid hipMalloc) and
// hold their CSR arrays in their own device memory use this entry point
// instead: the same kernels, in the same order, with the scratch in one
// fixed file-scope device array of NOALLOC_SCRATCH_BYTES.  Every call on a
// device shares that device's array, so calls that overlap in time on one
// device (two streams, two host threads) race on it.
#include <mutex>

#include "matrix.hpp"
#include "spmv_kernels.hpp"

namespace bmfp {

namespace {

// Scratch layout, fixed offsets: carry values (8 bytes per wavefront, read as
// float or double), carry rows (4 bytes per wavefront), then the
// NOALLOC_MAX_WAVES + 1 tile boundaries (4 bytes each).  Every slot a call
// reads is written earlier in the same call (boundaries by
// k_csr_merge_bounds, carries by k_csr_merge for every wavefront of the
// path), so nothing needs clearing between calls.
constexpr int64_t NOALLOC_MAX_WAVES = int64_t(1) << 19;  // n_rows + nnz <= 2^28
static_assert(NOALLOC_MAX_WAVES * MERGE_TILE_ITEMS == BMFP_NOALLOC_MAX_ITEMS, "limit in bmfp.h disagrees");
constexpr size_t NOALLOC_VAL_OFF = 0;
constexpr size_t NOALLOC_ROW_OFF = (size_t) NOALLOC_MAX_WAVES * 8;
constexpr size_t NOALLOC_BOUNDS_OFF = NOALLOC_ROW_OFF + (size_t) NOALLOC_MAX_WAVES * 4;
constexpr size_t NOALLOC_SCRATCH_BYTES = (NOALLOC_BOUNDS_OFF + (size_t) (NOALLOC_MAX_WAVES + 1) * 4 + 255) / 256 * 256;
static_assert(NOALLOC_SCRATCH_BYTES <= BMFP_NOALLOC_SCRATCH_BYTES, "scratch size in bmfp.h disagrees");

}  // namespace

// The scratch itself: one copy per device, resolved to a device address by
// hipGetSymbolAddress (host code never passes the name itself to a kernel).
static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
// Device address of `scratch` per device (each device holds its own copy of
// the array at its own address); filled once per device.
void *g_scratch[NOALLOC_MAX_DEVICES];
std::mutex g_scratch_mu;  // guards g_scratch

template <class T>
int noalloc_merge(char *base, int n_rows, int nnz, const int *row_ptr, const int *columns, const T *values,
                  const T *x, T *y, hipStream_t s)
{
  const int64_t waves = merge_waves(n_rows, nnz);
  CsrArgs<T, T> a;
  a.n_rows = n_rows;
  a.xcd_per = 0;
  a.rp = row_ptr;
  a.cols = columns;
  a.vals = values;
  a.x = x;
  a.y = y;
  T *carry_val = reinterpret_cast<T *>(base + NOALLOC_VAL_OFF);
  int *carry_row = reinterpret_cast<int *>(base + NOALLOC_ROW_OFF);
  int *bounds = reinterpret_cast<int *>(base + NOALLOC_BOUNDS_OFF);
  k_csr_merge_bounds<MERGE_TILE><<<(unsigned) ((waves + 1 + 255) / 256), 256, 0, s>>>(n_rows, nnz, waves, row_ptr,
                                                                                    bounds);
  BMFP_LAUNCHED();
  k_csr_merge<T, T><<<(unsigned) ((waves + 3) / 4), 256, 0, s>>>(a, nnz, waves, bounds, carry_row, carry_val);
  BMFP_LAUNCHED();
  k_csr_merge_fixup<T><<<(unsigned) ((waves + 255) / 256), 256, 0, s>>>(waves, n_rows, carry_row, carry_val, y);
  BMFP_LAUNCHED();
  return BMFP_OK;
}

}  // namespace

// Resolves, once per device, the device address of `scratch` on `device`,
// which must be the current device, and returns it in *out.  Called by
// bmfp_init and by the entry point (the first call on a device bmfp_init
// never saw resolves it there).
int noalloc_setup(int device, void **out)
{
  if (device < 0 || device >= NOALLOC_MAX_DEVICES)
    return fail(BMFP_ERR_INVALID, "device %d outside the %d devices the scratch table covers", device,
                NOALLOC_MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
    if (!ptr) return fail(BMFP_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch[device] = ptr;
  }
  if (out) *out = g_scratch[device];
  return BMFP_OK;
}

}  // namespace bmfp

using namespace bmfp;

extern "C" int bmfp_csr_spmv_noalloc(int dtype, int n_rows, int n_cols, int nnz, const int *row_ptr,
                                     const int *columns, const void *values, const void *x, void *y,
                                     bmfp_stream stream)
{
  return bmfp::guarded("bmfp_csr_spmv_noalloc", [&]() -> int {
    if (dtype != BMFP_F64 && dtype != BMFP_F32)
      return fail(BMFP_ERR_UNSUPPORTED, "allocation-free CSR SpMV supports F64 and F32");
    if (n_rows < 0 || n_cols < 0 || nnz < 0) return fail(BMFP_ERR_INVALID, "negative size");
    if (n_rows == 0) return BMFP_OK;
    if (!row_ptr || !y || (nnz > 0 && (!columns || !values || !x)))
      return fail(BMFP_ERR_INVALID, "NULL device array");
    const int64_t items = (int64_t) n_rows + nnz;
    if (merge_waves(n_rows, nnz) > NOALLOC_MAX_WAVES)
      return fail(BMFP_ERR_INVALID,
                  "n_rows + nnz = %lld exceeds the allocation-free limit of %lld (scratch of %zu bytes)",
                  (long long) items, (long long) (NOALLOC_MAX_WAVES * MERGE_TILE_ITEMS), NOALLOC_SCRATCH_BYTES);
    int dev = 0;
    BMFP_HIP(hipGetDevice(&dev));
    void *ptr = nullptr;
    BMFP_TRY(noalloc_setup(dev, &ptr));
    char *base = static_cast<char *>(ptr);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (dtype == BMFP_F64)
      return noalloc_merge<double>(base, n_rows, nnz, row_ptr, columns, static_cast<const double *>(values),
                                   static_cast<const double *>(x), static_cast<double *>(y), s);
    return noalloc_merge<float>(base, n_rows, nnz, row_ptr, columns, static_cast<const float *>(values),
                                static_cast<const float *>(x), static_cast<float *>(y), s);
  });
}
