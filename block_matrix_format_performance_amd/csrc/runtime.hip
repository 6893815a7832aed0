// runtime.hip -- C ABI: errors, device/memory/stream plumbing, matrix handles,
// SpMV entry points.
#include <algorithm>
#include <type_traits>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "matrix.hpp"
#include "stencil_kernels.hpp"

namespace bmfp {

namespace {
thread_local std::string g_last_error;
}

void set_error(const char *fmt, ...)
{
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

const char *get_error() { return g_last_error.c_str(); }

int fail(int code, const char *fmt, ...)
{
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

// ---- device fill -------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_fill(T *dst, int64_t n, int kind, uint64_t key, double value)
{
  const int64_t stride = (int64_t) gridDim.x * blockDim.x;
  for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double v = kind == 0 ? rng_unit(key, (uint64_t) i) : kind == 1 ? rng_sym(key, (uint64_t) i) : value;
    dst[i] = (T) v;
  }
}

// ---- streaming copy (bandwidth yardstick) -------------------------------------
typedef unsigned int u32x4c __attribute__((ext_vector_type(4)));
// Grid-stride loops with STREAM_U 16-byte loads in flight per lane (all
// issued before the first is used); consecutive lanes, consecutive addresses.
constexpr int STREAM_U = 8;

__global__ __launch_bounds__(256) void k_stream_copy(u32x4c *__restrict__ dst, const u32x4c *__restrict__ src,
                                                     int64_t n16)
{
  const int64_t stride = (int64_t) gridDim.x * blockDim.x;
  int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + (STREAM_U - 1) * stride < n16; i += STREAM_U * stride) {
    u32x4c v[STREAM_U];
#pragma unroll
    for (int k = 0; k < STREAM_U; ++k) v[k] = src[i + k * stride];
#pragma unroll
    for (int k = 0; k < STREAM_U; ++k) dst[i + k * stride] = v[k];
  }
  for (; i < n16; i += stride) dst[i] = src[i];
}

// Reads every 16-byte chunk once and folds it into a per-lane XOR; the fold
// is written only if it equals a fixed constant (a store the compiler cannot
// drop, which practically never happens).
__global__ __launch_bounds__(256) void k_stream_read(const u32x4c *__restrict__ src, int64_t n16,
                                                     unsigned *__restrict__ sink)
{
  const int64_t stride = (int64_t) gridDim.x * blockDim.x;
  int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  unsigned acc = 0;
  for (; i + (STREAM_U - 1) * stride < n16; i += STREAM_U * stride) {
    u32x4c v[STREAM_U];
#pragma unroll
    for (int k = 0; k < STREAM_U; ++k) v[k] = src[i + k * stride];
#pragma unroll
    for (int k = 0; k < STREAM_U; ++k) acc ^= v[k].x ^ v[k].y ^ v[k].z ^ v[k].w;
  }
  for (; i < n16; i += stride) {
    const u32x4c v = src[i];
    acc ^= v.x ^ v.y ^ v.z ^ v.w;
  }
  if (acc == 0x9e3779b9u) sink[0] = acc;
}

// ---- value conversion (F64 -> F64/F32/F16, optional block transpose) --------
// One wavefront per block row: first the row's max |entry| (F16 scale), then
// a strided pass that writes every element in the target layout.
template <class OT, bool SCALED>
__global__ __launch_bounds__(256) void k_convert(int n_brows, int bs, int src_layout, int dst_layout,
                                                 const int *__restrict__ rp, const double *__restrict__ src,
                                                 OT *__restrict__ dst, float *__restrict__ scale)
{
  const int64_t br = ((int64_t) blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (br >= n_brows) return;
  const int64_t bs2 = (int64_t) bs * bs;
  const int64_t e0 = (int64_t) rp[br] * bs2, e1 = (int64_t) rp[br + 1] * bs2;
  double inv = 1.0;
  if (SCALED) {
    double m = 0.0;
    for (int64_t e = e0 + lane; e < e1; e += 64) m = fmax(m, fabs(src[e]));
    for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if (m == 0.0) m = 1.0;
    inv = 1.0 / m;
    if (lane == 0) scale[br] = (float) m;
  }
  for (int64_t e = e0 + lane; e < e1; e += 64) {
    const int64_t b = e / bs2;
    const int w = (int) (e - b * bs2);
    int r, c;
    if (src_layout == BMFP_ROW_MAJOR) { r = w / bs; c = w % bs; }
    else { c = w / bs; r = w % bs; }
    const int64_t d = b * bs2 + (dst_layout == BMFP_ROW_MAJOR ? (int64_t) r * bs + c : (int64_t) c * bs + r);
    const double v = SCALED ? src[e] * inv : src[e];
    if constexpr (sizeof(OT) == 2) dst[d] = __float2half_rn((float) v);
    else dst[d] = (OT) v;
  }
}

// Same layout, no scale: a flat pass, two values per thread (one 16-byte
// load); nv values from a 256-byte aligned allocation.
template <class OT>
__global__ __launch_bounds__(256) void k_convert_flat(int64_t nv, const double *__restrict__ src,
                                                      OT *__restrict__ dst)
{
  const int64_t i = ((int64_t) blockIdx.x * 256 + threadIdx.x) * 2;
  if (i + 1 < nv) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    const d2 v = *reinterpret_cast<const d2 *>(src + i);
    dst[i] = (OT) v.x;
    dst[i + 1] = (OT) v.y;
  } else if (i < nv) {
    dst[i] = (OT) src[i];
  }
}

// Same layout with the per-block-row FP16 scale: the row's max |entry|, then
// every element scaled in place order (no index arithmetic).  A block row of
// up to 576 values (9 blocks of 8x8) is read once, 9 values per lane loaded
// together; longer rows are read a second time for the conversion.
__global__ __launch_bounds__(256) void k_convert_scaled(int n_brows, int64_t bs2, const int *__restrict__ rp,
                                                        const double *__restrict__ src, __half *__restrict__ dst,
                                                        float *__restrict__ scale)
{
  constexpr int K = 9;
  const int64_t br = ((int64_t) blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (br >= n_brows) return;
  const int64_t e0 = (int64_t) rp[br] * bs2, e1 = (int64_t) rp[br + 1] * bs2;
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = e0 + lane + 64 * k < e1 ? src[e0 + lane + 64 * k] : 0.0;
  double m = 0.0;
#pragma unroll
  for (int k = 0; k < K; ++k) m = fmax(m, fabs(v[k]));
  for (int64_t e = e0 + lane + 64 * K; e < e1; e += 64) m = fmax(m, fabs(src[e]));
  for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if (m == 0.0) m = 1.0;
  const double inv = 1.0 / m;
  if (lane == 0) scale[br] = (float) m;
#pragma unroll
  for (int k = 0; k < K; ++k)
    if (e0 + lane + 64 * k < e1) dst[e0 + lane + 64 * k] = __float2half_rn((float) (v[k] * inv));
  for (int64_t e = e0 + lane + 64 * K; e < e1; e += 64) dst[e] = __float2half_rn((float) (src[e] * inv));
}

// FP16 block-row scales of a pass-sliced copy: the row's max |entry| rounded
// to FP32 (1 for an all-zero row); 16 lanes per block row (4 rows per
// wavefront), 9 values per lane loaded together.
__global__ __launch_bounds__(256) void k_row_scale(int n_brows, int64_t bs2, const int *__restrict__ rp,
                                                   const double *__restrict__ src, float *__restrict__ scale)
{
  constexpr int K = 9, G = 16;
  const int64_t br = ((int64_t) blockIdx.x * 256 + threadIdx.x) / G;
  const int g = threadIdx.x & (G - 1);
  double m = 0.0;
  if (br < n_brows) {
    const int64_t e0 = (int64_t) rp[br] * bs2, e1 = (int64_t) rp[br + 1] * bs2;
    for (int64_t e = e0 + g; e < e1; e += G * K) {
      double v[K];
#pragma unroll
      for (int k = 0; k < K; ++k) v[k] = e + G * k < e1 ? src[e + G * k] : 0.0;
#pragma unroll
      for (int k = 0; k < K; ++k) m = fmax(m, fabs(v[k]));
    }
  }
  for (int off = G / 2; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if (br < n_brows && g == 0) scale[br] = m == 0.0 ? 1.0f : (float) m;
}

// Pass-sliced copy (layout: bmfp_matrix_t::sl_j): one thread per row of a
// sliced block (grid-stride), consecutive threads storing consecutive rows;
// each row read from its block in the FP64 source, zero in padding blocks;
// the thread of row 0 writes the block's column (padding: the row's last
// block column, 0 for an empty row).  FP16 values are FP16(FP32(v / scale)).
template <class OT, int BS>
__global__ __launch_bounds__(256) void k_slice_rows(int n_brows, int rpw, int sl_j, const int *__restrict__ rp,
                                                    const int *__restrict__ cols, const double *__restrict__ src,
                                                    const float *__restrict__ scale, OT *__restrict__ out,
                                                    int *__restrict__ cols_sl)
{
  constexpr int64_t BS2 = BS * BS;
  const int64_t n_rows = ((int64_t) n_brows + rpw - 1) / rpw * rpw * sl_j * BS;
  for (int64_t t = (int64_t) blockIdx.x * 256 + threadIdx.x; t < n_rows; t += (int64_t) gridDim.x * 256) {
    const int64_t sb = t / BS, pj = sb / rpw, pp = pj / sl_j;
    const int r = (int) (t - sb * BS), lbr = (int) (sb - pj * rpw), j = (int) (pj - pp * sl_j);
    const int64_t br = pp * rpw + lbr;
    int b0 = 0, len = 0;
    if (br < n_brows) {
      b0 = rp[br];
      len = rp[br + 1] - b0;
    }
    if (r == 0) cols_sl[sb] = j < len ? cols[b0 + j] : len > 0 ? cols[b0 + len - 1] : 0;
    OT o[BS];
    if (j < len) {
      double v[BS];
      load_n<double, BS, lowbit(8 * BS)>(src + (((int64_t) b0 + j) * BS2 + (int64_t) r * BS), v);
      if constexpr (sizeof(OT) == 2) {
        const double inv = 1.0 / (double) scale[br];
#pragma unroll
        for (int c = 0; c < BS; ++c) o[c] = __float2half_rn((float) (v[c] * inv));
      } else {
#pragma unroll
        for (int c = 0; c < BS; ++c) o[c] = (OT) v[c];
      }
    } else {
#pragma unroll
      for (int c = 0; c < BS; ++c) o[c] = (OT) 0.0f;
    }
    store_words<OT, BS>(out + t * BS, o);
  }
}

int make_sliced_copy(const bmfp_matrix_t &src, int dtype, int rpw, int j_slots, bmfp_matrix_t &m, hipStream_t s,
                     hipEvent_t e0, hipEvent_t e1)
{
  if (src.dtype != BMFP_F64 || src.layout != BMFP_ROW_MAJOR || !src.is_bcsr || rpw <= 0 || j_slots <= 0 ||
      (dtype != BMFP_F32 && dtype != BMFP_F16))
    return fail(BMFP_ERR_INVALID, "pass-sliced copy of an F64 row-major BCSR matrix into F32 / F16");
  const int bs = src.bs;
  const int64_t nsb = ((int64_t) src.n_rows + rpw - 1) / rpw * rpw * j_slots;
  m.sl_j = j_slots;
  m.sl_rpw = rpw;
  BMFP_TRY(m.values.alloc((size_t) std::max<int64_t>(nsb, 1) * bs * bs * dtype_size(dtype)));
  BMFP_TRY(m.cols_sl.alloc(sizeof(int) * (size_t) std::max<int64_t>(nsb, 1)));
  if (dtype == BMFP_F16) BMFP_TRY(m.scale.alloc(sizeof(float) * (size_t) std::max(src.n_rows, 1)));
  if (e0) BMFP_HIP(hipEventRecord(e0, s));
  if (src.n_rows == 0) {
    if (e1) BMFP_HIP(hipEventRecord(e1, s));
    return BMFP_OK;
  }
  const int *rp = src.row_ptr->as<int>(), *cols = src.columns->as<int>();
  const double *v64 = src.values.as<double>();
  if (dtype == BMFP_F16) {
    k_row_scale<<<(unsigned) (((int64_t) src.n_rows * 16 + 255) / 256), 256, 0, s>>>(src.n_rows, (int64_t) bs * bs,
                                                                                     rp, v64, m.scale.as<float>());
    BMFP_LAUNCHED();
  }
  const unsigned g = (unsigned) std::max<int64_t>(1, std::min<int64_t>((nsb * bs + 255) / 256, 1 << 20));
  auto launch_bs = [&](auto bsc) {
    constexpr int BS = decltype(bsc)::value;
    if (dtype == BMFP_F16)
      k_slice_rows<__half, BS><<<g, 256, 0, s>>>(src.n_rows, rpw, j_slots, rp, cols, v64, m.scale.as<float>(),
                                                 m.values.as<__half>(), m.cols_sl.as<int>());
    else
      k_slice_rows<float, BS><<<g, 256, 0, s>>>(src.n_rows, rpw, j_slots, rp, cols, v64, nullptr,
                                                m.values.as<float>(), m.cols_sl.as<int>());
  };
  switch (bs) {
  case 4: launch_bs(std::integral_constant<int, 4>{}); break;
  case 6: launch_bs(std::integral_constant<int, 6>{}); break;
  case 8: launch_bs(std::integral_constant<int, 8>{}); break;
  default: return fail(BMFP_ERR_UNSUPPORTED, "pass-sliced copies for bs 4, 6, 8 only");
  }
  BMFP_LAUNCHED();
  if (e1) BMFP_HIP(hipEventRecord(e1, s));
  return BMFP_OK;
}

}  // namespace bmfp

using namespace bmfp;

bmfp_matrix_t::~bmfp_matrix_t() { destroy_library(*this); }

// ---- diagnostics -------------------------------------------------------------
extern "C" const char *bmfp_last_error(void) { return get_error(); }

extern "C" const char *bmfp_status_string(int status)
{
  switch (status) {
  case BMFP_OK: return "ok";
  case BMFP_ERR_INVALID: return "invalid argument";
  case BMFP_ERR_UNSUPPORTED: return "unsupported";
  case BMFP_ERR_NO_MEMORY: return "out of device memory";
  case BMFP_ERR_BREAKDOWN: return "solver breakdown";
  case BMFP_ERR_NOT_CONVERGED: return "not converged";
  case BMFP_ERR_HIP: return "HIP runtime error";
  case BMFP_ERR_NONFINITE: return "non-finite value";
  case BMFP_ERR_BUDGET: return "memory budget exceeded";
  case BMFP_ERR_NO_DEVICE: return "no device";
  default: return "unknown status";
  }
}

extern "C" int bmfp_version(void) { return bmfp::guarded("bmfp_version", [&]() -> int { return BMFP_VERSION; }); }

// ---- device / memory / streams ---------------------------------------------------
extern "C" int bmfp_device_count(int *count)
{
  return bmfp::guarded("bmfp_device_count", [&]() -> int {
    if (!count) return fail(BMFP_ERR_INVALID, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return BMFP_OK;
  });
}

extern "C" int bmfp_init(int device)
{
  return bmfp::guarded("bmfp_init", [&]() -> int {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
      return fail(BMFP_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= n) return fail(BMFP_ERR_INVALID, "device %d out of range [0,%d)", device, n);
    BMFP_HIP(hipSetDevice(device));
    hipDeviceProp_t p;
    BMFP_HIP(hipGetDeviceProperties(&p, device));
    if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0)
      return fail(BMFP_ERR_NO_DEVICE, "device %d is %s; this library is built for gfx950", device, p.gcnArchName);
    return noalloc_setup(device, nullptr);
  });
}

extern "C" int bmfp_device_info(char *name, size_t name_len, size_t *total_mem, int *cus, int *arch)
{
  return bmfp::guarded("bmfp_device_info", [&]() -> int {
    int dev = 0;
    BMFP_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    BMFP_HIP(hipGetDeviceProperties(&p, dev));
    if (name && name_len) {
      std::snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
    }
    if (total_mem) *total_mem = p.totalGlobalMem;
    if (cus) *cus = p.multiProcessorCount;
    if (arch) *arch = p.major * 100 + p.minor;
    return BMFP_OK;
  });
}

extern "C" int bmfp_malloc(void **ptr, size_t bytes)
{
  return bmfp::guarded("bmfp_malloc", [&]() -> int {
    if (!ptr) return fail(BMFP_ERR_INVALID, "ptr is NULL");
    BMFP_HIP(hipMalloc(ptr, bytes ? bytes : 256));
    return BMFP_OK;
  });
}

extern "C" int bmfp_free(void *ptr)
{
  return bmfp::guarded("bmfp_free", [&]() -> int {
    if (ptr) BMFP_HIP(hipFree(ptr));
    return BMFP_OK;
  });
}

extern "C" int bmfp_malloc_host(void **ptr, size_t bytes)
{
  return bmfp::guarded("bmfp_malloc_host", [&]() -> int {
    if (!ptr) return fail(BMFP_ERR_INVALID, "ptr is NULL");
    BMFP_HIP(hipHostMalloc(ptr, bytes ? bytes : 64, hipHostMallocDefault));
    return BMFP_OK;
  });
}

extern "C" int bmfp_free_host(void *ptr)
{
  return bmfp::guarded("bmfp_free_host", [&]() -> int {
    if (ptr) BMFP_HIP(hipHostFree(ptr));
    return BMFP_OK;
  });
}

extern "C" int bmfp_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
  return bmfp::guarded("bmfp_memcpy_h2d", [&]() -> int {
    if (bytes) BMFP_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return BMFP_OK;
  });
}

extern "C" int bmfp_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
  return bmfp::guarded("bmfp_memcpy_d2h", [&]() -> int {
    if (bytes) BMFP_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return BMFP_OK;
  });
}

extern "C" int bmfp_memcpy_d2d(void *dst, const void *src, size_t bytes)
{
  return bmfp::guarded("bmfp_memcpy_d2d", [&]() -> int {
    if (bytes) BMFP_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return BMFP_OK;
  });
}

extern "C" int bmfp_memset(void *dst, int value, size_t bytes)
{
  return bmfp::guarded("bmfp_memset", [&]() -> int {
    if (bytes) BMFP_HIP(hipMemset(dst, value, bytes));
    return BMFP_OK;
  });
}

extern "C" int bmfp_stream_create(bmfp_stream *stream)
{
  return bmfp::guarded("bmfp_stream_create", [&]() -> int {
    if (!stream) return fail(BMFP_ERR_INVALID, "stream is NULL");
    hipStream_t s;
    BMFP_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return BMFP_OK;
  });
}

extern "C" int bmfp_stream_destroy(bmfp_stream stream)
{
  return bmfp::guarded("bmfp_stream_destroy", [&]() -> int {
    if (stream) BMFP_HIP(hipStreamDestroy(static_cast<hipStream_t>(stream)));
    return BMFP_OK;
  });
}

extern "C" int bmfp_stream_synchronize(bmfp_stream stream)
{
  return bmfp::guarded("bmfp_stream_synchronize", [&]() -> int {
    BMFP_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    return BMFP_OK;
  });
}

extern "C" int bmfp_device_synchronize(void)
{
  return bmfp::guarded("bmfp_device_synchronize", [&]() -> int {
    BMFP_HIP(hipDeviceSynchronize());
    return BMFP_OK;
  });
}

extern "C" int bmfp_event_create(bmfp_event *event)
{
  return bmfp::guarded("bmfp_event_create", [&]() -> int {
    if (!event) return fail(BMFP_ERR_INVALID, "event is NULL");
    hipEvent_t e;
    BMFP_HIP(hipEventCreate(&e));
    *event = e;
    return BMFP_OK;
  });
}

extern "C" int bmfp_event_destroy(bmfp_event event)
{
  return bmfp::guarded("bmfp_event_destroy", [&]() -> int {
    if (event) BMFP_HIP(hipEventDestroy(static_cast<hipEvent_t>(event)));
    return BMFP_OK;
  });
}

extern "C" int bmfp_event_record(bmfp_event event, bmfp_stream stream)
{
  return bmfp::guarded("bmfp_event_record", [&]() -> int {
    if (!event) return fail(BMFP_ERR_INVALID, "event is NULL");
    BMFP_HIP(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)));
    return BMFP_OK;
  });
}

extern "C" int bmfp_event_elapsed(bmfp_event start, bmfp_event stop, double *seconds)
{
  return bmfp::guarded("bmfp_event_elapsed", [&]() -> int {
    if (!start || !stop || !seconds) return fail(BMFP_ERR_INVALID, "NULL argument");
    BMFP_HIP(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
    float ms = 0.f;
    BMFP_HIP(hipEventElapsedTime(&ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
    *seconds = ms * 1e-3;
    return BMFP_OK;
  });
}

extern "C" int bmfp_stream_copy(void *dst, const void *src, size_t bytes, bmfp_stream stream)
{
  return bmfp::guarded("bmfp_stream_copy", [&]() -> int {
    if (bytes == 0) return BMFP_OK;
    if (!dst || !src || bytes % 16 || ((uintptr_t) dst | (uintptr_t) src) % 16)
      return fail(BMFP_ERR_INVALID, "bmfp_stream_copy: needs 16-byte aligned pointers and size");
    const int64_t n16 = (int64_t) (bytes / 16);
    const unsigned groups = (unsigned) std::min<int64_t>((n16 + 255) / 256, 256 * 32);
    k_stream_copy<<<groups, 256, 0, static_cast<hipStream_t>(stream)>>>(static_cast<u32x4c *>(dst),
                                                                          static_cast<const u32x4c *>(src), n16);
    BMFP_LAUNCHED();
    return BMFP_OK;
  });
}

extern "C" int bmfp_stream_read(const void *src, size_t bytes, void *sink, bmfp_stream stream)
{
  return bmfp::guarded("bmfp_stream_read", [&]() -> int {
    if (bytes == 0) return BMFP_OK;
    if (!src || !sink || bytes % 16 || (uintptr_t) src % 16 || (uintptr_t) sink % 4)
      return fail(BMFP_ERR_INVALID, "bmfp_stream_read: needs a 16-byte aligned source and size and a sink");
    const int64_t n16 = (int64_t) (bytes / 16);
    const unsigned groups = (unsigned) std::min<int64_t>((n16 + 255) / 256, 256 * 32);
    k_stream_read<<<groups, 256, 0, static_cast<hipStream_t>(stream)>>>(static_cast<const u32x4c *>(src), n16,
                                                                          static_cast<unsigned *>(sink));
    BMFP_LAUNCHED();
    return BMFP_OK;
  });
}

extern "C" int bmfp_fill(void *dst, int dtype, size_t n, int kind, uint64_t seed, uint64_t stream_id,
                         double value, bmfp_stream stream)
{
  return bmfp::guarded("bmfp_fill", [&]() -> int {
    if (n == 0) return BMFP_OK;
    if (!dst || kind < 0 || kind > 2) return fail(BMFP_ERR_INVALID, "bmfp_fill: bad arguments");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint64_t key = rng_key(seed, stream_id);
    const unsigned groups = (unsigned) std::min<size_t>((n + 255) / 256, 8192);
    if (dtype == BMFP_F64)
      k_fill<double><<<groups, 256, 0, s>>>(static_cast<double *>(dst), (int64_t) n, kind, key, value);
    else if (dtype == BMFP_F32)
      k_fill<float><<<groups, 256, 0, s>>>(static_cast<float *>(dst), (int64_t) n, kind, key, value);
    else
      return fail(BMFP_ERR_INVALID, "bmfp_fill: dtype must be F64 or F32");
    BMFP_LAUNCHED();
    return BMFP_OK;
  });
}

// ---- matrices -------------------------------------------------------------------
namespace {

int check_structure(int n_rows, int n_cols, int bs, int nnzb, const int *row_ptr, const int *columns,
                    int *max_row)
{
  if (n_rows < 0 || n_cols < 0 || bs < 1 || nnzb < 0)
    return fail(BMFP_ERR_INVALID, "negative shape (n_rows %d, n_cols %d, bs %d, nnzb %d)", n_rows, n_cols, bs, nnzb);
  if ((int64_t) nnzb * bs * bs > INT_MAX || (int64_t) n_rows * bs > INT_MAX || (int64_t) n_cols * bs > INT_MAX)
    return fail(BMFP_ERR_INVALID, "matrix too large for int32 indices");
  if (!row_ptr || (nnzb && !columns)) return fail(BMFP_ERR_INVALID, "NULL index array");
  if (row_ptr[0] != 0 || row_ptr[n_rows] != nnzb)
    return fail(BMFP_ERR_INVALID, "row_ptr[0]=%d, row_ptr[n_rows]=%d, nnz=%d", row_ptr[0], row_ptr[n_rows], nnzb);
  int mx = 0;
  for (int i = 0; i < n_rows; i++) {
    const int len = row_ptr[i + 1] - row_ptr[i];
    if (len < 0) return fail(BMFP_ERR_INVALID, "row_ptr decreases at row %d", i);
    mx = std::max(mx, len);
  }
  for (int k = 0; k < nnzb; k++)
    if (columns[k] < 0 || columns[k] >= n_cols)
      return fail(BMFP_ERR_INVALID, "column index %d at %d out of range [0,%d)", columns[k], k, n_cols);
  *max_row = mx;
  return BMFP_OK;
}

int upload_indices(bmfp_matrix_t &m, const int *row_ptr, const int *columns)
{
  m.row_ptr = std::make_shared<DeviceBuffer>();
  m.columns = std::make_shared<DeviceBuffer>();
  BMFP_TRY(m.row_ptr->alloc(sizeof(int) * ((size_t) m.n_rows + 1)));
  BMFP_TRY(m.columns->alloc(sizeof(int) * (size_t) m.nnzb));
  BMFP_HIP(hipMemcpy(m.row_ptr->ptr, row_ptr, sizeof(int) * ((size_t) m.n_rows + 1), hipMemcpyHostToDevice));
  if (m.nnzb)
    BMFP_HIP(hipMemcpy(m.columns->ptr, columns, sizeof(int) * (size_t) m.nnzb, hipMemcpyHostToDevice));
  return BMFP_OK;
}

// Converts device F64 values `src` (layout src_layout) into m.values / m.scale.
// BCSR -> CSR in device memory (csr_matrix_class(const bcsr_matrix_class&),
// matrix_converters.h:88-117): scalar row br*bs + r lists, block by block in
// the block row's order, the bs entries (col*bs + c, value of (r, c)).  One
// thread per scalar row writes its len*bs entries.
template <class VT>
__global__ __launch_bounds__(256) void k_bcsr_to_csr(int n_brows, int bs, int layout, const int *__restrict__ rp,
                                                     const int *__restrict__ cols, const double *__restrict__ vals,
                                                     int *__restrict__ out_rp, int *__restrict__ out_cols,
                                                     VT *__restrict__ out_vals)
{
  const int64_t i = (int64_t) blockIdx.x * 256 + threadIdx.x;
  const int64_t rows = (int64_t) n_brows * bs;
  if (i > rows) return;
  if (i == rows) {
    out_rp[rows] = (int) ((int64_t) rp[n_brows] * bs * bs);
    return;
  }
  const int br = (int) (i / bs), r = (int) (i - (int64_t) br * bs);
  const int b0 = rp[br], len = rp[br + 1] - b0;
  const int64_t bs2 = (int64_t) bs * bs;
  const int64_t base = (int64_t) b0 * bs2 + (int64_t) r * len * bs;
  out_rp[i] = (int) base;
  for (int k = 0; k < len; k++) {
    const int64_t b = b0 + k;
    const int col = cols[b];
    for (int c = 0; c < bs; c++) {
      out_cols[base + (int64_t) k * bs + c] = col * bs + c;
      out_vals[base + (int64_t) k * bs + c] =
          (VT) vals[b * bs2 + (layout == BMFP_ROW_MAJOR ? (int64_t) r * bs + c : (int64_t) c * bs + r)];
    }
  }
}

int convert_values(bmfp_matrix_t &m, const double *src, int src_layout, hipStream_t s)
{
  const int64_t nv = m.n_vals();
  BMFP_TRY(m.values.alloc((size_t) nv * dtype_size(m.dtype)));
  if (m.dtype == BMFP_F16) BMFP_TRY(m.scale.alloc(sizeof(float) * (size_t) std::max(m.n_rows, 1)));
  if (m.n_rows == 0) return BMFP_OK;
  const unsigned groups = (unsigned) (((int64_t) m.n_rows * 64 + 255) / 256);
  const int *rp = m.row_ptr->as<int>();
  if (src_layout == m.layout && m.dtype != BMFP_F16) {
    const unsigned flat = (unsigned) std::max<int64_t>((nv + 511) / 512, 1);
    if (m.dtype == BMFP_F64)
      k_convert_flat<double><<<flat, 256, 0, s>>>(nv, src, m.values.as<double>());
    else
      k_convert_flat<float><<<flat, 256, 0, s>>>(nv, src, m.values.as<float>());
    BMFP_LAUNCHED();
    return BMFP_OK;
  }
  if (src_layout == m.layout) {
    k_convert_scaled<<<groups, 256, 0, s>>>(m.n_rows, (int64_t) m.bs * m.bs, rp, src, m.values.as<__half>(),
                                             m.scale.as<float>());
    BMFP_LAUNCHED();
    return BMFP_OK;
  }
  switch (m.dtype) {
  case BMFP_F64:
    k_convert<double, false><<<groups, 256, 0, s>>>(m.n_rows, m.bs, src_layout, m.layout, rp, src,
                                                   m.values.as<double>(), nullptr);
    break;
  case BMFP_F32:
    k_convert<float, false><<<groups, 256, 0, s>>>(m.n_rows, m.bs, src_layout, m.layout, rp, src,
                                                  m.values.as<float>(), nullptr);
    break;
  case BMFP_F16:
    k_convert<__half, true><<<groups, 256, 0, s>>>(m.n_rows, m.bs, src_layout, m.layout, rp, src,
                                                  m.values.as<__half>(), m.scale.as<float>());
    break;
  default: return fail(BMFP_ERR_INVALID, "bad dtype %d", m.dtype);
  }
  BMFP_LAUNCHED();
  return BMFP_OK;
}

}  // namespace

extern "C" int bmfp_bcsr_create(bmfp_matrix *out, int dtype, int layout, int n_rows, int n_cols, int bs, int nnzb,
                                const int *row_ptr, const int *columns, const double *values)
{
  return bmfp::guarded("bmfp_bcsr_create", [&]() -> int {
    if (!out) return fail(BMFP_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (dtype < BMFP_F64 || dtype > BMFP_F16) return fail(BMFP_ERR_INVALID, "bad dtype %d", dtype);
    if (layout != BMFP_ROW_MAJOR && layout != BMFP_COL_MAJOR) return fail(BMFP_ERR_INVALID, "bad layout %d", layout);
    int mx = 0;
    BMFP_TRY(check_structure(n_rows, n_cols, bs, nnzb, row_ptr, columns, &mx));
    if (nnzb && !values) return fail(BMFP_ERR_INVALID, "values is NULL");
    std::unique_ptr<bmfp_matrix_t> m(new bmfp_matrix_t);
    m->is_bcsr = 1;
    m->dtype = dtype;
    m->layout = layout;
    m->n_rows = n_rows;
    m->n_cols = n_cols;
    m->bs = bs;
    m->nnzb = nnzb;
    m->max_row_blocks = mx;
    m->mean_row_blocks = n_rows ? (double) nnzb / n_rows : 0.0;
    BMFP_HIP(hipGetDevice(&m->device));
    BMFP_TRY(upload_indices(*m, row_ptr, columns));
    const size_t nv = (size_t) m->n_vals();
    if (dtype == BMFP_F64) {
      BMFP_TRY(m->values.alloc(nv * sizeof(double)));
      if (nv) BMFP_HIP(hipMemcpy(m->values.ptr, values, nv * sizeof(double), hipMemcpyHostToDevice));
    } else {
      DeviceBuffer staging;
      BMFP_TRY(staging.alloc(nv * sizeof(double)));
      if (nv) BMFP_HIP(hipMemcpy(staging.ptr, values, nv * sizeof(double), hipMemcpyHostToDevice));
      BMFP_TRY(convert_values(*m, staging.as<double>(), layout, nullptr));
      BMFP_HIP(hipDeviceSynchronize());
    }
    *out = m.release();
    return BMFP_OK;
  });
}

extern "C" int bmfp_csr_create(bmfp_matrix *out, int dtype, int n_rows, int n_cols, int nnz, const int *row_ptr,
                               const int *columns, const double *values)
{
  return bmfp::guarded("bmfp_csr_create", [&]() -> int {
    if (!out) return fail(BMFP_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (dtype != BMFP_F64 && dtype != BMFP_F32) return fail(BMFP_ERR_INVALID, "CSR dtype must be F64 or F32");
    int mx = 0;
    BMFP_TRY(check_structure(n_rows, n_cols, 1, nnz, row_ptr, columns, &mx));
    if (nnz && !values) return fail(BMFP_ERR_INVALID, "values is NULL");
    std::unique_ptr<bmfp_matrix_t> m(new bmfp_matrix_t);
    m->is_bcsr = 0;
    m->dtype = dtype;
    m->n_rows = n_rows;
    m->n_cols = n_cols;
    m->bs = 1;
    m->nnzb = nnz;
    m->max_row_blocks = mx;
    m->mean_row_blocks = n_rows ? (double) nnz / n_rows : 0.0;
    m->csr_group = csr_group_for(m->mean_row_blocks);
    BMFP_HIP(hipGetDevice(&m->device));
    BMFP_TRY(upload_indices(*m, row_ptr, columns));
    BMFP_TRY(m->merge_carry.alloc(merge_carry_bytes(n_rows, nnz)));
    if (dtype == BMFP_F64) {
      BMFP_TRY(m->values.alloc(sizeof(double) * (size_t) nnz));
      if (nnz) BMFP_HIP(hipMemcpy(m->values.ptr, values, sizeof(double) * (size_t) nnz, hipMemcpyHostToDevice));
    } else {
      std::vector<float> f(values, values + nnz);
      BMFP_TRY(m->values.alloc(sizeof(float) * (size_t) nnz));
      if (nnz) BMFP_HIP(hipMemcpy(m->values.ptr, f.data(), sizeof(float) * (size_t) nnz, hipMemcpyHostToDevice));
    }
    *out = m.release();
    return BMFP_OK;
  });
}

extern "C" int bmfp_stencil_generate(bmfp_matrix m, int kind, uint64_t seed, bmfp_stream stream)
{
  return bmfp::guarded("bmfp_stencil_generate", [&]() -> int {
    if (!m) return fail(BMFP_ERR_INVALID, "matrix is NULL");
    if (m->stencil_nx < 1 || m->dtype != BMFP_F64 || m->layout != BMFP_ROW_MAJOR)
      return fail(BMFP_ERR_INVALID, "needs an F64 row-major matrix made by bmfp_stencil_create");
    if (kind != 0 && kind != 1) return fail(BMFP_ERR_INVALID, "kind must be 0 or 1");
    launch_gen_stencil(m->stencil_nx, m->stencil_ny, m->bs, kind, seed, m->row_ptr->as<int>(), m->columns->as<int>(),
                       m->values.as<double>(), nullptr, static_cast<hipStream_t>(stream));
    BMFP_LAUNCHED();
    return BMFP_OK;
  });
}

extern "C" int bmfp_stencil_create(bmfp_matrix *out, int nx, int ny, int bs, int kind, uint64_t seed, int dtype,
                                   int layout)
{
  return bmfp::guarded("bmfp_stencil_create", [&]() -> int {
    if (!out) return fail(BMFP_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (nx < 1 || ny < 1 || bs < 1 || (kind != 0 && kind != 1))
      return fail(BMFP_ERR_INVALID, "bad stencil (nx %d, ny %d, bs %d, kind %d)", nx, ny, bs, kind);
    if (dtype < BMFP_F64 || dtype > BMFP_F16) return fail(BMFP_ERR_INVALID, "bad dtype %d", dtype);
    if (layout != BMFP_ROW_MAJOR && layout != BMFP_COL_MAJOR) return fail(BMFP_ERR_INVALID, "bad layout %d", layout);
    const int64_t nodes = (int64_t) nx * ny, nnzb = stencil_block_count(nx, ny);
    if (nnzb * bs * bs > INT_MAX || nodes * bs > INT_MAX) return fail(BMFP_ERR_INVALID, "stencil too large for int32 indices");
    std::unique_ptr<bmfp_matrix_t> m(new bmfp_matrix_t);
    m->is_bcsr = 1;
    m->dtype = dtype;
    m->layout = layout;
    m->n_rows = m->n_cols = (int) nodes;
    m->bs = bs;
    m->nnzb = (int) nnzb;
    m->max_row_blocks = std::min(nx, 3) * std::min(ny, 3);
    m->mean_row_blocks = (double) nnzb / (double) nodes;
    m->stencil_nx = nx;
    m->stencil_ny = ny;
    BMFP_HIP(hipGetDevice(&m->device));
    std::vector<int> rp;
    stencil_row_ptr(nx, ny, rp);
    m->row_ptr = std::make_shared<DeviceBuffer>();
    m->columns = std::make_shared<DeviceBuffer>();
    BMFP_TRY(m->row_ptr->alloc(sizeof(int) * rp.size()));
    BMFP_TRY(m->columns->alloc(sizeof(int) * (size_t) nnzb));
    BMFP_HIP(hipMemcpy(m->row_ptr->ptr, rp.data(), sizeof(int) * rp.size(), hipMemcpyHostToDevice));
    const size_t nv = (size_t) m->n_vals();
    const bool direct = dtype == BMFP_F64 && layout == BMFP_ROW_MAJOR;
    DeviceBuffer staging;
    double *gen = nullptr;
    if (direct) {
      BMFP_TRY(m->values.alloc(nv * sizeof(double)));
      gen = m->values.as<double>();
    } else {
      BMFP_TRY(staging.alloc(nv * sizeof(double)));
      gen = staging.as<double>();
    }
    launch_gen_stencil(nx, ny, bs, kind, seed, m->row_ptr->as<int>(), m->columns->as<int>(), gen, nullptr, nullptr);
    BMFP_LAUNCHED();
    if (!direct) BMFP_TRY(convert_values(*m, gen, BMFP_ROW_MAJOR, nullptr));
    BMFP_HIP(hipDeviceSynchronize());
    *out = m.release();
    return BMFP_OK;
  });
}

extern "C" int bmfp_matrix_convert(bmfp_matrix src, int dtype, int layout, bmfp_matrix *out, bmfp_stream stream)
{
  return bmfp::guarded("bmfp_matrix_convert", [&]() -> int {
    if (!src || !out) return fail(BMFP_ERR_INVALID, "NULL argument");
    *out = nullptr;
    if (!src->is_bcsr || src->dtype != BMFP_F64)
      return fail(BMFP_ERR_UNSUPPORTED, "bmfp_matrix_convert needs an F64 BCSR source");
    if (dtype < BMFP_F64 || dtype > BMFP_F16 || (layout != BMFP_ROW_MAJOR && layout != BMFP_COL_MAJOR))
      return fail(BMFP_ERR_INVALID, "bad dtype/layout");
    std::unique_ptr<bmfp_matrix_t> m(new bmfp_matrix_t);
    m->is_bcsr = 1;
    m->dtype = dtype;
    m->layout = layout;
    m->n_rows = src->n_rows;
    m->n_cols = src->n_cols;
    m->bs = src->bs;
    m->nnzb = src->nnzb;
    m->max_row_blocks = src->max_row_blocks;
    m->mean_row_blocks = src->mean_row_blocks;
    m->device = src->device;
    m->row_ptr = src->row_ptr;
    m->columns = src->columns;
    BMFP_TRY(convert_values(*m, src->values.as<double>(), src->layout, static_cast<hipStream_t>(stream)));
    *out = m.release();
    return BMFP_OK;
  });
}

extern "C" int bmfp_matrix_destroy(bmfp_matrix m)
{
  return bmfp::guarded("bmfp_matrix_destroy", [&]() -> int {
    delete m;
    return BMFP_OK;
  });
}

extern "C" int bmfp_matrix_shape(bmfp_matrix m, int *is_bcsr, int *dtype, int *layout, int *n_rows, int *n_cols,
                                 int *bs, int *nnzb)
{
  return bmfp::guarded("bmfp_matrix_shape", [&]() -> int {
    if (!m) return fail(BMFP_ERR_INVALID, "matrix is NULL");
    if (is_bcsr) *is_bcsr = m->is_bcsr;
    if (dtype) *dtype = m->dtype;
    if (layout) *layout = m->layout;
    if (n_rows) *n_rows = m->n_rows;
    if (n_cols) *n_cols = m->n_cols;
    if (bs) *bs = m->bs;
    if (nnzb) *nnzb = m->nnzb;
    return BMFP_OK;
  });
}

extern "C" int bmfp_spmv_traffic(bmfp_matrix m, double *bytes, double *flops)
{
  return bmfp::guarded("bmfp_spmv_traffic", [&]() -> int {
    if (!m) return fail(BMFP_ERR_INVALID, "matrix is NULL");
    const double nv = (double) m->n_vals();
    const double vb = (double) dtype_size(m->dtype), xb = (double) vec_size(m->dtype);
    double b = nv * vb + 4.0 * m->nnzb + 4.0 * (m->n_rows + 1.0) + xb * (double) m->x_len() + xb * (double) m->y_len();
    if (m->dtype == BMFP_F16) b += 4.0 * m->n_rows;
    if (bytes) *bytes = b;
    if (flops) *flops = 2.0 * nv;
    return BMFP_OK;
  });
}

extern "C" int bmfp_matrix_to_csr(bmfp_matrix src, int dtype, bmfp_matrix *out, bmfp_stream stream)
{
  return bmfp::guarded("bmfp_matrix_to_csr", [&]() -> int {
    if (!src || !out) return fail(BMFP_ERR_INVALID, "NULL argument");
    *out = nullptr;
    if (!src->is_bcsr || src->dtype != BMFP_F64)
      return fail(BMFP_ERR_UNSUPPORTED, "bmfp_matrix_to_csr needs an F64 BCSR source");
    if (dtype != BMFP_F64 && dtype != BMFP_F32) return fail(BMFP_ERR_INVALID, "CSR dtype must be F64 or F32");
    const int64_t nnz = (int64_t) src->nnzb * src->bs * src->bs, rows = src->y_len();
    if (nnz > INT_MAX || rows > INT_MAX) return fail(BMFP_ERR_INVALID, "expanded matrix too large for int32");
    std::unique_ptr<bmfp_matrix_t> m(new bmfp_matrix_t);
    m->is_bcsr = 0;
    m->dtype = dtype;
    m->n_rows = (int) rows;
    m->n_cols = (int) src->x_len();
    m->bs = 1;
    m->nnzb = (int) nnz;
    m->max_row_blocks = src->max_row_blocks * src->bs;
    m->mean_row_blocks = src->mean_row_blocks * src->bs;
    m->csr_group = csr_group_for(m->mean_row_blocks);
    m->device = src->device;
    m->row_ptr = std::make_shared<DeviceBuffer>();
    m->columns = std::make_shared<DeviceBuffer>();
    BMFP_TRY(m->row_ptr->alloc(sizeof(int) * (size_t) (rows + 1)));
    BMFP_TRY(m->columns->alloc(sizeof(int) * (size_t) nnz));
    BMFP_TRY(m->values.alloc((dtype == BMFP_F64 ? sizeof(double) : sizeof(float)) * (size_t) nnz));
    BMFP_TRY(m->merge_carry.alloc(merge_carry_bytes(rows, nnz)));
    hipStream_t s = static_cast<hipStream_t>(stream);
    const unsigned groups = (unsigned) ((rows + 1 + 255) / 256);
    const int *rp = src->row_ptr->as<int>(), *cols = src->columns->as<int>();
    if (dtype == BMFP_F64)
      k_bcsr_to_csr<double><<<groups, 256, 0, s>>>(src->n_rows, src->bs, src->layout, rp, cols, src->values.as<double>(),
                                                   m->row_ptr->as<int>(), m->columns->as<int>(), m->values.as<double>());
    else
      k_bcsr_to_csr<float><<<groups, 256, 0, s>>>(src->n_rows, src->bs, src->layout, rp, cols, src->values.as<double>(),
                                                  m->row_ptr->as<int>(), m->columns->as<int>(), m->values.as<float>());
    BMFP_LAUNCHED();
    BMFP_HIP(hipStreamSynchronize(s));
    *out = m.release();
    return BMFP_OK;
  });
}

extern "C" int bmfp_matrix_download_structure(bmfp_matrix m, int *row_ptr, int *columns)
{
  return bmfp::guarded("bmfp_matrix_download_structure", [&]() -> int {
    if (!m || !row_ptr || (m->nnzb && !columns)) return fail(BMFP_ERR_INVALID, "NULL argument");
    BMFP_HIP(hipMemcpy(row_ptr, m->row_ptr->ptr, sizeof(int) * ((size_t) m->n_rows + 1), hipMemcpyDeviceToHost));
    if (m->nnzb) BMFP_HIP(hipMemcpy(columns, m->columns->ptr, sizeof(int) * (size_t) m->nnzb, hipMemcpyDeviceToHost));
    return BMFP_OK;
  });
}

extern "C" int bmfp_matrix_download_values(bmfp_matrix m, double *values)
{
  return bmfp::guarded("bmfp_matrix_download_values", [&]() -> int {
    if (!m || !values) return fail(BMFP_ERR_INVALID, "NULL argument");
    const size_t nv = (size_t) m->n_vals();
    if (!nv) return BMFP_OK;
    if (m->dtype == BMFP_F64) {
      BMFP_HIP(hipMemcpy(values, m->values.ptr, nv * sizeof(double), hipMemcpyDeviceToHost));
    } else if (m->dtype == BMFP_F32) {
      std::vector<float> f(nv);
      BMFP_HIP(hipMemcpy(f.data(), m->values.ptr, nv * sizeof(float), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < nv; i++) values[i] = f[i];
    } else {
      std::vector<__half> h(nv);
      std::vector<float> sc((size_t) m->n_rows);
      std::vector<int> rp((size_t) m->n_rows + 1);
      BMFP_HIP(hipMemcpy(h.data(), m->values.ptr, nv * sizeof(__half), hipMemcpyDeviceToHost));
      BMFP_HIP(hipMemcpy(sc.data(), m->scale.ptr, sc.size() * sizeof(float), hipMemcpyDeviceToHost));
      BMFP_HIP(hipMemcpy(rp.data(), m->row_ptr->ptr, rp.size() * sizeof(int), hipMemcpyDeviceToHost));
      const size_t bs2 = (size_t) m->bs * m->bs;
      for (int br = 0; br < m->n_rows; br++)
        for (size_t e = (size_t) rp[br] * bs2; e < (size_t) rp[br + 1] * bs2; e++)
          values[e] = (double) __half2float(h[e]) * (double) sc[br];
    }
    return BMFP_OK;
  });
}

// ---- SpMV ---------------------------------------------------------------------------
static int spmv_dispatch(bmfp_matrix A, int variant, const void *x, void *y, hipStream_t s)
{
  if (variant == BMFP_SPMV_AUTO) variant = auto_variant(*A);
  if (variant == BMFP_BCSR_LIBRARY) return launch_library(*A, x, y, s);
  const bool csr_family = variant == BMFP_CSR_SCALAR || variant == BMFP_CSR_VECTOR || variant == BMFP_CSR_MERGE;
  if (A->is_bcsr) {
    if (csr_family) return fail(BMFP_ERR_INVALID, "CSR variant on a BCSR matrix");
    return launch_bcsr(*A, variant, x, y, s);
  }
  if (!csr_family) return fail(BMFP_ERR_INVALID, "BCSR variant on a CSR matrix");
  return launch_csr(*A, variant, x, y, s);
}

extern "C" int bmfp_spmv(bmfp_matrix A, int variant, const void *x, void *y, bmfp_stream stream)
{
  return bmfp::guarded("bmfp_spmv", [&]() -> int {
    if (!A || (A->n_rows && (!x || !y))) return fail(BMFP_ERR_INVALID, "NULL argument");
    return spmv_dispatch(A, variant, x, y, static_cast<hipStream_t>(stream));
  });
}

extern "C" int bmfp_spmv_timed(bmfp_matrix A, int variant, const void *x, void *y, int warmup, int reps,
                               bmfp_stream stream, double *avg_seconds)
{
  return bmfp::guarded("bmfp_spmv_timed", [&]() -> int {
    if (!A || !avg_seconds || reps <= 0 || warmup < 0) return fail(BMFP_ERR_INVALID, "bad arguments");
    hipStream_t s = static_cast<hipStream_t>(stream);
    for (int i = 0; i < warmup; i++) BMFP_TRY(spmv_dispatch(A, variant, x, y, s));
    hipEvent_t e0, e1;
    BMFP_HIP(hipEventCreate(&e0));
    BMFP_HIP(hipEventCreate(&e1));
    int st = BMFP_OK;
    if (hipEventRecord(e0, s) != hipSuccess) st = fail(BMFP_ERR_HIP, "hipEventRecord failed");
    for (int i = 0; i < reps && st == BMFP_OK; i++) st = spmv_dispatch(A, variant, x, y, s);
    float ms = 0.f;
    if (st == BMFP_OK) {
      hipError_t e = hipEventRecord(e1, s);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess) st = hip_fail(e, "event timing", __FILE__, __LINE__);
    }
    (void) hipEventDestroy(e0);
    (void) hipEventDestroy(e1);
    if (st == BMFP_OK) *avg_seconds = ms * 1e-3 / reps;
    return st;
  });
}

extern "C" int bmfp_spmv_kernel_symbol(bmfp_matrix A, int variant, char *buf, size_t len)
{
  return bmfp::guarded("bmfp_spmv_kernel_symbol", [&]() -> int {
    if (!A || !buf || len == 0) return fail(BMFP_ERR_INVALID, "NULL argument");
    const void *fn = nullptr;
    BMFP_TRY(spmv_kernel_fn(*A, variant, &fn));
    const char *name = hipKernelNameRefByPtr(fn, nullptr);
    if (!name) return fail(BMFP_ERR_HIP, "hipKernelNameRefByPtr found no kernel name");
    const size_t n = std::strlen(name);
    if (n + 1 > len) return fail(BMFP_ERR_INVALID, "symbol of %zu bytes does not fit a %zu-byte buffer", n, len);
    std::memcpy(buf, name, n + 1);
    return BMFP_OK;
  });
}

extern "C" int bmfp_spmv_auto_variant(bmfp_matrix A, int *variant)
{
  return bmfp::guarded("bmfp_spmv_auto_variant", [&]() -> int {
    if (!A || !variant) return fail(BMFP_ERR_INVALID, "NULL argument");
    *variant = auto_variant(*A);
    return BMFP_OK;
  });
}
