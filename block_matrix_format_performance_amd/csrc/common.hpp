// common.hpp -- internal helpers shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <atomic>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>

#include "bmfp.h"

namespace bmfp {

// Thread-local last-error text, returned by bmfp_last_error().
void set_error(const char *fmt, ...);
const char *get_error();

// Records a formatted error and returns `code`.
int fail(int code, const char *fmt, ...);

// Runs the body of a C entry point: a host allocation that fails inside it
// (arguments or files declaring more than the host can hold) returns
// BMFP_ERR_NO_MEMORY instead of throwing across the C ABI.
template <class F> int guarded(const char *what, F &&body)
{
  try {
    return body();
  } catch (const std::bad_alloc &) {
    return fail(BMFP_ERR_NO_MEMORY, "%s: host allocation failed", what);
  } catch (const std::length_error &) {
    return fail(BMFP_ERR_NO_MEMORY, "%s: host allocation failed", what);
  }
}

inline int hip_fail(hipError_t e, const char *what, const char *file, int line)
{
  return fail(e == hipErrorOutOfMemory ? BMFP_ERR_NO_MEMORY : BMFP_ERR_HIP,
              "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
}

#define BMFP_HIP(call)                                                         \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) return ::bmfp::hip_fail(e_, #call, __FILE__, __LINE__); \
  } while (0)

#define BMFP_TRY(call)                                                         \
  do {                                                                         \
    int s_ = (call);                                                           \
    if (s_ != BMFP_OK) return s_;                                              \
  } while (0)

// Launch-error check after a <<<>>> launch.
#define BMFP_LAUNCHED() BMFP_HIP(hipGetLastError())

inline size_t dtype_size(int dtype)
{
  return dtype == BMFP_F64 ? 8 : dtype == BMFP_F32 ? 4 : 2;
}
// Vector (x, y) element size for a matrix storage dtype.
inline size_t vec_size(int dtype) { return dtype == BMFP_F64 ? 8 : 4; }

// Owning device buffer.
struct DeviceBuffer {
  void *ptr = nullptr;
  size_t bytes = 0;
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
  ~DeviceBuffer() { reset(); }
  int alloc(size_t n)
  {
    reset();
    if (n == 0) n = 256;
    BMFP_HIP(hipMalloc(&ptr, n));
    bytes = n;
    return BMFP_OK;
  }
  void reset()
  {
    if (ptr) (void) hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// Counter-based RNG (SplitMix64 finaliser over a (seed, stream)-keyed
// counter).  Host and device produce identical bits.
__host__ __device__ inline uint64_t rng_mix(uint64_t z)
{
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t rng_key(uint64_t seed, uint64_t stream)
{
  return rng_mix(seed * 0x9E3779B97F4A7C15ULL + stream);
}
__host__ __device__ inline uint64_t rng_at(uint64_t key, uint64_t index)
{
  return rng_mix(key + (index + 1) * 0x9E3779B97F4A7C15ULL);
}
__host__ __device__ inline double rng_unit(uint64_t key, uint64_t index)
{
  return (double) (rng_at(key, index) >> 11) * 0x1.0p-53;
}
__host__ __device__ inline double rng_sym(uint64_t key, uint64_t index)
{
  return 2.0 * rng_unit(key, index) - 1.0;
}

// RNG stream ids used by the synthetic workloads.
enum : uint64_t {
  STREAM_DD = 1,     // diagonally dominant stencil entries
  STREAM_K = 2,      // convective stencil: symmetric K
  STREAM_C = 3,      // convective stencil: skew C
  STREAM_X = 4,      // SpMV x vectors
  STREAM_XSTAR = 5,  // manufactured solutions
  STREAM_RHS = 6     // serving right-hand sides
};

}  // namespace bmfp
