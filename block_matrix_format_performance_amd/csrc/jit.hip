// jit.hip -- BCSR SpMV kernels specialised at run time by hipRTC for block
// sizes without a precompiled instantiation (the role of the reference's
// NVRTC-compiled bcsr_jit kernel, main.cpp:170-294).
//
// The kernel source below is compiled for the device's architecture with the
// block size, element type and block layout as compile-time constants: a
// lane group of G lanes (a power of two chosen from bs) per scalar row, lane
// g summing columns g, g+G, ... of every block of the row, then a shuffle
// reduction inside the group.  Modules are cached per (device, bs, dtype,
// layout) for the life of the process.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <chrono>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "matrix.hpp"

namespace bmfp {
namespace {

const char *const JIT_SOURCE = R"(
typedef BMFP_VT vt;
extern "C" __global__ void __launch_bounds__(256)
bmfp_bcsr_jit(long long rows, const int *__restrict__ rp, const int *__restrict__ cols,
              const vt *__restrict__ vals, const vt *__restrict__ x, vt *__restrict__ y)
{
  constexpr int BS = BMFP_BS, G = BMFP_G;
  const long long t = (long long) blockIdx.x * 256 + threadIdx.x;
  const long long row = t / G;
  const int g = (int) (t - row * G);
  const bool on = row < rows;
  vt acc = 0;
  if (on) {
    const long long br = row / BS;
    const int r = (int) (row - br * BS);
    const int b1 = rp[br + 1];
    for (int b = rp[br]; b < b1; ++b) {
      const vt *blk = vals + (long long) b * (BS * BS);
      const vt *xs = x + (long long) cols[b] * BS;
#pragma unroll
      for (int c0 = 0; c0 < BS; c0 += G) {
        const int c = c0 + g;
        if (c < BS) {
          const vt v = BMFP_COL_MAJOR ? blk[c * BS + r] : blk[r * BS + c];
          acc = __builtin_fma(v, xs[c], acc);
        }
      }
    }
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, G);
  if (on && g == 0) y[row] = acc;
}
)";

// Lanes per scalar row: one for small blocks, up to 16 for wide ones.
int jit_group(int bs) { return bs <= 8 ? 1 : bs <= 16 ? 2 : bs <= 32 ? 4 : bs <= 64 ? 8 : 16; }

struct JitKernel {
  hipModule_t module = nullptr;
  hipFunction_t fn = nullptr;
};

std::mutex jit_mutex;
std::map<std::tuple<int, int, int, int>, JitKernel> jit_cache;  // (device, bs, dtype, layout)

int jit_compile(int bs, int dtype, int layout, const std::string &arch, std::vector<char> *code, double *seconds)
{
  if (bs < 1) return fail(BMFP_ERR_INVALID, "bs must be >= 1");
  if (dtype != BMFP_F64 && dtype != BMFP_F32) return fail(BMFP_ERR_UNSUPPORTED, "JIT kernels are F64 or F32");
  if (layout != BMFP_ROW_MAJOR && layout != BMFP_COL_MAJOR) return fail(BMFP_ERR_INVALID, "bad layout %d", layout);
  const auto t0 = std::chrono::steady_clock::now();
  hiprtcProgram prog = nullptr;
  if (hiprtcCreateProgram(&prog, JIT_SOURCE, "bmfp_bcsr_jit.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
    return fail(BMFP_ERR_HIP, "hiprtcCreateProgram failed");
  const std::string o_arch = "--offload-arch=" + arch, o_bs = "-DBMFP_BS=" + std::to_string(bs),
                    o_g = "-DBMFP_G=" + std::to_string(jit_group(bs)),
                    o_vt = std::string("-DBMFP_VT=") + (dtype == BMFP_F64 ? "double" : "float"),
                    o_lay = "-DBMFP_COL_MAJOR=" + std::to_string(layout == BMFP_COL_MAJOR ? 1 : 0);
  // packed FP32 instructions stay off, as for the precompiled kernels
  const char *opts[] = {o_arch.c_str(), "-O3",     o_bs.c_str(),      o_g.c_str(),       o_vt.c_str(),
                        o_lay.c_str(),  "-Xclang", "-target-feature", "-Xclang",         "-packed-fp32-ops"};
  const hiprtcResult rc = hiprtcCompileProgram(prog, (int) (sizeof(opts) / sizeof(opts[0])), opts);
  if (rc != HIPRTC_SUCCESS) {
    size_t n = 0;
    (void) hiprtcGetProgramLogSize(prog, &n);
    std::string log(n + 1, '\0');
    (void) hiprtcGetProgramLog(prog, &log[0]);
    (void) hiprtcDestroyProgram(&prog);
    return fail(BMFP_ERR_HIP, "hipRTC compile of the bs=%d kernel failed: %s", bs, log.c_str());
  }
  size_t bytes = 0;
  (void) hiprtcGetCodeSize(prog, &bytes);
  code->resize(bytes);
  (void) hiprtcGetCode(prog, code->data());
  (void) hiprtcDestroyProgram(&prog);
  if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return BMFP_OK;
}

int jit_kernel(int bs, int dtype, int layout, JitKernel *out)
{
  int dev = 0;
  BMFP_HIP(hipGetDevice(&dev));
  const auto key = std::make_tuple(dev, bs, dtype, layout);
  std::lock_guard<std::mutex> lock(jit_mutex);
  auto it = jit_cache.find(key);
  if (it != jit_cache.end()) {
    *out = it->second;
    return BMFP_OK;
  }
  hipDeviceProp_t prop;
  BMFP_HIP(hipGetDeviceProperties(&prop, dev));
  std::vector<char> code;
  BMFP_TRY(jit_compile(bs, dtype, layout, prop.gcnArchName, &code, nullptr));
  JitKernel k;
  BMFP_HIP(hipModuleLoadData(&k.module, code.data()));
  const hipError_t e = hipModuleGetFunction(&k.fn, k.module, "bmfp_bcsr_jit");
  if (e != hipSuccess) {
    (void) hipModuleUnload(k.module);
    return hip_fail(e, "hipModuleGetFunction", __FILE__, __LINE__);
  }
  jit_cache[key] = k;
  *out = k;
  return BMFP_OK;
}

}  // namespace

int launch_bcsr_jit(const bmfp_matrix_t &A, const void *x, void *y, hipStream_t s)
{
  if (!A.is_bcsr) return fail(BMFP_ERR_INVALID, "the JIT kernel is a BCSR kernel");
  JitKernel k;
  BMFP_TRY(jit_kernel(A.bs, A.dtype, A.layout, &k));
  long long rows = (long long) A.n_rows * A.bs;
  const long long threads = rows * jit_group(A.bs);
  const unsigned groups = (unsigned) ((threads + 255) / 256);
  const int *rp = A.row_ptr->as<int>(), *cols = A.columns->as<int>();
  const void *vals = A.values.ptr;
  void *args[] = {&rows, &rp, &cols, &vals, &x, &y};
  BMFP_HIP(hipModuleLaunchKernel(k.fn, groups, 1, 1, 256, 1, 1, 0, s, args, nullptr));
  return BMFP_OK;
}

}  // namespace bmfp

using namespace bmfp;

extern "C" int bmfp_jit_build(int bs, int dtype, int layout, size_t *code_bytes, double *seconds)
{
  return bmfp::guarded("bmfp_jit_build", [&]() -> int {
    std::vector<char> code;
    BMFP_TRY(jit_compile(bs, dtype, layout, "gfx950", &code, seconds));
    if (code_bytes) *code_bytes = code.size();
    return BMFP_OK;
  });
}
