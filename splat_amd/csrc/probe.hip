// probe.hip -- calibration kernels for the hardware counters (not part of the tracing library):
// a kernel of exactly known FP32 operation count, against which a rocprofv3 --pmc run reads what
// SQ_INSTS_VALU_FLOPS_FP32 (and the other VALU counters) report per operation.
#include <hip/hip_runtime.h>

#include <cstdint>

// Each thread starts four chains (one multiply, three adds), runs them as independent FMA chains of
// `iters` steps (4 * iters v_fma_f32), then combines them (two adds, one multiply):
// 8 * iters + 7 floating-point operations per thread, an FMA counted as two.
__global__ __launch_bounds__(256) void k_probe_fma(float* __restrict__ out, int64_t n, int32_t iters, float b,
                                                   float c) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float x0 = (float)(i & 255) * 1e-3f, x1 = x0 + 1.0f, x2 = x0 + 2.0f, x3 = x0 + 3.0f;
  for (int32_t k = 0; k < iters; ++k) {
    x0 = __builtin_fmaf(x0, b, c);
    x1 = __builtin_fmaf(x1, b, c);
    x2 = __builtin_fmaf(x2, b, c);
    x3 = __builtin_fmaf(x3, b, c);
  }
  out[i] = (x0 + x1) * (x2 + x3);
}

extern "C" {

// n threads (a multiple of 256), iters FMA steps per chain; returns 0 or the HIP error code.
int32_t splat_probe_fma(float* out, int64_t n, int32_t iters, void* stream) {
  if (!out || n < 256 || n % 256 != 0 || iters < 1) return -1;
  hipLaunchKernelGGL(k_probe_fma, dim3((unsigned)(n / 256)), dim3(256), 0, static_cast<hipStream_t>(stream), out, n,
                     iters, 0.999f, 1e-3f);
  return (int32_t)hipGetLastError();
}

// FP32 operations one splat_probe_fma call performs
int64_t splat_probe_fma_flops(int64_t n, int32_t iters) { return n * (8 * (int64_t)iters + 7); }

}  // extern "C"
