// Measured ceilings of the device: scalar (non-packed) FP32 FMA and add
// rates, and a plain copy's bandwidth.  They sit beside the published peaks
// as the denominators of the roofline fractions the benchmark reports: the
// flat filter's arithmetic is mostly adds and subtracts with a few
// multiplies, and this library is built without packed FP32 instructions,
// so the rate that bounds it is measured here, not taken from the data sheet.
#include <hip/hip_runtime.h>

#include "mfhip.h"
#include "mfh_internal.h"

namespace {

// W independent chains per lane, `iters` x 16 dependent steps each.  The
// operands are kernel arguments, so nothing folds; the chains are summed and
// stored (to a slot no launch reads unless the sum is a sentinel value) so
// nothing is dead.
template <int OP, int W>
__global__ __launch_bounds__(256) void valu_probe_kernel(float *out, int iters, float a, float b) {
  float x[W];
#pragma unroll
  for (int i = 0; i < W; i++) x[i] = (float)(threadIdx.x + i) * 1e-3f;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int r = 0; r < 16; r++) {
#pragma unroll
      for (int i = 0; i < W; i++) {
        if (OP == 0)
          x[i] = __builtin_fmaf(x[i], a, b);
        else
          x[i] = x[i] + b;
      }
    }
  }
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < W; i++) s += x[i];
  if (s == 1234.5f) out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void copy_probe_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                         size_t n4) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) dst[i] = src[i];
}

template <int OP, int W>
float time_valu(hipStream_t s, float *out, int blocks, int iters, hipEvent_t e0, hipEvent_t e1) {
  hipLaunchKernelGGL((valu_probe_kernel<OP, W>), dim3(blocks), dim3(256), 0, s, out, iters, 0.999f, 1e-7f);
  hipEventRecord(e0, s);
  hipLaunchKernelGGL((valu_probe_kernel<OP, W>), dim3(blocks), dim3(256), 0, s, out, iters, 0.999f, 1e-7f);
  hipEventRecord(e1, s);
  hipEventSynchronize(e1);
  float ms = 0.f;
  hipEventElapsedTime(&ms, e0, e1);
  return ms;
}

}  // namespace

using mfh::fail;
using mfh::hip_ok;

extern "C" int mfh_probe_valu(int device, int op, int chains, int blocks, int iters, double *ms_out,
                              double *flop_out) {
  if (op != 0 && op != 1) return fail("op must be 0 (FMA) or 1 (add)");
  if (blocks < 1 || iters < 1) return fail("blocks and iters must be >= 1");
  if (!hip_ok(hipSetDevice(device), "hipSetDevice")) return -1;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float *out = nullptr;
  if (!hip_ok(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate")) return -1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  int rc = 0;
  float ms = -1.f;
  if (!hip_ok(hipMalloc(&out, (size_t)blocks * 256 * sizeof(float)), "hipMalloc")) {
    rc = -1;
  } else {
    switch (op * 100 + chains) {
      case 4: ms = time_valu<0, 4>(s, out, blocks, iters, e0, e1); break;
      case 8: ms = time_valu<0, 8>(s, out, blocks, iters, e0, e1); break;
      case 16: ms = time_valu<0, 16>(s, out, blocks, iters, e0, e1); break;
      case 32: ms = time_valu<0, 32>(s, out, blocks, iters, e0, e1); break;
      case 104: ms = time_valu<1, 4>(s, out, blocks, iters, e0, e1); break;
      case 108: ms = time_valu<1, 8>(s, out, blocks, iters, e0, e1); break;
      case 116: ms = time_valu<1, 16>(s, out, blocks, iters, e0, e1); break;
      case 132: ms = time_valu<1, 32>(s, out, blocks, iters, e0, e1); break;
      default: rc = fail("chains must be 4, 8, 16 or 32");
    }
    if (rc == 0 && !hip_ok(hipGetLastError(), "valu probe")) rc = -1;
  }
  if (out) hipFree(out);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipStreamDestroy(s);
  if (rc < 0) return -1;
  if (ms_out) *ms_out = ms;
  if (flop_out) *flop_out = (double)blocks * 256 * (double)iters * 16 * chains * (op == 0 ? 2.0 : 1.0);
  return 0;
}

extern "C" int mfh_probe_copy(int device, size_t bytes, double *ms_out) {
  if (bytes < 16 || bytes % 16) return fail("bytes must be a positive multiple of 16");
  if (!hip_ok(hipSetDevice(device), "hipSetDevice")) return -1;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float4 *src = nullptr, *dst = nullptr;
  if (!hip_ok(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate")) return -1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  int rc = 0;
  float ms = -1.f;
  if (!hip_ok(hipMalloc(&src, bytes), "hipMalloc") || !hip_ok(hipMalloc(&dst, bytes), "hipMalloc")) {
    rc = -1;
  } else {
    const size_t n4 = bytes / 16;
    hipMemsetAsync(src, 0, bytes, s);
    const int grid = 256 * 32;
    hipLaunchKernelGGL(copy_probe_kernel, dim3(grid), dim3(256), 0, s, src, dst, n4);
    hipEventRecord(e0, s);
    hipLaunchKernelGGL(copy_probe_kernel, dim3(grid), dim3(256), 0, s, src, dst, n4);
    hipEventRecord(e1, s);
    hipEventSynchronize(e1);
    hipEventElapsedTime(&ms, e0, e1);
    if (!hip_ok(hipGetLastError(), "copy probe")) rc = -1;
  }
  if (src) hipFree(src);
  if (dst) hipFree(dst);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipStreamDestroy(s);
  if (rc < 0) return -1;
  if (ms_out) *ms_out = ms;
  return 0;
}
