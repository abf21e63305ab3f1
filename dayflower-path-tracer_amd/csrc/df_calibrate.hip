// Measured roofs of the GPU the renderer runs on: microbenchmarks of the
// rates the renderer's kernels are bound by, so that a kernel's achieved
// work can be set against what this chip sustains rather than against a
// datasheet figure alone.
//
//   FP32 VALU   fused multiply-adds, eight independent chains per lane,
//               issued as single v_fma_f32 (packed FP32 is disabled for
//               this library), at eight waves per SIMD: FLOP/s and wave
//               instructions/s;
//   L2 read     every workgroup streams a 256 KiB window of a 2 MiB buffer
//               (each XCD's L2 holds it; far larger than a CU's L1) in
//               coalesced 16-byte loads: bytes/s requested from L2;
//   MALL read   grid-stride passes over a 128 MiB buffer, beyond the L2s
//               and within the 256 MiB Infinity Cache; the rate of the
//               repeated passes (five passes minus one, to drop the first
//               pass's HBM reads);
//   HBM read    one grid-stride pass over 2 GiB.
//
// No counterpart in the reference: measurement only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "../../include/dayflower_hip.h"

namespace dfk_roof {

constexpr int kThreads = 256;

__global__ __launch_bounds__(kThreads) void roof_valu(float* out, int iters, float a, float b) {
  const float s = static_cast<float>(threadIdx.x) * 1.0e-3f;
  float x0 = s, x1 = s + 1.0f, x2 = s + 2.0f, x3 = s + 3.0f, x4 = s + 4.0f, x5 = s + 5.0f, x6 = s + 6.0f,
        x7 = s + 7.0f;
  for (int i = 0; i < iters; ++i) {
    x0 = fmaf(x0, a, b);
    x1 = fmaf(x1, a, b);
    x2 = fmaf(x2, a, b);
    x3 = fmaf(x3, a, b);
    x4 = fmaf(x4, a, b);
    x5 = fmaf(x5, a, b);
    x6 = fmaf(x6, a, b);
    x7 = fmaf(x7, a, b);
  }
  const float sum = ((x0 + x1) + (x2 + x3)) + ((x4 + x5) + (x6 + x7));
  if (sum == -1.0f) out[blockIdx.x] = sum;  // never true for these inputs: keeps the chains live, writes nothing
}

// Each workgroup reads `window_f4` float4s starting at a window chosen by its
// index, `passes` times.
__global__ __launch_bounds__(kThreads) void roof_window(const float4* __restrict__ buf, int64_t n_f4,
                                                         int64_t window_f4, int passes, float* out) {
  const int64_t windows = n_f4 / window_f4;
  const float4* w = buf + (blockIdx.x % windows) * window_f4;
  float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  for (int p = 0; p < passes; ++p) {
    for (int64_t i = threadIdx.x; i < window_f4; i += kThreads) {
      const float4 v = w[i];
      acc.x += v.x;
      acc.y += v.y;
      acc.z += v.z;
      acc.w += v.w;
    }
  }
  const float sum = (acc.x + acc.y) + (acc.z + acc.w);
  if (sum == -1.0f) out[blockIdx.x] = sum;  // the buffer holds zeros: never true
}

// Grid-stride passes over the whole buffer.
__global__ __launch_bounds__(kThreads) void roof_stream(const float4* __restrict__ buf, int64_t n_f4, int passes,
                                                         float* out) {
  float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kThreads;
  for (int p = 0; p < passes; ++p) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; i < n_f4; i += stride) {
      const float4 v = buf[i];
      acc.x += v.x;
      acc.y += v.y;
      acc.z += v.z;
      acc.w += v.w;
    }
  }
  const float sum = (acc.x + acc.y) + (acc.z + acc.w);
  if (sum == -1.0f) out[blockIdx.x] = sum;
}

struct Timer {
  hipStream_t s = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  bool ok = false;
  Timer() {
    ok = hipStreamCreate(&s) == hipSuccess && hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess;
  }
  ~Timer() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (s) (void)hipStreamDestroy(s);
  }
  // Best of `reps` timed launches after one untimed, in seconds (< 0 on error).
  template <class F>
  double best(F&& launch, int reps) {
    launch();
    if (hipStreamSynchronize(s) != hipSuccess) return -1.0;
    double t = 1e30;
    for (int r = 0; r < reps; ++r) {
      if (hipEventRecord(a, s) != hipSuccess) return -1.0;
      launch();
      if (hipEventRecord(b, s) != hipSuccess || hipEventSynchronize(b) != hipSuccess) return -1.0;
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return -1.0;
      t = std::min(t, static_cast<double>(ms) * 1e-3);
    }
    return hipGetLastError() == hipSuccess ? t : -1.0;
  }
};

}  // namespace dfk_roof

using namespace dfk_roof;

extern "C" int df_measure_roof(int32_t device, int32_t kind, double* rate, double* work_per_launch) {
  if (!rate) return DF_EINVAL;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return DF_ENODEVICE;
  if (device < 0 || device >= n) return DF_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return DF_EHIP;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return DF_EHIP;
  Timer tm;
  if (!tm.ok) return DF_EHIP;
  float* out = nullptr;
  if (hipMalloc(&out, sizeof(float) * 65536) != hipSuccess) return DF_ENOMEM;
  float4* buf = nullptr;
  double work = 0.0, secs = -1.0;
  int rc = DF_OK;
  if (kind == DF_ROOF_VALU_FP32) {
    const int blocks = cus * 4 * 8 * 64 / kThreads;  // eight waves per SIMD
    const int iters = 16384;
    work = 2.0 * 8.0 * iters * static_cast<double>(blocks) * kThreads;
    secs = tm.best([&] { hipLaunchKernelGGL(roof_valu, dim3(blocks), dim3(kThreads), 0, tm.s, out, iters,
                                            0.9999999f, 1.0e-7f); }, 3);
  } else if (kind == DF_ROOF_L2_READ) {
    const int64_t bytes = int64_t{2} << 20, window = int64_t{256} << 10;
    const int blocks = cus * 8, passes = 16;
    if (hipMalloc(&buf, bytes) != hipSuccess || hipMemsetAsync(buf, 0, bytes, tm.s) != hipSuccess) {
      rc = DF_ENOMEM;
    } else {
      work = static_cast<double>(window) * passes * blocks;
      secs = tm.best([&] { hipLaunchKernelGGL(roof_window, dim3(blocks), dim3(kThreads), 0, tm.s, buf, bytes / 16,
                                              window / 16, passes, out); }, 3);
    }
  } else if (kind == DF_ROOF_MALL_READ) {
    const int64_t bytes = int64_t{128} << 20;
    const int blocks = cus * 8;
    if (hipMalloc(&buf, bytes) != hipSuccess || hipMemsetAsync(buf, 0, bytes, tm.s) != hipSuccess) {
      rc = DF_ENOMEM;
    } else {
      const double t1 = tm.best([&] { hipLaunchKernelGGL(roof_stream, dim3(blocks), dim3(kThreads), 0, tm.s, buf,
                                                         bytes / 16, 1, out); }, 3);
      const double t5 = tm.best([&] { hipLaunchKernelGGL(roof_stream, dim3(blocks), dim3(kThreads), 0, tm.s, buf,
                                                         bytes / 16, 5, out); }, 3);
      work = 4.0 * static_cast<double>(bytes);
      secs = (t1 < 0.0 || t5 < 0.0) ? -1.0 : t5 - t1;
    }
  } else if (kind == DF_ROOF_HBM_READ) {
    const int64_t bytes = int64_t{2} << 30;
    const int blocks = cus * 8;
    if (hipMalloc(&buf, bytes) != hipSuccess || hipMemsetAsync(buf, 0, bytes, tm.s) != hipSuccess) {
      rc = DF_ENOMEM;
    } else {
      work = static_cast<double>(bytes);
      secs = tm.best([&] { hipLaunchKernelGGL(roof_stream, dim3(blocks), dim3(kThreads), 0, tm.s, buf, bytes / 16,
                                              1, out); }, 3);
    }
  } else {
    rc = DF_EINVAL;
  }
  if (rc == DF_OK && secs <= 0.0) rc = DF_EHIP;
  if (rc == DF_OK) {
    *rate = work / secs;
    if (work_per_launch) *work_per_launch = work;
  }
  if (buf) (void)hipFree(buf);
  (void)hipFree(out);
  return rc;
}
