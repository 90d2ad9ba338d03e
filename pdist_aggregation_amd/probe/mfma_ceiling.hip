// Sustained matrix-core rate of the GPU this runs on, for the roofline of the
// search kernels: every wave issues NACC independent chains of one MFMA
// instruction from registers (no memory traffic in the timed loop), with
// per_cu workgroups of 4 waves per CU (pinned by a dynamic LDS allocation of
// 160 KiB / per_cu), timed with HIP events over several launches.  The
// datasheet peak assumes the maximum clock; under sustained MFMA load the
// clock, and with it this rate, is lower.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kNAcc = 4;

// Format 0: v_mfma_f32_32x32x16_bf16 (32*32*16 MACs); 1: v_mfma_f32_32x32x2_f32.
template <int FMT>
__global__ __launch_bounds__(256, 4) void ceiling_kernel(float* out, int iters, float seed) {
  f32x16 acc[kNAcc];
#pragma unroll
  for (int g = 0; g < kNAcc; ++g)
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[g][i] = 0.f;
  if constexpr (FMT == 0) {
    bf16x8 a[4], b[4];
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        a[s][e] = (__bf16)(seed * (float)(threadIdx.x + s + e));
        b[s][e] = (__bf16)(seed * (float)(threadIdx.x + 2 * s - e));
      }
    for (int it = 0; it < iters; ++it) {
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int g = 0; g < kNAcc; ++g)
          acc[g] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[s], b[(s + g) & 3], acc[g], 0, 0, 0);
    }
  } else {
    float a[4], b[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      a[s] = seed * (float)(threadIdx.x + s);
      b[s] = seed * (float)(threadIdx.x - s);
    }
    for (int it = 0; it < iters; ++it) {
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int g = 0; g < kNAcc; ++g)
          acc[g] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[s], b[(s + g) & 3], acc[g], 0, 0, 0);
    }
  }
  float r = 0.f;
#pragma unroll
  for (int g = 0; g < kNAcc; ++g)
#pragma unroll
    for (int i = 0; i < 16; ++i) r += acc[g][i];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = r;  // keeps the chains live
}

}  // namespace

// Measured MFMA rate in TFLOP/s (2 flops per MAC) for format fmt (0 = bf16
// 32x32x16, 1 = fp32 32x32x2) with per_cu (1..4) workgroups per CU and iters
// loop trips per wave; the mean over `reps` launches after one warm-up.
// Returns 0, or a negative value when a HIP call fails.
extern "C" int pdist_probe_mfma_tflops(int fmt, int per_cu, int iters, int reps, double* tflops,
                                       double* ms_per_launch) {
  if (fmt < 0 || fmt > 1 || per_cu < 1 || per_cu > 4 || iters < 1 || reps < 1) return -1;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -2;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return -2;
  const int blocks = per_cu * cus;
  const size_t lds = (size_t)(160 * 1024 / per_cu) - 1024;
  auto kern = fmt == 0 ? ceiling_kernel<0> : ceiling_kernel<1>;
  if (hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
    return -2;
  float* out = nullptr;
  if (hipMalloc(&out, (size_t)blocks * 256 * sizeof(float)) != hipSuccess) return -3;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = -2;
  float ms = 0.f;
  if (rc == 0) {
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), lds, 0, out, iters, 1e-3f);
    bool ok = hipEventRecord(e0, 0) == hipSuccess;
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), lds, 0, out, iters, 1e-3f);
    ok = ok && hipEventRecord(e1, 0) == hipSuccess;
    if (!ok || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess || hipGetLastError() != hipSuccess)
      rc = -4;
  }
  if (rc == 0) {
    const double macs = (fmt == 0 ? 32.0 * 32 * 16 : 32.0 * 32 * 2);
    const double flops = 2.0 * macs * blocks * 4 /*waves*/ * (double)iters * 4 /*steps*/ * kNAcc;
    *ms_per_launch = ms / reps;
    *tflops = flops / (*ms_per_launch * 1e-3) / 1e12;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(out);
  return rc;
}
