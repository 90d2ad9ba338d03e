// Preparation of the low-precision candidate search.
//
// Rows are centred on a common vector mu (the feature means of up to 65,536
// train rows), rounded to bf16 and zero-padded to the kernel's feature count.
// Distances are translation invariant, and centring shrinks |x.y| and so the
// rounding error.  For every row we keep |q(r)|^2 (the C operand of the MFMA
// search) and an upper bound of the rounding error |r - mu - q(r)|, which the
// merge kernel uses to prove that the exact re-rank saw every true neighbour:
// by the triangle inequality | |q(x)-q(y)| - |x-y| | <= err(x) + err(y).
// There is no counterpart in the reference (its search is float64 only).
#include "knn_common.hpp"
#include "knn_kernels.hpp"

namespace knn {

namespace {

constexpr int kSampleRowsPerBlock = 256;

__global__ __launch_bounds__(kThreads) void column_sum_kernel(const float* __restrict__ Y,
                                                              int64_t n, int d,
                                                              float* __restrict__ mu_sum) {
  const int64_t r0 = (int64_t)blockIdx.x * kSampleRowsPerBlock;
  const int64_t r1 = r0 + kSampleRowsPerBlock < n ? r0 + kSampleRowsPerBlock : n;
  for (int f = threadIdx.x; f < d; f += kThreads) {
    float s = 0.f;
    for (int64_t r = r0; r < r1; ++r) s += Y[r * d + f];
    atomicAdd(&mu_sum[f], s);
  }
}

__global__ void column_scale_kernel(const float* __restrict__ mu_sum, int d, float inv,
                                    float* __restrict__ mu) {
  for (int f = threadIdx.x; f < d; f += blockDim.x) mu[f] = mu_sum[f] * inv;
}

__global__ void column_zero_kernel(float* __restrict__ p, int d) {
  for (int f = threadIdx.x; f < d; f += blockDim.x) p[f] = 0.f;
}

// One wave per row.
__global__ __launch_bounds__(kThreads) void quantize_bf16_kernel(
    const float* __restrict__ src, int64_t n, int d, int dm, const float* __restrict__ mu,
    __bf16* __restrict__ dst, float* __restrict__ norm, float* __restrict__ err,
    DeviceScalars* sc, int track_max) {
  __shared__ float wn[kThreads / 64];
  __shared__ float we[kThreads / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  float best_n = 0.f, best_e = 0.f;
  for (int64_t r = (int64_t)blockIdx.x * 4 + w; r < n; r += (int64_t)gridDim.x * 4) {
    float ns = 0.f;
    double es = 0.0;
    for (int f = lane; f < dm; f += 64) {
      float q = 0.f;
      if (f < d) {
        const float x = src[r * d + f];
        const float cen = x - mu[f];
        const __bf16 qb = (__bf16)cen;
        q = (float)qb;
        dst[r * dm + f] = qb;
        const double e = ((double)x - (double)mu[f]) - (double)q;
        es = fma(e, e, es);
      } else {
        dst[r * dm + f] = (__bf16)0.f;
      }
      ns = fmaf(q, q, ns);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      ns += __shfl_xor(ns, o, 64);
      es += __shfl_xor(es, o, 64);
    }
    // Round the error bound up: sqrt in double, then one float ulp of margin.
    const float eb = __uint_as_float(__float_as_uint((float)sqrt(es)) + 1u);  // next float up
    if (lane == 0) {
      norm[r] = ns;
      if (err) err[r] = eb;
    }
    best_n = fmaxf(best_n, ns);
    best_e = fmaxf(best_e, eb);
  }
  if (!track_max) return;
  if (lane == 0) {
    wn[w] = best_n;
    we[w] = best_e;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float mn = wn[0], me = we[0];
    for (int i = 1; i < kThreads / 64; ++i) {
      mn = fmaxf(mn, wn[i]);
      me = fmaxf(me, we[i]);
    }
    atomicMax(&sc->max_ynorm_key, ordered_key((double)mn));
    atomicMax(&sc->max_yerr_key, ordered_key((double)me));
  }
}

}  // namespace

void launch_column_mean(const float* Y, int64_t n_sample, int d, float* mu_sum, float* mu,
                        hipStream_t s) {
  hipLaunchKernelGGL(column_zero_kernel, dim3(1), dim3(kThreads), 0, s, mu_sum, d);
  const int64_t blocks = ceil_div(n_sample, kSampleRowsPerBlock);
  hipLaunchKernelGGL(column_sum_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, s, Y, n_sample,
                     d, mu_sum);
  hipLaunchKernelGGL(column_scale_kernel, dim3(1), dim3(kThreads), 0, s, mu_sum, d,
                     1.0f / (float)n_sample, mu);
}

void launch_quantize_bf16(const float* src, int64_t n, int d, int dm, const float* mu, void* dst,
                          float* norm, float* err, DeviceScalars* sc, bool track_max,
                          hipStream_t s) {
  if (n <= 0) return;
  const int64_t blocks = ceil_div(n, 4) < 16384 ? ceil_div(n, 4) : 16384;
  hipLaunchKernelGGL(quantize_bf16_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, s, src, n, d,
                     dm, mu, reinterpret_cast<__bf16*>(dst), norm, err, sc, track_max ? 1 : 0);
}

}  // namespace knn
