// Dense squared-distance matrix D[i][j] = sum_f (x_i[f] - y_j[f])^2.
//
// Counterpart of the SYCL backend's distance_squared_matrix / pw_d2m
// (sycl_ext/_pdist_aggregation_sycl.pyx:73-99,
// sycl_ext/parallel_knn_helper.hpp:26-55) and its kern_pwdist2 tile kernel
// (sycl_ext/parallel_knn.hpp:199-289): direct differences, no GEMM
// expansion, so no cancellation.  64 x 64 output tile per workgroup, both
// operand slices staged through LDS 16 features at a time, 4 x 4 outputs per
// thread; row-major output (n_test x n_train).  The differences and sums run
// in A (the input type for the dense matrix; double for the large-k search,
// whose distances are its final exact values).
#include "knn_common.hpp"
#include "knn_kernels.hpp"

namespace knn {

namespace {

constexpr int kT = 64, kF = 16;

template <class T, class A>
__global__ __launch_bounds__(kThreads) void sqdist_kernel(const T* __restrict__ X, int64_t n_q,
                                                          const T* __restrict__ Y, int64_t n_t,
                                                          int d, A* __restrict__ out) {
  __shared__ A xs[kF][kT + 1];
  __shared__ A ys[kF][kT + 1];
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const int64_t q0 = (int64_t)blockIdx.y * kT, y0 = (int64_t)blockIdx.x * kT;
  A acc[4][4] = {};
  for (int p0 = 0; p0 < d; p0 += kF) {
    for (int e = tid; e < kT * kF; e += kThreads) {
      const int r = e / kF, c = e % kF, p = p0 + c;
      xs[c][r] = (q0 + r < n_q && p < d) ? (A)X[(q0 + r) * d + p] : A(0);
      ys[c][r] = (y0 + r < n_t && p < d) ? (A)Y[(y0 + r) * d + p] : A(0);
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kF; ++c) {
      A xv[4], yv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) xv[i] = xs[c][ty + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) yv[j] = ys[c][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const A df = xv[i] - yv[j];
          acc[i][j] = fma(df, df, acc[i][j]);
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int64_t q = q0 + ty + 16 * i;
    if (q >= n_q) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t y = y0 + tx + 16 * j;
      if (y < n_t) out[q * n_t + y] = acc[i][j];
    }
  }
}

}  // namespace

template <class T>
void launch_sqdist(const T* X, int64_t n_q, const T* Y, int64_t n_t, int d, T* out,
                   hipStream_t s) {
  if (n_q <= 0 || n_t <= 0) return;
  dim3 grid((unsigned)ceil_div(n_t, kT), (unsigned)ceil_div(n_q, kT));
  hipLaunchKernelGGL((sqdist_kernel<T, T>), grid, dim3(kThreads), 0, s, X, n_q, Y, n_t, d, out);
}

template <class T>
void launch_sqdist_f64(const T* X, int64_t n_q, const T* Y, int64_t n_t, int d, double* out,
                       hipStream_t s) {
  if (n_q <= 0 || n_t <= 0) return;
  dim3 grid((unsigned)ceil_div(n_t, kT), (unsigned)ceil_div(n_q, kT));
  hipLaunchKernelGGL((sqdist_kernel<T, double>), grid, dim3(kThreads), 0, s, X, n_q, Y, n_t, d,
                     out);
}

template void launch_sqdist<float>(const float*, int64_t, const float*, int64_t, int, float*,
                                   hipStream_t);
template void launch_sqdist<double>(const double*, int64_t, const double*, int64_t, int, double*,
                                    hipStream_t);
template void launch_sqdist_f64<float>(const float*, int64_t, const float*, int64_t, int, double*,
                                       hipStream_t);
template void launch_sqdist_f64<double>(const double*, int64_t, const double*, int64_t, int,
                                        double*, hipStream_t);

}  // namespace knn
