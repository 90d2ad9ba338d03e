// Backward of the k-NN distances with respect to both point sets.
//
// The forward pass fixes idx (n_q x k).  For d_ij = |x_i - y_idx(i,j)|^2
//   dL/dx_i        =  sum_j g_ij * c_ij * (x_i - y_idx(i,j))
//   dL/dy_idx(i,j) -= g_ij * c_ij * (x_i - y_idx(i,j))
// with c = 2 for squared distances and c = 1/|x_i - y| for Euclidean ones (0
// at coincident points).  Differences and distances are recomputed here, not
// stored by the forward pass.  dX has one writer per element; dY is scattered
// by neighbour index with atomics, so its summation order (and last bits) is
// not deterministic.  There is no counterpart in the reference, whose search
// is not differentiable.
#include "knn_common.hpp"
#include "knn_kernels.hpp"

namespace knn {

namespace {

template <class T>
__global__ void zero_kernel(T* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = T(0);
}

// G lanes per query (G = power of two <= 64), kThreads / G queries per block.
template <class T, int G>
__global__ __launch_bounds__(kThreads) void backward_kernel(BackwardArgs<T> a) {
  const int lane_g = threadIdx.x & (G - 1);
  const int64_t i = (int64_t)blockIdx.x * (kThreads / G) + threadIdx.x / G;
  if (i >= a.n_q) return;  // whole groups leave together (G divides 64)
  const T* xr = a.X + i * a.d;
  T* gx = a.grad_X ? a.grad_X + i * a.d : nullptr;
  constexpr int kMaxF = 4;  // features per lane kept in registers when d <= G * kMaxF
  const bool small = a.d <= G * kMaxF;
  T accx[kMaxF] = {};
  for (int j = 0; j < a.k; ++j) {
    const int64_t m = a.idx[i * a.k + j];
    const T g = a.grad[i * a.k + j];
    if (m < 0 || m >= a.n_t || g == T(0)) continue;  // uniform within the group
    const T* yr = a.Y + m * a.d;
    T c = T(2);
    if (!a.squared) {
      T s = T(0);
      for (int f = lane_g; f < a.d; f += G) {
        const T df = xr[f] - yr[f];
        s = fma(df, df, s);
      }
#pragma unroll
      for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
      const T e = sqrt(s);
      c = e > T(0) ? T(1) / e : T(0);
    }
    const T w = g * c;
    int slot = 0;
    for (int f = lane_g; f < a.d; f += G, ++slot) {
      const T df = w * (xr[f] - yr[f]);
      if (a.grad_Y) atomicAdd(&a.grad_Y[m * a.d + f], -df);
      if (gx) {
        if (small) {
#pragma unroll
          for (int s2 = 0; s2 < kMaxF; ++s2)
            if (s2 == slot) accx[s2] += df;
        } else {
          gx[f] += df;  // this group is the only writer of row i
        }
      }
    }
  }
  if (gx && small) {
    int slot = 0;
    for (int f = lane_g; f < a.d; f += G, ++slot) {
#pragma unroll
      for (int s2 = 0; s2 < kMaxF; ++s2)
        if (s2 == slot) gx[f] = accx[s2];
    }
  }
}

template <class T, int G>
void launch_g(const BackwardArgs<T>& a, hipStream_t s) {
  const int64_t blocks = ceil_div(a.n_q, kThreads / G);
  hipLaunchKernelGGL((backward_kernel<T, G>), dim3((unsigned)blocks), dim3(kThreads), 0, s, a);
}

}  // namespace

template <class T>
void launch_backward(const BackwardArgs<T>& a, hipStream_t s) {
  const int64_t nz_y = a.grad_Y ? a.n_t * a.d : 0;
  if (nz_y > 0) {
    const int64_t b = ceil_div(nz_y, kThreads) < 8192 ? ceil_div(nz_y, kThreads) : 8192;
    hipLaunchKernelGGL(zero_kernel<T>, dim3((unsigned)b), dim3(kThreads), 0, s, a.grad_Y, nz_y);
  }
  if (a.grad_X && a.d > 64 * 4) {  // accumulated in place: start from zero
    const int64_t nz_x = a.n_q * a.d;
    const int64_t b = ceil_div(nz_x, kThreads) < 8192 ? ceil_div(nz_x, kThreads) : 8192;
    hipLaunchKernelGGL(zero_kernel<T>, dim3((unsigned)b), dim3(kThreads), 0, s, a.grad_X, nz_x);
  }
  if (a.n_q <= 0) return;
  if (a.d <= 4) launch_g<T, 4>(a, s);
  else if (a.d <= 16) launch_g<T, 16>(a, s);
  else launch_g<T, 64>(a, s);
}

template void launch_backward<float>(const BackwardArgs<float>&, hipStream_t);
template void launch_backward<double>(const BackwardArgs<double>&, hipStream_t);

}  // namespace knn
