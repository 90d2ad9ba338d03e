// Large k (k above the merge kernel's kMaxCandidates): exact selection and
// ordering from a chunk of dense FP64 distance rows.
//
// The reference keeps a k-entry max-heap per query whatever k is
// (src/_pdist_aggregation.pyx:54-105, :503-507) and sorts it at the end
// (:126-188).  Here one workgroup per query finds the k-th smallest key of
// its row by a radix select (eight passes of 8 bits over the ordered bits of
// the FP64 distances; distances are >= +0, so their bit patterns order like
// the values), keeps every entry below it and the lowest-index entries equal
// to it (the lower index wins a tie, as everywhere in this library), and a
// second kernel ranks the k kept entries by (distance, index): rank = the
// number of entries ordered before, a permutation because the (bits, index)
// keys are distinct.
#include "knn_common.hpp"
#include "knn_kernels.hpp"

namespace knn {

namespace {

constexpr int kRankLds = 4096;  // kept entries ranked from LDS up to this k

__device__ inline unsigned long long dkey(double v) {
  return __builtin_bit_cast(unsigned long long, v);
}

// Exclusive prefix over the workgroup of a 0/1 flag per thread, in thread
// order; returns this thread's offset and sets *total.
__device__ inline int block_prefix(bool f, int* wave_counts, int* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const unsigned long long m = __ballot(f);
  const int below = __builtin_popcountll(m & ((1ull << lane) - 1ull));
  if (lane == 0) wave_counts[w] = __builtin_popcountll(m);
  __syncthreads();
  int off = 0, tot = 0;
  for (int i = 0; i < kThreads / 64; ++i) {
    const int c = wave_counts[i];
    off += i < w ? c : 0;
    tot += c;
  }
  __syncthreads();  // wave_counts reused by the next call
  *total = tot;
  return off + below;
}

__global__ __launch_bounds__(kThreads) void bigk_select_kernel(const double* __restrict__ D,
                                                               int64_t n_t, int k,
                                                               double* __restrict__ sel_d,
                                                               int* __restrict__ sel_i) {
  __shared__ unsigned int hist[256];
  __shared__ int wave_counts[kThreads / 64];
  __shared__ unsigned long long prefix_s;
  __shared__ int need_s;
  const int64_t q = blockIdx.x;
  const double* row = D + q * n_t;
  const int tid = threadIdx.x;
  unsigned long long prefix = 0ull, mask = 0ull;
  int need = k;  // entries still to take among those matching the prefix
  for (int shift = 56; shift >= 0; shift -= 8) {
    hist[tid] = 0u;  // kThreads == 256 bins
    __syncthreads();
    for (int64_t j = tid; j < n_t; j += kThreads) {
      const unsigned long long key = dkey(row[j]);
      if ((key & mask) == prefix) atomicAdd(&hist[(key >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (tid == 0) {
      unsigned int below = 0;
      int b = 0;
      for (; b < 255; ++b) {
        if (below + hist[b] >= (unsigned)need) break;
        below += hist[b];
      }
      prefix_s = prefix | ((unsigned long long)b << shift);
      need_s = need - (int)below;
    }
    __syncthreads();
    prefix = prefix_s;
    need = need_s;
    mask |= 255ull << shift;
  }
  // prefix is the k-th smallest key; `need` entries equal to it are taken
  // (the lowest indices), every smaller one is taken.
  const unsigned long long kth = prefix;
  const int n_lt = k - need;
  double* od = sel_d + q * (int64_t)k;
  int* oi = sel_i + q * (int64_t)k;
  int lt_base = 0, eq_base = 0;
  for (int64_t j0 = 0; j0 < n_t; j0 += kThreads) {
    const int64_t j = j0 + tid;
    const double v = j < n_t ? row[j] : 0.0;
    const unsigned long long key = dkey(v);
    const bool lt = j < n_t && key < kth;
    const bool eq = j < n_t && key == kth;
    int lt_tot, eq_tot;
    const int lo = block_prefix(lt, wave_counts, &lt_tot);
    const int eo = block_prefix(eq, wave_counts, &eq_tot);
    if (lt) {
      od[lt_base + lo] = v;
      oi[lt_base + lo] = (int)j;
    }
    if (eq && eq_base + eo < need) {
      od[n_lt + eq_base + eo] = v;
      oi[n_lt + eq_base + eo] = (int)j;
    }
    lt_base += lt_tot;
    eq_base += eq_tot;
  }
}

template <class T>
__global__ __launch_bounds__(kThreads) void bigk_rank_kernel(const double* __restrict__ sel_d,
                                                             const int* __restrict__ sel_i, int k,
                                                             int squared, T* __restrict__ dist_out,
                                                             int64_t* __restrict__ idx_out) {
  __shared__ double ld[kRankLds];
  __shared__ int li[kRankLds];
  const int64_t q = blockIdx.x;
  const double* sd = sel_d + q * (int64_t)k;
  const int* si = sel_i + q * (int64_t)k;
  const bool lds = k <= kRankLds;
  if (lds) {
    for (int e = threadIdx.x; e < k; e += kThreads) {
      ld[e] = sd[e];
      li[e] = si[e];
    }
    __syncthreads();
  }
  const double* rd = lds ? ld : sd;
  const int* ri = lds ? li : si;
  // Keys compared as the select compared them (ordered bit patterns, then
  // index): a total order, so ranks stay a permutation whatever the values.
  for (int e = threadIdx.x; e < k; e += kThreads) {
    const double x = rd[e];
    const unsigned long long xk = dkey(x);
    const int xi = ri[e];
    int rank = 0;
    for (int j = 0; j < k; ++j) {
      const unsigned long long jk = dkey(rd[j]);
      rank += (jk < xk || (jk == xk && ri[j] < xi)) ? 1 : 0;
    }
    dist_out[q * k + rank] = (T)(squared ? x : sqrt(x));
    idx_out[q * k + rank] = (int64_t)xi;
  }
}

}  // namespace

template <class T>
void launch_bigk_select_sort(const double* D, int64_t n_q, int64_t n_t, int k, int squared,
                             double* sel_d, int* sel_i, T* dist_out, int64_t* idx_out,
                             hipStream_t s) {
  if (n_q <= 0) return;
  static_assert(kThreads == 256, "one histogram bin per thread");
  hipLaunchKernelGGL(bigk_select_kernel, dim3((unsigned)n_q), dim3(kThreads), 0, s, D, n_t, k,
                     sel_d, sel_i);
  hipLaunchKernelGGL(bigk_rank_kernel<T>, dim3((unsigned)n_q), dim3(kThreads), 0, s, sel_d, sel_i,
                     k, squared, dist_out, idx_out);
}

template void launch_bigk_select_sort<float>(const double*, int64_t, int64_t, int, int, double*,
                                             int*, float*, int64_t*, hipStream_t);
template void launch_bigk_select_sort<double>(const double*, int64_t, int64_t, int, int, double*,
                                              int*, double*, int64_t*, hipStream_t);

}  // namespace knn
