// Shared device helpers for the gfx950 argkmin kernels.
//
// Candidate keys are (distance, index) pairs ordered lexicographically: the
// lower index wins a tie.  That is the tie contract of this library (the
// reference leaves the order of equal distances unspecified:
// src/_pdist_aggregation.pyx:66 replaces the heap root on equality and
// :126-188 is an unstable quicksort).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cfloat>
#include <climits>
#include <cmath>

namespace knn {

constexpr int kThreads = 256;  // 4 wave64 per workgroup

// ---------------------------------------------------------------------------
// Order-preserving maps from floating point to unsigned integers, so that a
// per-query threshold can be shared between workgroups with atomicMin.
// ---------------------------------------------------------------------------
__device__ __host__ inline unsigned long long ordered_key(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __host__ inline double from_ordered_key(unsigned long long k) {
  const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __builtin_bit_cast(double, b);
}

template <class T>
__device__ inline T pos_inf();
template <>
__device__ inline float pos_inf<float>() {
  return __builtin_huge_valf();
}
template <>
__device__ inline double pos_inf<double>() {
  return __builtin_huge_val();
}

template <class T>
__device__ inline T max_finite();
template <>
__device__ inline float max_finite<float>() {
  return FLT_MAX;
}
template <>
__device__ inline double max_finite<double>() {
  return DBL_MAX;
}

template <class T>
__device__ inline bool key_less(T da, int ia, T db, int ib) {
  return da < db || (da == db && ia < ib);
}

// ---------------------------------------------------------------------------
// Workgroup-cooperative bitonic sort of n (a power of two) (dist, idx) pairs
// held in LDS, ascending by (dist, idx).  Every thread of the workgroup must
// call it; it starts and ends with a barrier.
// ---------------------------------------------------------------------------
template <class T>
__device__ void block_bitonic_sort(T* sd, int* si, int n) {
  __syncthreads();
  for (int size = 2; size <= n; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = threadIdx.x; i < (n >> 1); i += blockDim.x) {
        const int lo = ((i & ~(stride - 1)) << 1) | (i & (stride - 1));  // stride: power of 2
        const int hi = lo + stride;
        const bool up = (lo & size) == 0;
        const T dl = sd[lo], dh = sd[hi];
        const int il = si[lo], ih = si[hi];
        const bool swap = up ? key_less(dh, ih, dl, il) : key_less(dl, il, dh, ih);
        if (swap) {
          sd[lo] = dh;
          sd[hi] = dl;
          si[lo] = ih;
          si[hi] = il;
        }
      }
      __syncthreads();
    }
  }
}

// Relaxed loads at device (agent) scope: coherent in L2 without going to
// memory, for values other workgroups update (shared thresholds).
__device__ inline unsigned long long load_agent(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline float load_agent(const float* p) {
  return __uint_as_float(__hip_atomic_load(reinterpret_cast<const unsigned*>(p), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT));
}

// Orders LDS (and same-CU global) accesses between the lanes of one wave.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__host__ __device__ inline int next_pow2(int v);

// Wave-cooperative sort of one query's candidate list (cd, ci: n entries in
// global scratch) through wave-private LDS scratch (sd, si: >= next_pow2(n)
// entries), keeping the best min(n, kc) entries in cd/ci.  Called by all 64
// lanes of one wave with the same arguments; no workgroup barrier.
template <class T>
__device__ void wave_sort_cut(T* cd, int* ci, int n, int kc, T* sd, int* si, int lane) {
  const int P = next_pow2(n);
  for (int e = lane; e < P; e += 64) {
    const bool ok = e < n;
    sd[e] = ok ? cd[e] : pos_inf<T>();
    si[e] = ok ? ci[e] : INT_MAX;
  }
  wave_sync();
  for (int size = 2; size <= P; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = lane; i < (P >> 1); i += 64) {
        const int lo = ((i & ~(stride - 1)) << 1) | (i & (stride - 1));  // stride: power of 2
        const int hi = lo + stride;
        const bool up = (lo & size) == 0;
        const T dl = sd[lo], dh = sd[hi];
        const int il = si[lo], ih = si[hi];
        const bool swap = up ? key_less(dh, ih, dl, il) : key_less(dl, il, dh, ih);
        if (swap) {
          sd[lo] = dh;
          sd[hi] = dl;
          si[lo] = ih;
          si[hi] = il;
        }
      }
      wave_sync();
    }
  }
  const int m = n < kc ? n : kc;
  for (int e = lane; e < m; e += 64) {
    cd[e] = sd[e];
    ci[e] = si[e];
  }
  wave_sync();
}

__host__ __device__ inline int next_pow2(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

__host__ __device__ inline size_t align16(size_t v) { return (v + 15) & ~(size_t)15; }
__host__ __device__ inline int align8(int v) { return (v + 7) & ~7; }

}  // namespace knn
