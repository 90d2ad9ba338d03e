// Row norms, split-list merge, exact re-rank and certification.
//
// Reference counterparts:
//  * row norms: np.einsum('ij,ij->i', X_train, X_train)
//    (src/_pdist_aggregation.pyx:508);
//  * merging per-split lists: the GIL-serialised heap merge of
//    _parallel_knn_on_X_train (:420-430) and the SYCL kern_aggregate
//    (sycl_ext/parallel_knn.hpp:474-509);
//  * exact distances of the chosen neighbours: _compute_exact_distances
//    (:450-487).  Unlike the reference, rows are re-sorted on the exact
//    distance with the lower index winning ties.
#include "knn_common.hpp"
#include "knn_kernels.hpp"

namespace knn {

namespace {

// The smallest float above a finite v.
__device__ inline float next_up(float v) {
  const unsigned b = __float_as_uint(v);
  if (v == 0.f) return __uint_as_float(1u);  // +-0 -> smallest positive subnormal
  return __uint_as_float(v > 0.f ? b + 1u : b - 1u);
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <class T>
__global__ void fill_keys_kernel(unsigned long long* p, int64_t n, unsigned long long v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void fill_batch_kernel(FillBatch b) {
  const int s = blockIdx.y;
  const int64_t n = b.n[s];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  if (b.wide[s]) {
    unsigned long long* p = static_cast<unsigned long long*>(b.p[s]);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = b.v[s];
  } else {
    unsigned* p = static_cast<unsigned*>(b.p[s]);
    const unsigned v = (unsigned)b.v[s];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
  }
}

__global__ void fill_u32_kernel(unsigned int* p, int64_t n, unsigned int v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// One wave per query: the j-th smallest of the query's sampled minima
// (n_runs ascending runs of kSampleJ values), used as a guessed threshold.
// Lane l keeps the kSampleJ smallest of runs l, l+64, ...; then j rounds of
// "take the smallest head among the lanes" pick the j-th smallest overall.
__global__ __launch_bounds__(kThreads) void sample_guess_kernel(
    const float* __restrict__ so, int64_t n_q, int n_runs, int j, unsigned long long* tau,
    float* guess, int fail_stride, int* failed, int* failed_qblocks, int qblock_queries) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  static_assert(kSampleJ % 4 == 0, "runs are read as float4");
  const int lane = threadIdx.x & 63;
  const int64_t q = (int64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  if (q >= n_q) return;  // whole wave
  if (lane == 0 && failed) {
    failed[q] = 0;
    if (q % qblock_queries == 0) failed_qblocks[q / qblock_queries] = 0;
  }
  float b[kSampleJ];
  const float* p = so + (size_t)q * n_runs * kSampleJ;
  // A lane's first run is its sorted list as it stands (runs are ascending);
  // later runs (more than 64 per query) are merged into it from registers.
#pragma unroll
  for (int i = 0; i < kSampleJ; i += 4) {
    f4 v = {__builtin_huge_valf(), __builtin_huge_valf(), __builtin_huge_valf(),
            __builtin_huge_valf()};
    if (lane < n_runs) v = *reinterpret_cast<const f4*>(p + (size_t)lane * kSampleJ + i);
    b[i] = v[0];
    b[i + 1] = v[1];
    b[i + 2] = v[2];
    b[i + 3] = v[3];
  }
  for (int r = lane + 64; r < n_runs; r += 64) {
    // The whole run in four 16-byte loads, then merged from registers.
    float run[kSampleJ];
#pragma unroll
    for (int i = 0; i < kSampleJ; i += 4) {
      const f4 v = *reinterpret_cast<const f4*>(p + (size_t)r * kSampleJ + i);
      run[i] = v[0];
      run[i + 1] = v[1];
      run[i + 2] = v[2];
      run[i + 3] = v[3];
    }
#pragma unroll
    for (int i = 0; i < kSampleJ; ++i) {
      const float m = run[i];
      if (!(m < b[kSampleJ - 1])) break;  // the rest of an ascending run is no better
#pragma unroll
      for (int t = kSampleJ - 1; t > 0; --t) b[t] = fminf(b[t], fmaxf(b[t - 1], m));
      b[0] = fminf(b[0], m);
    }
  }
  // j rounds: the smallest head across the wave (ties: lowest lane) is taken
  // and its lane shifts its list.
  float g = __builtin_huge_valf();
  for (int round = 0; round < j; ++round) {
    float m = b[0];
    int who = lane;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const float om = __shfl_xor(m, o, 64);
      const int ow = __shfl_xor(who, o, 64);
      if (om < m || (om == m && ow < who)) {
        m = om;
        who = ow;
      }
    }
    g = m;
    if (lane == who) {
#pragma unroll
      for (int t = 0; t < kSampleJ - 1; ++t) b[t] = b[t + 1];
      b[kSampleJ - 1] = __builtin_huge_valf();
    }
  }
  if (fail_stride > 0 && q % fail_stride == 0) g = -__builtin_huge_valf();
  if (lane == 0) {
    guess[q] = g;
    if (g < __builtin_huge_valf()) {
      const unsigned long long key = ordered_key((double)g);
      if (key < tau[q]) tau[q] = key;
    }
  }
}

// One wave per row: |y_j|^2 accumulated in T (the precision the search uses)
// and the maximum over all rows as an ordered key.
template <class T>
__global__ __launch_bounds__(kThreads) void row_norms_kernel(const T* __restrict__ Y, int64_t n,
                                                             int d, T* __restrict__ out,
                                                             DeviceScalars* sc) {
  __shared__ T wmax[kThreads / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  T best = T(0);
  for (int64_t r = (int64_t)blockIdx.x * 4 + w; r < n; r += (int64_t)gridDim.x * 4) {
    T s = T(0);
    for (int p = lane; p < d; p += 64) {
      const T v = Y[r * d + p];
      s = fma(v, v, s);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) out[r] = s;
    best = s > best ? s : best;
  }
  if (lane == 0) wmax[w] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    T m = wmax[0];
    for (int i = 1; i < kThreads / 64; ++i) m = wmax[i] > m ? wmax[i] : m;
    atomicMax(&sc->max_ynorm_key, ordered_key((double)m));
  }
}

template <class T>
struct Vec4;
template <>
struct Vec4<float> {
  using type = float4;
};
template <>
struct Vec4<double> {
  using type = double4;
};

// Candidates left after the threshold filter up to which the k' best are
// found by ranking (each thread counts the entries ahead of its own, no
// barriers); more than that are bitonic-sorted.
constexpr int kRankMax = 128;
// k' up to which the exact distances are ordered by ranking.
constexpr int kRankExact = 256;
// Merge workgroup: two waves per query (kMaxMerge entries, kPerThread per
// thread), or one wave when the query's entries fit kPerThread per lane of
// one: each phase holds at most a few entries per thread, so the smaller
// group does not lengthen a query and twice as many are in flight per CU
// (65,536 low-d queries x 5 splits x k'=24: merge 0.41 -> 0.21 ms; float32
// 1e4 x 19 x 18: 0.084 -> 0.056 ms; larger merges are held by LDS instead).
constexpr int kMergeThreads = 128;
// Split-list entries per thread held in registers (kMaxMerge / kMergeThreads).
constexpr int kPerThread = kMaxMerge / kMergeThreads;
static_assert(kPerThread * kMergeThreads == kMaxMerge, "merge staging");

// Stream compaction inside a wave over kPerThread rounds: at[i] = the slot
// of this lane's round-i entry among all entries with pred[i] of the wave,
// after ONE atomicAdd per wave on the LDS counter (the votes of every round
// are taken first).  Slots only place entries; every later step orders by
// key, so the order of waves and rounds does not matter.
template <int N>
__device__ inline void wave_slots(const bool (&pred)[N], int rounds, int* counter, int (&at)[N]) {
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (1ull << lane) - 1ull;
  unsigned long long m[N];
  int total = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    m[i] = i < rounds ? __ballot(pred[i]) : 0ull;
    total += __builtin_popcountll(m[i]);
  }
  int base = 0;
  if (lane == 0 && total) base = atomicAdd(counter, total);
  base = __shfl(base, 0, 64);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    at[i] = base + __builtin_popcountll(m[i] & below);
    base += __builtin_popcountll(m[i]);
  }
}

// Number of keys (kd[j], ki[j]), j < n, ordered before (x, xi): four keys per
// LDS read while four remain (kd, ki 16-byte aligned), then one at a time.
template <class T>
__device__ inline int count_less(const T* kd, const int* ki, int n, T x, int xi) {
  using V4 = typename Vec4<T>::type;
  int r = 0, j = 0;
#pragma unroll 4
  for (; j + 4 <= n; j += 4) {
    const V4 d = *reinterpret_cast<const V4*>(kd + j);
    const int4 i4 = *reinterpret_cast<const int4*>(ki + j);
    r += (key_less(d.x, i4.x, x, xi) ? 1 : 0) + (key_less(d.y, i4.y, x, xi) ? 1 : 0) +
         (key_less(d.z, i4.z, x, xi) ? 1 : 0) + (key_less(d.w, i4.w, x, xi) ? 1 : 0);
  }
  for (; j < n; ++j) r += key_less(kd[j], ki[j], x, xi) ? 1 : 0;
  return r;
}

// The kth-smallest (1-based) of the float values v[i] held by the workgroup's
// threads (entries with id[i] >= 0, rounds i < rounds), by a radix select
// over their order-preserving 32-bit keys: four passes of 8 bits, each an LDS
// histogram of the entries that match the digits chosen so far and one
// wave's prefix scan that picks the next digit.  The caller guarantees at
// least kth such entries.  Every thread calls it (it holds barriers).
template <int NT, int N>
__device__ float block_select_kth(const float (&v)[N], const int (&id)[N], int rounds, int kth,
                                  unsigned* hist) {
  __shared__ unsigned pick_s[2];  // digits chosen so far, rank left within them
  const int tid = threadIdx.x, lane = tid & 63;
  unsigned key[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const unsigned b = __float_as_uint(v[i]);
    key[i] = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
  }
  unsigned prefix = 0, mask = 0, rank = (unsigned)kth;
#pragma unroll
  for (int pass = 0; pass < 4; ++pass) {
    const int shift = 24 - 8 * pass;
    for (int b = tid; b < 256; b += NT) hist[b] = 0;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < N; ++i)
      if (i < rounds && id[i] >= 0 && (key[i] & mask) == prefix)
        atomicAdd(&hist[(key[i] >> shift) & 255u], 1u);
    __syncthreads();
    if (tid < 64) {
      // lane l holds bins 4l .. 4l+3; inclusive scan of the lane sums
      unsigned c[4], s = 0;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        c[j] = hist[4 * lane + j];
        s += c[j];
      }
      unsigned incl = s;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const unsigned t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
      }
      const unsigned before = incl - s;
      // the lane whose bins reach `rank` picks the digit
      if (before < rank && incl >= rank) {
        unsigned acc = before;
        int j = 0;
        while (acc + c[j] < rank) acc += c[j++];
        pick_s[0] = prefix | ((unsigned)(4 * lane + j) << shift);
        pick_s[1] = rank - acc;
      }
    }
    __syncthreads();
    prefix = pick_s[0];
    rank = pick_s[1];
    mask |= 255u << shift;
    __syncthreads();  // pick_s is rewritten by the next pass
  }
  const unsigned b = (prefix & 0x80000000u) ? (prefix & 0x7fffffffu) : ~prefix;
  return __uint_as_float(b);
}

// One workgroup of NT threads per query (final merge) or per (query, group of
// splits) (pre-merge, a.pre_d set); the launch guarantees lists * k' <= NT *
// PT.
// PT list entries per thread (the launch picks 12 when a query's lists fit
// 768 entries, else kPerThread).  The float one-wave forms are held to 72 /
// 80 VGPRs (7 / 6 waves per SIMD, 28 / 24 queries in flight per CU with the
// capped staging): the merge waits on memory round trips, so more queries in
// flight is what shortens it (BF16 1e4 x 1e6 x 128, 19 lists of k' = 40:
// 0.118 -> 0.112 ms with 12 entries per thread).
template <class T, int NT, int PT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(
    sizeof(T) == 4 && NT == 64 ? (PT <= 12 ? 7 : 6) : 1)))
void merge_rerank_kernel(MergeArgs<T> a) {
  extern __shared__ __align__(16) unsigned char dyn_smem[];
  const bool pre = a.pre_d != nullptr;
  const int64_t q = pre ? (int64_t)blockIdx.x / a.n_groups : (int64_t)blockIdx.x;
  const int grp = pre ? (int)(blockIdx.x % a.n_groups) : 0;
  // Lists merged here, and the list count the LDS layout was sized for.
  const int lists_alloc = pre ? a.group_size : a.n_splits;
  const int first = grp * lists_alloc;
  const int lists = pre ? min(a.group_size, a.n_splits - first) : a.n_splits;
  const int total = lists * a.kc;
  const int P_all = next_pow2(lists_alloc * a.kc);
  // (float lists only: the MFMA searches write them)
  const bool capped =
      sizeof(T) == 4 && a.unsorted && a.stage_cap > 0 && a.stage_cap < P_all;  // uniform
  const int P = capped ? a.stage_cap : P_all;
  const int KP = next_pow2(a.kc);
  const size_t o_si = align16((size_t)P * sizeof(T));
  const size_t o_ed = align16(o_si + (size_t)P * sizeof(int));
  const size_t o_ei = o_ed + (size_t)KP * sizeof(double);
  T* sd = reinterpret_cast<T*>(dyn_smem);
  int* si = reinterpret_cast<int*>(dyn_smem + o_si);
  double* ed = reinterpret_cast<double*>(dyn_smem + o_ed);
  int* ei = reinterpret_cast<int*>(dyn_smem + o_ei);
  __shared__ double xnorm_s, akc_s, kth_s;
  __shared__ int n_s, full_s;
  __shared__ T cut_s;

  if (a.only_queries && a.only_queries[q] == 0) return;  // re-search merge: not selected
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const T* pd = a.part_d + ((size_t)q * a.n_splits + first) * a.kc;
  const int* pi = a.part_i + ((size_t)q * a.n_splits + first) * a.kc;
  // The query's final shared threshold bounds its k'-th value whenever the
  // search kept k' entries at or below it (a split reached k' entries there,
  // or every entry passed the sampled guess): entries above it are dropped.
  // Without k' entries at or below it the query fails the check below.
  const T tq = a.tau ? (T)from_ordered_key(a.tau[q]) : pos_inf<T>();
  // When each split's list is sorted, the k'-th smallest of the first r
  // entries of every split (r * n_splits >= k') bounds the k'-th value from
  // above: entries past it are dropped too, which leaves a few more than k'.
  const int r0 = (a.kc + lists - 1) / lists + 1;
  const int r = a.kc < r0 ? a.kc : r0;
  const bool sorted = !a.unsorted;
  T v[PT];
  int id[PT];
  // List of entry e: (e + 0.5) / k' is at least 0.5 / k' from an integer, far
  // more than the float product's error for e < kMaxMerge, so truncation is
  // exact (an integer division per entry cost about 10 % of the merge).
  const float inv_kc = 1.f / (float)a.kc;
  if (tid == 0) n_s = 0;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < PT; ++i) {
    const int e = tid + i * NT;
    bool in = e < total;
    if (in && a.part_n) {  // unwritten slots past the list's entries
      const int l = (int)(((float)e + 0.5f) * inv_kc);
      in = e - l * a.kc < a.part_n[(size_t)q * a.n_splits + first + l];
    }
    id[i] = in ? pi[e] : -1;
    v[i] = in ? pd[e] : pos_inf<T>();
    if (!(id[i] >= 0 && id[i] < a.n_t && v[i] <= tq)) id[i] = -1;  // never index past Y
  }
  const int rounds = (total + NT - 1) / NT;  // uniform
  if (sorted) {
    bool head[PT];
    int at[PT];
#pragma unroll
    for (int i = 0; i < PT; ++i) {
      const int e = tid + i * NT;
      head[i] = id[i] >= 0 && e % a.kc < r;
    }
    wave_slots(head, rounds, &n_s, at);
#pragma unroll
    for (int i = 0; i < PT; ++i)
      if (head[i]) {
        sd[at[i]] = v[i];
        si[at[i]] = id[i];
      }
  }
  __syncthreads();
  const int nh = sorted ? n_s : 0;  // unsorted lists: no head cut
  if (tid == 0) cut_s = tq;
  __syncthreads();
  if constexpr (sizeof(T) == 4) {
    // Unsorted lists: the k'-th smallest value of all entries at or below
    // the threshold, selected exactly (every barrier is uniform: a.unsorted
    // is a launch argument and n_s a shared value).
    if (!sorted) {
      if (tid == 0) n_s = 0;
      __syncthreads();
      int cnt = 0;
#pragma unroll
      for (int i = 0; i < PT; ++i) cnt += (i < rounds && id[i] >= 0) ? 1 : 0;
      if (cnt) atomicAdd(&n_s, cnt);
      __syncthreads();
      const int nv = n_s;
      __syncthreads();
      if (nv >= a.kc) {
        __shared__ unsigned hist[256];
        const float c = block_select_kth<NT>(v, id, rounds, a.kc, hist);
        if (tid == 0) cut_s = c;
        __syncthreads();
      }
    }
  }
  if (sorted && nh >= a.kc) {
    for (int e = tid; e < nh; e += NT) {
      const T x = sd[e];
      const int xi = si[e];
      const int rank = count_less(sd, si, nh, x, xi);
      if (rank == a.kc - 1) cut_s = x;  // one thread: keys are distinct
    }
  }
  __syncthreads();
  const T cut = cut_s;
  if (tid == 0) n_s = 0;
  __syncthreads();
  {
    // Every entry at or below the cut, in as many slots as there are (capped
    // staging: the first P of them).
    bool keep[PT];
    int at[PT];
#pragma unroll
    for (int i = 0; i < PT; ++i) keep[i] = id[i] >= 0 && v[i] <= cut;
    wave_slots(keep, rounds, &n_s, at);
#pragma unroll
    for (int i = 0; i < PT; ++i)
      if (keep[i] && at[i] < P) {
        sd[at[i]] = v[i];
        si[at[i]] = id[i];
      }
    if constexpr (sizeof(T) == 4) {
      __syncthreads();
      if (capped && n_s > P) {  // uniform
        // Overflow (only ties at the cut can cause it: fewer than k' entries
        // lie below it).  Staged again: the entries below the cut, then the
        // ties of lowest row index that fit, so the staged entries hold the
        // k' smallest (value, index) keys.
        __syncthreads();
        if (tid == 0) n_s = 0;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < PT; ++i) keep[i] = id[i] >= 0 && v[i] < cut;
        wave_slots(keep, rounds, &n_s, at);
#pragma unroll
        for (int i = 0; i < PT; ++i)
          if (keep[i]) {
            sd[at[i]] = v[i];
            si[at[i]] = id[i];
          }
        __syncthreads();
        const int below = n_s;
        // v and id are not read after the staging: the ties' row indices
        // replace them (indices >= 0, so their bit order is index order).
#pragma unroll
        for (int i = 0; i < PT; ++i) {
          const bool tie = id[i] >= 0 && v[i] == cut;
          id[i] = tie ? id[i] : -1;
          v[i] = __int_as_float(tie ? id[i] : 0);
        }
        __shared__ unsigned hist_t[256];
        const int imax = __float_as_int(block_select_kth<NT>(v, id, rounds, P - below, hist_t));
#pragma unroll
        for (int i = 0; i < PT; ++i) keep[i] = id[i] >= 0 && id[i] <= imax;
        wave_slots(keep, rounds, &n_s, at);
#pragma unroll
        for (int i = 0; i < PT; ++i)
          if (keep[i]) {
            sd[at[i]] = cut;
            si[at[i]] = id[i];
          }
      }
    }
  }
  __syncthreads();
  const int n = n_s < P ? n_s : P;
  // The k' best, ascending: ranked into the (still unused) exact-distance
  // arrays, or the head of the sorted staging arrays.
  T* od = n <= kRankMax ? reinterpret_cast<T*>(ed) : sd;
  int* oi = n <= kRankMax ? ei : si;
  if (n <= kRankMax) {
    // (value, index) keys are distinct (a train row is in one split only), so
    // ranks are a permutation.
    for (int e = tid; e < n; e += NT) {
      const T x = sd[e];
      const int xi = si[e];
      const int rank = count_less(sd, si, n, x, xi);
      if (rank < a.kc) {
        od[rank] = x;
        oi[rank] = xi;
      }
    }
  } else {
    const int Pn = next_pow2(n);
    for (int e = n + tid; e < Pn; e += NT) {
      sd[e] = pos_inf<T>();
      si[e] = INT_MAX;
    }
    block_bitonic_sort(sd, si, Pn);
  }
  for (int j = n + tid; j < a.kc; j += NT) {
    od[j] = pos_inf<T>();
    oi[j] = INT_MAX;
  }
  __syncthreads();
  if (pre) {  // the group's k' best, sorted, for the final merge
    const size_t o = ((size_t)q * a.n_groups + grp) * a.kc;
    for (int j = tid; j < a.kc; j += NT) {
      a.pre_d[o + j] = od[j];
      a.pre_i[o + j] = oi[j] == INT_MAX ? -1 : oi[j];
    }
    return;
  }

  // Sampled-threshold check: with at least min(k', n_t) candidates every row
  // the guess dropped has a value above the k'-th kept one, so nothing is
  // lost; with fewer (fewer than min(k, n_t) when the guess was planned for
  // k: see guess_bound), the query is searched again from an infinite
  // threshold.
  if (a.failed) {
    const int want = a.guess_bound ? a.k : a.kc;
    const int need = (int64_t)want < a.n_t ? want : (int)a.n_t;
    if (n < need) {  // n_s: the same value for every thread
      if (tid == 0) {
        a.failed[q] = 1;
        a.failed_qblocks[q / a.qblock_queries] = 1;  // every writer stores 1
        a.tau[q] = ordered_key(__builtin_huge_val());
        atomicAdd(&a.sc->n_rescued, 1u);
      }
      return;
    }
  }
  if (tid == 0) {
    full_s = oi[a.kc - 1] != INT_MAX;
    akc_s = full_s ? (double)od[a.kc - 1] : 0.0;
    // Fewer than k' candidates under a guess planned for k (n >= k here):
    // every row left out has a value above the threshold.
    if (!full_s && a.failed && a.guess_bound && tq < pos_inf<T>()) {
      full_s = 1;
      akc_s = (double)tq;
    }
  }
  __syncthreads();  // od is overwritten below (it may share storage with ed)

  // Exact FP64 direct-difference distances of the best k' candidates: tpc
  // consecutive lanes per candidate (8 or 16: a load instruction reads whole
  // 128-byte lines of each row; 16 when that already covers k' in one pass),
  // kU candidates per lane group and pass with all their loads issued
  // together, then a shuffle reduction over the tpc lanes.  A slot without a
  // valid row reads row 0 (no divergent load) and is dropped.
  constexpr int kU = 4;
  const T* xr = a.X + q * a.d;
  using V4 = typename Vec4<T>::type;
  const bool vec4 = (a.d & 3) == 0 &&
                    ((reinterpret_cast<uintptr_t>(a.X) | reinterpret_cast<uintptr_t>(a.Y)) &
                     (4 * sizeof(T) - 1)) == 0;
  const int span = vec4 ? (a.d >> 2) : a.d;  // loads per row
  const int cover = NT * kU / KP;  // lanes per candidate for one pass
  int tpc = cover < 8 ? 8 : (cover > 16 ? 16 : cover);
  while (tpc > 1 && tpc / 2 >= span) tpc >>= 1;  // no more lanes than loads
  const int groups = NT / tpc;
  const int grp_j = tid / tpc, c = tid % tpc;
  for (int j0 = 0; j0 < KP; j0 += groups * kU) {
    int cid[kU];
    bool valid[kU];
    const T* yr[kU];
    double s[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int j = j0 + u * groups + grp_j;
      cid[u] = j < a.kc ? oi[j] : INT_MAX;
      // Every slot below min(n, k') holds a staged row when keys are
      // distinct; the range test keeps the gather inside Y whatever it holds.
      valid[u] = cid[u] >= 0 && cid[u] < a.n_t;
      yr[u] = a.Y + (int64_t)(valid[u] ? cid[u] : 0) * a.d;
      s[u] = 0.0;
    }
    if (vec4) {
#pragma unroll 2
      for (int p = c; p < span; p += tpc) {
        const V4 xv = reinterpret_cast<const V4*>(xr)[p];
        V4 yv[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) yv[u] = reinterpret_cast<const V4*>(yr[u])[p];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
          const double d0 = (double)xv.x - (double)yv[u].x, d1 = (double)xv.y - (double)yv[u].y;
          const double d2 = (double)xv.z - (double)yv[u].z, d3 = (double)xv.w - (double)yv[u].w;
          s[u] = fma(d0, d0, s[u]);
          s[u] = fma(d1, d1, s[u]);
          s[u] = fma(d2, d2, s[u]);
          s[u] = fma(d3, d3, s[u]);
        }
      }
    } else {
      for (int p = c; p < span; p += tpc) {
        const double x = (double)xr[p];
        T yv[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) yv[u] = yr[u][p];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
          const double df = x - (double)yv[u];
          s[u] = fma(df, df, s[u]);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kU; ++u)
      for (int o = tpc >> 1; o > 0; o >>= 1) s[u] += __shfl_xor(s[u], o, 64);
    __syncthreads();  // every oi[] read of this pass precedes the ed/ei writes (shared storage)
    if (c == 0) {
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int j = j0 + u * groups + grp_j;
        if (j < KP) {
          ed[j] = valid[u] ? s[u] : __builtin_huge_val();
          ei[j] = valid[u] ? cid[u] : INT_MAX;
        }
      }
    }
  }
  if (w == 0) {
    double s = 0.0;
    for (int p = lane; p < a.d; p += 64) {
      const double x = (double)xr[p];
      s = fma(x, x, s);
    }
    s = wave_sum(s);
    if (lane == 0) xnorm_s = s;
  }
  __syncthreads();

  // Order by (exact distance, index) and write the first k.  Up to
  // kRankExact candidates: each one's rank is the number of candidates ahead
  // of it (distinct keys); more: bitonic sort.
  if (a.kc <= kRankExact) {
    if (tid == 0) kth_s = __builtin_huge_val();
    for (int j = tid; j < a.k; j += NT) {  // slots no candidate fills
      a.idx_out[q * a.k + j] = -1;
      a.dist_out[q * a.k + j] = pos_inf<T>();
    }
    __syncthreads();
    for (int j = tid; j < a.kc; j += NT) {
      if (ei[j] == INT_MAX) continue;
      const double x = ed[j];
      const int xi = ei[j];
      const int rank = count_less(ed, ei, a.kc, x, xi);
      if (rank < a.k) {
        a.idx_out[q * a.k + rank] = (int64_t)xi;
        a.dist_out[q * a.k + rank] = (T)(a.squared ? x : sqrt(x));
        if (rank == a.k - 1) kth_s = x;
      }
    }
  } else {
    block_bitonic_sort(ed, ei, KP);
    for (int j = tid; j < a.k; j += NT) {
      const bool ok = j < a.kc && ei[j] != INT_MAX;
      const double x = ok ? ed[j] : __builtin_huge_val();
      a.idx_out[q * a.k + j] = ok ? (int64_t)ei[j] : -1;
      a.dist_out[q * a.k + j] = (T)(a.squared ? x : sqrt(x));
    }
    if (tid == 0) kth_s = a.k <= a.kc ? ed[a.k - 1] : __builtin_huge_val();
  }
  __syncthreads();

  // Certification: every train row left out of the k' candidates has an
  // approximate value >= a_kc (the k'-th).  Its true distance is then at
  // least sqrt(max(0, a_kc + |xq|^2 - delta)) - xerr - max_yerr; if that
  // exceeds the k-th exact distance the k returned rows are the exact answer.
  if (tid == 0) {
    const bool trivial = a.kc >= a.n_t;
    bool certified = trivial;
    const double nx = a.xnorm ? (double)a.xnorm[q] : xnorm_s;
    const double ny = from_ordered_key(a.sc->max_ynorm_key);
    const double nn = (double)a.acc_terms;
    const double gamma = nn * a.unit_roundoff / (1.0 - nn * a.unit_roundoff);
    const double ex = a.xerr ? (double)a.xerr[q] : 0.0;
    const double ey = a.xerr ? from_ordered_key(a.sc->max_yerr_key) : 0.0;
    if (!certified && a.kc >= a.k && full_s) {
      const double akc = akc_s;
      const double delta = gamma * (ny + 2.0 * sqrt(nx * ny) + fabs(akc) + nx);
      const double lo = akc + nx - delta;
      const double lower = sqrt(lo > 0.0 ? lo : 0.0) - ex - ey;
      certified = lower > sqrt(kth_s);
    }
    if (a.widen_all > 0 && q % a.widen_all == 0 && !trivial) certified = false;
    a.uncert[q] = certified ? 0 : 1;
    // Widening: a train row y within the k-th exact distance sqrt(U) of x
    // (every true neighbour is: U bounds the true k-th distance from above)
    // has |q(x) - q(y)| <= sqrt(U) + xerr + max_yerr = R, i.e. an exact
    // search value |q(y)|^2 - 2 q(x).q(y) <= R^2 - |q(x)|^2, and the computed
    // one exceeds that by at most gamma (2 |q(y)|^2 + 2 |q(x)||q(y)| +
    // |q(x)|^2) (the fp32 norms and the dot product).  T, rounded up to
    // float (thr below), is the collect launch's fixed threshold.
    if (!certified && a.widen && kth_s < __builtin_huge_val()) {
      const double R = sqrt(kth_s * (1.0 + 1e-12)) + ex + ey;
      const double thr = R * R - nx + gamma * (2.0 * ny + 2.0 * sqrt(nx * ny) + nx) +
                         1e-300;
      float tf = (float)thr;
      if ((double)tf < thr) tf = next_up(tf);
      // Every row with a value below the search's final threshold sits in
      // one of the split lists (see lists_complete); rows at or below tf
      // below it are among them.  (tau[q] is still that threshold: only
      // this workgroup writes it, below.)
      const bool from_lists =
          a.lists_complete && (double)tf < from_ordered_key(a.tau[q]);
      a.widen[q] = from_lists ? 2 : 1;
      // Plain stores: every writer of a block flag stores 1 (an atomic per
      // widened query on 40 flags and a shared list counter took 19 us per
      // query of the FP8 merge at 1e4 x 1e6 x 128, where every query widens).
      // A query taken from its lists stores into the sink entry past the
      // last block (a conditional store here cost the kernel a private
      // segment).
      const int64_t sink = (a.n_q + a.qblock_queries - 1) / a.qblock_queries;
      a.widen_qblocks[from_lists ? sink : q / a.qblock_queries] = 1;
      a.ubound[q] = kth_s;
      a.tau[q] = ordered_key((double)tf);
    }
  }
}

// One workgroup per query selected for widening: exact FP64 distances of
// every collected row (tpc lanes per row, kU rows per lane group in flight),
// those within the first answer's k-th distance U (+ 1e-12 relative: U may
// come from another summation order) compacted into LDS, bitonic-sorted by
// (distance, index), the first k written.  The collected set holds every
// row within sqrt(U), hence every true neighbour and every tie at the true
// k-th distance, so the result is exact; the query's flag is cleared.
template <class T, int NT>
__global__ __launch_bounds__(NT) void collect_rerank_kernel(CollectArgs<T> a) {
  extern __shared__ __align__(16) unsigned char dyn_smem[];
  int64_t q = blockIdx.x;
  if (a.order) {
    // Workgroups go to the 8 XCDs round robin: position p = the b-th slot of
    // XCD b % 8's contiguous range, so neighbours in `order` share an L2.
    const int64_t per_xcd = ((int64_t)gridDim.x + 7) / 8;
    const int64_t p = (int64_t)(blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (p >= a.n_q) return;
    q = a.order[p];
  }
  if (a.widen[q] == 0) return;
  const int cnt = a.coll_cnt[q];
  if (cnt > a.coll_cap) return;  // set overflowed: the first answer and its flag stay
  double* sd = reinterpret_cast<double*>(dyn_smem);
  int* si = reinterpret_cast<int*>(sd + a.surv_cap);
  __shared__ int n_s, over_s;
  const int tid = threadIdx.x;
  if (tid == 0) {
    n_s = 0;
    over_s = 0;
  }
  __syncthreads();
  const double cut = a.ubound[q] * (1.0 + 1e-12);
  const int* ci = a.coll_idx + (size_t)q * a.coll_cap;
  constexpr int kU = 4;
  const T* xr = a.X + q * a.d;
  using V4 = typename Vec4<T>::type;
  const bool vec4 = (a.d & 3) == 0 &&
                    ((reinterpret_cast<uintptr_t>(a.X) | reinterpret_cast<uintptr_t>(a.Y)) &
                     (4 * sizeof(T) - 1)) == 0;
  const int span = vec4 ? (a.d >> 2) : a.d;  // loads per row
  int tpc = 16;
  while (tpc > 1 && tpc / 2 >= span) tpc >>= 1;  // no more lanes than loads
  const int groups = NT / tpc;
  const int grp_j = tid / tpc, c = tid % tpc;
  for (int j0 = 0; j0 < cnt; j0 += groups * kU) {
    int cid[kU];
    bool valid[kU];
    const T* yr[kU];
    double sum[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int j = j0 + u * groups + grp_j;
      cid[u] = j < cnt ? ci[j] : -1;
      valid[u] = cid[u] >= 0 && cid[u] < a.n_t;
      yr[u] = a.Y + (int64_t)(valid[u] ? cid[u] : 0) * a.d;
      sum[u] = 0.0;
    }
    if (vec4) {
      for (int p = c; p < span; p += tpc) {
        const V4 xv = reinterpret_cast<const V4*>(xr)[p];
        V4 yv[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) yv[u] = reinterpret_cast<const V4*>(yr[u])[p];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
          const double d0 = (double)xv.x - (double)yv[u].x, d1 = (double)xv.y - (double)yv[u].y;
          const double d2 = (double)xv.z - (double)yv[u].z, d3 = (double)xv.w - (double)yv[u].w;
          sum[u] = fma(d0, d0, sum[u]);
          sum[u] = fma(d1, d1, sum[u]);
          sum[u] = fma(d2, d2, sum[u]);
          sum[u] = fma(d3, d3, sum[u]);
        }
      }
    } else {
      for (int p = c; p < span; p += tpc) {
        const double x = (double)xr[p];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
          const double df = x - (double)yr[u][p];
          sum[u] = fma(df, df, sum[u]);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kU; ++u)
      for (int o = tpc >> 1; o > 0; o >>= 1) sum[u] += __shfl_xor(sum[u], o, 64);
    if (c == 0) {
#pragma unroll
      for (int u = 0; u < kU; ++u)
        if (valid[u] && sum[u] <= cut) {
          const int slot = atomicAdd(&n_s, 1);
          if (slot < a.surv_cap) {
            sd[slot] = sum[u];
            si[slot] = cid[u];
          } else {
            over_s = 1;
          }
        }
    }
  }
  __syncthreads();
  const int n = n_s;
  if (over_s || n < a.k) return;  // cannot happen for a complete set; keep the flag
  const int P = next_pow2(n);
  for (int e = n + tid; e < P; e += NT) {
    sd[e] = __builtin_huge_val();
    si[e] = INT_MAX;
  }
  block_bitonic_sort(sd, si, P);
  for (int j = tid; j < a.k; j += NT) {
    const double x = sd[j];
    a.idx_out[q * a.k + j] = (int64_t)si[j];
    a.dist_out[q * a.k + j] = (T)(a.squared ? x : sqrt(x));
  }
  if (tid == 0) a.uncert[q] = 0;
}

__global__ __launch_bounds__(kThreads) void count_i32_kernel(const int* __restrict__ f, int64_t n,
                                                             int value, unsigned int* out) {
  __shared__ unsigned int wsum[kThreads / 64];
  unsigned int c = 0;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads)
    c += value ? f[i] == value : f[i] != 0;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned int t = 0;
    for (int i = 0; i < kThreads / 64; ++i) t += wsum[i];
    if (t) atomicAdd(out, t);
  }
}

// Sum of the per-query uncertified flags: 16 bytes per thread and pass, one
// atomic per workgroup.
__global__ __launch_bounds__(kThreads) void count_flags_kernel(const unsigned char* __restrict__ f,
                                                               int64_t n, unsigned int* out) {
  __shared__ unsigned int wsum[kThreads / 64];
  unsigned int c = 0;
  const int64_t n16 = n >> 4;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n16;
       i += (int64_t)gridDim.x * kThreads) {
    const uint4 v = reinterpret_cast<const uint4*>(f)[i];
    // flags are 0 or 1: the byte sums of a word fit in its bytes
    const unsigned int s = v.x + v.y + v.z + v.w;
    c += (s & 0xff) + ((s >> 8) & 0xff) + ((s >> 16) & 0xff) + (s >> 24);
  }
  if (blockIdx.x == 0)
    for (int64_t i = (n16 << 4) + threadIdx.x; i < n; i += kThreads) c += f[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned int t = 0;
    for (int i = 0; i < kThreads / 64; ++i) t += wsum[i];
    if (t) atomicAdd(out, t);
  }
}

}  // namespace

template <class T>
void launch_fill_keys(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t s) {
  if (n <= 0) return;
  const int64_t blocks = ceil_div(n, kThreads) < 4096 ? ceil_div(n, kThreads) : 4096;
  hipLaunchKernelGGL(fill_keys_kernel<T>, dim3((unsigned)blocks), dim3(kThreads), 0, s, p, n, v);
}

template <class T>
void launch_row_norms(const T* Y, int64_t n, int d, T* out, DeviceScalars* sc, hipStream_t s) {
  if (n <= 0) return;
  // 2048 workgroups looping over rows: each ends with an atomicMax on one
  // address (1e6 x 128 float32: 8192 workgroups took 25 us longer).
  const int64_t blocks = ceil_div(n, 4) < 2048 ? ceil_div(n, 4) : 2048;
  hipLaunchKernelGGL(row_norms_kernel<T>, dim3((unsigned)blocks), dim3(kThreads), 0, s, Y, n, d,
                     out, sc);
}

void launch_fill_u32(unsigned int* p, int64_t n, unsigned int v, hipStream_t s) {
  if (n <= 0) return;
  const int64_t blocks = ceil_div(n, kThreads) < 4096 ? ceil_div(n, kThreads) : 4096;
  hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, s, p, n, v);
}

void launch_fill_batch(FillBatch& b, hipStream_t s) {
  if (b.count == 0) return;
  int64_t most = 0;
  for (int i = 0; i < b.count; ++i) most = b.n[i] > most ? b.n[i] : most;
  const int64_t blocks = ceil_div(most, kThreads) < 1024 ? ceil_div(most, kThreads) : 1024;
  hipLaunchKernelGGL(fill_batch_kernel, dim3((unsigned)blocks, (unsigned)b.count), dim3(kThreads), 0,
                     s, b);
  b.count = 0;
}

void launch_sample_guess(const float* sample_out, int64_t n_q, int n_runs, int j,
                         unsigned long long* tau, float* guess, int fail_stride, int* failed,
                         int* failed_qblocks, int qblock_queries, hipStream_t s) {
  if (n_q <= 0) return;
  hipLaunchKernelGGL(sample_guess_kernel, dim3((unsigned)ceil_div(n_q, kThreads / 64)),
                     dim3(kThreads), 0, s, sample_out, n_q, n_runs, j, tau, guess, fail_stride,
                     failed, failed_qblocks, qblock_queries);
}

size_t merge_lds_bytes(int n_splits, int kc, size_t elem, int stage_cap) {
  const int p_all = next_pow2(n_splits * kc);
  const bool capped = elem == 4 && stage_cap > 0 && stage_cap < p_all;
  const size_t P = (size_t)(capped ? stage_cap : p_all);
  const size_t KP = (size_t)next_pow2(kc);
  return align16(align16(P * elem) + P * sizeof(int)) + KP * (sizeof(double) + sizeof(int));
}

template <class T>
void launch_merge_rerank(const MergeArgs<T>& a, hipStream_t s) {
  if (a.n_q <= 0) return;
  const size_t lds = merge_lds_bytes(a.pre_d ? a.group_size : a.n_splits, a.kc, sizeof(T),
                                     a.unsorted ? a.stage_cap : 0);
  const int64_t blocks = a.pre_d ? a.n_q * a.n_groups : a.n_q;
  const int lists = a.pre_d ? a.group_size : a.n_splits;
  if (lists * a.kc <= 64 * 12)
    hipLaunchKernelGGL((merge_rerank_kernel<T, 64, 12>), dim3((unsigned)blocks), dim3(64), lds, s, a);
  else if (lists * a.kc <= 64 * kPerThread)
    hipLaunchKernelGGL((merge_rerank_kernel<T, 64, kPerThread>), dim3((unsigned)blocks), dim3(64),
                       lds, s, a);
  else
    hipLaunchKernelGGL((merge_rerank_kernel<T, kMergeThreads, kPerThread>), dim3((unsigned)blocks),
                       dim3(kMergeThreads), lds, s, a);
}

template <class T>
void launch_collect_rerank(const CollectArgs<T>& a, hipStream_t s) {
  if (a.n_q <= 0) return;
  const size_t lds = (size_t)a.surv_cap * (sizeof(double) + sizeof(int));
  // With an order, the grid is padded to a multiple of 8 (whole XCD ranges).
  const int64_t blocks = a.order ? (a.n_q + 7) / 8 * 8 : a.n_q;
  hipLaunchKernelGGL((collect_rerank_kernel<T, kThreads>), dim3((unsigned)blocks), dim3(kThreads),
                     lds, s, a);
}

namespace {

// key[q] = the smallest row index widened query q collected (n_t for a query
// that is not widened or whose set overflowed: it sorts last); one wave per
// query.
__global__ __launch_bounds__(kThreads) void collect_key_kernel(const int* widen, const int* cnt,
                                                               const int* idx, int cap,
                                                               int64_t n_q, int64_t n_t,
                                                               int* key) {
  const int lane = threadIdx.x & 63;
  const int64_t q = (int64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  if (q >= n_q) return;  // whole wave
  const int c = widen[q] ? cnt[q] : cap + 1;
  int m = (int)n_t;
  if (c <= cap)
    for (int j = lane; j < c; j += 64) m = min(m, idx[(size_t)q * cap + j]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = min(m, __shfl_xor(m, o, 64));
  if (lane == 0) key[q] = m < 0 ? 0 : m;
}

// Buckets spread over the keys in use, [0, kmax]: the smallest row index of
// a set is small whatever the set (1,000 Gaussians over 2e6 rows: all below
// about 10,000), so buckets over all n_t rows would hold a dozen clusters
// each.  Keys n_t (not widened) take the last bucket.
__device__ inline int order_bucket(int key, int64_t n_t, int kmax) {
  if (key >= n_t) return kOrderBuckets - 1;
  return (int)((int64_t)key * (kOrderBuckets - 1) / ((int64_t)kmax + 1));
}

// *kmax = the largest key below n_t (0 if none): one workgroup.
__global__ __launch_bounds__(1024) void order_max_kernel(const int* key, int64_t n_q, int64_t n_t,
                                                         int* kmax) {
  __shared__ int part[1024 / 64];
  int m = 0;
  for (int64_t q = threadIdx.x; q < n_q; q += 1024) {
    const int k = key[q];
    if (k < n_t && k > m) m = k;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < 1024 / 64; ++i) m = max(m, part[i]);
    *kmax = m;
  }
}

__global__ void order_count_kernel(const int* key, int64_t n_q, int64_t n_t, const int* kmax,
                                   int* counts) {
  const int km = *kmax;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n_q;
       q += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&counts[order_bucket(key[q], n_t, km)], 1);
}

// Exclusive prefix sum of counts[0, kOrderBuckets) in place, one workgroup of
// 1024 threads, kOrderBuckets / 1024 consecutive buckets per thread.
__global__ __launch_bounds__(1024) void order_scan_kernel(int* counts) {
  constexpr int kPer = kOrderBuckets / 1024;
  __shared__ int part[1024];
  const int t = threadIdx.x;
  int v[kPer], s = 0;
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    v[i] = counts[t * kPer + i];
    s += v[i];
  }
  part[t] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {  // Hillis-Steele inclusive scan
    const int x = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += x;
    __syncthreads();
  }
  int run = part[t] - s;  // exclusive prefix of this thread's run
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    counts[t * kPer + i] = run;
    run += v[i];
  }
}

__global__ void order_scatter_kernel(const int* key, int64_t n_q, int64_t n_t, const int* kmax,
                                     int* offsets, int* order) {
  const int km = *kmax;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n_q;
       q += (int64_t)gridDim.x * blockDim.x)
    order[atomicAdd(&offsets[order_bucket(key[q], n_t, km)], 1)] = (int)q;
}

}  // namespace

void launch_collect_order(const int* widen, const int* coll_cnt, const int* coll_idx,
                          int coll_cap, int64_t n_q, int64_t n_t, int* key, int* counts,
                          int* order, hipStream_t s) {
  if (n_q <= 0) return;
  hipLaunchKernelGGL(collect_key_kernel, dim3((unsigned)ceil_div(n_q, kThreads / 64)),
                     dim3(kThreads), 0, s, widen, coll_cnt, coll_idx, coll_cap, n_q, n_t, key);
  launch_fill_u32(reinterpret_cast<unsigned*>(counts), kOrderBuckets, 0u, s);
  int* kmax = counts + kOrderBuckets;  // the scan leaves this entry alone
  hipLaunchKernelGGL(order_max_kernel, dim3(1), dim3(1024), 0, s, key, n_q, n_t, kmax);
  const unsigned g = (unsigned)std::min<int64_t>(1024, ceil_div(n_q, (int64_t)kThreads));
  hipLaunchKernelGGL(order_count_kernel, dim3(g), dim3(kThreads), 0, s, key, n_q, n_t, kmax,
                     counts);
  hipLaunchKernelGGL(order_scan_kernel, dim3(1), dim3(1024), 0, s, counts);
  hipLaunchKernelGGL(order_scatter_kernel, dim3(g), dim3(kThreads), 0, s, key, n_q, n_t, kmax,
                     counts, order);
}

namespace {

// One atomic per wave reserves the slots of its flagged entries.
__global__ __launch_bounds__(kThreads) void pack_flagged_kernel(const int* flags, int64_t n,
                                                                int value, int* list,
                                                                int* count) {
  const int lane = threadIdx.x & 63;
  for (int64_t base = (int64_t)blockIdx.x * kThreads; base < n;
       base += (int64_t)gridDim.x * kThreads) {
    const int64_t q = base + threadIdx.x;
    const bool on = q < n && flags[q] == value;
    const uint64_t m = __builtin_amdgcn_ballot_w64(on);
    if (m == 0) continue;  // uniform across the wave
    int at = 0;
    if (lane == 0) at = atomicAdd(count, __popcll(m));
    at = __shfl(at, 0, 64);
    const int rank = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                               __builtin_amdgcn_mbcnt_lo((unsigned)m, 0));
    if (on) list[at + rank] = (int)q;
  }
}

// One wave per query with widen[q] == 2: the entries of its split lists not
// above its collect threshold, appended in list order (ballot ranks, no
// atomics).  Splits cover disjoint row ranges, so no row comes twice.
__global__ __launch_bounds__(kThreads) void list_collect_kernel(
    const int* __restrict__ widen, const unsigned long long* __restrict__ tau,
    const float* __restrict__ part_d, const int* __restrict__ part_i,
    const int* __restrict__ part_n, int64_t n_q, int n_splits, int kc, int64_t n_t,
    int* __restrict__ coll_cnt, int* __restrict__ coll_idx, int coll_cap) {
  const int lane = threadIdx.x & 63;
  const int64_t q = (int64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  if (q >= n_q || widen[q] != 2) return;  // whole wave
  const float t = (float)from_ordered_key(tau[q]);
  const int total = n_splits * kc;
  const float* pd = part_d + (size_t)q * total;
  const int* pi = part_i + (size_t)q * total;
  const int* pn = part_n ? part_n + (size_t)q * n_splits : nullptr;
  int* out = coll_idx + (size_t)q * coll_cap;
  int n = 0;
  // As the collect launch's test: a NaN value is not above t.
  auto append = [&](bool keep, int id) {
    const uint64_t m = __builtin_amdgcn_ballot_w64(keep);
    const int at = n + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((unsigned)m, 0));
    if (keep && at < coll_cap) out[at] = id;
    n += __popcll(m);
  };
  if (kc >= 64 && pn) {
    // Long lists: list by list up to each one's entry count (FP8 on 1,000
    // Gaussians: about 190 of 400 slots filled per list).
    for (int l = 0; l < n_splits; ++l) {
      const int len = __builtin_amdgcn_readfirstlane(pn[l]);
      for (int j = lane; j - lane < len; j += 64) {
        bool keep = false;
        int id = -1;
        if (j < len) {
          id = pi[l * kc + j];
          keep = !(pd[l * kc + j] > t) && id >= 0 && id < n_t;
        }
        append(keep, id);
      }
    }
  } else {
    // Short lists: all slots of all lists in one flat sweep.
    for (int e0 = 0; e0 < total; e0 += 64) {
      const int e = e0 + lane;
      bool keep = false;
      int id = -1;
      if (e < total) {
        const int l = e / kc;
        if (!pn || e - l * kc < pn[l]) {
          id = pi[e];
          keep = !(pd[e] > t) && id >= 0 && id < n_t;
        }
      }
      append(keep, id);
    }
  }
  if (lane == 0) coll_cnt[q] = n;
}

}  // namespace

void launch_pack_flagged(const int* flags, int64_t n, int value, int* list, int* count,
                         hipStream_t s) {
  if (n <= 0) return;
  const unsigned g = (unsigned)std::min<int64_t>(1024, ceil_div(n, (int64_t)kThreads));
  hipLaunchKernelGGL(pack_flagged_kernel, dim3(g), dim3(kThreads), 0, s, flags, n, value, list,
                     count);
}

void launch_list_collect(const int* widen, const unsigned long long* tau, const float* part_d,
                         const int* part_i, const int* part_n, int64_t n_q, int n_splits, int kc,
                         int64_t n_t, int* coll_cnt, int* coll_idx, int coll_cap, hipStream_t s) {
  if (n_q <= 0) return;
  const int64_t blocks = ceil_div(n_q, (int64_t)(kThreads / 64));
  hipLaunchKernelGGL(list_collect_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, s, widen, tau,
                     part_d, part_i, part_n, n_q, n_splits, kc, n_t, coll_cnt, coll_idx, coll_cap);
}

void launch_count_i32(const int* f, int64_t n, int value, unsigned int* out, hipStream_t s) {
  if (n <= 0) return;
  const int64_t blocks = std::min<int64_t>(64, ceil_div(n, (int64_t)kThreads));
  hipLaunchKernelGGL(count_i32_kernel, dim3((unsigned)std::max<int64_t>(1, blocks)),
                     dim3(kThreads), 0, s, f, n, value, out);
}

void launch_count_uncertified(const unsigned char* uncert, int64_t n, DeviceScalars* sc,
                              hipStream_t s) {
  if (n <= 0) return;
  const int64_t blocks = std::min<int64_t>(64, ceil_div(ceil_div(n, 16), (int64_t)kThreads));
  hipLaunchKernelGGL(count_flags_kernel, dim3((unsigned)std::max<int64_t>(1, blocks)),
                     dim3(kThreads), 0, s, uncert, n, &sc->n_uncertified);
}

template void launch_fill_keys<float>(unsigned long long*, int64_t, unsigned long long, hipStream_t);
template void launch_row_norms<float>(const float*, int64_t, int, float*, DeviceScalars*, hipStream_t);
template void launch_row_norms<double>(const double*, int64_t, int, double*, DeviceScalars*, hipStream_t);
template void launch_merge_rerank<float>(const MergeArgs<float>&, hipStream_t);
template void launch_merge_rerank<double>(const MergeArgs<double>&, hipStream_t);
template void launch_collect_rerank<float>(const CollectArgs<float>&, hipStream_t);
template void launch_collect_rerank<double>(const CollectArgs<double>&, hipStream_t);

}  // namespace knn
