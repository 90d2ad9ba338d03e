// Generic fused distance + candidate-selection kernel (SIMT FMA, any d, any
// k' up to kMaxCandidates, float or double).
//
// Role in the reference: the GEMM + heap-push loop of _k_closest_on_chunk
// (src/_pdist_aggregation.pyx:192-232) and the SYCL kern_pwdist2/kern_topk
// pair (sycl_ext/parallel_knn.hpp:199-289, :397-466), without ever storing
// the query x train distance tile in HBM.
//
// Workgroup (qb, s) owns 64 queries and the train rows of split s.  Per step
// of 64 train rows it streams 16-feature slices of the queries (scaled by -2,
// exact) and of the rows through two LDS buffers -- the next slice is loaded
// into registers while the matrix cores work on the current one -- and runs
// the 64 x 64 tile on v_mfma_f64_16x16x4_f64 (double) or
// v_mfma_f32_16x16x4_f32 (float): wave w takes queries [16w, 16w + 16) and
// the four 16-row blocks, so lane l ends with the values of four queries
// (16w + Mfma16::row(l/16, i), i < 4) and rows 16j + l%16 (j < 4), the
// accumulators starting at |y|^2.  The reduced distance v = |y|^2 - 2 x.y (the reference's "reduced
// distance", :225-232) is compared with the query's threshold tau; survivors are
// appended to the query's candidate list in global scratch.  When a list
// could overflow on the next step, the workgroup sorts it by (v, index),
// keeps the best k' and lowers tau to the k'-th value.  tau is also shared
// between the workgroups of one query through an atomicMin on an ordered key,
// so every split filters with the best threshold found anywhere.
//
// Two more launches of the same kernel serve the sampled threshold (the
// guess): the sample launch reads a few regions of rows (split s starts at
// s * split_stride) and writes each query's kSampleJ smallest reduced
// distances per region, ascending, rounded up to float, to sample_out without
// touching tau; the re-search launch runs only the query blocks and queries
// whose guessed threshold proved too tight (qblock_active / query_active),
// from tau = +inf.  A plain search starts every query from the global tau,
// which holds the guess when there is one and +inf otherwise.
#include "knn_common.hpp"
#include "knn_kernels.hpp"

namespace knn {

namespace {

constexpr int kGenSearch = 0, kGenSample = 1, kGenResearch = 2;

// 16 x 16 x 4 MFMAs.  Lane l supplies A[l%16][l/16] and B[l/16][l%16] and
// holds D rows row(l/16, r), r < 4, of column l%16 -- measured on gfx950:
// blocks of four rows for f32 (4(l/16) + r), interleaved for f64 (l/16 + 4r).
// Raw buffer load of one element at byte offset voff + soff (0 outside the
// resource's range).
template <class T>
__device__ inline T buffer_load(__amdgpu_buffer_rsrc_t r, int voff, int soff);
template <>
__device__ inline float buffer_load<float>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <>
__device__ inline double buffer_load<double>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  typedef unsigned u2 __attribute__((ext_vector_type(2)));
  const u2 v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
  return __builtin_bit_cast(double, v);
}

template <class T>
struct Mfma16;
template <>
struct Mfma16<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  __device__ static v4 op(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int g, int r) { return g + 4 * r; }
};
template <>
struct Mfma16<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  __device__ static v4 op(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int g, int r) { return 4 * g + r; }
};

// Sort and cut to k' every list of the workgroup longer than min_count: each
// wave takes queries w, w+4, ... and sorts them alone (wave_sort_cut in its
// own LDS slice sd/si + w * cap), so four lists are sorted at once and no
// workgroup barrier runs inside a sort.
template <class T, bool SHARE>
__device__ void compact_lists(const SearchArgs<T>& a, T* cd, int* ci, T* tau_s, int* cnt_s,
                              int64_t q0, int min_count, T* sd_all, int* si_all) {
  __syncthreads();  // appends and counts visible
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  T* sd = sd_all + (size_t)w * a.cap;
  int* si = si_all + (size_t)w * a.cap;
  for (int qq = w; qq < kGenQ; qq += kThreads / 64) {
    const int n = cnt_s[qq];  // uniform: LDS value written before the barrier
    if (n <= min_count) continue;
    wave_sort_cut(cd + (size_t)qq * a.cap, ci + (size_t)qq * a.cap, n, a.kc, sd, si, lane);
    if (lane == 0) {
      const int m = n < a.kc ? n : a.kc;
      cnt_s[qq] = m;
      if (m == a.kc) {
        const T kth = sd[a.kc - 1];
        if (kth < tau_s[qq]) tau_s[qq] = kth;
        if (SHARE) atomicMin(&a.tau[q0 + qq], ordered_key((double)kth));
      }
    }
  }
  __syncthreads();  // counts and thresholds visible to the workgroup
}

// compact_lists for the sampling launch (no tau sharing).  Wave w takes the
// queries 4w..4w+3, 4w+16.., ... four at a time.  When none of the four lists
// holds more than 32 entries (every list of a region of up to 32 rows), each
// 16-lane quarter of the wave sorts one of them in its own 32-entry slice of
// the wave's LDS scratch (the capacity is 128), so the four sorts run
// together; longer lists are sorted one after another by the whole wave
// (wave_sort_cut).  Both are the same total order on (value, index), so the
// kept entries do not depend on which form ran.  Used by the sampling launch
// only: in the search launches the extra code costs the step loop registers
// and instruction-cache room (float64 1e5 x 1e3 x 500 search 4.0 -> 4.5 ms).
// The loops stay rolled for the same reason (sampling launch at 1e4 x 1e3 x
// 50 float64: 0.10 ms unrolled, 0.04 ms rolled).
template <class T>
__device__ void compact_lists_sample(const SearchArgs<T>& a, T* cd, int* ci, T* tau_s, int* cnt_s,
                                     int min_count, T* sd_all, int* si_all) {
  constexpr int kSmall = 32;  // entries per list a 16-lane quarter sorts
  __syncthreads();  // appends and counts visible
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  T* sd = sd_all + (size_t)w * a.cap;
  int* si = si_all + (size_t)w * a.cap;
#pragma unroll 1
  for (int base = 4 * w; base < kGenQ; base += 4 * (kThreads / 64)) {
    int nmax = 0;  // uniform: LDS values written before the barrier or by this wave
#pragma unroll
    for (int g = 0; g < 4; ++g) nmax = max(nmax, cnt_s[base + g]);
    if (nmax <= min_count) continue;
    if (nmax <= kSmall) {
      const int g = lane >> 4, gl = lane & 15;
      const int qq = base + g;
      const int n = cnt_s[qq];
      const bool act = n > min_count;
      const int P = next_pow2(nmax);
      T* gd = sd + g * kSmall;
      int* gi = si + g * kSmall;
      T* cq = cd + (size_t)qq * a.cap;
      int* iq = ci + (size_t)qq * a.cap;
      for (int e = gl; e < P; e += 16) {
        const bool ok = e < n;
        gd[e] = ok ? cq[e] : pos_inf<T>();
        gi[e] = ok ? iq[e] : INT_MAX;
      }
      wave_sync();
#pragma unroll 1
      for (int size = 2; size <= P; size <<= 1) {
#pragma unroll 1
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
          if (gl < (P >> 1)) {
            const int lo = ((gl & ~(stride - 1)) << 1) | (gl & (stride - 1));
            const int hi = lo + stride;
            const bool up = (lo & size) == 0;
            const T dl = gd[lo], dh = gd[hi];
            const int il = gi[lo], ih = gi[hi];
            const bool swap = up ? key_less(dh, ih, dl, il) : key_less(dl, il, dh, ih);
            if (swap) {
              gd[lo] = dh;
              gd[hi] = dl;
              gi[lo] = ih;
              gi[hi] = il;
            }
          }
          wave_sync();
        }
      }
      const int m = n < a.kc ? n : a.kc;
      if (act) {
        for (int e = gl; e < m; e += 16) {
          cq[e] = gd[e];
          iq[e] = gi[e];
        }
        if (gl == 0) {
          cnt_s[qq] = m;
          if (m == a.kc) {
            const T kth = gd[a.kc - 1];
            if (kth < tau_s[qq]) tau_s[qq] = kth;
          }
        }
      }
      wave_sync();
      continue;
    }
#pragma unroll 1
    for (int g = 0; g < 4; ++g) {
      const int qq = base + g;
      const int n = cnt_s[qq];
      if (n <= min_count) continue;
      wave_sort_cut(cd + (size_t)qq * a.cap, ci + (size_t)qq * a.cap, n, a.kc, sd, si, lane);
      if (lane == 0) {
        const int m = n < a.kc ? n : a.kc;
        cnt_s[qq] = m;
        if (m == a.kc) {
          const T kth = sd[a.kc - 1];
          if (kth < tau_s[qq]) tau_s[qq] = kth;
        }
      }
    }
  }
  __syncthreads();  // counts and thresholds visible to the workgroup
}

template <class T, int MODE>
__global__ __launch_bounds__(kThreads, 4) void generic_search_kernel(SearchArgs<T> a) {
  constexpr bool kShare = MODE != kGenSample;  // a sample's k-th value bounds nothing
  constexpr int kRowPad = 16 / sizeof(T);  // keeps every slice row 16-byte aligned
  using M = Mfma16<T>;
  using V4 = typename M::v4;
  // two slice buffers: slice t in buffer t & 1
  __shared__ __align__(16) T xs[2][kGenK][kGenQ + kRowPad];
  __shared__ __align__(16) T ys[2][kGenK][kGenY + kRowPad];
  __shared__ T tau_s[kGenQ];
  __shared__ int cnt_s[kGenQ];
  __shared__ int flag_s;
  extern __shared__ __align__(16) unsigned char dyn_smem[];  // 4 waves x cap x (T + int)
  T* sd = reinterpret_cast<T*>(dyn_smem);
  int* si = reinterpret_cast<int*>(dyn_smem + (size_t)(kThreads / 64) * a.cap * sizeof(T));

  if (MODE == kGenResearch && a.qblock_active[blockIdx.x] == 0) return;  // whole workgroup
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int64_t q0 = (int64_t)blockIdx.x * kGenQ;
  const int64_t ybeg =
      (int64_t)blockIdx.y * (a.split_stride > 0 ? a.split_stride : a.rows_per_split);
  const int64_t yend = ybeg + a.rows_per_split < a.n_t ? ybeg + a.rows_per_split : a.n_t;
  const size_t blk = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
  T* cd = a.cand_d + blk * kGenQ * a.cap;
  int* ci = a.cand_i + blk * kGenQ * a.cap;
  const int64_t n_q_here = a.n_q - q0 < kGenQ ? a.n_q - q0 : kGenQ;

  if (tid < kGenQ) {
    T t = pos_inf<T>();
    if (kShare && tid < n_q_here) {
      if (MODE == kGenResearch && a.query_active[q0 + tid] == 0) {
        t = -pos_inf<T>();  // not re-searched: appends nothing
      } else {
        const T g = (T)from_ordered_key(load_agent(&a.tau[q0 + tid]));
        if (g < t) t = g;
      }
    }
    tau_s[tid] = t;
    cnt_s[tid] = 0;
  }
  if (tid == 0) flag_s = 0;
  __syncthreads();

  // Slice loads: thread tid brings elements tid + 256 u (u < kPer) of the
  // 64 x kGenK query and row slices (kGenK consecutive features of one row per
  // kGenK lanes).
  // Features past d and rows past the query block / split read as zero.
  constexpr int kPer = kGenQ * kGenK / kThreads;
  static_assert(kPer * kThreads == kGenQ * kGenK && kGenQ == kGenY && kGenK % 4 == 0,
                "slice geometry");
  const int n_slices = (a.d + kGenK - 1) / kGenK;
  // Element u of a thread: row tid / kGenK + u * (kThreads / kGenK), feature
  // tid % kGenK of the slice.  Loads are buffer loads: a scalar base per query
  // block / row step, one lane offset for all u, the slice and u in scalar
  // offsets; rows past the block or the split fall outside the buffer range
  // and read as zero, features past d are zeroed here.
  constexpr int kRowsPerU = kThreads / kGenK;
  const int lr = tid / kGenK, lc = tid % kGenK;
  const int lane_off = (lr * a.d + lc) * (int)sizeof(T);
  const int u_off = kRowsPerU * a.d * (int)sizeof(T);
  const __amdgpu_buffer_rsrc_t xrs = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<T*>(a.X + q0 * a.d), (short)0, (int)(n_q_here * a.d * (int64_t)sizeof(T)),
      0x00020000);
  T xr[kPer], yr[kPer];
  auto load_slice = [&](int64_t y0, int p0) {
    const int64_t rows = yend - y0 < kGenY ? yend - y0 : kGenY;
    const __amdgpu_buffer_rsrc_t yrs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<T*>(a.Y + y0 * a.d), (short)0, (int)(rows * a.d * (int64_t)sizeof(T)),
        0x00020000);
    const bool in_d = lc + p0 < a.d;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int so = u * u_off + p0 * (int)sizeof(T);
      const T xv = buffer_load<T>(xrs, lane_off, so);
      const T yv = buffer_load<T>(yrs, lane_off, so);
      xr[u] = in_d ? T(-2) * xv : T(0);
      yr[u] = in_d ? yv : T(0);
    }
  };
  auto store_slice = [&](int b) {
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      xs[b][lc][lr + u * kRowsPerU] = xr[u];
      ys[b][lc][lr + u * kRowsPerU] = yr[u];
    }
  };
  // MFMA operands: A = queries (lane l: query l%16 of the wave's block,
  // feature l/16 of the 4-feature step), B = rows (row l%16 of block j).
  const int fk = lane >> 4, fm = lane & 15;
  if (yend > ybeg) load_slice(ybeg, 0);
  store_slice(0);
  __syncthreads();
  int t = 0;  // slices computed so far (buffer t & 1 holds slice t)

  int step = 0;
  for (int64_t y0 = ybeg; y0 < yend; y0 += kGenY, ++step) {
    V4 acc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t y = y0 + 16 * j + fm;
      const T n0 = y < yend ? a.ynorm[y] : pos_inf<T>();
      acc[j] = V4{n0, n0, n0, n0};
    }
    for (int sl = 0; sl < n_slices; ++sl, ++t) {
      const int b = t & 1;
      // prefetch the next slice (of this step, or the first of the next one)
      const bool more = sl + 1 < n_slices || y0 + kGenY < yend;
      if (more) load_slice(sl + 1 < n_slices ? y0 : y0 + kGenY,
                           sl + 1 < n_slices ? (sl + 1) * kGenK : 0);
#pragma unroll
      for (int k0 = 0; k0 < kGenK; k0 += 4) {
        const T av = xs[b][k0 + fk][16 * w + fm];
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[j] = M::op(av, ys[b][k0 + fk][16 * j + fm], acc[j]);
      }
      if (more) store_slice(b ^ 1);  // buffer b^1 was last read before the previous barrier
      __syncthreads();
    }

    // Filter: append every reduced distance <= tau to the query's list.
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int qq = 16 * w + M::row(fk, i);
      if (qq >= n_q_here) continue;
      const T tq = tau_s[qq];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t y = y0 + 16 * j + fm;
        if (y >= yend) continue;
        const T v = acc[j][i];
        if (v <= tq) {
          const int pos = atomicAdd(&cnt_s[qq], 1);
          cd[(size_t)qq * a.cap + pos] = v;
          ci[(size_t)qq * a.cap + pos] = (int)y;
          if (pos + 1 > a.cap - kGenY) flag_s = 1;
        }
      }
    }
    __syncthreads();
    if (flag_s) {
      if constexpr (MODE == kGenSample)
        compact_lists_sample<T>(a, cd, ci, tau_s, cnt_s, a.cap - kGenY, sd, si);
      else
        compact_lists<T, kShare>(a, cd, ci, tau_s, cnt_s, q0, a.cap - kGenY, sd, si);
      if (tid == 0) flag_s = 0;
    }
    if (kShare && (step & 7) == 7 && tid < n_q_here) {
      const T g = (T)from_ordered_key(load_agent(&a.tau[q0 + tid]));
      if (g < tau_s[tid]) tau_s[tid] = g;
    }
    __syncthreads();
  }

  // Final: every list sorted and cut to k'; written to this split's slot.
  if constexpr (MODE == kGenSample)
    compact_lists_sample<T>(a, cd, ci, tau_s, cnt_s, 0, sd, si);
  else
    compact_lists<T, kShare>(a, cd, ci, tau_s, cnt_s, q0, 0, sd, si);
  if (MODE == kGenSample) {
    // kc == kSampleJ here; rounding up keeps every sampled value >= its row's.
    for (int e = tid; e < kGenQ * kSampleJ; e += kThreads) {
      const int qq = e / kSampleJ, j = e % kSampleJ;
      if (qq >= n_q_here) continue;
      const int n = cnt_s[qq];
      float v = __builtin_huge_valf();
      if (j < n) v = sizeof(T) == 4 ? (float)cd[(size_t)qq * a.cap + j]
                                    : __double2float_ru((double)cd[(size_t)qq * a.cap + j]);
      a.sample_out[((size_t)(q0 + qq) * a.n_splits + blockIdx.y) * kSampleJ + j] = v;
    }
    return;
  }
  for (int e = tid; e < kGenQ * a.kc; e += kThreads) {
    const int qq = e / a.kc, j = e % a.kc;
    if (qq >= n_q_here) continue;
    const int n = cnt_s[qq];
    const size_t o = ((size_t)(q0 + qq) * a.n_splits + blockIdx.y) * a.kc + j;
    a.part_d[o] = j < n ? cd[(size_t)qq * a.cap + j] : pos_inf<T>();
    a.part_i[o] = j < n ? ci[(size_t)qq * a.cap + j] : -1;
  }
}

}  // namespace

template <class T, int MODE>
void launch_generic_mode(const SearchArgs<T>& a, hipStream_t s) {
  if (a.n_q <= 0) return;
  dim3 grid((unsigned)ceil_div(a.n_q, kGenQ), (unsigned)a.n_splits);
  const size_t lds = (size_t)(kThreads / 64) * a.cap * (sizeof(T) + sizeof(int));
  hipLaunchKernelGGL((generic_search_kernel<T, MODE>), grid, dim3(kThreads), lds, s, a);
}

template <class T>
void launch_generic_search(const SearchArgs<T>& a, hipStream_t s) {
  launch_generic_mode<T, kGenSearch>(a, s);
}

template <class T>
bool launch_generic_sample(const SearchArgs<T>& a, hipStream_t s) {
  if (!a.sample_out || a.kc != kSampleJ) return false;
  launch_generic_mode<T, kGenSample>(a, s);
  return true;
}

template <class T>
bool launch_generic_research(const SearchArgs<T>& a, hipStream_t s) {
  if (!a.qblock_active || !a.query_active) return false;
  launch_generic_mode<T, kGenResearch>(a, s);
  return true;
}

template void launch_generic_search<float>(const SearchArgs<float>&, hipStream_t);
template void launch_generic_search<double>(const SearchArgs<double>&, hipStream_t);
template bool launch_generic_sample<float>(const SearchArgs<float>&, hipStream_t);
template bool launch_generic_sample<double>(const SearchArgs<double>&, hipStream_t);
template bool launch_generic_research<float>(const SearchArgs<float>&, hipStream_t);
template bool launch_generic_research<double>(const SearchArgs<double>&, hipStream_t);

}  // namespace knn
