// Fused distance + candidate kernel for low-dimensional points (d < 16),
// float or double, VALU FMA.
//
// At d = 3 a GEMM formulation wastes the matrix cores (K would be padded
// 3 -> 32) and the generic kernel's 16-feature LDS slices waste 13/16 of its
// FMAs.  Here every lane owns ONE query whose -2 x sits in registers; the
// workgroup stages 1024 train elements (256 rows at d <= 4) per stage in LDS,
// double buffered, and every lane walks them 32 rows at a time (broadcast LDS
// reads), so a candidate costs d FMAs plus its share of a min-tree, and each
// stage's global loads have a whole stage of compute to arrive in.  Selection is the same contract as the MFMA kernel: append
// v = |y|^2 - 2 x.y <= tau to the query's list in global scratch, compact
// with a wave-local sort, share tau across splits with atomicMin.
#include "knn_common.hpp"
#include "knn_kernels.hpp"

namespace knn {

namespace {

constexpr int kRows = 32;  // train rows per step (one pass of the value registers)
// Train rows per LDS stage: 1024 elements, i.e. 4 per thread.
template <int DP>
constexpr int stage_rows() { return 1024 / DP; }

// MODE kLowdSearch: the candidate search.  kLowdResearch: the same for the
// query blocks / queries selected by qblock_active / query_active.
// kLowdSample: the same sweep over sampled row ranges, keeping each lane's
// kSampleJ smallest 32-row minima (a threshold estimate, no lists).
constexpr int kLowdSearch = 0, kLowdSample = 1, kLowdResearch = 2;

// NIR (d < DP): the row's first padding slot carries |y|^2, with weight 1 on
// the query side, so one LDS read per row brings features and norm together
// (65,536^2 x 3: search 2.13 -> 1.96 ms; the reads, 16 bytes broadcast to 64
// lanes per row, share the bound with the FMAs).
template <class T, int DP, int MODE, bool NIR>
__global__ __launch_bounds__(kThreads) void lowd_search_kernel(LowdArgs<T> a) {
  constexpr int SR = stage_rows<DP>();
  constexpr int EPT = SR * DP / kThreads;  // staged elements per thread
  static_assert(SR * DP == EPT * kThreads && SR <= kThreads && SR % kRows == 0, "stage shape");
  __shared__ T ys[2][SR][DP];
  __shared__ T yn_s[2][NIR ? 1 : SR];  // NIR: the norms ride in ys
  __shared__ int cnt_s[kThreads];
  extern __shared__ __align__(16) unsigned char dyn_smem[];  // 4 waves x cap x (T + int)

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  int qb, split;
  {  // same XCD-aware order as the MFMA kernel
    const int b = blockIdx.x;
    const int per_xcd = (int)(gridDim.x >> 3);
    const int w_item = (b & 7) * per_xcd + (b >> 3);
    if (w_item >= a.n_qblocks * a.n_splits) return;
    split = w_item / a.n_qblocks;
    qb = w_item % a.n_qblocks;
  }
  if (MODE == kLowdResearch && a.qblock_active[qb] == 0) return;  // block not selected
  const int64_t q0 = (int64_t)qb * kThreads;
  const int64_t q = q0 + tid;
  const bool qok = q < a.n_q;
  const int64_t ybeg = (int64_t)split * (a.split_stride > 0 ? a.split_stride : a.rows_per_split);
  const int64_t yend = ybeg + a.rows_per_split < a.n_t ? ybeg + a.rows_per_split : a.n_t;
  const size_t blk = (size_t)split * a.n_qblocks + qb;
  T* cand_d = a.cand_d + blk * kThreads * a.cap;
  int* cand_i = a.cand_i + blk * kThreads * a.cap;
  T* sd = reinterpret_cast<T*>(dyn_smem + (size_t)w * a.cap * (sizeof(T) + sizeof(int)));
  int* si = reinterpret_cast<int*>(sd + a.cap);
  T* my_d = cand_d + (size_t)tid * a.cap;
  int* my_i = cand_i + (size_t)tid * a.cap;

  T xm2[DP];
#pragma unroll
  for (int f = 0; f < DP; ++f)
    xm2[f] = (qok && f < a.d) ? T(-2) * a.X[q * a.d + f] : (NIR && f == a.d ? T(1) : T(0));
  // +inf marks rows past the end; lanes without a (selected) query take -inf.
  const bool active = qok && (MODE != kLowdResearch || a.query_active[q] != 0);
  T tau = active ? max_finite<T>() : -pos_inf<T>();
  if (MODE != kLowdSample && active) {  // start from the best threshold known (sampled guess)
    const T g = (T)from_ordered_key(load_agent(&a.tau[q]));
    tau = g < tau ? g : tau;
  }
  float best[MODE == kLowdSample ? kSampleJ : 1];  // sample: kSampleJ smallest minima, ascending
#pragma unroll
  for (int i = 0; i < (MODE == kLowdSample ? kSampleJ : 1); ++i) best[i] = __builtin_huge_valf();
  int cnt = 0;  // entries in this lane's list (published in cnt_s for wave-wide passes)

  // stage loader: SR rows x DP elements (zero padded) + norms (+inf past the end)
  T pre[EPT];
  T pre_n = T(0);
  auto load_tile = [&](int64_t y0) {
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
      const int e = tid + i * kThreads;
      const int r = e / DP, f = e % DP;
      const int64_t y = y0 + r;
      pre[i] = (y < yend && f < a.d) ? a.Y[y * a.d + f] : T(0);
      if (NIR && f == a.d) pre[i] = y < yend ? a.ynorm[y] : pos_inf<T>();
    }
    if (!NIR && tid < SR) pre_n = (y0 + tid < yend) ? a.ynorm[y0 + tid] : pos_inf<T>();
  };
  auto store_tile = [&](int buf) {
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
      const int e = tid + i * kThreads;
      ys[buf][e / DP][e % DP] = pre[i];
    }
    if (!NIR && tid < SR) yn_s[buf][tid] = pre_n;
  };
  load_tile(ybeg);
  store_tile(0);
  __syncthreads();

  int cur = 0, step = 0;
  for (int64_t ys0 = ybeg; ys0 < yend; ys0 += SR) {
    const bool more = ys0 + SR < yend;
    if (more) load_tile(ys0 + SR);
  for (int sub = 0; sub < SR && ys0 + sub < yend; sub += kRows, ++step) {
    const int64_t y0 = ys0 + sub;
    T v[kRows];
    T m = pos_inf<T>();
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      T acc = NIR ? xm2[0] * ys[cur][sub + r][0] : yn_s[cur][sub + r];
#pragma unroll
      for (int f = NIR ? 1 : 0; f < DP; ++f) acc = fma(xm2[f], ys[cur][sub + r][f], acc);
      v[r] = acc;
    }
    // min tree of 3-input minima (v_min3); a NaN value is passed over, so it
    // hides no other row of the step (and is itself never appended).
#pragma unroll
    for (int r = 0; r + 1 < kRows; r += 2) m = fmin(m, fmin(v[r], v[r + 1]));
    if constexpr (MODE == kLowdSample) {
      // Insert m into the sorted run: b[i] <- min(b[i], max(b[i-1], m)).
      // Rounded up for double (a sampled value never below its row's); a NaN
      // becomes +inf, which leaves the run unchanged.
      const float m32 = sizeof(T) == 4 ? (float)m : __double2float_ru((double)m);
      const float mf = m32 == m32 ? m32 : __builtin_huge_valf();
      if (__any(qok && mf < best[kSampleJ - 1])) {
#pragma unroll
        for (int i = kSampleJ - 1; i > 0; --i) best[i] = fminf(best[i], fmaxf(best[i - 1], mf));
        best[0] = fminf(best[0], mf);
      }
      continue;
    }
    bool overflow = false;
    if (__any(m <= tau)) {
      if (m <= tau) {
        int pos = cnt;  // this lane owns the query: no atomic
#pragma unroll
        for (int r = 0; r < kRows; ++r) {
          if (v[r] <= tau) {
            my_d[pos] = v[r];
            my_i[pos] = (int)(y0 + r);
            ++pos;
          }
        }
        cnt = pos;
        overflow = pos > a.cap - kRows;
      }
    }
    if (__any(overflow)) {
      cnt_s[tid] = cnt;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
      wave_sync();
      for (int qq = 0; qq < 64; ++qq) {
        const int ql = w * 64 + qq;
        const int n = cnt_s[ql];
        if (n <= a.cap - kRows) continue;
        T* cd = cand_d + (size_t)ql * a.cap;
        int* ci = cand_i + (size_t)ql * a.cap;
        wave_sort_cut(cd, ci, n, a.kc, sd, si, lane);
        if (lane == 0) {
          cnt_s[ql] = n < a.kc ? n : a.kc;
          if (n >= a.kc) atomicMin(&a.tau[q0 + ql], ordered_key((double)sd[a.kc - 1]));
        }
        if (lane == qq && n >= a.kc) tau = tau < sd[a.kc - 1] ? tau : sd[a.kc - 1];
        wave_sync();
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
      cnt = cnt_s[tid];
    }
    if ((step & 15) == 15 && qok) {
      const T g = (T)from_ordered_key(load_agent(&a.tau[q]));
      tau = g < tau ? g : tau;
    }
  }
    if (more) store_tile(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }
  if constexpr (MODE == kLowdSample) {
    if (qok) {
      float* o = a.sample_out + ((size_t)q * a.n_splits + split) * kSampleJ;
#pragma unroll
      for (int i = 0; i < kSampleJ; ++i) o[i] = best[i];
    }
    return;
  }

  cnt_s[tid] = cnt;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  wave_sync();
  for (int qq = 0; qq < 64; ++qq) {
    const int ql = w * 64 + qq;
    const int n = cnt_s[ql];
    if (n <= 0) continue;
    wave_sort_cut(cand_d + (size_t)ql * a.cap, cand_i + (size_t)ql * a.cap, n, a.kc, sd, si, lane);
    if (lane == 0) cnt_s[ql] = n < a.kc ? n : a.kc;
    wave_sync();
  }
  for (int qq = 0; qq < 64; ++qq) {
    const int ql = w * 64 + qq;
    if (q0 + ql >= a.n_q) break;
    if (MODE == kLowdResearch && a.query_active[q0 + ql] == 0) continue;  // uniform
    const int n = cnt_s[ql];
    for (int j = lane; j < a.kc; j += 64) {
      const size_t o = ((size_t)(q0 + ql) * a.n_splits + split) * a.kc + j;
      a.part_d[o] = j < n ? cand_d[(size_t)ql * a.cap + j] : pos_inf<T>();
      a.part_i[o] = j < n ? cand_i[(size_t)ql * a.cap + j] : -1;
    }
  }
}

template <class T, int DP, int MODE>
void launch_dp(const LowdArgs<T>& a, hipStream_t s) {
  const size_t lds = MODE == kLowdSample ? 0 : (size_t)4 * a.cap * (sizeof(T) + sizeof(int));
  const unsigned blocks = (unsigned)((a.n_qblocks * a.n_splits + 7) / 8 * 8);
  if (a.d < DP)
    hipLaunchKernelGGL((lowd_search_kernel<T, DP, MODE, true>), dim3(blocks), dim3(kThreads), lds,
                       s, a);
  else
    hipLaunchKernelGGL((lowd_search_kernel<T, DP, MODE, false>), dim3(blocks), dim3(kThreads), lds,
                       s, a);
}

template <class T, int MODE>
bool launch_mode(const LowdArgs<T>& a, hipStream_t s) {
  if (a.d <= 4) launch_dp<T, 4, MODE>(a, s);
  else if (a.d <= 8) launch_dp<T, 8, MODE>(a, s);
  else if (a.d <= 16) launch_dp<T, 16, MODE>(a, s);
  else return false;
  return true;
}

template <class T, int DP>
int occupancy_dp(int d, int cap) {
  int n = 0;
  const size_t lds = (size_t)4 * cap * (sizeof(T) + sizeof(int));
  const hipError_t e =
      d < DP ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                   &n, lowd_search_kernel<T, DP, kLowdSearch, true>, kThreads, lds)
             : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                   &n, lowd_search_kernel<T, DP, kLowdSearch, false>, kThreads, lds);
  if (e != hipSuccess || n <= 0) return 2;
  return n;
}

}  // namespace

template <class T>
bool launch_lowd_search(const LowdArgs<T>& a, hipStream_t s) {
  return launch_mode<T, kLowdSearch>(a, s);
}

template <class T>
bool launch_lowd_sample(const LowdArgs<T>& a, hipStream_t s) {
  return a.sample_out && launch_mode<T, kLowdSample>(a, s);
}

template <class T>
bool launch_lowd_research(const LowdArgs<T>& a, hipStream_t s) {
  return a.qblock_active && a.query_active && launch_mode<T, kLowdResearch>(a, s);
}

template <class T>
int lowd_blocks_per_cu(int d, int cap) {
  if (d <= 4) return occupancy_dp<T, 4>(d, cap);
  if (d <= 8) return occupancy_dp<T, 8>(d, cap);
  return occupancy_dp<T, 16>(d, cap);
}

template int lowd_blocks_per_cu<float>(int, int);
template int lowd_blocks_per_cu<double>(int, int);
template bool launch_lowd_search<float>(const LowdArgs<float>&, hipStream_t);
template bool launch_lowd_search<double>(const LowdArgs<double>&, hipStream_t);
template bool launch_lowd_sample<float>(const LowdArgs<float>&, hipStream_t);
template bool launch_lowd_sample<double>(const LowdArgs<double>&, hipStream_t);
template bool launch_lowd_research<float>(const LowdArgs<float>&, hipStream_t);
template bool launch_lowd_research<double>(const LowdArgs<double>&, hipStream_t);

}  // namespace knn
