// C ABI and the device schedule of one argkmin call.
//
// Schedule (all on the caller's stream, nothing synchronises unless stats are
// requested):
//   1. fill   : per-query thresholds <- +inf, call scalars <- 0
//   2. norms  : |y_j|^2 in the search precision, and their maximum
//   3. search : fused distance + candidate kernel over (query block, split)
//   4. merge  : per query, merge the split lists, exact FP64 re-rank, keep k
// The reference does the same work as a host loop over chunk pairs
// (src/_pdist_aggregation.pyx:236-448) or as a queue of SYCL kernels per
// chunk pair (sycl_ext/parallel_knn.hpp:337-534).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>

#include "../../include/pdist_knn.h"
#include "knn_common.hpp"
#include "knn_kernels.hpp"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

constexpr size_t kAlign = 256;
size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// Reference chunk arithmetic (src/_pdist_aggregation.pyx:247-269).
int64_t ref_n_chunks(int64_t n, int64_t chunk_size) {
  const int64_t nsc = std::max<int64_t>(20, chunk_size);
  const int64_t per = std::min(n, nsc);
  if (per <= 0) return 0;
  const int64_t full = n / per;
  return full + (n != full * per);
}

// HIP events recorded on the call's own stream between the stages.
struct StageTimer {
  bool on;
  hipStream_t stream;
  hipEvent_t ev[4] = {};
  StageTimer(bool enable, hipStream_t s) : on(enable), stream(s) {
    if (!on) return;
    for (auto& e : ev)
      if (hipEventCreate(&e) != hipSuccess) on = false;
  }
  ~StageTimer() {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
  }
  void mark(int i) {
    if (on) (void)hipEventRecord(ev[i], stream);
  }
  double elapsed(int a, int b) {
    float ms = 0.f;
    if (!on || hipEventElapsedTime(&ms, ev[a], ev[b]) != hipSuccess) return -1.0;
    return ms;
  }
};

struct Plan {
  int64_t n_train = 0, n_test = 0, dim = 0, k = 0;
  int kc = 0, cap = 0, n_splits = 1, path = 0;
  int64_t rows_per_split = 0, n_qblocks = 0;
  size_t elem = 4;
  // workspace offsets
  size_t off_scalars = 0, off_ynorm = 0, off_tau = 0, off_cand_d = 0, off_cand_i = 0,
         off_part_d = 0, off_part_i = 0, total = 0;
};

int make_plan(Plan& p, size_t elem, int64_t n_train, int64_t n_test, int64_t dim, int64_t k,
              int precision, int64_t oversample) {
  if (n_train <= 0 || n_test < 0 || dim <= 0)
    return fail(KNN_EINVAL, "n_train and dim must be positive and n_test non-negative");
  if (k <= 0) return fail(KNN_EINVAL, "k must be positive");
  if (k > n_train) return fail(KNN_EINVAL, "k must not exceed n_train");
  if (n_train > INT32_MAX) return fail(KNN_EUNSUPPORTED, "n_train must be below 2^31");
  if (dim > 65536) return fail(KNN_EUNSUPPORTED, "dim above 65536 is not supported");
  if (precision != KNN_PRECISION_EXACT) {
    if (elem != 4) return fail(KNN_EINVAL, "low precision search needs float32 inputs");
    return fail(KNN_EUNSUPPORTED, "bf16/fp8 candidate search is not built in this library");
  }
  int64_t kc;
  if (oversample > 0) {
    if (oversample < k) return fail(KNN_EINVAL, "oversample (k') must be >= k");
    kc = oversample;
  } else {
    kc = k + std::max<int64_t>(8, k / 8);
  }
  kc = std::min<int64_t>(kc, n_train);
  if (kc > knn::kMaxCandidates) {
    if (k > knn::kMaxCandidates)
      return fail(KNN_EUNSUPPORTED, "k above " + std::to_string(knn::kMaxCandidates) +
                                         " is not supported");
    kc = knn::kMaxCandidates;
  }
  p.n_train = n_train;
  p.n_test = n_test;
  p.dim = dim;
  p.k = k;
  p.elem = elem;
  p.kc = (int)kc;
  p.path = 0;
  p.cap = knn::next_pow2((int)kc + knn::kGenY);
  p.n_qblocks = knn::ceil_div(std::max<int64_t>(n_test, 1), knn::kGenQ);
  // Enough workgroups to cover 256 CUs several times, each split at least
  // four row steps long, and the merged list within the merge kernel's LDS.
  int64_t splits = knn::ceil_div(2048, p.n_qblocks);
  splits = std::min<int64_t>(splits, knn::ceil_div(n_train, 4 * knn::kGenY));
  splits = std::min<int64_t>(splits, knn::kMaxMerge / p.kc);
  splits = std::max<int64_t>(splits, 1);
  p.rows_per_split = knn::ceil_div(knn::ceil_div(n_train, splits), knn::kGenY) * knn::kGenY;
  p.n_splits = (int)knn::ceil_div(n_train, p.rows_per_split);

  size_t o = 0;
  p.off_scalars = o;
  o = align_up(o + sizeof(knn::DeviceScalars));
  p.off_ynorm = o;
  o = align_up(o + (size_t)n_train * elem);
  p.off_tau = o;
  o = align_up(o + (size_t)std::max<int64_t>(n_test, 1) * sizeof(unsigned long long));
  const size_t ncand = (size_t)p.n_qblocks * p.n_splits * knn::kGenQ * p.cap;
  p.off_cand_d = o;
  o = align_up(o + ncand * elem);
  p.off_cand_i = o;
  o = align_up(o + ncand * sizeof(int));
  const size_t npart = (size_t)std::max<int64_t>(n_test, 1) * p.n_splits * p.kc;
  p.off_part_d = o;
  o = align_up(o + npart * elem);
  p.off_part_i = o;
  o = align_up(o + npart * sizeof(int));
  p.total = o;
  return KNN_OK;
}

template <class T>
int run(const T* X_train, int64_t n_train, const T* X_test, int64_t n_test, int64_t dim,
        int64_t k, int64_t chunk_size, int strategy, int precision, int64_t oversample,
        int flags, int n_threads_hint, T* dist_out, int64_t* idx_out, void* workspace,
        size_t workspace_bytes, void* stream_ptr, knn_stats* stats) {
  g_last_error.clear();
  if (strategy < KNN_STRATEGY_AUTO || strategy > KNN_STRATEGY_CHUNK_ON_TEST)
    return fail(KNN_EINVAL, "strategy must be auto, chunk_on_train or chunk_on_test");
  if (chunk_size <= 0) return fail(KNN_EINVAL, "chunk_size must be positive");
  if (!X_train || (n_test > 0 && (!X_test || !dist_out || !idx_out)))
    return fail(KNN_EINVAL, "null array pointer");
  Plan p;
  int rc = make_plan(p, sizeof(T), n_train, n_test, dim, k, precision, oversample);
  if (rc != KNN_OK) return rc;

  int threads = n_threads_hint > 0 ? n_threads_hint : (int)std::thread::hardware_concurrency();
  threads = std::max(threads, 1);
  int resolved = strategy;
  if (resolved == KNN_STRATEGY_AUTO)  // src/_pdist_aggregation.pyx:511-515
    resolved = (4 * chunk_size * (int64_t)threads < n_test) ? KNN_STRATEGY_CHUNK_ON_TEST
                                                             : KNN_STRATEGY_CHUNK_ON_TRAIN;
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->n_parallel_chunks = resolved == KNN_STRATEGY_CHUNK_ON_TEST
                                   ? ref_n_chunks(n_test, chunk_size)
                                   : ref_n_chunks(n_train, chunk_size);
    stats->k_candidates = p.kc;
    stats->n_splits = p.n_splits;
    stats->strategy_used = resolved;
    stats->kernel_path = p.path;
  }
  if (n_test == 0) return KNN_OK;

  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_ptr);
  unsigned char* ws = reinterpret_cast<unsigned char*>(workspace);
  bool owned = false;
  if (ws) {
    if (workspace_bytes < p.total)
      return fail(KNN_EWORKSPACE, "workspace of " + std::to_string(workspace_bytes) +
                                      " bytes is smaller than the " + std::to_string(p.total) +
                                      " required");
  } else {
    hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&ws), p.total, stream);
    if (e != hipSuccess)
      return fail(KNN_EHIP, std::string("hipMallocAsync: ") + hipGetErrorString(e));
    owned = true;
  }
  auto* sc = reinterpret_cast<knn::DeviceScalars*>(ws + p.off_scalars);
  T* ynorm = reinterpret_cast<T*>(ws + p.off_ynorm);
  auto* tau = reinterpret_cast<unsigned long long*>(ws + p.off_tau);
  StageTimer timer(stats && (flags & KNN_FLAG_TIMING), stream);
  timer.mark(0);

  // 1. fill: scalars are 4 ordered-key slots (max keys start at key(-inf)).
  knn::launch_fill_keys<float>(reinterpret_cast<unsigned long long*>(sc),
                               sizeof(knn::DeviceScalars) / 8,
                               0ull, stream);
  knn::launch_fill_keys<float>(tau, n_test, knn::ordered_key(__builtin_huge_val()), stream);
  // 2. norms
  knn::launch_row_norms<T>(X_train, n_train, (int)dim, ynorm, sc, stream);
  // 3. search
  knn::SearchArgs<T> sa;
  sa.X = X_test;
  sa.Y = X_train;
  sa.ynorm = ynorm;
  sa.n_q = n_test;
  sa.n_t = n_train;
  sa.d = (int)dim;
  sa.kc = p.kc;
  sa.cap = p.cap;
  sa.n_splits = p.n_splits;
  sa.rows_per_split = p.rows_per_split;
  sa.cand_d = reinterpret_cast<T*>(ws + p.off_cand_d);
  sa.cand_i = reinterpret_cast<int*>(ws + p.off_cand_i);
  sa.tau = tau;
  sa.part_d = reinterpret_cast<T*>(ws + p.off_part_d);
  sa.part_i = reinterpret_cast<int*>(ws + p.off_part_i);
  timer.mark(1);
  knn::launch_generic_search<T>(sa, stream);
  timer.mark(2);
  // 4. merge + exact re-rank
  knn::MergeArgs<T> ma;
  ma.X = X_test;
  ma.Y = X_train;
  ma.n_q = n_test;
  ma.n_t = n_train;
  ma.d = (int)dim;
  ma.part_d = sa.part_d;
  ma.part_i = sa.part_i;
  ma.n_splits = p.n_splits;
  ma.kc = p.kc;
  ma.k = (int)k;
  ma.squared = (flags & KNN_FLAG_SQUARED) ? 1 : 0;
  ma.xnorm = nullptr;
  ma.xerr = nullptr;
  ma.unit_roundoff = sizeof(T) == 4 ? 0x1p-24 : 0x1p-53;
  ma.acc_terms = 2 * (int)dim + 4;
  ma.sc = sc;
  ma.dist_out = dist_out;
  ma.idx_out = idx_out;
  knn::launch_merge_rerank<T>(ma, stream);
  timer.mark(3);

  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    if (owned) (void)hipFreeAsync(ws, stream);
    return fail(KNN_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
  }
  if (stats) {
    knn::DeviceScalars host_sc;
    e = hipMemcpyAsync(&host_sc, sc, sizeof(host_sc), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
      if (owned) (void)hipFreeAsync(ws, stream);
      return fail(KNN_EHIP, std::string("stats copy: ") + hipGetErrorString(e));
    }
    stats->n_uncertified = host_sc.n_uncertified;
    if (timer.on) {
      stats->prep_ms = timer.elapsed(0, 1);
      stats->search_ms = timer.elapsed(1, 2);
      stats->merge_ms = timer.elapsed(2, 3);
    }
  }
  if (owned) {
    e = hipFreeAsync(ws, stream);
    if (e != hipSuccess) return fail(KNN_EHIP, std::string("hipFreeAsync: ") + hipGetErrorString(e));
  }
  return KNN_OK;
}

}  // namespace

extern "C" {

int knn_argkmin_f32(const float* X_train, int64_t n_train, const float* X_test, int64_t n_test,
                    int64_t dim, int64_t k, int64_t chunk_size, int strategy, int precision,
                    int64_t oversample, int flags, int n_threads_hint, float* dist_out,
                    int64_t* idx_out, void* workspace, size_t workspace_bytes, void* stream,
                    knn_stats* stats) {
  return run<float>(X_train, n_train, X_test, n_test, dim, k, chunk_size, strategy, precision,
                    oversample, flags, n_threads_hint, dist_out, idx_out, workspace,
                    workspace_bytes, stream, stats);
}

int knn_argkmin_f64(const double* X_train, int64_t n_train, const double* X_test,
                    int64_t n_test, int64_t dim, int64_t k, int64_t chunk_size, int strategy,
                    int precision, int64_t oversample, int flags, int n_threads_hint,
                    double* dist_out, int64_t* idx_out, void* workspace,
                    size_t workspace_bytes, void* stream, knn_stats* stats) {
  if (precision != KNN_PRECISION_EXACT)
    return fail(KNN_EINVAL, "float64 inputs support only the exact precision");
  return run<double>(X_train, n_train, X_test, n_test, dim, k, chunk_size, strategy, precision,
                     oversample, flags, n_threads_hint, dist_out, idx_out, workspace,
                     workspace_bytes, stream, stats);
}

size_t knn_argkmin_workspace_size(int dtype_bytes, int64_t n_train, int64_t n_test, int64_t dim,
                                  int64_t k, int precision, int64_t oversample) {
  if (dtype_bytes != 4 && dtype_bytes != 8) return 0;
  Plan p;
  if (make_plan(p, (size_t)dtype_bytes, n_train, n_test, dim, k, precision, oversample) != KNN_OK)
    return 0;
  return p.total;
}

const char* knn_last_error(void) { return g_last_error.c_str(); }

int knn_abi_version(void) { return 1; }

}  // extern "C"
