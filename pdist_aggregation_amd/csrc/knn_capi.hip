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
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <mutex>
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

// Testing switches (PDIST_KNN_GUESS, PDIST_KNN_SPLITS, PDIST_KNN_WIDEN) are
// read only when PDIST_KNN_TESTING=1 is set as well: they force rare code
// paths (re-search, two-level merge, widened re-rank) at sizes a test can
// check, and are not a supported interface.  Without the opt-in a call never
// looks at them.
const char* test_switch(const char* name) {
  const char* on = std::getenv("PDIST_KNN_TESTING");
  if (!on || std::strcmp(on, "1") != 0) return nullptr;
  return std::getenv(name);
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
  int kc = 0, cap = 0, n_splits = 1, path = 0, dm = 0;
  int dpad = 0;  // float32 MFMA, d % 4 != 0: row width of the zero-padded copies
  // Two-level merge: with more than kMaxMerge / k' splits, groups of
  // group_size split lists are first reduced to n_groups lists per query.
  int group_size = 1, n_groups = 1;
  size_t off_pre_d = 0, off_pre_i = 0;
  // Large k: queries per chunk and the chunk's distance rows / selections.
  int64_t bigk_chunk = 0;
  size_t off_bigk_d = 0, off_bigk_sel_d = 0, off_bigk_sel_i = 0;
  int64_t rows_per_split = 0, n_qblocks = 0;
  int64_t slots = 0;  // resident search workgroups the device holds (MFMA, low-d)
  size_t elem = 4;
  // Sampled threshold (MFMA paths): a first launch over sample_regions row
  // ranges of sample_rows rows, sample_stride apart, estimates each query's
  // k'-th value as the sample_j-th smallest sampled minimum.
  bool guess = false;
  int guess_fail_stride = 0;  // > 0: force the guess of every stride-th query to fail
  int sample_regions = 0, sample_j = 0;
  int64_t sample_rows = 0, sample_stride = 0;
  double sample_fraction = 0.0, guess_fail_bound = 0.0;
  // Widening (MFMA paths): per-query collect lists of coll_cap row indices
  // for queries the error bound does not certify.
  int coll_cap = 0;
  // workspace offsets
  size_t off_scalars = 0, off_ynorm = 0, off_tau = 0, off_cand_d = 0, off_cand_i = 0,
         off_part_d = 0, off_part_i = 0, off_mu = 0, off_xq = 0, off_yq = 0, off_xnorm = 0,
         off_xerr = 0, off_xscale = 0, off_yscale = 0, off_slots = 0, off_sample = 0,
         off_guess = 0, off_failed = 0, off_failed_qb = 0, off_uncert = 0, off_xpad = 0,
         off_ypad = 0, off_widen = 0, off_widen_qb = 0, off_ubound = 0, off_coll_cnt = 0,
         off_coll_idx = 0, off_part_n = 0, off_order = 0, off_qmap = 0, total = 0;
};

constexpr int kPathGeneric = 0, kPathMfmaF32 = 1, kPathMfmaBF16 = 2, kPathMfmaFP8 = 3,
              kPathLowd = 4, kPathBigK = 5;
// Float32 with d % 4 != 0 takes the 32 x 32 MFMA kernel through zero-padded
// copies of the rows only while the train copy stays within this size; larger
// train sets use the 16 x 16 x 4 MFMA kernel on the caller's rows (no copy).
constexpr size_t kMaxPadBytes = (size_t)1 << 30;
// Large-k path: dense FP64 distance rows for at most this many bytes of
// queries at a time (and at most kBigKSelBytes of selected entries).
constexpr size_t kBigKRowBytes = (size_t)2 << 30, kBigKSelBytes = (size_t)512 << 20;

// Lists one merge workgroup takes (its register staging holds kMaxMerge
// entries), and the most splits the two-level merge can take.
int64_t merge_lists(int kc) { return std::max(1, knn::kMaxMerge / kc); }
int64_t merge_split_limit(int kc) { return merge_lists(kc) * merge_lists(kc); }

// Compute units of the current device (256 on an MI355X; also the fallback
// when no device is visible, so plans made without a GPU match one).
int device_cus() {
  int dev = 0, n = 0;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
    return n;
  return 256;
}

// Split count for a grid of n_qblocks query blocks: enough workgroups to keep
// every CU busy several times over, each split at least 8 row steps -- or 2
// when 8 would leave the CUs with fewer than two workgroups each (1e3 x 1e4 x
// 50 float64, 16 query blocks: 18 -> 53 splits, k=10 0.69 -> 0.52 ms, k=100
// 1.33 -> 0.83 ms, even with the two-level merge) --, within what the
// two-level merge takes, and (MFMA) a multiple of 8 so a split's query blocks
// share one XCD.
int choose_splits(int64_t n_qblocks, int64_t n_train, int row_step, int kc, int64_t target,
                  bool multiple_of_8) {
  int64_t splits = std::max<int64_t>(1, (target + n_qblocks / 2) / n_qblocks);
  const int64_t s8 = std::max<int64_t>(1, n_train / (8 * row_step));
  const bool thin = n_qblocks * s8 < 2 * (int64_t)device_cus();
  splits = std::min<int64_t>(splits, thin ? std::max<int64_t>(1, n_train / (2 * row_step)) : s8);
  splits = std::min<int64_t>(splits, merge_split_limit(kc));
  if (multiple_of_8 && splits >= 8) splits = splits / 8 * 8;
  return (int)std::max<int64_t>(splits, 1);
}

// Split count for the one-wave-of-workgroups kernels (MFMA, low-d): run time
// is proportional to rounds(s) / s with rounds(s) = ceil(n_qblocks * s /
// slots); take the smallest s within 4% of the best, since every split adds
// its own warm-up and k' candidates to merge.  Counts past merge_lists(k')
// need the two-level merge, which roughly doubles the merge: they are only
// considered when the one-level range is modelled more than 25% slower
// (few query blocks; 1e5 queries at k' = 128: 17 splits instead of 11 gave
// no faster search and 2.7 -> 6.2 ms of merge).  PDIST_KNN_SPLITS overrides.
int choose_splits_slots(int64_t n_qblocks, int64_t n_train, int row_step, int kc,
                        int64_t slots) {
  if (const char* env = test_switch("PDIST_KNN_SPLITS")) {
    const int v = std::atoi(env);
    if (v > 0) return (int)std::min<int64_t>(v, merge_split_limit(kc));
  }
  int64_t smax = std::min<int64_t>(n_train / (8 * row_step), merge_split_limit(kc));
  smax = std::max<int64_t>(1, std::min<int64_t>(smax, 512));
  auto cost = [&](int64_t s) {
    return (double)knn::ceil_div(n_qblocks * s, std::max<int64_t>(slots, 1)) / (double)s;
  };
  auto best_upto = [&](int64_t hi) {
    double b = cost(1);
    for (int64_t s = 2; s <= hi; ++s) b = std::min(b, cost(s));
    return b;
  };
  double best = best_upto(smax);
  const int64_t s1 = std::min<int64_t>(smax, merge_lists(kc));
  const double best1 = best_upto(s1);
  if (best1 <= 1.25 * best) {
    smax = s1;
    best = best1;
  }
  for (int64_t s = 1; s <= smax; ++s)
    if (cost(s) <= 1.04 * best) return (int)s;
  return 1;
}

// P(Binomial(n, f) >= j).
double binom_tail(int n, double f, int j) {
  if (j <= 0) return 1.0;
  if (j > n) return 0.0;
  double pmf = std::pow(1.0 - f, n), cdf = 0.0;  // P(X = 0)
  for (int i = 0; i < j; ++i) {
    cdf += pmf;
    pmf *= (double)(n - i) / (double)(i + 1) * f / (1.0 - f);
  }
  return std::max(0.0, 1.0 - cdf);
}

// Sampled-threshold plan.  The guess must leave `need` rows at or below it:
// k' where the merge needs k' candidates (generic and low-d kernels), k for
// the MFMA kernels, whose merge certifies a query with k <= n < k'
// candidates against the guess itself (every row above it was skipped, none
// at or below it).  The need-1 rows ranked above a query's need-th value
// each fall in a sample of fraction f with probability f (rows in random
// order), so the j-th smallest sampled value lies below the need-th value --
// the guess is too tight -- with probability P(Bin(need-1, f) >= j).  Two
// reductions only raise the guess, never lower it: the sample keeps minima
// of 16 rows, and each sampling lane keeps only its kSampleJ smallest
// minima (so j may exceed kSampleJ: the j-th smallest of the kept values is
// at least the j-th smallest of all).  j is the smallest with that bound at
// most kGuessFail, up to kGuessMaxJ rounds of the guess kernel; a query whose
// guess fails is searched again, so the bound only governs speed.  A guess
// about j / f rows deep lets that many rows through the search: at
// 1e4 x 1e6 x 128 (k = 10, k' = 40, f = 1/32) j = 6 against k instead of
// j = 11 against k' cuts them from about 350 to about 190 per query.
// PDIST_KNN_GUESS: "0" off, "1" on at any size, "fail:S" on with the guess of
// every S-th query forced to fail (exercises the re-search; "fail" = every
// query), "j:J" on with j = J (a tight guess exercises the certification of
// queries with fewer than k' candidates).
constexpr int64_t kGuessMinRows = 32768;
// Generic kernel: workgroups its search grid aims for, and the list capacity
// of its sampling launch (kSampleJ kept + one 64-row step of appends).
constexpr int64_t kGenericGridTarget = 2048;
constexpr int kGenericSampleCap = 128;
static_assert(kGenericSampleCap >= knn::kSampleJ + knn::kGenY &&
                  (kGenericSampleCap & (kGenericSampleCap - 1)) == 0,
              "sampling lists: a power of two holding kSampleJ plus one step");
constexpr int kSampleShift = 5;  // sample 1/32 of the train rows
constexpr double kGuessFail = 1e-7;
constexpr int kGuessMaxJ = 128;
// Collect lists: room for 8 k' rows per query (FP8: 16 k'), at least 1024
// and at most 16384.  A CPU model of the quantisation puts the
// bound-following set at about k + 2 rows for BF16 on the 1e4 x 1e6 x 128
// U[0,1) benchmark inputs, 100-350 rows for FP8 there, and about 2,000 (a
// whole Gaussian) for FP8 on the 1,000-Gaussian mixture in d = 256 with
// k' = 400 (capacity 8192).  FP8 with k = 32 (k' = 128) on 1e5 U[0,1)
// queries overflowed 1,024 rows for 7 queries (left uncertified).
constexpr int kCollMin = 1024, kCollMax = 16384;
// Collect launch geometry: per-lane list capacity (two halves per query) and
// the shortest split it is given.
constexpr int kCollectCap = 64;
// Collect-list capacity from which the exact re-rank of widened FP8 queries
// runs in cluster order.
constexpr int kOrderMinColl = 4096;
constexpr int64_t kCollectMinRows = 4096;
// Packed collect launches (a few listed query blocks): shortest split.
constexpr int64_t kPackedMinRows = 1024;

// runs: sampled runs per (query, region) the kernel writes (2 for the MFMA
// kernel's two lane halves, 1 for the low-d kernel's one lane per query).
void plan_guess(Plan& p, int64_t slots, int runs, int64_t need) {
  const char* env = test_switch("PDIST_KNN_GUESS");
  const std::string mode = env ? env : "";
  int fail_stride = 0, forced_j = 0;
  if (mode == "fail") fail_stride = 1;
  if (mode.rfind("fail:", 0) == 0) fail_stride = std::max(1, std::atoi(mode.c_str() + 5));
  if (mode.rfind("j:", 0) == 0) forced_j = std::max(1, std::atoi(mode.c_str() + 2));
  const bool forced = mode == "1" || fail_stride > 0 || forced_j > 0;
  // Off below kGuessMinRows train rows (the sampling launch and the guess
  // cost more than they save), unless a split holds fewer than 64 k' rows:
  // started from +inf, its lists would fill and be compacted over and over.
  const bool small_splits = p.rows_per_split < 64 * (int64_t)p.kc;
  if (mode == "0" || (!forced && p.n_train < kGuessMinRows && !small_splits)) return;
  const int64_t tiles = knn::ceil_div(p.n_train, (int64_t)knn::kMfmaRowStep);
  const int64_t want = std::max<int64_t>(1, knn::ceil_div(tiles, (int64_t)1 << kSampleShift));
  // Regions: enough sampling workgroups to fill the GPU, and enough of them
  // that the kept minima (runs x regions x kSampleJ per query) hold at least
  // 2j values: the first pass finds j at the fraction of the GPU-filling
  // layout, the second settles the layout.
  int64_t regions = std::max<int64_t>(1, slots / std::max<int64_t>(p.n_qblocks, 1));
  int j = 1;
  for (int pass = 0; pass < 2; ++pass) {
    if (pass == 1)
      regions = std::max<int64_t>(regions, knn::ceil_div(2 * (int64_t)j, runs * knn::kSampleJ));
    regions = std::min<int64_t>({regions, want, 64});
    const int64_t per = knn::ceil_div(want, regions);                // tiles per region
    const int64_t stride = std::max<int64_t>(per, tiles / regions);  // tiles between regions
    p.sample_regions = (int)regions;
    p.sample_rows = per * knn::kMfmaRowStep;
    p.sample_stride = stride * knn::kMfmaRowStep;
    p.sample_fraction = std::min(1.0, (double)(regions * per) / (double)tiles);
    j = 1;
    while (j <= kGuessMaxJ && binom_tail(need - 1, p.sample_fraction, j) > kGuessFail) ++j;
  }
  const int jmax = (int)std::min<int64_t>(kGuessMaxJ, runs * regions * knn::kSampleJ);
  if (forced_j > 0) j = forced_j;
  if (j > jmax) {
    if (!forced) return;
    j = jmax;
  }
  p.sample_j = j;
  p.guess_fail_bound = binom_tail(need - 1, p.sample_fraction, j);
  p.guess = true;
  p.guess_fail_stride = fail_stride;
}

int make_plan(Plan& p, size_t elem, int64_t n_train, int64_t n_test, int64_t dim, int64_t k,
              int precision, int64_t oversample, bool force_generic = false) {
  if (n_train <= 0 || n_test < 0 || dim <= 0)
    return fail(KNN_EINVAL, "n_train and dim must be positive and n_test non-negative");
  if (k <= 0) return fail(KNN_EINVAL, "k must be positive");
  if (k > n_train) return fail(KNN_EINVAL, "k must not exceed n_train");
  if (n_train > INT32_MAX) return fail(KNN_EUNSUPPORTED, "n_train must be below 2^31");
  if (dim > 65536) return fail(KNN_EUNSUPPORTED, "dim above 65536 is not supported");
  if (precision < KNN_PRECISION_EXACT || precision > KNN_PRECISION_FP8)
    return fail(KNN_EINVAL, "unknown precision");
  if (precision != KNN_PRECISION_EXACT && elem != 4)
    return fail(KNN_EINVAL, "low precision search needs float32 inputs");
  int64_t kc;
  if (oversample > 0) {
    if (oversample < k) return fail(KNN_EINVAL, "oversample (k') must be >= k");
    kc = oversample;
  } else if (precision == KNN_PRECISION_EXACT) {
    kc = k + std::max<int64_t>(8, k / 8);
  } else {
    kc = 4 * k;
  }
  kc = std::min<int64_t>(kc, n_train);
  if (k > knn::kMaxCandidates) {
    // More neighbours than one merge workgroup sorts: exact FP64 distance
    // rows of a chunk of queries, radix select and rank sort per query.
    p = Plan();
    p.n_train = n_train;
    p.n_test = n_test;
    p.dim = dim;
    p.k = k;
    p.elem = elem;
    p.kc = (int)k;
    p.path = kPathBigK;
    const int64_t nq1 = std::max<int64_t>(n_test, 1);
    int64_t chunk = std::max<int64_t>(1, (int64_t)(kBigKRowBytes / ((size_t)n_train * 8)));
    chunk = std::min<int64_t>(chunk, (int64_t)(kBigKSelBytes / ((size_t)k * 12)));
    p.bigk_chunk = std::max<int64_t>(1, std::min<int64_t>(chunk, nq1));
    p.n_splits = (int)knn::ceil_div(nq1, p.bigk_chunk);
    size_t o = 0;
    auto take = [&](size_t bytes) {
      const size_t at = o;
      o = align_up(o + bytes);
      return at;
    };
    p.off_scalars = take(sizeof(knn::DeviceScalars));
    p.off_bigk_d = take((size_t)p.bigk_chunk * n_train * sizeof(double));
    p.off_bigk_sel_d = take((size_t)p.bigk_chunk * k * sizeof(double));
    p.off_bigk_sel_i = take((size_t)p.bigk_chunk * k * sizeof(int));
    p.total = o;
    return KNN_OK;
  }
  if (kc > knn::kMaxCandidates) kc = knn::kMaxCandidates;
  p = Plan();
  p.n_train = n_train;
  p.n_test = n_test;
  p.dim = dim;
  p.k = k;
  p.elem = elem;
  p.kc = (int)kc;

  const int mfma_cap = knn::next_pow2((int)kc + knn::kMfmaCapSlack);
  const bool mfma_cap_ok = mfma_cap <= knn::kMfmaMaxCap;
  const int lowd_cap = knn::next_pow2((int)kc + knn::kLowdRowStep);
  p.path = kPathGeneric;
  if (precision != KNN_PRECISION_EXACT && (dim > 256 || !mfma_cap_ok)) {
    // Beyond the bf16/fp8 kernels (d <= 256, k' <= kMfmaMaxCap - kMfmaCapSlack):
    // the exact search, whose answer the low-precision one only reproduces.
    return make_plan(p, elem, n_train, n_test, dim, k, KNN_PRECISION_EXACT,
                     oversample > knn::kMaxCandidates ? 0 : oversample, force_generic);
  }
  if (precision != KNN_PRECISION_EXACT) {
    p.path = precision == KNN_PRECISION_BF16 ? kPathMfmaBF16 : kPathMfmaFP8;
    p.dm = std::max(p.path == kPathMfmaBF16 ? 32 : 64, knn::next_pow2((int)dim));
  } else if (!force_generic && elem == 4 && dim >= 16 && dim <= 256 && mfma_cap_ok &&
             (dim % 4 == 0 || (size_t)n_train * ((dim + 3) & ~(int64_t)3) * 4 <= kMaxPadBytes)) {
    p.path = kPathMfmaF32;
    // 192 is a search width too (48 chunks of 16 bytes per row): d in
    // (128, 192] would otherwise pad to 256 and spend up to half the MFMAs on zeros.
    p.dm = dim <= 128 ? std::max(32, knn::next_pow2((int)dim)) : dim <= 192 ? 192 : 256;
    // d % 4 != 0: the search reads zero-padded copies with 16-byte aligned
    // rows (the exact re-rank still reads the caller's rows).
    if (dim % 4 != 0) p.dpad = (int)((dim + 3) & ~(int64_t)3);
  } else if (dim < 16 && lowd_cap <= knn::kMfmaMaxCap) {
    p.path = kPathLowd;
  }

  int bq, step;
  if (p.path == kPathGeneric) {
    bq = knn::kGenQ;
    step = knn::kGenY;
    p.cap = knn::next_pow2((int)kc + knn::kGenY);
  } else if (p.path == kPathLowd) {
    bq = knn::kLowdQueriesPerBlock;
    step = knn::kLowdRowStep;
    // Twice the minimum where it fits: a lane's list then absorbs a split's
    // appends with a compaction (global read, sort, write) far more rarely
    // (65,536^2 x 3, k = 16: search 3.7 -> 2.1 ms, workspace 261 -> 421 MiB).
    p.cap = 2 * lowd_cap <= knn::kMfmaMaxCap ? 2 * lowd_cap : lowd_cap;
  } else {
    const int fmt = p.path == kPathMfmaF32    ? knn::kFmtF32
                    : p.path == kPathMfmaBF16 ? knn::kFmtBF16
                                              : knn::kFmtFP8;
    bq = knn::mfma_block_queries(fmt, p.dm);
    step = knn::kMfmaRowStep;
    p.cap = mfma_cap;
  }
  p.n_qblocks = knn::ceil_div(std::max<int64_t>(n_test, 1), bq);
  int splits;
  int64_t slots = 0;
  if (p.path == kPathGeneric) {
    splits = choose_splits(p.n_qblocks, n_train, step, p.kc, kGenericGridTarget, false);
  } else {
    int per_cu = 2;
    if (p.path == kPathLowd)
      per_cu = elem == 4 ? knn::lowd_blocks_per_cu<float>((int)dim, p.cap)
                         : knn::lowd_blocks_per_cu<double>((int)dim, p.cap);
    else
      per_cu = knn::mfma_blocks_per_cu(p.path == kPathMfmaF32    ? knn::kFmtF32
                                       : p.path == kPathMfmaBF16 ? knn::kFmtBF16
                                                                 : knn::kFmtFP8,
                                       p.dm, p.cap);
    slots = (int64_t)device_cus() * per_cu;
    p.slots = slots;
    splits = choose_splits_slots(p.n_qblocks, n_train, step, p.kc, slots);
    // The MFMA kernel addresses one split with 32-bit buffer offsets.
    if (p.path != kPathLowd) {
      const int64_t row_bytes = p.path == kPathMfmaF32 ? (p.dpad ? p.dpad : dim) * 4
                                : p.path == kPathMfmaBF16 ? (int64_t)p.dm * 2
                                                          : (int64_t)p.dm;
      const int64_t min_splits = knn::ceil_div(n_train * row_bytes, (int64_t)1 << 30);
      splits = (int)std::max<int64_t>(splits, min_splits);
    }
  }
  p.rows_per_split = knn::ceil_div(knn::ceil_div(n_train, splits), step) * step;
  p.n_splits = (int)knn::ceil_div(n_train, p.rows_per_split);
  if (p.n_splits > merge_split_limit(p.kc))
    return fail(KNN_EUNSUPPORTED, "k' = " + std::to_string(p.kc) + " with " +
                                      std::to_string(p.n_splits) +
                                      " train splits exceeds the two-level merge");
  if (p.n_splits > merge_lists(p.kc)) {
    p.group_size = (int)knn::ceil_div((int64_t)p.n_splits, merge_lists(p.kc));
    p.n_groups = (int)knn::ceil_div((int64_t)p.n_splits, (int64_t)p.group_size);
  }
  const bool mfma = p.path == kPathMfmaF32 || p.path == kPathMfmaBF16 || p.path == kPathMfmaFP8;
  if (mfma) {
    // FP8: deep enough for the split lists to hold the bound-following sets
    // (about 16 k rows at e4m3 precision: 100-350 at 1e4 x 1e6 x 128 U[0,1)
    // with k = 10, a whole Gaussian of 2,000 rows for the 1,000-Gaussian
    // mixture with k = 100), so the widened queries skip the collect sweep:
    // 4 k' rows, j = 21 and 91 there (2.77 -> 2.05 ms; 65.6 -> about 55 ms).
    // A guess that falls short of the sets costs more than it saves (mixture,
    // j = 48: 3 % of the sets in the lists, 71 ms).  Other formats widen few
    // queries, by about k + 2 rows: planned for k.
    if (p.path == kPathMfmaFP8) plan_guess(p, slots, 2, 4 * (int64_t)p.kc);
    if (!p.guess) plan_guess(p, slots, 2, p.k);
  }
  if (p.path == kPathLowd) plan_guess(p, slots, 1, p.kc);
  // Generic kernel: one run per (query, region), from as many sampling
  // workgroups as the search's own grid target.  Only with 4 or more splits:
  // the guess saves each split's first fill and compaction, and with fewer
  // the sampling launch costs more than those (float64 1e5 queries x 1e3
  // rows x 100, k = 10, one split: 2.6 -> 2.9 ms with it).
  if (p.path == kPathGeneric && p.n_splits >= 4) plan_guess(p, kGenericGridTarget, 1, p.kc);

  const size_t celem = (p.path == kPathGeneric || p.path == kPathLowd) ? elem : sizeof(float);
  const int64_t nq1 = std::max<int64_t>(n_test, 1);
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o = align_up(o + bytes);
    return at;
  };
  p.off_scalars = take(sizeof(knn::DeviceScalars));
  p.off_ynorm = take((size_t)(n_train + knn::kNormPad) * celem);
  p.off_tau = take((size_t)nq1 * sizeof(unsigned long long));
  p.off_uncert = take((size_t)nq1);
  size_t ncand = (size_t)p.n_qblocks * p.n_splits * bq * p.cap;
  if (p.guess && p.path == kPathGeneric)  // the sampling launch's lists share this scratch
    ncand = std::max(ncand, (size_t)p.n_qblocks * p.sample_regions * bq * kGenericSampleCap);
  if (mfma) {  // (value, row) pairs; the collect launch reuses the space for indices
    p.off_cand_d = take(ncand * 2 * sizeof(float));
    p.off_cand_i = p.off_cand_d;
  } else {
    p.off_cand_d = take(ncand * celem);
    p.off_cand_i = take(ncand * sizeof(int));
  }
  const size_t npart = (size_t)nq1 * p.n_splits * p.kc;
  p.off_part_d = take(npart * celem);
  p.off_part_i = take(npart * sizeof(int));
  if (mfma) p.off_part_n = take((size_t)nq1 * p.n_splits * sizeof(int));
  if (p.group_size > 1) {
    const size_t npre = (size_t)nq1 * p.n_groups * p.kc;
    p.off_pre_d = take(npre * celem);
    p.off_pre_i = take(npre * sizeof(int));
  }
  if (mfma && !p.guess) p.off_slots = take((size_t)nq1 * knn::align8(p.n_splits) * sizeof(float));
  if (p.guess) {
    p.off_sample = take((size_t)nq1 * p.sample_regions * 2 * knn::kSampleJ * sizeof(float));
    p.off_guess = take((size_t)nq1 * sizeof(float));
    p.off_failed = take((size_t)nq1 * sizeof(int));
    p.off_failed_qb = take((size_t)p.n_qblocks * sizeof(int));
  }
  if (p.dpad) {
    p.off_xpad = take((size_t)nq1 * p.dpad * sizeof(float));
    p.off_ypad = take((size_t)n_train * p.dpad * sizeof(float));
  }
  if (mfma) {
    p.coll_cap = std::min(kCollMax, knn::next_pow2(std::max(
                                        kCollMin, (p.path == kPathMfmaFP8 ? 16 : 8) * p.kc)));
    p.off_widen = take((size_t)nq1 * sizeof(int));
    // one flag per query block, then the merge's sink entry
    p.off_widen_qb = take(((size_t)p.n_qblocks + 1) * sizeof(int));
    p.off_ubound = take((size_t)nq1 * sizeof(double));
    p.off_coll_cnt = take((size_t)nq1 * sizeof(int));
    p.off_coll_idx = take((size_t)nq1 * p.coll_cap * sizeof(int));
    p.off_qmap = take(((size_t)nq1 + 1) * sizeof(int));  // widened queries, then their count
    // FP8 with large collect lists (widening is the rule there, and sets of
    // thousands of rows overlap between queries of one cluster): the re-rank's
    // query order, its keys and bucket counts.  On the FP8 mixture
    // (5e4 x 2e6 x 256, k' = 400) the re-rank takes 13.4 instead of 15.1 ms
    // for 0.4 ms of ordering; at 1e4 x 1e6 x 128, k' = 40, it gains nothing.
    if (p.path == kPathMfmaFP8 && p.coll_cap >= kOrderMinColl)
      p.off_order = take((2 * (size_t)nq1 + knn::kOrderBuckets + 1) * sizeof(int));
  }
  if (p.path == kPathMfmaBF16 || p.path == kPathMfmaFP8) {
    const size_t qe = p.path == kPathMfmaBF16 ? 2 : 1;
    p.off_mu = take((1 + (size_t)knn::column_mean_blocks(std::min<int64_t>(n_train, 65536))) *
                    p.dm * sizeof(float));
    p.off_xq = take((size_t)nq1 * p.dm * qe);
    p.off_yq = take((size_t)n_train * p.dm * qe);
    p.off_xnorm = take((size_t)nq1 * sizeof(float));
    p.off_xerr = take((size_t)nq1 * sizeof(float));
    if (p.path == kPathMfmaFP8) {
      p.off_xscale = take((size_t)nq1);
      p.off_yscale = take((size_t)n_train);
    }
  }
  p.total = o;
  return KNN_OK;
}

// Workspace that covers the chosen plan and the generic fallback it may take
// at run time (misaligned float32 pointers).
size_t plan_workspace(size_t elem, int64_t n_train, int64_t n_test, int64_t dim, int64_t k,
                      int precision, int64_t oversample) {
  Plan p;
  if (make_plan(p, elem, n_train, n_test, dim, k, precision, oversample) != KNN_OK) return 0;
  size_t total = p.total;
  if (p.path == kPathMfmaF32) {
    Plan g;
    if (make_plan(g, elem, n_train, n_test, dim, k, precision, oversample, true) == KNN_OK)
      total = std::max(total, g.total);
  }
  return total;
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
  if (p.path == kPathMfmaF32 && !p.dpad && ((reinterpret_cast<uintptr_t>(X_train) |
                                  reinterpret_cast<uintptr_t>(X_test)) & 15) != 0) {
    rc = make_plan(p, sizeof(T), n_train, n_test, dim, k, precision, oversample, true);
    if (rc != KNN_OK) return rc;
  }

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
  // Launch check, stats read-back, workspace release: the end of every path.
  auto finish = [&]() -> int {
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
      stats->n_rescued = host_sc.n_rescued;
      stats->n_widened = host_sc.n_widened;
      stats->n_widened_from_lists = host_sc.n_widened_lists;
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
  };

  if (p.path == kPathBigK) {
    knn::launch_fill_keys<float>(reinterpret_cast<unsigned long long*>(sc),
                                 sizeof(knn::DeviceScalars) / 8, 0ull, stream);
    timer.mark(1);
    double* D = reinterpret_cast<double*>(ws + p.off_bigk_d);
    double* sel_d = reinterpret_cast<double*>(ws + p.off_bigk_sel_d);
    int* sel_i = reinterpret_cast<int*>(ws + p.off_bigk_sel_i);
    for (int64_t q0 = 0; q0 < n_test; q0 += p.bigk_chunk) {
      const int64_t nq = std::min<int64_t>(p.bigk_chunk, n_test - q0);
      knn::launch_sqdist_f64<T>(X_test + q0 * dim, nq, X_train, n_train, (int)dim, D, stream);
      knn::launch_bigk_select_sort<T>(D, nq, n_train, (int)k, (flags & KNN_FLAG_SQUARED) ? 1 : 0,
                                      sel_d, sel_i, dist_out + q0 * k, idx_out + q0 * k, stream);
    }
    timer.mark(2);
    timer.mark(3);
    return finish();
  }

  // 1. fills, gathered into one launch before the path's first kernel:
  // scalars are 4 ordered-key slots (max keys start at key(-inf)).
  knn::FillBatch fills;
  fills.add64(sc, sizeof(knn::DeviceScalars) / 8, 0ull);
  fills.add64(tau, n_test, knn::ordered_key(__builtin_huge_val()));
  knn::MergeArgs<T> ma;
  ma.unsorted = 0;
  ma.stage_cap = 0;
  ma.part_n = nullptr;
  ma.pre_d = nullptr;
  ma.pre_i = nullptr;
  ma.group_size = 1;
  ma.n_groups = 1;
  ma.failed = nullptr;
  ma.failed_qblocks = nullptr;
  ma.guess_bound = 0;
  ma.qblock_queries = 1;
  ma.tau = tau;
  ma.only_queries = nullptr;
  ma.widen = nullptr;
  ma.widen_qblocks = nullptr;
  ma.ubound = nullptr;
  ma.widen_all = 0;
  ma.lists_complete = 0;
  knn::MfmaArgs fa;
  ma.xnorm = nullptr;
  ma.xerr = nullptr;
  ma.unit_roundoff = sizeof(T) == 4 ? 0x1p-24 : 0x1p-53;
  ma.acc_terms = 2 * (int)dim + 4;
  knn::SearchArgs<T> sa{};  // generic / low-d search (value-initialised: a plain search)
  if (p.path == kPathGeneric || p.path == kPathLowd) {
    // 2. norms  3. search
    knn::launch_fill_batch(fills, stream);
    knn::launch_row_norms<T>(X_train, n_train, (int)dim, ynorm, sc, stream);
    sa.X = X_test;
    sa.Y = X_train;
    sa.ynorm = ynorm;
    sa.n_q = n_test;
    sa.n_t = n_train;
    sa.d = (int)dim;
    sa.kc = p.kc;
    sa.cap = p.cap;
    sa.n_splits = p.n_splits;
    sa.n_qblocks = (int)p.n_qblocks;
    sa.rows_per_split = p.rows_per_split;
    sa.cand_d = reinterpret_cast<T*>(ws + p.off_cand_d);
    sa.cand_i = reinterpret_cast<int*>(ws + p.off_cand_i);
    sa.tau = tau;
    sa.part_d = reinterpret_cast<T*>(ws + p.off_part_d);
    sa.part_i = reinterpret_cast<int*>(ws + p.off_part_i);
    // 2b. sampled threshold: the kSampleJ smallest over 1/32 of the rows
    // (generic; low-d: of 32-row minima; float64 values rounded up) -> guess -> tau, as for the
    // MFMA kernels (one run per query and region).
    if (p.guess && p.path == kPathGeneric) {
      auto* failed = reinterpret_cast<int*>(ws + p.off_failed);
      auto* failed_qb = reinterpret_cast<int*>(ws + p.off_failed_qb);
      knn::SearchArgs<T> sm = sa;
      sm.kc = knn::kSampleJ;
      sm.cap = kGenericSampleCap;
      sm.n_splits = p.sample_regions;
      sm.rows_per_split = p.sample_rows;
      sm.split_stride = p.sample_stride;
      sm.sample_out = reinterpret_cast<float*>(ws + p.off_sample);
      knn::launch_generic_sample<T>(sm, stream);
      ma.failed = failed;
      ma.failed_qblocks = failed_qb;
      ma.qblock_queries = knn::kGenQ;
      knn::launch_sample_guess(sm.sample_out, n_test, p.sample_regions, p.sample_j, tau,
                               reinterpret_cast<float*>(ws + p.off_guess), p.guess_fail_stride,
                               failed, failed_qb, ma.qblock_queries, stream);
    }
    if (p.guess && p.path == kPathLowd) {
      auto* failed = reinterpret_cast<int*>(ws + p.off_failed);
      auto* failed_qb = reinterpret_cast<int*>(ws + p.off_failed_qb);
      knn::SearchArgs<T> sm = sa;
      sm.n_splits = p.sample_regions;
      sm.rows_per_split = p.sample_rows;
      sm.split_stride = p.sample_stride;
      sm.sample_out = reinterpret_cast<float*>(ws + p.off_sample);
      knn::launch_lowd_sample<T>(sm, stream);
      ma.failed = failed;
      ma.failed_qblocks = failed_qb;
      ma.qblock_queries = knn::kThreads;
      knn::launch_sample_guess(sm.sample_out, n_test, p.sample_regions, p.sample_j, tau,
                               reinterpret_cast<float*>(ws + p.off_guess), p.guess_fail_stride,
                               failed, failed_qb, ma.qblock_queries, stream);
    }
    timer.mark(1);
    if (p.path == kPathLowd)
      knn::launch_lowd_search<T>(sa, stream);
    else
      knn::launch_generic_search<T>(sa, stream);
    timer.mark(2);
    ma.part_d = sa.part_d;
    ma.part_i = sa.part_i;
  } else if constexpr (sizeof(T) == 4) {
    fa.xscale = nullptr;
    fa.yscale = nullptr;
    fa.dm = p.dm;
    fa.n_q = n_test;
    fa.n_t = n_train;
    fa.kc = p.kc;
    fa.cap = p.cap;
    fa.n_qblocks = (int)p.n_qblocks;
    fa.n_splits = p.n_splits;
    fa.rows_per_split = p.rows_per_split;
    fa.split_stride = p.rows_per_split;
    fa.qblock_active = nullptr;
    fa.query_active = nullptr;
    fa.qmap = nullptr;
    fa.qmap_n = nullptr;
    fa.coll_group = 1;
    fa.pk_items = 0;
    fa.pk_splits = 0;
    fa.pk_slots = 0;
    fa.pk_grid = 0;
    fa.sample_out = nullptr;
    fa.coll_cnt = nullptr;
    fa.coll_idx = nullptr;
    fa.coll_cap = 0;
    // Widening flags start clear; the merges set them.
    ma.widen = reinterpret_cast<int*>(ws + p.off_widen);
    ma.widen_qblocks = reinterpret_cast<int*>(ws + p.off_widen_qb);
    ma.ubound = reinterpret_cast<double*>(ws + p.off_ubound);
    ma.qblock_queries = knn::mfma_block_queries(
        p.path == kPathMfmaF32 ? knn::kFmtF32 : p.path == kPathMfmaBF16 ? knn::kFmtBF16 : knn::kFmtFP8,
        p.dm);
    // "all" / "every:S" widen every (S-th) query; with "sweep" in the value
    // every widened query takes the collect launch, none its split lists.
    bool sweep_only = false;
    if (const char* w = test_switch("PDIST_KNN_WIDEN")) {
      if (std::strncmp(w, "all", 3) == 0) ma.widen_all = 1;
      if (std::strncmp(w, "every:", 6) == 0) ma.widen_all = std::max(1, std::atoi(w + 6));
      sweep_only = std::strstr(w, "sweep") != nullptr;
    }
    // The widened queries' packed list for the collect launch, its count
    // after the n_test entries.
    fills.add32(ma.widen, n_test, 0u);
    fills.add32(ma.widen_qblocks, p.n_qblocks, 0u);
    fills.add32(reinterpret_cast<int*>(ws + p.off_qmap) + n_test, 1, 0u);
    fills.add32(ws + p.off_coll_cnt, n_test, 0u);
    // |y|^2 = +inf past the last train row (the kernels' partial last tile)
    fills.add32(reinterpret_cast<float*>(ynorm) + n_train, knn::kNormPad, 0x7f800000u);
    fa.cand_d = reinterpret_cast<float*>(ws + p.off_cand_d);
    fa.cand_i = reinterpret_cast<int*>(ws + p.off_cand_i);
    fa.tau = tau;
    fa.part_d = reinterpret_cast<float*>(ws + p.off_part_d);
    fa.part_i = reinterpret_cast<int*>(ws + p.off_part_i);
    fa.part_n = reinterpret_cast<int*>(ws + p.off_part_n);
    fa.ynorm = reinterpret_cast<float*>(ynorm);
    // Shared-threshold slots (+inf = nothing published), used when at least
    // kSlotJ splits exist and no sampled threshold is in use: with one, lists
    // rarely reach k' entries, so slots are seldom published, and scanning
    // them (n_splits loads per query every 32 steps) costs more than it saves.
    fa.qslots = nullptr;
    fa.slot_rank = (int)knn::ceil_div(p.kc, knn::kSlotJ) - 1;
    fa.slot_stride = knn::align8(p.n_splits);
    if (p.n_splits >= knn::kSlotJ && !p.guess) {
      fa.qslots = reinterpret_cast<float*>(ws + p.off_slots);
      fills.add32(fa.qslots, n_test * fa.slot_stride, 0x7f800000u);  // +inf
    }
    // A slot bound filters below the shared threshold, so rows under the
    // latter may be missing from the lists; without slots none are.
    ma.lists_complete = fa.qslots == nullptr && !sweep_only ? 1 : 0;
    knn::launch_fill_batch(fills, stream);
    if (p.path == kPathMfmaF32) {
      // 2. norms
      knn::launch_row_norms<float>(reinterpret_cast<const float*>(X_train), n_train, (int)dim,
                                   reinterpret_cast<float*>(ynorm), sc, stream);
      fa.fmt = knn::kFmtF32;
      if (p.dpad) {
        float* xp = reinterpret_cast<float*>(ws + p.off_xpad);
        float* yp = reinterpret_cast<float*>(ws + p.off_ypad);
        knn::launch_pad_rows(reinterpret_cast<const float*>(X_train), n_train, (int)dim, p.dpad,
                             yp, stream);
        knn::launch_pad_rows(reinterpret_cast<const float*>(X_test), n_test, (int)dim, p.dpad, xp,
                             stream);
        fa.X = xp;
        fa.Y = yp;
        fa.ldx = p.dpad;
        fa.ldy = p.dpad;
      } else {
        fa.X = X_test;
        fa.Y = X_train;
        fa.ldx = dim;
        fa.ldy = dim;
      }
      fa.d = (int)dim;
    } else {
      // 2. centre + bf16 rounding of both sets, norms and error bounds
      float* mu = reinterpret_cast<float*>(ws + p.off_mu);
      float* xnorm = reinterpret_cast<float*>(ws + p.off_xnorm);
      float* xerr = reinterpret_cast<float*>(ws + p.off_xerr);
      const float* Ytr = reinterpret_cast<const float*>(X_train);
      const float* Xte = reinterpret_cast<const float*>(X_test);
      knn::launch_column_mean(Ytr, std::min<int64_t>(n_train, 65536), (int)dim, mu + p.dm, mu,
                              stream);
      if (p.path == kPathMfmaBF16) {
        knn::launch_quantize_bf16(Ytr, n_train, (int)dim, p.dm, mu, ws + p.off_yq,
                                  reinterpret_cast<float*>(ynorm), nullptr, sc, true, stream);
        knn::launch_quantize_bf16(Xte, n_test, (int)dim, p.dm, mu, ws + p.off_xq, xnorm, xerr,
                                  sc, false, stream);
        fa.fmt = knn::kFmtBF16;
      } else {
        auto* xs = reinterpret_cast<unsigned char*>(ws + p.off_xscale);
        auto* ys = reinterpret_cast<unsigned char*>(ws + p.off_yscale);
        knn::launch_quantize_fp8(Ytr, n_train, (int)dim, p.dm, mu, ws + p.off_yq, ys,
                                 reinterpret_cast<float*>(ynorm), nullptr, sc, true, stream);
        knn::launch_quantize_fp8(Xte, n_test, (int)dim, p.dm, mu, ws + p.off_xq, xs, xnorm, xerr,
                                 sc, false, stream);
        fa.fmt = knn::kFmtFP8;
        fa.xscale = xs;
        fa.yscale = ys;
      }
      fa.X = ws + p.off_xq;
      fa.Y = ws + p.off_yq;
      fa.ldx = p.dm;
      fa.ldy = p.dm;
      fa.d = p.dm;
      ma.xnorm = xnorm;
      ma.xerr = xerr;
      ma.unit_roundoff = 0x1p-24;
      ma.acc_terms = 2 * p.dm + 4;
    }
    // 2b. sampled threshold: minima over 1/32 of the rows -> guess -> tau
    if (p.guess) {
      auto* failed = reinterpret_cast<int*>(ws + p.off_failed);
      auto* failed_qb = reinterpret_cast<int*>(ws + p.off_failed_qb);
      knn::MfmaArgs sa = fa;
      sa.n_splits = p.sample_regions;
      sa.rows_per_split = p.sample_rows;
      sa.split_stride = p.sample_stride;
      sa.qslots = nullptr;
      sa.sample_out = reinterpret_cast<float*>(ws + p.off_sample);
      if (!knn::launch_mfma_sample(sa, stream)) {
        if (owned) (void)hipFreeAsync(ws, stream);
        return fail(KNN_EUNSUPPORTED, "no MFMA kernel for this feature count");
      }
      ma.failed = failed;
      ma.failed_qblocks = failed_qb;
      ma.guess_bound = 1;
      ma.qblock_queries = knn::mfma_block_queries(fa.fmt, p.dm);
      knn::launch_sample_guess(sa.sample_out, n_test, 2 * p.sample_regions, p.sample_j, tau,
                               reinterpret_cast<float*>(ws + p.off_guess), p.guess_fail_stride,
                               failed, failed_qb, ma.qblock_queries, stream);
    }
    // 3. search
    timer.mark(1);
    if (!knn::launch_mfma_search(fa, stream)) {
      if (owned) (void)hipFreeAsync(ws, stream);
      return fail(KNN_EUNSUPPORTED, "no MFMA kernel for this feature count");
    }
    timer.mark(2);
    ma.part_d = reinterpret_cast<const T*>(fa.part_d);
    ma.part_i = fa.part_i;
    ma.unsorted = 1;  // the MFMA search copies short split lists unsorted
    // Staging for 2 k' entries (at least 256) instead of every list entry:
    // 1e4 x 1e6 x 128 BF16, 19 lists x k' = 40: 10 KiB -> 4 KiB of LDS per
    // merge workgroup.  Widening covers the queries whose ties it drops.
    ma.stage_cap = knn::next_pow2(std::max(256, 2 * p.kc));
    ma.part_n = fa.part_n;  // and writes only their entries
  }
  // 4. merge + exact re-rank
  ma.X = X_test;
  ma.Y = X_train;
  ma.n_q = n_test;
  ma.n_t = n_train;
  ma.d = (int)dim;
  ma.n_splits = p.n_splits;
  ma.kc = p.kc;
  ma.k = (int)k;
  ma.squared = (flags & KNN_FLAG_SQUARED) ? 1 : 0;
  ma.sc = sc;
  ma.uncert = reinterpret_cast<unsigned char*>(ws + p.off_uncert);
  ma.dist_out = dist_out;
  ma.idx_out = idx_out;
  // With more splits than one merge workgroup takes, groups of split lists
  // are first reduced to one list each (pre-merge), then merged.
  auto merge = [&](const knn::MergeArgs<T>& m0) {
    knn::MergeArgs<T> m = m0;
    if (p.group_size > 1) {
      knn::MergeArgs<T> pm = m0;
      pm.pre_d = reinterpret_cast<T*>(ws + p.off_pre_d);
      pm.pre_i = reinterpret_cast<int*>(ws + p.off_pre_i);
      pm.group_size = p.group_size;
      pm.n_groups = p.n_groups;
      pm.failed = nullptr;
      knn::launch_merge_rerank<T>(pm, stream);
      m.part_d = pm.pre_d;
      m.part_i = pm.pre_i;
      m.n_splits = p.n_groups;
      m.unsorted = 0;  // the pre-merge writes its group lists sorted, all k' slots
      m.part_n = nullptr;
    }
    knn::launch_merge_rerank<T>(m, stream);
  };
  merge(ma);
  // 5. queries whose sampled guess proved too tight: search again (only
  // their query blocks, only them), from tau = +inf, and merge them.
  if (ma.failed && (p.path == kPathGeneric || p.path == kPathLowd)) {
    knn::SearchArgs<T> ra = sa;
    ra.qblock_active = ma.failed_qblocks;
    ra.query_active = ma.failed;
    if (p.path == kPathLowd)
      knn::launch_lowd_research<T>(ra, stream);
    else
      knn::launch_generic_research<T>(ra, stream);
    knn::MergeArgs<T> rm = ma;
    rm.failed = nullptr;
    rm.only_queries = ma.failed;
    merge(rm);
  }
  if constexpr (sizeof(T) == 4) {
    if (ma.failed && p.path != kPathGeneric && p.path != kPathLowd) {
      knn::MfmaArgs ra = fa;
      ra.qblock_active = ma.failed_qblocks;
      ra.query_active = ma.failed;
      knn::launch_mfma_research(ra, stream);
      knn::MergeArgs<T> rm = ma;
      rm.failed = nullptr;
      rm.only_queries = ma.failed;
      merge(rm);
    }
    // 6. queries the error bound does not certify: gather every row that can
    // lie within their k-th exact distance (collect launch at the fixed
    // threshold the merge set), then re-rank all of them exactly.
    if (ma.widen) {
      knn::MfmaArgs ca = fa;
      ca.qblock_active = ma.widen_qblocks;
      ca.query_active = ma.widen;
      ca.coll_cnt = reinterpret_cast<int*>(ws + p.off_coll_cnt);
      ca.coll_idx = reinterpret_cast<int*>(ws + p.off_coll_idx);
      ca.coll_cap = p.coll_cap;
      // The collect launch keeps no per-split lists, so it takes shorter
      // splits (its latency is one workgroup's sweep: with a few widened
      // queries, that sweep is all it costs): lane lists of kCollectCap
      // indices, flushed when half full, in the search's candidate scratch
      // (values and indices, idle by now), as many splits as that holds, up
      // to splits of kCollectMinRows rows.
      ca.cap = kCollectCap;
      ca.cand_i = reinterpret_cast<int*>(ws + p.off_cand_d);
      ca.cand_d = reinterpret_cast<float*>(ca.cand_i);  // collect lists hold indices only
      {
        const int bq = knn::mfma_block_queries(fa.fmt, p.dm);
        const int64_t room = (int64_t)((p.off_part_d - p.off_cand_d) / sizeof(int));
        const int64_t fit = room / ((int64_t)p.n_qblocks * bq * kCollectCap);
        const int64_t want = knn::ceil_div(n_train, kCollectMinRows);
        const int64_t splits_c = std::max<int64_t>(p.n_splits, std::min(fit, want));
        ca.rows_per_split = knn::ceil_div(knn::ceil_div(n_train, splits_c), knn::kMfmaRowStep) *
                            knn::kMfmaRowStep;
        ca.split_stride = ca.rows_per_split;
        ca.n_splits = (int)knn::ceil_div(n_train, ca.rows_per_split);
        // Packed (fewer than half the queries listed, so at most pb_max
        // blocks): the listed blocks share the whole scratch and one round
        // of workgroups, down to kPackedMinRows-row splits (FP8 at
        // 1e4 x 1e6 x 128: 977 instead of 50 splits for the few queries the
        // lists miss; FP8 on 1,000 Gaussians, d = 256: 768, where 1,954
        // splits of 1,024 rows took 2.5 rounds and 0.24 instead of 0.19 ms).
        ca.pk_items = (int)std::min<int64_t>(room / ((int64_t)bq * kCollectCap), INT_MAX / 2);
        ca.pk_splits = (int)std::min<int64_t>(knn::ceil_div(n_train, kPackedMinRows), INT_MAX / 2);
        ca.pk_slots = (int)std::min<int64_t>(std::max<int64_t>(p.slots, 1), INT_MAX / 2);
        const int64_t pb_max = knn::ceil_div(knn::ceil_div(n_test, 2), (int64_t)bq);
        ca.pk_grid = (int)std::min<int64_t>(
            ca.pk_items, pb_max * std::max<int64_t>(ca.pk_splits, ca.n_splits));
        // Unpacked (most queries widened), the search's own split count when
        // its workgroups fit one round (FP8 at 1e4 x 1e6 x 128: collect
        // 1.28 -> 1.09 ms); with several rounds the short splits balance the
        // rounds better (FP8 on 1,000 Gaussians, 3.8 rounds: 11 ms slower
        // grouped).
        if (p.n_qblocks * p.n_splits <= p.slots)
          ca.coll_group = (int)std::max<int64_t>(1, ca.n_splits / std::max(1, p.n_splits));
      }
      // The widened queries packed into the first query blocks (one atomic
      // per wave): a few of them take one block's sweep instead of one per
      // block they fall in (BF16, 100 Gaussians, 1e5 x 1e6 x 128: 54 widened
      // queries in about 50 blocks).
      int* qmap = reinterpret_cast<int*>(ws + p.off_qmap);
      knn::launch_pack_flagged(ma.widen, n_test, 1, qmap, qmap + n_test, stream);
      // Widened queries whose split lists hold the whole set (flag 2: the
      // collect threshold lies below the search's final one) skip the sweep.
      if (ma.lists_complete)
        knn::launch_list_collect(ma.widen, tau, fa.part_d, fa.part_i, fa.part_n, n_test,
                                 p.n_splits, p.kc, n_train, ca.coll_cnt, ca.coll_idx, p.coll_cap,
                                 stream);
      ca.qmap = qmap;
      ca.qmap_n = qmap + n_test;
      knn::launch_mfma_collect(ca, stream);
      knn::CollectArgs<T> cr;
      cr.X = X_test;
      cr.Y = X_train;
      cr.n_q = n_test;
      cr.n_t = n_train;
      cr.d = (int)dim;
      cr.k = (int)k;
      cr.squared = ma.squared;
      cr.widen = ma.widen;
      cr.coll_cnt = ca.coll_cnt;
      cr.coll_idx = ca.coll_idx;
      cr.coll_cap = p.coll_cap;
      cr.ubound = ma.ubound;
      cr.surv_cap = std::min(4096, knn::next_pow2(std::max(256, 2 * (int)k + 64)));
      cr.uncert = ma.uncert;
      cr.dist_out = dist_out;
      cr.idx_out = idx_out;
      cr.order = nullptr;
      if (p.off_order) {
        // Queries whose collected sets overlap (FP8: whole clusters) re-ranked
        // side by side on one XCD, which then serves the shared rows from L2.
        int* key = reinterpret_cast<int*>(ws + p.off_order);
        int* order = key + n_test;
        knn::launch_collect_order(ma.widen, ca.coll_cnt, ca.coll_idx, p.coll_cap, n_test,
                                  n_train, key, order + n_test, order, stream);
        cr.order = order;
      }
      knn::launch_collect_rerank<T>(cr, stream);
      if (stats) {
        knn::launch_count_i32(ma.widen, n_test, 0, &sc->n_widened, stream);
        knn::launch_count_i32(ma.widen, n_test, 2, &sc->n_widened_lists, stream);
      }
    }
  }
  // Every query's final merge (the first one, or the re-search's for a query
  // whose guess failed) has set its flag.
  if (stats) knn::launch_count_uncertified(ma.uncert, n_test, sc, stream);
  timer.mark(3);

  return finish();
}

template <class T>
int run_backward(const T* X_train, int64_t n_train, const T* X_test, int64_t n_test,
                 int64_t dim, int64_t k, const int64_t* indices, const T* grad_dist, int flags,
                 T* grad_X_train, T* grad_X_test, void* stream_ptr) {
  g_last_error.clear();
  if (n_train <= 0 || n_test < 0 || dim <= 0 || k <= 0)
    return fail(KNN_EINVAL, "n_train, dim and k must be positive and n_test non-negative");
  if (dim > INT32_MAX || k > INT32_MAX) return fail(KNN_EUNSUPPORTED, "dim or k too large");
  if (!X_train || (n_test > 0 && (!X_test || !indices || !grad_dist)))
    return fail(KNN_EINVAL, "null array pointer");
  if (!grad_X_train && !grad_X_test) return fail(KNN_EINVAL, "no gradient requested");
  knn::BackwardArgs<T> a;
  a.X = X_test;
  a.Y = X_train;
  a.n_q = n_test;
  a.n_t = n_train;
  a.d = (int)dim;
  a.k = (int)k;
  a.idx = indices;
  a.grad = grad_dist;
  a.squared = (flags & KNN_FLAG_SQUARED) ? 1 : 0;
  a.grad_X = grad_X_test;
  a.grad_Y = grad_X_train;
  knn::launch_backward<T>(a, reinterpret_cast<hipStream_t>(stream_ptr));
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(KNN_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return KNN_OK;
}

template <class T>
int run_sqdist(const T* X_train, int64_t n_train, const T* X_test, int64_t n_test, int64_t dim,
               T* out, void* stream_ptr) {
  g_last_error.clear();
  if (n_train < 0 || n_test < 0 || dim <= 0)
    return fail(KNN_EINVAL, "dim must be positive and the row counts non-negative");
  if (dim > INT32_MAX) return fail(KNN_EUNSUPPORTED, "dim too large");
  if (n_train > 0 && n_test > 0 && (!X_train || !X_test || !out))
    return fail(KNN_EINVAL, "null array pointer");
  knn::launch_sqdist<T>(X_test, n_test, X_train, n_train, (int)dim, out,
                        reinterpret_cast<hipStream_t>(stream_ptr));
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(KNN_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return KNN_OK;
}

// Scratch of knn_argkmin_f32_noalloc: a file-scope device array, so that
// entry point allocates nothing.  Every GPU the library runs that entry point
// on holds its own copy until the process ends.  Its device address is
// looked up once per device (hipGetSymbolAddress, with that device current)
// and only that address is ever handed to kernels or copies: in host code
// the array's name is a host-side stand-in, not a device address.
constexpr size_t kNoAllocScratchBytes = KNN_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's scratch address (nullptr and a message on failure).
void* device_scratch(std::string& err) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    err = std::string("hipGetDevice: ") + hipGetErrorString(e);
    return nullptr;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    err = "device ordinal " + std::to_string(dev) + " beyond the scratch table";
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return nullptr;
    }
    g_scratch_ptr[dev] = ptr;
  }
  return g_scratch_ptr[dev];
}

}  // namespace

extern "C" {

int knn_argkmin_f32_noalloc(const float* X_train, int64_t n_train, const float* X_test,
                            int64_t n_test, int64_t dim, int64_t k, int64_t chunk_size,
                            int strategy, int precision, int64_t oversample, int flags,
                            int n_threads_hint, float* dist_out, int64_t* idx_out, void* stream,
                            knn_stats* stats) {
  g_last_error.clear();
  if (flags & KNN_FLAG_TIMING)
    return fail(KNN_EINVAL, "KNN_FLAG_TIMING creates HIP events: not available in the "
                            "allocation-free entry point");
  const size_t need =
      plan_workspace(sizeof(float), n_train, n_test, dim, k, precision, oversample);
  if (need == 0) {  // an invalid request: let the shared path name the reason
    Plan p;
    const int rc = make_plan(p, sizeof(float), n_train, n_test, dim, k, precision, oversample);
    return rc != KNN_OK ? rc : fail(KNN_EINVAL, "invalid request");
  }
  if (need > kNoAllocScratchBytes)
    return fail(KNN_EWORKSPACE, "this call needs " + std::to_string(need) +
                                    " bytes of scratch; knn_argkmin_f32_noalloc has " +
                                    std::to_string(kNoAllocScratchBytes) +
                                    " (use knn_argkmin_f32 with a workspace)");
  std::string err;
  void* ws = device_scratch(err);
  if (!ws) return fail(KNN_EHIP, err);
  return run<float>(X_train, n_train, X_test, n_test, dim, k, chunk_size, strategy, precision,
                    oversample, flags, n_threads_hint, dist_out, idx_out, ws,
                    kNoAllocScratchBytes, stream, stats);
}

int knn_sqdist_matrix_f32(const float* X_train, int64_t n_train, const float* X_test,
                          int64_t n_test, int64_t dim, float* out, void* stream) {
  return run_sqdist<float>(X_train, n_train, X_test, n_test, dim, out, stream);
}

int knn_sqdist_matrix_f64(const double* X_train, int64_t n_train, const double* X_test,
                          int64_t n_test, int64_t dim, double* out, void* stream) {
  return run_sqdist<double>(X_train, n_train, X_test, n_test, dim, out, stream);
}

int knn_argkmin_backward_f32(const float* X_train, int64_t n_train, const float* X_test,
                             int64_t n_test, int64_t dim, int64_t k, const int64_t* indices,
                             const float* grad_dist, int flags, float* grad_X_train,
                             float* grad_X_test, void* stream) {
  return run_backward<float>(X_train, n_train, X_test, n_test, dim, k, indices, grad_dist, flags,
                             grad_X_train, grad_X_test, stream);
}

int knn_argkmin_backward_f64(const double* X_train, int64_t n_train, const double* X_test,
                             int64_t n_test, int64_t dim, int64_t k, const int64_t* indices,
                             const double* grad_dist, int flags, double* grad_X_train,
                             double* grad_X_test, void* stream) {
  return run_backward<double>(X_train, n_train, X_test, n_test, dim, k, indices, grad_dist,
                              flags, grad_X_train, grad_X_test, stream);
}

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
  return plan_workspace((size_t)dtype_bytes, n_train, n_test, dim, k, precision, oversample);
}

const char* knn_last_error(void) { return g_last_error.c_str(); }

int knn_abi_version(void) { return 4; }

}  // extern "C"
