// Internal launch interface between the scheduler (knn_capi.hip) and the
// kernels.  Every pointer is a device pointer; every launch is asynchronous on
// `stream`.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdlib>
#include <stdint.h>

namespace knn {

// Geometry of the generic (SIMT FMA) fused distance + candidate kernel.
constexpr int kGenQ = 64;  // queries per workgroup
constexpr int kGenY = 64;  // train rows per step
constexpr int kGenK = 16;  // feature slice staged in LDS per step

// Largest candidate list (k') a query may keep, and the largest merged
// (splits x k') list the merge kernel sorts in LDS.
constexpr int kMaxCandidates = 1984;
constexpr int kMaxMerge = 2048;

// Per-call scalars the kernels share (lives in the workspace).
struct DeviceScalars {
  unsigned long long max_ynorm_key;  // ordered key of max_j |y_j|^2 (as used by the search)
  unsigned long long max_yerr_key;   // ordered key of max_j |y_j - q(y_j)| (low precision)
  unsigned int n_uncertified;
  unsigned int n_rescued;            // queries searched again after a too-tight sampled threshold
  unsigned int n_widened;            // queries re-ranked over their bound-following row set
  unsigned int n_widened_lists;      // ... of which the set came from the search's lists
};

// Everything the fused search kernels need.  T is the element type of the
// search arithmetic inputs (float or double); candidate values are T.
template <class T>
struct SearchArgs {
  const T* X;        // queries, n_q x d row-major
  const T* Y;        // train rows, n_t x d row-major
  const T* ynorm;    // |y_j|^2, n_t
  int64_t n_q, n_t;
  int d;
  int kc;                  // candidates kept per query (k')
  int cap;                 // per-query capacity of the candidate scratch (power of two)
  int n_splits;
  int n_qblocks;           // query blocks (low-d kernel grid)
  int64_t rows_per_split;  // multiple of the kernel's row step
  T* cand_d;               // [blocks][queries per block][cap]
  int* cand_i;
  unsigned long long* tau;  // [n_q] ordered key of the best known k'-th value
  T* part_d;                // [n_q][n_splits][kc] sorted candidate lists
  int* part_i;
  // Low-d and generic kernels (zero / nullptr: a plain search).  split_stride: rows
  // between the starts of consecutive splits (0: rows_per_split).
  // sample_out: the sampling launch's per-query runs, [n_q][n_splits][kSampleJ]
  // smallest 32-row minima, ascending.  qblock_active / query_active: the
  // re-search launch's query blocks and queries.
  int64_t split_stride;
  float* sample_out;
  const int* qblock_active;
  const int* query_active;
};

// ---------------------------------------------------------------------------
// MFMA search (knn_mfma.hip)
// ---------------------------------------------------------------------------
constexpr int kFmtF32 = 0;   // v_mfma_f32_32x32x2_f32 on the caller's rows
constexpr int kFmtBF16 = 1;  // v_mfma_f32_32x32x16_bf16 on quantised rows
constexpr int kFmtFP8 = 2;   // v_mfma_scale_f32_32x32x64_f8f6f4, E4M3 + row scales
constexpr int kMfmaRowStep = 32;
// +inf entries after the n_t row norms the MFMA kernels read: rows of the
// last (partial) tile past n_t get |y|^2 = +inf and are never selected.
constexpr int kNormPad = 128;
constexpr int kMfmaMaxCap = 512;  // per-query list capacity (wave sort in LDS)
// A query's list is two halves of cap/2, one per owning lane; each half must
// hold ceil(k'/2) survivors plus one step's 16 appends: cap >= k' + 34.
constexpr int kMfmaCapSlack = 34;

struct MfmaArgs {
  int fmt;
  int dm;                // padded feature count the kernel is compiled for
  const void* X;         // queries (float or bf16), row stride ldx elements
  const void* Y;         // train rows (float or bf16), row stride ldy elements
  const float* ynorm;    // |y|^2 as used by the search
  const unsigned char* xscale;  // FP8: E8M0 row scales of X and Y
  const unsigned char* yscale;
  int64_t ldx, ldy;
  int64_t n_q, n_t;
  int d;                 // valid features (<= dm)
  int kc, cap;
  int n_qblocks, n_splits;
  int64_t rows_per_split;  // multiple of kMfmaRowStep
  float* cand_d;
  int* cand_i;
  unsigned long long* tau;
  float* part_d;
  int* part_i;
  // part_n[q * n_splits + s]: entries of (query q, split s) written to the
  // lists above (the rest of the k' slots are left unwritten).
  int* part_n;
  // Cross-split threshold sharing: qslots[q * slot_stride + split] (init
  // +inf; slot_stride = n_splits rounded up to 8) receives the split's
  // (slot_rank+1)-th smallest value at each list compaction, with
  // kSlotJ * (slot_rank+1) >= k'; nullptr disables it.
  float* qslots;
  int slot_rank;
  int slot_stride;
  // Split s covers rows [s * split_stride, s * split_stride + rows_per_split)
  // (split_stride == rows_per_split for the search; larger for sampling).
  int64_t split_stride;
  // Re-search and collect launches only: workgroups of query blocks whose
  // qblock_active entry is 0 exit at once, queries whose query_active entry
  // is 0 (collect launch: other than 1) neither collect candidates nor write
  // lists.
  const int* qblock_active;
  const int* query_active;
  // Sampling launch: [n_q][n_splits][2][kSampleJ] ascending smallest
  // per-tile minima seen by each lane.
  float* sample_out;
  // Collect launch (bound-following re-rank set): for the query blocks and
  // queries selected by qblock_active / query_active, the index of every
  // train row whose value is <= the query's fixed threshold tau[q] goes to
  // coll_idx[q * coll_cap + slot], slots taken from coll_cnt[q] (zeroed
  // before the launch); the count keeps growing past coll_cap, the entries
  // past it are dropped.
  int* coll_cnt;
  int* coll_idx;
  int coll_cap;
  // Packed query blocks, read by the collect launch only (nullptr: query
  // block qb holds queries qb * BQ + i).  While 2 * *qmap_n < n_q, query
  // block qb holds queries qmap[qb * BQ + i] for qb * BQ + i < *qmap_n and
  // blocks past *qmap_n exit at once; with more, the launch runs unpacked
  // (qblock_active and query_active must be set either way).  The collect
  // launch packs the widened queries this way, so a few of them cost one
  // block's sweep, not one per query block they fall in.
  const int* qmap;
  const int* qmap_n;
  // Collect launch, unpacked: workgroups take coll_group consecutive splits
  // each (1: one split); split_stride must equal rows_per_split then.
  int coll_group;
  // Collect launch, packed: the P = ceil(*qmap_n / BQ) blocks each sweep
  // S = max(min(pk_items / P, pk_slots / P, pk_splits), n_splits) splits of
  // equal whole row steps (pk_items: the block lists the candidate scratch
  // holds, at least n_qblocks * n_splits; pk_slots: workgroups resident at
  // once), from a grid of at least pk_grid workgroups (pk_grid >= P * S for
  // every packed P).
  int pk_items, pk_splits, pk_slots, pk_grid;
};

constexpr int kSlotJ = 4;
constexpr int kSampleJ = 16;  // per-lane minima kept by the sampling launch

// Returns false when no instantiation matches (fmt, dm).
bool launch_mfma_search(const MfmaArgs& a, hipStream_t s);
// The re-search launch (qblock_active and query_active must be set).
bool launch_mfma_research(const MfmaArgs& a, hipStream_t s);
// The collect launch (qblock_active, query_active and the coll_* fields set).
bool launch_mfma_collect(const MfmaArgs& a, hipStream_t s);
// The sampling launch of the same kernel: rows [s * split_stride, + rows_per_split)
// of each split s, writing sample_out; no candidate lists.
bool launch_mfma_sample(const MfmaArgs& a, hipStream_t s);
// guess[q] = the j-th smallest of the sampled minima of query q (ascending
// runs of kSampleJ, n_runs per query); tau[q] <- min(tau[q], key(guess[q])).
// fail_stride > 0 forces guess -inf on queries q % fail_stride == 0, which
// then fail the check and are searched again (exercises that path).  Also
// clears the failure flags the merge sets: failed[q] and, from the first
// query of each block of qblock_queries, failed_qblocks[q / qblock_queries].
void launch_sample_guess(const float* sample_out, int64_t n_q, int n_runs, int j,
                         unsigned long long* tau, float* guess, int fail_stride, int* failed,
                         int* failed_qblocks, int qblock_queries, hipStream_t s);
void launch_fill_u32(unsigned int* p, int64_t n, unsigned int v, hipStream_t s);

// Several fills in one launch (each kernel boundary on a stream costs a few
// microseconds of idle GPU): segment i sets n[i] elements of 4 or 8 bytes
// (wide[i]) at p[i] to v[i].
constexpr int kMaxFills = 10;
struct FillBatch {
  void* p[kMaxFills];
  int64_t n[kMaxFills];
  unsigned long long v[kMaxFills];
  int wide[kMaxFills];
  int count = 0;
  void add32(void* at, int64_t len, unsigned val) { add(at, len, val, 0); }
  void add64(void* at, int64_t len, unsigned long long val) { add(at, len, val, 1); }
  void add(void* at, int64_t len, unsigned long long val, int w) {
    if (len <= 0) return;
    if (count == kMaxFills) std::abort();  // the call sites set at most 8
    p[count] = at;
    n[count] = len;
    v[count] = val;
    wide[count] = w;
    ++count;
  }
};
// Launches the batch's fills (none: no launch) and empties it.
void launch_fill_batch(FillBatch& b, hipStream_t s);
// Queries per workgroup of the (fmt, dm) instantiation (4 waves x 32 x groups).
int mfma_block_queries(int fmt, int dm);
// Resident workgroups per CU of that instantiation (2 if it cannot be queried).
int mfma_blocks_per_cu(int fmt, int dm, int cap);

// ---------------------------------------------------------------------------
// Low-precision preparation (knn_quant.hip)
// ---------------------------------------------------------------------------
// mu[f] = mean of feature f over the first n_sample rows of Y (f < d), in a
// fixed summation order; part holds column_mean_blocks(n_sample) * d floats.
void launch_column_mean(const float* Y, int64_t n_sample, int d, float* part, float* mu,
                        hipStream_t s);
int64_t column_mean_blocks(int64_t n_sample);
// dst[r] = bf16(src[r] - mu) zero-padded to dm; norm[r] = |dst[r]|^2 (fp32);
// err[r] >= |src[r] - mu - dst[r]| (nullable).  With track_max the maxima of
// norm and err go to sc (max_ynorm_key, max_yerr_key).
void launch_quantize_bf16(const float* src, int64_t n, int d, int dm, const float* mu,
                          void* dst, float* norm, float* err, DeviceScalars* sc, bool track_max,
                          hipStream_t s);

// OCP E4M3 with a power-of-two scale per row (E8M0 byte in scale[r]).
void launch_quantize_fp8(const float* src, int64_t n, int d, int dm, const float* mu, void* dst,
                         unsigned char* scale, float* norm, float* err, DeviceScalars* sc,
                         bool track_max, hipStream_t s);

// Rows of d floats copied to rows of dp floats (dp >= d), zero-filled.
void launch_pad_rows(const float* src, int64_t n, int d, int dp, float* dst, hipStream_t s);

template <class T>
void launch_fill_keys(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t s);

template <class T>
void launch_row_norms(const T* Y, int64_t n, int d, T* out, DeviceScalars* sc, hipStream_t s);

template <class T>
void launch_generic_search(const SearchArgs<T>& a, hipStream_t s);
// Sampling launch (kc == kSampleJ; rows [s * split_stride, + rows_per_split)
// of each of the n_splits regions, writing sample_out) and re-search launch
// (qblock_active / query_active, 64 queries per block) of the generic kernel.
template <class T>
bool launch_generic_sample(const SearchArgs<T>& a, hipStream_t s);
template <class T>
bool launch_generic_research(const SearchArgs<T>& a, hipStream_t s);

// Low-dimensional (d < 16) VALU kernel (knn_lowd.hip): 256 queries per
// workgroup, 32 rows per step.  Returns false for d >= 16.
constexpr int kLowdQueriesPerBlock = 256;
constexpr int kLowdRowStep = 32;
template <class T>
using LowdArgs = SearchArgs<T>;
template <class T>
bool launch_lowd_search(const LowdArgs<T>& a, hipStream_t s);
// Sampling launch (rows [s * split_stride, + rows_per_split) of each of the
// n_splits regions, writing sample_out as float, double minima rounded up)
// and re-search launch (qblock_active / query_active) of the same kernel.
template <class T>
bool launch_lowd_sample(const LowdArgs<T>& a, hipStream_t s);
template <class T>
bool launch_lowd_research(const LowdArgs<T>& a, hipStream_t s);
template <class T>
int lowd_blocks_per_cu(int d, int cap);

// Merge the per-split candidate lists, re-rank them with exact FP64
// direct-difference distances from the original rows, keep k.
template <class T>
struct MergeArgs {
  const T* X;  // original queries (n_q x d)
  const T* Y;  // original train rows (n_t x d)
  int64_t n_q, n_t;
  int d;
  const T* part_d;
  const int* part_i;
  int n_splits, kc, k;
  // 1: the lists at part_d / part_i are not sorted (the MFMA search copies a
  // split's list as it stands unless it outgrew k'); every entry at or below
  // the query's threshold is then ranked.  0: each list is ascending.
  int unsorted;
  // Unsorted float lists only (0: off): LDS staging for at most stage_cap of
  // the entries at or below the k'-th value.  Only ties at it can overflow
  // that (fewer than k' entries lie below it); then the ties of lowest row
  // index are staged, so the staged entries still hold the k' smallest
  // (value, index) keys.  The smaller arrays let more merge workgroups share
  // a CU.
  int stage_cap;
  // Valid entries per (query, list), [n_q][n_splits] (nullptr: all k' slots
  // are written, the unused ones with +inf / -1).
  const int* part_n;
  int squared;  // output squared distances instead of Euclidean
  // Certification: the approximate values were computed on vectors xq, yq
  // with |x - xq| <= xerr[i] (nullptr: 0), |y - yq| <= max_yerr, rows norms
  // |xq|^2 = xnorm[i] (nullptr: computed from X) and |yq|^2 <= max_ynorm,
  // with arithmetic of unit roundoff `unit_roundoff` over `acc_terms` terms.
  const float* xnorm;
  const float* xerr;
  double unit_roundoff;
  int acc_terms;
  DeviceScalars* sc;
  // Per query, set by its final merge: 1 if not certified, else 0 (counted
  // into sc->n_uncertified afterwards by launch_count_uncertified; one
  // atomic per query on a single address costs ~40 ns each across the XCDs).
  unsigned char* uncert;
  T* dist_out;        // n_q x k
  int64_t* idx_out;   // n_q x k
  // Sampled-threshold check (nullptr: off).  The search started from a
  // guessed threshold; a query with fewer than min(k', n_t) merged candidates
  // may have lost rows to it: it gets failed[q] = 1, its query block
  // (q / qblock_queries) failed_qblocks = 1 and tau[q] = key(+inf), and its
  // outputs are left to the re-search.
  int* failed;
  int* failed_qblocks;
  // With failed set: 1 if the guess was planned for k rather than k' rows
  // (MFMA kernels).  A query then fails only below min(k, n_t) candidates;
  // one with k <= n < k' candidates lost no row at or below its threshold
  // tau[q] (the guess: with fewer than k' rows under it no list reached k'
  // entries, so no compaction or shared bound lowered it) and every row it
  // lost lies above it, so the certification takes that threshold in place
  // of the k'-th value.
  int guess_bound;
  int qblock_queries;
  unsigned long long* tau;
  // Re-search merge: only queries with only_queries[q] != 0 (nullptr: all).
  const int* only_queries;
  // Pre-merge (pre_d != nullptr): workgroup (q, g) reduces the lists of
  // splits [g * group_size, + group_size) of query q (n_splits in all) to its
  // k' best in ascending (value, index) order, written to
  // pre_d / pre_i[(q * n_groups + g) * kc + j] (padding +inf / -1); no exact
  // re-rank, no failure check.  A second (final) merge then reads n_groups
  // lists per query, which lifts the n_splits * k' <= kMaxMerge limit.
  T* pre_d;
  int* pre_i;
  int group_size, n_groups;
  // Widening (widen == nullptr: off).  A query its final merge does not
  // certify gets ubound[q] = its k-th exact squared distance U and tau[q] =
  // the key of T, the largest value the search can compute for a row within
  // distance sqrt(U) of the query, and then either
  //  * widen[q] = 2 when lists_complete and T lies below the query's final
  //    shared threshold: the split lists already hold every row at or below
  //    T, and launch_list_collect gathers them from there; or
  //  * widen[q] = 1 and widen_qblocks[q / qblock_queries] = 1: a collect
  //    launch sweeps the train rows for them.
  // widen_qblocks holds ceil(n_q / qblock_queries) + 1 entries: the last is
  // a sink the first kind of query stores into.
  // lists_complete: the search kept every row below the final tau[q] in its
  // split lists (no split slots: the sweep filtered with shared thresholds
  // only, each of which tau[q] ends at or below, and every list cut lowered
  // tau[q] to the k'-th value it kept).
  // widen_all (testing): S > 0 widens every S-th query (q % S == 0) that is
  // not trivially certified.
  int* widen;
  int* widen_qblocks;
  double* ubound;
  int widen_all;
  int lists_complete;
};

template <class T>
void launch_merge_rerank(const MergeArgs<T>& a, hipStream_t s);

// Exact re-rank of the collected sets (knn_merge.hip).  For every query with
// widen[q] != 0 whose coll_cnt[q] <= coll_cap: FP64 direct-difference
// distances of all its collected rows, those within ubound[q] (plus a
// relative 1e-12) kept -- at most surv_cap, a power of two -- sorted by
// (distance, index), the first k written to the outputs and uncert[q]
// cleared.  A query whose set overflowed keeps its first answer and flag.
template <class T>
struct CollectArgs {
  const T* X;
  const T* Y;
  int64_t n_q, n_t;
  int d, k;
  int squared;
  const int* widen;
  const int* coll_cnt;
  const int* coll_idx;
  int coll_cap;
  const double* ubound;
  int surv_cap;
  unsigned char* uncert;
  T* dist_out;
  int64_t* idx_out;
  // Order of the queries (nullptr: 0 .. n_q-1): workgroup b takes query
  // order[p], p = b's position in an XCD-major numbering (the XCD takes a
  // contiguous range of positions), so that queries next to each other in
  // `order` run together on one XCD and share its L2.
  const int* order;
};
template <class T>
void launch_collect_rerank(const CollectArgs<T>& a, hipStream_t s);
// order[0, n_q): the queries bucketed by the smallest row index each widened
// query collected (queries whose sets overlap mostly share it), widened ones
// first.  key, counts: scratch of n_q and kOrderBuckets + 1 ints (the last
// holds the largest key in use).
constexpr int kOrderBuckets = 16384;
void launch_collect_order(const int* widen, const int* coll_cnt, const int* coll_idx,
                          int coll_cap, int64_t n_q, int64_t n_t, int* key, int* counts,
                          int* order, hipStream_t s);
// *out += number of entries of f[0, n) that are nonzero (value 0) or equal
// to value (otherwise).
void launch_count_i32(const int* f, int64_t n, int value, unsigned int* out, hipStream_t s);
// list[*count ...] += the q < n with flags[q] == value, in no particular
// order (*count zeroed beforehand).
void launch_pack_flagged(const int* flags, int64_t n, int value, int* list, int* count,
                         hipStream_t s);
// Collect lists from the search's split lists, for every query with
// widen[q] == 2: coll_idx[q * coll_cap ...] = the rows of its n_splits lists
// (part_d / part_i[(q * n_splits + s) * kc + j], j < part_n[q * n_splits + s])
// whose value is not above the float threshold of key tau[q], and
// coll_cnt[q] = their number (written past coll_cap only as a count, like
// the collect launch's overflow).
void launch_list_collect(const int* widen, const unsigned long long* tau, const float* part_d,
                         const int* part_i, const int* part_n, int64_t n_q, int n_splits, int kc,
                         int64_t n_t, int* coll_cnt, int* coll_idx, int coll_cap, hipStream_t s);
// sc->n_uncertified += number of nonzero flags in uncert[0, n).
void launch_count_uncertified(const unsigned char* uncert, int64_t n, DeviceScalars* sc,
                              hipStream_t s);

// Backward of the k-NN distances (knn_backward.hip).
template <class T>
struct BackwardArgs {
  const T* X;  // queries (n_q x d)
  const T* Y;  // train rows (n_t x d)
  int64_t n_q, n_t;
  int d, k;
  const int64_t* idx;  // forward indices (n_q x k); entries outside [0, n_t) are skipped
  const T* grad;       // dL/d(dist) (n_q x k)
  int squared;         // distances were squared (else Euclidean)
  T* grad_X;           // n_q x d, or nullptr
  T* grad_Y;           // n_t x d, or nullptr (zeroed, then scattered with atomics)
};

template <class T>
void launch_backward(const BackwardArgs<T>& a, hipStream_t s);

// Dense direct-difference squared distances, out is n_q x n_t (knn_dense.hip).
template <class T>
void launch_sqdist(const T* X, int64_t n_q, const T* Y, int64_t n_t, int d, T* out,
                   hipStream_t s);

// The same in FP64 (differences and sums in double whatever T is).
template <class T>
void launch_sqdist_f64(const T* X, int64_t n_q, const T* Y, int64_t n_t, int d, double* out,
                       hipStream_t s);

// Large k (knn_bigk.hip): for the n_q queries of one chunk, D holds their
// exact squared distances to all n_t train rows (row-major).  Per query the k
// smallest (distance, index) keys are found by a radix select (ties at the
// k-th distance resolved by the lower index), then ordered by a rank sort
// into dist_out / idx_out rows q0 .. q0 + n_q - 1 (sqrt unless squared).
// sel_d / sel_i: n_q * k scratch.
template <class T>
void launch_bigk_select_sort(const double* D, int64_t n_q, int64_t n_t, int k, int squared,
                             double* sel_d, int* sel_i, T* dist_out, int64_t* idx_out,
                             hipStream_t s);

size_t merge_lds_bytes(int n_splits, int kc, size_t elem, int stage_cap);

}  // namespace knn
