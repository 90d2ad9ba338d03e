/*
 * pdist_knn.h - C ABI of the MI355X (gfx950) brute-force k-NN argkmin library
 * (libpdist_knn.so, built from pdist_aggregation_amd/csrc/).
 *
 * Conventions
 *  - X_train is the n_train x dim train set, X_test the n_test x dim query
 *    set, both row-major, C-contiguous.  The train set comes first, as in the
 *    reference's parallel_knn(X_train, X_test, k, ...).
 *  - Every array pointer is a DEVICE pointer (hipMalloc'd or a torch tensor's
 *    data_ptr); outputs are caller-allocated, row-major (n_test x k).
 *  - Calls are asynchronous on `stream` (NULL = the default stream) unless a
 *    knn_stats pointer is passed, in which case the call synchronises the
 *    stream to fill it.  Calls on different streams are independent.
 *  - Return value: 0 on success, a negative KNN_E* code otherwise;
 *    knn_last_error() gives the message of the last failure on this thread.
 *  - Result order: ascending exact distance, the lower train index first on
 *    equal distances.
 *  - Environment variables PDIST_KNN_GUESS, PDIST_KNN_SPLITS and
 *    PDIST_KNN_WIDEN are unsupported testing switches: the library reads them
 *    only when PDIST_KNN_TESTING=1 is also set.
 */
#ifndef PDIST_KNN_H
#define PDIST_KNN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KNN_OK 0
#define KNN_EINVAL (-1)       /* bad argument (shape, k, strategy, precision) */
#define KNN_EHIP (-2)         /* HIP runtime error */
#define KNN_EWORKSPACE (-3)   /* caller workspace smaller than required */
#define KNN_EUNSUPPORTED (-4) /* a valid request this build does not handle */

/* strategy: accepted for drop-in compatibility with the reference; the GPU
 * schedule does not depend on it.  It selects the axis n_parallel_chunks is
 * counted on (src/_pdist_aggregation.pyx:511-515, :330, :448). */
#define KNN_STRATEGY_AUTO 0
#define KNN_STRATEGY_CHUNK_ON_TRAIN 1
#define KNN_STRATEGY_CHUNK_ON_TEST 2

/* precision of the candidate search; the returned distances and the final
 * order always come from an exact FP64 re-rank of the candidates. */
#define KNN_PRECISION_EXACT 0 /* search in the input type (fp32 or fp64) */
#define KNN_PRECISION_BF16 1  /* bf16 MFMA candidates, fp32 accumulation */
#define KNN_PRECISION_FP8 2   /* fp8 e4m3 MFMA candidates, per-row scales */

/* flags */
#define KNN_FLAG_SQUARED 1 /* return squared distances (SYCL backend semantics) */
#define KNN_FLAG_TIMING 2  /* with stats: time each stage with HIP events on `stream` */

typedef struct knn_stats {
  int64_t n_parallel_chunks; /* the reference's closed form for the strategy */
  int64_t n_uncertified;     /* queries whose exactness could not be proven */
  int64_t k_candidates;      /* k' kept per query before the exact re-rank */
  int64_t n_splits;          /* train splits searched in parallel */
  int64_t strategy_used;     /* KNN_STRATEGY_CHUNK_ON_TRAIN or _ON_TEST */
  int64_t kernel_path;       /* 0 16x16x4 MFMA (float64, float32 d > 256), 1 MFMA fp32,
                                2 MFMA bf16, 3 MFMA fp8, 4 low-d SIMT, 5 large k (k above
                                1,984: dense FP64 rows, radix select, rank sort) */
  double prep_ms;            /* KNN_FLAG_TIMING: fills, norms, quantisation, sampling */
  double search_ms;          /* KNN_FLAG_TIMING: fused distance + candidate kernel */
  double merge_ms;           /* KNN_FLAG_TIMING: merge + exact re-rank (+ re-search) */
  int64_t n_rescued;         /* queries searched a second time because the threshold
                                guessed from a row sample was below their k'-th value */
  int64_t n_widened;         /* queries (MFMA paths) the error bound did not certify from
                                their k' candidates, re-ranked exactly over every train
                                row that can lie within their k-th exact distance; such
                                a query counts in n_uncertified only if that row set
                                overflowed its capacity */
  int64_t n_widened_from_lists; /* ... of n_widened, the queries whose row set the
                                   search's split lists already held (no second sweep
                                   of the train rows) */
} knn_stats;

/*
 * k nearest train rows of every query row.
 * Replaces the reference's Cython parallel_knn
 * (src/_pdist_aggregation.pyx:492-540, fused over float/double) and the SYCL
 * C shim p_knn_search<floatT, intT> (sycl_ext/parallel_knn_helper.hpp:7-23).
 *
 * oversample: k' candidates kept before the exact re-rank (0 = default:
 * k + max(8, k/8) for EXACT, 4k for BF16/FP8).  Any k <= n_train is accepted;
 * BF16/FP8 requests beyond those kernels (dim > 256 or k' > 478) run the
 * exact search.  n_threads_hint: the host
 * thread count the reference's "auto" rule would use (<= 0: this machine's).
 * workspace: device scratch of knn_argkmin_workspace_size() bytes, or NULL to
 * let the call allocate (stream-ordered) and free it.
 */
int knn_argkmin_f32(const float* X_train, int64_t n_train, const float* X_test, int64_t n_test,
                    int64_t dim, int64_t k, int64_t chunk_size, int strategy, int precision,
                    int64_t oversample, int flags, int n_threads_hint, float* dist_out,
                    int64_t* idx_out, void* workspace, size_t workspace_bytes, void* stream,
                    knn_stats* stats);

/* float64 inputs; precision must be KNN_PRECISION_EXACT. */
int knn_argkmin_f64(const double* X_train, int64_t n_train, const double* X_test,
                    int64_t n_test, int64_t dim, int64_t k, int64_t chunk_size, int strategy,
                    int precision, int64_t oversample, int flags, int n_threads_hint,
                    double* dist_out, int64_t* idx_out, void* workspace,
                    size_t workspace_bytes, void* stream, knn_stats* stats);

/*
 * knn_argkmin_f32 without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.
 * It calls no hipMalloc, hipFree or other allocating routine: its scratch is
 * a static device array of KNN_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends.  A request needing more (knn_argkmin_workspace_size() above
 * that size) fails with KNN_EWORKSPACE; KNN_FLAG_TIMING is rejected
 * (KNN_EINVAL: it would create HIP events).  Every call on one GPU shares
 * that GPU's array, so calls that overlap on one GPU (on different streams,
 * or from different threads without ordering) race and must not be made.
 * Arguments and results are those of knn_argkmin_f32.
 */
#define KNN_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int knn_argkmin_f32_noalloc(const float* X_train, int64_t n_train, const float* X_test,
                            int64_t n_test, int64_t dim, int64_t k, int64_t chunk_size,
                            int strategy, int precision, int64_t oversample, int flags,
                            int n_threads_hint, float* dist_out, int64_t* idx_out, void* stream,
                            knn_stats* stats);

/*
 * Gradients of the k-NN distances returned by knn_argkmin_* with respect to
 * the train and query rows.  No counterpart in the reference (its search is
 * not differentiable); used by the torch.autograd.Function of
 * pdist_aggregation_amd.autograd.
 *
 * indices: the forward's (n_test x k) result; grad_dist: dL/d(dist), same
 * shape; flags: KNN_FLAG_SQUARED if the forward returned squared distances.
 * grad_X_train (n_train x dim) and/or grad_X_test (n_test x dim) may be NULL.
 * grad_X_train is overwritten (zeroed, then accumulated with atomics: the
 * summation order, hence the last bits, is not deterministic).
 */
int knn_argkmin_backward_f32(const float* X_train, int64_t n_train, const float* X_test,
                             int64_t n_test, int64_t dim, int64_t k, const int64_t* indices,
                             const float* grad_dist, int flags, float* grad_X_train,
                             float* grad_X_test, void* stream);

int knn_argkmin_backward_f64(const double* X_train, int64_t n_train, const double* X_test,
                             int64_t n_test, int64_t dim, int64_t k, const int64_t* indices,
                             const double* grad_dist, int flags, double* grad_X_train,
                             double* grad_X_test, void* stream);

/*
 * Dense squared-distance matrix out[i][j] = sum_f (X_test[i][f] -
 * X_train[j][f])^2, (n_test x n_train) row-major, by direct differences.
 * Replaces the SYCL backend's distance_squared_matrix / pw_d2m
 * (sycl_ext/_pdist_aggregation_sycl.pyx:73-99,
 * sycl_ext/parallel_knn_helper.hpp:26-55).
 */
int knn_sqdist_matrix_f32(const float* X_train, int64_t n_train, const float* X_test,
                          int64_t n_test, int64_t dim, float* out, void* stream);
int knn_sqdist_matrix_f64(const double* X_train, int64_t n_train, const double* X_test,
                          int64_t n_test, int64_t dim, double* out, void* stream);

/* dtype_bytes: 4 (f32) or 8 (f64).  Returns 0 for an invalid request. */
size_t knn_argkmin_workspace_size(int dtype_bytes, int64_t n_train, int64_t n_test, int64_t dim,
                                  int64_t k, int precision, int64_t oversample);

/* Message of the last failed call on this thread ("" if none). */
const char* knn_last_error(void);

/* ABI version, bumped on any signature change. */
int knn_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PDIST_KNN_H */
