// CPU restatement of the reference's Cython/OpenMP argkmin engine.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker and the CPU baseline:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// it.  The product path (pdist_aggregation_amd) never links or calls it.
//
// It follows src/_pdist_aggregation.pyx of the reference step by step:
//   constants                    :24-28   (CHUNK_SIZE, MIN_CHUNK_SAMPLES, FLOAT_INF)
//   _push                        :54-105  (max-heap, rejects only val > root)
//   dual_swap / _simultaneous_sort :108-188 (median-of-3 quicksort)
//   _k_closest_on_chunk          :192-232 (-2 X_c Y_c^T GEMM + push of reduced dist)
//   _parallel_knn_on_X_test      :236-330
//   _parallel_knn_on_X_train     :333-448
//   _euclidean_dist / _compute_exact_distances :450-487
//   parallel_knn                 :492-540 (init values, strategy rule)
// Deliberate difference: the index heap of the chunk_on_train strategy is
// allocated with sizeof(int64) (the reference uses sizeof(floating) at
// :392-393, which under-allocates for float32).
// The reference calls an external BLAS GEMM (sklearn _cython_blas._gemm);
// here the same product C = -2 X_c Y_c^T is a cache-blocked loop nest.  The
// GIL-serialised heap merge of :420-430 is an OpenMP critical section.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>
#include <omp.h>

namespace {

constexpr double FLOAT_INF = 1e36;      // :28
constexpr int64_t MIN_CHUNK_SAMPLES = 20;  // :26

// _push (:54-105): push (val, i_val) into the max-heap (dist, idx) of `size`.
template <class T>
inline void heap_push(T* dist, int64_t* idx, int64_t size, T val, int64_t i_val) {
  if (val > dist[0]) return;  // an equal value replaces the root (:66)
  int64_t cur = 0;
  for (;;) {
    const int64_t l = 2 * cur + 1, r = l + 1;
    int64_t sw;
    if (l >= size) break;
    if (r >= size) {
      if (dist[l] > val) sw = l; else break;
    } else if (dist[l] >= dist[r]) {
      if (val < dist[l]) sw = l; else break;
    } else {
      if (val < dist[r]) sw = r; else break;
    }
    dist[cur] = dist[sw];
    idx[cur] = idx[sw];
    cur = sw;
  }
  dist[cur] = val;
  idx[cur] = i_val;
}

template <class T>
inline void dual_swap(T* dist, int64_t* idx, int64_t a, int64_t b) {
  std::swap(dist[a], dist[b]);
  std::swap(idx[a], idx[b]);
}

// _simultaneous_sort (:126-188): quicksort on dist with matching swaps on idx.
template <class T>
void simultaneous_sort(T* dist, int64_t* idx, int64_t size) {
  if (size <= 1) return;
  if (size == 2) {
    if (dist[0] > dist[1]) dual_swap(dist, idx, 0, 1);
    return;
  }
  if (size == 3) {
    if (dist[0] > dist[1]) dual_swap(dist, idx, 0, 1);
    if (dist[1] > dist[2]) {
      dual_swap(dist, idx, 1, 2);
      if (dist[0] > dist[1]) dual_swap(dist, idx, 0, 1);
    }
    return;
  }
  int64_t pivot_idx = size / 2;
  if (dist[0] > dist[size - 1]) dual_swap(dist, idx, 0, size - 1);
  if (dist[size - 1] > dist[pivot_idx]) {
    dual_swap(dist, idx, size - 1, pivot_idx);
    if (dist[0] > dist[size - 1]) dual_swap(dist, idx, 0, size - 1);
  }
  const T pivot_val = dist[size - 1];
  int64_t store = 0;
  for (int64_t i = 0; i < size - 1; ++i)
    if (dist[i] < pivot_val) dual_swap(dist, idx, i, store++);
  dual_swap(dist, idx, store, size - 1);
  pivot_idx = store;
  if (pivot_idx > 1) simultaneous_sort(dist, idx, pivot_idx);
  if (pivot_idx + 2 < size)
    simultaneous_sort(dist + pivot_idx + 1, idx + pivot_idx + 1, size - pivot_idx - 1);
}

// C[cq x ct] (row stride ct) = -2 * Xc[cq x d] . Yc[ct x d]^T   (:217-222)
// Cache-blocked: a panel of JP train rows is transposed into yt (d x JP) and
// swept by blocks of 4 query rows x 16 train rows held in registers.
template <class T>
void gemm_neg2(const T* Xc, int64_t cq, const T* Yc, int64_t ct, int64_t d, T* C,
               std::vector<T>& yt) {
  constexpr int64_t JP = 256, JB = 16, IB = 4;
  yt.resize((size_t)d * JP);
  for (int64_t j0 = 0; j0 < ct; j0 += JP) {
    const int64_t jn = std::min(JP, ct - j0);
    const int64_t jpad = (jn + JB - 1) / JB * JB;
    for (int64_t p = 0; p < d; ++p) {
      T* row = yt.data() + p * JP;
      for (int64_t j = 0; j < jn; ++j) row[j] = Yc[(j0 + j) * d + p];
      for (int64_t j = jn; j < jpad; ++j) row[j] = T(0);
    }
    for (int64_t i0 = 0; i0 < cq; i0 += IB) {
      const T* xr[IB];
      for (int a = 0; a < IB; ++a) xr[a] = Xc + std::min(i0 + a, cq - 1) * d;
      for (int64_t jb = 0; jb < jpad; jb += JB) {
        T acc[IB][JB] = {};
        for (int64_t p = 0; p < d; ++p) {
          const T* yrow = yt.data() + p * JP + jb;
          for (int a = 0; a < IB; ++a) {
            const T xa = xr[a][p];
#pragma omp simd
            for (int b = 0; b < JB; ++b) acc[a][b] += xa * yrow[b];
          }
        }
        for (int a = 0; a < IB && i0 + a < cq; ++a)
          for (int b = 0; b < JB && jb + b < jn; ++b)
            C[(i0 + a) * ct + j0 + jb + b] = T(-2) * acc[a][b];
      }
    }
  }
}

// _k_closest_on_chunk (:192-232)
template <class T>
void k_closest_on_chunk(const T* Ytr_c, int64_t ct, const T* Xte_c, int64_t cq, int64_t d,
                        const T* ysq_c, T* mid, T* heap_d, int64_t* heap_i, int64_t k,
                        int64_t offset, std::vector<T>& yt) {
  gemm_neg2(Xte_c, cq, Ytr_c, ct, d, mid, yt);
  for (int64_t i = 0; i < cq; ++i)
    for (int64_t j = 0; j < ct; ++j)
      heap_push(heap_d + i * k, heap_i + i * k, k, mid[i * ct + j] + ysq_c[j], j + offset);
}

struct Chunking {
  int64_t n_chunk, n_full, rem, n_chunks;
};
// chunk arithmetic of :247-269
inline Chunking chunking(int64_t n, int64_t chunk_size) {
  Chunking c;
  const int64_t nsc = std::max(MIN_CHUNK_SAMPLES, chunk_size);
  c.n_chunk = std::min(n, nsc);
  c.n_full = n / c.n_chunk;
  c.rem = n % c.n_chunk;
  c.n_chunks = c.n_full + (n != c.n_full * c.n_chunk);
  return c;
}

// _parallel_knn_on_X_test (:236-330)
template <class T>
int64_t knn_on_X_test(const T* Ytr, int64_t n_train, const T* Xte, int64_t n_test, int64_t d,
                      const T* ysq, int64_t k, int64_t chunk_size, int n_threads,
                      int64_t* knn_idx) {
  const Chunking tr = chunking(n_train, chunk_size), te = chunking(n_test, chunk_size);
  const int num_threads = (int)std::min<int64_t>(tr.n_chunks, n_threads);
#pragma omp parallel num_threads(num_threads)
  {
    std::vector<T> mid((size_t)tr.n_chunk * te.n_chunk);
    std::vector<T> heap_d((size_t)te.n_chunk * k);
    std::vector<T> yt;
#pragma omp for schedule(static)
    for (int64_t qc = 0; qc < te.n_chunks; ++qc) {
      std::fill(heap_d.begin(), heap_d.end(), T(FLOAT_INF));
      const int64_t q0 = qc * te.n_chunk;
      const int64_t q1 = (qc == te.n_chunks - 1 && te.rem > 0) ? q0 + te.rem : q0 + te.n_chunk;
      for (int64_t tc = 0; tc < tr.n_chunks; ++tc) {
        const int64_t t0 = tc * tr.n_chunk;
        const int64_t t1 = (tc == tr.n_chunks - 1 && tr.rem > 0) ? t0 + tr.rem : t0 + tr.n_chunk;
        k_closest_on_chunk(Ytr + t0 * d, t1 - t0, Xte + q0 * d, q1 - q0, d, ysq + t0, mid.data(),
                           heap_d.data(), knn_idx + q0 * k, k, t0, yt);
      }
      for (int64_t i = 0; i < q1 - q0; ++i)
        simultaneous_sort(heap_d.data() + i * k, knn_idx + (q0 + i) * k, k);
    }
  }
  return te.n_chunks;
}

// _parallel_knn_on_X_train (:333-448)
template <class T>
int64_t knn_on_X_train(const T* Ytr, int64_t n_train, const T* Xte, int64_t n_test, int64_t d,
                       const T* ysq, int64_t k, int64_t chunk_size, int n_threads,
                       int64_t* knn_idx, T* knn_red) {
  const Chunking tr = chunking(n_train, chunk_size), te = chunking(n_test, chunk_size);
  const int num_threads = (int)std::min<int64_t>(tr.n_chunks, n_threads);
  for (int64_t qc = 0; qc < te.n_chunks; ++qc) {
    const int64_t q0 = qc * te.n_chunk;
    const int64_t q1 = (qc == te.n_chunks - 1 && te.rem > 0) ? q0 + te.rem : q0 + te.n_chunk;
#pragma omp parallel num_threads(num_threads)
    {
      std::vector<T> mid((size_t)tr.n_chunk * te.n_chunk);
      std::vector<T> heap_d((size_t)te.n_chunk * k, T(FLOAT_INF));
      std::vector<int64_t> heap_i((size_t)te.n_chunk * k, -1);
      std::vector<T> yt;
#pragma omp for schedule(static) nowait
      for (int64_t tc = 0; tc < tr.n_chunks; ++tc) {
        const int64_t t0 = tc * tr.n_chunk;
        const int64_t t1 = (tc == tr.n_chunks - 1 && tr.rem > 0) ? t0 + tr.rem : t0 + tr.n_chunk;
        k_closest_on_chunk(Ytr + t0 * d, t1 - t0, Xte + q0 * d, q1 - q0, d, ysq + t0, mid.data(),
                           heap_d.data(), heap_i.data(), k, t0, yt);
      }
#pragma omp critical(knn_port_merge)
      {
        for (int64_t i = 0; i < q1 - q0; ++i)
          for (int64_t j = 0; j < k; ++j)
            heap_push(knn_red + (q0 + i) * k, knn_idx + (q0 + i) * k, k, heap_d[i * k + j],
                      heap_i[i * k + j]);
      }
    }
#pragma omp parallel for schedule(static) num_threads(num_threads)
    for (int64_t i = 0; i < n_test; ++i) simultaneous_sort(knn_red + i * k, knn_idx + i * k, k);
  }
  return tr.n_chunks;
}

// _euclidean_dist (:450-471): sequential sum of squared differences, then sqrt.
template <class T>
inline T euclidean_dist(const T* x, const T* y, int64_t d) {
  T s = 0;
  for (int64_t p = 0; p < d; ++p) s += (x[p] - y[p]) * (x[p] - y[p]);
  return std::sqrt(s);
}

// parallel_knn (:492-540).  strategy: 0 auto, 1 chunk_on_train, 2 chunk_on_test.
template <class T>
int64_t parallel_knn(const T* Ytr, int64_t n_train, const T* Xte, int64_t n_test, int64_t d,
                     int64_t k, int64_t chunk_size, int strategy, int return_distance,
                     int n_threads, int64_t* knn_idx, T* knn_dist, int* strategy_used) {
  if (n_threads <= 0) n_threads = omp_get_max_threads();
  for (int64_t i = 0; i < n_test * k; ++i) {  // :503-507
    knn_idx[i] = 0;
    knn_dist[i] = T(FLOAT_INF);
  }
  std::vector<T> ysq((size_t)n_train);  // :508
  for (int64_t j = 0; j < n_train; ++j) {
    T s = 0;
    for (int64_t p = 0; p < d; ++p) s += Ytr[j * d + p] * Ytr[j * d + p];
    ysq[j] = s;
  }
  if (strategy == 0)  // :511-515
    strategy = (4 * chunk_size * (int64_t)n_threads < n_test) ? 2 : 1;
  if (strategy_used) *strategy_used = strategy;
  int64_t n_chunks;
  if (strategy == 1)
    n_chunks = knn_on_X_train(Ytr, n_train, Xte, n_test, d, ysq.data(), k, chunk_size, n_threads,
                              knn_idx, knn_dist);
  else if (strategy == 2)
    n_chunks = knn_on_X_test(Ytr, n_train, Xte, n_test, d, ysq.data(), k, chunk_size, n_threads,
                             knn_idx);
  else
    return -1;
  if (return_distance) {  // :473-487, rows are not re-sorted
#pragma omp parallel for schedule(static) num_threads(n_threads)
    for (int64_t i = 0; i < n_test; ++i)
      for (int64_t j = 0; j < k; ++j)
        knn_dist[i * k + j] = euclidean_dist(Xte + i * d, Ytr + knn_idx[i * k + j] * d, d);
  }
  return n_chunks;
}

}  // namespace

extern "C" {

int64_t port_parallel_knn_f64(const double* X_train, int64_t n_train, const double* X_test,
                              int64_t n_test, int64_t d, int64_t k, int64_t chunk_size,
                              int strategy, int return_distance, int n_threads,
                              int64_t* knn_indices, double* knn_distances, int* strategy_used) {
  return parallel_knn<double>(X_train, n_train, X_test, n_test, d, k, chunk_size, strategy,
                              return_distance, n_threads, knn_indices, knn_distances,
                              strategy_used);
}

int64_t port_parallel_knn_f32(const float* X_train, int64_t n_train, const float* X_test,
                              int64_t n_test, int64_t d, int64_t k, int64_t chunk_size,
                              int strategy, int return_distance, int n_threads,
                              int64_t* knn_indices, float* knn_distances, int* strategy_used) {
  return parallel_knn<float>(X_train, n_train, X_test, n_test, d, k, chunk_size, strategy,
                             return_distance, n_threads, knn_indices, knn_distances,
                             strategy_used);
}

int port_omp_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
