// Preparation of the low-precision candidate search.
//
// Rows are centred on a common vector mu (the feature means of up to 65,536
// train rows), rounded to bf16 and zero-padded to the kernel's feature count.
// Distances are translation invariant, and centring shrinks |x.y| and so the
// rounding error.  For every row we keep |q(r)|^2 (the C operand of the MFMA
// search) and an upper bound of the rounding error |r - mu - q(r)|, which the
// merge kernel uses to prove that the exact re-rank saw every true neighbour:
// by the triangle inequality | |q(x)-q(y)| - |x-y| | <= err(x) + err(y).
// There is no counterpart in the reference (its search is float64 only).
#include "knn_common.hpp"
#include "knn_kernels.hpp"

namespace knn {

namespace {

// Rows per column-sum workgroup: 64 keeps each thread's chain of dependent
// adds short (256 rows per workgroup: 17 us for 65,536 x 128 floats).
constexpr int kSampleRowsPerBlock = 64;
typedef float f4 __attribute__((ext_vector_type(4)));

// Column means without floating-point atomics, so the centring vector --
// and with it every quantised value -- is the same on every call: block b
// writes the sums of its kSampleRowsPerBlock rows to part[b][f], then one
// thread per feature adds the blocks in order.
__global__ __launch_bounds__(kThreads) void column_sum_kernel(const float* __restrict__ Y,
                                                              int64_t n, int d,
                                                              float* __restrict__ part) {
  const int64_t r0 = (int64_t)blockIdx.x * kSampleRowsPerBlock;
  const int64_t r1 = r0 + kSampleRowsPerBlock < n ? r0 + kSampleRowsPerBlock : n;
  for (int f = threadIdx.x; f < d; f += kThreads) {
    float s = 0.f;
    for (int64_t r = r0; r < r1; ++r) s += Y[r * d + f];
    part[(size_t)blockIdx.x * d + f] = s;
  }
}

// One workgroup per feature: thread t adds blocks t, t + kThreads, ..., then a
// fixed tree over the threads.
__global__ __launch_bounds__(kThreads) void column_mean_kernel(const float* __restrict__ part,
                                                               int n_blocks, int d, double inv,
                                                               float* __restrict__ mu) {
  __shared__ double red[kThreads];
  const int f = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < n_blocks; b += kThreads) s += part[(size_t)b * d + f];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) mu[f] = (float)(red[0] * inv);
}

// One wave per row.
__global__ __launch_bounds__(kThreads) void quantize_bf16_kernel(
    const float* __restrict__ src, int64_t n, int d, int dm, const float* __restrict__ mu,
    __bf16* __restrict__ dst, float* __restrict__ norm, float* __restrict__ err,
    DeviceScalars* sc, int track_max) {
  __shared__ float wn[kThreads / 64];
  __shared__ float we[kThreads / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  float best_n = 0.f, best_e = 0.f;
  for (int64_t r = (int64_t)blockIdx.x * 4 + w; r < n; r += (int64_t)gridDim.x * 4) {
    float ns = 0.f;
    double es = 0.0;
    for (int f = lane; f < dm; f += 64) {
      float q = 0.f;
      if (f < d) {
        const float x = src[r * d + f];
        const float cen = x - mu[f];
        const __bf16 qb = (__bf16)cen;
        q = (float)qb;
        dst[r * dm + f] = qb;
        const double e = ((double)x - (double)mu[f]) - (double)q;
        es = fma(e, e, es);
      } else {
        dst[r * dm + f] = (__bf16)0.f;
      }
      ns = fmaf(q, q, ns);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      ns += __shfl_xor(ns, o, 64);
      es += __shfl_xor(es, o, 64);
    }
    // Round the error bound up: sqrt in double, then one float ulp of margin.
    const float eb = __uint_as_float(__float_as_uint((float)sqrt(es)) + 1u);  // next float up
    if (lane == 0) {
      norm[r] = ns;
      if (err) err[r] = eb;
    }
    best_n = fmaxf(best_n, ns);
    best_e = fmaxf(best_e, eb);
  }
  if (!track_max) return;
  if (lane == 0) {
    wn[w] = best_n;
    we[w] = best_e;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float mn = wn[0], me = we[0];
    for (int i = 1; i < kThreads / 64; ++i) {
      mn = fmaxf(mn, wn[i]);
      me = fmaxf(me, we[i]);
    }
    atomicMax(&sc->max_ynorm_key, ordered_key((double)mn));
    atomicMax(&sc->max_yerr_key, ordered_key((double)me));
  }
}

// OCP FP8 E4M3 (bias 7, no infinities, max normal 448), round to nearest
// even, with the gfx950 conversion instructions (v_cvt_pk_fp8_f32 /
// v_cvt_f32_fp8).  |v| <= 448 on entry.  Their bytes equal those of a
// software round-to-nearest-even encoder (frexp, 3-bit mantissa, subnormals
// m * 2^-9, saturation at 448) for every exponent from 2^-20 to 448 and all
// 16-bit rounding patterns below the kept bits (3.9e6 values checked on the
// GPU), except that -0 keeps its sign bit (0x80, the value is the same).
__device__ inline unsigned encode_e4m3(float v) {
  return (unsigned)__builtin_amdgcn_cvt_pk_fp8_f32(v, 0.f, 0, false) & 0xffu;
}

__device__ inline float decode_e4m3(unsigned b) { return __builtin_amdgcn_cvt_f32_fp8((int)b, 0); }

constexpr int kFp8MaxPerLane = 4;  // dm <= 256

// One wave per row: centre, choose a power-of-two row scale 2^e with
// |r - mu| / 2^e <= 448, encode, and keep |deq|^2 and the error bound.
__global__ __launch_bounds__(kThreads) void quantize_fp8_kernel(
    const float* __restrict__ src, int64_t n, int d, int dm, const float* __restrict__ mu,
    unsigned char* __restrict__ dst, unsigned char* __restrict__ scale, float* __restrict__ norm,
    float* __restrict__ err, DeviceScalars* sc, int track_max) {
  __shared__ float wn[kThreads / 64];
  __shared__ float we[kThreads / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  float best_n = 0.f, best_e = 0.f;
  for (int64_t r = (int64_t)blockIdx.x * 4 + w; r < n; r += (int64_t)gridDim.x * 4) {
    float cen[kFp8MaxPerLane];
    float amax = 0.f;
#pragma unroll
    for (int i = 0; i < kFp8MaxPerLane; ++i) {
      const int f = lane + 64 * i;
      cen[i] = (f < d) ? src[r * d + f] - mu[f] : 0.f;
      amax = fmaxf(amax, fabsf(cen[i]));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) amax = fmaxf(amax, __shfl_xor(amax, o, 64));
    int e = 0;
    if (amax > 0.f) {
      int E;
      const float fr = frexpf(amax / 448.f, &E);  // amax/448 = fr * 2^E
      e = fr > 0.5f ? E : E - 1;                   // smallest e with amax/448 <= 2^e
      e = e < -120 ? -120 : (e > 120 ? 120 : e);
    }
    float ns = 0.f;
    double es = 0.0;
#pragma unroll
    for (int i = 0; i < kFp8MaxPerLane; ++i) {
      const int f = lane + 64 * i;
      if (f >= dm) break;
      const unsigned b = f < d ? encode_e4m3(ldexpf(cen[i], -e)) : 0u;
      dst[r * dm + f] = (unsigned char)b;
      const float q = ldexpf(decode_e4m3(b), e);
      ns = fmaf(q, q, ns);
      if (f < d) {
        const double ed = ((double)src[r * d + f] - (double)mu[f]) - (double)q;
        es = fma(ed, ed, es);
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      ns += __shfl_xor(ns, o, 64);
      es += __shfl_xor(es, o, 64);
    }
    const float eb = __uint_as_float(__float_as_uint((float)sqrt(es)) + 1u);
    if (lane == 0) {
      norm[r] = ns;
      scale[r] = (unsigned char)(e + 127);  // E8M0
      if (err) err[r] = eb;
    }
    best_n = fmaxf(best_n, ns);
    best_e = fmaxf(best_e, eb);
  }
  if (!track_max) return;
  if (lane == 0) {
    wn[w] = best_n;
    we[w] = best_e;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float mn = wn[0], me = we[0];
    for (int i = 1; i < kThreads / 64; ++i) {
      mn = fmaxf(mn, wn[i]);
      me = fmaxf(me, we[i]);
    }
    atomicMax(&sc->max_ynorm_key, ordered_key((double)mn));
    atomicMax(&sc->max_yerr_key, ordered_key((double)me));
  }
}

// ---- vectorised forms: d % 4 == 0, 16-byte aligned rows --------------------
// LPR lanes cover one row of dm = 4 * LPR features (4 per lane, float4 loads),
// so a wave holds 64 / LPR rows at once and each lane keeps kVecRows rows'
// loads in flight.
constexpr int kVecRows = 4;

__global__ __launch_bounds__(kThreads) void column_sum_vec_kernel(const float* __restrict__ Y,
                                                                  int64_t n, int d,
                                                                  float* __restrict__ part) {
  __shared__ f4 red[kThreads];
  const int groups = d >> 2;                 // float4 columns (<= 64)
  const int lanes_r = kThreads / groups;     // rows summed in parallel
  const int cg = threadIdx.x % groups, rl = threadIdx.x / groups;
  const int64_t r0 = (int64_t)blockIdx.x * kSampleRowsPerBlock;
  const int64_t r1 = r0 + kSampleRowsPerBlock < n ? r0 + kSampleRowsPerBlock : n;
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  if (rl < lanes_r)
    for (int64_t r = r0 + rl; r < r1; r += lanes_r) {
      const f4 v = *reinterpret_cast<const f4*>(Y + r * d + 4 * cg);
      acc.x += v.x;
      acc.y += v.y;
      acc.z += v.z;
      acc.w += v.w;
    }
  red[threadIdx.x] = acc;
  __syncthreads();
  if (rl == 0) {
    for (int i = 1; i < lanes_r; ++i) {
      const f4 v = red[i * groups + cg];
      acc.x += v.x;
      acc.y += v.y;
      acc.z += v.z;
      acc.w += v.w;
    }
    *reinterpret_cast<f4*>(part + (size_t)blockIdx.x * d + 4 * cg) = acc;
  }
}

// Block-wide maxima of the row norms and error bounds -> sc.
__device__ inline void track_maxima(float best_n, float best_e, DeviceScalars* sc) {
  __shared__ float wn[kThreads / 64];
  __shared__ float we[kThreads / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    best_n = fmaxf(best_n, __shfl_xor(best_n, o, 64));
    best_e = fmaxf(best_e, __shfl_xor(best_e, o, 64));
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    wn[w] = best_n;
    we[w] = best_e;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float mn = wn[0], me = we[0];
    for (int i = 1; i < kThreads / 64; ++i) {
      mn = fmaxf(mn, wn[i]);
      me = fmaxf(me, we[i]);
    }
    atomicMax(&sc->max_ynorm_key, ordered_key((double)mn));
    atomicMax(&sc->max_yerr_key, ordered_key((double)me));
  }
}

template <int LPR>
__global__ __launch_bounds__(kThreads) void quantize_bf16_vec_kernel(
    const float* __restrict__ src, int64_t n, int d, const float* __restrict__ mu,
    __bf16* __restrict__ dst, float* __restrict__ norm, float* __restrict__ err,
    DeviceScalars* sc, int track_max) {
  constexpr int DM = 4 * LPR;
  constexpr int RPB = (kThreads / LPR);  // rows per block and pass
  const int l = threadIdx.x % LPR, rsub = threadIdx.x / LPR;
  const int f0 = 4 * l;
  const bool fok = f0 < d;
  const f4 mu4 = fok ? *reinterpret_cast<const f4*>(mu + f0) : f4{0.f, 0.f, 0.f, 0.f};
  float best_n = 0.f, best_e = 0.f;
  for (int64_t base = (int64_t)blockIdx.x * RPB * kVecRows; base < n;
       base += (int64_t)gridDim.x * RPB * kVecRows) {
    f4 x[kVecRows];
#pragma unroll
    for (int u = 0; u < kVecRows; ++u) {
      const int64_t r = base + u * RPB + rsub;
      x[u] = (fok && r < n) ? *reinterpret_cast<const f4*>(src + r * d + f0)
                            : f4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int u = 0; u < kVecRows; ++u) {
      const int64_t r = base + u * RPB + rsub;
      const float xv[4] = {x[u].x, x[u].y, x[u].z, x[u].w};
      const float mv[4] = {mu4.x, mu4.y, mu4.z, mu4.w};
      __bf16 qb[4];
      float ns = 0.f;
      double es = 0.0;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        qb[i] = fok ? (__bf16)(xv[i] - mv[i]) : (__bf16)0.f;
        const float q = (float)qb[i];
        ns = fmaf(q, q, ns);
        const double e = fok ? ((double)xv[i] - (double)mv[i]) - (double)q : 0.0;
        es = fma(e, e, es);
      }
#pragma unroll
      for (int o = LPR / 2; o > 0; o >>= 1) {
        ns += __shfl_xor(ns, o, 64);
        es += __shfl_xor(es, o, 64);
      }
      if (r < n) {
        unsigned short h[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) h[i] = __builtin_bit_cast(unsigned short, qb[i]);
        const uint2 packed = {(unsigned)h[0] | ((unsigned)h[1] << 16),
                              (unsigned)h[2] | ((unsigned)h[3] << 16)};
        *reinterpret_cast<uint2*>(dst + r * DM + f0) = packed;
        const float eb = __uint_as_float(__float_as_uint((float)sqrt(es)) + 1u);  // next float up
        if (l == 0) {
          norm[r] = ns;
          if (err) err[r] = eb;
        }
        best_n = fmaxf(best_n, ns);
        best_e = fmaxf(best_e, eb);
      }
    }
  }
  if (track_max) track_maxima(best_n, best_e, sc);
}

template <int LPR>
__global__ __launch_bounds__(kThreads) void quantize_fp8_vec_kernel(
    const float* __restrict__ src, int64_t n, int d, const float* __restrict__ mu,
    unsigned char* __restrict__ dst, unsigned char* __restrict__ scale, float* __restrict__ norm,
    float* __restrict__ err, DeviceScalars* sc, int track_max) {
  constexpr int DM = 4 * LPR;
  constexpr int RPB = (kThreads / LPR);
  const int l = threadIdx.x % LPR, rsub = threadIdx.x / LPR;
  const int f0 = 4 * l;
  const bool fok = f0 < d;
  const f4 mu4 = fok ? *reinterpret_cast<const f4*>(mu + f0) : f4{0.f, 0.f, 0.f, 0.f};
  float best_n = 0.f, best_e = 0.f;
  for (int64_t base = (int64_t)blockIdx.x * RPB * kVecRows; base < n;
       base += (int64_t)gridDim.x * RPB * kVecRows) {
    f4 x[kVecRows];
#pragma unroll
    for (int u = 0; u < kVecRows; ++u) {
      const int64_t r = base + u * RPB + rsub;
      x[u] = (fok && r < n) ? *reinterpret_cast<const f4*>(src + r * d + f0)
                            : f4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int u = 0; u < kVecRows; ++u) {
      const int64_t r = base + u * RPB + rsub;
      const float xv[4] = {x[u].x, x[u].y, x[u].z, x[u].w};
      const float mv[4] = {mu4.x, mu4.y, mu4.z, mu4.w};
      float cen[4], amax = 0.f;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        cen[i] = fok ? xv[i] - mv[i] : 0.f;
        amax = fmaxf(amax, fabsf(cen[i]));
      }
#pragma unroll
      for (int o = LPR / 2; o > 0; o >>= 1) amax = fmaxf(amax, __shfl_xor(amax, o, 64));
      int e = 0;
      if (amax > 0.f) {
        int E;
        const float fr = frexpf(amax / 448.f, &E);  // amax/448 = fr * 2^E
        e = fr > 0.5f ? E : E - 1;                   // smallest e with amax/448 <= 2^e
        e = e < -120 ? -120 : (e > 120 ? 120 : e);
      }
      float ns = 0.f;
      double es = 0.0;
      // four features in two packed conversions: bytes 0, 1 (low half), 2, 3
      int packed = __builtin_amdgcn_cvt_pk_fp8_f32(ldexpf(cen[0], -e), ldexpf(cen[1], -e), 0,
                                                   false);
      packed = __builtin_amdgcn_cvt_pk_fp8_f32(ldexpf(cen[2], -e), ldexpf(cen[3], -e), packed,
                                               true);
      if (!fok) packed = 0;
      const unsigned word = (unsigned)packed;
      const float dq[4] = {__builtin_amdgcn_cvt_f32_fp8(packed, 0),
                           __builtin_amdgcn_cvt_f32_fp8(packed, 1),
                           __builtin_amdgcn_cvt_f32_fp8(packed, 2),
                           __builtin_amdgcn_cvt_f32_fp8(packed, 3)};
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float q = ldexpf(dq[i], e);
        ns = fmaf(q, q, ns);
        const double ed = fok ? ((double)xv[i] - (double)mv[i]) - (double)q : 0.0;
        es = fma(ed, ed, es);
      }
#pragma unroll
      for (int o = LPR / 2; o > 0; o >>= 1) {
        ns += __shfl_xor(ns, o, 64);
        es += __shfl_xor(es, o, 64);
      }
      if (r < n) {
        *reinterpret_cast<unsigned*>(dst + r * DM + f0) = word;
        const float eb = __uint_as_float(__float_as_uint((float)sqrt(es)) + 1u);
        if (l == 0) {
          norm[r] = ns;
          scale[r] = (unsigned char)(e + 127);  // E8M0
          if (err) err[r] = eb;
        }
        best_n = fmaxf(best_n, ns);
        best_e = fmaxf(best_e, eb);
      }
    }
  }
  if (track_max) track_maxima(best_n, best_e, sc);
}

bool vec_ok(const float* src, int d, int dm) {
  return (d & 3) == 0 && (reinterpret_cast<uintptr_t>(src) & 15) == 0 && dm >= 32 && dm <= 256;
}

// Every workgroup ends with one atomicMax per tracked maximum, on one address
// each: 2048 workgroups (8 per CU, all resident) loop over the rows rather
// than 8192 finishing one pass each (1e6 x 128 to bf16: 0.40 -> 0.31 ms of
// preparation; 1024 is slower again).
constexpr int64_t kMaxRowBlocks = 2048;

int64_t vec_blocks(int64_t n, int dm) {
  const int64_t rows_per_block = (int64_t)(kThreads / (dm / 4)) * kVecRows;
  const int64_t b = ceil_div(n, rows_per_block);
  return b < kMaxRowBlocks ? b : kMaxRowBlocks;
}

}  // namespace

void launch_quantize_fp8(const float* src, int64_t n, int d, int dm, const float* mu, void* dst,
                         unsigned char* scale, float* norm, float* err, DeviceScalars* sc,
                         bool track_max, hipStream_t s) {
  if (n <= 0) return;
  if (vec_ok(src, d, dm)) {
    auto* out = reinterpret_cast<unsigned char*>(dst);
    const dim3 g((unsigned)vec_blocks(n, dm)), b(kThreads);
    const int tm = track_max ? 1 : 0;
    if (dm == 64)
      hipLaunchKernelGGL(quantize_fp8_vec_kernel<16>, g, b, 0, s, src, n, d, mu, out, scale, norm,
                         err, sc, tm);
    else if (dm == 128)
      hipLaunchKernelGGL(quantize_fp8_vec_kernel<32>, g, b, 0, s, src, n, d, mu, out, scale, norm,
                         err, sc, tm);
    else if (dm == 256)
      hipLaunchKernelGGL(quantize_fp8_vec_kernel<64>, g, b, 0, s, src, n, d, mu, out, scale, norm,
                         err, sc, tm);
    if (dm == 64 || dm == 128 || dm == 256) return;
  }
  const int64_t blocks = ceil_div(n, 4) < kMaxRowBlocks ? ceil_div(n, 4) : kMaxRowBlocks;
  hipLaunchKernelGGL(quantize_fp8_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, s, src, n, d,
                     dm, mu, reinterpret_cast<unsigned char*>(dst), scale, norm, err, sc,
                     track_max ? 1 : 0);
}

namespace {

// n rows of d floats copied into rows of dp >= d floats, zero-filled: the
// float32 MFMA search then reads 16-byte aligned rows when d % 4 != 0.
__global__ __launch_bounds__(kThreads) void pad_rows_kernel(const float* __restrict__ src,
                                                            int64_t n, int d, int dp,
                                                            float* __restrict__ dst) {
  const int64_t total = n * dp;
  for (int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * kThreads) {
    const int64_t r = e / dp;
    const int c = (int)(e - r * dp);
    dst[e] = c < d ? src[r * d + c] : 0.f;
  }
}

}  // namespace

void launch_pad_rows(const float* src, int64_t n, int d, int dp, float* dst, hipStream_t s) {
  if (n <= 0) return;
  const int64_t want = ceil_div(n * dp, (int64_t)kThreads);
  const int64_t blocks = want < 8192 ? want : 8192;
  hipLaunchKernelGGL(pad_rows_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, s, src, n, d, dp,
                     dst);
}

void launch_column_mean(const float* Y, int64_t n_sample, int d, float* part, float* mu,
                        hipStream_t s) {
  const int64_t blocks = column_mean_blocks(n_sample);
  if ((d & 3) == 0 && d <= 4 * 64 && (reinterpret_cast<uintptr_t>(Y) & 15) == 0)
    hipLaunchKernelGGL(column_sum_vec_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, s, Y,
                       n_sample, d, part);
  else
    hipLaunchKernelGGL(column_sum_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, s, Y,
                       n_sample, d, part);
  hipLaunchKernelGGL(column_mean_kernel, dim3((unsigned)d), dim3(kThreads), 0, s, part,
                     (int)blocks, d, 1.0 / (double)n_sample, mu);
}

int64_t column_mean_blocks(int64_t n_sample) { return ceil_div(n_sample, kSampleRowsPerBlock); }

void launch_quantize_bf16(const float* src, int64_t n, int d, int dm, const float* mu, void* dst,
                          float* norm, float* err, DeviceScalars* sc, bool track_max,
                          hipStream_t s) {
  if (n <= 0) return;
  if (vec_ok(src, d, dm)) {
    auto* out = reinterpret_cast<__bf16*>(dst);
    const dim3 g((unsigned)vec_blocks(n, dm)), b(kThreads);
    const int tm = track_max ? 1 : 0;
    if (dm == 32)
      hipLaunchKernelGGL(quantize_bf16_vec_kernel<8>, g, b, 0, s, src, n, d, mu, out, norm, err,
                         sc, tm);
    else if (dm == 64)
      hipLaunchKernelGGL(quantize_bf16_vec_kernel<16>, g, b, 0, s, src, n, d, mu, out, norm, err,
                         sc, tm);
    else if (dm == 128)
      hipLaunchKernelGGL(quantize_bf16_vec_kernel<32>, g, b, 0, s, src, n, d, mu, out, norm, err,
                         sc, tm);
    else if (dm == 256)
      hipLaunchKernelGGL(quantize_bf16_vec_kernel<64>, g, b, 0, s, src, n, d, mu, out, norm, err,
                         sc, tm);
    if (dm == 32 || dm == 64 || dm == 128 || dm == 256) return;
  }
  const int64_t blocks = ceil_div(n, 4) < kMaxRowBlocks ? ceil_div(n, 4) : kMaxRowBlocks;
  hipLaunchKernelGGL(quantize_bf16_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, s, src, n, d,
                     dm, mu, reinterpret_cast<__bf16*>(dst), norm, err, sc, track_max ? 1 : 0);
}

}  // namespace knn
