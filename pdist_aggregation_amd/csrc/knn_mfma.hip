// MFMA fused distance + candidate-selection kernel for gfx950.
//
// Same contract as the generic kernel (knn_generic.hip) -- per (query block,
// train split) workgroup, keep every train row whose reduced distance
// v = |y|^2 - 2 x.y is <= the query's threshold, compact lists to the best k'
// -- but the dot products run on the matrix cores:
//
//   * one 32x32 MFMA output tile = 32 train rows (A operand) x 32 queries
//     (B operand); lane l owns query l%32 and 16 train rows of the tile, so
//     each lane compares its 16 values against ONE threshold;
//   * the queries of a wave stay in registers for the whole sweep, already
//     multiplied by -2 (exact), as the B operand;
//   * train rows stream global -> registers -> LDS (double buffered, one
//     barrier per 32-row step) and are read back in the A-operand layout;
//   * |y|^2 enters as the C operand of the first MFMA of each tile, so the
//     accumulator IS the reduced distance and selection costs a min-tree of
//     16 values and one compare per lane in the common case;
//   * survivors are appended to per-query lists in global scratch; lists
//     are sorted and cut by the owning wave alone (no workgroup barrier);
//   * the per-query threshold is shared across splits through an atomicMin.
//
// The feature axis may be permuted freely inside a dot product, so lane l
// (half h = l/32) feeds the features [h*DM/2, (h+1)*DM/2) of its row/query in
// both operands, which keeps every LDS read a contiguous 16 bytes.
//
// Formats: F32 = v_mfma_f32_32x32x2_f32 on the caller's float32 rows;
//          BF16 = v_mfma_f32_32x32x16_bf16 on rows pre-quantised by
//          knn_quant.hip (centred, zero-padded to DM);
//          FP8 = v_mfma_scale_f32_32x32x64_f8f6f4 (OCP E4M3 operands) with the
//          per-row power-of-two scales applied by the instruction itself.
#include "knn_common.hpp"
#include "knn_kernels.hpp"

namespace knn {

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x8 __attribute__((ext_vector_type(8)));

constexpr int kBY = 32;  // train rows per step (one MFMA row block)

template <int FMT>
struct Fmt;
template <>
struct Fmt<kFmtF32> {
  static constexpr int esize = 4;
  static constexpr int KL = 1;  // elements per lane per MFMA
  using breg = float;
};
template <>
struct Fmt<kFmtBF16> {
  static constexpr int esize = 2;
  static constexpr int KL = 8;
  using breg = bf16x8;
};
template <>
struct Fmt<kFmtFP8> {
  static constexpr int esize = 1;
  static constexpr int KL = 32;
  using breg = i32x8;
};

// LDS row pitch in bytes: DM elements + 16 bytes, so the 16-byte reads of 8
// (or 16) consecutive rows start in distinct banks.
template <int FMT, int DM>
constexpr int row_pitch() {
  return DM * Fmt<FMT>::esize + 16;
}

template <int FMT, int DM, int QG>
__global__ __launch_bounds__(kThreads, 2) void mfma_search_kernel(MfmaArgs a) {
  using F = Fmt<FMT>;
  constexpr int PITCH = row_pitch<FMT, DM>();
  constexpr int WQ = QG * 32;           // queries per wave
  constexpr int BQ = 4 * WQ;            // queries per workgroup
  constexpr int NSTEP = DM / (2 * F::KL);  // MFMAs per 32x32 tile
  constexpr int CHUNKS = kBY * DM * F::esize / 16;      // 16-byte loads per tile
  constexpr int CPT = (CHUNKS + kThreads - 1) / kThreads;  // per thread

  __shared__ __align__(16) unsigned char ytile[2][kBY * PITCH];
  __shared__ __align__(16) float yn_s[2][kBY];
  __shared__ int ye_s[2][kBY];  // FP8: E8M0 row scales
  __shared__ int cnt_s[BQ];
  __shared__ float tau_s[BQ];
  extern __shared__ __align__(16) unsigned char dyn_smem[];  // 4 waves x cap x 8 B

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, h = lane >> 5, c = lane & 31;

  // Workgroup -> (query block, split).  When the split count is a multiple of
  // 8, all query blocks of one split land on one XCD (workgroups are dealt to
  // the 8 XCDs round-robin) and run back to back, so they share that XCD's L2
  // while streaming the same train rows.
  int qb, split;
  {
    const int b = blockIdx.x;
    if ((a.n_splits & 7) == 0) {
      const int xcd = b & 7, t = b >> 3;
      split = xcd + 8 * (t / a.n_qblocks);
      qb = t % a.n_qblocks;
    } else {
      split = b / a.n_qblocks;
      qb = b % a.n_qblocks;
    }
  }
  const int64_t q0 = (int64_t)qb * BQ;
  const int64_t ybeg = (int64_t)split * a.rows_per_split;
  const int64_t yend = ybeg + a.rows_per_split < a.n_t ? ybeg + a.rows_per_split : a.n_t;
  const size_t blk = (size_t)split * a.n_qblocks + qb;
  float* cand_d = a.cand_d + blk * BQ * a.cap;
  int* cand_i = a.cand_i + blk * BQ * a.cap;
  float* sd = reinterpret_cast<float*>(dyn_smem) + (size_t)w * 2 * a.cap;
  int* si = reinterpret_cast<int*>(sd + a.cap);
  const int wq0 = w * WQ;

  // ---- resident queries: B operand, scaled by -2 ------------------------------
  typename F::breg xr[QG][NSTEP];
  int xsc[QG];  // FP8: E8M0 scale of -2 x (the row scale times 2)
  float tau[QG];
#pragma unroll
  for (int g = 0; g < QG; ++g) {
    const int64_t q = q0 + wq0 + g * 32 + c;
    const bool qok = q < a.n_q;
    tau[g] = qok ? FLT_MAX : -__builtin_huge_valf();
    if constexpr (FMT == kFmtF32) {
      const float* xrow = reinterpret_cast<const float*>(a.X) + q * a.ldx;
#pragma unroll
      for (int s = 0; s < NSTEP; s += 4) {
        const int f = h * (DM / 2) + s;
        f32x4 v = {0.f, 0.f, 0.f, 0.f};
        if (qok && f < a.d) v = *reinterpret_cast<const f32x4*>(xrow + f);
        xr[g][s + 0] = -2.f * v[0];
        xr[g][s + 1] = -2.f * v[1];
        xr[g][s + 2] = -2.f * v[2];
        xr[g][s + 3] = -2.f * v[3];
      }
    } else if constexpr (FMT == kFmtBF16) {
      const __bf16* xrow = reinterpret_cast<const __bf16*>(a.X) + q * a.ldx;
#pragma unroll
      for (int s = 0; s < NSTEP; ++s) {
        bf16x8 v = {};
        if (qok) v = *reinterpret_cast<const bf16x8*>(xrow + h * (DM / 2) + s * 8);
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = (__bf16)(-2.f * (float)v[e]);
        xr[g][s] = v;
      }
    } else {
      // -2 x: flip every sign bit, and double the power-of-two row scale.
      const unsigned char* xrow = reinterpret_cast<const unsigned char*>(a.X) + q * a.ldx;
#pragma unroll
      for (int s = 0; s < NSTEP; ++s) {
        i32x8 v = {};
        if (qok) {
          const i32x4 lo = *reinterpret_cast<const i32x4*>(xrow + h * (DM / 2) + s * 32);
          const i32x4 hi = *reinterpret_cast<const i32x4*>(xrow + h * (DM / 2) + s * 32 + 16);
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            v[e] = lo[e] ^ (int)0x80808080u;
            v[e + 4] = hi[e] ^ (int)0x80808080u;
          }
        }
        xr[g][s] = v;
      }
      xsc[g] = qok ? (int)a.xscale[q] + 1 : 127;
    }
  }
  for (int i = tid; i < BQ; i += kThreads) {
    cnt_s[i] = 0;
    tau_s[i] = FLT_MAX;
  }

  // ---- tile loader: global -> registers -> LDS -------------------------------
  i32x4 pre[CPT];
  float pre_n = 0.f;
  int pre_e = 127;
  auto load_tile = [&](int64_t y0) {
#pragma unroll
    for (int i = 0; i < CPT; ++i) {
      const int ch = tid + i * kThreads;
      const int row = ch / (DM * F::esize / 16);
      const int col = ch % (DM * F::esize / 16);
      const int64_t y = y0 + row;
      const int f = col * (16 / F::esize);
      i32x4 v = {0, 0, 0, 0};
      if (ch < CHUNKS && y < yend && f < a.d)
        v = *reinterpret_cast<const i32x4*>(reinterpret_cast<const unsigned char*>(a.Y) +
                                           ((size_t)y * a.ldy + f) * F::esize);
      pre[i] = v;
    }
    if (tid < kBY) {
      const int64_t y = y0 + tid;
      pre_n = y < yend ? a.ynorm[y] : __builtin_huge_valf();
      if constexpr (FMT == kFmtFP8) pre_e = y < yend ? (int)a.yscale[y] : 127;
    }
  };
  auto store_tile = [&](int buf) {
#pragma unroll
    for (int i = 0; i < CPT; ++i) {
      const int ch = tid + i * kThreads;
      const int row = ch / (DM * F::esize / 16);
      const int col = ch % (DM * F::esize / 16);
      if (ch < CHUNKS) *reinterpret_cast<i32x4*>(&ytile[buf][row * PITCH + col * 16]) = pre[i];
    }
    if (tid < kBY) {
      yn_s[buf][tid] = pre_n;
      ye_s[buf][tid] = pre_e;
    }
  };

  load_tile(ybeg);
  store_tile(0);
  __syncthreads();

  int cur = 0, step = 0;
  for (int64_t y0 = ybeg; y0 < yend; y0 += kBY, ++step) {
    const bool more = y0 + kBY < yend;
    if (more) load_tile(y0 + kBY);

    // C operand: |y|^2 of this lane's 16 rows (4 runs of 4 consecutive rows).
    f32x16 acc[QG];
    {
      f32x16 yn;
#pragma unroll
      for (int r4 = 0; r4 < 4; ++r4) {
        const f32x4 v = *reinterpret_cast<const f32x4*>(&yn_s[cur][r4 * 8 + h * 4]);
        yn[r4 * 4 + 0] = v[0];
        yn[r4 * 4 + 1] = v[1];
        yn[r4 * 4 + 2] = v[2];
        yn[r4 * 4 + 3] = v[3];
      }
#pragma unroll
      for (int g = 0; g < QG; ++g) acc[g] = yn;
    }
    const unsigned char* arow = &ytile[cur][c * PITCH + h * (DM / 2) * F::esize];
    if constexpr (FMT == kFmtF32) {
#pragma unroll
      for (int s = 0; s < NSTEP; s += 4) {
        const f32x4 av = *reinterpret_cast<const f32x4*>(arow + s * 4);
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
          for (int g = 0; g < QG; ++g)
            acc[g] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[e], xr[g][s + e], acc[g], 0, 0, 0);
      }
    } else if constexpr (FMT == kFmtBF16) {
#pragma unroll
      for (int s = 0; s < NSTEP; ++s) {
        const bf16x8 av = *reinterpret_cast<const bf16x8*>(arow + s * 16);
#pragma unroll
        for (int g = 0; g < QG; ++g)
          acc[g] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, xr[g][s], acc[g], 0, 0, 0);
      }
    } else {
      // Row scales ride in the MFMA (E8M0 per 32-element block): A's block is
      // one train row, B's one query, so D is already |y|^2 - 2 x.y.
      const int sa = ye_s[cur][c];
#pragma unroll
      for (int s = 0; s < NSTEP; ++s) {
        const i32x4 lo = *reinterpret_cast<const i32x4*>(arow + s * 32);
        const i32x4 hi = *reinterpret_cast<const i32x4*>(arow + s * 32 + 16);
        const i32x8 av = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
#pragma unroll
        for (int g = 0; g < QG; ++g)
          acc[g] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, xr[g][s], acc[g], 0, 0, 0,
                                                                   sa, 0, xsc[g]);
      }
    }

    // ---- selection ----------------------------------------------------------
    bool overflow = false;
#pragma unroll
    for (int g = 0; g < QG; ++g) {
      float m = acc[g][0];
#pragma unroll
      for (int i = 1; i < 16; ++i) m = fminf(m, acc[g][i]);
      if (__any(m <= tau[g])) {
        const int ql = wq0 + g * 32 + c;
        float* cd = cand_d + (size_t)ql * a.cap;
        int* ci = cand_i + (size_t)ql * a.cap;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const float v = acc[g][i];
          if (v <= tau[g]) {
            const int pos = atomicAdd(&cnt_s[ql], 1);
            cd[pos] = v;
            ci[pos] = (int)(y0 + (i >> 2) * 8 + h * 4 + (i & 3));
            overflow |= pos + 1 > a.cap - kBY;
          }
        }
      }
    }
    if (__any(overflow)) {
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
      for (int qq = 0; qq < WQ; ++qq) {
        const int ql = wq0 + qq;
        const int n = cnt_s[ql];
        if (n <= a.cap - kBY) continue;
        float* cd = cand_d + (size_t)ql * a.cap;
        int* ci = cand_i + (size_t)ql * a.cap;
        wave_sort_cut(cd, ci, n, a.kc, sd, si, lane);
        if (lane == 0) {
          cnt_s[ql] = n < a.kc ? n : a.kc;
          if (n >= a.kc) {
            const float kth = sd[a.kc - 1];
            if (kth < tau_s[ql]) tau_s[ql] = kth;
            atomicMin(&a.tau[q0 + ql], ordered_key((double)kth));
          }
        }
        wave_sync();
      }
#pragma unroll
      for (int g = 0; g < QG; ++g) tau[g] = fminf(tau[g], tau_s[wq0 + g * 32 + c]);
    }
    if ((step & 15) == 15) {
#pragma unroll
      for (int g = 0; g < QG; ++g) {
        const int64_t q = q0 + wq0 + g * 32 + c;
        if (q < a.n_q) {
          const float t =
              (float)from_ordered_key(__atomic_load_n(&a.tau[q], __ATOMIC_RELAXED));
          tau[g] = fminf(tau[g], t);
        }
      }
    }

    if (more) store_tile(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }

  // ---- final: sort and cut every list of this wave, write the split slot ----
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  for (int qq = 0; qq < WQ; ++qq) {
    const int ql = wq0 + qq;
    const int n = cnt_s[ql];
    if (n <= 0) continue;
    wave_sort_cut(cand_d + (size_t)ql * a.cap, cand_i + (size_t)ql * a.cap, n, a.kc, sd, si,
                  lane);
    if (lane == 0) cnt_s[ql] = n < a.kc ? n : a.kc;
    wave_sync();
  }
  for (int qq = 0; qq < WQ; ++qq) {
    const int ql = wq0 + qq;
    const int64_t q = q0 + ql;
    if (q >= a.n_q) break;
    const int n = cnt_s[ql];
    for (int j = lane; j < a.kc; j += 64) {
      const size_t o = ((size_t)q * a.n_splits + split) * a.kc + j;
      a.part_d[o] = j < n ? cand_d[(size_t)ql * a.cap + j] : __builtin_huge_valf();
      a.part_i[o] = j < n ? cand_i[(size_t)ql * a.cap + j] : -1;
    }
  }
}

template <int FMT, int DM, int QG>
void launch_one(const MfmaArgs& a, hipStream_t s) {
  const size_t lds = (size_t)4 * a.cap * 8;
  hipLaunchKernelGGL((mfma_search_kernel<FMT, DM, QG>), dim3((unsigned)(a.n_qblocks * a.n_splits)),
                     dim3(kThreads), lds, s, a);
}

}  // namespace

static_assert(kMfmaQueriesPerBlock == 4 * 2 * 32, "instantiations use 2 query groups per wave");
static_assert(kMfmaRowStep == kBY, "row step");

bool launch_mfma_search(const MfmaArgs& a, hipStream_t s) {
  if (a.fmt == kFmtF32) {
    switch (a.dm) {
      case 32: launch_one<kFmtF32, 32, 2>(a, s); return true;
      case 64: launch_one<kFmtF32, 64, 2>(a, s); return true;
      case 128: launch_one<kFmtF32, 128, 2>(a, s); return true;
      default: return false;
    }
  }
  if (a.fmt == kFmtFP8) {
    switch (a.dm) {
      case 64: launch_one<kFmtFP8, 64, 2>(a, s); return true;
      case 128: launch_one<kFmtFP8, 128, 2>(a, s); return true;
      case 256: launch_one<kFmtFP8, 256, 2>(a, s); return true;
      default: return false;
    }
  }
  if (a.fmt == kFmtBF16) {
    switch (a.dm) {
      case 32: launch_one<kFmtBF16, 32, 2>(a, s); return true;
      case 64: launch_one<kFmtBF16, 64, 2>(a, s); return true;
      case 128: launch_one<kFmtBF16, 128, 2>(a, s); return true;
      case 256: launch_one<kFmtBF16, 256, 2>(a, s); return true;
      default: return false;
    }
  }
  return false;
}

}  // namespace knn
