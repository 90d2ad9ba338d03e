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
//   * train rows stream global -> LDS by direct-to-LDS buffer loads (three
//     buffers, one wait + barrier per 32-row step) and are read back in the
//     A-operand layout;
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

__device__ inline float fmin_ieee(float a, float b) { return __builtin_elementwise_minimum(a, b); }

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

// Direct-to-LDS buffer loads (lane l's bytes land at lds_addr + size * l),
// written out so that the compiler does not order its own LDS reads behind
// them: the sweep waits for them itself (vmcnt, then a barrier).  M0 takes
// the LDS address, one wait state before the load.  The compiler warns that
// it may not preserve a clobbered M0: nothing it generates for these kernels
// keeps a value in M0 (gfx9 LDS instructions do not use it; the generated
// code has no M0 reference outside these statements), so none is lost.
__device__ inline unsigned lds_address(const void* p) {
  return (unsigned)(uintptr_t)(const __attribute__((address_space(3))) void*)p;
}
// The LDS address is a scalar base plus an immediate, so the per-buffer
// addresses need no registers of their own.
template <int IMM>
__device__ inline void lds_load16(unsigned lds_base, int voff, i32x4 rsrc) {
  asm volatile("s_add_u32 m0, %0, %3\n\ts_nop 0\n\tbuffer_load_dwordx4 %1, %2, 0 offen lds"
               ::"s"(lds_base), "v"(voff), "s"(rsrc), "i"(IMM) : "m0", "scc");
}
template <int IMM>
__device__ inline void lds_load4(unsigned lds_base, int voff, i32x4 rsrc) {
  asm volatile("s_add_u32 m0, %0, %3\n\ts_nop 0\n\tbuffer_load_dword %1, %2, 0 offen lds"
               ::"s"(lds_base), "v"(voff), "s"(rsrc), "i"(IMM) : "m0", "scc");
}
// Buffer resource (raw, stride 0, range-checked against `bytes`).
__device__ inline i32x4 make_rsrc(const void* base, int bytes) {
  const uint64_t b = (uint64_t)(uintptr_t)base;
  return i32x4{__builtin_amdgcn_readfirstlane((int)(unsigned)b),
               __builtin_amdgcn_readfirstlane((int)((b >> 32) & 0xffff)),
               __builtin_amdgcn_readfirstlane(bytes), 0x00020000};
}
// The same resource moved `off` bytes on (scalar arithmetic only): the range
// shrinks by as much, down to zero, so the hardware check still ends at the
// original end.
__device__ inline i32x4 rsrc_advance(i32x4 r, int off) {
  const uint64_t b = ((uint64_t)(uint32_t)r[1] << 32 | (uint32_t)r[0]) + (uint32_t)off;
  const int left = r[2] > off ? r[2] - off : 0;
  return i32x4{(int)(uint32_t)b, (int)(uint32_t)(b >> 32), left, r[3]};
}

// Tiles reach LDS by direct-to-LDS buffer loads (each lane's 16 bytes land at
// the wave's LDS base + 16 * lane), three buffers deep.  Rows are not padded;
// instead 16-byte chunk j of row r sits at chunk position j ^ swz(r).  The
// LDS has 64 banks of 4 bytes (256 bytes, 16 chunks), and a 16-byte read
// serves 16 lanes at a time: the A-operand reads of rows r..r+15 (same
// logical chunk) land on 16 different chunk columns when swz(r) is r's index
// among the rows that share a column offset, i.e. r / R with R = 16 / NCH
// rows per 256 bytes (R = 1 for rows of 256 bytes or more).
template <int FMT, int DM>
struct TileGeom {
  static constexpr int ROWB = DM * Fmt<FMT>::esize;  // bytes per train row
  static constexpr int NCH = ROWB / 16;               // 16-byte chunks per row
  static constexpr int R = NCH < 16 ? 16 / NCH : 1;
  static constexpr int MASK = (NCH < 16 ? NCH : 16) - 1;
  __device__ static int swz(int r) { return (r / R) & MASK; }
  // j ^ swz stays in the row: NCH is a power of two below 16, else a
  // multiple of 16 (the swizzle flips bits inside an aligned run of 16).
  static_assert(NCH < 16 ? (NCH & (NCH - 1)) == 0 : NCH % 16 == 0, "row chunks");
};
constexpr int kNBuf = 3;

// fn(std::integral_constant<int, I>()) for I = 0 .. N-1.
template <typename Fn, int... I>
__device__ inline void static_for_impl(Fn& fn, std::integer_sequence<int, I...>) {
  (fn(std::integral_constant<int, I>()), ...);
}
template <int N, typename Fn>
__device__ inline void static_for(Fn&& fn) {
  static_for_impl(fn, std::make_integer_sequence<int, N>());
}

// Wait until at most K vector-memory operations of this wave are
// outstanding (lgkm: none), then the workgroup barrier.  vmcnt takes an
// immediate, so a run-time count picks one of 16 forms; counts above B+15
// wait for B+15, i.e. a little longer than needed.
template <int K>
__device__ inline void wait_barrier() {
  asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(0)\n\ts_barrier" ::"n"(K) : "memory");
}
template <int B, int... I>
__device__ inline void wait_barrier_upto(int n, std::integer_sequence<int, I...>) {
  const int m = __builtin_amdgcn_readfirstlane(n < B ? B : (n > B + 15 ? B + 15 : n));
  ((m == B + I ? wait_barrier<B + I>() : void()), ...);
}

// MODE kModeSearch: the candidate search.  MODE kModeResearch: the same for
// the query blocks / queries selected by qblock_active / query_active (a
// separate instantiation, so profiles tell the two launches apart).  MODE
// kModeSample: the same sweep over sampled row ranges, keeping only each
// lane's kSampleJ smallest per-tile minima in registers (a threshold
// estimate, no lists).  MODE kModeCollect: the selected queries keep their
// threshold fixed (tau[q], set by the merge) and every row at or below it is
// gathered into the query's global collect list: no compaction, no threshold
// refresh, no per-split lists.
constexpr int kModeSearch = 0;
constexpr int kModeSample = 1;
constexpr int kModeResearch = 2;
constexpr int kModeCollect = 3;

template <int FMT, int DM, int QG, int MODE>
__global__ __launch_bounds__(kThreads, FMT != kFmtF32 ? 2 : DM > 128 ? 1 : 3) void mfma_search_kernel(MfmaArgs a) {
  using F = Fmt<FMT>;
  using G = TileGeom<FMT, DM>;
  constexpr int WQ = QG * 32;           // queries per wave
  constexpr int BQ = 4 * WQ;            // queries per workgroup
  constexpr int NSTEP = DM / (2 * F::KL);  // MFMAs per 32x32 tile
  constexpr int CHUNKS = kBY * DM * F::esize / 16;      // 16-byte loads per tile
  constexpr int CPT = (CHUNKS + kThreads - 1) / kThreads;  // per thread

  // Three tile buffers as distinct objects: the compiler can tell the loads
  // into one from the reads of another, and waits only where this code does.
  struct alignas(16) TileBuf {
    unsigned char y[kBY * G::ROWB];
    float n[64];                                  // |y|^2 of the tile's rows (32 used)
    unsigned char e[FMT == kFmtFP8 ? 256 : 16];   // FP8: E8M0 row scales (32 used)
  };
  __shared__ TileBuf tbuf[kNBuf];
  __shared__ int cnt_s[2 * BQ];
  __shared__ float tau_s[BQ];
  extern __shared__ __align__(16) unsigned char dyn_smem[];  // 4 waves x cap x 8 B

  const int tid = threadIdx.x, lane = tid & 63, h = lane >> 5, c = lane & 31;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave index (uniform)
  // Lets the compiler drop the per-load range tests of waves that always
  // load (two scalar branches per tile load at BF16 d = 128).
  __builtin_assume(w >= 0 && w < kThreads / 64);
  int n_stores = 0;  // candidate store instructions this wave issued (uniform)

  // Workgroup -> (query block, split).  Workgroups are dealt to the 8 XCDs
  // round-robin; work items are numbered split-major and each XCD takes a
  // contiguous range of them, so the query blocks of one split run together
  // on one XCD and share its L2 while they stream the same train rows.
  // Packed blocks (collect launch only): slot q0 + i holds query
  // qmap[q0 + i], while fewer than half of all queries are listed (the count
  // is the same for every workgroup, so all take the same layout); with more,
  // the unpacked blocks and their flags, which keep neighbouring queries
  // together (FP8 on its 1,000-Gaussian mixture, every query widened: 0.7 ms
  // slower packed).
  const int* qmap = MODE == kModeCollect ? a.qmap : nullptr;
  const int64_t n_mapped = qmap ? (int64_t)__builtin_amdgcn_readfirstlane(*a.qmap_n) : 0;
  if (qmap && 2 * n_mapped >= a.n_q) qmap = nullptr;
  // Unpacked collect launch: the grid is sized for short splits (the latency
  // of a few packed blocks), so consecutive groups of coll_group splits go
  // to one workgroup each, which keeps the query loads per row what the
  // search's are (the leading workgroups, dealt to the XCDs as below).
  const int group = MODE == kModeCollect && !qmap && a.coll_group > 1 ? a.coll_group : 1;
  int qb, split, nsplit = 1;
  int nqb = a.n_qblocks;  // query blocks of this launch's layout
  int64_t stride = a.split_stride, rows = a.rows_per_split;
  {
    const int b = blockIdx.x;
    int nsp = a.n_splits;
    if (MODE == kModeCollect && qmap) {
      // Packed: only the blocks of listed queries, each over as many splits
      // as the collect scratch holds (pk_items blocks' lists) and one round
      // of workgroups takes (pk_slots), down to pk_splits of them, never
      // fewer than the unpacked layout's: a few listed queries cost a short
      // sweep.  n_mapped is the same for every
      // workgroup, so all take the same layout (the host sizes the grid for
      // the largest).
      nqb = (int)((n_mapped + BQ - 1) / BQ);
      if (nqb == 0) return;
      int sp = a.pk_items / nqb;
      sp = sp < a.pk_splits ? sp : a.pk_splits;
      const int one_round = a.pk_slots / nqb;  // more splits would queue a second round
      sp = sp < one_round ? sp : one_round;
      sp = sp > a.n_splits ? sp : a.n_splits;
      rows = ((a.n_t + sp - 1) / sp + kMfmaRowStep - 1) / kMfmaRowStep * kMfmaRowStep;
      stride = rows;
      nsp = (int)((a.n_t + rows - 1) / rows);
    }
    const int groups = (nsp + group - 1) / group;
    const int items = nqb * groups;
    const int grid = (items + 7) & ~7;  // the launch's own grid, or a part of it
    if (b >= grid) return;  // past this layout's grid (whole workgroup)
    const int per_xcd = grid >> 3;
    const int w_item = (b & 7) * per_xcd + (b >> 3);
    if (w_item >= items) return;  // grid padding (whole workgroup)
    split = (w_item / nqb) * group;
    qb = w_item % nqb;
    nsplit = nsp - split < group ? nsp - split : group;
  }
  constexpr bool kSelected = MODE == kModeResearch || MODE == kModeCollect;
  const int64_t q0 = (int64_t)qb * BQ;
  if (qmap ? q0 >= n_mapped : kSelected && a.qblock_active[qb] == 0) return;  // not selected
  // Query of local slot ql (0 <= ql < BQ), or -1 for an empty packed slot.
  auto query_of = [&](int ql) -> int64_t {
    if (!qmap) return q0 + ql;
    return q0 + ql < n_mapped ? (int64_t)qmap[q0 + ql] : -1;
  };
  // A group of splits is contiguous (split_stride == rows_per_split there).
  const int64_t ybeg = (int64_t)split * stride;
  const int64_t ylen = (int64_t)(nsplit - 1) * stride + rows;
  const int64_t yend = ybeg + ylen < a.n_t ? ybeg + ylen : a.n_t;
  const size_t blk = (size_t)split * nqb + qb;
  // Candidate lists of the search launches: (value bits, row) pairs, one
  // 8-byte store per survivor (two 4-byte stores cost about 0.13 ms more at
  // 1e4 x 1e6 x 128).  The collect launch keeps row indices only, in cand_i.
  uint2* cand = reinterpret_cast<uint2*>(a.cand_d) + blk * BQ * a.cap;
  int* cand_i = a.cand_i + blk * BQ * a.cap;
  float* sd = reinterpret_cast<float*>(dyn_smem) + (size_t)w * 2 * a.cap;
  int* si = reinterpret_cast<int*>(sd + a.cap);
  const int wq0 = w * WQ;

  // ---- tile loader: global -> LDS (direct) ---------------------------------
  // Raw buffer loads over this split's rows: rows past the split read as
  // zero (range-checked by the hardware), so no load needs a branch.  Splits
  // hold whole tiles except the last one of the train set, whose rows past
  // n_t take |y|^2 = +inf from the padding after the norms (kNormPad
  // entries): their values are +inf, never selected.  The look-ahead tiles
  // past a split are loaded but never used.
  const int row_bytes = a.ldy * F::esize;
  const int64_t nrows = yend - ybeg;
  const i32x4 yrs = make_rsrc(reinterpret_cast<const unsigned char*>(a.Y) +
                                  ybeg * (int64_t)row_bytes,
                              (int)(nrows * row_bytes));
  const i32x4 nrs = make_rsrc(a.ynorm + ybeg, (int)((nrows + 3 * kBY) * 4));
  i32x4 ers = nrs;
  if constexpr (FMT == kFmtFP8) ers = make_rsrc(a.yscale + ybeg, (int)nrows);
  // Load instruction i of wave w fills LDS chunks [(4i + w) * 64, + 64); lane
  // l brings the chunk for position (4i + w) * 64 + l.  Each tile's loads go
  // through resources moved to the tile (scalar work only), so a lane's
  // offsets are the same for every tile.  Float32 features past d come from
  // an offset past the buffer, i.e. zero.
  constexpr int kOOB = 0x7ffffff0;
  int voff[CPT];
#pragma unroll
  for (int i = 0; i < CPT; ++i) {
    const int slot = (i * 4 + w) * 64 + lane;
    const int r = slot / G::NCH, j = (slot % G::NCH) ^ G::swz(r);
    voff[i] = (FMT == kFmtF32 && j * 4 >= a.d) ? kOOB : r * row_bytes + j * 16;
  }
  const int noff = lane * 4, eoff = lane * 4;
  // LDS byte address of this wave's first chunk in buffer 0.
  const unsigned lds_w = lds_address(&tbuf[0].y[0]) + (unsigned)w * 64 * 16;
  constexpr int kBufBytes = (int)sizeof(TileBuf);
  constexpr int kNOff = (int)offsetof(TileBuf, n), kEOff = (int)offsetof(TileBuf, e);
  auto load_tile = [&](int t, auto buf_c) {  // tile t of this split -> tbuf[buf_c]
    constexpr int B = decltype(buf_c)::value;
    const int tb = t * kBY;
    const i32x4 yt = rsrc_advance(yrs, tb * row_bytes);
    static_for<CPT>([&](auto ic) {
      constexpr int I = decltype(ic)::value;
      if ((I * 4 + w) * 64 >= CHUNKS) return;  // wave-uniform
      lds_load16<B * kBufBytes + I * 4 * 64 * 16>(lds_w, voff[I], yt);
    });
    if (w == 0)  // 64 norms: the tile's 32 rows and 32 unused
      lds_load4<B * kBufBytes + kNOff>(lds_w, noff, rsrc_advance(nrs, tb * 4));
    if constexpr (FMT == kFmtFP8)
      if (w == 1)  // 4 scale bytes per lane: the tile's 32 in lanes 0..7
        lds_load4<B * kBufBytes + kEOff - 1 * 64 * 16>(lds_w, eoff, rsrc_advance(ers, tb));
  };
  // The first two tiles are in flight while the queries load.
  load_tile(0, std::integral_constant<int, 0>());
  load_tile(1, std::integral_constant<int, 1>());

  // ---- resident queries: B operand, scaled by -2 ------------------------------
  typename F::breg xr[QG][NSTEP];
  int xsc[QG];  // FP8: E8M0 scale of -2 x (the row scale times 2)
  float tau[QG];
  bool qact[QG];  // query exists and takes part in this launch
#pragma unroll
  for (int g = 0; g < QG; ++g) {
    const int64_t q = query_of(wq0 + g * 32 + c);
    const bool qok = q >= 0 && q < a.n_q;
    // (collect launch: flag 1; flag 2 marks widened queries whose set comes
    // from the search's lists)
    qact[g] = qok && (!kSelected || (MODE == kModeCollect ? a.query_active[q] == 1
                                                          : a.query_active[q] != 0));
    tau[g] = qact[g] ? FLT_MAX : -__builtin_huge_valf();
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
  // Per-query candidate list = two halves of cap/2 entries; lane (c, h)
  // appends to half h only, so its count lives in a register and appends
  // need no atomics.  A step adds at most 16 entries per half.
  const int half = a.cap >> 1;
  const int thr = half - 16;
  int cnt[QG];
#pragma unroll
  for (int g = 0; g < QG; ++g) {
    cnt[g] = 0;
    const int64_t q = query_of(wq0 + g * 32 + c);
    if (MODE != kModeSample && qact[g])  // start from the best threshold known
      tau[g] = fminf(tau[g],
                     (float)from_ordered_key(load_agent(&a.tau[q])));
  }
  for (int i = tid; i < BQ; i += kThreads) tau_s[i] = FLT_MAX;

  // Loads each wave issues per tile, counting only what every wave issues:
  // waiting until at most that many are outstanding completes every load of
  // the tile before (in-order completion), then the barrier makes all waves'
  // loads of that tile visible and frees the buffer read in the previous step.
  // Wave w issues the load instructions of chunk groups w, w+4, ... of the
  // CHUNKS/64 groups: the fewest, CHUNKS/256, are wave 3's.
  constexpr int kMinLoads = CHUNKS / 256;
  static_assert(CHUNKS % 64 == 0 && kMinLoads < 16, "tile load accounting");
  // A-operand chunk addresses of this lane (row c), fixed for the sweep.
  constexpr int NFRAG = FMT == kFmtF32 ? NSTEP / 4 : FMT == kFmtBF16 ? NSTEP : 2 * NSTEP;
  int aoff[NFRAG];
#pragma unroll
  for (int f = 0; f < NFRAG; ++f) {
    const int j = h * (G::NCH / 2) + f;
    aoff[f] = (c * G::NCH + (j ^ G::swz(c))) * 16;
  }

  // ---- one 32-row tile on the matrix cores ------------------------------------
  // acc = |y|^2 (C operand of the first MFMA) - 2 x.y
  auto compute = [&](f32x16(&acc)[QG], const TileBuf& b) {
    f32x16 yn;
#pragma unroll
    for (int r4 = 0; r4 < 4; ++r4) {
      const f32x4 v = *reinterpret_cast<const f32x4*>(&b.n[r4 * 8 + h * 4]);
      yn[r4 * 4 + 0] = v[0];
      yn[r4 * 4 + 1] = v[1];
      yn[r4 * 4 + 2] = v[2];
      yn[r4 * 4 + 3] = v[3];
    }
#pragma unroll
    for (int g = 0; g < QG; ++g) acc[g] = yn;
    const unsigned char* tile = &b.y[0];
    if constexpr (FMT == kFmtF32) {
#pragma unroll
      for (int s = 0; s < NSTEP; s += 4) {
        const f32x4 av = *reinterpret_cast<const f32x4*>(tile + aoff[s / 4]);
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
          for (int g = 0; g < QG; ++g)
            acc[g] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[e], xr[g][s + e], acc[g], 0, 0, 0);
      }
    } else if constexpr (FMT == kFmtBF16) {
      // Fragment reads three steps ahead of their MFMAs, kept there by
      // scheduling barriers (left alone, the scheduler moves each read next
      // to its first use and the MFMAs wait on LDS latency).
      constexpr int PF = 3 < NSTEP ? 3 : NSTEP;
      bf16x8 af[NSTEP];
#pragma unroll
      for (int s = 0; s < PF; ++s) af[s] = *reinterpret_cast<const bf16x8*>(tile + aoff[s]);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int s = 0; s < NSTEP; ++s) {
        if (s + PF < NSTEP) af[s + PF] = *reinterpret_cast<const bf16x8*>(tile + aoff[s + PF]);
#pragma unroll
        for (int g = 0; g < QG; ++g)
          acc[g] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[s], xr[g][s], acc[g], 0, 0, 0);
        __builtin_amdgcn_sched_barrier(0);
      }
    } else {
      // Row scales ride in the MFMA (E8M0 per 32-element block): A's block is
      // one train row, B's one query, so D is already |y|^2 - 2 x.y.
      const int sa = b.e[c];
#pragma unroll
      for (int s = 0; s < NSTEP; ++s) {
        const i32x4 lo = *reinterpret_cast<const i32x4*>(tile + aoff[2 * s]);
        const i32x4 hi = *reinterpret_cast<const i32x4*>(tile + aoff[2 * s + 1]);
        const i32x8 av = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
#pragma unroll
        for (int g = 0; g < QG; ++g)
          acc[g] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, xr[g][s], acc[g], 0, 0, 0,
                                                                   sa, 0, xsc[g]);
      }
    }
  };

  // ---- list maintenance (wave-local) ----------------------------------------
  // Sort the two halves of query ql together, keep the best min(n, kc).
  auto sort_query = [&](int ql, int n0, int n1) {
    const uint2* cp = cand + (size_t)ql * a.cap;
    const int n = n0 + n1;
    const int P = next_pow2(n);
    for (int e = lane; e < P; e += 64) {
      const bool ok = e < n;
      const int src = e < n0 ? e : half + (e - n0);
      const uint2 pr = ok ? cp[src] : make_uint2(0x7f800000u, (unsigned)INT_MAX);
      sd[e] = __uint_as_float(pr.x);
      si[e] = (int)pr.y;
    }
    wave_sync();
    for (int size = 2; size <= P; size <<= 1) {
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        for (int i = lane; i < (P >> 1); i += 64) {
          const int lo = ((i & ~(stride - 1)) << 1) | (i & (stride - 1));  // stride: power of 2
          const int hi = lo + stride;
          const bool up = (lo & size) == 0;
          const float dl = sd[lo], dh = sd[hi];
          const int il = si[lo], ih = si[hi];
          const bool swap = up ? key_less(dh, ih, dl, il) : key_less(dl, il, dh, ih);
          if (swap) {
            sd[lo] = dh;
            sd[hi] = dl;
            si[lo] = ih;
            si[hi] = il;
          }
        }
        wave_sync();
      }
    }
  };
  // Compact every query of this wave whose half exceeds `limit`.
  auto compact = [&](int limit) {
#pragma unroll
    for (int g = 0; g < QG; ++g) cnt_s[2 * (wq0 + g * 32 + c) + h] = cnt[g];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    wave_sync();
    for (int qq = 0; qq < WQ; ++qq) {
      const int ql = wq0 + qq;
      const int n0 = cnt_s[2 * ql], n1 = cnt_s[2 * ql + 1];
      if (n0 <= limit && n1 <= limit) continue;
      sort_query(ql, n0, n1);
      const int m = n0 + n1 < a.kc ? n0 + n1 : a.kc;
      const int m0 = (m + 1) >> 1;
      uint2* cp = cand + (size_t)ql * a.cap;
      for (int e = lane; e < m; e += 64) {
        const int dst = e < m0 ? e : half + (e - m0);
        cp[dst] = make_uint2(__float_as_uint(sd[e]), (unsigned)si[e]);
      }
      if (lane == 0) {
        cnt_s[2 * ql] = m0;
        cnt_s[2 * ql + 1] = m - m0;
        if (m == a.kc) {
          const float kth = sd[a.kc - 1];
          if (kth < tau_s[ql]) tau_s[ql] = kth;
          atomicMin(&a.tau[q0 + ql], ordered_key((double)kth));
        }
        // Publish this split's (slot_rank+1)-th smallest value (see select()).
        if (a.qslots && m > a.slot_rank)
          __atomic_store_n(reinterpret_cast<unsigned*>(a.qslots) + (size_t)(q0 + ql) * a.slot_stride +
                               split,
                           __float_as_uint(sd[a.slot_rank]), __ATOMIC_RELAXED);
      }
      wave_sync();
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
#pragma unroll
    for (int g = 0; g < QG; ++g) {
      const int ql = wq0 + g * 32 + c;
      cnt[g] = cnt_s[2 * ql + h];
      tau[g] = fminf(tau[g], tau_s[ql]);
    }
  };
  // Collect launch: move the entries of every query of this wave with a half
  // above `limit` to the query's global collect list (one atomic per query
  // and flush reserves the slots) and empty its halves.
  auto flush = [&](int limit) {
#pragma unroll
    for (int g = 0; g < QG; ++g) cnt_s[2 * (wq0 + g * 32 + c) + h] = cnt[g];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    wave_sync();
    for (int qq = 0; qq < WQ; ++qq) {
      const int ql = wq0 + qq;
      const int n0 = cnt_s[2 * ql], n1 = cnt_s[2 * ql + 1];
      if (n0 <= limit && n1 <= limit) continue;  // uniform across the wave
      // Only queries of the launch collect (a padding lane past n_q, threshold
      // -inf, appends only -inf values, i.e. never for finite inputs).
      const int64_t q = query_of(ql);
      if (q >= 0 && q < a.n_q) {
        int base = 0;
        if (lane == 0) base = atomicAdd(&a.coll_cnt[q], n0 + n1);
        base = __shfl(base, 0, 64);
        const int* ci = cand_i + (size_t)ql * a.cap;
        int* dst = a.coll_idx + (size_t)q * a.coll_cap;
        for (int e = lane; e < n0 + n1; e += 64) {
          const int src = e < n0 ? e : half + (e - n0);
          if (base + e < a.coll_cap) dst[base + e] = ci[src];
        }
      }
      if (lane == 0) {
        cnt_s[2 * ql] = 0;
        cnt_s[2 * ql + 1] = 0;
      }
      wave_sync();
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
#pragma unroll
    for (int g = 0; g < QG; ++g) cnt[g] = cnt_s[2 * (wq0 + g * 32 + c) + h];
  };
  // Appends go through buffer stores on this workgroup's candidate block, so
  // the address of an append is one 32-bit offset.
  // Appends are non-temporal (nt): they are read back only at compactions and
  // at the end of the split, and each would otherwise take an L2 line from
  // the row stream (search reads 0.72 -> 0.65 GB per launch at 1e4 x 1e6 x 128).
  constexpr int kAppendNT = 2;
  const __amdgpu_buffer_rsrc_t cprs =
      __builtin_amdgcn_make_buffer_rsrc(cand, (short)0, BQ * a.cap * 8, 0x00020000);
  const __amdgpu_buffer_rsrc_t cirs =
      __builtin_amdgcn_make_buffer_rsrc(cand_i, (short)0, BQ * a.cap * 4, 0x00020000);
  // Selection of one finished tile.  Lane values come in 4 quads of 4
  // consecutive train rows; the common case costs a min-tree and one
  // compare, a tile with survivors tests the 4 quad minima and then only the
  // values of quads that hold a survivor.
  auto select = [&](f32x16(&acc)[QG], int64_t y0) {
    bool overflow = false;
    // IEEE minimum (NaN-propagating, v_minimum3_f32 on gfx950: no operand
    // canonicalisation).  A NaN makes a quad "not above tau", so its values
    // are tested one by one and a NaN hides no other row.  One vote covers
    // every query group of the wave (the common case is no survivor at all).
    float m4q[QG][4];
    bool hot[QG], any_hot = false;
#pragma unroll
    for (int g = 0; g < QG; ++g) {
#pragma unroll
      for (int r4 = 0; r4 < 4; ++r4)
        m4q[g][r4] = fmin_ieee(fmin_ieee(acc[g][4 * r4], acc[g][4 * r4 + 1]),
                               fmin_ieee(acc[g][4 * r4 + 2], acc[g][4 * r4 + 3]));
      const float m = fmin_ieee(fmin_ieee(m4q[g][0], m4q[g][1]), fmin_ieee(m4q[g][2], m4q[g][3]));
      hot[g] = !(m > tau[g]);
      any_hot |= hot[g];
    }
    if (!__any(any_hot)) return;
#pragma unroll
    for (int g = 0; g < QG; ++g) {
      const float* m4 = m4q[g];
      if (__any(hot[g])) {
        // Wave-wide votes (independent compares into scalar masks) pick the
        // quads, then the values, that hold a survivor in any lane; a value
        // with one is stored by every lane, the lanes without one at an
        // offset past the buffer, which the range check drops (no divergent
        // branch).
        const int ql = wq0 + g * 32 + c;
        const int base = ql * a.cap + h * half;  // entries
        const int yb = (int)y0 + h * 4;
        int n = cnt[g];
        uint64_t qm[4];
#pragma unroll
        for (int r4 = 0; r4 < 4; ++r4) qm[r4] = __builtin_amdgcn_ballot_w64(!(m4[r4] > tau[g]));
#pragma unroll
        for (int r4 = 0; r4 < 4; ++r4) {
          if (qm[r4] == 0) continue;
          uint64_t vm[4];
#pragma unroll
          for (int e = 0; e < 4; ++e) vm[e] = __builtin_amdgcn_ballot_w64(acc[g][4 * r4 + e] <= tau[g]);
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            if (vm[e] == 0) continue;
            const float v = acc[g][4 * r4 + e];
            const bool hit = v <= tau[g];
            if constexpr (MODE == kModeCollect) {  // a collect list keeps indices only
              __builtin_amdgcn_raw_buffer_store_b32(yb + r4 * 8 + e, cirs,
                                                    hit ? (base + n) * 4 : kOOB, 0, 0);
            } else {
              typedef int i32x2 __attribute__((ext_vector_type(2)));
              const i32x2 pr = {(int)__float_as_uint(v), yb + r4 * 8 + e};
              __builtin_amdgcn_raw_buffer_store_b64(pr, cprs, hit ? (base + n) * 8 : kOOB, 0,
                                                    kAppendNT);
            }
            n += hit;
            n_stores += 1;
          }
        }
        cnt[g] = n;
        overflow |= n > thr;
      }
    }
    if (__any(overflow)) {
      if constexpr (MODE == kModeCollect)
        flush(thr);
      else
        compact(thr);
    }
  };
  // Threshold refresh (search launches; the collect launch keeps its fixed
  // threshold), called by the sweep after ring cycles 1, 2, 5 and then every
  // 11 (every 33 rows steps: cheap early sharing, rare later) when splits
  // publish slots, else after cycles 1, 3, 7, 15, ...: with a sampled guess
  // only a compaction that reaches k' entries lowers the shared threshold,
  // and each refresh waits for the wave's tile loads in flight (the load's
  // wait cannot leave them out), which at 1e4 x 1e6 x 128 cost 1.6 % of the
  // search at the denser schedule.
  auto refresh = [&]() {
    if constexpr (MODE == kModeSearch || MODE == kModeResearch) {
      // Thresholds shared across the splits of a query:
      //  * the best k'-th value any split has reached (atomicMin key);
      //  * slot s holds split s's (slot_rank+1)-th smallest value, so the
      //    kSlotJ-th smallest slot value T has at least
      //    kSlotJ * (slot_rank+1) >= k' rows with v <= T: T bounds the
      //    global k'-th value.  Sorted insertion keeps the kSlotJ smallest.
#pragma unroll
      for (int g = 0; g < QG; ++g) {
        const int64_t q = q0 + wq0 + g * 32 + c;
        if (qact[g]) {
          float t = (float)from_ordered_key(load_agent(&a.tau[q]));
          if (a.qslots) {
            // Rows of slot_stride (a multiple of 8, padding +inf): eight
            // loads in flight per round trip.
            float r[kSlotJ];
#pragma unroll
            for (int j = 0; j < kSlotJ; ++j) r[j] = __builtin_huge_valf();
            const float* sl = a.qslots + (size_t)q * a.slot_stride;
            for (int s0 = 0; s0 < a.slot_stride; s0 += 8) {
              float x[8];
#pragma unroll
              for (int u = 0; u < 8; ++u) x[u] = load_agent(sl + s0 + u);
#pragma unroll
              for (int u = 0; u < 8; ++u) {
#pragma unroll
                for (int j = kSlotJ - 1; j > 0; --j) {
                  const float hi = r[j - 1] > x[u] ? r[j - 1] : x[u];
                  r[j] = r[j] < hi ? r[j] : hi;
                }
                r[0] = r[0] < x[u] ? r[0] : x[u];
              }
            }
            t = t < r[kSlotJ - 1] ? t : r[kSlotJ - 1];
          }
          tau[g] = fminf(tau[g], t);
        }
      }
    }
  };

  // ---- main sweep --------------------------------------------------------
  // Step t (tile t in LDS buffer t%3, tile t+1 in flight):
  //   wait for tile t + barrier; issue the loads of tile t+2 into the buffer
  //   read in step t-1; MFMAs of tile t; select tile t.
  // The loads are issued unconditionally (tiles past the split read as zero
  // through the buffer range check), so every step issues the same number of
  // them and the wait for tile t leaves kMinLoads later loads in flight.
  // Offsets stay below 2^31: a split holds at most 1 GiB of rows, plus the 95 rows
  // of look-ahead.
  const int ntiles = yend > ybeg ? (int)((yend - ybeg + kBY - 1) / kBY) : 0;
  // One accumulator set: each tile is selected right after its own MFMAs.
  // Keeping a second set to select tile t-1 during tile t's MFMAs costs
  // 16*QG registers, which is the difference between 2 and 3 resident waves
  // per SIMD at bf16 d=128; with 3 the other waves fill the wait for the
  // MFMA results (measured: 3.15 -> 2.48 ms at 1e4 x 1e6 x 128).
  // `sel(acc, y0, step)` consumes one finished tile.
  //
  // Stores count in vmcnt too, and complete in order with the loads: the
  // candidate stores of steps t-2 and t-1 were issued after tile t's loads,
  // so the wait for tile t leaves them (and tile t+1's loads) outstanding
  // instead of waiting for stores issued moments ago.  Only stores counted
  // in n_stores are allowed for; any other load or store issued since only
  // makes the wait longer than needed, never too short.
  auto sweep = [&](auto&& sel, auto&& after_cycle) {
    using B0 = std::integral_constant<int, 0>;
    using B1 = std::integral_constant<int, 1>;
    using B2 = std::integral_constant<int, 2>;
    f32x16 acc[QG];
    // At step t's wait: n_stores when steps t-2 and t-1 began.
    int st_m2 = 0, st_m1 = 0;
    // Wave priority rotates with the ring cycle (3 steps), offset by the
    // wave's slot on its SIMD, so the (up to four) waves sharing a SIMD take
    // turns at issue priority.  Left to age order, the oldest workgroups run
    // ahead and the last ones finish alone on their CUs (at 1e4 x 1e6 x 128
    // the sweeps spread over 1.7 - 2.2 ms instead of 1.9 - 2.1 ms).
    unsigned wslot;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_HW_ID, 0, 4)" : "=s"(wslot));
    auto step = [&](int t, auto cur, auto ahead) {
      // Tile t (loaded two steps ago) complete and visible; `ahead`, read in
      // step t-1, free for tile t+2.
      wait_barrier_upto<kMinLoads>(kMinLoads + n_stores - st_m2,
                                   std::make_integer_sequence<int, 16>());
      st_m2 = st_m1;
      st_m1 = n_stores;
      load_tile(t + 2, ahead);
      compute(acc, tbuf[decltype(cur)::value]);
      sel(acc, ybeg + (int64_t)t * kBY);
    };
    for (int t = 0, cyc = 0; t < ntiles; t += kNBuf, ++cyc) {
      switch ((cyc + (int)wslot) & 3) {
        case 0: __builtin_amdgcn_s_setprio(0); break;
        case 1: __builtin_amdgcn_s_setprio(1); break;
        case 2: __builtin_amdgcn_s_setprio(2); break;
        default: __builtin_amdgcn_s_setprio(3); break;
      }
      step(t, B0(), B2());
      if (t + 1 < ntiles) step(t + 1, B1(), B0());
      if (t + 2 < ntiles) step(t + 2, B2(), B1());
      after_cycle(cyc);
    }
    // No load may still be writing LDS when the workgroup ends.
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  };

  if constexpr (MODE == kModeSample) {
    // Each lane keeps the kSampleJ smallest of its per-tile minima (16 rows
    // of one query per tile), sorted ascending in registers.  Inserting m
    // into the sorted run: b[i] <- min(b[i], max(b[i-1], m)), b[0] <- min(b[0], m).
    float best[QG][kSampleJ];
#pragma unroll
    for (int g = 0; g < QG; ++g)
#pragma unroll
      for (int i = 0; i < kSampleJ; ++i) best[g][i] = __builtin_huge_valf();
    sweep([&](f32x16(&acc)[QG], int64_t) {
#pragma unroll
      for (int g = 0; g < QG; ++g) {
        float m = acc[g][0];  // a NaN value drops this step's minimum (m < b fails)
#pragma unroll
        for (int i = 1; i < 16; ++i) m = fmin_ieee(m, acc[g][i]);
        if (__any(qact[g] && m < best[g][kSampleJ - 1])) {
#pragma unroll
          for (int i = kSampleJ - 1; i > 0; --i)
            best[g][i] = fminf(best[g][i], fmaxf(best[g][i - 1], m));
          best[g][0] = fminf(best[g][0], m);
        }
      }
    }, [](int) {});
#pragma unroll
    for (int g = 0; g < QG; ++g) {
      const int64_t q = q0 + wq0 + g * 32 + c;
      if (!qact[g]) continue;
      float* o = a.sample_out + (((size_t)q * a.n_splits + split) * 2 + h) * kSampleJ;
#pragma unroll
      for (int i = 0; i < kSampleJ; i += 4)
        *reinterpret_cast<f32x4*>(o + i) =
            f32x4{best[g][i], best[g][i + 1], best[g][i + 2], best[g][i + 3]};
    }
    return;
  }

  int next_refresh = 1;
  sweep(select, [&](int cyc) {
    if (cyc == next_refresh) {
      refresh();
      next_refresh = !a.qslots ? 2 * cyc + 1 : cyc == 1 ? 2 : cyc == 2 ? 5 : cyc + 11;
    }
  });
  if constexpr (MODE == kModeCollect) {
    flush(0);  // whatever is left
    return;
  }

  // ---- final: this split's list of every query of this wave ----------------
  // A list longer than k' is sorted and cut to its k' best (rare once a
  // threshold is known).  The others go out as they stand, unsorted (the
  // merge ranks every entry), for all queries of the wave at once: entry j
  // of query qq is item qq * k' + j, items are dealt to the lanes, and each
  // lane has kCopyBatch loads in flight before it stores (20: two rounds
  // for k' = 40, the sweep's registers are free by now; 8 cost 1.3 % of the
  // BF16 search).  Done one query after another, each copy waited for its
  // own loads (about 0.4 ms of the launch at 1e4 x 1e6 x 128).  Only the
  // list's entries are written, and their count (the merge reads no further).
#pragma unroll
  for (int g = 0; g < QG; ++g) cnt_s[2 * (wq0 + g * 32 + c) + h] = cnt[g];
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  wave_sync();
  for (int qq = 0; qq < WQ; ++qq) {
    const int ql = wq0 + qq;
    const int64_t q = q0 + ql;
    if (q >= a.n_q) break;
    if (kSelected && a.query_active[q] == 0) continue;  // uniform across the wave
    const int n0 = cnt_s[2 * ql], n1 = cnt_s[2 * ql + 1];
    if (n0 + n1 <= a.kc) continue;  // copied below
    sort_query(ql, n0, n1);
    for (int j = lane; j < a.kc; j += 64) {
      const size_t o = ((size_t)q * a.n_splits + split) * a.kc + j;
      a.part_d[o] = sd[j];
      a.part_i[o] = si[j];
    }
    if (lane == 0) {
      a.part_n[(size_t)q * a.n_splits + split] = a.kc;
      // As in a compaction: the shared threshold ends at or below the k'-th
      // value kept, so every row below it stays in some split's list.
      atomicMin(&a.tau[q], ordered_key((double)sd[a.kc - 1]));
    }
    wave_sync();
  }
  constexpr int kCopyBatch = 20;
  const int items = WQ * a.kc;
  for (int base = 0; base < items; base += 64 * kCopyBatch) {
    float vd[kCopyBatch];
    int vi[kCopyBatch];
    int64_t dst[kCopyBatch];
#pragma unroll
    for (int u = 0; u < kCopyBatch; ++u) {
      const int it = base + u * 64 + lane;
      const int qq = it / a.kc, j = it - qq * a.kc;
      const int ql = wq0 + (qq < WQ ? qq : 0);
      const int64_t q = q0 + ql;
      const int n0 = cnt_s[2 * ql], n1 = cnt_s[2 * ql + 1];
      const bool mine = it < items && q < a.n_q && n0 + n1 <= a.kc &&
                        (!kSelected || a.query_active[q] != 0);
      if (mine && j == 0) a.part_n[(size_t)q * a.n_splits + split] = n0 + n1;
      const bool has = mine && j < n0 + n1;
      dst[u] = has ? ((int64_t)q * a.n_splits + split) * a.kc + j : -1;
      const int src = ql * a.cap + (j < n0 ? j : half + (j - n0));
      const uint2 pr = has ? cand[src] : make_uint2(0x7f800000u, 0xffffffffu);
      vd[u] = __uint_as_float(pr.x);
      vi[u] = (int)pr.y;
    }
#pragma unroll
    for (int u = 0; u < kCopyBatch; ++u)
      if (dst[u] >= 0) {
        a.part_d[dst[u]] = vd[u];
        a.part_i[dst[u]] = vi[u];
      }
  }
}

template <int FMT, int DM, int QG, int MODE>
void launch_one(const MfmaArgs& a, hipStream_t s) {
  // the list sorts' LDS (collect lists are never sorted in the kernel)
  const size_t lds = MODE == kModeSample || MODE == kModeCollect ? 0 : (size_t)4 * a.cap * 8;
  int64_t items = (int64_t)a.n_qblocks * a.n_splits;
  if (MODE == kModeCollect && a.qmap && a.pk_grid > items) items = a.pk_grid;
  const unsigned blocks = (unsigned)((items + 7) / 8 * 8);
  hipLaunchKernelGGL((mfma_search_kernel<FMT, DM, QG, MODE>), dim3(blocks), dim3(kThreads), lds,
                     s, a);
}

template <int FMT, int DM, int QG>
int occupancy_one(int cap) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, mfma_search_kernel<FMT, DM, QG, kModeSearch>,
                                                   kThreads, (size_t)4 * cap * 8) != hipSuccess ||
      n <= 0)
    return 2;
  return n;
}

}  // namespace

static_assert(kMfmaRowStep == kBY, "row step");

// Query groups (of 32) per wave: two where the resident queries (B operand,
// 2 x DM x esize bytes per lane pair) leave room for 3 waves per SIMD, else
// one.  FP8 at DM = 256 takes two (161 VGPRs; 1e4 x 1e6 x 256 search 3.22 ->
// 2.39 ms against one group); BF16 at 256 and float32 from 128 would not fit.
constexpr int groups_for(int fmt, int dm) {
  return ((fmt == kFmtF32 && dm >= 128) || (fmt == kFmtBF16 && dm >= 256)) ? 1 : 2;
}

int mfma_block_queries(int fmt, int dm) { return 4 * 32 * groups_for(fmt, dm); }

int mfma_blocks_per_cu(int fmt, int dm, int cap) {
#define KNN_OCC_CASE(F, D) \
  if (fmt == F && dm == D) return occupancy_one<F, D, groups_for(F, D)>(cap);
  KNN_OCC_CASE(kFmtF32, 32)
  KNN_OCC_CASE(kFmtF32, 64)
  KNN_OCC_CASE(kFmtF32, 128)
  KNN_OCC_CASE(kFmtF32, 192)
  KNN_OCC_CASE(kFmtF32, 256)
  KNN_OCC_CASE(kFmtBF16, 32)
  KNN_OCC_CASE(kFmtBF16, 64)
  KNN_OCC_CASE(kFmtBF16, 128)
  KNN_OCC_CASE(kFmtBF16, 256)
  KNN_OCC_CASE(kFmtFP8, 64)
  KNN_OCC_CASE(kFmtFP8, 128)
  KNN_OCC_CASE(kFmtFP8, 256)
#undef KNN_OCC_CASE
  return 2;
}

template <int MODE>
static bool launch_mode(const MfmaArgs& a, hipStream_t s) {
#define KNN_MFMA_CASE(F, D)                              \
  if (a.fmt == F && a.dm == D) {                         \
    launch_one<F, D, groups_for(F, D), MODE>(a, s);      \
    return true;                                         \
  }
  KNN_MFMA_CASE(kFmtF32, 32)
  KNN_MFMA_CASE(kFmtF32, 64)
  KNN_MFMA_CASE(kFmtF32, 128)
  KNN_MFMA_CASE(kFmtF32, 192)
  KNN_MFMA_CASE(kFmtF32, 256)
  KNN_MFMA_CASE(kFmtBF16, 32)
  KNN_MFMA_CASE(kFmtBF16, 64)
  KNN_MFMA_CASE(kFmtBF16, 128)
  KNN_MFMA_CASE(kFmtBF16, 256)
  KNN_MFMA_CASE(kFmtFP8, 64)
  KNN_MFMA_CASE(kFmtFP8, 128)
  KNN_MFMA_CASE(kFmtFP8, 256)
#undef KNN_MFMA_CASE
  return false;
}

bool launch_mfma_search(const MfmaArgs& a, hipStream_t s) { return launch_mode<kModeSearch>(a, s); }

bool launch_mfma_sample(const MfmaArgs& a, hipStream_t s) { return launch_mode<kModeSample>(a, s); }

bool launch_mfma_research(const MfmaArgs& a, hipStream_t s) {
  return a.qblock_active && a.query_active && launch_mode<kModeResearch>(a, s);
}

bool launch_mfma_collect(const MfmaArgs& a, hipStream_t s) {
  return a.qblock_active && a.query_active && a.coll_cnt && a.coll_idx && a.coll_cap > 0 &&
         launch_mode<kModeCollect>(a, s);
}

}  // namespace knn
