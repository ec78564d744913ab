// Candidate sweep: MFMA Gram tiles for candidate selection.
//
// Replaces the O(N^2 d) distance loop of the reference's exact kNN
// (src/knn.rs:31-47) for candidate *selection* only; the kept candidates are
// re-ranked in the reference's own arithmetic afterwards (rerank_kernel in
// hm_knn.hip), so the selection precision never reaches the output.
//
// Work item = 128 query rows of one dataset x one column split. Workgroup =
// 4 waves; wave w owns queries [32w, 32w+32). The split's columns are swept in
// 128-column blocks; every wave computes its 32 queries x 128 columns with
// four 32x32 MFMA tiles (candidates as the M operand, queries as N, so an
// output lane holds a single query: lane l -> query l%32, register a of tile
// t -> column 32t + 8*(a/4) + 4*(l/32) + a%4 of the block).
//
// Operand staging: a "stage" is one 128-byte slice of every operand row of a
// block (64 bf16 or 32 f32 features) for the 128 query rows and the 128
// candidate rows, plus the next block's 128 column biases. Stages are copied
// global -> LDS by the LDS-DMA path (global_load_lds, 16 bytes per lane, no
// registers held while in flight) into a ring of NSTAGE slots, NSTAGE-1
// stages ahead of the MFMAs; a wave waits for its own copies with
// s_waitcnt vmcnt and a barrier publishes them. The ring runs across block
// boundaries, so the pipeline never drains inside an item. Rows in LDS are
// 128 bytes with the eight 16-byte chunks XOR-swizzled per row, which keeps
// the MFMA operand reads (16 bytes per lane, 32 rows) free of bank conflicts.
//
// Selection key per (query i, column j):
//   euclidean  key = ||x_j||^2 - 2 <x_i, x_j>   (row constant ||x_i||^2 dropped)
//   cosine     key = -<w_i, w_j>, w = x/||x||; zero columns carry key -inf
// The MFMAs accumulate key' = key / s (s = -2 euclidean, -1 cosine) directly:
// each block's accumulators start at bias_j / s (exact: s is a power of two)
// and add the dot products, so key = s * key' exactly and smaller keys are
// larger key'. Columns outside the split start at -inf (key +inf). The
// epilogue then needs no arithmetic on the key: a max (pass 0) or a compare
// (pass 1) per position.
//
// Two passes over the same tiles, each with a one-instruction epilogue:
//  PASS 0 (threshold): lane (q, half) keeps the running minimum key of each
//    of its 64 register positions (the maximum key') -- 128 "groups" per
//    row, each a fixed column subset. T = the C-th smallest group minimum.
//    C distinct groups each contain a key <= T, so at least C keys of the
//    split are <= T.
//  PASS 1 (collect): with T the smallest of the row's split thresholds,
//    every column with key <= T is appended to its lane's list in global
//    memory (the row's candidate slots; a few dozen appends per lane over a
//    whole split). All other columns have key > T: that is
//    the threshold the certification bound uses. A list that overflows is
//    reported (threshold 0) and its row is recomputed by the exact scan.
//    The keys of both passes come from identical MFMA sequences;
//    correctness only relies on pass 1.
// No atomics; nothing depends on scheduling.
#include "hm_internal.hpp"

namespace hm {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __attribute__((address_space(3))) void lds_void_t;
// A global (not generic) pointer: its stores are global_store, counted by
// vmcnt only. A flat store would also count in lgkmcnt, which every LDS
// barrier waits for.
typedef __attribute__((address_space(1))) uint64_t gptr_u64;

// Workgroup barrier that orders LDS only. __syncthreads() also waits for
// every outstanding global access (vmcnt 0), which would drain the copy
// pipeline at each barrier; the copies are waited for explicitly instead.
// The lgkm wait is the builtin (0xc07f: lgkmcnt 0, vmcnt and expcnt at
// their maxima, i.e. not waited for), which the compiler's own wait
// insertion sees: after it, no LDS read counts as outstanding, so the waits
// it places before later MFMAs can be partial.
#define HM_LDS_BARRIER()                 \
  do {                                   \
    __builtin_amdgcn_s_waitcnt(0xc07f);  \
    asm volatile("s_barrier" ::: "memory"); \
  } while (0)

template <int N>
__device__ __forceinline__ void vm_wait() {
  asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

// MODE 0: bf16 operand (op16); MODE 1: f32 operand (op32, fp32 MFMA);
// MODE 2: bf16x3 -- hi (op16) and lo (op16lo) bf16 parts, three MFMAs per
//         step (hi.hi + lo.hi + hi.lo).
// A workgroup has QW query rows that share every candidate slice (256 in
// modes 0 and 1; 128 in mode 2, whose stages are twice as big) and one wave
// per 32 of them. A wave computes its 32 queries x CW columns (CW / 32 tiles
// of 32x32). With eight waves, two share each SIMD, so one computes while
// the other waits. (A 64-query wave tile halves the LDS reads per MFMA but
// leaves one wave per SIMD; measured slower.)
// CW = 256 (the bf16 collect pass): twice the MFMAs per stage and barrier,
// and the query slices are re-read once per 256 columns instead of 128; its
// 128 accumulators leave no room for the threshold pass's group maxima.
template <int MODE, int CW = 128>
struct SweepCfg {
  static constexpr int NT = MODE == 2 ? 2 : 1;   // parts per operand (hi, lo)
  static constexpr int ES = MODE == 1 ? 4 : 2;   // element bytes
  static constexpr int KC = 128 / ES;            // features per stage
  static constexpr int QW = MODE == 2 ? 128 : 256;  // query rows per item
  static constexpr int NU = 1;                   // 32-query tiles per wave
  static constexpr int NW = QW / (32 * NU);      // waves per workgroup
  static constexpr int NCT = CW / 32;            // 32-column tiles per wave
  static constexpr int QG = QW / 8;              // 8-row copy groups of a query slice
  static constexpr int CG = CW / 8;              // ... and of a candidate slice
  static constexpr int RG = NT * (QG + CG);      // copy groups per stage
  static constexpr int NI = RG / NW;             // operand copies per wave per stage
  static constexpr int QT = 0;                   // stage offsets: query hi,
  static constexpr int CT = QW * 128;            //   candidate hi,
  static constexpr int QL = (QW + CW) * 128;     //   query lo, candidate lo
  static constexpr int CL = (2 * QW + CW) * 128;
  static constexpr int NBC = CW / 64;            // waves per copy of a block's CW biases
  static constexpr int BIAS = NT * (QW + CW) * 128;  // NW / NBC copies of the biases
  static constexpr int STAGE = BIAS + NW * 256;
  static constexpr int NSTAGE = (MODE == 2 || CW == 256) ? 2 : 3;
  static constexpr int VM = NI + 1;              // vmcnt entries per stage and wave
  static constexpr int BIAS0 = NSTAGE * STAGE;   // the item's first block's biases
  static constexpr int LDS = BIAS0 + CW * 4;
  static_assert(RG % NW == 0, "copy groups split evenly over the waves");
  static_assert(NW % NBC == 0, "whole copies of the biases");
  static_assert(LDS <= 160 * 1024, "LDS");
  static_assert(NW * 64 == sweep_threads(MODE), "the runtime launches sweep_threads(MODE)");
};

// Chunk slot of logical 16-byte chunk c in LDS row r: c ^ swz(r).
__device__ __forceinline__ int row_swz(int r) { return (r & 7) ^ ((r >> 3) & 1); }

// Per-item copy plan of one lane. A stage is QG + 16 (x NT) groups of 8 rows
// in the order query hi, candidate hi, query lo, candidate lo, then the
// biases. Wave w copies the groups rb = w*NI + i, i < NI (1 KB each: lane l
// -> row rb*8 + l/8 of the stage, chunk slot l%8) and 64 biases. src[i] is
// the lane's source for feature slice 0 (and, for candidate rows, block 0 of
// the split); cand[i] marks candidate groups, which also advance by block.
template <int MODE, int CW = 128>
struct CopyPlan {
  const uint8_t* src[SweepCfg<MODE, CW>::NI];
  bool cand[SweepCfg<MODE, CW>::NI];
  const float* bias;
  int64_t blim;  // the lane's bias column is in the split iff boff < blim
};

// Bias source of the columns past a split's end: key' = -inf (key +inf), so
// the accumulators need no per-column check.
__device__ const float kBiasOut = -INFINITY;

template <int MODE, int CW>
__device__ __forceinline__ void sweep_issue(uint8_t* stage, const CopyPlan<MODE, CW>& P,
                                            int64_t boff_bytes, int koff, int64_t boff, int wave) {
  using Cfg = SweepCfg<MODE, CW>;
#pragma unroll
  for (int i = 0; i < Cfg::NI; ++i) {
    const int rb = wave * Cfg::NI + i;
    const uint8_t* src = P.src[i] + (P.cand[i] ? boff_bytes + koff : (int64_t)koff);
    __builtin_amdgcn_global_load_lds(src, (lds_void_t*)(stage + rb * 1024), 16, 0, 0);
  }
  __builtin_amdgcn_global_load_lds(
      boff < P.blim ? P.bias + boff : &kBiasOut,
      (lds_void_t*)(stage + Cfg::BIAS + (wave / Cfg::NBC) * (Cfg::NBC * 256) +
                    (wave % Cfg::NBC) * 256),
      4, 0, 0);
}

// Wait until this wave's copies of the stage computed next have landed;
// `after` stages were issued after it.
template <int MODE, int CW>
__device__ __forceinline__ void sweep_wait(int after) {
  constexpr int V = SweepCfg<MODE, CW>::VM;
  static_assert(SweepCfg<MODE, CW>::NSTAGE <= 3, "waits for up to two later stages");
  if (after >= 2) vm_wait<2 * V>();
  else if (after == 1) vm_wait<V>();
  else vm_wait<0>();
}

// Accumulate one stage: the wave's NU x 32 queries x 128 columns x one
// 128-byte slice. Operand reads go in batches ahead of the MFMAs that use
// them (sched_group_barrier keeps that order; the default schedule
// interleaves reads and MFMAs one by one and exposes the LDS latency each
// time); all of a stage's reads are issued before its first MFMA.
template <int MODE, int NCT = 4>
__device__ __forceinline__ void sweep_mma(floatx16 (&acc)[SweepCfg<MODE>::NU][NCT],
                                          const uint8_t* cand, const uint8_t* qrow_p, int q,
                                          int half, int swz) {
  using Cfg = SweepCfg<MODE>;
  constexpr int NU = Cfg::NU;
  static_assert(NCT == 4 || MODE == 0, "wide tiles in the bf16 mode only");
  if constexpr (MODE == 1) {
    // chunk c = 2m + half holds features 8m + 4*half + j (j < 4); MFMA j
    // of step m sums features 8m + j and 8m + 4 + j over the two halves
#pragma unroll
    for (int mh = 0; mh < 2; ++mh) {
      floatx4 b4[2][NU], a4[2][4];
#pragma unroll
      for (int v = 0; v < 2; ++v) {
        const int cs = ((2 * (2 * mh + v) + half) ^ swz) * 16;
#pragma unroll
        for (int u = 0; u < NU; ++u) b4[v][u] = *(const floatx4*)(qrow_p + u * 32 * 128 + cs);
#pragma unroll
        for (int t = 0; t < 4; ++t) a4[v][t] = *(const floatx4*)(cand + (t * 32 + q) * 128 + cs);
      }
#pragma unroll
      for (int v = 0; v < 2; ++v)
#pragma unroll
        for (int u = 0; u < NU; ++u)
#pragma unroll
          for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int j = 0; j < 4; ++j)
              acc[u][t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a4[v][t][j], b4[v][u][j], acc[u][t],
                                                               0, 0, 0);
    }
    __builtin_amdgcn_sched_group_barrier(0x100, 2 * (NU + 4), 0);
    __builtin_amdgcn_sched_group_barrier(0x100, 2 * (NU + 4), 0);
    __builtin_amdgcn_sched_group_barrier(0x008, 2 * NU * 16, 0);
    __builtin_amdgcn_sched_group_barrier(0x008, 2 * NU * 16, 0);
  } else if constexpr (MODE == 0) {
    // four k-steps of 16 features. The reads of step s + LA go out with the
    // MFMAs of step s; before step s an explicit wait leaves only the reads
    // of the later steps in flight (LDS reads complete in order). The
    // compiler still strengthens some of these waits to 0: it cannot tell
    // the reads apart from the LDS copies in flight. Wide tiles (8 column
    // tiles: 8 MFMAs per step) look one step ahead, which keeps their
    // operands beside 128 accumulators.
    constexpr int RS = NU + NCT;  // LDS reads per step
    constexpr int LA = NCT > 4 ? 1 : 2;
    static_assert(RS < 16, "lgkmcnt");
    bf16x8_t bh[4][NU], ah[4][NCT];
#pragma unroll
    for (int s = 0; s < 4 + LA; ++s) {
      if (s >= LA) {
        // reads of steps s - LA + 1 .. min(s, 4) - 1 may stay in flight
        const int keep = (s < 4 ? s : 4) - (s - LA) - 1;
        if (keep >= 1) __builtin_amdgcn_s_waitcnt(0xc07f | (RS << 8));  // lgkmcnt(RS)
        else __builtin_amdgcn_s_waitcnt(0xc07f);                        // lgkmcnt(0)
      }
      __builtin_amdgcn_sched_barrier(0);
      if (s < 4) {
        const int cs = ((2 * s + half) ^ swz) * 16;
#pragma unroll
        for (int u = 0; u < NU; ++u) bh[s][u] = *(const bf16x8_t*)(qrow_p + u * 32 * 128 + cs);
#pragma unroll
        for (int t = 0; t < NCT; ++t) ah[s][t] = *(const bf16x8_t*)(cand + (t * 32 + q) * 128 + cs);
      }
      if (s >= LA) {
#pragma unroll
        for (int u = 0; u < NU; ++u)
#pragma unroll
          for (int t = 0; t < NCT; ++t)
            acc[u][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[s - LA][t], bh[s - LA][u],
                                                                acc[u][t], 0, 0, 0);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  } else {
    static_assert(NU == 1, "bf16x3 waves hold 32 queries");
    // two halves of two k-steps: hi and lo parts of both operands
#pragma unroll
    for (int kh = 0; kh < 2; ++kh) {
      bf16x8_t bh[2], bl[2], ah[2][4], al[2][4];
#pragma unroll
      for (int v = 0; v < 2; ++v) {
        const int cs = ((2 * (2 * kh + v) + half) ^ swz) * 16;
        bh[v] = *(const bf16x8_t*)(qrow_p + cs);
        bl[v] = *(const bf16x8_t*)(qrow_p + Cfg::QL + cs);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          ah[v][t] = *(const bf16x8_t*)(cand + (t * 32 + q) * 128 + cs);
          al[v][t] = *(const bf16x8_t*)(cand + (Cfg::CL - Cfg::CT) + (t * 32 + q) * 128 + cs);
        }
      }
#pragma unroll
      for (int v = 0; v < 2; ++v)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          acc[0][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[v][t], bh[v], acc[0][t], 0, 0, 0);
          acc[0][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[v][t], bh[v], acc[0][t], 0, 0, 0);
          acc[0][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[v][t], bl[v], acc[0][t], 0, 0, 0);
        }
      __builtin_amdgcn_sched_group_barrier(0x100, 20, 0);  // 20 LDS reads
      __builtin_amdgcn_sched_group_barrier(0x008, 24, 0);  // 24 MFMAs
    }
  }
}

// Accumulator start of a block: bias_j / s for the block's columns, as
// copied with the stage (prep stores the bias already divided by s; columns
// past the split's end were copied from kBiasOut).
template <int NCT>
__device__ __forceinline__ void sweep_acc_start(floatx16 (&acc)[NCT], const uint8_t* bias,
                                                int half) {
  const float* colb = (const float*)bias + 4 * half;
#pragma unroll
  for (int t = 0; t < NCT; ++t)
#pragma unroll
    for (int a4 = 0; a4 < 4; ++a4) {
      const floatx4 v = *(const floatx4*)(colb + t * 32 + 8 * a4);
#pragma unroll
      for (int e = 0; e < 4; ++e) acc[t][a4 * 4 + e] = v[e];
    }
}

// Epilogue of one block (NCT x 32 columns) for one 32-query tile. acc holds
// key' (see the top of the file). Pass 0: gmax = max(gmax, key') per position
// (v_med3 with +inf: a max without NaN canonicalisation). Pass 1: a hit is
// key' >= Tp (Tp = T / s; NaN for lanes without a query: never); hits are
// appended as (key' bits, column) to dst[min(cnt, cout-1)] and all counted
// -- cnt > cout afterwards means the list overflowed (its contents are then
// unused).
template <int PASS, int NCT>
__device__ __forceinline__ void sweep_epilogue(const floatx16 (&acc)[NCT], float (&gmax)[4][16],
                                               uint32_t colbase, float Tp, int cout,
                                               gptr_u64* dst, int& cnt) {
  static_assert(PASS == 1 || NCT == 4, "group maxima for 128-column blocks");
  if constexpr (PASS == 0) {
    // +inf in a register: with a literal the compiler turns the med3 into a
    // NaN-canonicalising max (three instructions)
    float pinf = INFINITY;
    asm volatile("" : "+v"(pinf));
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int a = 0; a < 16; ++a) gmax[t][a] = __builtin_amdgcn_fmed3f(gmax[t][a], acc[t][a], pinf);
  } else {
    // A branch per position (compare, exec save, branch, exec restore in
    // sequence) costs ~75 cycles each, although hits are rare (about one
    // per lane and block); so does a scalar branch per position on the
    // wave's hit mask (measured slower). Instead: a tile's 16 hit bits first
    // (plain VALU), then one pass per hit of the lane (positions ascending,
    // the order the lists keep), the value picked by a bitwise select tree
    // on the position (a tree of conditional moves on its bits; no register
    // indexing).
#pragma unroll
    for (int t = 0; t < NCT; ++t) {
      uint32_t m = 0;
#pragma unroll
      for (int a = 0; a < 16; ++a) m |= acc[t][a] >= Tp ? 1u << a : 0u;
      while (m) {
        const uint32_t a = (uint32_t)__builtin_ctz(m);
        m &= m - 1;
        const uint32_t m0 = 0u - (a & 1u), m1 = 0u - ((a >> 1) & 1u);
        const uint32_t m2 = 0u - ((a >> 2) & 1u), m3 = 0u - ((a >> 3) & 1u);
        uint32_t l1[8], l2[4], l3[2];
#pragma unroll
        for (int q = 0; q < 8; ++q)
          l1[q] = (__float_as_uint(acc[t][2 * q + 1]) & m0) | (__float_as_uint(acc[t][2 * q]) & ~m0);
#pragma unroll
        for (int q = 0; q < 4; ++q) l2[q] = (l1[2 * q + 1] & m1) | (l1[2 * q] & ~m1);
#pragma unroll
        for (int q = 0; q < 2; ++q) l3[q] = (l2[2 * q + 1] & m2) | (l2[2 * q] & ~m2);
        const uint32_t bits = (l3[1] & m3) | (l3[0] & ~m3);
        const uint32_t jl = t * 32 + 8 * (a >> 2) + (a & 3);  // column at lane half 0
        dst[cnt < cout ? cnt : cout - 1] = ((uint64_t)bits << 32) | (colbase + jl);
        ++cnt;
      }
    }
  }
}

// Pass-1 threshold of a row, as T / s: the smallest of the pass-0
// thresholds of its column splits. Every split collects the keys <= T, and
// the split whose own threshold is T holds at least C of them, so the row
// gets >= C candidates in all (a single split's threshold would give each
// split about C). T = +inf (every split has fewer than C columns) must not
// admit the columns outside a split, whose key' is -inf: T / s >= -FLT_MAX.
__device__ __forceinline__ float sweep_threshold(const DsDesc& D, int64_t row, float scale) {
  float t = INFINITY;
  for (int sp = 0; sp < D.nsplit; ++sp) t = fminf(t, D.thr[(size_t)row * D.nsplit + sp]);
  return fmaxf(t / scale, -0x1.fffffep127f);
}

// End of an item for one lane. Pass 0: T = C-th smallest of the row's 128
// group minima (both lanes) = s * (C-th largest group maximum of key'),
// found by bisection on the order-preserving u32 image of -key'. Pass 1:
// the list's threshold and count.
template <int PASS>
__device__ __forceinline__ void sweep_finish(const float (&gmax)[4][16], float T, int cnt,
                                             bool myvalid, int64_t myrow, const DsDesc& D,
                                             int split, int half, float scale, int C) {
  if constexpr (PASS == 0) {
    if (C <= 8) {
      // few to select (a sampled threshold pass): the C-th largest group
      // maximum by extracting the largest value of the lane pair C times
      // at most (with multiplicity); -inf when fewer than C values exist
      float g[4][16];
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int a = 0; a < 16; ++a) g[t][a] = gmax[t][a];
      float m = -INFINITY;
      int need = C;
      for (int it = 0; it < C; ++it) {  // uniform over the wave
        m = -INFINITY;
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
          for (int a = 0; a < 16; ++a) m = fmaxf(m, g[t][a]);
        m = fmaxf(m, __shfl_xor(m, 32, 64));
        int c = 0;
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
          for (int a = 0; a < 16; ++a) c += g[t][a] == m ? 1 : 0;
        c += __shfl_xor(c, 32, 64);
        if (c >= need || !(m > -INFINITY)) break;
        need -= c;
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
          for (int a = 0; a < 16; ++a) g[t][a] = g[t][a] == m ? -INFINITY : g[t][a];
      }
      if (myvalid && half == 0) D.thr[(size_t)myrow * D.nsplit + split] = scale * m;
    } else {
      // bisection on the order-preserving images of -key' (ascending in the
      // key), bracketed by the pair's smallest and largest image: hi keeps
      // >= C values at or below it, the C-th smallest stays >= lo. It stops
      // once the bracket is 2^10 images wide; T from hi is then at most
      // that many ulps above the C-th smallest group minimum (a few more
      // keys listed) and still has >= C groups at or below it.
      uint32_t img[4][16];
      uint32_t lo = 0xffffffffu, hi = 0u;
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int a = 0; a < 16; ++a) {
          img[t][a] = ordered_u32(-gmax[t][a]);
          lo = img[t][a] < lo ? img[t][a] : lo;
          hi = img[t][a] > hi ? img[t][a] : hi;
        }
      {
        const uint32_t plo = (uint32_t)__shfl_xor((int)lo, 32, 64);
        const uint32_t phi = (uint32_t)__shfl_xor((int)hi, 32, 64);
        lo = plo < lo ? plo : lo;
        hi = phi > hi ? phi : hi;
      }
      // both lanes of a pair hold the same bracket: they iterate together
      while (hi - lo > 1024u) {
        const uint32_t mid = lo + (hi - lo) / 2;
        int c = 0;
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
          for (int a = 0; a < 16; ++a) c += img[t][a] <= mid ? 1 : 0;
        c += __shfl_xor(c, 32, 64);
        if (c >= C) hi = mid; else lo = mid + 1;
      }
      // hi: the image of +inf when the split has fewer than C columns;
      // T = -s * the value
      if (myvalid && half == 0)
        D.thr[(size_t)myrow * D.nsplit + split] = -scale * unordered_f32(hi);
    }
  } else {
    if (myvalid) {
      // every column not listed has key > T; an overflowing list reports 0
      // (no threshold) so the row is left to the exact scan
      const int list = split * 2 + half;
      D.ctau[(size_t)myrow * D.nlist + list] =
          cnt > D.cout ? 0ull : (((uint64_t)ordered_u32(scale * T) << 32) | 0xffffffffull);
      D.ccnt[(size_t)myrow * D.nlist + list] = (uint32_t)(cnt < D.cout ? cnt : D.cout);
    }
  }
}

template <int MODE, int PASS, int CW>
__global__ __launch_bounds__((SweepCfg<MODE, CW>::NW * 64), 1) void gram_sweep_kernel(
    const DsDesc* __restrict__ ds_arr, const WorkItem* __restrict__ items,
    const int* __restrict__ nitems_dev, int nitems_host, const uint32_t* __restrict__ qlist,
    int sample) {
  using Cfg = SweepCfg<MODE, CW>;
  constexpr int NI = Cfg::NI, NSTAGE = Cfg::NSTAGE, STAGE = Cfg::STAGE, QG = Cfg::QG;
  constexpr int NCT = Cfg::NCT, NBC = Cfg::NBC;
  static_assert(CW == 128 || PASS == 1, "the threshold pass keeps 128-column group maxima");
  __shared__ __attribute__((aligned(1024))) uint8_t smem[Cfg::LDS];

  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: LDS bases stay scalar
  const int q = lane & 31, half = lane >> 5;
  const int nitems = nitems_dev ? *nitems_dev : nitems_host;
  // copy side: the swizzled global chunk of a copy is cb_even (even group)
  // or cb_even ^ 16 (odd group); tiles start at even groups
  const int cb_even = ((lane & 7) ^ (lane >> 3)) * 16;  // byte offset of the chunk (rb even)
  // read side: rows 32t + q and 32w + q share swz(q)
  const int swz = row_swz(q);

  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const WorkItem W = items[it];
    const DsDesc& D = ds_arr[W.ds];
    const int64_t n = D.n;
    const int kp = D.kp;
    // operand rows are a multiple of 128 bytes (bf16 rows padded to kp16)
    const int rowbytes = MODE == 1 ? kp * 4 : D.kp16 * 2;
    const float scale = D.metric == 1 ? -1.0f : -2.0f;
    // column range of this split (aligned to 128 columns)
    const int64_t span = ((n + (int64_t)D.nsplit * 128 - 1) / ((int64_t)D.nsplit * 128)) * 128;
    const int64_t cbeg = (int64_t)W.split * span;
    const int64_t cend = cbeg + span < n ? cbeg + span : n;
    const int nchunk = rowbytes / 128;
    // the threshold pass of a sampled sweep visits every bstride-th block
    const int bstride = (PASS == 0 && sample) ? D.pstride : 1;
    const int csel = (PASS == 0 && sample) ? D.C0 : D.C;
    const int nblk_all = cend > cbeg ? (int)((cend - cbeg + CW - 1) / CW) : 0;
    const int nblk = (nblk_all + bstride - 1) / bstride;
    const int total = nblk * nchunk;
    const int64_t blkbytes = (int64_t)CW * rowbytes * bstride;  // between swept blocks
    const int64_t bcols = (int64_t)CW * bstride;

    const uint8_t* ophi = (MODE == 1) ? (const uint8_t*)D.op32 : (const uint8_t*)D.op16;
    const uint8_t* oplo = (MODE == 2) ? (const uint8_t*)D.op16lo : ophi;
    CopyPlan<MODE, CW> P;
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int rb = wave * NI + i;  // uniform
      const bool lo = rb >= QG + Cfg::CG;
      const int rl = lo ? rb - (QG + Cfg::CG) : rb;  // group within the hi or lo part
      const bool cand = rl >= QG;
      const int r = ((cand ? rl - QG : rl) << 3) + (lane >> 3);  // row within its tile
      int64_t row;
      if (!cand) {
        row = 0;  // missing queries read row 0 (results unused)
        if (r < W.qcnt) {
          const int p = W.qbeg + r;
          row = qlist ? (int64_t)qlist[p] - D.row_off : (int64_t)p;
        }
      } else {
        row = cbeg + r;
      }
      const int cbyte = (i & 1) ? (cb_even ^ 16) : cb_even;
      P.src[i] = (lo ? oplo : ophi) + row * rowbytes + cbyte;
      P.cand[i] = cand;
    }
    P.bias = D.colbias + cbeg + (wave % NBC) * 64 + lane;
    P.blim = cend - cbeg - ((wave % NBC) * 64 + lane);

    // this lane's own queries: slot 32*(NU*wave + u) + q of the item, u < NU
    constexpr int NU = Cfg::NU;
    bool myvalid[NU];
    int64_t myrow[NU];
    float T[NU];
    int cnt[NU];
    gptr_u64* dst[NU];
    float gmax[NU][4][16];
    const int list = W.split * 2 + half;
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int slot = 32 * (NU * wave + u) + q;
      myvalid[u] = slot < W.qcnt;
      myrow[u] = myvalid[u] ? (qlist ? (int64_t)qlist[W.qbeg + slot] - D.row_off
                                     : (int64_t)(W.qbeg + slot))
                            : 0;
      T[u] = __builtin_nanf("");  // no query: compares false (T[u] holds T / s)
      cnt[u] = 0;
      dst[u] = (gptr_u64*)(D.cand + ((size_t)myrow[u] * D.nlist + list) * D.cout);
      if constexpr (PASS == 0) {
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
          for (int a = 0; a < 16; ++a) gmax[u][t][a] = -INFINITY;
      } else {
        if (myvalid[u]) T[u] = sweep_threshold(D, myrow[u], scale);
      }
    }

    HM_LDS_BARRIER();  // the previous item's last stage is no longer read
    // the first block's biases (waves 0 .. NBC-1, 64 each), then stages
    // 0 .. NSTAGE-2; every stage carries the biases of the block after its own
    if (wave < NBC)
      __builtin_amdgcn_global_load_lds(0 < P.blim ? P.bias : &kBiasOut,
                                       (lds_void_t*)(smem + Cfg::BIAS0 + wave * 256), 4, 0, 0);
    // (bi, ci): block and slice of the next stage to issue
    int bi = 0, ci = 0;
#pragma unroll
    for (int g = 0; g < NSTAGE - 1; ++g) {
      if (g < total) {
        sweep_issue<MODE, CW>(smem + g * STAGE, P, bi * blkbytes, ci * 128,
                              (int64_t)(bi + 1) * bcols, wave);
        if (++ci == nchunk) { ci = 0; ++bi; }
      }
    }
    {
      const int issued = total < NSTAGE - 1 ? total : NSTAGE - 1;
      if (issued >= 2) vm_wait<2 * Cfg::VM>();
      else if (issued == 1) vm_wait<Cfg::VM>();
      else vm_wait<0>();
    }
    HM_LDS_BARRIER();  // the first block's biases are in LDS
    floatx16 acc[NU][NCT];
    {
      floatx16 a0[NCT];
      sweep_acc_start(a0, smem + Cfg::BIAS0, half);
#pragma unroll
      for (int u = 0; u < NU; ++u)
#pragma unroll
        for (int t = 0; t < NCT; ++t) acc[u][t] = a0[t];
    }

    // stage g: wait for this wave's copies of stage g; one barrier then
    // publishes stage g and retires stage g-1 in every wave, whose slot is
    // refilled with stage g+NSTAGE-1
    int ch = 0;                // slice of stage g within its block
    uint32_t colbase = (uint32_t)cbeg + 4 * half;  // this lane's columns of the block
    for (int g = 0; g < total; ++g) {
      sweep_wait<MODE, CW>(total - 1 - g < NSTAGE - 2 ? total - 1 - g : NSTAGE - 2);
      HM_LDS_BARRIER();
      const int gi = g + NSTAGE - 1;
      if (gi < total) {
        sweep_issue<MODE, CW>(smem + (gi % NSTAGE) * STAGE, P, bi * blkbytes, ci * 128,
                              (int64_t)(bi + 1) * bcols, wave);
        if (++ci == nchunk) { ci = 0; ++bi; }
      }
      const uint8_t* st = smem + (g % NSTAGE) * STAGE;
      sweep_mma<MODE, NCT>(acc, st + Cfg::CT, st + Cfg::QT + (NU * wave * 32 + q) * 128, q, half,
                           swz);
      if (++ch == nchunk) {
        // ---- epilogue of the block; the next block starts at its biases ----
#pragma unroll
        for (int u = 0; u < NU; ++u)
          sweep_epilogue<PASS>(acc[u], gmax[u], colbase, T[u], D.cout, dst[u], cnt[u]);
        floatx16 a0[NCT];
        sweep_acc_start(a0, st + Cfg::BIAS, half);
#pragma unroll
        for (int u = 0; u < NU; ++u)
#pragma unroll
          for (int t = 0; t < NCT; ++t) acc[u][t] = a0[t];
        ch = 0;
        colbase += (uint32_t)bcols;
      }
    }

#pragma unroll
    for (int u = 0; u < NU; ++u)
      sweep_finish<PASS>(gmax[u], T[u], cnt[u], myvalid[u], myrow[u], D, W.split, half, scale,
                         csel);
  }
}

// ---------------------------------------------------------------------------
// bf16 sweep with the item's query rows resident in LDS, for rows of at most
// QresCfg::QMAXB bytes (192 bf16 features). The 256 query rows are copied
// once per item; the ring stages then carry only the 128 candidate rows and
// the biases (18 KB instead of 48 KB), a third of the copy traffic per MFMA.
// Same tiles, MFMA sequence, epilogue and lists as gram_sweep_kernel<0, *>.
// ---------------------------------------------------------------------------
// Copies of one stage: wave w copies candidate groups w*NI + i (8 rows of
// the block at byte offset `off` of their rows) and 64 of the biases at boff.
template <class Cfg>
__device__ __forceinline__ void qres_issue(uint8_t* stage, const uint8_t* const (&csrc)[Cfg::NI],
                                           const float* bsrc, int64_t blim, int64_t off,
                                           int64_t boff, int wave) {
#pragma unroll
  for (int i = 0; i < Cfg::NI; ++i)
    __builtin_amdgcn_global_load_lds(csrc[i] + off,
                                     (lds_void_t*)(stage + (wave * Cfg::NI + i) * 1024), 16, 0, 0);
  __builtin_amdgcn_global_load_lds(
      boff < blim ? bsrc + boff : &kBiasOut,
      (lds_void_t*)(stage + Cfg::BIAS + (wave >> 1) * 512 + (wave & 1) * 256), 4, 0, 0);
}

template <class Cfg, int PASS>
__global__ __launch_bounds__(Cfg::NW * 64, 2) void gram_sweep_qres_kernel(
    const DsDesc* __restrict__ ds_arr, const WorkItem* __restrict__ items,
    const int* __restrict__ nitems_dev, int nitems_host, const uint32_t* __restrict__ qlist,
    int sample) {
  constexpr int NI = Cfg::NI, NSTAGE = Cfg::NSTAGE, STAGE = Cfg::STAGE, QW = Cfg::QW;
  __shared__ __attribute__((aligned(1024))) uint8_t smem[Cfg::LDS];

  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int q = lane & 31, half = lane >> 5;
  const int nitems = nitems_dev ? *nitems_dev : nitems_host;
  const int cb_even = ((lane & 7) ^ (lane >> 3)) * 16;  // copy side, even 8-row group
  const int swz = row_swz(q);                           // read side

  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const WorkItem W = items[it];
    const DsDesc& D = ds_arr[W.ds];
    const int64_t n = D.n;
    const int rowbytes = D.kp16 * 2;  // <= QMAXB (the runtime routes longer rows elsewhere)
    const float scale = D.metric == 1 ? -1.0f : -2.0f;
    const int64_t span = ((n + (int64_t)D.nsplit * 128 - 1) / ((int64_t)D.nsplit * 128)) * 128;
    const int64_t cbeg = (int64_t)W.split * span;
    const int64_t cend = cbeg + span < n ? cbeg + span : n;
    const int nchunk = rowbytes / 128;
    // the threshold pass of a sampled sweep visits every bstride-th block
    const int bstride = (PASS == 0 && sample) ? D.pstride : 1;
    const int csel = (PASS == 0 && sample) ? D.C0 : D.C;
    const int nblk_all = cend > cbeg ? (int)((cend - cbeg + 127) / 128) : 0;
    const int nblk = (nblk_all + bstride - 1) / bstride;
    const int total = nblk * nchunk;
    const int64_t blkbytes = (int64_t)128 * rowbytes * bstride;  // between swept blocks
    const int64_t bcols = (int64_t)128 * bstride;
    const uint8_t* op = (const uint8_t*)D.op16;

    // candidate copies: wave w copies groups w*NI + i (8 rows each) of a block
    const uint8_t* csrc[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int rb = wave * NI + i;  // uniform
      csrc[i] = op + (cbeg + rb * 8 + (lane >> 3)) * rowbytes + ((rb & 1) ? (cb_even ^ 16) : cb_even);
    }
    const float* bsrc = D.colbias + cbeg + (wave & 1) * 64 + lane;
    const int64_t blim = cend - cbeg - ((wave & 1) * 64 + lane);

    const int slot = 32 * wave + q;
    const bool myvalid = slot < W.qcnt;
    const int64_t myrow =
        myvalid ? (qlist ? (int64_t)qlist[W.qbeg + slot] - D.row_off : (int64_t)(W.qbeg + slot))
                : 0;
    const int list = W.split * 2 + half;
    gptr_u64* dst = (gptr_u64*)(D.cand + ((size_t)myrow * D.nlist + list) * D.cout);
    float T = __builtin_nanf("");  // no query: compares false (T holds T / s)
    int cnt = 0;
    float gmax[4][16];
    if constexpr (PASS == 0) {
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int a = 0; a < 16; ++a) gmax[t][a] = -INFINITY;
    } else {
      if (myvalid) T = sweep_threshold(D, myrow, scale);
    }

    HM_LDS_BARRIER();  // the previous item's stages and query rows are no longer read
    // the item's query rows, slice by slice: wave w copies groups 4w + i
    // (missing queries read row 0; their results are unused)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int qg = wave * 4 + i;  // uniform
      const int r = qg * 8 + (lane >> 3);
      int64_t row = 0;
      if (r < W.qcnt) row = qlist ? (int64_t)qlist[W.qbeg + r] - D.row_off : (int64_t)(W.qbeg + r);
      const uint8_t* src = op + row * rowbytes + ((qg & 1) ? (cb_even ^ 16) : cb_even);
      for (int c = 0; c < nchunk; ++c)
        __builtin_amdgcn_global_load_lds(src + c * 128,
                                         (lds_void_t*)(smem + Cfg::QOFF + c * QW * 128 + qg * 1024),
                                         16, 0, 0);
    }
    if (wave < 2)
      __builtin_amdgcn_global_load_lds(0 < blim ? bsrc : &kBiasOut,
                                       (lds_void_t*)(smem + Cfg::BIAS0 + wave * 256), 4, 0, 0);
    // stage copies: the block's candidate slice and the next block's biases
    int bi = 0, ci = 0;  // block and slice of the next stage to issue
#pragma unroll
    for (int g = 0; g < NSTAGE - 1; ++g) {
      if (g < total) {
        qres_issue<Cfg>(smem + g * STAGE, csrc, bsrc, blim, bi * blkbytes + ci * 128,
                        (int64_t)(bi + 1) * bcols, wave);
        if (++ci == nchunk) { ci = 0; ++bi; }
      }
    }
    {
      const int issued = total < NSTAGE - 1 ? total : NSTAGE - 1;
      if (issued >= 2) vm_wait<2 * Cfg::VM>();
      else if (issued == 1) vm_wait<Cfg::VM>();
      else vm_wait<0>();
    }
    HM_LDS_BARRIER();  // query rows and the first block's biases are in LDS
    floatx16 acc[1][4];
    sweep_acc_start(acc[0], smem + Cfg::BIAS0, half);

    int ch = 0;
    uint32_t colbase = (uint32_t)cbeg + 4 * half;
    const uint8_t* qrow = smem + Cfg::QOFF + slot * 128;
    for (int g = 0; g < total; ++g) {
      // this wave's copies of stage g have landed; one barrier publishes
      // stage g and retires stage g-1, whose slot gets stage g+NSTAGE-1
      if (NSTAGE > 2 && total - 1 - g >= 1) vm_wait<Cfg::VM>();
      else vm_wait<0>();
      HM_LDS_BARRIER();
      if (g + NSTAGE - 1 < total) {
        qres_issue<Cfg>(smem + ((g + NSTAGE - 1) % NSTAGE) * STAGE, csrc, bsrc, blim,
                        bi * blkbytes + ci * 128, (int64_t)(bi + 1) * bcols, wave);
        if (++ci == nchunk) { ci = 0; ++bi; }
      }
      const uint8_t* st = smem + (g % NSTAGE) * STAGE;
      sweep_mma<0>(acc, st, qrow + ch * (QW * 128), q, half, swz);
      if (++ch == nchunk) {
        sweep_epilogue<PASS>(acc[0], gmax, colbase, T, D.cout, dst, cnt);
        sweep_acc_start(acc[0], st + Cfg::BIAS, half);
        ch = 0;
        colbase += (uint32_t)bcols;
      }
    }
    sweep_finish<PASS>(gmax, T, cnt, myvalid, myrow, D, W.split, half, scale, csel);
  }
}

// ---------------------------------------------------------------------------
// bf16 sweep with the query operand in registers, for rows of at most
// QregCfg::QMAXB bytes (128 bf16 features). A wave's B operand (its 32
// queries, chunk 2s + half of slice c for k-step s) is read from LDS once per
// item into eight registers of 16 bytes, so a stage's LDS reads are the
// candidate slices only, and the LDS the query rows would occupy holds two
// more ring slots: three stages are in flight while one is computed (a
// 2-slot ring leaves each copy one stage's MFMAs to land, less than its
// round trip). Same tiles, MFMA order, epilogue and lists as
// gram_sweep_qres_kernel.
// ---------------------------------------------------------------------------
// One 128-byte slice of a block: the wave's 32 queries (registers) x 128
// candidate columns (LDS), four k-steps of 16 features. The candidate reads
// of step s + 2 go out with the MFMAs of step s (as in sweep_mma<0>).
__device__ __forceinline__ void qreg_mma(floatx16 (&acc)[4], const uint8_t* cand,
                                         const bf16x8_t (&qr)[4], int q, int half, int swz) {
  constexpr int LA = 2;
  bf16x8_t ah[4][4];
#pragma unroll
  for (int s = 0; s < 4 + LA; ++s) {
    if (s >= LA) {
      // reads of steps s - LA + 1 .. min(s, 4) - 1 may stay in flight
      const int keep = (s < 4 ? s : 4) - (s - LA) - 1;
      if (keep >= 1) __builtin_amdgcn_s_waitcnt(0xc07f | (4 << 8));  // lgkmcnt(4)
      else __builtin_amdgcn_s_waitcnt(0xc07f);                       // lgkmcnt(0)
    }
    __builtin_amdgcn_sched_barrier(0);
    if (s < 4) {
      const int cs = ((2 * s + half) ^ swz) * 16;
#pragma unroll
      for (int t = 0; t < 4; ++t) ah[s][t] = *(const bf16x8_t*)(cand + (t * 32 + q) * 128 + cs);
    }
    if (s >= LA) {
#pragma unroll
      for (int t = 0; t < 4; ++t)
        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[s - LA][t], qr[s - LA], acc[t], 0, 0, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <int PASS>
__global__ __launch_bounds__(QregCfg::NW * 64, QregCfg::PER) void gram_sweep_qreg_kernel(
    const DsDesc* __restrict__ ds_arr, const WorkItem* __restrict__ items,
    const int* __restrict__ nitems_dev, int nitems_host, const uint32_t* __restrict__ qlist,
    int sample) {
  using Cfg = QregCfg;
  constexpr int NI = Cfg::NI, STAGE = Cfg::STAGE, QW = Cfg::QW, VM = Cfg::VM;
  constexpr int NS = Cfg::NSTAGE;
  __shared__ __attribute__((aligned(1024))) uint8_t smem[Cfg::LDS];

  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int q = lane & 31, half = lane >> 5;
  const int nitems = nitems_dev ? *nitems_dev : nitems_host;
  const int cb_even = ((lane & 7) ^ (lane >> 3)) * 16;  // copy side, even 8-row group
  const int swz = row_swz(q);                           // read side
  // stage g lives in ring slot (g + 2) % NS: stages 0 .. NS - 3 are copied
  // into slots 2 .. NS - 1 while the query rows occupy slots 0 and 1
  auto slot = [](int g) { return (g + 2) % NS; };

  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const WorkItem W = items[it];
    const DsDesc& D = ds_arr[W.ds];
    const int64_t n = D.n;
    const int rowbytes = D.kp16 * 2;  // <= QMAXB (the runtime routes longer rows elsewhere)
    const float scale = D.metric == 1 ? -1.0f : -2.0f;
    const int64_t span = ((n + (int64_t)D.nsplit * 128 - 1) / ((int64_t)D.nsplit * 128)) * 128;
    const int64_t cbeg = (int64_t)W.split * span;
    const int64_t cend = cbeg + span < n ? cbeg + span : n;
    const int nchunk = rowbytes / 128;  // 1 or 2
    // the threshold pass of a sampled sweep visits every bstride-th block
    const int bstride = (PASS == 0 && sample) ? D.pstride : 1;
    const int csel = (PASS == 0 && sample) ? D.C0 : D.C;
    const int nblk_all = cend > cbeg ? (int)((cend - cbeg + 127) / 128) : 0;
    const int nblk = (nblk_all + bstride - 1) / bstride;
    const int total = nblk * nchunk;
    const int64_t blkbytes = (int64_t)128 * rowbytes * bstride;  // between swept blocks
    const int64_t bcols = (int64_t)128 * bstride;
    const uint8_t* op = (const uint8_t*)D.op16;

    // candidate copies: wave w copies groups w*NI + i (8 rows each) of a block
    const uint8_t* csrc[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int rb = wave * NI + i;  // uniform
      csrc[i] = op + (cbeg + rb * 8 + (lane >> 3)) * rowbytes + ((rb & 1) ? (cb_even ^ 16) : cb_even);
    }
    const float* bsrc = D.colbias + cbeg + (wave & 1) * 64 + lane;
    const int64_t blim = cend - cbeg - ((wave & 1) * 64 + lane);

    const int qslot = 32 * wave + q;
    const bool myvalid = qslot < W.qcnt;
    const int64_t myrow =
        myvalid ? (qlist ? (int64_t)qlist[W.qbeg + qslot] - D.row_off : (int64_t)(W.qbeg + qslot))
                : 0;
    const int list = W.split * 2 + half;
    gptr_u64* dst = (gptr_u64*)(D.cand + ((size_t)myrow * D.nlist + list) * D.cout);
    float T = __builtin_nanf("");  // no query: compares false (T holds T / s)
    int cnt = 0;
    float gmax[4][16];
    if constexpr (PASS == 0) {
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int a = 0; a < 16; ++a) gmax[t][a] = -INFINITY;
    } else {
      if (myvalid) T = sweep_threshold(D, myrow, scale);
    }

    HM_LDS_BARRIER();  // the previous item's ring slots are no longer read
    // the item's query rows into slots 0 and 1 (slice c of query row r at
    // c*QW*128 + r*128; missing queries read row 0, their results unused):
    // wave w copies groups 4w + i
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int qg = wave * 4 + i;  // uniform
      const int r = qg * 8 + (lane >> 3);
      int64_t row = 0;
      if (r < W.qcnt) row = qlist ? (int64_t)qlist[W.qbeg + r] - D.row_off : (int64_t)(W.qbeg + r);
      const uint8_t* src = op + row * rowbytes + ((qg & 1) ? (cb_even ^ 16) : cb_even);
      for (int c = 0; c < nchunk; ++c)
        __builtin_amdgcn_global_load_lds(src + c * 128, (lds_void_t*)(smem + c * QW * 128 + qg * 1024),
                                         16, 0, 0);
    }
    if (wave < 2)
      __builtin_amdgcn_global_load_lds(0 < blim ? bsrc : &kBiasOut,
                                       (lds_void_t*)(smem + Cfg::BIAS0 + wave * 256), 4, 0, 0);
    // stages 0 .. NS - 3 (the block's candidate slice and the next block's
    // biases)
    int bi = 0, ci = 0;  // block and slice of the next stage to issue
#pragma unroll
    for (int g = 0; g < NS - 2; ++g) {
      if (g < total) {
        qres_issue<Cfg>(smem + slot(g) * STAGE, csrc, bsrc, blim, bi * blkbytes + ci * 128,
                        (int64_t)(bi + 1) * bcols, wave);
        if (++ci == nchunk) { ci = 0; ++bi; }
      }
    }
    // the query rows and first biases were issued before the stages
    {
      const int pre = total < NS - 2 ? total : NS - 2;
      if (pre >= 2) vm_wait<2 * VM>();
      else if (pre == 1) vm_wait<VM>();
      else vm_wait<0>();
    }
    HM_LDS_BARRIER();  // query rows and the first block's biases are in LDS
    bf16x8_t qreg[2][4];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int cs = ((2 * s + half) ^ swz) * 16;
        qreg[c][s] = c < nchunk ? *(const bf16x8_t*)(smem + c * QW * 128 + qslot * 128 + cs)
                                : bf16x8_t{};
      }
    floatx16 acc[4];
    sweep_acc_start(acc, smem + Cfg::BIAS0, half);
    HM_LDS_BARRIER();  // every wave holds its queries: slots 0 and 1 are free
    if (NS - 2 < total) {
      qres_issue<Cfg>(smem + slot(NS - 2) * STAGE, csrc, bsrc, blim, bi * blkbytes + ci * 128,
                      (int64_t)(bi + 1) * bcols, wave);
      if (++ci == nchunk) { ci = 0; ++bi; }
    }

    int g = 0;
    uint32_t colbase = (uint32_t)cbeg + 4 * half;
    for (int b = 0; b < nblk; ++b) {
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        if (c < nchunk) {
          // this wave's copies of stage g have landed (stages g+1 ..
          // g+NS-2 were issued after it); one barrier publishes stage g and
          // retires stage g-1, whose slot gets stage g+NS-1
          const int after = total - 1 - g < NS - 2 ? total - 1 - g : NS - 2;
          if (after >= 2) vm_wait<2 * VM>();
          else if (after == 1) vm_wait<VM>();
          else vm_wait<0>();
          HM_LDS_BARRIER();
          if (g + NS - 1 < total) {
            qres_issue<Cfg>(smem + slot(g + NS - 1) * STAGE, csrc, bsrc, blim,
                            bi * blkbytes + ci * 128, (int64_t)(bi + 1) * bcols, wave);
            if (++ci == nchunk) { ci = 0; ++bi; }
          }
          qreg_mma(acc, smem + slot(g) * STAGE, qreg[c], q, half, swz);
          ++g;
        }
      }
      sweep_epilogue<PASS>(acc, gmax, colbase, T, D.cout, dst, cnt);
      sweep_acc_start(acc, smem + slot(g - 1) * STAGE + Cfg::BIAS, half);
      colbase += (uint32_t)bcols;
    }
    sweep_finish<PASS>(gmax, T, cnt, myvalid, myrow, D, W.split, half, scale, csel);
  }
}

template __global__ void gram_sweep_qreg_kernel<0>(const DsDesc*, const WorkItem*, const int*, int,
                                                   const uint32_t*, int);
template __global__ void gram_sweep_qreg_kernel<1>(const DsDesc*, const WorkItem*, const int*, int,
                                                   const uint32_t*, int);

#define HM_QRES_INST(CFG, PASS)                                                                \
  template __global__ void gram_sweep_qres_kernel<CFG, PASS>(const DsDesc*, const WorkItem*,      \
                                                             const int*, int, const uint32_t*,   \
                                                             int);
HM_QRES_INST(QresWide, 0)
HM_QRES_INST(QresWide, 1)
#undef HM_QRES_INST

#define HM_INST(MODE, PASS, CW)                                                       \
  template __global__ void gram_sweep_kernel<MODE, PASS, CW>(const DsDesc*, const WorkItem*, \
                                                             const int*, int, const uint32_t*, int);
HM_INST(0, 0, 128)
HM_INST(0, 1, 128)
HM_INST(0, 1, 256)
HM_INST(1, 0, 128)
HM_INST(1, 1, 128)
HM_INST(2, 0, 128)
HM_INST(2, 1, 128)

}  // namespace hm
