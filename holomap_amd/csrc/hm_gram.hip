// Candidate sweep: MFMA Gram tiles for candidate selection.
//
// Replaces the O(N^2 d) distance loop of the reference's exact kNN
// (src/knn.rs:31-47) for candidate *selection* only; the kept candidates are
// re-ranked in the reference's own arithmetic afterwards (rerank_kernel in
// hm_knn.hip), so the selection precision never reaches the output.
//
// Work item = 128 query rows of one dataset x one column split. Workgroup =
// 4 waves; wave w owns queries [32w, 32w+32). Each step the workgroup stages
// a 128-column block and the 128 query rows (KC features at a time) in LDS
// and every wave computes its 32 queries x 128 columns with four 32x32 MFMA
// tiles (candidates as the M operand, queries as N, so an output lane holds a
// single query: lane l -> query l%32, register a of tile t -> column
// 32t + 8*(a/4) + 4*(l/32) + a%4 of the block).
//
// Selection key per (query i, column j):
//   euclidean  key = ||x_j||^2 - 2 <x_i, x_j>   (row constant ||x_i||^2 dropped)
//   cosine     key = -<w_i, w_j>, w = x/||x||; zero columns carry key -inf
//
// Two passes over the same tiles, each with a two-instruction epilogue:
//  PASS 0 (threshold): lane (q, half) keeps the running minimum key of each
//    of its 64 register positions -- 128 "groups" per row, each a fixed
//    column subset. T = the C-th smallest group minimum. C distinct groups
//    each contain a key <= T, so at least C keys of the split are <= T.
//  PASS 1 (collect): every column with key <= T is appended to its lane's
//    list (f32 key, u16 column offset in LDS). All other columns have
//    key > T: that is the threshold the certification bound uses. A list
//    that overflows is reported (threshold 0) and its row is recomputed by
//    the exact scan. The keys of both passes come from identical MFMA
//    sequences; correctness only relies on pass 1.
// No atomics; nothing depends on scheduling.
#include "hm_internal.hpp"

namespace hm {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

#define HM_QB 48  // pass-1 list capacity per lane

__device__ __forceinline__ uint16_t bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  uint32_t r = ((u >> 16) & 1u) + 0x7fffu;
  return (uint16_t)((u + r) >> 16);
}
__device__ __forceinline__ float bf16_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// MODE 0: bf16 operand (op16); MODE 1: f32 operand (op32, fp32 MFMA);
// MODE 2: bf16x3 -- op32 split on load into hi + lo bf16, three MFMAs per
//         step (hi.hi + lo.hi + hi.lo).
// LDS: the two operand tiles take ~70-74 KB, the pass-1 lists 72 KB.
template <int MODE>
struct GramCfg {
  static constexpr int KC = (MODE == 0) ? 128 : 64;     // features per chunk
  static constexpr int GES = (MODE == 0) ? 2 : 4;       // global element bytes
  static constexpr int ES = (MODE == 1) ? 4 : 2;        // LDS element bytes
  static constexpr int LD = KC + (MODE == 1 ? 4 : 8);   // LDS row stride (elements)
  static constexpr int NT = (MODE == 2) ? 2 : 1;        // tiles per operand (hi/lo)
  static constexpr int U = KC * GES / 16;               // 16-byte loads per row
  static constexpr int S = 128 * U / 256;               // loads per thread per operand
};

// Global -> register prefetch of one KC-feature chunk of the 128 query rows
// and the 128 candidate rows (16-byte loads, zero beyond the padded width).
template <int S, int U, int KC, int GES>
__device__ __forceinline__ void gram_gload(u32x4 (&gq)[S], u32x4 (&gc)[S], const void* opbase,
                                           const int64_t (&qrow)[S], int64_t cb, int kp, int k0,
                                           int tid) {
  const int valid_bytes = (kp - k0 < KC ? kp - k0 : KC) * GES;
  const size_t rowbytes = (size_t)kp * GES;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int idx = tid + 256 * s;
    const int r = idx / U, c16 = idx % U;
    const bool inb = c16 * 16 < valid_bytes;
    const size_t koff = (size_t)k0 * GES + (size_t)c16 * 16;
    const u32x4 zero = {0u, 0u, 0u, 0u};
    gq[s] = (qrow[s] >= 0 && inb)
                ? *(const u32x4*)((const uint8_t*)opbase + (size_t)qrow[s] * rowbytes + koff)
                : zero;
    gc[s] = inb ? *(const u32x4*)((const uint8_t*)opbase + (size_t)(cb + r) * rowbytes + koff)
                : zero;
  }
}

__device__ __forceinline__ uint32_t pack_hi(uint32_t a, uint32_t b, uint32_t& lo_pack) {
  // hi/lo bf16 split of two f32 bit patterns; returns packed hi, sets packed lo
  const float fa = __uint_as_float(a), fb = __uint_as_float(b);
  const uint16_t ha = bf16_rne(fa), hb = bf16_rne(fb);
  const uint16_t la = bf16_rne(fa - bf16_f32(ha)), lb = bf16_rne(fb - bf16_f32(hb));
  lo_pack = (uint32_t)la | ((uint32_t)lb << 16);
  return (uint32_t)ha | ((uint32_t)hb << 16);
}

// Register -> LDS store of the prefetched chunk (MODE 2 splits into hi/lo).
template <int MODE, int S, int U, int LD>
__device__ __forceinline__ void gram_lstore(const u32x4 (&gq)[S], const u32x4 (&gc)[S],
                                            uint8_t* qh, uint8_t* ql, uint8_t* ch,
                                            uint8_t* cl, int tid) {
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int idx = tid + 256 * s;
    const int r = idx / U, c16 = idx % U;
    if constexpr (MODE == 0) {
      *(u32x4*)(qh + (r * LD + c16 * 8) * 2) = gq[s];
      *(u32x4*)(ch + (r * LD + c16 * 8) * 2) = gc[s];
    } else if constexpr (MODE == 1) {
      *(u32x4*)(qh + (r * LD + c16 * 4) * 4) = gq[s];
      *(u32x4*)(ch + (r * LD + c16 * 4) * 4) = gc[s];
    } else {
      uint32_t l0, l1, l2, l3;
      const uint32_t h0 = pack_hi(gq[s].x, gq[s].y, l0), h1 = pack_hi(gq[s].z, gq[s].w, l1);
      const uint32_t h2 = pack_hi(gc[s].x, gc[s].y, l2), h3 = pack_hi(gc[s].z, gc[s].w, l3);
      *(u32x2*)(qh + (r * LD + c16 * 4) * 2) = u32x2{h0, h1};
      *(u32x2*)(ql + (r * LD + c16 * 4) * 2) = u32x2{l0, l1};
      *(u32x2*)(ch + (r * LD + c16 * 4) * 2) = u32x2{h2, h3};
      *(u32x2*)(cl + (r * LD + c16 * 4) * 2) = u32x2{l2, l3};
    }
  }
}

template <int MODE, int PASS>
__global__ __launch_bounds__(256, 1) void gram_sweep_kernel(
    const DsDesc* __restrict__ ds_arr, const WorkItem* __restrict__ items,
    const int* __restrict__ nitems_dev, int nitems_host, const uint32_t* __restrict__ qlist) {
  using Cfg = GramCfg<MODE>;
  constexpr int KC = Cfg::KC, LD = Cfg::LD, NT = Cfg::NT, U = Cfg::U, S = Cfg::S,
                GES = Cfg::GES;
  constexpr int QB = (PASS == 1) ? HM_QB : 1;
  __shared__ __attribute__((aligned(16))) uint8_t q_s[NT][128 * LD * Cfg::ES];
  __shared__ __attribute__((aligned(16))) uint8_t c_s[NT][128 * LD * Cfg::ES];
  __shared__ float colb_s[128];
  __shared__ float lkf[QB][256];
  __shared__ uint16_t lkj[QB][256];  // column - cbeg (split span <= 65536)

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int q = lane & 31, half = lane >> 5;
  const int nitems = nitems_dev ? *nitems_dev : nitems_host;

  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const WorkItem W = items[it];
    const DsDesc& D = ds_arr[W.ds];
    const int64_t n = D.n;
    const int kp = D.kp;
    const float scale = D.metric == 1 ? -1.0f : -2.0f;
    // column range of this split (aligned to 128 columns)
    const int64_t span = ((n + (int64_t)D.nsplit * 128 - 1) / ((int64_t)D.nsplit * 128)) * 128;
    const int64_t cbeg = (int64_t)W.split * span;
    const int64_t cend = cbeg + span < n ? cbeg + span : n;

    // row ids of this item's queries, for the loader (S rows per thread)
    int64_t qrow[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int r = (tid + 256 * s) / U;
      if (r < W.qcnt) {
        const int p = W.qbeg + r;
        qrow[s] = qlist ? (int64_t)qlist[p] - D.row_off : (int64_t)p;
      } else {
        qrow[s] = -1;
      }
    }
    // this lane's own query
    const int slot = wave * 32 + q;
    const bool myvalid = slot < W.qcnt;
    const int64_t myrow =
        myvalid ? (qlist ? (int64_t)qlist[W.qbeg + slot] - D.row_off : (int64_t)(W.qbeg + slot))
                : 0;

    float gmin[4][16];
    float T = -INFINITY;
    int cnt = 0;
    bool ovf = false;
    if constexpr (PASS == 0) {
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int a = 0; a < 16; ++a) gmin[t][a] = INFINITY;
    } else {
      if (myvalid) T = D.thr[(size_t)myrow * D.nsplit + W.split];
    }

    const void* opbase = (MODE == 0) ? (const void*)D.op16 : (const void*)D.op32;
    const int nchunk = (kp + KC - 1) / KC;

    for (int64_t cb = cbeg; cb < cend; cb += 128) {
      floatx16 acc[4];
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[t][e] = 0.0f;

      u32x4 gq[S], gc[S];
      gram_gload<S, U, KC, GES>(gq, gc, opbase, qrow, cb, kp, 0, tid);
      for (int ch = 0; ch < nchunk; ++ch) {
        const int k0 = ch * KC;
        __syncthreads();  // all waves are done with the previous tiles / colb_s
        gram_lstore<MODE, S, U, LD>(gq, gc, q_s[0], q_s[NT - 1], c_s[0], c_s[NT - 1], tid);
        if (ch == 0 && tid < 128) {
          const int64_t j = cb + tid;
          colb_s[tid] = j < n ? D.colbias[j] : 0.0f;
        }
        __syncthreads();
        if (ch + 1 < nchunk) gram_gload<S, U, KC, GES>(gq, gc, opbase, qrow, cb, kp, k0 + KC, tid);
        const int kvalid = kp - k0 < KC ? kp - k0 : KC;
        if constexpr (MODE == 1) {
          const float* qf = (const float*)q_s[0];
          const float* cf = (const float*)c_s[0];
          for (int ks = 0; ks < kvalid / 2; ++ks) {
            const int kk = ks * 2 + half;
            const float b = qf[(wave * 32 + q) * LD + kk];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
              const float a = cf[(t * 32 + q) * LD + kk];
              acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc[t], 0, 0, 0);
            }
          }
        } else {
          for (int ks = 0; ks < kvalid / 16; ++ks) {
            const int kk = ks * 16 + half * 8;
            const bf16x8_t bh = *(const bf16x8_t*)(q_s[0] + ((wave * 32 + q) * LD + kk) * 2);
            bf16x8_t bl;
            if constexpr (MODE == 2) bl = *(const bf16x8_t*)(q_s[1] + ((wave * 32 + q) * LD + kk) * 2);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
              const bf16x8_t ah = *(const bf16x8_t*)(c_s[0] + ((t * 32 + q) * LD + kk) * 2);
              acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc[t], 0, 0, 0);
              if constexpr (MODE == 2) {
                const bf16x8_t al = *(const bf16x8_t*)(c_s[1] + ((t * 32 + q) * LD + kk) * 2);
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc[t], 0, 0, 0);
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc[t], 0, 0, 0);
              }
            }
          }
        }
      }

      // ---- epilogue ----
#pragma unroll
      for (int t = 0; t < 4; ++t) {
#pragma unroll
        for (int a = 0; a < 16; ++a) {
          const int jl = t * 32 + 8 * (a >> 2) + 4 * half + (a & 3);
          const bool valid = cb + jl < cend;
          const float key = __builtin_fmaf(scale, acc[t][a], colb_s[jl]);
          if constexpr (PASS == 0) {
            gmin[t][a] = valid ? fminf(gmin[t][a], key) : gmin[t][a];
          } else {
            const bool hit = valid && key <= T;
            if (__ballot(hit) != 0ull) {
              if (hit) {
                if (cnt < QB) {
                  lkf[cnt][tid] = key;
                  lkj[cnt][tid] = (uint16_t)(cb + jl - cbeg);
                  ++cnt;
                } else {
                  ovf = true;
                }
              }
            }
          }
        }
      }
    }

    if constexpr (PASS == 0) {
      // T = C-th smallest of the row's 128 group minima (both lanes), found
      // by bisection on the order-preserving u32 image of the key.
      uint32_t o[4][16];
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int a = 0; a < 16; ++a) o[t][a] = ordered_u32(gmin[t][a]);
      const int C = D.C;
      uint32_t lo = 0u, hi = 0xffffffffu;
      for (int itb = 0; itb < 32; ++itb) {
        const uint32_t mid = lo + (hi - lo) / 2;
        int c = 0;
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
          for (int a = 0; a < 16; ++a) c += o[t][a] <= mid ? 1 : 0;
        c += __shfl_xor(c, 32, 64);
        if (c >= C) hi = mid; else lo = mid + 1;
      }
      // lo == hi: the smallest image with >= C group minima at or below it
      // (the image of +inf when the split has fewer than C columns)
      if (myvalid && half == 0) D.thr[(size_t)myrow * D.nsplit + W.split] = unordered_f32(lo);
    } else {
      if (myvalid) {
        const int list = W.split * 2 + half;
        uint64_t* dst = D.cand + ((size_t)myrow * D.nlist + list) * D.cout;
        for (int e = 0; e < D.cout; ++e) {
          uint64_t v = ~0ull;
          if (e < cnt)
            v = ((uint64_t)ordered_u32(lkf[e][tid]) << 32) | (uint32_t)(lkj[e][tid] + cbeg);
          dst[e] = v;
        }
        // every column not listed has key > T; an overflowing list reports 0
        // (no threshold) so the row is left to the exact scan
        D.ctau[(size_t)myrow * D.nlist + list] =
            ovf ? 0ull : (((uint64_t)ordered_u32(T) << 32) | 0xffffffffull);
      }
    }
  }
}

#define HM_INST(MODE, PASS)                                                       \
  template __global__ void gram_sweep_kernel<MODE, PASS>(const DsDesc*, const WorkItem*, \
                                                         const int*, int, const uint32_t*);
HM_INST(0, 0)
HM_INST(0, 1)
HM_INST(1, 0)
HM_INST(1, 1)
HM_INST(2, 0)
HM_INST(2, 1)

}  // namespace hm
