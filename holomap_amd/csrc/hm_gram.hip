// Candidate sweep: MFMA Gram tiles fused with a per-row top-C selection.
//
// Replaces the O(N^2 d) distance loop of the reference's exact kNN
// (src/knn.rs:31-47) for candidate *selection* only; the kept candidates are
// re-ranked in the reference's own arithmetic afterwards (hm_rerank.hip).
//
// Work item = 128 query rows of one dataset x one column split. Workgroup =
// 4 waves; wave w owns queries [32w, 32w+32). Each step the workgroup stages
// a 128-column block and the 128 query rows (KC features at a time) in LDS
// and every wave computes its 32 queries x 128 columns with four 32x32 MFMA
// tiles (candidates as the M operand, queries as N, so an output lane holds a
// single query: lane l -> query l%32, register a -> column
// 8*(a/4) + 4*(l/32) + a%4 of the tile).
//
// Selection key per (query i, column j):
//   euclidean  key = ||x_j||^2 - 2 <x_i, x_j>   (row constant ||x_i||^2 dropped)
//   cosine     key = -<w_i, w_j>, w = x/||x||; zero columns carry key -inf
// ordered as the u64 (ordered(key) << 32 | j), a strict total order.
//
// Selection. Lanes l and l+32 hold the same query but disjoint columns, so
// each lane keeps its own list: an LDS buffer of QL (key, j) entries and a
// threshold tau. A value is appended (a masked store, no cross-lane work)
// only if its key is below min(tau, tau of the partner lane). Before each 32-column subtile (at most 16
// values per lane) a lane with fewer than 16 free slots is compacted: a
// bisection on the key value finds a cut p keeping between ceil(C/2) and
// QL-16 entries, entries above p are dropped and tau := the largest kept
// key. Invariant: every column of the lane's half that is not in its buffer
// has key > min(tau_0, tau_1) (both only decrease). A row's two lists hold
// at least C keys between them, every column missing from both has key >
// min(tau_0, tau_1): the candidate set and the threshold the certification
// bound uses. No atomics anywhere.
#include "hm_internal.hpp"
#include "hm_select.hpp"

namespace hm {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint16_t bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  uint32_t r = ((u >> 16) & 1u) + 0x7fffu;
  return (uint16_t)((u + r) >> 16);
}
__device__ __forceinline__ float bf16_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// MODE 0: bf16 operand (op16); MODE 1: f32 operand (op32, fp32 MFMA);
// MODE 2: bf16x3 -- op32 split on load into hi + lo bf16, three MFMAs per
//         step (hi.hi + lo.hi + hi.lo).
// LDS budget (160 KB): the lane lists take 256*QL*6 bytes (f32 key + u16
// column offset; 96 KB for QL=64) and the two operand tiles ~36-40 KB.
template <int MODE, int QL>
struct GramCfg {
  static constexpr int KC = (MODE == 0) ? 64 : 32;
  static constexpr int GES = (MODE == 0) ? 2 : 4;          // global element bytes
  static constexpr int ES = (MODE == 1) ? 4 : 2;           // LDS element bytes
  static constexpr int LD = KC + (MODE == 1 ? 4 : 8);      // LDS row stride (elements)
  static constexpr int NT = (MODE == 2) ? 2 : 1;           // tiles per operand (hi/lo)
  static constexpr int U = KC * GES / 16;                  // 16-byte loads per row
  static constexpr int S = 128 * U / 256;                  // loads per thread per operand
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
template <int MODE, int S, int U, int LD, int NT>
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

template <int MODE, int QL>
__global__ __launch_bounds__(256, 1) void gram_topc_kernel(
    const DsDesc* __restrict__ ds_arr, const WorkItem* __restrict__ items,
    const int* __restrict__ nitems_dev, int nitems_host, const uint32_t* __restrict__ qlist) {
  using Cfg = GramCfg<MODE, QL>;
  constexpr int KC = Cfg::KC, LD = Cfg::LD, NT = Cfg::NT, U = Cfg::U, S = Cfg::S,
                GES = Cfg::GES;
  // LDS: operand tiles (queries, candidates), column bias, lane lists.
  __shared__ __attribute__((aligned(16))) uint8_t q_s[NT][128 * LD * Cfg::ES];
  __shared__ __attribute__((aligned(16))) uint8_t c_s[NT][128 * LD * Cfg::ES];
  __shared__ float colb_s[128];
  __shared__ float lkf[QL][256];
  __shared__ uint16_t lkj[QL][256];  // column - cbeg (split span <= 65536)

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int q = lane & 31, half = lane >> 5;
  const int nitems = nitems_dev ? *nitems_dev : nitems_host;

  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const WorkItem W = items[it];
    const DsDesc& D = ds_arr[W.ds];
    const int64_t n = D.n;
    const int kp = D.kp;
    // list sizes: a compaction triggered by a full lane list (> QL - 16)
    // takes every lane holding more than keep_lo + win down to
    // [keep_lo, keep_lo + win] (so lanes stay in step and a compaction buys
    // many free slots); the final trim keeps at most cout = QL - 16
    const int keep_lo = (D.C + 1) / 2, keep_out = D.cout;
    const int win = (keep_out - keep_lo) / 2;
    const float scale = D.metric == 1 ? -1.0f : -2.0f;
    // column range of this split (aligned to 128 columns)
    const int64_t span = ((n + (int64_t)D.nsplit * 128 - 1) / ((int64_t)D.nsplit * 128)) * 128;
    const int64_t cbeg = (int64_t)W.split * span;
    const int64_t cend = cbeg + span < n ? cbeg + span : n;

    // row ids of this item's queries, for the loader (S rows per thread)
    int64_t qrow[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      int r = (tid + 256 * s) / U;
      if (r < W.qcnt) {
        int p = W.qbeg + r;
        qrow[s] = qlist ? (int64_t)qlist[p] - D.row_off : (int64_t)p;
      } else {
        qrow[s] = -1;
      }
    }

    uint64_t tau = ~0ull;
    int cnt = 0;

    const void* opbase = (MODE == 0) ? (const void*)D.op16 : (const void*)D.op32;
    const int nchunk = (kp + KC - 1) / KC;

    for (int64_t cb = cbeg; cb < cend; cb += 128) {
      floatx16 acc[4];
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[t][e] = 0.0f;

      // register prefetch of one chunk: S x 16B per operand per thread
      u32x4 gq[S], gc[S];
      gram_gload<S, U, KC, GES>(gq, gc, opbase, qrow, cb, kp, 0, tid);
      for (int ch = 0; ch < nchunk; ++ch) {
        const int k0 = ch * KC;
        __syncthreads();  // all waves are done with the previous tiles / colb_s
        gram_lstore<MODE, S, U, LD, NT>(gq, gc, q_s[0], q_s[NT - 1], c_s[0], c_s[NT - 1], tid);
        if (ch == 0 && tid < 128) {
          int64_t j = cb + tid;
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
            float b = qf[(wave * 32 + q) * LD + kk];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
              float a = cf[(t * 32 + q) * LD + kk];
              acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc[t], 0, 0, 0);
            }
          }
        } else {
          for (int ks = 0; ks < kvalid / 16; ++ks) {
            const int kk = ks * 16 + half * 8;
            bf16x8_t bh = *(const bf16x8_t*)(q_s[0] + ((wave * 32 + q) * LD + kk) * 2);
            bf16x8_t bl;
            if constexpr (MODE == 2) bl = *(const bf16x8_t*)(q_s[1] + ((wave * 32 + q) * LD + kk) * 2);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
              bf16x8_t ah = *(const bf16x8_t*)(c_s[0] + ((t * 32 + q) * LD + kk) * 2);
              acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc[t], 0, 0, 0);
              if constexpr (MODE == 2) {
                bf16x8_t al = *(const bf16x8_t*)(c_s[1] + ((t * 32 + q) * LD + kk) * 2);
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc[t], 0, 0, 0);
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc[t], 0, 0, 0);
              }
            }
          }
        }
      }

      // ---- epilogue: keys, fast threshold filter, list appends ----
      // The two lanes of a row filter with the smaller of their thresholds:
      // certification uses that minimum anyway (see the file header).
      uint64_t teff = tau;
      {
        const uint64_t tp = shfl_xor_u64(tau, 32);
        teff = tp < teff ? tp : teff;
      }
      const float tauf = teff == ~0ull ? INFINITY : unordered_f32((uint32_t)(teff >> 32));
      bool anyhit = false;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
#pragma unroll
        for (int a = 0; a < 16; ++a) {
          const int jl = t * 32 + 8 * (a >> 2) + 4 * half + (a & 3);
          const float key = __builtin_fmaf(scale, acc[t][a], colb_s[jl]);
          acc[t][a] = key;
          anyhit |= (cb + jl < cend) && (key <= tauf);
        }
      }
      if (__ballot(anyhit) != 0ull) {
        // Each 32-column subtile adds at most 16 values to a lane's list;
        // before it, a lane list with fewer than 16 free slots triggers a
        // compaction of every lane list holding more than keep_lo + win.
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          if (__ballot(cnt > QL - 16) != 0ull) {
            lane_compact<QL>(&lkf[0][tid], &lkj[0][tid], 256, cnt, tau, keep_lo + win, keep_lo,
                             keep_lo + win);
            const uint64_t tp = shfl_xor_u64(tau, 32);
            teff = tp < tau ? tp : tau;
          }
#pragma unroll
          for (int a = 0; a < 16; ++a) {
            const int jl = t * 32 + 8 * (a >> 2) + 4 * half + (a & 3);
            const int64_t j = cb + jl;
            const float key = acc[t][a];
            // list keys carry the column offset within the split; teff is
            // in the same space (both lists of a row share cbeg)
            const uint64_t k64 = ((uint64_t)ordered_u32(key) << 32) | (uint64_t)(uint32_t)(j - cbeg);
            const bool hit = j < cend && k64 < teff;
            if (__ballot(hit) != 0ull) {
              if (hit) {
                lkf[cnt][tid] = key;
                lkj[cnt][tid] = (uint16_t)(j - cbeg);
                ++cnt;
              }
            }
          }
        }
      }
    }

    // ---- final: trim each lane list to <= cout keys, write it with tau ----
    if (__ballot(cnt > keep_out) != 0ull)
      lane_compact<QL>(&lkf[0][tid], &lkj[0][tid], 256, cnt, tau, keep_out, keep_lo, keep_out);
    const int slot = wave * 32 + q;
    if (slot < W.qcnt) {
      const int p = W.qbeg + slot;
      const int64_t row = qlist ? (int64_t)qlist[p] - D.row_off : (int64_t)p;
      const int list = W.split * 2 + half;
      uint64_t* dst = D.cand + ((size_t)row * D.nlist + list) * D.cout;
      for (int e = 0; e < D.cout; ++e) {
        uint64_t v = ~0ull;
        if (e < cnt)
          v = ((uint64_t)ordered_u32(lkf[e][tid]) << 32) | (uint32_t)(lkj[e][tid] + cbeg);
        dst[e] = v;
      }
      // back to absolute column ids (order preserved: same offset)
      D.ctau[(size_t)row * D.nlist + list] = tau == ~0ull ? tau : tau + (uint64_t)cbeg;
    }
    cnt = 0;
  }
}

#define HM_INST(MODE, QL)                                                          \
  template __global__ void gram_topc_kernel<MODE, QL>(const DsDesc*, const WorkItem*,    \
                                                      const int*, int, const uint32_t*);
HM_INST(0, 64)
HM_INST(1, 64)
HM_INST(2, 64)

}  // namespace hm
