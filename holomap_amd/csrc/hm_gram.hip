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
// keyed as the u64 (ordered(key) << 32 | j), a strict total order. Each row
// keeps an LDS buffer of QCAP keys and a threshold tau: a key is appended only
// if it is below tau; when a buffer fills, the wave sorts it (bitonic across
// the 64 lanes) and keeps the C smallest, tau := the C-th. Every column never
// kept therefore has key > tau_final, which is what the certification bound
// uses. No atomics: append positions come from ballots.
#include "hm_internal.hpp"

namespace hm {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  uint32_t r = ((u >> 16) & 1u) + 0x7fffu;
  return (uint16_t)((u + r) >> 16);
}
__device__ __forceinline__ float bf16_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Ascending bitonic sort of 128 u64 keys, two per lane (index lane, lane+64).
__device__ __forceinline__ void wave_sort128(uint64_t& v0, uint64_t& v1) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 2; k <= 128; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j == 64) {
        // partner of index lane is lane+64 in the same lane; k == 128 here
        uint64_t mn = v0 < v1 ? v0 : v1, mx = v0 < v1 ? v1 : v0;
        v0 = mn;
        v1 = mx;
      } else {
        uint64_t o0 = shfl_xor_u64(v0, j), o1 = shfl_xor_u64(v1, j);
        const int i0 = lane, i1 = lane + 64;
        bool up0 = (i0 & k) == 0, up1 = (i1 & k) == 0;
        bool lw = (lane & j) == 0;
        uint64_t mn0 = v0 < o0 ? v0 : o0, mx0 = v0 < o0 ? o0 : v0;
        uint64_t mn1 = v1 < o1 ? v1 : o1, mx1 = v1 < o1 ? o1 : v1;
        v0 = (lw == up0) ? mn0 : mx0;
        v1 = (lw == up1) ? mn1 : mx1;
      }
    }
  }
}

// Sort row r's `c` buffered keys, keep the C smallest in place; returns the
// C-th smallest (the new threshold) or ~0 if c < C.
template <int QCAP>
__device__ __forceinline__ uint64_t compact_row(uint64_t* row, int c, int C) {
  const int lane = threadIdx.x & 63;
  if constexpr (QCAP == 64) {
    uint64_t v = lane < c ? row[lane] : ~0ull;
    v = wave_sort64(v);
    if (lane < C) row[lane] = v;
    return readlane_u64(v, C - 1);
  } else {
    uint64_t v0 = lane < c ? row[lane] : ~0ull;
    uint64_t v1 = lane + 64 < c ? row[lane + 64] : ~0ull;
    wave_sort128(v0, v1);
    if (lane < C) row[lane] = v0;
    if (lane + 64 < C) row[lane + 64] = v1;
    return (C - 1) < 64 ? readlane_u64(v0, C - 1) : readlane_u64(v1, C - 1 - 64);
  }
}

// MODE 0: bf16 operand (op16); MODE 1: f32 operand (op32, fp32 MFMA);
// MODE 2: bf16x3 -- op32 split on load into hi + lo bf16, three MFMAs per
//         step (hi.hi + lo.hi + hi.lo).
// LDS budget (160 KB): row buffers take 128*QCAP*8 bytes (64, 96 or 128 KB),
// so the QCAP=128 variants stage half-size feature chunks.
template <int MODE, int QCAP>
struct GramCfg {
  static constexpr int KC = (MODE == 0) ? (QCAP <= 96 ? 64 : 32) : (QCAP <= 96 ? 32 : 16);
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
__device__ __forceinline__ void gram_gload(uint4 (&gq)[S], uint4 (&gc)[S], const void* opbase,
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
    const uint4 zero = make_uint4(0, 0, 0, 0);
    gq[s] = (qrow[s] >= 0 && inb)
                ? *(const uint4*)((const uint8_t*)opbase + (size_t)qrow[s] * rowbytes + koff)
                : zero;
    gc[s] = inb ? *(const uint4*)((const uint8_t*)opbase + (size_t)(cb + r) * rowbytes + koff)
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
__device__ __forceinline__ void gram_lstore(const uint4 (&gq)[S], const uint4 (&gc)[S],
                                            uint8_t* qh, uint8_t* ql, uint8_t* ch,
                                            uint8_t* cl, int tid) {
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int idx = tid + 256 * s;
    const int r = idx / U, c16 = idx % U;
    if constexpr (MODE == 0) {
      *(uint4*)(qh + (r * LD + c16 * 8) * 2) = gq[s];
      *(uint4*)(ch + (r * LD + c16 * 8) * 2) = gc[s];
    } else if constexpr (MODE == 1) {
      *(uint4*)(qh + (r * LD + c16 * 4) * 4) = gq[s];
      *(uint4*)(ch + (r * LD + c16 * 4) * 4) = gc[s];
    } else {
      uint32_t l0, l1, l2, l3;
      const uint32_t h0 = pack_hi(gq[s].x, gq[s].y, l0), h1 = pack_hi(gq[s].z, gq[s].w, l1);
      const uint32_t h2 = pack_hi(gc[s].x, gc[s].y, l2), h3 = pack_hi(gc[s].z, gc[s].w, l3);
      *(uint2*)(qh + (r * LD + c16 * 4) * 2) = make_uint2(h0, h1);
      *(uint2*)(ql + (r * LD + c16 * 4) * 2) = make_uint2(l0, l1);
      *(uint2*)(ch + (r * LD + c16 * 4) * 2) = make_uint2(h2, h3);
      *(uint2*)(cl + (r * LD + c16 * 4) * 2) = make_uint2(l2, l3);
    }
  }
}

template <int MODE, int QCAP>
__global__ __launch_bounds__(256, 1) void gram_topc_kernel(
    const DsDesc* __restrict__ ds_arr, const WorkItem* __restrict__ items,
    const int* __restrict__ nitems_dev, int nitems_host, const uint32_t* __restrict__ qlist) {
  using Cfg = GramCfg<MODE, QCAP>;
  constexpr int KC = Cfg::KC, LD = Cfg::LD, NT = Cfg::NT, U = Cfg::U, S = Cfg::S,
                GES = Cfg::GES;
  // LDS: operand tiles (queries, candidates), column bias, row buffers.
  __shared__ __attribute__((aligned(16))) uint8_t q_s[NT][128 * LD * Cfg::ES];
  __shared__ __attribute__((aligned(16))) uint8_t c_s[NT][128 * LD * Cfg::ES];
  __shared__ float colb_s[128];
  __shared__ uint64_t buf[128][QCAP];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int q = lane & 31, half = lane >> 5;
  const int nitems = nitems_dev ? *nitems_dev : nitems_host;

  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const WorkItem W = items[it];
    const DsDesc& D = ds_arr[W.ds];
    const int64_t n = D.n;
    const int kp = D.kp, C = D.C;
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
        qrow[s] = qlist ? (int64_t)qlist[p] : (int64_t)p;
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
      uint4 gq[S], gc[S];
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

      // ---- epilogue: keys, fast threshold filter, ballot insertion ----
      const float tauf = tau == ~0ull ? INFINITY : unordered_f32((uint32_t)(tau >> 32));
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
        uint64_t* mybuf = buf[wave * 32 + q];
        // A 32-column subtile adds at most 32 keys to a row. Before each
        // subtile every row is given >= 32 free slots (compacting the rows
        // that lack them; QCAP - C >= 32 by construction), so the unrolled
        // insertion below never overflows and needs no branch.
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          if (__ballot(cnt > QCAP - 32) != 0ull) {
            for (int r = 0; r < 32; ++r) {
              const int c = __builtin_amdgcn_readlane(cnt, r);
              if (c > QCAP - 32) {
                uint64_t nt = compact_row<QCAP>(buf[wave * 32 + r], c, C);
                if (q == r) { tau = nt; cnt = C; }
              }
            }
          }
#pragma unroll
          for (int a = 0; a < 16; ++a) {
            const int jl = t * 32 + 8 * (a >> 2) + 4 * half + (a & 3);
            const int64_t j = cb + jl;
            const uint64_t k64 = ((uint64_t)ordered_u32(acc[t][a]) << 32) | (uint64_t)(uint32_t)j;
            const bool hit = (j < cend) && (k64 < tau);
            const uint64_t mask = __ballot(hit);
            const int h0 = (int)((mask >> q) & 1ull), h1 = (int)((mask >> (q + 32)) & 1ull);
            const int pos = cnt + (half ? h0 : 0);
            if (hit) mybuf[pos] = k64;
            cnt += h0 + h1;
          }
        }
      }
    }

    // ---- final: trim to C, write candidates and threshold ----
    for (int r = 0; r < 32; ++r) {
      const int c = __builtin_amdgcn_readlane(cnt, r);
      if (c > C) {
        uint64_t nt = compact_row<QCAP>(buf[wave * 32 + r], c, C);
        if (q == r) { tau = nt; cnt = C; }
      }
    }
    for (int r = 0; r < 32; ++r) {
      const int slot = wave * 32 + r;
      if (slot >= W.qcnt) break;
      const int p = W.qbeg + slot;
      const int64_t row = qlist ? (int64_t)qlist[p] : (int64_t)p;
      const int c = __builtin_amdgcn_readlane(cnt, r);
      uint64_t* dst = D.cand + ((size_t)row * D.nsplit + W.split) * C;
      for (int e = lane; e < C; e += 64) dst[e] = e < c ? buf[slot][e] : ~0ull;
      if (lane == r) D.ctau[(size_t)row * D.nsplit + W.split] = tau;
    }
  }
}

#define HM_INST(MODE, QCAP)                                                        \
  template __global__ void gram_topc_kernel<MODE, QCAP>(const DsDesc*, const WorkItem*, \
                                                        const int*, int, const uint32_t*);
HM_INST(0, 64)
HM_INST(0, 96)
HM_INST(0, 128)
HM_INST(1, 64)
HM_INST(1, 96)
HM_INST(1, 128)
HM_INST(2, 64)
HM_INST(2, 96)
HM_INST(2, 128)

}  // namespace hm
