// Fuzzy-union symmetrisation (src/fuzzy.rs:49-84) and CSR (src/sparse.rs:21-35).
//
// The reference sorts the directed edges by (row, col), builds the sorted,
// de-duplicated union of A and A^T coordinates and fuses each pair with the
// probabilistic t-conorm w = (v + vt) - v*vt in f32. Here one stable radix
// sort of the kept edges by destination (hm_sort.hip) gives every row r its
// in-edges i -> r, sources ascending, at positions [lo_r, hi_r) of the
// sorted keys. A_r: the row's kept out-edge columns (at most k), I_r: its
// in-edge sources.
//   union_count_kernel (a group of G lanes per row, lane t = out-edge slot t;
//     k > 64 rows: lane t takes slots t, t + 64, ...): lo_r / hi_r by binary
//     search; for each out-edge its count of
//     smaller in-edge sources and its reverse weight; the union size
//     |A_r| + |I_r| - |A_r n I_r|;
//   (exclusive scan of the sizes -> global row pointer)
//   union_write_kernel: every entry goes to its rank in the sorted union,
//     #{a in A_r: a < x} + #{i in I_r: i < x} - #{m in A_r n I_r: m < x}, so
//     no row is sorted; out-edges by their row's group, in-edges one thread
//     each (so no row's in-degree serialises), plus the local row pointer.
// For an entry (r, c): v = A[r][c], vt = A[c][r] (0 if absent), w evaluated
// with the reference's operand order. With v, vt in [0, 1] and v + vt > 0,
// w >= max(v, vt) - min(v, vt) and w == 0 would need v == vt == v*vt, so
// w > 0 always and the reference's drop of zero entries (src/fuzzy.rs:73)
// never applies.
#include "hm_internal.hpp"

namespace hm {

__device__ __forceinline__ int ds_of_row(const DsDesc* __restrict__ ds, int nds, int64_t r) {
  int lo = 0, hi = nds - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ds[mid].row_off <= r) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// first position p in [0, n) with a[p] >= v (a sorted)
__device__ __forceinline__ int64_t lower_bound_u32(const uint32_t* __restrict__ a, int64_t n,
                                                   uint32_t v) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Value of lane U of the calling lane's G-lane group. A 16-lane group is
// one DPP row: a row_newbcast move (VALU); wider groups use an LDS permute.
template <int G, int U>
__device__ __forceinline__ uint32_t grp_get(uint32_t x) {
  if constexpr (G == 16)
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x150 + U, 0xf, 0xf, false);
  else
    return (uint32_t)__shfl((int)x, U, G);
}
// b(x_u, y_u) for u = 0 .. G-1 in order (unrolled, so U is a constant).
template <int G, int U = 0, class B>
__device__ __forceinline__ void grp_for(uint32_t x, uint32_t y, B& b) {
  if constexpr (U < G) {
    b(grp_get<G, U>(x), grp_get<G, U>(y));
    grp_for<G, U + 1>(x, y, b);
  }
}

// Loop bodies over a group's broadcast values.
struct ScanIn {  // sources below the column; the reverse weight
  uint32_t ca, less;
  bool found;
  float vt;
  __device__ void operator()(uint32_t su, uint32_t wu) {
    less += su < ca ? 1u : 0u;
    if (su == ca) { found = true; vt = __uint_as_float(wu); }
  }
};
struct ScanOut {  // out-edge columns (and mutual ones) below x; x among them
  uint32_t x, less, mut;
  bool in;
  __device__ void operator()(uint32_t cu, uint32_t fu) {
    less += cu < x ? 1u : 0u;
    mut += (fu && cu < x) ? 1u : 0u;
    in |= cu == x;
  }
};

// Row group (G lanes, lane t = out-edge slot t): in-edge range [lo, hi) by
// binary search (inptr[r] = lo; inptr[R] = hi of the last row); per kept
// out-edge r -> c: ai[e] = #{in-edge sources < c} | (c is one: bit 31) and
// av[e] = the weight of c -> r (0 if absent); cnt[r] = union size. Sources
// are compared G at a time through group broadcasts when the row has at most
// 4G in-edges, else each lane binary-searches them, so a row of very high
// in-degree costs O(k log in-degree).
template <int G>
__global__ __launch_bounds__(256) void union_count_kernel(
    const DsDesc* __restrict__ ds, int nds, int64_t total_rows, int64_t total_edges,
    const uint32_t* __restrict__ key1, const float* __restrict__ wval,
    const uint32_t* __restrict__ k1s, const uint32_t* __restrict__ p1s,
    const uint32_t* __restrict__ srcrow, uint32_t sentinel, uint32_t* __restrict__ inptr,
    uint64_t* __restrict__ cnt, uint32_t* __restrict__ ai, float* __restrict__ av) {
  static_assert(G >= 2 && G <= 64 && (G & (G - 1)) == 0, "group size");
  const int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int t = (int)(threadIdx.x % G);
  if (r >= total_rows) return;  // whole groups leave together
  const DsDesc& D = ds[ds_of_row(ds, nds, r)];
  const int k = D.k;
  const int64_t e0 = D.edge_off + (r - D.row_off) * k;
  // lane 0 finds lo, lane 1 hi; the group reads both
  const int64_t b = t == 1 ? lower_bound_u32(k1s, total_edges, (uint32_t)(r + 1))
                           : lower_bound_u32(k1s, total_edges, (uint32_t)r);
  const int64_t lo = (uint32_t)__shfl((int)(uint32_t)b, 0, G);
  const int64_t hi = (uint32_t)__shfl((int)(uint32_t)b, 1, G);
  if (k > G) {
    // more out-edges than lanes (k > 64, uniform over the group): lane t
    // takes slots t, t + G, ..., each by binary search of the in-edge sources
    uint32_t na = 0, nm = 0;
    for (int s = t; s < k; s += G) {
      const uint32_t c = key1[e0 + s];
      uint32_t less = 0, fbit = 0;
      float vt = 0.0f;
      if (c != sentinel) {
        int64_t a = lo, z = hi;  // first source >= c
        while (a < z) {
          const int64_t mid = (a + z) >> 1;
          if (srcrow[p1s[mid]] < c) a = mid + 1; else z = mid;
        }
        less = (uint32_t)(a - lo);
        if (a < hi) {
          const uint32_t e = p1s[a];
          if (srcrow[e] == c) { fbit = 0x80000000u; vt = wval[e]; ++nm; }
        }
        ++na;
      }
      ai[e0 + s] = less | fbit;
      av[e0 + s] = vt;
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
      na += __shfl_xor(na, o, G);
      nm += __shfl_xor(nm, o, G);
    }
    if (t == 0) {
      inptr[r] = (uint32_t)lo;
      if (r == total_rows - 1) inptr[total_rows] = (uint32_t)hi;
      cnt[r] = (uint64_t)na + (uint64_t)(hi - lo) - nm;
    }
    return;
  }
  const uint32_t ca = t < k ? key1[e0 + t] : sentinel;
  ScanIn si{ca, 0u, false, 0.0f};
  if (hi - lo <= 4 * G) {
    for (int64_t p0 = lo; p0 < hi; p0 += G) {
      const int64_t p = p0 + t;
      uint32_t src = 0xffffffffu;
      float w = 0.0f;
      if (p < hi) {
        const uint32_t e = p1s[p];
        src = srcrow[e];
        w = wval[e];
      }
      grp_for<G>(src, __float_as_uint(w), si);
    }
  } else if (ca != sentinel) {
    int64_t a = lo, z = hi;  // first source >= ca
    while (a < z) {
      const int64_t mid = (a + z) >> 1;
      if (srcrow[p1s[mid]] < ca) a = mid + 1; else z = mid;
    }
    si.less = (uint32_t)(a - lo);
    if (a < hi) {
      const uint32_t e = p1s[a];
      if (srcrow[e] == ca) { si.found = true; si.vt = wval[e]; }
    }
  }
  if (t < k) {
    ai[e0 + t] = si.less | (si.found ? 0x80000000u : 0u);
    av[e0 + t] = si.vt;
  }
  uint32_t na = ca != sentinel ? 1u : 0u, nm = si.found ? 1u : 0u;
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) {
    na += __shfl_xor(na, o, G);
    nm += __shfl_xor(nm, o, G);
  }
  if (t == 0) {
    inptr[r] = (uint32_t)lo;
    if (r == total_rows - 1) inptr[total_rows] = (uint32_t)hi;
    cnt[r] = (uint64_t)na + (uint64_t)(hi - lo) - nm;
  }
}

// The union's entries at indptr_g[r] + their rank in row r's sorted union,
// #{a in A_r: a < x} + #{i in I_r: i < x} - #{mutual m: m < x}:
//  * threads [0, R*G): row groups place the row's kept out-edges (and write
//    the dataset-local row pointer);
//  * threads [R*G, R*G + E): one per destination-sorted in-edge i -> r
//    without the out-edge r -> i (entry (r, i), v = 0, vt = w), counting
//    against the row's <= k out-edges, so rows of any in-degree spread over
//    as many threads as they have in-edges.
template <int G>
__global__ __launch_bounds__(256) void union_write_kernel(
    const DsDesc* __restrict__ ds, int nds, int64_t total_rows, int64_t total_edges,
    const uint32_t* __restrict__ key1, const float* __restrict__ wval,
    const uint32_t* __restrict__ k1s, const uint32_t* __restrict__ p1s,
    const uint32_t* __restrict__ srcrow, const uint32_t* __restrict__ inptr,
    const uint32_t* __restrict__ ai, const float* __restrict__ av,
    const uint64_t* __restrict__ indptr_g, const uint64_t* __restrict__ cnt,
    const int64_t* __restrict__ iptr_off, uint32_t sentinel, uint32_t* __restrict__ rows,
    uint32_t* __restrict__ cols, float* __restrict__ vals, uint64_t* __restrict__ indptr_l) {
  const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gt < total_rows * G) {
    const int64_t r = gt / G;
    const int t = (int)(gt % G);
    const int di = ds_of_row(ds, nds, r);
    const DsDesc& D = ds[di];
    const int k = D.k;
    const int64_t roff = D.row_off;
    const int64_t e0 = D.edge_off + (r - roff) * k;
    const uint64_t base = indptr_g[r];
    if (k > G) {
      // more out-edges than lanes: lane t places slots t, t + G, ..., each
      // ranked against the row's k columns read from memory
      for (int s = t; s < k; s += G) {
        const uint32_t ca = key1[e0 + s];
        if (ca == sentinel) continue;
        const float va = wval[e0 + s];
        const uint32_t a = ai[e0 + s];
        const float vt = av[e0 + s];
        uint32_t less = 0, mut = 0;
        for (int u = 0; u < k; ++u) {
          const uint32_t cu = key1[e0 + u];
          less += cu < ca ? 1u : 0u;
          mut += ((ai[e0 + u] >> 31) != 0u && cu < ca) ? 1u : 0u;
        }
        const uint64_t pos = base + less + (a & 0x7fffffffu) - mut;
        rows[pos] = (uint32_t)(r - roff);
        cols[pos] = (uint32_t)((int64_t)ca - roff);
        vals[pos] = (va + vt) - va * vt;  // f32, the reference's operand order
      }
    } else {
      const uint32_t ca = t < k ? key1[e0 + t] : sentinel;
      const float va = t < k ? wval[e0 + t] : 0.0f;
      const uint32_t a = t < k ? ai[e0 + t] : 0u;
      const float vt = t < k ? av[e0 + t] : 0.0f;
      // rank among the row's out-edges, and mutual columns below
      ScanOut so{ca, 0u, 0u, false};
      grp_for<G>(ca, a >> 31, so);
      if (ca != sentinel) {
        const uint64_t pos = base + so.less + (a & 0x7fffffffu) - so.mut;
        rows[pos] = (uint32_t)(r - roff);
        cols[pos] = (uint32_t)((int64_t)ca - roff);
        vals[pos] = (va + vt) - va * vt;  // f32, the reference's operand order
      }
    }
    if (t == 0) {
      // dataset di owns indptr_l[iptr_off[di] .. + n]
      const int64_t i = r - roff;
      const uint64_t b0 = indptr_g[roff];
      indptr_l[iptr_off[di] + i] = base - b0;
      if (i == D.n - 1) indptr_l[iptr_off[di] + D.n] = base + cnt[r] - b0;
    }
    return;
  }
  const int64_t p = gt - total_rows * G;
  if (p >= total_edges) return;
  const uint32_t r = k1s[p];
  if (r == sentinel) return;  // dropped edges sort last
  const uint32_t e = p1s[p];
  const uint32_t src = srcrow[e];
  const float w = wval[e];
  const DsDesc& D = ds[ds_of_row(ds, nds, r)];
  const int k = D.k;
  const int64_t roff = D.row_off;
  const int64_t e0 = D.edge_off + ((int64_t)r - roff) * k;
  bool in_a = false;
  uint32_t less_a = 0, mut = 0;
  for (int t = 0; t < k; ++t) {
    const uint32_t c = key1[e0 + t];  // sentinel (> every row) when dropped
    const bool m = (ai[e0 + t] >> 31) != 0u;
    in_a |= c == src;
    less_a += c < src ? 1u : 0u;
    mut += (m && c < src) ? 1u : 0u;
  }
  if (in_a) return;  // placed with the out-edge r -> src
  const uint64_t pos = indptr_g[r] + less_a + (uint64_t)(p - inptr[r]) - mut;
  const float v = 0.0f;
  rows[pos] = (uint32_t)((int64_t)r - roff);
  cols[pos] = (uint32_t)((int64_t)src - roff);
  vals[pos] = (v + w) - v * w;
}

#define HM_UNION_INST(G)                                                                        \
  template __global__ void union_count_kernel<G>(                                               \
      const DsDesc*, int, int64_t, int64_t, const uint32_t*, const float*, const uint32_t*,     \
      const uint32_t*, const uint32_t*, uint32_t, uint32_t*, uint64_t*, uint32_t*, float*);     \
  template __global__ void union_write_kernel<G>(                                               \
      const DsDesc*, int, int64_t, int64_t, const uint32_t*, const float*, const uint32_t*,     \
      const uint32_t*, const uint32_t*, const uint32_t*, const uint32_t*, const float*,         \
      const uint64_t*, const uint64_t*, const int64_t*, uint32_t, uint32_t*, uint32_t*, float*, \
      uint64_t*);
HM_UNION_INST(16)
HM_UNION_INST(32)
HM_UNION_INST(64)
#undef HM_UNION_INST

}  // namespace hm
