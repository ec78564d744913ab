// Fuzzy-union symmetrisation (src/fuzzy.rs:49-84) and CSR (src/sparse.rs:21-35).
//
// The reference sorts the directed edges by (row, col), builds the sorted,
// de-duplicated union of A and A^T coordinates and fuses each pair with the
// probabilistic t-conorm w = (v + vt) - v*vt in f32. Here the transpose is a
// counting sort by destination: the calibration counts every row's kept
// in-edges with integer atomics, each edge keeping the count it saw as its
// slot in the destination's bucket; an exclusive scan turns the counts into
// bucket offsets (in-edges of r: [off[r], off[r+1])), and
// transpose_scatter_kernel moves every kept edge i -> r to off[r] + its
// slot. The order inside a bucket therefore depends on scheduling, and
// nothing downstream reads it: every rank below is a count over the
// bucket's contents, so the output is a fixed function of the input.
// A_r: the row's kept out-edge columns (at most k), I_r: its in-edge
// sources.
//   union_count_kernel (a group of G lanes per row, lane t = out-edge slot t;
//     k > 64 rows: lane t takes slots t, t + 64, ...): for each out-edge its
//     count of smaller in-edge sources and its reverse weight; the union size
//     |A_r| + |I_r| - |A_r n I_r|;
//   (exclusive scan of the sizes -> global row pointer)
//   union_write_kernel: every entry goes to its rank in the sorted union,
//     #{a in A_r: a < x} + #{i in I_r: i < x} - #{m in A_r n I_r: m < x}, so
//     no row is sorted; out-edges and the in-edges of rows with few of them
//     by the row's group, the other in-edges one thread each (so no row's
//     in-degree serialises), plus the row pointers.
// A row of in-degree s costs O(s) per in-edge (its rank is a count over the
// bucket), O(s^2) in all: negligible at the in-degrees of kNN graphs (a few
// times k), and bounded by N^2 work for a row every point names.
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
// Kept edge e (key1[e] = its destination r, sentinel when dropped; erank[e]
// its slot in r's bucket) to in_dst / in_src / in_w[off[r] + erank[e]].
__global__ __launch_bounds__(256) void transpose_scatter_kernel(
    const DsDesc* __restrict__ ds, int nds, int64_t total_edges, const uint32_t* __restrict__ key1,
    const uint32_t* __restrict__ erank, const float* __restrict__ wval, uint32_t sentinel,
    const uint32_t* __restrict__ off, uint32_t* __restrict__ in_dst, uint32_t* __restrict__ in_src,
    float* __restrict__ in_w) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total_edges) return;
  const uint32_t r = key1[e];
  if (r == sentinel) return;
  int lo = 0, hi = nds - 1;  // the edge's dataset (edge_off ascending)
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ds[mid].edge_off <= e) lo = mid; else hi = mid - 1;
  }
  const DsDesc& D = ds[lo];
  const uint32_t slot = off[r] + erank[e];
  in_dst[slot] = r;
  in_src[slot] = (uint32_t)(D.row_off + (e - D.edge_off) / D.k);
  in_w[slot] = wval[e];
}

// Up to 16 consecutive u32 of a[q0 .. hi) (~0 past hi), loaded together so
// a short bucket costs one memory round trip.
__device__ __forceinline__ void load16(const uint32_t* __restrict__ a, int64_t q0, int64_t hi,
                                       uint32_t (&v)[16]) {
#pragma unroll
  for (int u = 0; u < 16; ++u) v[u] = q0 + u < hi ? a[q0 + u] : 0xffffffffu;
}

// Row group (G lanes, lane t = out-edge slot t): per kept out-edge r -> c,
// ai[e] = #{in-edge sources < c} | (c is one: bit 31) and av[e] = the
// weight of c -> r (0 if absent); cnt[r] = union size
// |A_r| + |I_r| - |A_r n I_r|. The bucket is read 4G slots at a time (one
// memory round trip) and its sources compared through group broadcasts.
template <int G>
__global__ __launch_bounds__(256) void union_count_kernel(
    const DsDesc* __restrict__ ds, int nds, int64_t total_rows, const uint32_t* __restrict__ key1,
    const uint32_t* __restrict__ off, const uint32_t* __restrict__ in_src,
    const float* __restrict__ in_w, uint32_t sentinel, uint64_t* __restrict__ cnt,
    uint32_t* __restrict__ ai, float* __restrict__ av) {
  static_assert(G >= 2 && G <= 64 && (G & (G - 1)) == 0, "group size");
  const int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int t = (int)(threadIdx.x % G);
  if (r >= total_rows) return;  // whole groups leave together
  const DsDesc& D = ds[ds_of_row(ds, nds, r)];
  const int k = D.k;
  const int64_t e0 = D.edge_off + (r - D.row_off) * k;
  const int64_t lo = off[r], hi = off[r + 1];
  uint32_t na = 0, nm = 0;
  if (k > G) {
    // more out-edges than lanes (k > 64, uniform over the group): lane t
    // takes slots t, t + G, ..., each against the whole bucket
    for (int s = t; s < k; s += G) {
      const uint32_t ca = key1[e0 + s];
      uint32_t less = 0, fbit = 0;
      float vt = 0.0f;
      if (ca != sentinel) {
        for (int64_t p = lo; p < hi; ++p) {
          const uint32_t src = in_src[p];
          less += src < ca ? 1u : 0u;
          if (src == ca) { fbit = 0x80000000u; vt = in_w[p]; }
        }
        nm += fbit ? 1u : 0u;
        ++na;
      }
      ai[e0 + s] = less | fbit;
      av[e0 + s] = vt;
    }
  } else {
    const uint32_t ca = t < k ? key1[e0 + t] : sentinel;
    ScanIn si{ca, 0u, false, 0.0f};
    const int nch = (int)((hi - lo + G - 1) / G);  // chunks of G slots (uniform)
    for (int c0 = 0; c0 < nch; c0 += 4) {
      uint32_t bs[4];  // slots lo + G (c0 + j) + t
      float bw[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t q = lo + (int64_t)G * (c0 + j) + t;
        bs[j] = q < hi ? in_src[q] : 0xffffffffu;
        bw[j] = q < hi ? in_w[q] : 0.0f;
      }
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (c0 + j < nch) grp_for<G>(bs[j], __float_as_uint(bw[j]), si);
    }
    if (t < k) {
      ai[e0 + t] = si.less | (si.found ? 0x80000000u : 0u);
      av[e0 + t] = si.vt;
    }
    na = ca != sentinel ? 1u : 0u;
    nm = si.found ? 1u : 0u;
  }
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) {
    na += __shfl_xor(na, o, G);
    nm += __shfl_xor(nm, o, G);
  }
  if (t == 0) cnt[r] = (uint64_t)na + (uint64_t)(hi - lo) - nm;
}

// The union's entries at indptr_g[r] + their rank in row r's sorted union,
// #{a in A_r: a < x} + #{i in I_r: i < x} - #{mutual m: m < x}. Equal
// sources in one bucket (a row naming a column twice) rank by slot. Entries
// at or past cap (the arrays' length, 2E) are never written.
//  * threads [0, R*G): row groups place the row's kept out-edges and write
//    the dataset-local row pointer;
//  * threads [R*G, R*G + E): one per in-edge slot p < off[R] (in-edge
//    i -> r) without the out-edge r -> i (entry (r, i), v = 0, vt = w),
//    counting against the row's out-edges and its bucket, so rows of any
//    in-degree spread over as many threads as they have in-edges;
//  * workgroups from fix_wg on (when zany is set): the calibration's
//    global-mean floor of dataset blockIdx.x - fix_wg (calib_fixup_ds), one
//    launch fewer than a kernel of its own.
template <int G>
__global__ __launch_bounds__(256) void union_write_kernel(
    const DsDesc* __restrict__ ds, int nds, int64_t total_rows, int64_t total_edges,
    const uint32_t* __restrict__ key1, const float* __restrict__ wval,
    const uint32_t* __restrict__ off, const uint32_t* __restrict__ in_dst,
    const uint32_t* __restrict__ in_src, const float* __restrict__ in_w,
    const uint32_t* __restrict__ ai, const float* __restrict__ av,
    const uint64_t* __restrict__ indptr_g, const int64_t* __restrict__ iptr_off,
    uint32_t sentinel, uint64_t cap, uint32_t* __restrict__ rows, uint32_t* __restrict__ cols,
    float* __restrict__ vals, uint64_t* __restrict__ indptr_l, uint32_t fix_wg, int calib_mode,
    float* __restrict__ sigma, double* __restrict__ sigma64, const float* __restrict__ rho,
    uint32_t* __restrict__ zany) {
  if (zany && blockIdx.x >= fix_wg) {  // uniform over the workgroup
    const int dsi = (int)(blockIdx.x - fix_wg);
    calib_fixup_ds(ds[dsi], dsi, calib_mode, sigma, sigma64, rho, zany);
    return;
  }
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
        if (pos >= cap) continue;
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
      const uint64_t pos = base + so.less + (a & 0x7fffffffu) - so.mut;
      if (ca != sentinel && pos < cap) {
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
      if (i == D.n - 1) indptr_l[iptr_off[di] + D.n] = indptr_g[r + 1] - b0;
    }
    return;
  }
  const int64_t p = gt - total_rows * G;
  if (p >= total_edges) return;
  // (slots past off[R] hold no in-edge; their loads stay inside the arrays)
  const uint32_t r = in_dst[p], src = in_src[p];
  const float w = in_w[p];
  if (p >= (int64_t)off[total_rows]) return;
  const DsDesc& D = ds[ds_of_row(ds, nds, r)];
  const int64_t roff = D.row_off;
  const int k = D.k;
  const int64_t lo = off[r], hi = off[r + 1];
  const int64_t e0 = D.edge_off + ((int64_t)r - roff) * k;
  bool in_a = false;
  uint32_t less_a = 0, mut = 0;
  for (int t0 = 0; t0 < k; t0 += 16) {
    uint32_t c[16], m[16];  // columns: sentinel (> every row) when dropped
    load16(key1 + e0, t0, k, c);
    load16(ai + e0, t0, k, m);
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      in_a |= c[u] == src;
      less_a += c[u] < src ? 1u : 0u;
      mut += (c[u] < src && (m[u] >> 31) != 0u) ? 1u : 0u;
    }
  }
  if (in_a) return;  // placed with the out-edge r -> src
  uint64_t less_i = 0;  // bucket sources below src (equal ones by slot), 32 a trip
  for (int64_t q0 = lo; q0 < hi; q0 += 32) {
    uint32_t v0[16], v1[16];
    load16(in_src, q0, hi, v0);
    load16(in_src, q0 + 16, hi, v1);
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      less_i += (v0[u] < src || (v0[u] == src && q0 + u < p)) ? 1u : 0u;
      less_i += (v1[u] < src || (v1[u] == src && q0 + 16 + u < p)) ? 1u : 0u;
    }
  }
  const uint64_t pos = indptr_g[r] + less_a + less_i - mut;
  if (pos >= cap) return;
  const float v = 0.0f;
  rows[pos] = (uint32_t)((int64_t)r - roff);
  cols[pos] = (uint32_t)((int64_t)src - roff);
  vals[pos] = (v + w) - v * w;
}

#define HM_UNION_INST(G)                                                                        \
  template __global__ void union_count_kernel<G>(                                               \
      const DsDesc*, int, int64_t, const uint32_t*, const uint32_t*, const uint32_t*,           \
      const float*, uint32_t, uint64_t*, uint32_t*, float*);                                    \
  template __global__ void union_write_kernel<G>(                                               \
      const DsDesc*, int, int64_t, int64_t, const uint32_t*, const float*, const uint32_t*,     \
      const uint32_t*, const uint32_t*, const float*, const uint32_t*, const float*,            \
      const uint64_t*, const int64_t*, uint32_t, uint64_t, uint32_t*, uint32_t*, float*,        \
      uint64_t*, uint32_t, int, float*, double*, const float*, uint32_t*);
HM_UNION_INST(16)
HM_UNION_INST(32)
HM_UNION_INST(64)
#undef HM_UNION_INST

}  // namespace hm
