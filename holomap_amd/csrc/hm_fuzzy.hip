// Fuzzy-union symmetrisation (src/fuzzy.rs:49-84) and CSR (src/sparse.rs:21-35).
//
// The reference sorts the directed edges by (row, col), builds the sorted,
// de-duplicated union of A and A^T coordinates and fuses each pair with the
// probabilistic t-conorm w = (v + vt) - v*vt in f32. Here the same COO comes
// from one stable radix sort (hm_sort.hip) and a per-row rank sort:
//   sort 1 by global column -> every row c's in-edges, sources ascending;
//   each row's kept out-edges ranked by column within the row (a lane group
//            per row: a row has at most k of them) -> list A_r, columns
//            ascending, at outptr[r] = exclusive scan of the kept counts;
// then, load-balanced (one thread per element, no per-row loops, so rows of
// very high in-degree cost nothing extra):
//   * each out-edge r->c looks up its reverse c->r in A_c by binary search
//     (vt, or 0 if absent);
//   * the in-edges i->c without a reverse form list B_c (sources ascending);
//   * row r of the union is A_r merged with B_r (disjoint column sets), and
//     every element finds its output slot as its rank in its own list plus
//     the number of smaller columns in the other list ("merge path").
// For an entry (r, c): v = A[r][c], vt = A[c][r] (0 if absent), w evaluated
// with the reference's operand order. With v, vt in [0, 1] and v + vt > 0,
// w >= max(v, vt) - min(v, vt) and w == 0 would need v == vt == v*vt, so
// w > 0 always and the reference's drop of zero entries (src/fuzzy.rs:73)
// never applies.
#include "hm_internal.hpp"

namespace hm {

// A lists, G lanes per row (G a power of two <= 64; lane t holds out-edge
// slots t, t + G, ... of the row). PLACE 0: cnt[r] = kept out-edges of row
// r. PLACE 1: each kept edge goes to outptr[r] + (number of kept edges of
// the row with a smaller column): ocol/oval/p2s (column, weight, edge id)
// and k2s (= r) at that slot.
template <int G, int PLACE>
__global__ __launch_bounds__(256) void union_rows_a(
    const DsDesc* __restrict__ ds, int nds, int64_t total_rows, const uint32_t* __restrict__ key1,
    const float* __restrict__ wval, uint32_t sentinel, uint32_t* __restrict__ cnt,
    const uint32_t* __restrict__ outptr, uint32_t* __restrict__ ocol, float* __restrict__ oval,
    uint32_t* __restrict__ p2s, uint32_t* __restrict__ k2s) {
  static_assert(G >= 2 && G <= 64 && (G & (G - 1)) == 0, "group size");
  const int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int t = (int)(threadIdx.x % G);
  if (r >= total_rows) return;  // whole groups leave together
  int lo = 0, hi = nds - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ds[mid].row_off <= r) lo = mid; else hi = mid - 1;
  }
  const DsDesc& D = ds[lo];
  const int k = D.k;
  const int64_t e0 = D.edge_off + (r - D.row_off) * k;  // the row's first slot
  if constexpr (PLACE == 0) {
    uint32_t n_kept = 0;
    for (int s = t; s < k; s += G) n_kept += key1[e0 + s] != sentinel ? 1u : 0u;
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) n_kept += __shfl_xor(n_kept, o, G);
    if (t == 0) cnt[r] = n_kept;
  } else {
    auto place = [&](int s, uint32_t c, uint32_t rank) {
      const uint32_t pos = outptr[r] + rank;
      ocol[pos] = c;
      oval[pos] = wval[e0 + s];
      p2s[pos] = (uint32_t)(e0 + s);
      k2s[pos] = (uint32_t)r;
    };
    if (k <= G) {
      // one slot per lane; every lane of the group takes part in the shuffles
      const uint32_t c = t < k ? key1[e0 + t] : sentinel;
      uint32_t rank = 0;
      for (int u = 0; u < k; ++u) {
        const uint32_t cu = __shfl(c, u, G);
        rank += (cu != sentinel && cu < c) ? 1u : 0u;
      }
      if (c != sentinel) place(t, c, rank);
    } else {
      for (int s = t; s < k; s += G) {
        const uint32_t c = key1[e0 + s];
        if (c == sentinel) continue;
        uint32_t rank = 0;
        for (int u = 0; u < k; ++u) {
          const uint32_t cu = key1[e0 + u];
          rank += (cu != sentinel && cu < c) ? 1u : 0u;
        }
        place(s, c, rank);
      }
    }
  }
}

template __global__ void union_rows_a<16, 0>(const DsDesc*, int, int64_t, const uint32_t*,
                                             const float*, uint32_t, uint32_t*, const uint32_t*,
                                             uint32_t*, float*, uint32_t*, uint32_t*);
template __global__ void union_rows_a<16, 1>(const DsDesc*, int, int64_t, const uint32_t*,
                                             const float*, uint32_t, uint32_t*, const uint32_t*,
                                             uint32_t*, float*, uint32_t*, uint32_t*);
template __global__ void union_rows_a<32, 0>(const DsDesc*, int, int64_t, const uint32_t*,
                                             const float*, uint32_t, uint32_t*, const uint32_t*,
                                             uint32_t*, float*, uint32_t*, uint32_t*);
template __global__ void union_rows_a<32, 1>(const DsDesc*, int, int64_t, const uint32_t*,
                                             const float*, uint32_t, uint32_t*, const uint32_t*,
                                             uint32_t*, float*, uint32_t*, uint32_t*);
template __global__ void union_rows_a<64, 0>(const DsDesc*, int, int64_t, const uint32_t*,
                                             const float*, uint32_t, uint32_t*, const uint32_t*,
                                             uint32_t*, float*, uint32_t*, uint32_t*);
template __global__ void union_rows_a<64, 1>(const DsDesc*, int, int64_t, const uint32_t*,
                                             const float*, uint32_t, uint32_t*, const uint32_t*,
                                             uint32_t*, float*, uint32_t*, uint32_t*);

__device__ __forceinline__ uint32_t lb_u32(const uint32_t* __restrict__ a, uint32_t lo, uint32_t hi,
                                           uint32_t v) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Reverse lookup of every out-edge r->c in A_c; marks mutual edges by id.
__global__ void union_mutual(const uint32_t* __restrict__ k2s, const uint32_t* __restrict__ p2s,
                             const uint32_t* __restrict__ ocol, const float* __restrict__ oval,
                             const uint32_t* __restrict__ outptr, int64_t total_rows,
                             float* __restrict__ orev, uint32_t* __restrict__ mflag) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= (int64_t)outptr[total_rows]) return;
  const uint32_t r = k2s[a], c = ocol[a];
  const uint32_t lo = outptr[c], hi = outptr[c + 1];
  const uint32_t f = lb_u32(ocol, lo, hi, r);
  const bool found = f < hi && ocol[f] == r;
  orev[a] = found ? oval[f] : 0.0f;
  mflag[p2s[a]] = found ? 1u : 0u;
}

// nm[p] = 1 for a valid in-edge (sort-1 position p) without a reverse edge.
__global__ void union_nonmutual(const uint32_t* __restrict__ p1s, const uint32_t* __restrict__ mflag,
                                const uint32_t* __restrict__ inptr, int64_t total_rows,
                                int64_t total_edges, uint32_t* __restrict__ nm) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p > total_edges) return;
  nm[p] = (p < (int64_t)inptr[total_rows] && !mflag[p1s[p]]) ? 1u : 0u;
}

// B lists (order kept) and their row pointers.
__global__ void union_gather_b(const uint32_t* __restrict__ k1s, const uint32_t* __restrict__ p1s,
                               const uint32_t* __restrict__ srcrow, const float* __restrict__ wval,
                               const uint32_t* __restrict__ nm, const uint32_t* __restrict__ bpos,
                               const uint32_t* __restrict__ inptr, int64_t total_rows,
                               uint32_t* __restrict__ bsrc, uint32_t* __restrict__ brow,
                               float* __restrict__ bval) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (int64_t)inptr[total_rows] || !nm[p]) return;
  const uint32_t b = bpos[p], e = p1s[p];
  bsrc[b] = srcrow[e];
  brow[b] = k1s[p];
  bval[b] = wval[e];
}

// bptr[c] = first B slot of row c; cnt[r] = |A_r| + |B_r|.
__global__ void union_rowcount(const uint32_t* __restrict__ inptr, const uint32_t* __restrict__ outptr,
                               const uint32_t* __restrict__ bpos, int64_t total_rows,
                               uint32_t* __restrict__ bptr, uint64_t* __restrict__ cnt) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r > total_rows) return;
  const uint32_t b0 = bpos[inptr[r]];
  bptr[r] = b0;
  if (r < total_rows) {
    const uint32_t b1 = bpos[inptr[r + 1]];
    cnt[r] = (uint64_t)(outptr[r + 1] - outptr[r]) + (uint64_t)(b1 - b0);
  }
}

__device__ __forceinline__ int64_t row_offset_of(const DsDesc* __restrict__ ds, int nds, int64_t r) {
  int lo = 0, hi = nds - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ds[mid].row_off <= r) lo = mid; else hi = mid - 1;
  }
  return ds[lo].row_off;
}

// A elements: entry (r, c) with v = A[r][c], vt = A[c][r].
__global__ void union_scatter_a(const DsDesc* __restrict__ ds, int nds,
                                const uint32_t* __restrict__ k2s, const uint32_t* __restrict__ ocol,
                                const float* __restrict__ oval, const float* __restrict__ orev,
                                const uint32_t* __restrict__ outptr, const uint32_t* __restrict__ bsrc,
                                const uint32_t* __restrict__ bptr,
                                const uint64_t* __restrict__ indptr_g, int64_t total_rows,
                                uint32_t* __restrict__ rows, uint32_t* __restrict__ cols,
                                float* __restrict__ vals) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= (int64_t)outptr[total_rows]) return;
  const uint32_t r = k2s[a], c = ocol[a];
  const uint32_t nb = lb_u32(bsrc, bptr[r], bptr[r + 1], c) - bptr[r];
  const uint64_t pos = indptr_g[r] + (uint64_t)(a - outptr[r]) + nb;
  const float v = oval[a], vt = orev[a];
  const int64_t roff = row_offset_of(ds, nds, r);
  rows[pos] = (uint32_t)(r - roff);
  cols[pos] = (uint32_t)((int64_t)c - roff);
  vals[pos] = (v + vt) - v * vt;  // f32, the reference's operand order
}

// B elements: entry (c, i) for an edge i->c without reverse: v = 0, vt = w.
__global__ void union_scatter_b(const DsDesc* __restrict__ ds, int nds,
                                const uint32_t* __restrict__ brow, const uint32_t* __restrict__ bsrc,
                                const float* __restrict__ bval, const uint32_t* __restrict__ bptr,
                                const uint32_t* __restrict__ ocol, const uint32_t* __restrict__ outptr,
                                const uint64_t* __restrict__ indptr_g, int64_t total_rows,
                                uint32_t* __restrict__ rows, uint32_t* __restrict__ cols,
                                float* __restrict__ vals) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= (int64_t)bptr[total_rows]) return;
  const uint32_t c = brow[b], i = bsrc[b];
  const uint32_t na = lb_u32(ocol, outptr[c], outptr[c + 1], i) - outptr[c];
  const uint64_t pos = indptr_g[c] + (uint64_t)(b - bptr[c]) + na;
  const float v = 0.0f, vt = bval[b];
  const int64_t roff = row_offset_of(ds, nds, c);
  rows[pos] = (uint32_t)(c - roff);
  cols[pos] = (uint32_t)((int64_t)i - roff);
  vals[pos] = (v + vt) - v * vt;
}

// Per-dataset CSR row pointer (dataset-local) from the global offsets.
__global__ void local_indptr_kernel(const DsDesc* __restrict__ ds, int nds, int64_t total_rows,
                                    const uint64_t* __restrict__ indptr_g,
                                    const uint64_t* __restrict__ total_nnz,
                                    const int64_t* __restrict__ iptr_off,
                                    uint64_t* __restrict__ indptr_l) {
  // thread per global row slot g in [0, total_rows + nds): each dataset owns
  // n+1 entries starting at iptr_off[ds] = row_off + ds
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total_rows + nds) return;
  int lo = 0, hi = nds - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (iptr_off[mid] <= t) lo = mid; else hi = mid - 1;
  }
  const DsDesc& D = ds[lo];
  const int64_t i = t - iptr_off[lo];  // 0..n
  const int64_t g = D.row_off + i;
  const uint64_t base = indptr_g[D.row_off];
  const uint64_t at = g < total_rows ? indptr_g[g] : *total_nnz;
  indptr_l[t] = at - base;
}

// CSR values widened to f64 (src/sparse.rs:32).
__global__ void widen_kernel(const float* __restrict__ v, int64_t n, double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = (double)v[t];
}

}  // namespace hm
