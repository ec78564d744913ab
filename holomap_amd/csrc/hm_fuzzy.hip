// Fuzzy-union symmetrisation (src/fuzzy.rs:49-84) and CSR (src/sparse.rs:21-35).
//
// The reference sorts the directed edges by (row, col), builds the sorted,
// de-duplicated union of A and A^T coordinates and fuses each pair with the
// probabilistic t-conorm w = (v + vt) - v*vt in f32. Here the same COO comes
// from two stable radix sorts (hm_sort.hip):
//   sort 1 by global column -> for every row c, its in-edges (sources
//            ascending, since edges enter in source order);
//   sort 2 of that sequence by source row -> every row's out-edges with
//            columns ascending;
// then one thread per row merges its out-list and in-list by column id,
// counting (pass 0) or writing (pass 1, at the exclusive-scan offset) the
// fused entries. The merge visits exactly the union coordinates in (row,
// col) order and w is evaluated with the reference's operand order, so the
// output is the reference's COO entry for entry.
#include "hm_internal.hpp"

namespace hm {

__global__ void gather_key2(const uint32_t* __restrict__ pay1s, const uint32_t* __restrict__ key1,
                            const uint32_t* __restrict__ srcrow, int64_t n, uint32_t sentinel,
                            uint32_t* __restrict__ key2) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t e = pay1s[p];
  key2[p] = key1[e] != sentinel ? srcrow[e] : sentinel;
}

// pass 0: cnt[r] = number of union entries of row r
// pass 1: write rows/cols (dataset-local ids) and vals at indptr_g[r]
__global__ void union_kernel(int pass, const DsDesc* __restrict__ ds, int nds, int64_t total_rows,
                             const uint32_t* __restrict__ inptr, const uint32_t* __restrict__ pay1s,
                             const uint32_t* __restrict__ outptr, const uint32_t* __restrict__ pay2s,
                             const uint32_t* __restrict__ key1, const uint32_t* __restrict__ srcrow,
                             const float* __restrict__ wval, uint64_t* __restrict__ cnt,
                             const uint64_t* __restrict__ indptr_g, uint32_t* __restrict__ rows,
                             uint32_t* __restrict__ cols, float* __restrict__ vals) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= total_rows) return;
  uint32_t a = outptr[r], ae = outptr[r + 1];
  uint32_t b = inptr[r], be = inptr[r + 1];
  int64_t roff = 0;
  uint64_t o = 0;
  if (pass == 1) {
    int lo = 0, hi = nds - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (ds[mid].row_off <= r) lo = mid; else hi = mid - 1;
    }
    roff = ds[lo].row_off;
    o = indptr_g[r];
  }
  uint64_t c_out = 0;
  const uint32_t INF = 0xffffffffu;
  while (a < ae || b < be) {
    const uint32_t ea = a < ae ? pay2s[a] : 0u;
    const uint32_t eb = b < be ? pay1s[b] : 0u;
    const uint32_t ca = a < ae ? key1[ea] : INF;
    const uint32_t cb = b < be ? srcrow[eb] : INF;
    uint32_t c;
    float v = 0.0f, vt = 0.0f;
    if (ca < cb) {
      c = ca; v = wval[ea]; ++a;
    } else if (cb < ca) {
      c = cb; vt = wval[eb]; ++b;
    } else {
      c = ca; v = wval[ea]; vt = wval[eb]; ++a; ++b;
    }
    const float w = (v + vt) - v * vt;  // f32, the reference's operand order
    if (w != 0.0f) {
      if (pass == 1) {
        rows[o + c_out] = (uint32_t)(r - roff);
        cols[o + c_out] = (uint32_t)((int64_t)c - roff);
        vals[o + c_out] = w;
      }
      ++c_out;
    }
  }
  if (pass == 0) cnt[r] = c_out;
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
