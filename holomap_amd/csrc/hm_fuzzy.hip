// Fuzzy-union symmetrisation (src/fuzzy.rs:49-84) and CSR (src/sparse.rs:21-35).
//
// The reference sorts the directed edges by (row, col), builds the sorted,
// de-duplicated union of A and A^T coordinates and fuses each pair with the
// probabilistic t-conorm w = (v + vt) - v*vt in f32. Here the transpose is a
// counting sort by destination: the calibration counts every row's kept
// in-edges with integer atomics, each edge keeping the count it saw as its
// slot in the destination's bucket; an exclusive scan turns the counts into
// bucket offsets (in-edges of r: [off[r], off[r+1])), and
// sym_scatter_kernel moves every kept edge i -> r to off[r] + its slot. The
// order inside a bucket therefore depends on scheduling, and nothing
// downstream reads it: every rank below is a count over the bucket's
// contents, so the output is a fixed function of the input.
// A_r: the row's kept out-edge columns (at most k), I_r: its in-edge
// sources; an in-edge i -> r is mutual when r -> i is kept too. The union of
// row r is A_r plus its non-mutual in-edge sources:
//   sym_scatter_kernel (a group of G lanes per row): the transpose, each
//     edge marked mutual or not as it is placed (a look at the destination's
//     k columns), and the union size |A_r| + |I_r| - #mutual, no bucket
//     scanned;
//   (exclusive scan of the sizes -> global row pointer)
//   sym_write_kernel: every entry at its rank in the sorted union,
//     #{a in A_r: a < x} + #{non-mutual i in I_r: i < x}, so no row is
//     sorted; the buckets a workgroup's entries rank against pass through
//     LDS a chunk at a time (a row every point names costs its in-degree per
//     entry, spread over the workgroups that hold its entries).
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

// Up to 16 consecutive u32 of a[q0 .. hi) (~0 past hi), loaded together
// (one memory round trip for a row of up to 16 columns).
__device__ __forceinline__ void load16(const uint32_t* __restrict__ a, int64_t q0, int64_t hi,
                                       uint32_t (&v)[16]) {
#pragma unroll
  for (int u = 0; u < 16; ++u) v[u] = q0 + u < hi ? a[q0 + u] : 0xffffffffu;
}

__device__ __forceinline__ int ds_of_edge(const DsDesc* __restrict__ ds, int nds, int64_t e) {
  int lo = 0, hi = nds - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ds[mid].edge_off <= e) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Position of x among the k columns key1[ec .. ec + k), or -1: the row
// read as the 16-byte quads that cover it, 8 quads a trip (one memory round
// trip for k <= 29; the quad past the array's end read as single words).
__device__ __forceinline__ int find_in_row(const uint32_t* __restrict__ key1, int64_t ec, int k,
                                           int64_t E, uint32_t x) {
  const uint4* k4 = (const uint4*)key1;
  const int64_t q0 = ec >> 2, q1 = (ec + k + 3) >> 2;
  int found = -1;
  for (int64_t q = q0; q < q1; q += 8) {
    uint32_t v[32];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int64_t qq = q + u;
      if (qq < q1 && qq * 4 + 4 <= E) {
        const uint4 t = k4[qq];
        v[4 * u] = t.x; v[4 * u + 1] = t.y; v[4 * u + 2] = t.z; v[4 * u + 3] = t.w;
      } else {
#pragma unroll
        for (int c = 0; c < 4; ++c)
          v[4 * u + c] = (qq < q1 && qq * 4 + c < E) ? key1[qq * 4 + c] : 0xffffffffu;
      }
    }
#pragma unroll
    for (int m = 0; m < 32; ++m) {
      const int64_t pos = q * 4 + m - ec;
      if (pos >= 0 && pos < k && v[m] == x && found < 0) found = (int)pos;
    }
  }
  return found;
}

// A group of G lanes per row i, lane t taking out-edge slots t, t + G, ...:
// kept edge e = i -> c goes to slot off[c] + erank[e] of c's bucket
// (in_dst = c; in_key = i, or ~0 when c -> i is kept too: the pair is placed
// from c's own out-edge; in_w = its weight) and revw[e] = the weight of
// c -> i when that edge is kept, else 0; cnt[i] = the union size. A slot at
// or past E (never, with the counts the calibration made) is not written
// and sets *err.
template <int G>
__global__ __launch_bounds__(256) void sym_scatter_kernel(
    const DsDesc* __restrict__ ds, int nds, int64_t R, int64_t E,
    const uint32_t* __restrict__ key1, const uint32_t* __restrict__ erank,
    const float* __restrict__ wval, const uint32_t* __restrict__ off,
    uint32_t* __restrict__ in_dst, uint32_t* __restrict__ in_key, float* __restrict__ in_w,
    float* __restrict__ revw, uint64_t* __restrict__ cnt, uint32_t* __restrict__ err) {
  static_assert(G >= 2 && G <= 64 && (G & (G - 1)) == 0, "group size");
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int t = (int)(threadIdx.x % G);
  if (i >= R) return;  // whole groups leave together
  const DsDesc& D = ds[ds_of_row(ds, nds, i)];
  const int k = D.k;
  const int64_t e0 = D.edge_off + (i - D.row_off) * k;
  // the row's bucket size (read up front: it is only needed at the end)
  const uint32_t si = t == 0 ? off[i + 1] - off[i] : 0u;
  uint32_t na = 0, nm = 0;
  for (int s = t; s < k; s += G) {
    const int64_t e = e0 + s;
    const uint32_t c = key1[e];
    if (c >= (uint64_t)R) continue;  // dropped (sentinel R)
    const int64_t ec = D.edge_off + ((int64_t)c - D.row_off) * k;  // c's out-edges
    const float w = wval[e];
    const uint32_t rank = erank[e];
    const uint32_t oc = off[c];
    const int found = find_in_row(key1, ec, k, E, (uint32_t)i);
    revw[e] = found >= 0 ? wval[ec + found] : 0.0f;
    ++na;
    nm += found >= 0 ? 1u : 0u;
    const uint64_t slot = (uint64_t)oc + rank;
    if (slot >= (uint64_t)E) {
      *err = 1u;
      continue;
    }
    in_dst[slot] = c;
    in_key[slot] = found >= 0 ? 0xffffffffu : (uint32_t)i;
    in_w[slot] = w;
  }
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) {
    na += __shfl_xor(na, o, G);
    nm += __shfl_xor(nm, o, G);
  }
  if (t == 0) cnt[i] = (uint64_t)na + si - nm;
}

// The end H of a workgroup's bucket span [L, H), checked against the kept
// edges (a bad bound -- never, with consistent counts -- sets *err and
// empties the span, so no load leaves the bucket arrays).
__device__ __forceinline__ int64_t span_end(int64_t H, int64_t L, int64_t ptot,
                                            uint32_t* __restrict__ err) {
  if (H > ptot || H < L) {
    *err = 1u;
    return L;
  }
  return H;
}

// Keys in sk[a, b) below x (LE: at most x), 16 a trip from 16-byte LDS
// reads once aligned (four reads in flight: the loop is bound by LDS
// latency otherwise).
template <bool LE>
__device__ __forceinline__ uint32_t count_keys(const uint32_t* __restrict__ sk, int a, int b,
                                               uint32_t x) {
  auto c1 = [&](uint32_t v) { return (LE ? v <= x : v < x) ? 1u : 0u; };
  auto c4 = [&](uint4 q) { return c1(q.x) + c1(q.y) + c1(q.z) + c1(q.w); };
  uint32_t n = 0;
  int j = a;
  for (; j < b && (j & 3) != 0; ++j) n += c1(sk[j]);
  const uint4* s4 = (const uint4*)sk;
  for (; j + 16 <= b; j += 16) {
    const uint4 q0 = s4[j >> 2], q1 = s4[(j >> 2) + 1], q2 = s4[(j >> 2) + 2], q3 = s4[(j >> 2) + 3];
    n += c4(q0) + c4(q1) + c4(q2) + c4(q3);
  }
  for (; j + 4 <= b; j += 4) n += c4(s4[j >> 2]);
  for (; j < b; ++j) n += c1(sk[j]);
  return n;
}

// Number of keys in bucket slots [lo, hi) below x (equal ones at slots
// below p: at most x before p, below x from p on), for every thread of the
// workgroup at once: the slots [L, H) that the workgroup's rows' buckets
// span pass through LDS (sk, kSymChunk entries, 16-byte aligned) a chunk at
// a time, each thread counting its own range. L and H are uniform over the
// workgroup; a thread with nothing to count passes an empty range.
__device__ __forceinline__ uint32_t span_count(const uint32_t* __restrict__ in_key, int64_t L,
                                               int64_t H, int64_t lo, int64_t hi, uint32_t x,
                                               int64_t p, uint32_t* __restrict__ sk) {
  uint32_t cnt = 0;
  for (int64_t c0 = L; c0 < H; c0 += kSymChunk) {
#pragma unroll
    for (int u = 0; u < kSymChunk / 256; ++u) {
      const int64_t q = c0 + u * 256 + threadIdx.x;
      sk[u * 256 + threadIdx.x] = q < H ? in_key[q] : 0xffffffffu;
    }
    __syncthreads();
    auto rel = [&](int64_t v) {  // v - c0 clamped to [0, kSymChunk]
      const int64_t d = v - c0;
      return (int)(d < 0 ? 0 : (d > kSymChunk ? kSymChunk : d));
    };
    const int a = rel(lo), b = rel(hi), pt = rel(p);
    const int m = pt < a ? a : (pt > b ? b : pt);
    cnt += count_keys<true>(sk, a, m, x) + count_keys<false>(sk, m, b, x);
    __syncthreads();
  }
  return cnt;
}

// #{a in A_r: a < x}: the row's out-edge columns from key1 (the sentinel R,
// for a dropped edge, is above every column).
__device__ __forceinline__ uint32_t less_in_row(const uint32_t* __restrict__ key1, int64_t e0,
                                                int k, uint32_t x) {
  uint32_t n = 0;
  for (int t0 = 0; t0 < k; t0 += 16) {
    uint32_t cc[16];
    load16(key1 + e0, t0, k, cc);
#pragma unroll
    for (int u = 0; u < 16; ++u) n += cc[u] < x ? 1u : 0u;
  }
  return n;
}

// The union's entries, by workgroup (indptr_g: the exclusive scan of the
// union sizes, its last entry the total):
//  * [0, nrt): 256 rows each: the dataset-local row pointers indptr_l;
//  * [nrt, nrt + nslot): 256 directed edges each: kept edge r -> c to
//    indptr_g[r] + #{a in A_r: a < c} + #{non-mutual in-edge sources of r
//    below c}, value (v + vt) - v*vt with vt = revw (0 if absent);
//  * [nrt + nslot, nrt + 2 nslot): 256 bucket slots each: non-mutual in-edge
//    i -> r to indptr_g[r] + #{a in A_r: a < i} + #{non-mutual sources of r's
//    bucket below i, equal ones at lower slots}, value (0 + w) - 0*w;
//  * workgroups from nrt + 2 nslot (when zany is set): the calibration's
//    global-mean floor of dataset blockIdx.x - (nrt + 2 nslot)
//    (calib_fixup_ds), one launch fewer than a kernel of its own.
// Entries at or past cap (2E) are never written; a row id read from the
// buckets at or past R (never, with consistent counts) sets *err and is
// skipped.
__global__ __launch_bounds__(256) void sym_write_kernel(
    const DsDesc* __restrict__ ds, int nds, int64_t R, int64_t E,
    const uint32_t* __restrict__ key1, const float* __restrict__ wval,
    const float* __restrict__ revw, const uint32_t* __restrict__ off,
    const uint32_t* __restrict__ in_dst, const uint32_t* __restrict__ in_key,
    const float* __restrict__ in_w, const uint64_t* __restrict__ indptr_g,
    const int64_t* __restrict__ iptr_off, uint64_t cap, uint32_t* __restrict__ rows,
    uint32_t* __restrict__ cols, float* __restrict__ vals, uint64_t* __restrict__ indptr_l,
    uint32_t nrt, uint32_t nslot, uint32_t* __restrict__ err, int calib_mode,
    float* __restrict__ sigma, double* __restrict__ sigma64, const float* __restrict__ rho,
    uint32_t* __restrict__ zany) {
  const uint32_t b = blockIdx.x;
  const int tid = threadIdx.x;
  if (b < nrt) {
    const int64_t r = (int64_t)b * 256 + tid;
    if (r >= R) return;
    const int di = ds_of_row(ds, nds, r);
    const DsDesc& D = ds[di];
    const int64_t i = r - D.row_off;
    const uint64_t b0 = indptr_g[D.row_off];
    indptr_l[iptr_off[di] + i] = indptr_g[r] - b0;
    if (i == D.n - 1) indptr_l[iptr_off[di] + D.n] = indptr_g[r + 1] - b0;
    return;
  }
  if (b >= nrt + 2 * nslot) {
    if (zany) {
      const int dsi = (int)(b - nrt - 2 * nslot);
      calib_fixup_ds(ds[dsi], dsi, calib_mode, sigma, sigma64, rho, zany);
    }
    return;
  }
  __shared__ __attribute__((aligned(16))) uint32_t sk[kSymChunk];
  __shared__ int64_t span_s[2];
  const int64_t ptot = off[R] < (uint64_t)E ? (int64_t)off[R] : E;
  if (b < nrt + nslot) {
    const int64_t eb = (int64_t)(b - nrt) * 256;
    const int64_t ee = eb + 256 < E ? eb + 256 : E;
    auto row_of = [&](int64_t e) {
      const DsDesc& D = ds[ds_of_edge(ds, nds, e)];
      return D.row_off + (int64_t)((uint32_t)(e - D.edge_off) / (uint32_t)D.k);
    };
    if (tid == 0) {
      span_s[0] = off[row_of(eb)];
      span_s[1] = span_end(off[row_of(ee - 1) + 1], span_s[0], ptot, err);
    }
    const int64_t e = eb + tid;
    bool act = false;
    uint32_t c = 0, la = 0;
    int64_t r = 0, roff = 0, lo = 0, hi = 0;
    uint64_t base = 0;
    float va = 0.0f, vt = 0.0f;
    if (e < E) {
      c = key1[e];
      if (c < (uint64_t)R) {  // kept
        const DsDesc& D = ds[ds_of_edge(ds, nds, e)];
        act = true;
        roff = D.row_off;
        r = roff + (int64_t)((uint32_t)(e - D.edge_off) / (uint32_t)D.k);
        lo = off[r];
        hi = off[r + 1];
        base = indptr_g[r];
        va = wval[e];
        vt = revw[e];
        la = less_in_row(key1, D.edge_off + (r - roff) * D.k, D.k, c);
      }
    }
    __syncthreads();
    const int64_t L = span_s[0], H = span_s[1];
    if (!act) lo = hi = L;
    const uint32_t li = span_count(in_key, L, H, lo, hi, c, lo, sk);
    if (!act) return;
    const uint64_t pos = base + la + li;
    if (pos >= cap) return;
    rows[pos] = (uint32_t)(r - roff);
    cols[pos] = (uint32_t)((int64_t)c - roff);
    vals[pos] = (va + vt) - va * vt;  // f32, the reference's operand order
    return;
  }
  const int64_t pb = (int64_t)(b - nrt - nslot) * 256;
  if (pb >= ptot) return;  // uniform
  const int64_t pe = pb + 256 < ptot ? pb + 256 : ptot;
  if (tid == 0) {
    const uint32_t rf = in_dst[pb], rl = in_dst[pe - 1];
    const bool ok = rf < (uint64_t)R && rl < (uint64_t)R;
    if (!ok) *err = 1u;
    span_s[0] = ok ? off[rf] : 0;
    span_s[1] = ok ? span_end(off[rl + 1], span_s[0], ptot, err) : 0;
  }
  const int64_t p = pb + tid;
  bool act = false;
  uint32_t src = 0xffffffffu, la = 0;
  int64_t r = 0, roff = 0, lo = 0, hi = 0;
  uint64_t base = 0;
  float w = 0.0f;
  if (p < pe) {
    src = in_key[p];
    r = in_dst[p];
    if (src != 0xffffffffu && (src >= (uint64_t)R || r >= R)) {
      *err = 1u;
      src = 0xffffffffu;
    }
    if (src != 0xffffffffu) {  // non-mutual
      act = true;
      w = in_w[p];
      const DsDesc& D = ds[ds_of_row(ds, nds, r)];
      roff = D.row_off;
      lo = off[r];
      hi = off[r + 1];
      base = indptr_g[r];
      la = less_in_row(key1, D.edge_off + (r - roff) * D.k, D.k, src);
    }
  }
  __syncthreads();
  const int64_t L = span_s[0], H = span_s[1];
  if (!act) lo = hi = L;
  const uint32_t li = span_count(in_key, L, H, lo, hi, src, p, sk);
  if (!act) return;
  const uint64_t pos = base + la + li;
  if (pos >= cap) return;
  const float v = 0.0f;
  rows[pos] = (uint32_t)(r - roff);
  cols[pos] = (uint32_t)((int64_t)src - roff);
  vals[pos] = (v + w) - v * w;
}

#define HM_SYM_INST(G)                                                                          \
  template __global__ void sym_scatter_kernel<G>(const DsDesc*, int, int64_t, int64_t,          \
                                                 const uint32_t*, const uint32_t*, const float*, \
                                                 const uint32_t*, uint32_t*, uint32_t*, float*,  \
                                                 float*, uint64_t*, uint32_t*);
HM_SYM_INST(16)
HM_SYM_INST(32)
HM_SYM_INST(64)
#undef HM_SYM_INST


}  // namespace hm
