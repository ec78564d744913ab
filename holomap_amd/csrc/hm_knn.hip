// Exact kNN in the reference's arithmetic.
//
//  * exact_scan_kernel  -- brute force over a column split for a block of up
//    to 32 query rows: every pair distance is the reference's sequential f64
//    fold (src/metric.rs:42-53 / 58-70), self forced to 0 (src/knn.rs:38-40),
//    and each row keeps its k smallest (f32 dist, idx) keys -- the order of
//    src/knn.rs:48, since dist >= +0 makes the f32 bit pattern monotone.
//    Serves the exact precision mode and the fallback for uncertified rows.
//  * exact_merge_kernel -- merges the per-split top-k lists of a row (two
//    levels when there are many splits).
//  * rerank_kernel      -- after the MFMA candidate sweep: recomputes the
//    plausible candidates exactly, takes the top-k and certifies that no
//    column outside the candidate set can enter it.
#include "hm_scan.hpp"

namespace hm {

__device__ __forceinline__ double f64max0(double v) {
  // f64::max(v, 0.0): a NaN operand yields the other one
  return (v != v) ? 0.0 : (v > 0.0 ? v : 0.0);
}

// Reference distance of query row a (uniform) to row b, over the zero-padded
// copies (padding adds exact zeros: (0-0)^2 and 0*0 change no partial sum),
// in f64 before the reference's f32 rounding.
__device__ __forceinline__ double ref_dist64(const float* __restrict__ a,
                                             const float* __restrict__ b, int d4, int metric,
                                             double rna, double rnb) {
  double s = 0.0;
  if (metric == 0) {
    for (int f = 0; f < d4; f += 4) {
      const float4 x = *(const float4*)(a + f);
      const float4 y = *(const float4*)(b + f);
      double t;
      t = (double)x.x - (double)y.x; s = s + t * t;
      t = (double)x.y - (double)y.y; s = s + t * t;
      t = (double)x.z - (double)y.z; s = s + t * t;
      t = (double)x.w - (double)y.w; s = s + t * t;
    }
    return sqrt(s);
  } else {
    for (int f = 0; f < d4; f += 4) {
      const float4 x = *(const float4*)(a + f);
      const float4 y = *(const float4*)(b + f);
      s = s + (double)x.x * (double)y.x;
      s = s + (double)x.y * (double)y.y;
      s = s + (double)x.z * (double)y.z;
      s = s + (double)x.w * (double)y.w;
    }
    return f64max0(1.0 - s / (rna * rnb));
  }
}
__device__ __forceinline__ float ref_dist(const float* __restrict__ a, const float* __restrict__ b,
                                          int d4, int metric, double rna, double rnb) {
  return (float)ref_dist64(a, b, d4, metric, rna, rnb);
}

// The f64 distance of output pair (row i, column j) of dataset D, for the
// FP64 calibration: the fold whose f32 rounding is the output distance (0
// for the row itself, as the exact kNN sets it).
__device__ __forceinline__ double pair_dist64(const DsDesc& D, int64_t i, uint32_t j) {
  if ((int64_t)j == i) return 0.0;
  const int d4 = (D.d + 3) & ~3;
  return ref_dist64(D.xr + (size_t)i * D.kp, D.xr + (size_t)j * D.kp, d4, D.metric,
                    D.metric == 1 ? D.rnorm[i] : 0.0, D.metric == 1 ? D.rnorm[j] : 0.0);
}

__device__ __forceinline__ uint64_t dist_key(float dist, uint32_t j) {
  return ((uint64_t)__float_as_uint(dist) << 32) | j;
}

// Ascending bitonic sort of 128 keys, two per lane (indices lane, lane+64).
__device__ __forceinline__ void sort128(uint64_t& v0, uint64_t& v1) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 2; k <= 128; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j == 64) {
        uint64_t mn = v0 < v1 ? v0 : v1, mx = v0 < v1 ? v1 : v0;
        v0 = mn;
        v1 = mx;
      } else {
        uint64_t o0 = shfl_xor_u64(v0, j), o1 = shfl_xor_u64(v1, j);
        bool up0 = (lane & k) == 0, up1 = ((lane + 64) & k) == 0;
        bool lw = (lane & j) == 0;
        uint64_t mn0 = v0 < o0 ? v0 : o0, mx0 = v0 < o0 ? o0 : v0;
        uint64_t mn1 = v1 < o1 ? v1 : o1, mx1 = v1 < o1 ? o1 : v1;
        v0 = (lw == up0) ? mn0 : mx0;
        v1 = (lw == up1) ? mn1 : mx1;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// exact_scan_kernel<QPW, QCAP>
// Work item: (ds, split, qbeg, qcnt <= 4*QPW). Queries are qlist[qbeg + r]
// (global row ids) or, without a list, local rows qbeg + r. Output: for
// query slot p (= list position, or local row in identity mode) the split's
// k smallest keys, sorted, at xpart[(p*xs + split)*kmax ...].
// Each of the four waves owns QPW queries and keeps, per query, a buffer of
// QCAP keys below the running k-th smallest (tau): a step appends at most
// 64 keys, and a full buffer is sorted down to its k smallest, so
// QCAP >= k + 64. Three shapes cover k <= 64 (registers sort), k <= 448 and
// k <= 1024 (LDS sorts).
// ---------------------------------------------------------------------------
#define XS_CB 64   // candidate columns per step
#define XS_KF 64   // features per LDS chunk

template <int QPW, int QCAP>
__global__ __launch_bounds__(256) void exact_scan_kernel(
    const DsDesc* __restrict__ ds_arr, const WorkItem* __restrict__ items,
    const int* __restrict__ nitems_dev, int nitems_host, const uint32_t* __restrict__ qlist,
    const int* __restrict__ xs_dev, int xs_host, int kmax, uint64_t* __restrict__ xpart) {
  constexpr int QB = 4 * QPW;
  static_assert(QCAP >= 128 && (QCAP & (QCAP - 1)) == 0, "buffer");
  __shared__ double qf[QB][XS_KF];
  __shared__ float cf[XS_CB][XS_KF + 1];
  __shared__ uint64_t buf[QB][QCAP];
  __shared__ int64_t qrow_s[QB];
  __shared__ double qrn_s[QB];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nitems = nitems_dev ? *nitems_dev : nitems_host;
  const int xs = xs_dev ? *xs_dev : xs_host;
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const WorkItem W = items[it];
    const DsDesc& D = ds_arr[W.ds];
    const int64_t n = D.n;
    const int kp = D.kp, d = D.d, k = D.k, metric = D.metric;
    // waves whose query slots are all empty skip the arithmetic
    const bool wact = wave * QPW < W.qcnt;
    const int64_t span = ((n + (int64_t)xs * XS_CB - 1) / ((int64_t)xs * XS_CB)) * XS_CB;
    const int64_t cbeg = (int64_t)W.split * span;
    const int64_t cend = cbeg + span < n ? cbeg + span : n;
    __syncthreads();  // previous item fully done with the shared arrays
    if (tid < QB) {
      int64_t row = -1;
      if (tid < W.qcnt) {
        const int p = W.qbeg + tid;
        row = qlist ? (int64_t)qlist[p] - D.row_off : (int64_t)p;
      }
      qrow_s[tid] = row;
      qrn_s[tid] = (row >= 0 && metric == 1) ? D.rnorm[row] : 0.0;
    }
    __syncthreads();
    uint64_t tau[QPW];
    int cnt[QPW];
    int64_t myq[QPW];
    double myrn[QPW];
#pragma unroll
    for (int qq = 0; qq < QPW; ++qq) {
      tau[qq] = ~0ull;
      cnt[qq] = 0;
      myq[qq] = qrow_s[wave * QPW + qq];
      myrn[qq] = qrn_s[wave * QPW + qq];
    }
    for (int64_t cb = cbeg; cb < cend; cb += XS_CB) {
      double acc[QPW];
#pragma unroll
      for (int qq = 0; qq < QPW; ++qq) acc[qq] = 0.0;
      for (int f0 = 0; f0 < d; f0 += XS_KF) {
        __syncthreads();
        {  // candidates: 64 rows x 64 features, 16 per thread
          const int r = tid >> 2, c = (tid & 3) * 16;
          const int64_t j = cb + r;
#pragma unroll
          for (int e = 0; e < 16; e += 4) {
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (j < n && f0 + c + e < kp) v = *(const float4*)(D.xr + (size_t)j * kp + f0 + c + e);
            cf[r][c + e] = v.x;
            cf[r][c + e + 1] = v.y;
            cf[r][c + e + 2] = v.z;
            cf[r][c + e + 3] = v.w;
          }
        }
        // queries: QB rows x 64 features as f64, a float4 per thread and step
        for (int e = tid; e < QB * (XS_KF / 4); e += 256) {
          const int r = e / (XS_KF / 4), c = (e % (XS_KF / 4)) * 4;
          const int64_t row = qrow_s[r];
          float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
          if (row >= 0 && f0 + c < kp) v = *(const float4*)(D.xr + (size_t)row * kp + f0 + c);
          qf[r][c] = (double)v.x;
          qf[r][c + 1] = (double)v.y;
          qf[r][c + 2] = (double)v.z;
          qf[r][c + 3] = (double)v.w;
        }
        __syncthreads();
        const int nf = d - f0 < XS_KF ? d - f0 : XS_KF;
        if (!wact) {
        } else if (metric == 0) {
          for (int f = 0; f < nf; ++f) {
            const double b = (double)cf[lane][f];
#pragma unroll
            for (int qq = 0; qq < QPW; ++qq) {
              const double t = qf[wave * QPW + qq][f] - b;
              acc[qq] = acc[qq] + t * t;
            }
          }
        } else {
          for (int f = 0; f < nf; ++f) {
            const double b = (double)cf[lane][f];
#pragma unroll
            for (int qq = 0; qq < QPW; ++qq) acc[qq] = acc[qq] + qf[wave * QPW + qq][f] * b;
          }
        }
      }
      const int64_t j = cb + lane;
      const double rnc = (metric == 1 && j < n) ? D.rnorm[j] : 0.0;
#pragma unroll
      for (int qq = 0; qq < QPW; ++qq) {
        const int slot = wave * QPW + qq;
        if (myq[qq] < 0) continue;  // uniform
        float dist;
        if (metric == 0) dist = (float)sqrt(acc[qq]);
        else dist = (float)f64max0(1.0 - acc[qq] / (myrn[qq] * rnc));
        if (j == myq[qq]) dist = 0.0f;
        const uint64_t key = dist_key(dist, (uint32_t)j);
        bool hit = (j < cend) && key < tau[qq];
        uint64_t mask = __ballot(hit);
        if (mask == 0ull) continue;
        if (cnt[qq] + __popcll(mask) > QCAP) {
          // full: keep the k smallest, tau = the k-th
          if constexpr (QCAP == 128) {
            uint64_t v0 = lane < cnt[qq] ? buf[slot][lane] : ~0ull;
            uint64_t v1 = lane + 64 < cnt[qq] ? buf[slot][lane + 64] : ~0ull;
            sort128(v0, v1);
            if (lane < k) buf[slot][lane] = v0;
            if (lane + 64 < k) buf[slot][lane + 64] = v1;
            tau[qq] = (k - 1) < 64 ? readlane_u64(v0, k - 1) : readlane_u64(v1, k - 65);
          } else {
            for (int e = cnt[qq] + lane; e < QCAP; e += 64) buf[slot][e] = ~0ull;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            wave_sort_lds(buf[slot], QCAP);
            tau[qq] = buf[slot][k - 1];
          }
          cnt[qq] = k;
          hit = (j < cend) && key < tau[qq];
          mask = __ballot(hit);
        }
        const int pos = cnt[qq] + (int)__popcll(mask & ((1ull << lane) - 1ull));
        if (hit) buf[slot][pos] = key;
        cnt[qq] += (int)__popcll(mask);
      }
    }
    // final: sort each row's buffer, emit its k smallest
#pragma unroll
    for (int qq = 0; qq < QPW; ++qq) {
      const int slot = wave * QPW + qq;
      if (myq[qq] < 0) continue;
      const int p = W.qbeg + slot;
      uint64_t* dst = xpart + ((size_t)p * xs + W.split) * kmax;
      if constexpr (QCAP == 128) {
        uint64_t v0 = lane < cnt[qq] ? buf[slot][lane] : ~0ull;
        uint64_t v1 = lane + 64 < cnt[qq] ? buf[slot][lane + 64] : ~0ull;
        sort128(v0, v1);
        if (lane < k) dst[lane] = v0;
        if (lane + 64 < k) dst[lane + 64] = v1;
      } else {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        for (int e = cnt[qq] + lane; e < QCAP; e += 64) buf[slot][e] = ~0ull;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        wave_sort_lds(buf[slot], QCAP);
        for (int t = lane; t < k; t += 64) dst[t] = buf[slot][t];
      }
    }
  }
}

#define HM_XS_INST(QPW, QCAP)                                                                    \
  template __global__ void exact_scan_kernel<QPW, QCAP>(const DsDesc*, const WorkItem*,          \
                                                        const int*, int, const uint32_t*,       \
                                                        const int*, int, int, uint64_t*);
HM_XS_INST(8, 128)
HM_XS_INST(4, 512)
HM_XS_INST(1, 2048)
#undef HM_XS_INST

// Merging the xs sorted per-split lists of every listed query (or local
// row in identity mode) into its knn row, one wave per row, in two levels so
// a wave never sorts more than 1024 keys: with xs * k > 1024 each group of G
// consecutive splits (G * k <= 1024) is first reduced to its k smallest
// keys, gathered in a second LDS area (at most G groups: G * k <= 1024
// keys), which is then sorted; else all the lists are sorted at once.
// The last kernel of the kNN stage: workgroup 0 also clears clr[0 .. nclr)
// (the statistics maxima and the flagged-row count, which the next run's
// prep and re-rank accumulate into; their readers have run).
__global__ __launch_bounds__(256) void exact_merge_kernel(
    const DsDesc* __restrict__ ds_arr, int nds, const uint32_t* __restrict__ qlist,
    const int* __restrict__ count_dev, int64_t count_host, const int* __restrict__ xs_dev,
    int xs_host, int kmax, int G, const uint64_t* __restrict__ xpart, uint32_t* __restrict__ clr,
    int nclr) {
  __shared__ uint64_t s[4][2][1024];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (blockIdx.x == 0 && clr)
    for (int t = threadIdx.x; t < nclr; t += 256) clr[t] = 0u;
  const int64_t count = count_dev ? (int64_t)*count_dev : count_host;
  const int xs = xs_dev ? *xs_dev : xs_host;
  const bool two = xs * kmax > 1024;
  const int ngrp = two ? (xs + G - 1) / G : 1;
  for (int64_t p = (int64_t)blockIdx.x * 4 + wave; p < count; p += (int64_t)gridDim.x * 4) {
    const int64_t grow = qlist ? (int64_t)qlist[p] : p;
    int lo = 0, hi = nds - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (ds_arr[mid].row_off <= grow) lo = mid; else hi = mid - 1;
    }
    const DsDesc& D = ds_arr[lo];
    const int64_t row = grow - D.row_off;
    const int k = D.k;
    const uint64_t* src = xpart + (size_t)p * xs * kmax;
    uint64_t* a = s[wave][0];
    uint64_t* m = s[wave][1];
    // the lists of splits [s0, s0 + ns) into a[0 .. len), padded
    auto gather = [&](int s0, int ns) {
      int len = 64;
      while (len < ns * k) len <<= 1;
      for (int e = lane; e < len; e += 64) {
        const int sp = e / k, t = e % k;
        a[e] = (sp < ns) ? src[(size_t)(s0 + sp) * kmax + t] : ~0ull;
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      return len;
    };
    uint64_t* fin = a;
    int nm = xs * k;
    if (two) {
      for (int g = 0; g < ngrp; ++g) {
        const int s0 = g * G;
        wave_sort_lds(a, gather(s0, xs - s0 < G ? xs - s0 : G));
        for (int t = lane; t < k; t += 64) m[g * k + t] = a[t];
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      }
      fin = m;
      nm = ngrp * k;
    } else {
      gather(0, xs);
    }
    int len = 64;
    while (len < nm) len <<= 1;
    if (two) {
      for (int e = nm + lane; e < len; e += 64) m[e] = ~0ull;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    wave_sort_lds(fin, len);
    for (int t = lane; t < k; t += 64) {
      const uint64_t key = fin[t];
      D.knn_idx[row * k + t] = (uint32_t)key;
      D.knn_dist[row * k + t] = __uint_as_float((uint32_t)(key >> 32));
      if (D.knn_dist64) D.knn_dist64[row * k + t] = pair_dist64(D, row, (uint32_t)key);
    }
  }
}

// ---------------------------------------------------------------------------
// rerank_kernel: one wave per global row (all rows of all datasets).
//   mode: 0 bf16, 1 fp32, 2 bf16x3 -- selects the error bound of the keys.
// Writes knn rows for every row, and flag[grow] = 1 where the candidate set
// could not be certified to contain the reference's top-k (those rows are
// recomputed by the exact scan).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float next_up(float x) {
  // next float toward +inf for finite x >= 0
  return __uint_as_float(__float_as_uint(x) + 1u);
}
__device__ __forceinline__ float f32_round_up(double v) {
  // smallest float >= v, for v >= 0
  float f = (float)v;
  if ((double)f < v) f = next_up(f);
  return f;
}

#define RR_LISTS 10   // lists per row at most (two per column split)
#define RR_QROW 512   // query rows of at most this many floats are staged in LDS
// LDS words per wave: candidate columns and distances (512 each, 32-bit),
// the query row, 64 sums (f64); six four-wave workgroups fill a CU's LDS
#define RR_LDS ((512 + 512 + RR_QROW) / 2 + 64)

// ref_dist64 for the re-rank: the query row qa (read at the same address by
// every lane: LDS when staged there, else global memory through the scalar
// path) against the lane's candidate row b, fetched 8 floats at a time with
// the next 8 in flight while these are folded (two buffers of 8: the
// register cost of one 16-float batch, which six waves per SIMD allow; 32
// floats spill). Same operations in the same order as ref_dist64.
template <class X>  // float4, or double4 holding f32 values
__device__ __forceinline__ void rr_fold4(double& s, const X x, const float4 y, int metric) {
  if (metric == 0) {
    double t;
    t = (double)x.x - (double)y.x; s = s + t * t;
    t = (double)x.y - (double)y.y; s = s + t * t;
    t = (double)x.z - (double)y.z; s = s + t * t;
    t = (double)x.w - (double)y.w; s = s + t * t;
  } else {
    s = s + (double)x.x * (double)y.x;
    s = s + (double)x.y * (double)y.y;
    s = s + (double)x.z * (double)y.z;
    s = s + (double)x.w * (double)y.w;
  }
}
// The candidate row is read through a global-address-space pointer
// (global_load rather than flat_load: the row pointer comes from the
// descriptor, whose address space the compiler cannot infer); the query
// pointer keeps the address space it is called with (LDS or global).
typedef __attribute__((address_space(1))) const float gfloat;
typedef float f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld4(gfloat* p) {
  const f4v v = *(__attribute__((address_space(1))) const f4v*)p;
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float4 ld4(const float* p) { return *(const float4*)p; }
// Query sources of the fold: f32 values read at the same address by every
// lane, in LDS (staged) or in global memory.
struct QueryL {
  const float* q;
  __device__ __forceinline__ float4 x4(int f) const { return *(const float4*)(q + f); }
};
struct QueryG {
  gfloat* q;
  __device__ __forceinline__ float4 x4(int f) const { return ld4(q + f); }
};
// the query's f32 values staged in LDS as f64 (exact), saving a conversion
// per element and candidate
struct QueryL64 {
  const double* q;
  __device__ __forceinline__ double4 x4(int f) const {
    const double2 a = *(const double2*)(q + f), b = *(const double2*)(q + f + 2);
    return make_double4(a.x, a.y, b.x, b.y);
  }
};
// Accumulators of the fold: the reference's sequential one, or four partial
// sums of fused multiply-adds over the same f64 differences (the fast path:
// a quarter of the dependent chain and one f64 operation fewer per element;
// its distance from the reference's sum is bounded in rr_end).
struct AccExact {
  double s = 0.0;
  template <class X>
  __device__ __forceinline__ void add4(const X x, const float4 y, int metric) {
    rr_fold4(s, x, y, metric);
  }
  __device__ __forceinline__ double sum() const { return s; }
};
struct AccFast {
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  template <class X>
  __device__ __forceinline__ void add4(const X x, const float4 y, int metric) {
    if (metric == 0) {
      const double t0 = (double)x.x - (double)y.x, t1 = (double)x.y - (double)y.y;
      const double t2 = (double)x.z - (double)y.z, t3 = (double)x.w - (double)y.w;
      a0 = __builtin_fma(t0, t0, a0);
      a1 = __builtin_fma(t1, t1, a1);
      a2 = __builtin_fma(t2, t2, a2);
      a3 = __builtin_fma(t3, t3, a3);
    } else {
      a0 = __builtin_fma((double)x.x, (double)y.x, a0);
      a1 = __builtin_fma((double)x.y, (double)y.y, a1);
      a2 = __builtin_fma((double)x.z, (double)y.z, a2);
      a3 = __builtin_fma((double)x.w, (double)y.w, a3);
    }
  }
  __device__ __forceinline__ double sum() const { return (a0 + a1) + (a2 + a3); }
};
template <class A, class Q>
__device__ __forceinline__ double rr_sum(const Q qa, const float* __restrict__ bp, int d4,
                                         int metric) {
  gfloat* b = (gfloat*)bp;
  A acc;
  const int d8 = d4 & ~7;
  float4 y0[2], y1[2];
  if (d8 > 0) {
    y0[0] = ld4(b);
    y0[1] = ld4(b + 4);
  }
  int f = 0;
  while (f < d8) {
    if (f + 8 < d8) {
      y1[0] = ld4(b + f + 8);
      y1[1] = ld4(b + f + 12);
    }
    acc.add4(qa.x4(f), y0[0], metric);
    acc.add4(qa.x4(f + 4), y0[1], metric);
    f += 8;
    if (f >= d8) break;
    if (f + 8 < d8) {
      y0[0] = ld4(b + f + 8);
      y0[1] = ld4(b + f + 12);
    }
    acc.add4(qa.x4(f), y1[0], metric);
    acc.add4(qa.x4(f + 4), y1[1], metric);
    f += 8;
  }
  if (f < d4) acc.add4(qa.x4(f), ld4(b + f), metric);  // d4 % 8 == 4
  return acc.sum();
}
template <class Q>
__device__ __forceinline__ double rr_dist64(const Q qa, const float* __restrict__ bp, int d4,
                                            int metric, double rna, double rnb) {
  const double s = rr_sum<AccExact>(qa, bp, d4, metric);
  if (metric == 0) return sqrt(s);
  return f64max0(1.0 - s / (rna * rnb));
}
// The fast sum of one candidate row, spread over the G = 2^lg lanes of a
// group: lane g folds the row's 16-byte chunks g, g + G, ... (of nch) into
// four fma accumulators, and the group adds its lanes' sums by a butterfly
// (each pair adds the same two values, so every lane ends with the same
// sum). The query chunks come from Q.
template <class Q>
__device__ __forceinline__ double rr_group_sum(const Q qa, gfloat* b, int g, int lg, int nch,
                                               int metric) {
  AccFast acc;
  const int G = 1 << lg;
  int f = g;
  for (; f + G < nch; f += 2 * G) {
    const float4 y0 = ld4(b + 4 * f), y1 = ld4(b + 4 * (f + G));
    acc.add4(qa.x4(4 * f), y0, metric);
    acc.add4(qa.x4(4 * (f + G)), y1, metric);
  }
  for (; f < nch; f += G) acc.add4(qa.x4(4 * f), ld4(b + 4 * f), metric);
  return acc.sum();
}
template <int O>
__device__ __forceinline__ double rr_xor_add(double s) {
  return s + __longlong_as_double((long long)shfl_xor_u64((uint64_t)__double_as_longlong(s), O));
}
__device__ __forceinline__ double rr_group_reduce(double s, int lg) {  // lg uniform
  if (lg >= 1) s = rr_xor_add<1>(s);
  if (lg >= 2) s = rr_xor_add<2>(s);
  if (lg >= 3) s = rr_xor_add<4>(s);
  if (lg >= 4) s = rr_xor_add<8>(s);
  return s;
}
// One end of the interval of f32 distances a fast sum s can stand for:
// hi = false the low end, true the high end. Every term of a sum reaches it
// through at most m roundings (its product or fma and the additions after
// it): m <= d4 for the reference's sequential fold, m <= d4/4 + 6 for the
// group sum (a chain of at most nch fmas, two additions of the
// accumulators, at most four butterfly levels). The two sums are thus at
// most (1.25 d4 + 6) u apart, u = 2^-53, relative to sum |p_f| <= s / (1 -
// m u) (Euclidean, p_f = t_f^2 >= 0 over the same differences t_f) or
// sum |x_f y_f| <= ||x|| ||y|| <= rna rnb (1 + (d4/2 + 2) u) (cosine).
// del = (3 d4 + 16) u also covers the roundings of the ends themselves. The
// f32 distance is a non-decreasing function of the sum (Euclidean: the f64
// sqrt, then f32) or a non-increasing one (cosine: 1 - s / r, max with 0,
// then f32), all steps correctly rounded, so where both ends give the same
// f32 value the reference's sum gives it too.
__device__ __forceinline__ float rr_end(double s, bool hi, int d4, int metric, double r) {
  const double del = (3.0 * d4 + 16.0) * 0x1p-53;
  if (metric == 0) return (float)sqrt(s * (hi ? 1.0 + del : 1.0 - del));
  const double err = del * r;
  return (float)f64max0(1.0 - (hi ? s - err : s + err) / r);
}

// One row (global id grow) by one wave; lds: the wave's 6 KB LDS area.
//  1. the stored keys of all lists, up to 8 per lane, loaded at once;
//  2. an upper bound of the k-th smallest key by bisection on the keys'
//     order-preserving u32 images (ballot counts), instead of sorting;
//  3. the candidates whose lower bound can beat the k-th upper bound are
//     compacted (ballot prefix) into a column list;
//  4. their exact distances, one lane per candidate (LDS: the columns in
//     the first 2 KB, the distance bits in the second);
//  5. the k smallest exact keys (distance bits, column) by rank: each key
//     counts the smaller ones; keys are distinct (distinct columns).
// Sorting all keys and scoring the plausible prefix would score a subset of
// these candidates; the extra ones have lower bounds above that prefix's
// threshold, so they cannot enter the top k: the output is the same.
// An uncertified row: its flag, and (nflag set) a count of such rows that
// lets the list builders skip their scan when it is zero (an atomic per
// flagged row, which is rare; the count itself is all that is read).
__device__ __forceinline__ void flag_row(__attribute__((address_space(1))) uint32_t* gflag,
                                         int64_t grow, uint32_t* __restrict__ nflag) {
  gflag[grow] = 1u;
  if (nflag) atomicAdd(nflag, 1u);
}

__device__ __forceinline__ void rerank_row(const DsDesc* __restrict__ ds_arr, int nds,
                                           int64_t grow, int mode, uint32_t* __restrict__ flag,
                                           uint32_t* __restrict__ nflag, uint64_t* lds) {
  const int lane = threadIdx.x & 63;
  int lo = 0, hi = nds - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (ds_arr[mid].row_off <= grow) lo = mid; else hi = mid - 1;
  }
  const DsDesc& D = ds_arr[lo];
  const int64_t i = grow - D.row_off;
  const int k = D.k, L = D.nlist, CO = D.cout, d = D.d, kp = D.kp, metric = D.metric;
  const int d4 = (d + 3) & ~3;
  // outputs and row data through global-address-space pointers (global
  // rather than flat accesses, which would also count in lgkmcnt)
  __attribute__((address_space(1))) uint32_t* oidx =
      (__attribute__((address_space(1))) uint32_t*)(D.knn_idx + i * k);
  __attribute__((address_space(1))) float* odst =
      (__attribute__((address_space(1))) float*)(D.knn_dist + i * k);
  __attribute__((address_space(1))) double* od64 =
      (__attribute__((address_space(1))) double*)(D.knn_dist64 ? D.knn_dist64 + i * k : nullptr);
  __attribute__((address_space(1))) uint32_t* gflag = (__attribute__((address_space(1))) uint32_t*)flag;
  typedef __attribute__((address_space(1))) const double gf64;
  const double rna = metric == 1 ? ((gf64*)D.rnorm)[i] : 0.0;

  if (metric == 1 && rna == 0.0) {
    // zero query: every cosine distance is max(NaN, 0) = 0, so the list is
    // the k lowest indices (self included) at distance 0
    for (int t = lane; t < k; t += 64) {
      oidx[t] = (uint32_t)t;
      odst[t] = 0.0f;
      if (od64) od64[t] = 0.0;
    }
    if (lane == 0) gflag[grow] = 0;
    return;
  }

  // loads that do not depend on the keys go out first, so their round trips
  // overlap the list counts' and the keys': the row's bound statistics and
  // the dataset's maxima, and the query row staged in LDS
  float st[HM_NSTAT], mx[HM_NSTAT];
  {
    gfloat* sp = (gfloat*)(D.rstat + i * HM_NSTAT);
    gfloat* mp = (gfloat*)D.smax;
#pragma unroll
    for (int h = 0; h < HM_NSTAT / 4; ++h) {
      const float4 a = ld4(sp + 4 * h), b = ld4(mp + 4 * h);
      st[4 * h] = a.x; st[4 * h + 1] = a.y; st[4 * h + 2] = a.z; st[4 * h + 3] = a.w;
      mx[4 * h] = b.x; mx[4 * h + 1] = b.y; mx[4 * h + 2] = b.z; mx[4 * h + 3] = b.w;
    }
  }
  uint32_t* cl = (uint32_t*)lds;  // <= 512 plausible candidate columns
  uint32_t* xd = cl + 512;        // their exact distances (f32 bits)
  float* qs = (float*)(xd + 512);  // the query row (d4 <= RR_QROW)
  double* ss = (double*)(qs + RR_QROW);  // fast sums of 64 candidates
  const float* qa = D.xr + (size_t)i * kp;
  const bool qlds = d4 <= RR_QROW;
  // without the FP64 outputs a query of at most RR_QROW / 2 floats is
  // staged as f64 (the fast sums and their fallback fold read it)
  double* qd = (double*)qs;
  const bool q64 = !od64 && d4 <= RR_QROW / 2;
  if (q64) {
    for (int f = 4 * lane; f < d4; f += 256) {
      const float4 v = ld4((gfloat*)qa + f);
      *(double2*)(qd + f) = make_double2(v.x, v.y);
      *(double2*)(qd + f + 2) = make_double2(v.z, v.w);
    }
  } else if (qlds) {
    for (int f = 4 * lane; f < d4; f += 256) *(float4*)(qs + f) = ld4((gfloat*)qa + f);
  }

  // ---- error bound of the candidate keys for this row ----
  const double u = 0x1p-24;
  double E;
  double dotmax;
  if (mode == 2) {
    const double gam = 2.0 * 3.0 * kp * u;
    const double dot_err = (double)st[ST_LN] * mx[ST_LN] + (double)st[ST_HLN] * mx[ST_RN] +
                           (double)st[ST_RN] * mx[ST_XO];
    const double acc_err = gam * ((double)st[ST_HN] * mx[ST_HN] + (double)st[ST_HN] * mx[ST_LN] +
                                  (double)st[ST_LN] * mx[ST_HN]);
    E = dot_err + acc_err;
    dotmax = (double)st[ST_XO] * mx[ST_XO] * 1.01;
  } else if (mode == 1) {
    // f32 operand w: ||w|| = XO (rounded up), its rounding error
    // ||w_exact - w|| <= u/(1-u) ||w|| componentwise (whatever the job's
    // primary mode stored in the XN/EN slots)
    const double gam = 2.0 * kp * u;
    const double xn = (double)st[ST_XO], mxn = (double)mx[ST_XO];
    const double en = xn * u * 1.0001, men = mxn * u * 1.0001;
    const double dot_err = en * (mxn + men) + xn * men;
    const double acc_err = gam * xn * mxn;
    E = dot_err + acc_err;
    dotmax = xn * mxn * (1.0 + gam);
  } else {
    const double gam = 2.0 * kp * u;
    const double dot_err = (double)st[ST_EN] * ((double)mx[ST_XN] + mx[ST_EN]) +
                           (double)st[ST_XN] * mx[ST_EN];
    const double acc_err = gam * (double)st[ST_XN] * mx[ST_XN];
    E = dot_err + acc_err;
    dotmax = (double)st[ST_XN] * mx[ST_XN] * (1.0 + gam);
  }
  const double eps_exact = (d + 8) * 0x1p-52;  // f64 fold of the exact distance
  double nzi = 0.0;
  if (metric == 0) {
    nzi = (double)st[ST_NZ];
    E = 2.0 * E + (double)st[ST_NZ] * 0x1p-23 + (double)mx[ST_NZ] * 0x1p-23 +
        u * ((double)mx[ST_NZ] + 2.0 * dotmax);
    // the sweep's accumulators start at -||x_j||^2 / 2: every accumulation
    // step (at most one per product, three products per element for bf16x3)
    // also rounds at that magnitude
    E += (mode == 2 ? 3.0 : 1.0) * 2.0 * (kp + 1) * u * (double)mx[ST_NZ];
    E = E * (1.0 + 1e-3) + 1e-30;
  } else {
    E = E * (1.0 + 1e-3) + 1e-12;
  }


  // ---- 1. the stored keys of all lists: slot e = lane + 64 m of the packed
  // order (list after list) holds list l = the last with off[l] <= e ----
  typedef __attribute__((address_space(1))) const uint64_t gu64;
  typedef __attribute__((address_space(1))) const uint32_t gu32;
  gu64* src = (gu64*)(D.cand + (size_t)i * L * CO);
  gu32* cnt = (gu32*)(D.ccnt + (size_t)i * L);
  int off[RR_LISTS + 1];  // uniform
  off[0] = 0;
#pragma unroll
  for (int l = 0; l < RR_LISTS; ++l) off[l + 1] = off[l] + (l < L ? (int)cnt[l] : 0);
  const int nvalid = off[RR_LISTS];  // <= L * CO <= 512
  uint64_t taumin = ~0ull;
  for (int s = 0; s < L; ++s) {
    const uint64_t t = ((gu64*)D.ctau)[(size_t)i * L + s];
    taumin = t < taumin ? t : taumin;
  }
  if (nvalid < k) {
    // cannot happen for n >= k (the lists keep min(C, columns) keys);
    // leave the row to the exact scan rather than decode an empty slot
    if (lane == 0) flag_row(gflag, grow, nflag);
    return;
  }
  // the sweep stores (key' bits, column) with key' = key / s; kh holds the
  // order-preserving image of key = s * key', kl the column
  const float sc = metric == 1 ? -1.0f : -2.0f;
  const int mc = (nvalid + 63) >> 6;  // slot rounds in use (uniform)
  uint32_t kh[8], kl[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    kh[m] = ~0u;
    kl[m] = 0;
    const int e = lane + 64 * m;
    if (m < mc && e < nvalid) {
      int l = 0, base = 0;
#pragma unroll
      for (int q = 1; q < RR_LISTS; ++q)
        if (e >= off[q]) { l = q; base = off[q]; }
      const uint64_t v = src[(size_t)l * CO + (e - base)];
      kh[m] = ordered_u32(sc * __uint_as_float((uint32_t)(v >> 32)));
      kl[m] = (uint32_t)v;
    }
  }

  // ---- 2. an upper bound of the k-th smallest key: bisection keeps
  // #{kh <= bhi} >= k; any such bhi gives a valid (at most slightly larger)
  // pruning threshold, so it stops once the bracket is 2^10 ulps wide ----
  uint32_t blo = 0xffffffffu, bhi = 0;
#pragma unroll
  for (int m = 0; m < 8; ++m)
    if (m < mc && lane + 64 * m < nvalid) {
      blo = kh[m] < blo ? kh[m] : blo;
      bhi = kh[m] > bhi ? kh[m] : bhi;
    }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t a = (uint32_t)__shfl_xor((int)blo, o, 64), b = (uint32_t)__shfl_xor((int)bhi, o, 64);
    blo = a < blo ? a : blo;
    bhi = b > bhi ? b : bhi;
  }
  while (bhi - blo > 1024u) {
    const uint32_t mid = blo + ((bhi - blo) >> 1);
    int c = 0;
#pragma unroll
    for (int m = 0; m < 8; ++m)
      if (m < mc) c += __popcll(__ballot(lane + 64 * m < nvalid && kh[m] <= mid));
    if (c >= k) bhi = mid; else blo = mid + 1;
  }

  // ---- prune: candidates that cannot beat the k key-best ones ----
  const double kk = (double)unordered_f32(bhi);  // >= the k-th smallest key
  double thr;  // a candidate with lower bound >= thr is strictly beyond the top-k
  if (metric == 0) {
    double ub = nzi + kk + E;
    ub = ub > 0.0 ? ub : 0.0;
    const float dup = next_up(f32_round_up(sqrt(ub * (1.0 + eps_exact))));
    thr = (double)dup * (double)dup * (1.0 + eps_exact) + 1e-300;
  } else {
    double ub = 1.0 + kk + E + eps_exact;
    ub = ub > 0.0 ? ub : 0.0;
    thr = (double)next_up(f32_round_up(ub));
  }
  // ---- 3. the plausible candidates, compacted (column list in lds) ----
  int np = 0;
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    if (m < mc) {
      bool pass = false;
      if (lane + 64 * m < nvalid) {
        const double kv = (double)unordered_f32(kh[m]);
        const double lb = metric == 0 ? (nzi + kv - E) * (1.0 - eps_exact) : 1.0 + kv - E - eps_exact;
        pass = lb < thr;
      }
      const uint64_t bal = __ballot(pass);
      const int pos = np + (int)__builtin_amdgcn_mbcnt_hi(
                               (uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
      if (pass) cl[pos] = kl[m];
      np += __popcll(bal);
    }
  }
  if (np < k) {
    // cannot happen (the k key-best candidates pass)
    if (lane == 0) flag_row(gflag, grow, nflag);
    return;
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // the query row is in LDS

  // ---- 4. exact distances of the candidates. With the FP64 outputs: the
  // reference's fold, one lane per candidate (the lane's first candidate's
  // f64 value kept for them). Without: G = 2^lg lanes per candidate, 64 / G
  // candidates per pass, so each load instruction covers whole cache lines
  // of a few rows (one lane per candidate touched 64 rows' lines per
  // instruction, more than the CU's L1 keeps for its waves); the groups'
  // fast sums of 64 candidates are staged in LDS, and one lane per candidate
  // decides its f32 distance from its sum (rr_end) or, in the rare cases
  // the sum leaves open, by the reference's fold. ----
  bool badid = false;
  double first64 = 0.0;
  if (od64) {
    for (int e = lane; e < np; e += 64) {
      const uint32_t j = cl[e];
      double dist = 0.0;
      if ((int64_t)j >= D.n) {
        badid = true;  // never produced by the sweep; guard the row read
      } else if ((int64_t)j != i) {
        const double rnb = metric == 1 ? ((gf64*)D.rnorm)[j] : 0.0;
        const float* bp = D.xr + (size_t)j * kp;
        dist = qlds ? rr_dist64(QueryL{qs}, bp, d4, metric, rna, rnb)
                    : rr_dist64(QueryG{(gfloat*)qa}, bp, d4, metric, rna, rnb);
      }
      if (e == lane) first64 = dist;
      xd[e] = __float_as_uint((float)dist);
    }
  }
  const int nch = d4 >> 2;  // 16-byte chunks per row
  const int lg = nch >= 64 ? 4 : nch >= 32 ? 3 : nch >= 16 ? 2 : nch >= 8 ? 1 : 0;  // uniform
  const int g = lane & ((1 << lg) - 1);
  for (int c00 = 0; !od64 && c00 < np; c00 += 64) {
    // the sums of candidates [c00, c00 + 64) into ss, by lane groups ...
    const int cend = np - c00 < 64 ? np : c00 + 64;
    for (int c0 = c00; c0 < cend; c0 += 64 >> lg) {
      const int c = c0 + (lane >> lg);
      const uint32_t j = c < cend ? cl[c] : 0u;
      const bool live = c < cend && (int64_t)j < D.n && (int64_t)j != i;  // same in a group
      double sum = 0.0;
      if (live) {
        gfloat* bp = (gfloat*)(D.xr + (size_t)j * kp);
        // inlined copies, so the query reads keep their address space
        sum = q64    ? rr_group_sum(QueryL64{qd}, bp, g, lg, nch, metric)
              : qlds ? rr_group_sum(QueryL{qs}, bp, g, lg, nch, metric)
                     : rr_group_sum(QueryG{(gfloat*)qa}, bp, g, lg, nch, metric);
      }
      sum = rr_group_reduce(sum, lg);
      if (c < cend && g == 0) ss[c - c00] = sum;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // ... then each decided in the lane of its own: both ends of the
    // interval, and the reference's fold where they differ (rare)
    const int c = c00 + lane;
    if (c < cend) {
      const uint32_t j = cl[c];
      if ((int64_t)j >= D.n) badid = true;  // never produced by the sweep
      float d32 = 0.0f;
      if ((int64_t)j < D.n && (int64_t)j != i) {
        const double rnb = metric == 1 ? ((gf64*)D.rnorm)[j] : 0.0, r = rna * rnb;
        const double sum = ss[lane];
        d32 = rr_end(sum, false, d4, metric, r);
        if (d32 != rr_end(sum, true, d4, metric, r)) {
          const float* bp = D.xr + (size_t)j * kp;
          d32 = (float)(q64    ? rr_dist64(QueryL64{qd}, bp, d4, metric, rna, rnb)
                        : qlds ? rr_dist64(QueryL{qs}, bp, d4, metric, rna, rnb)
                               : rr_dist64(QueryG{(gfloat*)qa}, bp, d4, metric, rna, rnb));
        }
      }
      xd[c] = __float_as_uint(d32);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // ss is reused
  }
  if (__ballot(badid) != 0ull) {
    if (lane == 0) flag_row(gflag, grow, nflag);  // recomputed by the exact scan
    return;
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");

  // ---- 5. the k smallest exact keys ----
  uint64_t kth = 0;  // the k-th smallest exact key (its distance bits decide below)
  if (np <= 64 || k > 64) {
    // by rank counts (the lists padded to a multiple of 4 with keys above
    // every real one, so the counts read 4 keys per 16-byte LDS read; np <=
    // 512 slots); kth found by the lane that holds it
    const int np4 = (np + 3) & ~3;
    if (lane < np4 - np) {
      xd[np + lane] = 0xffffffffu;
      cl[np + lane] = 0xffffffffu;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const uint4* xd4 = (const uint4*)xd;
    const uint4* cl4 = (const uint4*)cl;
    for (int m = 0; 64 * m < np; ++m) {
      const int e = lane + 64 * m;
      const uint64_t mine = e < np ? dist_key(__uint_as_float(xd[e]), cl[e]) : ~0ull;
      int rk = 0;
      for (int f = 0; f < np4; f += 4) {  // same address in every lane: broadcasts
        const uint4 xv = xd4[f >> 2], cv = cl4[f >> 2];
        rk += (((uint64_t)xv.x << 32) | cv.x) < mine ? 1 : 0;
        rk += (((uint64_t)xv.y << 32) | cv.y) < mine ? 1 : 0;
        rk += (((uint64_t)xv.z << 32) | cv.z) < mine ? 1 : 0;
        rk += (((uint64_t)xv.w << 32) | cv.w) < mine ? 1 : 0;
      }
      if (e < np && rk < k) {
        oidx[rk] = (uint32_t)mine;
        odst[rk] = __uint_as_float((uint32_t)(mine >> 32));
        if (od64) od64[rk] = m == 0 ? first64 : pair_dist64(D, i, (uint32_t)mine);
        if (rk == k - 1) kth = mine;
      }
    }
    // the one lane holding the k-th key: reduce by max (the others hold 0)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const uint64_t t = shfl_xor_u64(kth, o);
      kth = t > kth ? t : kth;
    }
  } else {
    // more than 64 candidates: the k-th smallest distance dk by bisection on
    // the distance bits (ballot counts; the bits of non-negative floats
    // order as the values, as the keys compare them), then, if more than
    // the k - #(dist < dk) still needed share dk, the largest column among
    // them to take, by bisection on the columns; the k keys at or below
    // (dk, that column) are compacted into ss and ranked among themselves
    // (k counts of k instead of np of np)
    const int mr = (np + 63) >> 6;  // rounds in use (uniform, <= 8)
    uint32_t xv[8];
    uint32_t dlo = 0xffffffffu, dhi = 0u;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int e = lane + 64 * m;
      xv[m] = (m < mr && e < np) ? xd[e] : 0xffffffffu;
      if (m < mr && e < np) {
        dlo = xv[m] < dlo ? xv[m] : dlo;
        dhi = xv[m] > dhi ? xv[m] : dhi;
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const uint32_t a = (uint32_t)__shfl_xor((int)dlo, o, 64), b = (uint32_t)__shfl_xor((int)dhi, o, 64);
      dlo = a < dlo ? a : dlo;
      dhi = b > dhi ? b : dhi;
    }
    auto count_le = [&](uint32_t v) {  // #(dist <= v); slots past np hold ~0 and are not counted
      int c = 0;
#pragma unroll
      for (int m = 0; m < 8; ++m)
        if (m < mr) c += __popcll(__ballot(lane + 64 * m < np && xv[m] <= v));
      return c;
    };
    while (dlo < dhi) {  // uniform; #(dist <= dhi) >= k throughout
      const uint32_t mid = dlo + ((dhi - dlo) >> 1);
      if (count_le(mid) >= k) dhi = mid; else dlo = mid + 1;
    }
    const uint32_t dk = dlo;
    const int nlt = dk == 0u ? 0 : count_le(dk - 1u);
    const int need = k - nlt, neq = count_le(dk) - nlt;  // 1 <= need <= neq
    uint32_t cmax = 0xffffffffu;
    if (neq > need) {
      uint32_t clo = 0u, chi = 0xffffffffu;
      while (clo < chi) {  // uniform: the least column bound taking `need` of the ties
        const uint32_t mid = clo + ((chi - clo) >> 1);
        int c = 0;
#pragma unroll
        for (int m = 0; m < 8; ++m)
          if (m < mr) {
            const int e = lane + 64 * m;
            c += __popcll(__ballot(e < np && xv[m] == dk && cl[e] <= mid));
          }
        if (c >= need) chi = mid; else clo = mid + 1;
      }
      cmax = clo;
    }
    kth = ((uint64_t)dk << 32) | cmax;
    uint64_t* wk = (uint64_t*)ss;  // the k <= 64 winners
    int nw = 0;
#pragma unroll
    for (int m = 0; m < 8; ++m)
      if (m < mr) {
        const int e = lane + 64 * m;
        const uint64_t key = e < np ? dist_key(__uint_as_float(xv[m]), cl[e]) : ~0ull;
        const bool win = e < np && key <= kth;
        const uint64_t bal = __ballot(win);
        const int pos = nw + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
        if (win && pos < 64) wk[pos] = key;  // (keys are distinct: nw == k <= 64)
        nw += __popcll(bal);
      }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (lane < nw && lane < k) {  // nw == k
      const uint64_t mine = wk[lane];
      int rk = 0;
      for (int f = 0; f < k; ++f) rk += wk[f] < mine ? 1 : 0;  // broadcasts
      if (rk >= k) rk = k - 1;  // (unreachable with distinct keys)
      oidx[rk] = (uint32_t)mine;
      odst[rk] = __uint_as_float((uint32_t)(mine >> 32));
      if (od64) od64[rk] = pair_dist64(D, i, (uint32_t)mine);
    }
  }

  // ---- certification against every column outside the candidate set ----
  bool cert = taumin != 0ull;  // 0: a candidate list overflowed (incomplete)
  if (cert && taumin != ~0ull) {
    const float dk = __uint_as_float((uint32_t)(kth >> 32));
    const double tf = (double)unordered_f32((uint32_t)(taumin >> 32));
    const double nu = (double)next_up(dk);
    if (metric == 0) {
      const double lb = (nzi + tf - E) * (1.0 - eps_exact);
      cert = lb >= nu * nu * (1.0 + eps_exact) + 1e-300;
    } else {
      cert = (1.0 + tf - E - eps_exact) >= nu;
    }
  }
  if (lane == 0) {
    if (cert) gflag[grow] = 0u;
    else flag_row(gflag, grow, nflag);
  }
}

// All rows (rlist == nullptr: rows 0 .. total_rows-1, one per wave), or the
// *rcount rows listed in rlist (global ids), grid-stride.
// 6 KB of LDS and at most 80 VGPRs per wave: six waves per SIMD keep
// enough rows in flight to cover their dependent memory round trips.
template <int NT, int PER>
__device__ void list_compact_block(const uint32_t* __restrict__ flag, uint32_t* __restrict__ list,
                                   int* __restrict__ count);
template <int NT>
__device__ void fallback_items_block(const DsDesc* __restrict__ ds_arr, int nds,
                                     const uint32_t* __restrict__ list,
                                     const int* __restrict__ count, int qb, int xs, int xs_cap,
                                     int64_t item_cap, int64_t row_cap, int target,
                                     WorkItem* __restrict__ items, int* __restrict__ nitems,
                                     int* __restrict__ xs_out);
template <bool TAIL>
__global__ __launch_bounds__(256, 6) void rerank_kernel(const DsDesc* __restrict__ ds_arr, int nds,
                                                        int64_t total_rows, int mode,
                                                        uint32_t* __restrict__ flag,
                                                        const uint32_t* __restrict__ rlist,
                                                        const int* __restrict__ rcount,
                                                        uint32_t* __restrict__ nflag,
                                                        const FbTail tail) {
  __shared__ __attribute__((aligned(16))) uint64_t cs[4][RR_LDS];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: row data in SGPRs
  const int64_t cnt = rlist ? (int64_t)*rcount : total_rows;
  for (int64_t p = (int64_t)blockIdx.x * 4 + wave; p < cnt; p += (int64_t)gridDim.x * 4) {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // previous row's LDS use is over
    rerank_row(ds_arr, nds, rlist ? (int64_t)rlist[p] : p, mode, flag, nflag, cs[wave]);
  }
  if constexpr (TAIL) {
    // The last workgroup with rows to finish builds the exact scan's work
    // list (it waits for nobody: it is last by its ticket; workgroups
    // without rows write and read nothing, so take none, and with no rows
    // at all workgroup 0 builds it). The barrier's release waits for the
    // workgroup's stores; the ticket's thread then releases them at agent
    // scope and the last workgroup acquires them all before reading the
    // flags. (A flag read stale would still hold the 1 the escalated rows
    // start with: the row would be recomputed exactly.)
    __shared__ int last;
    const int64_t nwork = cnt < 4 * (int64_t)gridDim.x ? (cnt + 3) / 4 : (int64_t)gridDim.x;
    if ((int64_t)blockIdx.x >= (nwork > 0 ? nwork : 1)) return;
    __syncthreads();
    if (threadIdx.x == 0) {
      if (nwork > 0) {
        __threadfence();
        last = atomicAdd(tail.done, 1u) == (uint32_t)nwork - 1u ? 1 : 0;
      } else {
        last = 1;
      }
    }
    __syncthreads();
    if (last) {  // uniform
      __threadfence();
      list_compact_block<256, 16>(flag, tail.list, tail.count);
      __syncthreads();
      fallback_items_block<256>(ds_arr, nds, tail.list, tail.count, tail.qb, tail.xs,
                                tail.xs_cap, tail.item_cap, tail.row_cap, tail.target, tail.items,
                                tail.nitems, tail.xs_out);
      if (threadIdx.x == 0 && nwork > 0) *tail.done = 0u;  // the next run's count starts at zero
    }
  }
}

// First position in the sorted list [0, cnt) holding a row id >= v.
__device__ __forceinline__ int list_lower_bound(const uint32_t* __restrict__ list, int cnt,
                                                int64_t v) {
  int lo = 0, hi = cnt;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if ((int64_t)list[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Block-wide helpers of the work-list builders (every thread of the
// workgroup calls them; NT is a multiple of 64). sh: NT / 64 words.
template <int NT>
__device__ __forceinline__ uint64_t block_excl_scan_u64(uint64_t v, uint64_t& total,
                                                        uint64_t* sh) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t inc = wave_scan_incl(v);
  if (lane == 63) sh[wave] = inc;
  __syncthreads();
  uint64_t base = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) {
    const uint64_t x = sh[w];
    base += w < wave ? x : 0ull;
    tot += x;
  }
  __syncthreads();  // sh is rewritten by the next call
  total = tot;
  return base + inc - v;
}
// Sum (MAX = false) or maximum of a double over the workgroup, in a fixed
// order (the same result on every run).
template <int NT, bool MAX>
__device__ __forceinline__ double block_reduce_f64(double v, double* sh) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double t = __shfl_xor(v, o, 64);
    v = MAX ? (t > v ? t : v) : v + t;
  }
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double r = sh[0];
#pragma unroll
  for (int w = 1; w < NT / 64; ++w) r = MAX ? (sh[w] > r ? sh[w] : r) : r + sh[w];
  __syncthreads();
  return r;
}

// Work items of a candidate sweep over the listed rows only (sorted global
// row ids): per dataset, blocks of qw listed rows x its column splits, in
// dataset order. One workgroup of NT threads, a thread per dataset (its
// range of the list by binary search); the item offsets are a scan over the
// datasets.
//
// The re-sweep pays a whole item (qw rows x a column split, both passes)
// even for one listed row, while the exact scan pays per row, so a handful of
// rows goes straight to the exact scan: the kernel compares the two
// estimated times (the re-sweep at `rate` Gram flop/s per CU -- measured
// ~1.1e12 for bf16x3, ~0.45e12 for FP32 -- in rounds of `cus` items; the
// exact f64 scan at ~2e13 flop/s over the device) and emits no items
// (*nrows = 0: the re-rank of the escalated rows is skipped too) when the
// scan is cheaper.
template <int NT>
__device__ void sweep_items_block(const DsDesc* __restrict__ ds_arr, int nds,
                                  const uint32_t* __restrict__ list, const int* __restrict__ count,
                                  int qw, int cus, double rate, WorkItem* __restrict__ items,
                                  int* __restrict__ nitems, int* __restrict__ nrows) {
  __shared__ uint64_t sh[NT / 64];
  __shared__ double shd[NT / 64];
  const int cnt = *count;
  double te = 0.0, ti = 0.0;
  uint64_t nq = 0;
  for (int di = threadIdx.x; di < nds; di += NT) {
    const DsDesc& D = ds_arr[di];
    const int p = list_lower_bound(list, cnt, D.row_off);
    const int q = list_lower_bound(list, cnt, D.row_off + D.n);
    if (q == p) continue;
    te += (double)(q - p) * (double)D.n * D.kp * 3.0 / 2e13;
    if (D.nsplit == 0) continue;  // k beyond the sweep: its rows go to the exact scan
    const double span = (double)D.n / D.nsplit;
    const double t = 4.0 * span * D.kp * qw / rate;
    ti = t > ti ? t : ti;
    nq += (uint64_t)((q - p + qw - 1) / qw) * D.nsplit;
  }
  const double t_exact = block_reduce_f64<NT, false>(te, shd);
  const double t_item = block_reduce_f64<NT, true>(ti, shd);
  uint64_t nit;
  (void)block_excl_scan_u64<NT>(nq, nit, sh);
  const double t_sweep = (double)((nit + cus - 1) / cus) * t_item;
  if (cnt == 0 || nit == 0 || t_exact <= t_sweep) {  // uniform
    if (threadIdx.x == 0) {
      *nitems = 0;
      *nrows = 0;
    }
    return;
  }
  uint64_t base = 0;
  for (int c0 = 0; c0 < nds; c0 += NT) {  // uniform
    const int di = c0 + (int)threadIdx.x;
    int p = 0, q = 0, ns = 0;
    if (di < nds) {
      p = list_lower_bound(list, cnt, ds_arr[di].row_off);
      q = list_lower_bound(list, cnt, ds_arr[di].row_off + ds_arr[di].n);
      ns = ds_arr[di].nsplit;
    }
    uint64_t tot;
    const uint64_t off = block_excl_scan_u64<NT>((uint64_t)((q - p + qw - 1) / qw) * ns, tot, sh);
    int w = (int)(base + off);
    for (int b = p; b < q; b += qw) {
      const int qc = q - b < qw ? q - b : qw;
      for (int s = 0; s < ns; ++s) items[w++] = WorkItem{di, s, b, qc};
    }
    base += tot;
  }
  if (threadIdx.x == 0) {
    *nrows = cnt;
    *nitems = (int)base;
  }
}
__global__ __launch_bounds__(256) void sweep_items_kernel(
    const DsDesc* __restrict__ ds_arr, int nds, const uint32_t* __restrict__ list,
    const int* __restrict__ count, int qw, int cus, double rate, WorkItem* __restrict__ items,
    int* __restrict__ nitems, int* __restrict__ nrows) {
  sweep_items_block<256>(ds_arr, nds, list, count, qw, cus, rate, items, nitems, nrows);
}

// The column split count starts at the planned xs and grows when few rows are
// flagged (a handful of rows would otherwise run on a handful of CUs), up to
// xs_cap (which keeps the merge at two levels) and within the item and
// partial-list capacities, which hold item_cap items and row_cap * xs lists.
// Items: per dataset, blocks of <= qb listed rows (the exact scan's queries
// per item) x the splits, in dataset order; one workgroup, a thread per
// dataset, as above.
template <int NT>
__device__ void fallback_items_block(const DsDesc* __restrict__ ds_arr, int nds,
                                     const uint32_t* __restrict__ list,
                                     const int* __restrict__ count, int qb, int xs, int xs_cap,
                                     int64_t item_cap, int64_t row_cap, int target,
                                     WorkItem* __restrict__ items, int* __restrict__ nitems,
                                     int* __restrict__ xs_out) {
  __shared__ uint64_t sh[NT / 64];
  const int cnt = *count;
  uint64_t part = 0;
  for (int di = threadIdx.x; di < nds; di += NT) {
    const int p = list_lower_bound(list, cnt, ds_arr[di].row_off);
    const int q = list_lower_bound(list, cnt, ds_arr[di].row_off + ds_arr[di].n);
    part += (uint64_t)((q - p + qb - 1) / qb);
  }
  uint64_t nqb_u;
  (void)block_excl_scan_u64<NT>(part, nqb_u, sh);
  const int64_t nqb = (int64_t)nqb_u;
  int64_t xd = xs;
  if (nqb > 0 && nqb * xs < target) {
    int64_t g = target / nqb;
    g = g < xs_cap ? g : xs_cap;
    g = g < item_cap / nqb ? g : item_cap / nqb;
    g = g < row_cap * xs / cnt ? g : row_cap * xs / cnt;
    xd = g > xd ? g : xd;
  }
  uint64_t base = 0;
  for (int c0 = 0; c0 < nds; c0 += NT) {  // uniform
    const int di = c0 + (int)threadIdx.x;
    int p = 0, q = 0;
    if (di < nds) {
      p = list_lower_bound(list, cnt, ds_arr[di].row_off);
      q = list_lower_bound(list, cnt, ds_arr[di].row_off + ds_arr[di].n);
    }
    uint64_t tot;
    const uint64_t off =
        block_excl_scan_u64<NT>((uint64_t)((q - p + qb - 1) / qb) * (uint64_t)xd, tot, sh);
    int w = (int)(base + off);
    for (int b = p; b < q; b += qb) {
      const int qc = q - b < qb ? q - b : qb;
      for (int s = 0; s < (int)xd; ++s) items[w++] = WorkItem{di, s, b, qc};
    }
    base += tot;
  }
  if (threadIdx.x == 0) {
    *nitems = (int)base;
    *xs_out = (int)xd;
  }
}
__global__ __launch_bounds__(256) void fallback_items_kernel(
    const DsDesc* __restrict__ ds_arr, int nds, const uint32_t* __restrict__ list,
    const int* __restrict__ count, int qb, int xs, int xs_cap, int64_t item_cap, int64_t row_cap,
    int target, WorkItem* __restrict__ items, int* __restrict__ nitems, int* __restrict__ xs_out) {
  fallback_items_block<256>(ds_arr, nds, list, count, qb, xs, xs_cap, item_cap, row_cap, target,
                            items, nitems, xs_out);
}

// Flag compaction fused with the work list it feeds (flag arrays of at most
// kScanSmall entries: one workgroup of 512 threads compacts, then builds the
// items from the list, which the barrier has made visible to it).
// nflag: the number of flagged rows (a count the re-rank keeps): none, the
// usual case, skips the scan.
__global__ __launch_bounds__(512) void flag_sweep_items_kernel(
    const uint32_t* __restrict__ flag, int64_t n, uint32_t* __restrict__ list,
    int* __restrict__ count, const DsDesc* __restrict__ ds_arr, int nds, int qw, int cus,
    double rate, WorkItem* __restrict__ items, int* __restrict__ nitems, int* __restrict__ nrows,
    const uint32_t* __restrict__ nflag) {
  if (*nflag == 0u) {  // uniform
    if (threadIdx.x == 0) {
      *count = 0;
      *nitems = 0;
      *nrows = 0;
    }
    return;
  }
  scan_small_body<uint32_t, 512, 32>(flag, n, nullptr, nullptr, list, count);
  __syncthreads();
  sweep_items_block<512>(ds_arr, nds, list, count, qw, cus, rate, items, nitems, nrows);
}
__global__ __launch_bounds__(512) void flag_fallback_items_kernel(
    const uint32_t* __restrict__ flag, int64_t n, uint32_t* __restrict__ list,
    int* __restrict__ count, const DsDesc* __restrict__ ds_arr, int nds, int qb, int xs,
    int xs_cap, int64_t item_cap, int64_t row_cap, int target, WorkItem* __restrict__ items,
    int* __restrict__ nitems, int* __restrict__ xs_out, const uint32_t* __restrict__ nflag) {
  if (nflag && *nflag == 0u) {  // uniform: no flagged row
    if (threadIdx.x == 0) {
      *count = 0;
      *nitems = 0;
      *xs_out = xs;
    }
    return;
  }
  scan_small_body<uint32_t, 512, 32>(flag, n, nullptr, nullptr, list, count);
  __syncthreads();
  fallback_items_block<512>(ds_arr, nds, list, count, qb, xs, xs_cap, item_cap, row_cap, target,
                            items, nitems, xs_out);
}

// The exact fallback's list when an escalation ran: the rows still flagged
// are among the escalated rows (the list its flag compaction left), so the
// list is compacted in place -- chunks of NT * PER entries, each read whole
// before any of its kept entries moves down -- and the work list built from
// it. Cheaper than compacting the whole flag array when few rows escalated.
template <int NT, int PER>
__device__ void list_compact_block(const uint32_t* __restrict__ flag, uint32_t* __restrict__ list,
                                   int* __restrict__ count) {
  __shared__ uint64_t sh[NT / 64];
  const int cnt = *count;
  uint64_t base = 0;
  for (int c0 = 0; c0 < cnt; c0 += NT * PER) {  // uniform
    uint32_t v[PER];
    bool keep[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int i = c0 + (int)threadIdx.x * PER + u;
      v[u] = i < cnt ? list[i] : 0u;
    }
    uint64_t m = 0;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int i = c0 + (int)threadIdx.x * PER + u;
      keep[u] = i < cnt && flag[v[u]] != 0u;
      m += keep[u] ? 1u : 0u;
    }
    uint64_t tot;
    uint64_t pos = base + block_excl_scan_u64<NT>(m, tot, sh);  // (its barriers: chunk read)
#pragma unroll
    for (int u = 0; u < PER; ++u)
      if (keep[u]) list[pos++] = v[u];
    base += tot;
  }
  __syncthreads();
  if (threadIdx.x == 0) *count = (int)base;
}

template __global__ void rerank_kernel<false>(const DsDesc*, int, int64_t, int, uint32_t*,
                                              const uint32_t*, const int*, uint32_t*, FbTail);
template __global__ void rerank_kernel<true>(const DsDesc*, int, int64_t, int, uint32_t*,
                                             const uint32_t*, const int*, uint32_t*, FbTail);

}  // namespace hm
