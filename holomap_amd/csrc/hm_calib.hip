// Smooth-kNN calibration (src/smooth_knn.rs:12-84) on the device.
//
// Thread per row. Mode CALIB_REF_F32 repeats the reference's float32
// bisection step for step (state in f32, psum folded left to right over
// row[1..], the f64 comparisons against log2(k), f32 mid-point updates,
// hi seeded at f32::MAX with doubling), with expf evaluated by the same
// algorithm the reference's libm crate uses (see hm_expf). The result is
// bit-identical to the reference arithmetic. Mode CALIB_F64 runs the same
// bracket search in f64 with f64 exp until the bracket stops moving; when
// the kNN ran on the device it keeps the f64 distances (knn_dist64), and the
// whole chain -- rho, the search, the floors and the directed weights
// exp(-(d - rho) / sigma) -- is evaluated in f64 and rounded to f32 once at
// the end, with each row's largest relative rounding error (distances,
// sigma, weights) stated in err3. Without them it searches on the f32
// distances widened.
//
// The floor for rows without a non-zero distance uses mean_all, a
// sequential f64 mean over all n*k distances (src/smooth_knn.rs:26). That
// fold is inherently serial, so it is evaluated lazily by one lane in
// calib_fixup_kernel, and only for datasets with such a row: every such row
// sets its dataset's flag zany[ds] (a plain store of 1; all writers agree).
#include "hm_internal.hpp"

namespace hm {

// f32 exp by the method of libm 0.2.16 expf (src/fmath.rs:119-122): argument
// reduction by k*ln2 (ln2 split hi/lo), the degree-2 rational Remez form
// exp(r) = 1 + (r*c/(2-c) - lo + hi), c = r - r^2 (P1 + r^2 P2), then an exact
// power-of-two scaling. Every operation is a single f32 rounding (this file
// is compiled without contraction).
// The method and its constants are those of FreeBSD msun's e_expf.c, which
// carries this notice:
//   Conversion to float by Ian Lance Taylor, Cygnus Support, ian@cygnus.com.
//   Copyright (C) 1993 by Sun Microsystems, Inc. All rights reserved.
//   Developed at SunPro, a Sun Microsystems, Inc. business.
//   Permission to use, copy, modify, and distribute this software is freely
//   granted, provided that this notice is preserved.
__device__ __forceinline__ float hm_scalbnf(float x, int n) {
  float y = x;
  if (n > 127) {
    y *= 0x1p127f;
    n -= 127;
    if (n > 127) {
      y *= 0x1p127f;
      n -= 127;
      if (n > 127) n = 127;
    }
  } else if (n < -126) {
    y *= 0x1p-126f * 0x1p24f;
    n += 126 - 24;
    if (n < -126) {
      y *= 0x1p-126f * 0x1p24f;
      n += 126 - 24;
      if (n < -126) n = -126;
    }
  }
  return y * __uint_as_float((uint32_t)(0x7f + n) << 23);
}

__device__ __forceinline__ float hm_expf(float x) {
  const float ln2hi = 6.9314575195e-01f, ln2lo = 1.4286067653e-06f;
  const float invln2 = 1.4426950216e+00f;
  const float P1 = 1.6666625440e-1f, P2 = -2.7667332906e-3f;
  uint32_t hx = __float_as_uint(x);
  const int sign = (int)(hx >> 31);
  hx &= 0x7fffffffu;
  if (hx >= 0x42aeac50u) {
    if (hx > 0x7f800000u) return x;
    if (hx >= 0x42b17218u && !sign) return x * 0x1p127f;
    if (sign && hx >= 0x42cff1b5u) return 0.0f;
  }
  int k;
  float hi, lo;
  if (hx > 0x3eb17218u) {
    if (hx > 0x3f851592u)
      k = (int)(invln2 * x + (sign ? -0.5f : 0.5f));
    else
      k = 1 - sign - sign;
    const float kf = (float)k;
    hi = x - kf * ln2hi;
    lo = kf * ln2lo;
    x = hi - lo;
  } else if (hx > 0x39000000u) {
    k = 0;
    hi = x;
    lo = 0.0f;
  } else {
    return 1.0f + x;
  }
  const float xx = x * x;
  const float c = x - xx * (P1 + xx * P2);
  const float y = 1.0f + (x * c / (2.0f - c) - lo + hi);
  return k == 0 ? y : hm_scalbnf(y, k);
}

__device__ __forceinline__ int calib_find_ds(const DsDesc* ds, int nds, int64_t grow) {
  int lo = 0, hi = nds - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (ds[mid].row_off <= grow) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Directed membership strength of edge t of row i (global row g) of
// dataset D (src/fuzzy.rs:31-48): w = 1 if d - rho <= 0 or sigma == 0, else
// exp(-((d - rho) / sigma)); self edges and w == 0 are dropped (key1 =
// sentinel). Written for the symmetrisation's transpose: key1 = global
// column, wval = w, and each kept edge takes the next slot of its column's
// in-edge count, erank = the count before (an integer atomic: which edge
// gets which slot depends on scheduling, the counts do not).
__device__ __forceinline__ void directed_edge(const DsDesc& D, int64_t i, int64_t g, int t,
                                              float d, float rho, float sigma, uint32_t sentinel,
                                              uint32_t* __restrict__ key1,
                                              uint32_t* __restrict__ indeg,
                                              float* __restrict__ wval,
                                              uint32_t* __restrict__ erank) {
  const int64_t le = i * D.k + t;
  const int64_t e = D.edge_off + le;
  const uint32_t col = D.knn_idx[le];
  const float dd = d - rho;
  const float w = (dd <= 0.0f || sigma == 0.0f) ? 1.0f : hm_expf(-(dd / sigma));
  const bool keep = (int64_t)col != i && w != 0.0f;
  key1[e] = keep ? (uint32_t)(D.row_off + col) : sentinel;
  wval[e] = w;
  erank[e] = keep ? atomicAdd(indeg + D.row_off + col, 1u) : 0u;
}

// The same edge from the f64 chain: w = exp(-((d - rho) / sigma)) in f64,
// rounded to f32 once (dropped when that is 0). Returns the relative error
// of the f32 weight (0 for dropped or unit weights).
__device__ __forceinline__ double directed_edge64(const DsDesc& D, int64_t i, int64_t g, int t,
                                                  double d, double rho, double sigma,
                                                  uint32_t sentinel, uint32_t* __restrict__ key1,
                                                  uint32_t* __restrict__ indeg,
                                                  float* __restrict__ wval,
                                                  uint32_t* __restrict__ erank) {
  const int64_t le = i * D.k + t;
  const int64_t e = D.edge_off + le;
  const uint32_t col = D.knn_idx[le];
  const double dd = d - rho;
  const double w = (dd <= 0.0 || sigma == 0.0) ? 1.0 : exp(-(dd / sigma));
  const float wf = (float)w;
  const bool keep = (int64_t)col != i && wf != 0.0f;
  if (key1) {
    key1[e] = keep ? (uint32_t)(D.row_off + col) : sentinel;
    wval[e] = wf;
    erank[e] = keep ? atomicAdd(indeg + D.row_off + col, 1u) : 0u;
  }
  return keep ? fabs((double)wf - w) / w : 0.0;
}


// The calibration kernels also write the row's directed edges when key1 is
// set (the symmetrisation follows): a row's sigma is final here except for
// the global-mean floor of rows with rho == 0, whose distances are all 0, so
// their weights are 1 whatever sigma is.
//
// target_log2k[ds] is log2(k) computed on the host (glibc), as the
// reference's std f64::log2 does.
//
// Thread per row. KM > 0: the row's k <= KM distances are held in registers
// (loops unrolled to KM with a t < k guard), for jobs with enough rows that
// the VALU work, not latency, bounds the calibration: a lane evaluates all
// k - 1 terms of its row, where the lane-group kernel spends a lane per term
// and repeats the bracket updates in every lane. KM = 0: rows of any k, read
// from memory.
#define HM_ROW_LOOP(T0, BODY)                                  \
  do {                                                         \
    if constexpr (KM > 0) {                                    \
      _Pragma("unroll") for (int t = (T0); t < KM; ++t) {      \
        if (t < k) { BODY }                                    \
      }                                                        \
    } else {                                                   \
      for (int t = (T0); t < k; ++t) { BODY }                  \
    }                                                          \
  } while (0)
template <int KM>
__global__ __launch_bounds__(256) void smooth_knn_kernel(
    const DsDesc* __restrict__ ds, int nds, int64_t total_rows,
    const double* __restrict__ target_log2k, int mode, float* __restrict__ sigma,
    float* __restrict__ rho, double* __restrict__ sigma64, uint32_t* __restrict__ zany,
    uint32_t sentinel, uint32_t* __restrict__ key1, uint32_t* __restrict__ indeg,
    float* __restrict__ wval, uint32_t* __restrict__ erank) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total_rows) return;
  const int di = calib_find_ds(ds, nds, g);
  const DsDesc& D = ds[di];
  const int64_t i = g - D.row_off;
  const int k = D.k;
  const float* row = D.knn_dist + i * k;
  const double target = target_log2k[di];
  float xr[KM > 0 ? KM : 1];
  if constexpr (KM > 0) {
#pragma unroll
    for (int t = 0; t < KM; ++t) xr[t] = t < k ? row[t] : 0.0f;
  }
#define X(t) (KM > 0 ? xr[(KM > 0 ? (t) : 0)] : row[t])

  // rho: the first positive distance in row order
  float r = 0.0f;
  if constexpr (KM > 0) {
#pragma unroll
    for (int t = KM - 1; t >= 0; --t)
      if (t < k && xr[t] > 0.0f) r = xr[t];
  } else {
    for (int t = 0; t < k; ++t) {
      if (row[t] > 0.0f) { r = row[t]; break; }
    }
  }
  rho[g] = r;
  if (mode == CALIB_REF_F32) {
    const float fmax = 3.40282347e+38f;
    float lo = 0.0f, hi = fmax, mid = 1.0f;
    for (int it = 0; it < 64; ++it) {
      float psum = 0.0f;
      HM_ROW_LOOP(1, {
        const float dd = X(t) - r;
        if (dd > 0.0f)
          psum += hm_expf(-(dd / mid));
        else
          psum += 1.0f;
      });
      if (fabs((double)psum - target) < 1e-5) break;
      if ((double)psum > target) {
        hi = mid;
        mid = (lo + hi) / 2.0f;
      } else {
        lo = mid;
        if (hi >= fmax)
          mid *= 2.0f;
        else
          mid = (lo + hi) / 2.0f;
      }
    }
    float s = mid;
    if (r > 0.0f) {
      double acc = 0.0;
      HM_ROW_LOOP(0, { acc = acc + (double)X(t); });
      const float fl = 1e-3f * (float)(acc / (double)k);
      if (s < fl) s = fl;
    } else {
      zany[di] = 1u;  // floor needs mean_all: applied by calib_fixup_kernel
    }
    sigma[g] = s;
    if (sigma64) sigma64[g] = (double)s;
    if (key1) HM_ROW_LOOP(0, { directed_edge(D, i, g, t, X(t), r, s, sentinel, key1, indeg, wval, erank); });
  } else {
    // the f64 chain from the f64 distances when they were kept, else from
    // the f32 distances widened
    const double* row64 = D.knn_dist64 ? D.knn_dist64 + i * k : nullptr;
#define X64(t) (row64 ? row64[t] : (double)X(t))
    double rr = (double)r;
    if (row64) {
      rr = 0.0;
      for (int t = 0; t < k; ++t) {
        if (row64[t] > 0.0) { rr = row64[t]; break; }
      }
      rho[g] = (float)rr;
    }
    double lo = 0.0, hi = INFINITY, mid = 1.0;
    for (int it = 0; it < 2200; ++it) {
      double psum = 0.0;
      HM_ROW_LOOP(1, {
        const double dd = X64(t) - rr;
        psum += dd > 0.0 ? exp(-(dd / mid)) : 1.0;
      });
      if (psum == target) break;
      double nmid;
      if (psum > target) {
        hi = mid;
        nmid = 0.5 * (lo + hi);
      } else {
        lo = mid;
        nmid = isinf(hi) ? mid * 2.0 : 0.5 * (lo + hi);
      }
      if (nmid == mid || isinf(nmid)) break;
      mid = nmid;
    }
    double s = mid;
    if (rr > 0.0) {
      double acc = 0.0;
      HM_ROW_LOOP(0, { acc += X64(t); });
      const double fl = 1e-3 * (acc / (double)k);
      if (s < fl) s = fl;
    } else {
      zany[di] = 1u;
    }
    if (sigma64) sigma64[g] = s;
    sigma[g] = (float)s;
    if (row64) {
      double ed = 0.0, ew = 0.0;
      HM_ROW_LOOP(0, {
        ed = fmax(ed, rel_err(X(t), row64[t]));
        ew = fmax(ew, directed_edge64(D, i, g, t, row64[t], rr, s, sentinel, key1, indeg, wval,
                                      erank));
      });
      D.err3[i * 3 + 0] = (float)ed;
      D.err3[i * 3 + 1] = (float)rel_err((float)s, s);
      D.err3[i * 3 + 2] = (float)ew;
    } else {
      const float sf = (float)s;
      if (key1) HM_ROW_LOOP(0, { directed_edge(D, i, g, t, X(t), r, sf, sentinel, key1, indeg, wval, erank); });
    }
#undef X64
  }
#undef X
}
#undef HM_ROW_LOOP
#define HM_CALIB_T_INST(KM)                                                                    \
  template __global__ void smooth_knn_kernel<KM>(const DsDesc*, int, int64_t, const double*, int, \
                                                 float*, float*, double*, uint32_t*, uint32_t,   \
                                                 uint32_t*, uint32_t*, float*, uint32_t*);
HM_CALIB_T_INST(0)
HM_CALIB_T_INST(16)
HM_CALIB_T_INST(32)
HM_CALIB_T_INST(64)
#undef HM_CALIB_T_INST

// Same calibration with G lanes per row (k <= G <= 64): lane t holds the
// row's distance t and evaluates term t of every bisection step (the exp
// and its division), and every lane of the group folds the terms in the
// reference's order t = 1 .. k-1 from the group's registers, so the sums,
// the bracket updates and the outputs are those of smooth_knn_kernel bit
// for bit. Loop bounds depend on the row only, so a group stays converged.
// The value of lane U of the calling lane's G-lane group: a DPP row
// broadcast (one VALU move) for 16-lane groups, an LDS permute otherwise.
// (The kernel folds with these for 16- and 32-lane groups; 64-lane groups
// keep a loop of permutes, the unrolled form being too long.)
template <int G, int U>
__device__ __forceinline__ float grp_lane_f32(float x) {
  if constexpr (G == 16)
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x150 + U, 0xf, 0xf, false));
  else
    return __shfl(x, U, G);
}
template <int G, int U>
__device__ __forceinline__ double grp_lane_f64(double x) {
  if constexpr (G == 16) {
    const uint64_t u = (uint64_t)__double_as_longlong(x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, 0x150 + U, 0xf, 0xf, false);
    const uint32_t hi =
        (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), 0x150 + U, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
  } else {
    return __shfl(x, U, G);
  }
}
// psum + e_1 + e_2 + ... + e_{k-1}, left to right (e_t: lane t's value).
// The lanes t >= k hold +0.0 (the callers compute terms only in lanes 1 ..
// k-1), and adding +0.0 leaves the non-negative partial sums as they are, so
// every lane is added unmasked and the result is the reference's fold bit
// for bit (one broadcast and one add per term).
template <int G, int U = 1, typename T>
__device__ __forceinline__ T grp_fold(T e, int k, T psum) {
  if constexpr (U < G) {
    T v;
    if constexpr (sizeof(T) == 4) v = grp_lane_f32<G, U>(e);
    else v = grp_lane_f64<G, U>(e);
    psum += v;
    return grp_fold<G, U + 1>(e, k, psum);
  } else {
    return psum;
  }
}
// Both 16-lane rows of every 32-lane group hold row 0's values (lo) and
// row 1's values (hi): v_permlane16_swap exchanges the odd rows of its first
// operand with the even rows of its second, within each pair of rows.
__device__ __forceinline__ void rows_of_pair(uint32_t x, uint32_t& lo, uint32_t& hi) {
  const auto r = __builtin_amdgcn_permlane16_swap(x, x, false, false);
  lo = r[0];
  hi = r[1];
}
template <typename T>
__device__ __forceinline__ void rows_of_pair_t(T x, T& lo, T& hi) {
  if constexpr (sizeof(T) == 4) {
    uint32_t a, b;
    rows_of_pair(__float_as_uint(x), a, b);
    lo = __uint_as_float(a);
    hi = __uint_as_float(b);
  } else {
    const uint64_t u = (uint64_t)__double_as_longlong(x);
    uint32_t la, lb, ha, hb;
    rows_of_pair((uint32_t)u, la, lb);
    rows_of_pair((uint32_t)(u >> 32), ha, hb);
    lo = __longlong_as_double((long long)(((uint64_t)ha << 32) | la));
    hi = __longlong_as_double((long long)(((uint64_t)hb << 32) | lb));
  }
}
// The same fold for 32-lane groups: terms 1..15 from row 0's values, then
// 16..31 from row 1's, each by a DPP row broadcast.
template <int U, typename T>
__device__ __forceinline__ T row_fold(T v, int k, int base, T psum) {
  if constexpr (U < 16) {
    T w;
    if constexpr (sizeof(T) == 4) w = grp_lane_f32<16, U>(v);
    else w = grp_lane_f64<16, U>(v);
    psum += w;  // (+0.0 from the lanes past k, as in grp_fold)
    return row_fold<U + 1>(v, k, base, psum);
  } else {
    return psum;
  }
}
template <typename T>
__device__ __forceinline__ T grp_fold32(T e, int k, T psum) {
  T lo, hi;
  rows_of_pair_t(e, lo, hi);
  psum = row_fold<1>(lo, k, 0, psum);
  return row_fold<0>(hi, k, 16, psum);
}

template <int G>
__global__ __launch_bounds__(256) void smooth_knn_group_kernel(
    const DsDesc* __restrict__ ds, int nds, int64_t total_rows,
    const double* __restrict__ target_log2k, int mode, float* __restrict__ sigma,
    float* __restrict__ rho, double* __restrict__ sigma64, uint32_t* __restrict__ zany,
    uint32_t sentinel, uint32_t* __restrict__ key1, uint32_t* __restrict__ indeg,
    float* __restrict__ wval, uint32_t* __restrict__ erank) {
  static_assert(G >= 2 && G <= 64 && (G & (G - 1)) == 0 && 256 % G == 0, "group size");
  const int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int l = (int)(threadIdx.x % G);
  if (g >= total_rows) return;  // whole groups leave together
  const int di = calib_find_ds(ds, nds, g);
  const DsDesc& D = ds[di];
  const int64_t i = g - D.row_off;
  const int k = D.k;
  const float x = l < k ? D.knn_dist[i * k + l] : 0.0f;
  const double target = target_log2k[di];

  // rho: the first positive distance in row order
  float r = 0.0f;
  for (int t = 0; t < k; ++t) {
    const float v = __shfl(x, t, G);
    if (v > 0.0f) { r = v; break; }
  }
  const bool term = l >= 1 && l < k;  // lanes holding a term of the sum
  if (mode == CALIB_REF_F32) {
    const float fmax = 3.40282347e+38f;
    float lo = 0.0f, hi = fmax, mid = 1.0f;
    for (int it = 0; it < 64; ++it) {
      float e = 0.0f;
      if (term) {
        const float dd = x - r;
        e = dd > 0.0f ? hm_expf(-(dd / mid)) : 1.0f;
      }
      float psum = 0.0f;
      if constexpr (G == 16) {
        psum = grp_fold<G>(e, k, 0.0f);
      } else if constexpr (G == 32) {
        psum = grp_fold32(e, k, 0.0f);
      } else {
        for (int t = 1; t < k; ++t) psum += __shfl(e, t, G);
      }
      if (fabs((double)psum - target) < 1e-5) break;
      if ((double)psum > target) {
        hi = mid;
        mid = (lo + hi) / 2.0f;
      } else {
        lo = mid;
        if (hi >= fmax)
          mid *= 2.0f;
        else
          mid = (lo + hi) / 2.0f;
      }
    }
    float s = mid;
    bool z = true;
    if (r > 0.0f) {
      double acc = 0.0;
      for (int t = 0; t < k; ++t) acc = acc + (double)__shfl(x, t, G);
      const float fl = 1e-3f * (float)(acc / (double)k);
      if (s < fl) s = fl;
      z = false;
    }
    if (l == 0) {
      rho[g] = r;
      if (z) zany[di] = 1u;  // the floor needs mean_all (calib_fixup_kernel)
      sigma[g] = s;
      if (sigma64) sigma64[g] = (double)s;
    }
    if (key1 && l < k) directed_edge(D, i, g, l, x, r, s, sentinel, key1, indeg, wval, erank);
  } else {
    // the f64 chain from the f64 distances when they were kept (lane t
    // holds distance t in f64), else from the f32 distances widened
    const bool chain = D.knn_dist64 != nullptr;
    const double x64 = chain ? (l < k ? D.knn_dist64[i * k + l] : 0.0) : (double)x;
    double rr = (double)r;
    if (chain) {
      rr = 0.0;
      for (int t = 0; t < k; ++t) {
        const double v = __shfl(x64, t, G);
        if (v > 0.0) { rr = v; break; }
      }
    }
    double lo = 0.0, hi = INFINITY, mid = 1.0;
    for (int it = 0; it < 2200; ++it) {
      double e = 0.0;
      if (term) {
        const double dd = x64 - rr;
        e = dd > 0.0 ? exp(-(dd / mid)) : 1.0;
      }
      double psum = 0.0;
      if constexpr (G == 16) {
        psum = grp_fold<G>(e, k, 0.0);
      } else if constexpr (G == 32) {
        psum = grp_fold32(e, k, 0.0);
      } else {
        for (int t = 1; t < k; ++t) psum += __shfl(e, t, G);
      }
      if (psum == target) break;
      double nmid;
      if (psum > target) {
        hi = mid;
        nmid = 0.5 * (lo + hi);
      } else {
        lo = mid;
        nmid = isinf(hi) ? mid * 2.0 : 0.5 * (lo + hi);
      }
      if (nmid == mid || isinf(nmid)) break;
      mid = nmid;
    }
    double s = mid;
    bool z = true;
    if (rr > 0.0) {
      double acc = 0.0;
      for (int t = 0; t < k; ++t) acc += __shfl(x64, t, G);
      const double fl = 1e-3 * (acc / (double)k);
      if (s < fl) s = fl;
      z = false;
    }
    if (l == 0) {
      rho[g] = chain ? (float)rr : r;
      if (z) zany[di] = 1u;
      if (sigma64) sigma64[g] = s;
      sigma[g] = (float)s;
    }
    if (chain) {
      double ed = 0.0, ew = 0.0;
      if (l < k) {
        ed = rel_err(x, x64);
        ew = directed_edge64(D, i, g, l, x64, rr, s, sentinel, key1, indeg, wval, erank);
      }
      for (int o = G / 2; o > 0; o >>= 1) {
        ed = fmax(ed, __shfl_xor(ed, o, G));
        ew = fmax(ew, __shfl_xor(ew, o, G));
      }
      if (l == 0) {
        D.err3[i * 3 + 0] = (float)ed;
        D.err3[i * 3 + 1] = (float)rel_err((float)s, s);
        D.err3[i * 3 + 2] = (float)ew;
      }
    } else if (key1 && l < k) {
      directed_edge(D, i, g, l, x, r, (float)s, sentinel, key1, indeg, wval, erank);
    }
  }
}

#define HM_CALIB_G_INST(G)                                                                    \
  template __global__ void smooth_knn_group_kernel<G>(const DsDesc*, int, int64_t,              \
                                                      const double*, int, float*, float*,       \
                                                      double*, uint32_t*, uint32_t, uint32_t*,  \
                                                      uint32_t*, float*, uint32_t*);
HM_CALIB_G_INST(16)
HM_CALIB_G_INST(32)
HM_CALIB_G_INST(64)
#undef HM_CALIB_G_INST

// Applies the global-mean floor to rows with rho == 0 (calib_fixup_ds), one
// workgroup per dataset, when no symmetrisation follows (else its last
// kernel carries these workgroups).
__global__ __launch_bounds__(256) void calib_fixup_kernel(const DsDesc* __restrict__ ds, int mode,
                                                          float* __restrict__ sigma,
                                                          double* __restrict__ sigma64,
                                                          const float* __restrict__ rho,
                                                          uint32_t* __restrict__ zany) {
  calib_fixup_ds(ds[blockIdx.x], (int)blockIdx.x, mode, sigma, sigma64, rho, zany);
}

// Directed membership strengths (src/fuzzy.rs:31-48), one thread per edge,
// for jobs whose sigma / rho are given rather than calibrated.
__global__ void directed_kernel(const DsDesc* __restrict__ ds, int nds, int64_t total_edges,
                                const float* __restrict__ sigma, const float* __restrict__ rho,
                                uint32_t sentinel, uint32_t* __restrict__ key1,
                                uint32_t* __restrict__ indeg, float* __restrict__ wval,
                                uint32_t* __restrict__ erank) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total_edges) return;
  int lo = 0, hi = nds - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (ds[mid].edge_off <= e) lo = mid; else hi = mid - 1;
  }
  const DsDesc& D = ds[lo];
  const int64_t le = e - D.edge_off;
  const int64_t i = le / D.k;
  const int64_t g = D.row_off + i;
  directed_edge(D, i, g, (int)(le - i * D.k), D.knn_dist[le], rho[g], sigma[g], sentinel, key1,
                indeg, wval, erank);
}

// Standalone expf over an array (parity tests of the device expf).
__global__ void expf_kernel(const float* __restrict__ x, float* __restrict__ y, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] = hm_expf(x[t]);
}

}  // namespace hm
