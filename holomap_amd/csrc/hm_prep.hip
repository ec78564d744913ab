// Data preparation for the candidate sweep.
//
// For every row of every dataset of a batch:
//   * prep_kernel     (16 lanes per row): the row's norm, the reference's
//     sequential f64 fold sum_f x_f^2 (src/metric.rs:61-66 keeps na/nb as
//     such folds), every lane of the row folding the 16 values of a slice
//     in order from DPP broadcasts: the cosine denominator factor sqrt(na)
//     bit-exactly and the Euclidean key bias ||x||^2 rounded to f32; a
//     zero-padded f32 copy (n_pad x kp), the candidate-selection operand
//     (euclid: the copy itself; cosine: the row scaled to unit length), its
//     bf16 (RNE) image, the statistics the certification bound needs
//     (representation error norms), and their per-dataset maxima: each
//     workgroup folds the maxima of its rows into smax by integer atomic
//     maxima of the bit patterns (the statistics are non-negative floats,
//     whose patterns order as their values; max is exact and order-free, so
//     the result is the same on every run). smax is clear when prep runs:
//     zero-filled when the job is made, cleared again by the kNN stage's
//     last kernel after the re-ranks have read it.
#include "hm_internal.hpp"

namespace hm {

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  uint32_t r = ((u >> 16) & 1u) + 0x7fffu;
  return (uint16_t)((u + r) >> 16);  // inputs are finite
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
  return __uint_as_float((uint32_t)h << 16);
}

// f64 DPP move (two 32-bit halves); lanes without a source read +0.0.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_f64(double x) {
  const uint64_t u = (uint64_t)__double_as_longlong(x);
  const uint32_t lo =
      (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, ROWS, 0xf, false);
  const uint32_t hi =
      (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, ROWS, 0xf, false);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
// Inclusive scan of each 16-lane DPP row (row shifts 1, 2, 4, 8); lane 15
// of the row ends with the row's sum.
__device__ __forceinline__ double row_sum15(double x) {
  x += dpp_f64<0x111, 0xf>(x);
  x += dpp_f64<0x112, 0xf>(x);
  x += dpp_f64<0x114, 0xf>(x);
  x += dpp_f64<0x118, 0xf>(x);
  return x;
}

__device__ __forceinline__ int find_ds(const DsDesc* ds, int nds, int64_t grow) {
  int lo = 0, hi = nds - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (ds[mid].row_off <= grow) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Thread per global row: sequential f64 fold of x_f^2 in feature order.
// zero_g (optional): a per-row count cleared for a later kernel of the run;
// smax_g: the nds x HM_NSTAT statistics maxima, cleared for prep_kernel,
// and the word after them (the re-rank's count of flagged rows).
__device__ __forceinline__ void prep_row(const DsDesc* __restrict__ ds, int di, int64_t pr,
                                         const int64_t* __restrict__ pad_off,
                                         uint32_t* __restrict__ zero_g, int mode_bits, int lane,
                                         float (&stout)[HM_NSTAT]);

// 16 lanes (one DPP row) per (dataset, padded row), four rows per wave, so
// a wave overlaps the memory round trips of four rows. Writes rnorm, xr,
// op32 (cosine), op16 (+ op16lo), colbias and rstat, and clears the row's
// in-degree count zero_g[row] when a symmetrisation follows (null
// otherwise). mode_bits: 1 = produce op16 (bf16 sweep), 2 = bf16x3 stats. A
// workgroup's 16 rows belong to one dataset (padded row counts are
// multiples of 128); it folds the maxima of their statistics into the
// dataset's smax (atomic maxima of the non-negative values' bit patterns).
__global__ __launch_bounds__(256) void prep_kernel(
    const DsDesc* __restrict__ ds, int nds, int64_t total_pad_rows,
    const int64_t* __restrict__ pad_off /* nds+1 prefix of n_pad */,
    uint32_t* __restrict__ zero_g, int mode_bits) {
  __shared__ float red[16][HM_NSTAT];
  __shared__ int di_s;
  if (threadIdx.x == 0) {  // the workgroup's dataset, by padded-row prefix
    const int64_t pr0 = (int64_t)blockIdx.x * 16;
    int lo = 0, hi = nds - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (pad_off[mid] <= pr0) lo = mid; else hi = mid - 1;
    }
    di_s = lo;
  }
  __syncthreads();
  const int di = di_s;
  const int64_t pr = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4;
  float st[HM_NSTAT];
#pragma unroll
  for (int s = 0; s < HM_NSTAT; ++s) st[s] = 0.0f;
  if (pr < total_pad_rows)  // (always: the grid covers whole 16-row groups)
    prep_row(ds, di, pr, pad_off, zero_g, mode_bits, (int)(threadIdx.x & 15), st);
  if ((threadIdx.x & 15) == 15) {
#pragma unroll
    for (int s = 0; s < HM_NSTAT; ++s) red[threadIdx.x >> 4][s] = st[s];
  }
  __syncthreads();
  if (threadIdx.x < HM_NSTAT) {
    float m = 0.0f;  // fmaxf: a NaN never enters (none arises from finite rows)
#pragma unroll
    for (int r = 0; r < 16; ++r) m = fmaxf(m, red[r][threadIdx.x]);
    // every workgroup of the dataset updates the same eight words, where
    // atomics serialise: read the current maximum (coherently; it only
    // grows during the run) and update only when this one exceeds it
    uint32_t* p = (uint32_t*)const_cast<float*>(ds[di].smax) + threadIdx.x;
    const uint32_t cur = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (__float_as_uint(m) > cur) atomicMax(p, __float_as_uint(m));
  }
}

// s + v_0^2 + v_1^2 + ... + v_15^2 in f64, left to right (v_u: lane u of
// the calling lane's 16-lane DPP row, by a row broadcast): every lane of the
// row returns the same sum.
template <int U = 0>
__device__ __forceinline__ double row_fold_sq(float v, double s) {
  if constexpr (U < 16) {
    const float vu =
        __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x150 + U, 0xf, 0xf, false));
    const double t = (double)vu;
    return row_fold_sq<U + 1>(v, s + t * t);
  } else {
    return s;
  }
}

// One padded row of dataset di by 16 lanes (prep_kernel's body).
__device__ __forceinline__ void prep_row(const DsDesc* __restrict__ ds, int di, int64_t pr,
                                         const int64_t* __restrict__ pad_off,
                                         uint32_t* __restrict__ zero_g, int mode_bits, int lane,
                                         float (&stout)[HM_NSTAT]) {
  const DsDesc& D = ds[di];
  const int64_t i = pr - pad_off[di];
  const int kp = D.kp, d = D.d;
  float* xr = const_cast<float*>(D.xr) + i * kp;
  float* op32 = const_cast<float*>(D.op32) + i * kp;
  uint16_t* op16 = D.op16 ? const_cast<uint16_t*>(D.op16) + i * D.kp16 : nullptr;
  uint16_t* op16lo = D.op16lo ? const_cast<uint16_t*>(D.op16lo) + i * D.kp16 : nullptr;
  for (int f = kp + lane; f < D.kp16; f += 16) {  // bf16 rows are padded further
    if (op16) op16[f] = 0;
    if (op16lo) op16lo[f] = 0;
  }
  const bool cosine = D.metric == 1;
  const bool real = i < D.n;
  // the norm fold, d values in order (padding lanes add +0.0, which leaves
  // the non-negative partial sums as they are); the loop is uniform over the
  // workgroup (one dataset)
  double s = 0.0;
  for (int f0 = 0; f0 < d; f0 += 16) {
    const int f = f0 + lane;
    s = row_fold_sq((real && f < d) ? D.x[i * d + f] : 0.0f, s);
  }
  const double rn = real ? sqrt(s) : 0.0;
  const float nz = real ? (float)s : 0.0f;
  if (real && lane == 0) {
    const_cast<double*>(D.rnorm)[i] = rn;
    if (zero_g) zero_g[D.row_off + i] = 0u;
  }
  const bool zero_row = real && rn == 0.0;
  // statistics accumulated per lane, reduced by a fixed butterfly
  double s_xn = 0, s_en = 0, s_xo = 0, s_hn = 0, s_ln = 0, s_hln = 0, s_rn = 0;
  for (int f = lane; f < kp; f += 16) {
    float v = (real && f < d) ? D.x[i * d + f] : 0.0f;
    xr[f] = v;
    float w = v;
    double wexact = (double)v;
    if (cosine) {
      if (real && !zero_row) {
        wexact = (double)v / rn;
        w = (float)wexact;
      } else {
        wexact = 0.0;
        w = 0.0f;
      }
      op32[f] = w;
    }
    s_xo += (double)w * (double)w;
    uint16_t h = f32_to_bf16_rne(w);
    float hf = bf16_to_f32(h);
    if (op16) op16[f] = h;
    if (op16lo) op16lo[f] = f32_to_bf16_rne(w - hf);  // w - hf is exact in f32
    if (mode_bits & 1) {
      double e = wexact - (double)hf;
      s_xn += (double)hf * (double)hf;
      s_en += e * e;
    } else {
      double e = wexact - (double)w;
      s_xn += (double)w * (double)w;
      s_en += e * e;
    }
    if (mode_bits & 2) {
      float lof = w - hf;  // exact in f32
      float lf = bf16_to_f32(f32_to_bf16_rne(lof));
      double r = wexact - (double)hf - (double)lf;
      s_hn += (double)hf * hf;
      s_ln += (double)lf * lf;
      double hl = (double)hf + (double)lf;
      s_hln += hl * hl;
      s_rn += r * r;
    }
  }
  // the row's sums in its lane 15 (the statistics feed error bounds, which
  // hold for any summation order)
  s_xn = row_sum15(s_xn);
  s_en = row_sum15(s_en);
  s_xo = row_sum15(s_xo);
  if (mode_bits & 2) {  // the bf16x3 statistics (zero otherwise)
    s_hn = row_sum15(s_hn);
    s_ln = row_sum15(s_ln);
    s_hln = row_sum15(s_hln);
    s_rn = row_sum15(s_rn);
  }
  if (lane == 15) {
    float* cb = const_cast<float*>(D.colbias);
    // the sweep's accumulator start bias_j / s (s = -2 euclidean, -1 cosine;
    // exact): key bias ||x_j||^2, or -inf for a zero row under cosine
    if (!real) {
      cb[i] = 0.0f;
    } else if (cosine) {
      cb[i] = zero_row ? INFINITY : 0.0f;
    } else {
      cb[i] = -0.5f * nz;
    }
    if (real) {
      // norms rounded up by a relative 2^-20 so the stored f32 bounds the
      // f64 value; the f64 partial sums themselves are within d*2^-53.
      const double up = 1.0 + 0x1p-20;
      float* st = const_cast<float*>(D.rstat) + i * HM_NSTAT;
      st[ST_NZ] = cosine ? 0.0f : nz;
      st[ST_XN] = (float)(sqrt(s_xn) * up);
      st[ST_EN] = (float)(sqrt(s_en) * up);
      st[ST_XO] = (float)(sqrt(s_xo) * up);
      st[ST_HN] = (float)(sqrt(s_hn) * up);
      st[ST_LN] = (float)(sqrt(s_ln) * up);
      st[ST_HLN] = (float)(sqrt(s_hln) * up);
      st[ST_RN] = (float)(sqrt(s_rn) * up);
#pragma unroll
      for (int k = 0; k < HM_NSTAT; ++k) stout[k] = st[k];
    }
  }
}

}  // namespace hm

namespace hm {

// Index of the first non-finite value of x[0 .. len), or ~0 (input
// validation of src/api.rs:142-144 on device data), in two launches without
// atomics: nonfinite_part_kernel writes each workgroup's smallest such
// index, nonfinite_min_kernel reduces the partial results. Thread t visits
// t, t + stride, ... in increasing order, so its first hit is its smallest.
__device__ __forceinline__ uint64_t block_min_u64(uint64_t v) {
  __shared__ uint64_t red[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t t = shfl_xor_u64(v, o);
    v = t < v ? t : v;
  }
  if (lane == 0) red[wave] = v;
  __syncthreads();
  uint64_t r = red[0];
#pragma unroll
  for (int w = 1; w < 4; ++w) r = red[w] < r ? red[w] : r;
  return r;
}

__global__ __launch_bounds__(256) void nonfinite_part_kernel(const float* __restrict__ x,
                                                             int64_t len,
                                                             uint64_t* __restrict__ part) {
  uint64_t best = ~0ull;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < len; i += (int64_t)gridDim.x * 256) {
    if ((__float_as_uint(x[i]) & 0x7f800000u) == 0x7f800000u) {
      best = (uint64_t)i;
      break;
    }
  }
  const uint64_t m = block_min_u64(best);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// f32 L2 normalisation of each row (crates/holomap-clusterer/src/
// pipeline.rs:151-162): norm = sqrt of the squares folded left to right from
// -0.0 (Rust's f32 Sum), a zero row left as it is, else every value divided
// by the norm; a thread per row keeps the fold's order.
__global__ void l2_normalise_kernel(const float* __restrict__ x, int64_t n, int d,
                                    float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* r = x + i * d;
  float s = -0.0f;
  for (int f = 0; f < d; ++f) s = s + r[f] * r[f];
  const float norm = sqrtf(s);
  float* o = out + i * d;
  if (norm == 0.0f) {
    for (int f = 0; f < d; ++f) o[f] = r[f];
  } else {
    for (int f = 0; f < d; ++f) o[f] = r[f] / norm;
  }
}

__global__ __launch_bounds__(256) void nonfinite_min_kernel(const uint64_t* __restrict__ part,
                                                            int nparts,
                                                            uint64_t* __restrict__ out) {
  uint64_t best = ~0ull;
  for (int i = threadIdx.x; i < nparts; i += 256) best = part[i] < best ? part[i] : best;
  const uint64_t m = block_min_u64(best);
  if (threadIdx.x == 0) *out = m;
}

}  // namespace hm
