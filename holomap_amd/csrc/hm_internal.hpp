// Internal definitions shared by the HIP kernels and the host runtime of the
// holomap kNN-graph library (gfx950 / CDNA4, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define HM_WAVE 64

namespace hm {

// ---------------------------------------------------------------------------
// Per-dataset descriptor, resident in device memory. One per dataset of a
// batch; a single fit uses a batch of one. Row ids inside a dataset are local
// (0..n-1); `row_off` places the dataset's rows in the batch-global row space
// that per-row arrays (flags, sigma, rho, ...) are indexed by.
// ---------------------------------------------------------------------------
struct DsDesc {
  const float* x;        // raw input, n x d, row-major (caller's layout)
  const float* xr;       // raw input padded: n_pad x kp f32 (zero-filled)
  const float* op32;     // candidate-selection operand, f32: n_pad x kp
                         //   (euclidean: == xr; cosine: rows / ||row||)
  const uint16_t* op16;  // same operand in bf16 (RNE), n_pad x kp16, or null
  const uint16_t* op16lo;  // bf16 of (operand - op16), n_pad x kp16, or null
  const float* colbias;  // n_pad: key bias / s: euclid -||x_j||^2/2 (f32); cosine 0 / +inf
  const double* rnorm;   // n: cosine sqrt(sum x^2) (f64, reference fold)
  const float* rstat;    // n x HM_NSTAT per-row bound statistics
  const float* smax;     // HM_NSTAT maxima over rows (per dataset)
  uint64_t* cand;        // candidate keys: n x nlist x cout
  uint64_t* ctau;        // list thresholds: n x nlist
  uint32_t* ccnt;        // keys stored per list (<= cout): n x nlist
  float* thr;            // sweep pass-0 thresholds: n x nsplit
  uint32_t* knn_idx;     // n x k (output)
  float* knn_dist;       // n x k (output)
  double* knn_dist64;    // n x k f64 distances before the f32 rounding, or null:
                         //   kept (and the calibration runs on them) in the
                         //   FP64 calibration mode when the kNN runs on the device
  float* err3;           // n x 3 per-row relative error of the f32 outputs against
                         //   their f64 values (distances, sigma, directed
                         //   weights: the largest of the row), with knn_dist64
  int64_t n;             // rows
  int64_t n_pad;         // rows padded to the column-block size
  int64_t row_off;       // first global row id
  int64_t edge_off;      // first global edge id (= sum of n*k before)
  int32_t d;             // features
  int32_t kp;            // padded feature stride (elements, multiple of 32)
  int32_t kp16;          // op16 / op16lo row stride (multiple of 64: 128-byte rows)
  int32_t k;             // neighbours
  int32_t metric;        // 0 euclidean, 1 cosine
  int32_t C;             // candidates targeted per (row, split)
  int32_t nsplit;        // column splits of the candidate sweep
  int32_t nlist;         // candidate lists per row = 2 * nsplit (one per
                         //   half of the 64-lane wave, see hm_gram.hip)
  int32_t cout;          // key slots per list
  int32_t C0;            // threshold pass: group maxima selected per (row, split)
  int32_t pstride;       // threshold pass: every pstride-th 128-column block
                         //   of a split is swept (a sample of its columns)
};

// Query rows per work item of the candidate sweep (hm_gram.hip): 256 in the
// bf16 and fp32 modes, 128 in bf16x3.
__host__ __device__ constexpr int sweep_rows(int mode) { return mode == 2 ? 128 : 256; }
// Threads per sweep workgroup: one wave per 32 query rows.
__host__ __device__ constexpr int sweep_threads(int mode) { return sweep_rows(mode) / 32 * 64; }
// Ring slots and workgroups per CU of the register-query sweep (a 3-slot
// ring at three workgroups per CU leaves 168 VGPRs and spills: measured
// slower on N = 20,000 x d = 128).
#ifndef HM_QREG_NSTAGE
#define HM_QREG_NSTAGE 4
#define HM_QREG_PER 2
#endif
// bf16 operand rows (kp16 * 2 bytes) up to this length are swept with the
// item's query rows resident in LDS (gram_sweep_qres_kernel).
constexpr int kQresRowBytes = 384;
// ... and up to this length with the queries in registers instead
// (gram_sweep_qreg_kernel, two 128-query workgroups per CU).
constexpr int kQresPairRowBytes = 256;
// Neighbour counts: the MFMA candidate sweep serves k <= kSweepMaxK (more
// neighbours than its 128 selection groups per row can certify are left to
// the exact scan); the exact scan and its merges serve k <= kMaxK.
constexpr int kSweepMaxK = 64;
constexpr int kMaxK = 1024;
// Bucket slots a workgroup of the symmetrisation's write kernel stages in
// LDS at a time.
constexpr int kSymChunk = 1024;
// Queries per wave of the exact-scan shape for a job whose largest k is k
// (exact_scan_kernel<QPW, QCAP>, QCAP >= k + 64; items hold 4 * QPW queries).
__host__ __device__ constexpr int xs_qpw_for_k(int k) { return k <= 64 ? 8 : (k <= 448 ? 4 : 1); }
// Calibration: a thread per row (rather than a lane group per row) for jobs
// of at least this many rows per CU (four waves per SIMD).
constexpr int kCalibThreadRowsPerCu = 4 * 4 * 64;

// NW waves (32 queries each), NSTAGE ring slots, rows of up to QMAXB bytes.
template <int NW_, int NSTAGE_, int QMAXB_>
struct QresCfgT {
  static constexpr int NW = NW_, QW = 32 * NW_;
  static constexpr int BIAS = 128 * 128;             // after the candidate rows
  static constexpr int STAGE = BIAS + NW / 2 * 512;  // + NW/2 copies of the biases
  static constexpr int NSTAGE = NSTAGE_;
  static constexpr int NI = 16 / NW;                 // candidate copy groups per wave
  static constexpr int VM = NI + 1;                  // vmcnt entries per stage and wave
  static constexpr int BIAS0 = NSTAGE * STAGE;       // the item's first block's biases
  static constexpr int QOFF = BIAS0 + 512;           // slice c of query row r: QOFF + c*QW*128 + r*128
  static constexpr int QMAXB = QMAXB_;               // row bytes held
  static constexpr int LDS = QOFF + QMAXB / 128 * QW * 128;
  static_assert(QW / 8 == 4 * NW, "four query copy groups per wave and slice");
  static_assert(QOFF % 128 == 0, "query rows are 128-byte aligned");
  static_assert(NSTAGE >= 2 && NSTAGE <= 3, "waits cover at most one later stage");
};
// Query-resident bf16 sweep configurations (gram_sweep_qres_kernel).
// One workgroup per CU: 256 queries, 3-slot ring, rows up to 384 bytes.
using QresWide = QresCfgT<8, 3, kQresRowBytes>;
static_assert(QresWide::LDS <= 160 * 1024, "LDS");

// bf16 sweep with the query operand in registers (gram_sweep_qreg_kernel),
// for rows of at most 256 bytes: PER 128-query workgroups per CU, each with
// a ring of NSTAGE candidate stages (NSTAGE - 1 in flight while one is
// computed); the item's query rows pass through ring slots 0 and 1 on their
// way to the registers.
template <int NSTAGE_, int PER_>
struct QregCfgT {
  static constexpr int NW = 4, QW = 32 * NW;
  static constexpr int BIAS = 128 * 128;             // after the candidate rows
  static constexpr int STAGE = BIAS + NW / 2 * 512;  // + NW/2 copies of the biases
  static constexpr int NSTAGE = NSTAGE_;
  static constexpr int PER = PER_;                   // workgroups per CU
  static constexpr int NI = 16 / NW;                 // candidate copy groups per wave
  static constexpr int VM = NI + 1;                  // vmcnt entries per stage and wave
  static constexpr int BIAS0 = NSTAGE * STAGE;       // the item's first block's biases
  static constexpr int LDS = BIAS0 + 512;
  static constexpr int QMAXB = kQresPairRowBytes;    // row bytes held
  static_assert(NSTAGE >= 3 && NSTAGE <= 4, "waits cover at most two later stages");
  static_assert(QMAXB / 128 * QW * 128 <= 2 * STAGE, "query rows fit ring slots 0 and 1");
  static_assert(QW / 8 == 4 * NW, "four query copy groups per wave and slice");
  static_assert(PER * LDS <= 160 * 1024, "PER workgroups per CU");
};
using QregCfg = QregCfgT<HM_QREG_NSTAGE, HM_QREG_PER>;
static_assert(QresWide::NW * 64 == sweep_threads(0), "the runtime launches sweep_threads(0)");


// Number of float statistics kept per row for the candidate error bound.
#define HM_NSTAT 8
// Stat slots.
enum : int {
  ST_NZ = 0,   // f32 ||x||^2 as used in keys (euclid), 0 for cosine
  ST_XN = 1,   // ||xhat|| of the bf16 (or f32) representation
  ST_EN = 2,   // ||x - xhat||  (representation error, rounded up)
  ST_XO = 3,   // ||x|| of the (normalised) f32 operand
  ST_HN = 4,   // bf16x3: ||hi||
  ST_LN = 5,   // bf16x3: ||lo||
  ST_HLN = 6,  // bf16x3: ||hi + lo||
  ST_RN = 7,   // bf16x3: ||x - hi - lo||
};

// Work item of the candidate sweep / exact scan: a block of query rows of
// one dataset against one column split.
struct WorkItem {
  int32_t ds;      // dataset
  int32_t split;   // column split
  int32_t qbeg;    // first query (row id, or index into a row list)
  int32_t qcnt;    // number of queries in the block
};

// The exact scan's work list, built by the last workgroup of the escalation
// re-rank (rerank_kernel<true>): the escalated rows still flagged, compacted
// in place in list / count, and their work items (items, nitems, xs_out: the
// column splits); done counts the finished workgroups (zero between runs).
struct FbTail {
  uint32_t* list;
  int* count;
  int qb, xs, xs_cap;
  int64_t item_cap, row_cap;
  int target;
  WorkItem* items;
  int* nitems;
  int* xs_out;
  uint32_t* done;
};

// Precision / mode codes (mirrored in the public header).
enum : int { PREC_EXACT = 0, PREC_FP32 = 1, PREC_BF16 = 2, PREC_BF16X3 = 3 };
enum : int { CALIB_REF_F32 = 0, CALIB_F64 = 1 };

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t f2u(float x) { return __float_as_uint(x); }
__device__ __forceinline__ float u2f(uint32_t u) { return __uint_as_float(u); }

// Relative error of an f32 value against the f64 value it rounds (0 for 0).
__device__ __forceinline__ double rel_err(float f, double v) {
  return v != 0.0 ? fabs((double)f - v) / fabs(v) : 0.0;
}

// The calibration's global-mean floor for one dataset (src/smooth_knn.rs:
// 70-75), by one 256-thread workgroup: sigma = max(sigma, 1e-3 * mean of all
// kNN distances) for the rows with rho == 0, the mean being the reference's
// left-to-right f64 fold by one lane, computed only if the dataset flagged
// such a row (zany, cleared here for the next run). The rows' directed
// weights do not depend on it (their distances all equal rho = 0, so every
// strength is exp(0) = 1), which lets it run after the symmetrisation.
__device__ __forceinline__ void calib_fixup_ds(const DsDesc& D, int dsi, int mode,
                                               float* __restrict__ sigma,
                                               double* __restrict__ sigma64,
                                               const float* __restrict__ rho,
                                               uint32_t* __restrict__ zany) {
  if (!zany[dsi]) return;  // no row of this dataset needs the floor (uniform)
  __shared__ double mean_s;
  if (threadIdx.x == 0) {
    const int64_t m = D.n * D.k;
    double s = 0.0;
    if (D.knn_dist64)
      for (int64_t t = 0; t < m; ++t) s = s + D.knn_dist64[t];
    else
      for (int64_t t = 0; t < m; ++t) s = s + (double)D.knn_dist[t];
    mean_s = s / (double)m;
  }
  __syncthreads();
  const double mean_all = mean_s;
  for (int64_t i = threadIdx.x; i < D.n; i += 256) {
    const int64_t g = D.row_off + i;
    if (rho[g] > 0.0f) continue;  // rows with a non-zero distance have their floor
    if (mode == CALIB_REF_F32) {
      const float fl = 1e-3f * (float)mean_all;
      float s = sigma[g];
      if (s < fl) s = fl;
      sigma[g] = s;
      if (sigma64) sigma64[g] = (double)s;
    } else {
      double s = sigma64 ? sigma64[g] : (double)sigma[g];
      const double fl = 1e-3 * mean_all;
      if (s < fl) s = fl;
      if (sigma64) sigma64[g] = s;
      sigma[g] = (float)s;
      if (D.err3) D.err3[i * 3 + 1] = (float)rel_err((float)s, s);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) zany[dsi] = 0u;  // cleared for the next run
}

// Order-preserving map of a float to uint32 (total order, -0 < +0).
__device__ __forceinline__ uint32_t ordered_u32(float x) {
  uint32_t u = __float_as_uint(x);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unordered_f32(uint32_t o) {
  uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
  return __uint_as_float(u);
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
  uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
  lo = __shfl_xor(lo, m, 64);
  hi = __shfl_xor(hi, m, 64);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int lane) {
  uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)v, lane);
  uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), lane);
  return ((uint64_t)hi << 32) | lo;
}

// Ascending bitonic sort of one u64 per lane across the 64-lane wave.
__device__ __forceinline__ uint64_t wave_sort64(uint64_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      uint64_t o = shfl_xor_u64(v, j);
      bool up = (lane & k) == 0;
      bool lower = (lane & j) == 0;
      uint64_t mn = v < o ? v : o, mx = v < o ? o : v;
      v = (lower == up) ? mn : mx;
    }
  }
  return v;
}

// Ascending bitonic sort of `len` (power of two, 64..1024) u64 keys held in
// LDS by one wave. Caller provides lane-synchronous access (single wave).
__device__ __forceinline__ void wave_sort_lds(uint64_t* a, int len) {
  const int lane = threadIdx.x & 63;
  for (int k = 2; k <= len; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = lane; i < len; i += 64) {
        int p = i ^ j;
        if (p > i) {
          uint64_t x = a[i], y = a[p];
          bool up = (i & k) == 0;
          if ((x > y) == up) { a[i] = y; a[p] = x; }
        }
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
  }
}

}  // namespace hm
