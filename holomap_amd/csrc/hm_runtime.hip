// Host runtime of the holomap kNN-graph library: contexts, job planning and
// allocation, the device-driven pipeline (no host synchronisation between
// stages, so a job can be captured into a HIP graph and replayed), and the
// C ABI declared in include/holomap_gpu.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/holomap_gpu.h"
#include "hm_internal.hpp"

namespace hm {
// kernels (defined in the other translation units)
__global__ void prep_kernel(const DsDesc*, int, int64_t, const int64_t*, uint32_t*, int);
template <int MODE, int PASS, int CW>
__global__ void gram_sweep_kernel(const DsDesc*, const WorkItem*, const int*, int, const uint32_t*,
                                  int);
template <class Cfg, int PASS>
__global__ void gram_sweep_qres_kernel(const DsDesc*, const WorkItem*, const int*, int,
                                       const uint32_t*, int);
template <int PASS>
__global__ void gram_sweep_qreg_kernel(const DsDesc*, const WorkItem*, const int*, int,
                                       const uint32_t*, int);
template <int QPW, int QCAP>
__global__ void exact_scan_kernel(const DsDesc*, const WorkItem*, const int*, int,
                                  const uint32_t*, const int*, int, int, uint64_t*);
__global__ void exact_merge_kernel(const DsDesc*, int, const uint32_t*, const int*, int64_t,
                                   const int*, int, int, int, const uint64_t*, uint32_t*, int);
template <bool TAIL>
__global__ void rerank_kernel(const DsDesc*, int, int64_t, int, uint32_t*, const uint32_t*,
                              const int*, uint32_t*, FbTail);
__global__ void sweep_items_kernel(const DsDesc*, int, const uint32_t*, const int*, int, int,
                                   double, WorkItem*, int*, int*);
__global__ void fallback_items_kernel(const DsDesc*, int, const uint32_t*, const int*, int, int,
                                      int, int64_t, int64_t, int, WorkItem*, int*, int*);
__global__ void flag_sweep_items_kernel(const uint32_t*, int64_t, uint32_t*, int*, const DsDesc*,
                                        int, int, int, double, WorkItem*, int*, int*,
                                        const uint32_t*);
__global__ void flag_fallback_items_kernel(const uint32_t*, int64_t, uint32_t*, int*,
                                           const DsDesc*, int, int, int, int, int64_t, int64_t,
                                           int, WorkItem*, int*, int*, const uint32_t*);
template <int KM>
__global__ void smooth_knn_kernel(const DsDesc*, int, int64_t, const double*, int, float*, float*,
                                  double*, uint32_t*, uint32_t, uint32_t*, uint32_t*, float*,
                                  uint32_t*);
__global__ void calib_fixup_kernel(const DsDesc*, int, float*, double*, const float*, uint32_t*);
template <int G>
__global__ void smooth_knn_group_kernel(const DsDesc*, int, int64_t, const double*, int, float*,
                                        float*, double*, uint32_t*, uint32_t, uint32_t*, uint32_t*,
                                        float*, uint32_t*);
__global__ void directed_kernel(const DsDesc*, int, int64_t, const float*, const float*, uint32_t,
                                uint32_t*, uint32_t*, float*, uint32_t*);
__global__ void expf_kernel(const float*, float*, int64_t);
template <typename T>
__global__ void scan_tile_sums(const T*, int64_t, T*);
template <typename T, int NT, int MS, bool LIST>
__global__ void scan_small(const T*, int64_t, T*, T*, uint32_t*, int*);
template <typename T>
__global__ void scan_sums(T*, int64_t, T*);
template <typename T>
__global__ void scan_tiles(const T*, int64_t, const T*, T*);
__global__ void lower_bounds(const uint32_t*, int64_t, int64_t, uint32_t*);
__global__ void compact_kernel(const uint32_t*, const uint32_t*, int64_t, uint32_t*, int*);
__global__ void fill_u32(uint32_t*, int64_t, uint32_t);
__global__ void sum_u32_part_kernel(const uint32_t*, int64_t, uint64_t*);
__global__ void sum_u64_kernel(const uint64_t*, int, uint64_t*);
__global__ void nonfinite_part_kernel(const float*, int64_t, uint64_t*);
__global__ void cc_init_kernel(uint32_t*, int64_t);
__global__ void cc_step_kernel(const uint64_t*, const uint32_t*, int64_t, uint32_t*, uint32_t*);
__global__ void cc_jump_kernel(uint32_t*, int64_t);
__global__ void cc_root_kernel(const uint32_t*, int64_t, uint32_t*);
__global__ void cc_label_kernel(const uint32_t*, const uint32_t*, int64_t, uint32_t*);
__global__ void lap_degree_kernel(const uint64_t*, const float*, int64_t, double*, double*);
__global__ void lap_scale_kernel(const uint64_t*, const uint32_t*, const float*, const double*,
                                 int64_t, double*);
__global__ void lap_matvec_kernel(const uint64_t*, const uint32_t*, const double*, const float*,
                                  const double*, int64_t, const double*, double*);
__global__ void sched_max_part_kernel(const float*, int64_t, float*);
__global__ void sched_max_kernel(const float*, int, float*);
__global__ void sched_flag_kernel(const float*, int64_t, const float*, uint32_t, uint32_t*);
__global__ void sched_scatter_kernel(const uint32_t*, const uint32_t*, const float*, int64_t,
                                     const uint32_t*, const uint32_t*, const float*, uint32_t*,
                                     uint32_t*, double*);
__global__ void nonfinite_min_kernel(const uint64_t*, int, uint64_t*);
__global__ void l2_normalise_kernel(const float*, int64_t, int, float*);
template <int G>
__global__ void sym_scatter_kernel(const DsDesc*, int, int64_t, int64_t, const uint32_t*,
                                   const uint32_t*, const float*, const uint32_t*, uint32_t*,
                                   uint32_t*, float*, float*, uint64_t*, uint32_t*);
__global__ void sym_write_kernel(const DsDesc*, int, int64_t, int64_t, const uint32_t*,
                                 const float*, const float*, const uint32_t*, const uint32_t*,
                                 const uint32_t*, const float*, const uint64_t*, const int64_t*,
                                 uint64_t, uint32_t*, uint32_t*, float*, uint64_t*, uint32_t,
                                 uint32_t, uint32_t*, int, float*, double*, const float*,
                                 uint32_t*);
}  // namespace hm

using namespace hm;
static_assert(kMaxK == HM_MAX_NEIGHBORS, "the public limit is the exact scan's");

struct hm_ctx {
  int device = 0;
  int num_cus = 256;
  hipStream_t stream = nullptr;
  uint64_t budget = 0;
  uint64_t in_use = 0;
  uint8_t* scratch = nullptr;  // the device's copy of `scratch` (hm_knn_graph_noalloc)
  uint8_t* stage = nullptr;    // pinned host staging of downloads (stage_cap bytes)
  uint64_t stage_cap = 0;
  std::vector<hipStream_t> spare_streams;  // streams of destroyed jobs, for reuse
  std::string err;
};

// Device memory of the allocation-free entry point: every intermediate
// buffer of hm_knn_graph_noalloc is carved from this array. Each GPU holds
// its own copy (at its own address) from the library's load until the
// process ends. Host code never uses the array's name as an address: the
// device address comes from hipGetSymbolAddress, once per device
// (scratch_address), and only that address reaches kernels and copies.
static __device__ __align__(256) uint8_t scratch[HM_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
// The last 16 KiB hold the input validation's partial results; the rest is
// the job's.
constexpr uint64_t kScratchTail = 16384;
std::mutex g_scratch_mu[kMaxDevices];  // serialises the calls that share a device's array
uint8_t* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_init_mu;

// The scratch array's address on `device` (the current device), looked up
// once per device; nullptr with *err set if the lookup fails.
uint8_t* scratch_address(int device, hipError_t* err) {
  std::lock_guard<std::mutex> lock(g_scratch_init_mu);
  *err = hipSuccess;
  if (device < 0 || device >= kMaxDevices) return nullptr;
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    *err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (*err != hipSuccess) return nullptr;
    g_scratch_ptr[device] = (uint8_t*)p;
  }
  return g_scratch_ptr[device];
}
}  // namespace

namespace {

constexpr int kScanTile = 4096;
// Arrays scanned by one workgroup (scan_small: 512 threads x kScanSmallQuads
// quads of 4 elements held per lane): 32 quads for 32-bit values, 16 for
// 64-bit ones. The plain scans (scan_excl, scan_u32_excl) run the same
// capacity on 1024 threads with half the quads per lane (half the serial
// steps); the flag compactions keep 512 threads, whose item building
// follows in the same workgroup.
constexpr int kScanSmallThreads = 512;
template <typename T>
constexpr int kScanSmallQuads = sizeof(T) == 4 ? 32 : 16;
template <typename T>
constexpr int64_t kScanSmall = (int64_t)kScanSmallThreads * 4 * kScanSmallQuads<T>;
static_assert(kScanSmallThreads == 512 && kScanSmallQuads<uint32_t> == 32,
              "flag_*_items_kernel compact with scan_small_body<uint32_t, 512, 32>");
constexpr int kScanWideThreads = 1024;
template <typename T>
constexpr int kScanWideQuads = kScanSmallQuads<T> / 2;
static_assert((int64_t)kScanWideThreads * 4 * kScanWideQuads<uint32_t> == kScanSmall<uint32_t> &&
                  (int64_t)kScanWideThreads * 4 * kScanWideQuads<uint64_t> == kScanSmall<uint64_t>,
              "the wide scans hold the same arrays");

struct DsHost {
  int64_t n, n_pad, row_off, edge_off;
  int d, kp, kp16, k, metric, C, nsplit, nlist, cout, C0, pstride;
};

int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }
int bitlen(uint64_t v) {
  int b = 0;
  while (v) { ++b; v >>= 1; }
  return b;
}

}  // namespace

struct hm_job {
  hm_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  hm_options opt{};
  uint32_t stages = HM_STAGE_ALL;
  int nds = 0;
  std::vector<DsHost> dsh;
  std::vector<DsDesc> hdesc;
  int64_t total_rows = 0, total_pad = 0, total_edges = 0;
  int kmax = 0, mode = 0, xs = 1, fb_cap = 0, n_items_l1 = 0;
  // items_l1 holds three consecutive lists: queries in registers with two
  // workgroups per CU, query-resident with one, streaming (kSweepLists)
  int n_items_list[3] = {0, 0, 0};
  int esc_mode = -1;  // sweep mode of the escalation pass (-1: none)
  bool xcd_perm = true;  // permute sweep items for XCD-local L2 sharing
  bool tuned = false;    // autotuning done (hm_options.autotune)
  float tune_ms = 0.f;   // best pipeline time the autotuner measured
  bool counted = false;  // bytes added to the context's in_use
  // allocation-free jobs (hm_knn_graph_noalloc): buffers are carved from the
  // device's scratch array and the stream is the context's
  uint8_t* arena = nullptr;
  uint64_t arena_cap = 0;
  bool own_stream = true;
  std::vector<WorkItem> items_l1_h;
  std::vector<void*> allocs;
  uint64_t bytes = 0;
  // device buffers
  DsDesc* d_desc = nullptr;
  float *x = nullptr, *xr = nullptr, *op32c = nullptr, *colbias = nullptr, *rstat = nullptr,
        *thr = nullptr,
        *smax = nullptr, *knn_dist = nullptr, *sigma = nullptr, *rho = nullptr,
        *wval = nullptr, *vals = nullptr, *in_w = nullptr;
  uint16_t *op16 = nullptr, *op16lo = nullptr;
  double *rnorm = nullptr, *sigma64 = nullptr, *target = nullptr, *dist64 = nullptr;
  float* err3 = nullptr;  // FP64 chain: per-row rounding errors (n x 3)
  uint64_t *stat64 = nullptr;  // partial sums of hm_job_counters
  uint64_t *cand = nullptr, *ctau = nullptr, *xpart = nullptr, *cntu = nullptr,
           *indptr_g = nullptr, *total_nnz = nullptr, *indptr_l = nullptr, *sc64 = nullptr;
  uint32_t *ccnt = nullptr;
  uint32_t *knn_idx = nullptr, *flag = nullptr, *flagpos = nullptr, *list = nullptr,
           *zany = nullptr, *key1 = nullptr, *erank = nullptr, *k1s = nullptr, *p1s = nullptr,
           *tmpv = nullptr, *inptr = nullptr, *rows = nullptr, *cols = nullptr,
           *sc32 = nullptr;
  uint32_t* nflag = nullptr;  // rows the primary re-rank flagged (the word after smax)
  uint32_t* serr = nullptr;   // set by a symmetrisation kernel that met inconsistent counts
  int *count = nullptr, *nitems_fb = nullptr, *xs_fb = nullptr;
  WorkItem *items_l1 = nullptr, *items_fb = nullptr, *items_l2 = nullptr;
  int* nitems_l2 = nullptr;
  int64_t cap_l2 = 0;  // items of a sweep over any subset of rows
  int64_t *pad_off = nullptr, *iptr_off = nullptr;
  std::vector<int64_t> iptr_off_h;
  // Laplacian of the resident graphs (hm_job_laplacian_matvec): 1/sqrt of
  // the degrees, all datasets, recomputed after each run
  double* lap_inv = nullptr;
  bool lap_valid = false;
  // graph and timing
  hipGraphExec_t gexec = nullptr;
  bool captured = false;
  hipEvent_t ev[HM_NTIMES + 1] = {};
  float times[HM_NTIMES] = {};
  bool timed = false;
};

namespace {

hm_status fail(hm_ctx* ctx, hm_status st, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return st;
}

#define HM_HIP(ctx, call)                                                              \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail((ctx), HM_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define HM_LAUNCH(ctx)                                                                    \
  do {                                                                                    \
    hipError_t e_ = hipGetLastError();                                                    \
    if (e_ != hipSuccess)                                                                 \
      return fail((ctx), HM_EDEVICE, std::string("kernel launch: ") + hipGetErrorString(e_)); \
  } while (0)


// Escalation sweep of uncertified rows: bf16 -> bf16x3, bf16x3 -> FP32.
int esc_mode_of(int mode) { return mode == 0 ? 2 : (mode == 2 ? 1 : -1); }

int gram_mode(int precision) {
  switch (precision) {
    case HM_PREC_FP32: return 1;
    case HM_PREC_BF16: return 0;
    case HM_PREC_BF16X3: return 2;
    default: return -1;
  }
}

int choose_nsplit(int64_t n, int qw, int C, int cus, int64_t other_items) {
  const int64_t rb = cdiv(n, qw);
  int best = 1;
  double bestc = 1e300;
  const int smax = std::max(1, std::min<int>(8, std::min<int>(512 / std::max(C, 1), (int)(n / 256))));
  for (int s = 1; s <= smax; ++s) {
    const double rounds = (double)cdiv(rb * s + other_items, cus);
    const double c = rounds / s + 0.02 * s;
    if (c < bestc - 1e-12) { bestc = c; best = s; }
  }
  return best;
}

// Splits merged per 1024-key sort of the exact scan's partial lists.
int merge_group(int kmax) {
  int kpow = 1;
  while (kpow < kmax) kpow <<= 1;
  return std::max(1, 1024 / kpow);
}

struct Plan {
  std::vector<DsHost> dsh;
  int64_t total_rows = 0, total_pad = 0, total_edges = 0, sum_xr = 0, sum_x16 = 0, sum_cos = 0,
          sum_cand = 0, fb_cap = 0;
  int kmax = 0, cmax = 0, xs = 1;
};

// Sweep configuration of every dataset: candidates C = k + margin (capped at
// 64), column splits and list capacity. margin < 0: 2k; nsplit <= 0: chosen
// so that the launch has enough work items for every CU. Buffers are sized
// for any configuration (kSplitCap splits, kCandCap key slots per row), so
// the autotuner can change it after allocation.
// The primary bf16 collect pass over rows longer than the query-resident
// limit sweeps 256-column blocks (twice the MFMAs per stage and barrier,
// half the query-slice re-reads); its last block may read up to 128
// operand rows past a split's end (masked by their -inf biases), which the
// operand's allocation covers.
constexpr bool kWideCollect = true;
constexpr int kWideSlackRows = 128;

constexpr int kSplitCap = 5;
constexpr int kCandCap = 512;  // the re-rank sorts at most 512 keys per row
constexpr int kSampleStride = 4;  // threshold pass: at most every 4th block
constexpr int kSampleBlocks = 4;  // ... and at least this many blocks per split
void plan_sweep(std::vector<DsHost>& dsh, int margin_opt, int nsplit_opt, int mode, int cus,
                int margin_mult = 2) {
  const int qw = sweep_rows(mode);
  int64_t rb_total = 0;
  for (auto& h : dsh)
    if (h.k <= kSweepMaxK) rb_total += cdiv(h.n, qw);
  for (auto& h : dsh) {
    if (h.k > kSweepMaxK) {
      // no sweep items: the re-rank finds no candidates and flags every
      // row, which the exact scan then computes
      h.C = h.k;
      h.nsplit = 0;
      h.nlist = 0;
      h.cout = 0;
      h.C0 = h.k;
      h.pstride = 1;
      continue;
    }
    const int margin = margin_opt < 0 ? margin_mult * h.k : margin_opt;
    h.C = std::max(h.k, std::min(h.k + margin, 64));
    // A list (one lane: half of a split's column groups) collects the keys
    // <= T, T = C-th smallest of the row's 128 group minima: about
    // 64 * -ln(1 - C/128) keys. Capacity with 40 % headroom; the re-rank
    // holds <= kCandCap keys per row, which bounds the split count.
    const double expect = 64.0 * -std::log(1.0 - h.C / 128.0);
    const int cneed = std::max(48, (int)((1.4 * expect + 8.0 + 7.0) / 8.0) * 8);
    const int smax = std::max(1, std::min(kSplitCap, kCandCap / 2 / cneed));
    if (nsplit_opt > 0) {
      h.nsplit = std::max(1, std::min<int>(nsplit_opt, smax));
    } else if (rb_total >= 2 * cus) {
      h.nsplit = 1;
    } else {
      h.nsplit = std::min(smax, choose_nsplit(h.n, qw, h.C, cus, rb_total - cdiv(h.n, qw)));
    }
    h.nlist = 2 * h.nsplit;
    // the lists share the re-rank's key slots: fewer splits, longer lists
    h.cout = std::max(cneed, std::min(256, (kCandCap / h.nlist) & ~7));
    // The threshold pass sweeps a sample of each split: every pstride-th
    // 128-column block (at least kSampleBlocks of them), and selects the
    // C0-th of the 128 group maxima, C0 ~ 4.6 C / (pstride * max(nsplit, 2))
    // (at most C). The row's threshold, the least over its splits, then
    // lists ~120-160 keys per row on the benchmark data, with the 0.1 %
    // tail above 50 (simulated on the N = 20,000 and N = 50,000 workloads
    // with this group structure), at 1 / pstride of the threshold pass's
    // blocks. The threshold only steers how many keys are listed:
    // correctness rests on the collect pass and the certification.
    const int64_t span = cdiv(h.n, (int64_t)h.nsplit * 128) * 128;
    const int64_t nb = span / 128;
    h.pstride = (int)std::max<int64_t>(1, std::min<int64_t>(kSampleStride, nb / kSampleBlocks));
    h.C0 = h.pstride == 1
               ? h.C
               : std::min(h.C, std::max(4, (int)std::ceil(4.6 * h.C /
                                                          (h.pstride * std::max(h.nsplit, 2)))));
  }
}

Plan make_plan(uint32_t nds, const uint64_t* n, const uint32_t* d, const uint32_t* k,
               const int32_t* metric, const hm_options& opt, int cus) {
  Plan P;
  P.dsh.resize(nds);
  int64_t nmax = 0;
  for (uint32_t i = 0; i < nds; ++i) {
    DsHost& h = P.dsh[i];
    h.n = (int64_t)n[i];
    h.d = (int)d[i];
    h.k = (int)k[i];
    h.metric = metric ? metric[i] : 0;
    h.kp = (int)cdiv(h.d, 32) * 32;
    h.kp16 = (int)cdiv(h.d, 64) * 64;
    h.n_pad = cdiv(h.n, 128) * 128;
    h.row_off = P.total_rows;
    h.edge_off = P.total_edges;
    P.total_rows += h.n;
    P.total_pad += h.n_pad;
    P.total_edges += h.n * h.k;
    P.sum_xr += h.n_pad * h.kp;
    P.sum_x16 += h.n_pad * h.kp16;
    if (h.metric == 1) P.sum_cos += h.n_pad * h.kp;
    P.kmax = std::max(P.kmax, h.k);
    nmax = std::max(nmax, h.n);
  }
  plan_sweep(P.dsh, opt.margin, opt.nsplit, gram_mode(opt.precision), cus);
  for (auto& h : P.dsh) P.cmax = std::max(P.cmax, h.C);
  P.sum_cand = P.total_rows * kCandCap;
  // exact-scan column splits: the two merge levels hold G * G splits of k
  // keys (G splits per 1024-key sort)
  P.xs = (int)std::max<int64_t>(1, std::min<int64_t>(8, nmax / 2048));
  P.xs = std::min(P.xs, merge_group(P.kmax) * merge_group(P.kmax));
  const int qb = 4 * xs_qpw_for_k(P.kmax);
  for (auto& h : P.dsh) P.fb_cap += cdiv(h.n, qb) * P.xs;
  return P;
}

uint64_t plan_bytes(const Plan& P, const hm_options& opt, uint32_t stages) {
  const int mode = gram_mode(opt.precision);
  const int64_t R = P.total_rows, E = P.total_edges, nds = (int64_t)P.dsh.size();
  uint64_t b = 0;
  int64_t sum_nd = 0;
  for (auto& h : P.dsh) sum_nd += h.n * h.d;
  b += sum_nd * 4;                                   // x
  if (stages & HM_STAGE_KNN) {
    b += P.sum_xr * 4 + P.sum_cos * 4;               // xr, op32 (cosine)
    if (mode == 0 || mode == 2) b += P.sum_x16 * 4;  // op16 + op16lo (bf16x3 or its escalation)
    int64_t kp16_max = 0;
    for (auto& h : P.dsh) kp16_max = std::max<int64_t>(kp16_max, h.kp16);
    b += kWideSlackRows * kp16_max * 2;  // rows the wide collect pass may read past the end
    b += P.total_pad * 4 + R * HM_NSTAT * 4 + nds * HM_NSTAT * 4 + R * 8;
    b += P.sum_cand * 8 + R * 2 * kSplitCap * 12 + R * kSplitCap * 4;  // cand, ctau+ccnt, thr
    b += 520 * 8;                                                          // counter sums
    b += R * 12 + P.fb_cap * 16 + R * P.xs * P.kmax * 8;
  }
  b += E * 8;                                        // knn idx + dist
  b += R * 20 + nds * 12;                            // sigma, rho, sigma64; flags, targets
  if ((stages & HM_STAGE_KNN) && opt.calib == HM_CALIB_F64) b += E * 8 + R * 12;  // dist64, err3
  if (stages & HM_STAGE_FUZZY) {
    // key1 erank wval in_dst in_key in_w revw; inptr + indptr_g (its last
    // entry the total); union sizes; rows cols vals; indptr_l; iptr_off
    b += E * 4 * 7 + (R + 1) * 12 + R * 8 + 2 * E * 12 + (R + nds) * 8 + nds * 8 + 4;
  }
  b += 2 * (cdiv(std::max<int64_t>(R + 1, E + 1), kScanTile) + 2) * 8;
  return b + (1u << 20);
}

template <typename T>
hm_status dalloc(hm_job* J, T** p, int64_t count) {
  if (count <= 0) count = 1;
  const size_t bytes = (size_t)count * sizeof(T);
  if (J->arena) {
    const uint64_t off = (J->bytes + 255) & ~(uint64_t)255;
    if (off + bytes > J->arena_cap)
      return fail(J->ctx, HM_EBUDGET, "the job needs more than the " +
                                          std::to_string(J->arena_cap) + "-byte scratch array");
    J->bytes = off + bytes;
    *p = (T*)(J->arena + off);
    return HM_OK;
  }
  void* q = nullptr;
  // stream-ordered: creating and destroying jobs never synchronises the
  // device, so jobs on other streams keep running (the context's pool keeps
  // freed memory for reuse)
  hipError_t e = hipMallocAsync(&q, bytes, J->stream);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(J->ctx, HM_ENOMEM, "hipMallocAsync of " + std::to_string(bytes) + " bytes failed");
  }
  J->allocs.push_back(q);
  J->bytes += bytes;
  *p = (T*)q;
  return HM_OK;
}

#define HM_TRY(x)                 \
  do {                            \
    hm_status s_ = (x);           \
    if (s_ != HM_OK) return s_;   \
  } while (0)

// The symmetrisation's kernels check the bucket counts they are given (a
// slot or row id out of range is skipped and flagged, never written): a
// flag means the graph arrays are not the union, reported by every reader.
static hm_status sym_check(hm_job* J) {
  if (!(J->stages & HM_STAGE_FUZZY) || !J->serr) return HM_OK;
  uint32_t e = 0;
  HM_HIP(J->ctx, hipMemcpyAsync(&e, J->serr, 4, hipMemcpyDeviceToHost, J->stream));
  HM_HIP(J->ctx, hipStreamSynchronize(J->stream));
  if (e) return fail(J->ctx, HM_EDEVICE, "symmetrisation: inconsistent in-degree counts");
  return HM_OK;
}

// ---- device building blocks ----

unsigned grid_for(int64_t n, int per) { return (unsigned)std::max<int64_t>(1, cdiv(n, per)); }

template <typename T>
hm_status scan_excl(hm_job* J, const T* in, int64_t n, T* out, T* total, T* tmp) {
  if (n <= 0) return HM_OK;
  if (n <= kScanSmall<T>) {
    constexpr int NT = kScanWideThreads, MS = kScanWideQuads<T>;
    hipLaunchKernelGGL((scan_small<T, NT, MS, false>), dim3(1), dim3(NT), 0, J->stream, in, n, out, total,
                       (uint32_t*)nullptr, (int*)nullptr);
    HM_LAUNCH(J->ctx);
    return HM_OK;
  }
  const int64_t nt = cdiv(n, kScanTile);
  hipLaunchKernelGGL(scan_tile_sums<T>, dim3((unsigned)nt), dim3(256), 0, J->stream, in, n, tmp);
  hipLaunchKernelGGL(scan_sums<T>, dim3(1), dim3(1024), 0, J->stream, tmp, nt, total);
  hipLaunchKernelGGL(scan_tiles<T>, dim3((unsigned)nt), dim3(256), 0, J->stream, in, n,
                     (const T*)tmp, out);
  HM_LAUNCH(J->ctx);
  return HM_OK;
}

// list[0 .. count) = indices i < n with flag[i] set, in order; count on the device.
hm_status compact_flags(hm_job* J, const uint32_t* flag, int64_t n, uint32_t* list, int* count) {
  if (n <= kScanSmall<uint32_t>) {
    hipLaunchKernelGGL((scan_small<uint32_t, kScanSmallThreads, kScanSmallQuads<uint32_t>, true>), dim3(1),
                       dim3(kScanSmallThreads), 0, J->stream, flag, n,
                       (uint32_t*)nullptr, (uint32_t*)nullptr, list, count);
  } else {
    HM_TRY(scan_excl<uint32_t>(J, flag, n, J->flagpos, (uint32_t*)nullptr, J->sc32));
    hipLaunchKernelGGL(compact_kernel, dim3(grid_for(n, 256)), dim3(256), 0, J->stream, flag,
                       (const uint32_t*)J->flagpos, n, list, count);
  }
  HM_LAUNCH(J->ctx);
  return HM_OK;
}

// Persistent grid of the candidate sweep: `per_cu` workgroups per CU, at
// most one per item.
int64_t sweep_grid(const hm_ctx* ctx, int mode, int64_t items, int per_cu = 1) {
  (void)mode;
  return std::max<int64_t>(1, std::min<int64_t>(items, (int64_t)ctx->num_cus * per_cu));
}

// The three lists of the primary sweep: bf16 rows of <= kQresPairRowBytes
// (two 128-query workgroups per CU, queries in registers), of <=
// kQresRowBytes (one 256-query workgroup, queries in LDS), and the rest
// (streaming).
int sweep_list_of(int mode, int kp16) {
  if (mode == 0 && kp16 * 2 <= kQresPairRowBytes) return 0;
  if (mode == 0 && kp16 * 2 <= kQresRowBytes) return 1;
  return 2;
}
constexpr int kListRows[3] = {128, 256, 0};  // query rows per item (0: sweep_rows(mode))
constexpr int kListPerCu[3] = {QregCfg::PER, 1, 1};  // workgroups per CU


// Passes of the candidate sweep in `mode` over `items` (a host count, or a
// device count when nitems_dev is set; qlist maps item rows to global row
// ids when set): passes 0 (threshold only), 1 (collect only) or 2 (both).
// sample: the threshold pass sweeps each dataset's sample of blocks (the
// primary sweep) rather than every block (the escalation's re-sweep, whose
// rows need the tightest threshold).
hm_status launch_sweep(hm_job* J, int mode, const WorkItem* items, const int* nitems_dev,
                       int nitems_host, const uint32_t* qlist, unsigned grid, int sample,
                       int passes) {
#define HM_GRAM(M, P, CW)                                                                    \
  hipLaunchKernelGGL((gram_sweep_kernel<M, P, CW>), dim3(grid), dim3(sweep_threads(M)), 0,     \
                     J->stream, J->d_desc, items, nitems_dev, nitems_host, qlist, sample)
  // the primary bf16 collect pass sweeps 256-column blocks (kWideCollect)
#define HM_GRAM2(M)                                               \
  do {                                                            \
    if (passes != 0 && M == 0 && sample && kWideCollect) {        \
      if (passes != 1) HM_GRAM(M, 0, 128);                        \
      HM_GRAM(M, 1, 256);                                         \
    } else {                                                      \
      if (passes != 1) HM_GRAM(M, 0, 128);                        \
      if (passes != 0) HM_GRAM(M, 1, 128);                        \
    }                                                             \
  } while (0)
  switch (mode) {
    case 0: HM_GRAM2(0); break;
    case 1: HM_GRAM2(1); break;
    case 2: HM_GRAM2(2); break;
    default: break;
  }
#undef HM_GRAM2
#undef HM_GRAM
  HM_LAUNCH(J->ctx);
  return HM_OK;
}

void mark(hm_job* J, int i);

// The primary sweep over its three item lists (datasets with short bf16
// rows keep their query rows in LDS, the rest stream them with the
// candidates): the threshold passes of every list, then the collect passes
// (a list's collect pass reads only its own thresholds). Grids match the
// item permutation of each list.
hm_status launch_sweep_l1(hm_job* J) {
  hm_ctx* ctx = J->ctx;
  const WorkItem* it0 = J->items_l1;
  const int n0 = J->n_items_list[0], n1 = J->n_items_list[1], n2 = J->n_items_list[2];
  for (int pass = 0; pass < 2; ++pass) {
    const WorkItem* it = it0;
#define HM_QRES(CFG, N, PER)                                                                   \
  do {                                                                                         \
    const unsigned grid = (unsigned)sweep_grid(ctx, 0, N, PER);                                \
    if (pass == 0)                                                                             \
      hipLaunchKernelGGL((gram_sweep_qres_kernel<CFG, 0>), dim3(grid), dim3(CFG::NW * 64), 0,    \
                         J->stream, J->d_desc, it, nullptr, N, nullptr, 1);                    \
    else                                                                                       \
      hipLaunchKernelGGL((gram_sweep_qres_kernel<CFG, 1>), dim3(grid), dim3(CFG::NW * 64), 0,    \
                         J->stream, J->d_desc, it, nullptr, N, nullptr, 1);                    \
    HM_LAUNCH(ctx);                                                                            \
  } while (0)
    if (n0 > 0) {
      const unsigned grid = (unsigned)sweep_grid(ctx, 0, n0, QregCfg::PER);
      if (pass == 0)
        hipLaunchKernelGGL(gram_sweep_qreg_kernel<0>, dim3(grid), dim3(QregCfg::NW * 64), 0,
                           J->stream, J->d_desc, it, nullptr, n0, nullptr, 1);
      else
        hipLaunchKernelGGL(gram_sweep_qreg_kernel<1>, dim3(grid), dim3(QregCfg::NW * 64), 0,
                           J->stream, J->d_desc, it, nullptr, n0, nullptr, 1);
      HM_LAUNCH(ctx);
    }
    it += n0;
    if (n1 > 0) HM_QRES(QresWide, n1, kListPerCu[1]);
    it += n1;
#undef HM_QRES
    if (n2 > 0)
      HM_TRY(launch_sweep(J, J->mode, it, nullptr, n2, nullptr,
                          (unsigned)sweep_grid(ctx, J->mode, n2), 1, pass));
    if (pass == 0) mark(J, 7);
  }
  return HM_OK;
}

void mark(hm_job* J, int i) {
  if (J->timed) (void)hipEventRecord(J->ev[i], J->stream);
}

hm_status enqueue_pipeline(hm_job* J) {
  hm_ctx* ctx = J->ctx;
  const int nds = J->nds;
  const int64_t R = J->total_rows, E = J->total_edges;
  const int cus = ctx->num_cus;
  mark(J, 0);
  if (J->stages & HM_STAGE_KNN) {
    // ---- prep ----
    // (also clears the in-degree counts of the symmetrisation to follow; the
    // statistics maxima and the flagged-row count are clear: zero-filled at
    // creation, cleared by the stage's last kernel)
    const int mode_bits = (J->mode == 0 ? 1 : 0) | 2;
    hipLaunchKernelGGL(prep_kernel, dim3(grid_for(J->total_pad, 16)), dim3(256), 0, J->stream,
                       J->d_desc, nds, J->total_pad, (const int64_t*)J->pad_off,
                       (J->stages & HM_STAGE_FUZZY) ? J->inptr : nullptr, mode_bits);
    HM_LAUNCH(ctx);
    mark(J, 1);
    // the exact scan's configuration (its work list may be built by the
    // escalation re-rank's last workgroup)
    const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>(J->fb_cap, 2 * cus));
    // merge groups: G splits of k keys fit one 1024-key sort; two merge
    // levels cover up to G * G splits
    const int G = merge_group(J->kmax);
    const int xs_cap = std::max(J->xs, std::min(512, G * G));
    const int qpw = xs_qpw_for_k(J->kmax), qb = 4 * qpw;
    // ---- candidate sweep + re-rank ----
    if (J->mode >= 0) {
      // grid and item order go together (see the item permutation in hm_job_create)
      HM_TRY(launch_sweep_l1(J));
      mark(J, 2);
      hipLaunchKernelGGL(rerank_kernel<false>,
                         dim3(std::min<int64_t>(grid_for(R, 4), 6 * (int64_t)cus)), dim3(256), 0,
                         J->stream, J->d_desc, nds, R, J->mode, J->flag, (const uint32_t*)nullptr,
                         (const int*)nullptr, J->nflag, FbTail{});
      HM_LAUNCH(ctx);
      mark(J, 3);
      if (J->esc_mode >= 0) {
        // escalation: rows the primary keys could not certify are swept
        // again with more precise keys (bf16 -> bf16x3, bf16x3 -> FP32; a far
        // smaller error bound) and re-ranked; only rows still uncertified
        // reach the exact scan below
        const int em = J->esc_mode;
        const double rate = em == 2 ? 1.1e12 : 0.45e12;
        if (R <= kScanSmall<uint32_t>) {  // compaction and items in one launch
          hipLaunchKernelGGL(flag_sweep_items_kernel, dim3(1), dim3(kScanSmallThreads), 0,
                             J->stream, (const uint32_t*)J->flag, R, J->list, J->count, J->d_desc,
                             nds, sweep_rows(em), cus, rate, J->items_l2, J->nitems_l2,
                             J->nitems_l2 + 1, (const uint32_t*)J->nflag);
        } else {
          HM_TRY(compact_flags(J, J->flag, R, J->list, J->count));
          hipLaunchKernelGGL(sweep_items_kernel, dim3(1), dim3(256), 0, J->stream, J->d_desc, nds,
                             (const uint32_t*)J->list, (const int*)J->count, sweep_rows(em), cus,
                             rate, J->items_l2, J->nitems_l2, J->nitems_l2 + 1);
        }
        HM_LAUNCH(ctx);
        HM_TRY(launch_sweep(J, em, J->items_l2, J->nitems_l2, 0, J->list, (unsigned)cus, 0, 2));
        // its last workgroup also compacts the list to the rows still
        // flagged and builds the exact scan's items
        const FbTail tail{J->list, J->count, qb, J->xs, xs_cap, (int64_t)J->fb_cap, R, (int)gx,
                          J->items_fb, J->nitems_fb, J->xs_fb, (uint32_t*)(J->nitems_l2 + 2)};
        hipLaunchKernelGGL(rerank_kernel<true>,
                           dim3(std::max<int64_t>(1, std::min<int64_t>(cdiv(R, 4), 4 * cus))),
                           dim3(256), 0, J->stream, J->d_desc, nds, R, em, J->flag,
                           (const uint32_t*)J->list, (const int*)(J->nitems_l2 + 1),
                           (uint32_t*)nullptr, tail);
        HM_LAUNCH(ctx);
      }
    } else {
      mark(J, 7);
      mark(J, 2);
      hipLaunchKernelGGL(fill_u32, dim3(grid_for(R, 256)), dim3(256), 0, J->stream, J->flag, R, 1u);
      HM_LAUNCH(ctx);
      mark(J, 3);
    }
    // ---- exact scan of the uncertified rows ----
    // (after an escalation its work list is built: the rows still flagged
    // are among the escalated ones)
    if (!(J->mode >= 0 && J->esc_mode >= 0)) {
      if (R <= kScanSmall<uint32_t>) {  // compaction and items in one launch
        hipLaunchKernelGGL(flag_fallback_items_kernel, dim3(1), dim3(kScanSmallThreads), 0,
                           J->stream, (const uint32_t*)J->flag, R, J->list, J->count, J->d_desc,
                           nds, qb, J->xs, xs_cap, (int64_t)J->fb_cap, R, (int)gx, J->items_fb,
                           J->nitems_fb, J->xs_fb,
                           J->mode >= 0 ? (const uint32_t*)J->nflag : (const uint32_t*)nullptr);
      } else {
        HM_TRY(compact_flags(J, J->flag, R, J->list, J->count));
        hipLaunchKernelGGL(fallback_items_kernel, dim3(1), dim3(256), 0, J->stream, J->d_desc,
                           nds, (const uint32_t*)J->list, (const int*)J->count, qb, J->xs, xs_cap,
                           (int64_t)J->fb_cap, R, (int)gx, J->items_fb, J->nitems_fb, J->xs_fb);
      }
    }
#define HM_XSCAN(QPW, QCAP)                                                                    \
  hipLaunchKernelGGL((exact_scan_kernel<QPW, QCAP>), dim3(gx), dim3(256), 0, J->stream,         \
                     J->d_desc, (const WorkItem*)J->items_fb, (const int*)J->nitems_fb, 0,     \
                     (const uint32_t*)J->list, (const int*)J->xs_fb, J->xs, J->kmax, J->xpart)
    if (qpw == 8) HM_XSCAN(8, 128);
    else if (qpw == 4) HM_XSCAN(4, 512);
    else HM_XSCAN(1, 2048);
#undef HM_XSCAN
    const unsigned gm = (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(R, 4), 8 * cus));
    hipLaunchKernelGGL(exact_merge_kernel, dim3(gm), dim3(256), 0, J->stream, J->d_desc, nds,
                       (const uint32_t*)J->list, (const int*)J->count, (int64_t)0,
                       (const int*)J->xs_fb, J->xs, J->kmax, G, (const uint64_t*)J->xpart,
                       (uint32_t*)J->smax, nds * HM_NSTAT + 1);
    HM_LAUNCH(ctx);
    mark(J, 4);
  } else {
    for (int i = 1; i <= 4; ++i) mark(J, i);
    mark(J, 7);
  }
  if (J->stages & HM_STAGE_CALIB) {
    // G lanes per row when the largest k fits a group, else a thread per row.
    // Jobs with enough rows for four waves per SIMD at a thread per row are
    // bound by VALU work rather than latency: a thread per row then, its
    // distances in registers (a lane evaluates every term of its row instead
    // of one).
    const int G = J->kmax <= 16 ? 16 : J->kmax <= 32 ? 32 : J->kmax <= 64 ? 64 : 0;
    const bool by_thread = G == 0 || R >= kCalibThreadRowsPerCu * (int64_t)cus;
    // with the symmetrisation to follow, the calibration also writes the
    // directed edges
    const bool edges = (J->stages & HM_STAGE_FUZZY) != 0;
    const uint32_t sent = (uint32_t)R;
    uint32_t* k1 = edges ? J->key1 : nullptr;
    if (edges && !(J->stages & HM_STAGE_KNN)) {  // else cleared by prep_kernel
      hipLaunchKernelGGL(fill_u32, dim3(grid_for(R, 256)), dim3(256), 0, J->stream, J->inptr, R, 0u);
      HM_LAUNCH(ctx);
    }
#define HM_CALIB_G(GS)                                                                         \
  hipLaunchKernelGGL((smooth_knn_group_kernel<GS>), dim3(grid_for(R * GS, 256)), dim3(256), 0, \
                     J->stream, J->d_desc, nds, R, (const double*)J->target, J->opt.calib,     \
                     J->sigma, J->rho, J->sigma64, J->zany, sent, k1, J->inptr, J->wval,       \
                     J->erank)
#define HM_CALIB_T(KM)                                                                         \
  hipLaunchKernelGGL((smooth_knn_kernel<KM>), dim3(grid_for(R, 256)), dim3(256), 0, J->stream, \
                     J->d_desc, nds, R, (const double*)J->target, J->opt.calib, J->sigma,       \
                     J->rho, J->sigma64, J->zany, sent, k1, J->inptr, J->wval, J->erank)
    if (by_thread) {
      if (G == 16) HM_CALIB_T(16);
      else if (G == 32) HM_CALIB_T(32);
      else if (G == 64) HM_CALIB_T(64);
      else HM_CALIB_T(0);
    } else if (G == 16) {
      HM_CALIB_G(16);
    } else if (G == 32) {
      HM_CALIB_G(32);
    } else {
      HM_CALIB_G(64);
    }
#undef HM_CALIB_G
#undef HM_CALIB_T
    // the global-mean floor (it leaves the directed weights unchanged) rides
    // on the symmetrisation's last kernel when one follows
    if (!(J->stages & HM_STAGE_FUZZY))
      hipLaunchKernelGGL(calib_fixup_kernel, dim3(nds), dim3(256), 0, J->stream, J->d_desc,
                         J->opt.calib, J->sigma, J->sigma64, (const float*)J->rho, J->zany);
    HM_LAUNCH(ctx);
  }
  mark(J, 5);
  if (J->stages & HM_STAGE_FUZZY) {
    const uint32_t sent = (uint32_t)R;
    if (!(J->stages & HM_STAGE_CALIB)) {  // else written by the calibration
      if (!(J->stages & HM_STAGE_KNN))  // else cleared by prep_kernel
        hipLaunchKernelGGL(fill_u32, dim3(grid_for(R, 256)), dim3(256), 0, J->stream, J->inptr, R,
                           0u);
      hipLaunchKernelGGL(directed_kernel, dim3(grid_for(E, 256)), dim3(256), 0, J->stream,
                         J->d_desc, nds, E, (const float*)J->sigma, (const float*)J->rho, sent,
                         J->key1, J->inptr, J->wval, J->erank);
      HM_LAUNCH(ctx);
    }
    // transpose: in-degree counts -> bucket offsets (in place; inptr[R] =
    // the kept edges), then every kept edge to its slot of its destination's
    // bucket, marked mutual or not, with the rows' union sizes (a group of
    // G >= k lanes per row); the sizes' scan; the entries and row pointers
    HM_TRY(scan_excl<uint32_t>(J, J->inptr, R, J->inptr, J->inptr + R, J->sc32));
    uint32_t* in_dst = J->k1s;
    uint32_t* in_key = J->p1s;
    float* in_w = J->in_w;
    float* revw = (float*)J->tmpv;
    const int GU = J->kmax <= 16 ? 16 : J->kmax <= 32 ? 32 : 64;
#define HM_SYM_SCATTER(GS)                                                                      \
  hipLaunchKernelGGL((sym_scatter_kernel<GS>), dim3(grid_for(R * GS, 256)), dim3(256), 0,        \
                     J->stream, J->d_desc, nds, R, E, (const uint32_t*)J->key1,                  \
                     (const uint32_t*)J->erank, (const float*)J->wval,                           \
                     (const uint32_t*)J->inptr, in_dst, in_key, in_w, revw, J->cntu, J->serr)
    if (GU == 16) HM_SYM_SCATTER(16); else if (GU == 32) HM_SYM_SCATTER(32); else HM_SYM_SCATTER(64);
#undef HM_SYM_SCATTER
    HM_LAUNCH(ctx);
    // (indptr_g[R] is the total: J->total_nnz points there)
    HM_TRY(scan_excl<uint64_t>(J, J->cntu, R, J->indptr_g, J->indptr_g + R, J->sc64));
    // workgroups past the entries' carry the calibration's global-mean floor
    const bool fix = (J->stages & HM_STAGE_CALIB) != 0;
    const uint32_t nrt = (uint32_t)cdiv(R, 256), nslot = (uint32_t)cdiv(E, 256);
    hipLaunchKernelGGL(sym_write_kernel, dim3(nrt + 2 * nslot + (fix ? (uint32_t)nds : 0u)),
                       dim3(256), 0, J->stream, J->d_desc, nds, R, E, (const uint32_t*)J->key1,
                       (const float*)J->wval, (const float*)revw, (const uint32_t*)J->inptr,
                       (const uint32_t*)in_dst, (const uint32_t*)in_key, (const float*)in_w,
                       (const uint64_t*)J->indptr_g, (const int64_t*)J->iptr_off,
                       (uint64_t)(2 * E), J->rows, J->cols, J->vals, J->indptr_l, nrt, nslot,
                       J->serr, J->opt.calib, J->sigma, J->sigma64, (const float*)J->rho,
                       fix ? J->zany : nullptr);
    HM_LAUNCH(ctx);
  }
  mark(J, 6);
  return HM_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

void hm_default_options(hm_options* o) {
  if (!o) return;
  o->precision = HM_PREC_BF16;
  o->calib = HM_CALIB_REF_F32;
  o->margin = -1;
  o->nsplit = 0;
  o->use_graph = 0;
  o->autotune = 0;
  o->timing = 0;
  o->reserved = 0;
}

const char* hm_version(void) { return "holomap_gpu 0.1 gfx950"; }

int hm_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

hm_status hm_ctx_create(hm_ctx** out, int device, uint64_t budget) {
  if (!out) return HM_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return HM_EDEVICE;
  if (device < 0 || device >= n) return HM_EINVAL;
  hm_ctx* c = new hm_ctx();
  c->device = device;
  c->budget = budget;
  if (hipSetDevice(device) != hipSuccess) { delete c; return HM_EDEVICE; }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return HM_EDEVICE;
  }
  {  // the allocation-free entry point's array on this device
    hipError_t e = hipSuccess;
    c->scratch = scratch_address(device, &e);
    if (!c->scratch) {
      (void)hipGetLastError();
      (void)hipStreamDestroy(c->stream);
      delete c;
      return HM_EDEVICE;
    }
  }
  // job buffers come from the device's default pool; keep freed blocks
  // cached instead of returning them at every synchronisation
  hipMemPool_t pool = nullptr;
  if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) {
    uint64_t keep = ~0ull;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
  }
  *out = c;
  return HM_OK;
}

void hm_ctx_destroy(hm_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  if (c->stage) (void)hipHostFree(c->stage);
  for (hipStream_t st : c->spare_streams) (void)hipStreamDestroy(st);
  delete c;
}

const char* hm_last_error(const hm_ctx* c) { return c ? c->err.c_str() : "no context"; }

uint64_t hm_ctx_budget(const hm_ctx* c) { return c ? c->budget : 0; }

uint64_t hm_ctx_in_use(const hm_ctx* c) { return c ? c->in_use : 0; }

hm_status hm_validate(const float* data, uint64_t len, uint32_t d, uint32_t k, uint64_t* bad) {
  if (d == 0 || len == 0 || len % d != 0) return HM_EBADSHAPE;
  if (k < 2) return HM_EINVAL;
  const uint64_t n = len / d;
  if (n <= k) return HM_ETOOFEW;
  if (data) {
    // blocks of 4096 values tested with a branch-free OR of the exponent
    // test (vectorised); only a block holding a non-finite value is searched
    for (uint64_t b = 0; b < len; b += 4096) {
      const uint64_t e = std::min<uint64_t>(len, b + 4096);
      uint32_t any = 0;
      for (uint64_t i = b; i < e; ++i) {
        uint32_t u;
        std::memcpy(&u, data + i, 4);
        any |= (u & 0x7f800000u) == 0x7f800000u ? 1u : 0u;
      }
      if (!any) continue;
      for (uint64_t i = b; i < e; ++i) {
        if (!std::isfinite(data[i])) {
          if (bad) *bad = i;
          return HM_ENONFINITE;
        }
      }
    }
  }
  return HM_OK;
}

uint64_t hm_job_estimate_bytes(uint32_t nds, const uint64_t* n, const uint32_t* d,
                               const uint32_t* k, const int32_t* metric, const hm_options* opt) {
  hm_options o;
  if (opt) o = *opt; else hm_default_options(&o);
  std::vector<int32_t> m;
  if (!metric) {
    m.assign(nds, HM_COSINE);  // the cosine operand makes it the larger estimate
    metric = m.data();
  }
  Plan P = make_plan(nds, n, d, k, metric, o, 256);
  return plan_bytes(P, o, HM_STAGE_ALL);
}

// Writes the sweep configuration of J->dsh (C, splits, lists) into the
// device descriptors and rebuilds the sweep's work items: (dataset, split,
// row block), longest first, permuted for the XCDs. Blocks until uploaded.
hm_status apply_sweep(hm_job* J) {
  hm_ctx* ctx = J->ctx;
  const int nds = J->nds;
  for (int i = 0; i < nds; ++i) {
    const DsHost& h = J->dsh[i];
    DsDesc& D = J->hdesc[i];
    D.C = h.C;
    D.nsplit = h.nsplit;
    D.nlist = h.nlist;
    D.cout = h.cout;
    D.C0 = h.C0;
    D.pstride = h.pstride;
  }
  if (hipMemcpyAsync(J->d_desc, J->hdesc.data(), sizeof(DsDesc) * nds, hipMemcpyHostToDevice,
                     J->stream) != hipSuccess)
    return fail(ctx, HM_EDEVICE, "descriptor upload failed");
  if ((J->stages & HM_STAGE_KNN) && J->mode >= 0) {
    // three item lists (sweep_list_of), each longest first
    std::vector<std::pair<int64_t, WorkItem>> it[3];
    for (int i = 0; i < nds; ++i) {
      const DsHost& h = J->dsh[i];
      if (h.nsplit == 0) continue;  // k beyond the sweep: exact scan only
      const int64_t cost = cdiv(h.n, h.nsplit) * h.kp;
      const int li = sweep_list_of(J->mode, h.kp16);
      const int qw = kListRows[li] ? kListRows[li] : sweep_rows(J->mode);
      for (int64_t rb = 0; rb < h.n; rb += qw)
        for (int s = 0; s < h.nsplit; ++s)
          it[li].push_back({cost, WorkItem{i, s, (int)rb, (int)std::min<int64_t>(qw, h.n - rb)}});
    }
    // The sweep grid is G persistent workgroups; workgroup b takes items b,
    // b + G, ... and runs on XCD b % X (round-robin dispatch). Item slots are
    // permuted so that in every round each XCD receives a contiguous run of
    // the (row block, split) order above: its workgroups then stream the same
    // few row blocks and column splits at the same time and share them in
    // that XCD's L2. Each list is one launch with its own grid.
    const int64_t X = (J->xcd_perm && ctx->num_cus % 8 == 0) ? 8 : 1;
    const int64_t ni = (int64_t)(it[0].size() + it[1].size() + it[2].size());
    J->items_l1_h.resize(ni);
    int64_t base = 0;
    for (int li = 0; li < 3; ++li) {
      auto& L = it[li];
      std::stable_sort(L.begin(), L.end(),
                       [](const auto& a, const auto& b) { return a.first > b.first; });
      const int64_t nl = (int64_t)L.size();
      const int64_t G = sweep_grid(ctx, J->mode, nl, kListPerCu[li]);
      for (int64_t r0 = 0; r0 < nl; r0 += G) {
        const int64_t rl = std::min(G, nl - r0);  // items in this round
        const bool perm = X > 1 && rl % X == 0;
        const int64_t per = rl / X;
        for (int64_t l = 0; l < rl; ++l) {
          const int64_t p = perm ? (l % per) * X + l / per : l;
          J->items_l1_h[base + r0 + p] = L[r0 + l].second;
        }
      }
      base += nl;
    }
    for (int li = 0; li < 3; ++li) J->n_items_list[li] = (int)it[li].size();
    J->n_items_l1 = (int)ni;
    if (hipMemcpyAsync(J->items_l1, J->items_l1_h.data(), sizeof(WorkItem) * ni,
                       hipMemcpyHostToDevice, J->stream) != hipSuccess)
      return fail(ctx, HM_EDEVICE, "work item upload failed");
  }
  if (hipStreamSynchronize(J->stream) != hipSuccess)
    return fail(ctx, HM_EDEVICE, "descriptor upload failed");
  return HM_OK;
}

// Job planning and allocation. arena == nullptr: buffers from the context's
// memory pool on a stream of the job's own (hm_job_create); otherwise every
// buffer is carved from `arena` (arena_cap bytes of device memory) and the
// job runs on the context's stream, with no device allocation, stream or
// event created (hm_knn_graph_noalloc).
static hm_status job_create(hm_ctx* ctx, uint32_t nds, const uint64_t* n, const uint32_t* d,
                            const uint32_t* k, const int32_t* metric, const hm_options* opt,
                            uint32_t stages, uint8_t* arena, uint64_t arena_cap, hm_job** out) {
  if (!ctx || !out || nds == 0 || !n || !d || !k) return fail(ctx, HM_EINVAL, "bad job arguments");
  *out = nullptr;
  hm_options o;
  if (opt) o = *opt; else hm_default_options(&o);
  const bool knn_stage = (stages & HM_STAGE_KNN) != 0;
  for (uint32_t i = 0; i < nds; ++i) {
    if (d[i] == 0) return fail(ctx, HM_EBADSHAPE, "n_features must be > 0");
    if (n[i] == 0 || k[i] == 0) return fail(ctx, HM_EBADSHAPE, "empty input");
    if (knn_stage) {
      // stage-level kNN (src/knn.rs:22) needs k <= n; fit_transform's
      // stricter n > k >= 2 is checked by hm_validate in hm_knn_graph
      if (n[i] < k[i]) return fail(ctx, HM_ETOOFEW, "need at least n_neighbors points");
      if (k[i] > (uint32_t)kMaxK)
        return fail(ctx, HM_EUNSUPPORTED,
                    "n_neighbors > " + std::to_string(kMaxK) + " is not supported on the GPU path");
    }
    if (n[i] >= (1ull << 31)) return fail(ctx, HM_EUNSUPPORTED, "n >= 2^31");
    if (metric && metric[i] != HM_EUCLIDEAN && metric[i] != HM_COSINE)
      return fail(ctx, HM_EINVAL, "metric must be \"euclidean\" or \"cosine\"");
  }
  if (o.precision < HM_PREC_EXACT || o.precision > HM_PREC_BF16X3)
    return fail(ctx, HM_EINVAL, "unknown precision");
  if (o.calib != HM_CALIB_REF_F32 && o.calib != HM_CALIB_F64)
    return fail(ctx, HM_EINVAL, "unknown calibration mode");
  if (o.nsplit < 0 || o.nsplit > kSplitCap)
    return fail(ctx, HM_EINVAL, "nsplit must be in 0.." + std::to_string(kSplitCap));
  HM_HIP(ctx, hipSetDevice(ctx->device));
  Plan P = make_plan(nds, n, d, k, metric, o, ctx->num_cus);
  if (P.total_rows >= 0xffffffffll || P.total_edges >= 0xffffffffll)
    return fail(ctx, HM_EUNSUPPORTED, "batch too large for 32-bit row/edge ids");
  const uint64_t need = plan_bytes(P, o, stages);
  if (arena) {
    if (need > arena_cap)
      return fail(ctx, HM_EBUDGET, "job needs " + std::to_string(need) + " bytes; the scratch "
                                   "array holds " + std::to_string(arena_cap));
  } else if (ctx->budget && ctx->in_use + need > ctx->budget) {
    return fail(ctx, HM_EBUDGET, "job needs " + std::to_string(need) + " bytes, budget left " +
                                     std::to_string(ctx->budget - ctx->in_use));
  }

  hm_job* J = new hm_job();
  J->ctx = ctx;
  if (arena) {
    J->arena = arena;
    J->arena_cap = arena_cap;
    J->own_stream = false;
    J->stream = ctx->stream;
  } else if (!ctx->spare_streams.empty()) {
    // every job has its own stream (a destroyed job's, when one is spare):
    // jobs of one context overlap on the device
    J->stream = ctx->spare_streams.back();
    ctx->spare_streams.pop_back();
  } else if (hipStreamCreateWithFlags(&J->stream, hipStreamNonBlocking) != hipSuccess) {
    delete J;
    return fail(ctx, HM_EDEVICE, "stream create failed");
  }
  J->opt = o;
  J->stages = stages;
  J->nds = (int)nds;
  J->dsh = P.dsh;
  J->total_rows = P.total_rows;
  J->total_pad = P.total_pad;
  J->total_edges = P.total_edges;
  J->kmax = P.kmax;
  J->mode = gram_mode(o.precision);
  J->esc_mode = esc_mode_of(J->mode);
  J->xs = P.xs;
  J->fb_cap = (int)P.fb_cap;
  auto bail = [&](hm_status s) {
    hm_job_destroy(J);
    return s;
  };
#define HM_A(ptr, cnt)                          \
  do {                                          \
    hm_status s_ = dalloc(J, &(ptr), (cnt));    \
    if (s_ != HM_OK) return bail(s_);           \
  } while (0)
  const int64_t R = P.total_rows, E = P.total_edges;
  int64_t sum_nd = 0, kp16_max = 0;
  for (auto& h : P.dsh) {
    sum_nd += h.n * h.d;
    kp16_max = std::max<int64_t>(kp16_max, h.kp16);
  }
  HM_A(J->d_desc, nds);
  HM_A(J->x, sum_nd);
  const bool knn = stages & HM_STAGE_KNN;
  if (knn) {
    HM_A(J->xr, P.sum_xr);
    if (P.sum_cos > 0) HM_A(J->op32c, P.sum_cos);
    if (J->mode == 0 || J->mode == 2) HM_A(J->op16, P.sum_x16 + kWideSlackRows * kp16_max);
    if (J->mode == 2 || J->esc_mode == 2) HM_A(J->op16lo, P.sum_x16);
    HM_A(J->colbias, P.total_pad);
    HM_A(J->rstat, R * HM_NSTAT);
    HM_A(J->smax, (int64_t)nds * HM_NSTAT + 1);  // + the flagged-row count
    J->nflag = (uint32_t*)(J->smax + (int64_t)nds * HM_NSTAT);
    // clear for the first run (each run's kNN stage clears it for the next)
    hipLaunchKernelGGL(fill_u32, dim3(grid_for((int64_t)nds * HM_NSTAT + 1, 256)), dim3(256), 0,
                       J->stream, (uint32_t*)J->smax, (int64_t)nds * HM_NSTAT + 1, 0u);
    HM_A(J->rnorm, R);
    // sized for any sweep configuration (see plan_sweep)
    HM_A(J->cand, std::max<int64_t>(P.sum_cand, 1));
    HM_A(J->ctau, R * 2 * kSplitCap);
    HM_A(J->ccnt, R * 2 * kSplitCap);
    HM_A(J->stat64, 520);
    HM_A(J->thr, R * kSplitCap);
    HM_A(J->flag, R);
    HM_A(J->flagpos, R);
    HM_A(J->list, R);
    HM_A(J->count, 1);
    HM_A(J->nitems_fb, 1);
    if (J->esc_mode >= 0) {
      int64_t cap = 0;
      for (auto& h : P.dsh) cap += cdiv(h.n, sweep_rows(J->esc_mode)) * kSplitCap;
      J->cap_l2 = cap;
      HM_A(J->items_l2, std::max<int64_t>(cap, 1));
      HM_A(J->nitems_l2, 3);  // items, rows escalated, the escalation re-rank's finished workgroups
      hipLaunchKernelGGL(fill_u32, dim3(1), dim3(256), 0, J->stream, (uint32_t*)J->nitems_l2,
                         (int64_t)3, 0u);
    }
    HM_A(J->xs_fb, 1);
    HM_A(J->items_fb, std::max<int64_t>(P.fb_cap, 1));
    HM_A(J->xpart, R * P.xs * P.kmax);
    HM_A(J->pad_off, nds + 1);
  }
  HM_A(J->knn_idx, E);
  HM_A(J->knn_dist, E);
  HM_A(J->sigma, R);
  HM_A(J->rho, R);
  HM_A(J->sigma64, R);
  // the FP64 calibration runs on the kNN's f64 distances when the kNN runs here
  if (knn && o.calib == HM_CALIB_F64) {
    HM_A(J->dist64, E);
    HM_A(J->err3, R * 3);
  }
  HM_A(J->zany, nds);  // per-dataset flags, kept clear between runs by calib_fixup_ds
  // (a fill kernel: the scratch array keeps the last call's contents)
  if (J->op16) {  // zero bf16 operand rows past the last dataset (read, never used)
    const int64_t words = kWideSlackRows * kp16_max / 2;
    hipLaunchKernelGGL(fill_u32, dim3(grid_for(words, 256)), dim3(256), 0, J->stream,
                       (uint32_t*)(J->op16 + P.sum_x16), words, 0u);
  }
  hipLaunchKernelGGL(fill_u32, dim3(grid_for(nds, 256)), dim3(256), 0, J->stream, J->zany,
                     (int64_t)nds, 0u);
  if (hipGetLastError() != hipSuccess) return bail(fail(ctx, HM_EDEVICE, "flag clear failed"));
  HM_A(J->target, nds);
  if (stages & HM_STAGE_FUZZY) {
    HM_A(J->key1, E); HM_A(J->in_w, E); HM_A(J->erank, E); HM_A(J->wval, E);
    HM_A(J->k1s, E); HM_A(J->p1s, E);
    HM_A(J->tmpv, E);
    HM_A(J->serr, 1);
    hipLaunchKernelGGL(fill_u32, dim3(1), dim3(256), 0, J->stream, J->serr, (int64_t)1, 0u);
    if (hipGetLastError() != hipSuccess) return bail(fail(ctx, HM_EDEVICE, "flag clear failed"));
    HM_A(J->inptr, R + 1);
    HM_A(J->cntu, R); HM_A(J->indptr_g, R + 1);
    J->total_nnz = J->indptr_g + R;
    HM_A(J->rows, 2 * E); HM_A(J->cols, 2 * E); HM_A(J->vals, 2 * E);
    HM_A(J->indptr_l, R + nds); HM_A(J->iptr_off, nds);
  }
  const int64_t scan_max = std::max<int64_t>(R + 1, E + 1);
  HM_A(J->sc32, cdiv(scan_max, kScanTile) + 2);
  HM_A(J->sc64, cdiv(scan_max, kScanTile) + 2);
#undef HM_A

  // descriptors
  J->hdesc.resize(nds);
  int64_t off_x16 = 0, off_x = 0, off_xr = 0, off_cos = 0, off_pad = 0;
  std::vector<int64_t> pad_off_h(nds + 1, 0);
  std::vector<double> target_h(nds);
  J->iptr_off_h.resize(nds);
  for (uint32_t i = 0; i < nds; ++i) {
    const DsHost& h = P.dsh[i];
    DsDesc& D = J->hdesc[i];
    memset(&D, 0, sizeof(D));
    D.x = J->x + off_x;
    if (knn) {
      D.xr = J->xr + off_xr;
      if (h.metric == 1) { D.op32 = J->op32c + off_cos; off_cos += h.n_pad * h.kp; }
      else D.op32 = D.xr;
      D.op16 = J->op16 ? J->op16 + off_x16 : nullptr;
      D.op16lo = J->op16lo ? J->op16lo + off_x16 : nullptr;
      D.colbias = J->colbias + off_pad;
      D.rnorm = J->rnorm + h.row_off;
      D.rstat = J->rstat + h.row_off * HM_NSTAT;
      D.smax = J->smax + (int64_t)i * HM_NSTAT;
      D.cand = J->cand + h.row_off * kCandCap;
      D.ctau = J->ctau + h.row_off * 2 * kSplitCap;
      D.ccnt = J->ccnt + h.row_off * 2 * kSplitCap;
      D.thr = J->thr + h.row_off * kSplitCap;
    }
    D.knn_idx = J->knn_idx + h.edge_off;
    D.knn_dist = J->knn_dist + h.edge_off;
    D.knn_dist64 = J->dist64 ? J->dist64 + h.edge_off : nullptr;
    D.err3 = J->err3 ? J->err3 + h.row_off * 3 : nullptr;
    D.n = h.n;
    D.n_pad = h.n_pad;
    D.row_off = h.row_off;
    D.edge_off = h.edge_off;
    D.d = h.d;
    D.kp = h.kp;
    D.kp16 = h.kp16;
    D.k = h.k;
    D.metric = h.metric;
    D.C = h.C;
    D.nsplit = h.nsplit;
    D.nlist = h.nlist;
    D.cout = h.cout;
    D.C0 = h.C0;
    D.pstride = h.pstride;
    pad_off_h[i] = off_pad;
    off_x += h.n * h.d;
    off_xr += h.n_pad * h.kp;
    off_x16 += h.n_pad * h.kp16;
    off_pad += h.n_pad;
    target_h[i] = std::log2((double)h.k);  // std f64::log2 (glibc), src/smooth_knn.rs:25
    J->iptr_off_h[i] = h.row_off + i;
  }
  pad_off_h[nds] = off_pad;
  hm_status st = HM_OK;
  auto up = [&](void* dst, const void* src, size_t bytes) {
    if (st == HM_OK &&
        hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, J->stream) != hipSuccess)
      st = fail(ctx, HM_EDEVICE, "descriptor upload failed");
  };
  up(J->d_desc, J->hdesc.data(), sizeof(DsDesc) * nds);
  up(J->target, target_h.data(), sizeof(double) * nds);
  if (knn) up(J->pad_off, pad_off_h.data(), sizeof(int64_t) * (nds + 1));
  if (stages & HM_STAGE_FUZZY) up(J->iptr_off, J->iptr_off_h.data(), sizeof(int64_t) * nds);
  if (knn && J->mode >= 0) {
    int64_t cap = 0;
    for (auto& h : P.dsh) {
      const int li = sweep_list_of(J->mode, h.kp16);
      cap += cdiv(h.n, kListRows[li] ? kListRows[li] : sweep_rows(J->mode)) * kSplitCap;
    }
    hm_status s2 = dalloc(J, &J->items_l1, std::max<int64_t>(cap, 1));
    if (s2 != HM_OK) return bail(s2);
  }
  if (st == HM_OK) st = apply_sweep(J);
  if (st == HM_OK && hipStreamSynchronize(J->stream) != hipSuccess)  // host sources are locals
    st = fail(ctx, HM_EDEVICE, "descriptor upload failed");
  if (st != HM_OK) return bail(st);
  if (!arena) {  // (the timing events are created by the first timed run)
    ctx->in_use += J->bytes;
    J->counted = true;
  }
  *out = J;
  return HM_OK;
}

hm_status hm_job_create(hm_ctx* ctx, uint32_t nds, const uint64_t* n, const uint32_t* d,
                        const uint32_t* k, const int32_t* metric, const hm_options* opt,
                        uint32_t stages, hm_job** out) {
  return job_create(ctx, nds, n, d, k, metric, opt, stages, nullptr, 0, out);
}

void hm_job_destroy(hm_job* J) {
  if (!J) return;
  (void)hipSetDevice(J->ctx->device);
  if (J->stream) (void)hipStreamSynchronize(J->stream);
  if (J->gexec) (void)hipGraphExecDestroy(J->gexec);
  for (auto e : J->ev)
    if (e) (void)hipEventDestroy(e);
  for (void* p : J->allocs) (void)hipFreeAsync(p, J->stream);
  if (J->stream && J->own_stream) {
    // the stream goes back to the context (up to 16 kept) once its frees ran
    if (hipStreamSynchronize(J->stream) == hipSuccess && J->ctx->spare_streams.size() < 16)
      J->ctx->spare_streams.push_back(J->stream);
    else
      (void)hipStreamDestroy(J->stream);
  }
  if (J->counted) J->ctx->in_use -= J->bytes;
  delete J;
}

uint64_t hm_job_bytes(const hm_job* J) { return J ? J->bytes : 0; }

hm_status hm_job_upload(hm_job* J, uint32_t ds, const float* data) {
  if (!J || ds >= (uint32_t)J->nds || !data) return HM_EINVAL;
  const DsHost& h = J->dsh[ds];
  HM_HIP(J->ctx, hipMemcpyAsync((void*)J->hdesc[ds].x, data, (size_t)h.n * h.d * 4,
                                hipMemcpyHostToDevice, J->stream));
  return HM_OK;
}

hm_status hm_job_upload_device(hm_job* J, uint32_t ds, const float* dev) {
  if (!J || ds >= (uint32_t)J->nds || !dev) return HM_EINVAL;
  const DsHost& h = J->dsh[ds];
  HM_HIP(J->ctx, hipMemcpyAsync((void*)J->hdesc[ds].x, dev, (size_t)h.n * h.d * 4,
                                hipMemcpyDeviceToDevice, J->stream));
  return HM_OK;
}

hm_status hm_job_set_knn(hm_job* J, uint32_t ds, const uint32_t* idx, const float* dist) {
  if (!J || ds >= (uint32_t)J->nds || !idx || !dist) return HM_EINVAL;
  const DsHost& h = J->dsh[ds];
  const size_t m = (size_t)h.n * h.k;
  HM_HIP(J->ctx, hipMemcpyAsync(J->knn_idx + h.edge_off, idx, m * 4, hipMemcpyHostToDevice, J->stream));
  HM_HIP(J->ctx, hipMemcpyAsync(J->knn_dist + h.edge_off, dist, m * 4, hipMemcpyHostToDevice, J->stream));
  return HM_OK;
}

hm_status hm_job_set_calib(hm_job* J, uint32_t ds, const float* sig, const float* rh) {
  if (!J || ds >= (uint32_t)J->nds || !sig || !rh) return HM_EINVAL;
  const DsHost& h = J->dsh[ds];
  HM_HIP(J->ctx, hipMemcpyAsync(J->sigma + h.row_off, sig, h.n * 4, hipMemcpyHostToDevice, J->stream));
  HM_HIP(J->ctx, hipMemcpyAsync(J->rho + h.row_off, rh, h.n * 4, hipMemcpyHostToDevice, J->stream));
  return HM_OK;
}

// Autotuner (hm_options.autotune): times the whole pipeline for candidate
// sweep configurations -- split count x XCD item permutation at the planned
// margin, then margins 0, k, 3k at the best of those -- and keeps the
// fastest. The outputs do not depend on the configuration (it only decides
// which columns are examined), so tuning changes time only.
hm_status autotune(hm_job* J) {
  hm_ctx* ctx = J->ctx;
  const bool timed0 = J->timed;
  J->timed = true;
  auto time_cfg = [&](float* ms) -> hm_status {
    HM_TRY(apply_sweep(J));
    float best = 1e30f;
    for (int rep = 0; rep < 2; ++rep) {  // the first run also warms caches
      HM_TRY(enqueue_pipeline(J));
      HM_HIP(ctx, hipStreamSynchronize(J->stream));
      float t = 0.f;
      HM_HIP(ctx, hipEventElapsedTime(&t, J->ev[0], J->ev[6]));
      best = std::min(best, t);
    }
    *ms = best;
    return HM_OK;
  };
  struct Cfg { int mult, nsplit; bool perm; std::vector<DsHost> dsh; float ms; };
  std::vector<Cfg> tried;
  auto same_plan = [](const std::vector<DsHost>& a, const std::vector<DsHost>& b) {
    for (size_t i = 0; i < a.size(); ++i)
      if (a[i].C != b[i].C || a[i].nsplit != b[i].nsplit) return false;
    return true;
  };
  auto try_cfg = [&](int mult, int nsplit, bool perm) -> hm_status {
    std::vector<DsHost> dsh = J->dsh;
    plan_sweep(dsh, J->opt.margin, nsplit, J->mode, ctx->num_cus, mult);
    for (auto& c : tried)
      if (c.perm == perm && same_plan(c.dsh, dsh)) return HM_OK;
    J->dsh = dsh;
    J->xcd_perm = perm;
    float ms = 0.f;
    HM_TRY(time_cfg(&ms));
    tried.push_back(Cfg{mult, nsplit, perm, dsh, ms});
    return HM_OK;
  };
  const int nsplits[] = {0, 1, 2, 3, 4, 5};
  for (int ns : nsplits) {
    if (J->opt.nsplit > 0 && ns != J->opt.nsplit) continue;
    for (bool perm : {true, false}) HM_TRY(try_cfg(2, ns, perm));
  }
  auto best_of = [&]() {
    size_t b = 0;
    for (size_t i = 1; i < tried.size(); ++i)
      if (tried[i].ms < tried[b].ms) b = i;
    return b;
  };
  if (tried.empty()) return fail(ctx, HM_EINVAL, "autotune: no sweep configuration to time");
  if (J->opt.margin < 0) {
    const Cfg b = tried[best_of()];
    for (int mult : {0, 1, 3}) HM_TRY(try_cfg(mult, b.nsplit, b.perm));
  }
  const Cfg& b = tried[best_of()];
  J->dsh = b.dsh;
  J->xcd_perm = b.perm;
  J->tune_ms = b.ms;
  HM_TRY(apply_sweep(J));
  J->timed = timed0;
  J->tuned = true;
  return HM_OK;
}

hm_status hm_job_run(hm_job* J) {
  if (!J) return HM_EINVAL;
  hm_ctx* ctx = J->ctx;
  HM_HIP(ctx, hipSetDevice(ctx->device));
  J->lap_valid = false;
  if ((J->opt.timing || J->opt.autotune) && !J->ev[0]) {  // timing events, on first need
    for (auto& e : J->ev)
      if (!e && hipEventCreate(&e) != hipSuccess) return fail(ctx, HM_EDEVICE, "event create");
  }
  if (J->opt.autotune && !J->tuned && (J->stages & HM_STAGE_KNN) && J->mode >= 0)
    HM_TRY(autotune(J));
  J->timed = J->opt.timing && !J->opt.use_graph;
  if (J->opt.use_graph) {
    if (!J->captured) {
      hipGraph_t g = nullptr;
      HM_HIP(ctx, hipStreamBeginCapture(J->stream, hipStreamCaptureModeThreadLocal));
      hm_status s = enqueue_pipeline(J);
      hipError_t e = hipStreamEndCapture(J->stream, &g);
      if (s != HM_OK) {
        if (g) (void)hipGraphDestroy(g);
        return s;
      }
      if (e != hipSuccess) return fail(ctx, HM_EDEVICE, std::string("capture: ") + hipGetErrorString(e));
      e = hipGraphInstantiate(&J->gexec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) return fail(ctx, HM_EDEVICE, std::string("instantiate: ") + hipGetErrorString(e));
      J->captured = true;
    }
    HM_HIP(ctx, hipGraphLaunch(J->gexec, J->stream));
    return HM_OK;
  }
  return enqueue_pipeline(J);
}

hm_status hm_job_sync(hm_job* J) {
  if (!J) return HM_EINVAL;
  HM_HIP(J->ctx, hipStreamSynchronize(J->stream));
  if (J->timed) {
    // stage boundaries: 0 start,1 prep,2 sweep,3 rerank,4 fallback,5 calib,
    // 6 fuzzy; 7 the sweep's threshold passes (inside the sweep stage)
    const int map[7][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 4}, {4, 5}, {5, 6}, {1, 7}};
    for (int s = 0; s < 7; ++s) {
      const int slot = s < 6 ? s : HM_T_THRESHOLD;
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, J->ev[map[s][0]], J->ev[map[s][1]]);
      J->times[slot] = ms;
    }
    float tot = 0.f;
    (void)hipEventElapsedTime(&tot, J->ev[0], J->ev[6]);
    J->times[HM_T_TOTAL] = tot;
  }
  return HM_OK;
}

int hm_job_done(hm_job* J) {
  if (!J) return HM_EINVAL;
  hipError_t e = hipStreamQuery(J->stream);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return HM_EDEVICE;
}

hm_status hm_job_times(const hm_job* J, float* ms) {
  if (!J || !ms) return HM_EINVAL;
  for (int i = 0; i < HM_NTIMES; ++i) ms[i] = J->times[i];
  return HM_OK;
}

hm_status hm_job_counters(hm_job* J, uint64_t* c) {
  if (!J || !c) return HM_EINVAL;
  for (int i = 0; i < HM_NCOUNTERS; ++i) c[i] = 0;
  if (J->stages & HM_STAGE_KNN) {
    int cnt = 0;
    HM_HIP(J->ctx, hipMemcpy(&cnt, J->count, sizeof(int), hipMemcpyDeviceToHost));
    c[HM_C_FALLBACK_ROWS] = (uint64_t)cnt;
    c[HM_C_CANDIDATES] = (uint64_t)J->dsh[0].nlist * J->dsh[0].cout;
    c[HM_C_NSPLIT] = (uint64_t)J->dsh[0].nsplit;
    c[HM_C_TARGET] = (uint64_t)J->dsh[0].C;
    c[HM_C_SAMPLE] = (uint64_t)J->dsh[0].pstride;
    if (J->mode >= 0) {
      // keys listed for dataset 0: its rows' per-list counts (n x nlist,
      // row-major from the dataset's base)
      const int64_t m = J->dsh[0].n * J->dsh[0].nlist;
      const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(512, cdiv(m, 256)));
      hipLaunchKernelGGL(sum_u32_part_kernel, dim3(g), dim3(256), 0, J->stream,
                         (const uint32_t*)J->ccnt, m, J->stat64);
      hipLaunchKernelGGL(sum_u64_kernel, dim3(1), dim3(256), 0, J->stream,
                         (const uint64_t*)J->stat64, (int)g, J->stat64 + 512);
      HM_LAUNCH(J->ctx);
      uint64_t listed = 0;
      HM_HIP(J->ctx, hipMemcpyAsync(&listed, J->stat64 + 512, 8, hipMemcpyDeviceToHost, J->stream));
      HM_HIP(J->ctx, hipStreamSynchronize(J->stream));
      c[HM_C_LISTED] = listed;
    }
    if (J->nitems_l2) {
      int esc = 0;
      HM_HIP(J->ctx, hipMemcpy(&esc, J->nitems_l2 + 1, sizeof(int), hipMemcpyDeviceToHost));
      c[HM_C_ESCALATED] = (uint64_t)esc;
    }
  }
  if (J->stages & HM_STAGE_FUZZY) {
    HM_TRY(sym_check(J));
    uint64_t nnz = 0;
    HM_HIP(J->ctx, hipMemcpy(&nnz, J->total_nnz, sizeof(uint64_t), hipMemcpyDeviceToHost));
    c[HM_C_NNZ] = nnz;
  }
  return HM_OK;
}

}  // extern "C"

// The largest relative rounding error of dataset ds's f32 distances, sigmas
// and directed membership strengths against their f64 values (the FP64
// chain), or -1 each without that chain.
static hm_status fp64_error(hm_job* J, uint32_t ds, double* err) {
  err[0] = err[1] = err[2] = -1.0;
  if (!J->err3) return HM_OK;
  const DsHost& h = J->dsh[ds];
  std::vector<float> e((size_t)h.n * 3);
  HM_HIP(J->ctx, hipMemcpy(e.data(), J->err3 + h.row_off * 3, e.size() * 4, hipMemcpyDeviceToHost));
  err[0] = err[1] = err[2] = 0.0;
  for (size_t i = 0; i < e.size(); ++i) err[i % 3] = std::max(err[i % 3], (double)e[i]);
  return HM_OK;
}

extern "C" {

hm_status hm_job_download(hm_job* J, uint32_t ds, hm_graph_out* o) {
  if (!J || !o || ds >= (uint32_t)J->nds) return HM_EINVAL;
  hm_ctx* ctx = J->ctx;
  HM_HIP(ctx, hipStreamSynchronize(J->stream));
  const DsHost& h = J->dsh[ds];
  const size_t m = (size_t)h.n * h.k;
  if (o->dists64 && !J->dist64)
    return fail(ctx, HM_EINVAL, "no f64 distances: they are kept in the FP64 calibration mode "
                                "with the kNN stage");
  struct Piece { void* dst; const void* src; size_t bytes; };
  std::vector<Piece> pc;
  auto want = [&](void* dst, const void* src, size_t bytes) {
    if (dst && bytes) pc.push_back(Piece{dst, src, bytes});
  };
  want(o->indices, J->knn_idx + h.edge_off, m * 4);
  want(o->dists, J->knn_dist + h.edge_off, m * 4);
  want(o->sigmas, J->sigma + h.row_off, h.n * 4);
  want(o->rhos, J->rho + h.row_off, h.n * 4);
  want(o->sigmas64, J->sigma64 + h.row_off, h.n * 8);
  if (J->dist64) want(o->dists64, J->dist64 + h.edge_off, m * 8);
  HM_TRY(fp64_error(J, ds, o->err64));
  o->nnz = 0;
  if (J->stages & HM_STAGE_FUZZY) {
    HM_TRY(sym_check(J));
    uint64_t se[2];
    HM_HIP(ctx, hipMemcpy(se, J->indptr_g + h.row_off, 8, hipMemcpyDeviceToHost));
    if (ds + 1 < (uint32_t)J->nds)
      HM_HIP(ctx, hipMemcpy(se + 1, J->indptr_g + h.row_off + h.n, 8, hipMemcpyDeviceToHost));
    else
      HM_HIP(ctx, hipMemcpy(se + 1, J->total_nnz, 8, hipMemcpyDeviceToHost));
    const uint64_t nnz = se[1] - se[0];
    o->nnz = nnz;
    if ((o->rows || o->cols || o->vals) && o->cap < nnz)
      return fail(ctx, HM_EINVAL, "COO capacity " + std::to_string(o->cap) + " < nnz " + std::to_string(nnz));
    want(o->rows, J->rows + se[0], nnz * 4);
    want(o->cols, J->cols + se[0], nnz * 4);
    want(o->vals, J->vals + se[0], nnz * 4);
    want(o->indptr, J->indptr_l + J->iptr_off_h[ds], (h.n + 1) * 8);
  }
  // The outputs go through the context's pinned staging buffer: asynchronous
  // DMA copies, one synchronisation, then host copies (a copy into pageable
  // memory is staged and synchronised piece by piece otherwise). If pinning
  // fails, the copies go directly.
  uint64_t total = 0;
  for (auto& p : pc) total += (p.bytes + 255) & ~(uint64_t)255;
  if (total > ctx->stage_cap) {
    uint64_t cap = 1ull << 20;
    while (cap < total) cap <<= 1;
    if (ctx->stage) (void)hipHostFree(ctx->stage);
    ctx->stage = nullptr;
    ctx->stage_cap = 0;
    if (hipHostMalloc((void**)&ctx->stage, cap, hipHostMallocDefault) == hipSuccess) {
      ctx->stage_cap = cap;
    } else {
      (void)hipGetLastError();
      ctx->stage = nullptr;
    }
  }
  if (!ctx->stage) {
    for (auto& p : pc) HM_HIP(ctx, hipMemcpy(p.dst, p.src, p.bytes, hipMemcpyDeviceToHost));
    return HM_OK;
  }
  uint64_t off = 0;
  for (auto& p : pc) {
    HM_HIP(ctx, hipMemcpyAsync(ctx->stage + off, p.src, p.bytes, hipMemcpyDeviceToHost, J->stream));
    off += (p.bytes + 255) & ~(uint64_t)255;
  }
  HM_HIP(ctx, hipStreamSynchronize(J->stream));
  off = 0;
  for (auto& p : pc) {
    std::memcpy(p.dst, ctx->stage + off, p.bytes);
    off += (p.bytes + 255) & ~(uint64_t)255;
  }
  return HM_OK;
}

hm_status hm_job_device_outputs(hm_job* J, uint32_t ds, const uint32_t** idx, const float** dist,
                                const uint32_t** rows, const uint32_t** cols, const float** vals,
                                const uint64_t** indptr) {
  if (!J || ds >= (uint32_t)J->nds) return HM_EINVAL;
  const DsHost& h = J->dsh[ds];
  if (idx) *idx = J->knn_idx + h.edge_off;
  if (dist) *dist = J->knn_dist + h.edge_off;
  uint64_t s0 = 0;
  if ((rows || cols || vals) && (J->stages & HM_STAGE_FUZZY)) {
    HM_TRY(sym_check(J));
    HM_HIP(J->ctx, hipMemcpy(&s0, J->indptr_g + h.row_off, 8, hipMemcpyDeviceToHost));
  }
  if (rows) *rows = J->rows ? J->rows + s0 : nullptr;
  if (cols) *cols = J->cols ? J->cols + s0 : nullptr;
  if (vals) *vals = J->vals ? J->vals + s0 : nullptr;
  if (indptr) *indptr = J->indptr_l ? J->indptr_l + J->iptr_off_h[ds] : nullptr;
  return HM_OK;
}

// ---- stage-level and fused host-buffer entry points ----

static hm_status one_job(hm_ctx* ctx, uint64_t n, uint32_t d, uint32_t k, int32_t metric,
                         const hm_options* opt, uint32_t stages, hm_job** J) {
  const uint64_t nn[1] = {n};
  const uint32_t dd[1] = {d}, kk[1] = {k};
  const int32_t mm[1] = {metric};
  return hm_job_create(ctx, 1, nn, dd, kk, mm, opt, stages, J);
}

hm_status hm_exact_knn(hm_ctx* ctx, const float* data, uint64_t n, uint32_t d, uint32_t k,
                       int32_t metric, const hm_options* opt, uint32_t* indices, float* dists) {
  if (!ctx || !data || !indices || !dists) return fail(ctx, HM_EINVAL, "null argument");
  if (d == 0 || n == 0) return fail(ctx, HM_EBADSHAPE, "empty input");
  if (k == 0) return fail(ctx, HM_EINVAL, "n_neighbors must be >= 1");
  if (n < k) return fail(ctx, HM_ETOOFEW, "need at least n_neighbors points");
  for (uint64_t t = 0; t < n * d; ++t)
    if (!std::isfinite(data[t])) return fail(ctx, HM_ENONFINITE, "non-finite input");
  hm_job* J = nullptr;
  HM_TRY(one_job(ctx, n, d, k, metric, opt, HM_STAGE_KNN, &J));
  hm_status s = hm_job_upload(J, 0, data);
  if (s == HM_OK) s = hm_job_run(J);
  if (s == HM_OK) s = hm_job_sync(J);
  hm_graph_out o{};
  o.indices = indices;
  o.dists = dists;
  if (s == HM_OK) s = hm_job_download(J, 0, &o);
  hm_job_destroy(J);
  return s;
}

hm_status hm_smooth_knn(hm_ctx* ctx, const float* knn_dists, uint64_t n, uint32_t k, int32_t calib,
                        float* sigmas, float* rhos) {
  if (!ctx || !knn_dists || !sigmas || !rhos || n == 0 || k == 0)
    return fail(ctx, HM_EINVAL, "null argument");
  hm_options o;
  hm_default_options(&o);
  o.calib = calib;
  hm_job* J = nullptr;
  // a calibration-only job: the kNN distances are uploaded, not computed
  const uint64_t nn[1] = {n};
  const uint32_t dd[1] = {1}, kk[1] = {k};
  hm_status s = hm_job_create(ctx, 1, nn, dd, kk, nullptr, &o, HM_STAGE_CALIB, &J);
  if (s != HM_OK) return s;
  std::vector<uint32_t> dummy((size_t)n * k, 0);
  s = hm_job_set_knn(J, 0, dummy.data(), knn_dists);
  if (s == HM_OK) s = hm_job_run(J);
  if (s == HM_OK) s = hm_job_sync(J);
  hm_graph_out out{};
  out.sigmas = sigmas;
  out.rhos = rhos;
  if (s == HM_OK) s = hm_job_download(J, 0, &out);
  hm_job_destroy(J);
  return s;
}

hm_status hm_fuzzy_simplicial_set(hm_ctx* ctx, const uint32_t* idx, const float* dists,
                                  const float* sigmas, const float* rhos, uint64_t n, uint32_t k,
                                  uint32_t* rows, uint32_t* cols, float* vals, uint64_t cap,
                                  uint64_t* nnz) {
  if (!ctx || !idx || !dists || !sigmas || !rhos || !nnz) return fail(ctx, HM_EINVAL, "null argument");
  for (uint64_t t = 0; t < n * k; ++t)
    if (idx[t] >= n) return fail(ctx, HM_EINVAL, "neighbour index out of range");
  hm_options o;
  hm_default_options(&o);
  const uint64_t nn[1] = {n};
  const uint32_t dd[1] = {1}, kk[1] = {k};
  hm_job* J = nullptr;
  HM_TRY(hm_job_create(ctx, 1, nn, dd, kk, nullptr, &o, HM_STAGE_FUZZY, &J));
  hm_status s = hm_job_set_knn(J, 0, idx, dists);
  if (s == HM_OK) s = hm_job_set_calib(J, 0, sigmas, rhos);
  if (s == HM_OK) s = hm_job_run(J);
  if (s == HM_OK) s = hm_job_sync(J);
  hm_graph_out out{};
  out.rows = rows;
  out.cols = cols;
  out.vals = vals;
  out.cap = cap;
  if (s == HM_OK) s = hm_job_download(J, 0, &out);
  *nnz = out.nnz;
  hm_job_destroy(J);
  return s;
}

hm_status hm_csr_from_fuzzy(hm_ctx* ctx, const uint32_t* rows, uint64_t nnz, uint64_t n,
                            uint64_t* indptr) {
  if (!ctx || !indptr || (!rows && nnz)) return fail(ctx, HM_EINVAL, "null argument");
  HM_HIP(ctx, hipSetDevice(ctx->device));
  uint32_t* drows = nullptr;
  uint32_t* dptr = nullptr;
  HM_HIP(ctx, hipMalloc(&drows, std::max<uint64_t>(nnz, 1) * 4));
  if (hipMalloc(&dptr, (n + 1) * 4) != hipSuccess) {
    (void)hipFree(drows);
    return fail(ctx, HM_ENOMEM, "hipMalloc failed");
  }
  hm_status s = HM_OK;
  if (nnz && hipMemcpy(drows, rows, nnz * 4, hipMemcpyHostToDevice) != hipSuccess)
    s = fail(ctx, HM_EDEVICE, "upload failed");
  if (s == HM_OK) {
    // rows are sorted: indptr[r] = lower_bound(rows, r), indptr[n] = nnz
    hipLaunchKernelGGL(lower_bounds, dim3(grid_for((int64_t)n + 1, 256)), dim3(256), 0, ctx->stream,
                       (const uint32_t*)drows, (int64_t)nnz, (int64_t)n, dptr);
    std::vector<uint32_t> h(n + 1);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess ||
        hipMemcpy(h.data(), dptr, (n + 1) * 4, hipMemcpyDeviceToHost) != hipSuccess)
      s = fail(ctx, HM_EDEVICE, "csr kernel failed");
    else
      for (uint64_t i = 0; i <= n; ++i) indptr[i] = h[i];
  }
  (void)hipFree(drows);
  (void)hipFree(dptr);
  return s;
}

hm_status hm_knn_graph(hm_ctx* ctx, const float* data, uint64_t n, uint32_t d, uint32_t k,
                       int32_t metric, const hm_options* opt, hm_graph_out* out) {
  if (!ctx || !data || !out) return fail(ctx, HM_EINVAL, "null argument");
  uint64_t bad = 0;
  hm_status v = hm_validate(data, n * d, d, k, &bad);
  if (v != HM_OK) return fail(ctx, v, "input validation failed");
  hm_job* J = nullptr;
  HM_TRY(one_job(ctx, n, d, k, metric, opt, HM_STAGE_ALL, &J));
  hm_status s = hm_job_upload(J, 0, data);
  if (s == HM_OK) s = hm_job_run(J);
  if (s == HM_OK) s = hm_job_sync(J);
  if (s == HM_OK) s = hm_job_download(J, 0, out);
  hm_job_destroy(J);
  return s;
}

hm_status hm_knn_graph_noalloc(hm_ctx* ctx, const float* data, uint64_t n, uint32_t d,
                               uint32_t k, int32_t metric, const hm_options* opt,
                               hm_graph_out* out) {
  if (!ctx || !data || !out) return fail(ctx, HM_EINVAL, "null argument");
  if (!ctx->scratch) return fail(ctx, HM_EDEVICE, "no scratch array on this device");
  const hm_status v = hm_validate(nullptr, n * d, d, k, nullptr);  // shape and k
  if (v != HM_OK) return fail(ctx, v, "input validation failed");
  // calls on one device share its array: one at a time
  std::lock_guard<std::mutex> lock(g_scratch_mu[ctx->device]);
  HM_HIP(ctx, hipSetDevice(ctx->device));
  hm_options o;
  if (opt) o = *opt; else hm_default_options(&o);
  o.use_graph = 0;  // one eager run; no events, no autotuning runs
  o.autotune = 0;
  o.timing = 0;
  hipStream_t st = ctx->stream;
  // finite input (src/api.rs:142-144), checked on the device
  uint64_t* part = (uint64_t*)(ctx->scratch + (HM_NOALLOC_SCRATCH_BYTES - kScratchTail));
  uint64_t* first = part + 1024;
  const int64_t len = (int64_t)(n * d);
  const unsigned g = (unsigned)std::min<int64_t>(1024, cdiv(len, 256));
  hipLaunchKernelGGL(nonfinite_part_kernel, dim3(g), dim3(256), 0, st, data, len, part);
  hipLaunchKernelGGL(nonfinite_min_kernel, dim3(1), dim3(256), 0, st, (const uint64_t*)part,
                     (int)g, first);
  HM_LAUNCH(ctx);
  uint64_t bad = 0;
  HM_HIP(ctx, hipMemcpyAsync(&bad, first, 8, hipMemcpyDeviceToHost, st));
  HM_HIP(ctx, hipStreamSynchronize(st));
  if (bad != ~0ull)
    return fail(ctx, HM_ENONFINITE, "non-finite input at index " + std::to_string(bad));
  hm_job* J = nullptr;
  HM_TRY(job_create(ctx, 1, &n, &d, &k, &metric, &o, HM_STAGE_ALL, ctx->scratch,
                    HM_NOALLOC_SCRATCH_BYTES - kScratchTail, &J));
  hm_status s = hm_job_upload_device(J, 0, data);
  if (s == HM_OK) s = hm_job_run(J);
  uint64_t nnz = 0;
  if (s == HM_OK && (hipMemcpyAsync(&nnz, J->total_nnz, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
                     hipStreamSynchronize(st) != hipSuccess))
    s = fail(ctx, HM_EDEVICE, "graph build failed");
  if (s == HM_OK) {
    out->nnz = nnz;
    if ((out->rows || out->cols || out->vals) && out->cap < nnz)
      s = fail(ctx, HM_EINVAL, "COO capacity " + std::to_string(out->cap) + " < nnz " +
                                   std::to_string(nnz));
  }
  auto copy = [&](void* dst, const void* src, uint64_t bytes) {
    if (s == HM_OK && dst && bytes &&
        hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess)
      s = fail(ctx, HM_EDEVICE, "output copy failed");
  };
  copy(out->indices, J->knn_idx, n * k * 4);
  copy(out->dists, J->knn_dist, n * k * 4);
  copy(out->sigmas, J->sigma, n * 4);
  copy(out->rhos, J->rho, n * 4);
  copy(out->sigmas64, J->sigma64, n * 8);
  if (s == HM_OK && out->dists64 && !J->dist64)
    s = fail(ctx, HM_EINVAL, "no f64 distances: they are kept in the FP64 calibration mode");
  copy(out->dists64, J->dist64, n * k * 8);
  // (the error statement needs a host reduction: not made here)
  out->err64[0] = out->err64[1] = out->err64[2] = -1.0;
  copy(out->rows, J->rows, nnz * 4);
  copy(out->cols, J->cols, nnz * 4);
  copy(out->vals, J->vals, nnz * 4);
  copy(out->indptr, J->indptr_l, (n + 1) * 8);
  if (s == HM_OK && hipStreamSynchronize(st) != hipSuccess)
    s = fail(ctx, HM_EDEVICE, "output copy failed");
  hm_job_destroy(J);
  return s;
}

// ---- consumers of the graph's CSR: components and Laplacian ----

}  // extern "C"

namespace {

// Exclusive scan of n u32 values on `st` (tmp: cdiv(n, kScanTile) + 2 values).
hm_status scan_u32_excl(hm_ctx* ctx, hipStream_t st, const uint32_t* in, int64_t n,
                        uint32_t* out, uint32_t* total, uint32_t* tmp) {
  if (n <= kScanSmall<uint32_t>) {
    constexpr int NT = kScanWideThreads, MS = kScanWideQuads<uint32_t>;
    hipLaunchKernelGGL((scan_small<uint32_t, NT, MS, false>), dim3(1), dim3(NT), 0, st, in, n, out, total,
                       (uint32_t*)nullptr, (int*)nullptr);
  } else {
    const int64_t nt = cdiv(n, kScanTile);
    hipLaunchKernelGGL(scan_tile_sums<uint32_t>, dim3((unsigned)nt), dim3(256), 0, st, in, n, tmp);
    hipLaunchKernelGGL(scan_sums<uint32_t>, dim3(1), dim3(1024), 0, st, tmp, nt, total);
    hipLaunchKernelGGL(scan_tiles<uint32_t>, dim3((unsigned)nt), dim3(256), 0, st, in, n,
                       (const uint32_t*)tmp, out);
  }
  HM_LAUNCH(ctx);
  return HM_OK;
}

// Stream-ordered device buffers released when the owner goes out of scope.
struct DevBufs {
  hipStream_t st;
  std::vector<void*> p;
  explicit DevBufs(hipStream_t s) : st(s) {}
  ~DevBufs() {
    for (void* q : p) (void)hipFreeAsync(q, st);
  }
  template <typename T>
  T* get(hm_ctx* ctx, int64_t count, hm_status* status) {
    void* q = nullptr;
    if (hipMallocAsync(&q, (size_t)std::max<int64_t>(count, 1) * sizeof(T), st) != hipSuccess) {
      (void)hipGetLastError();
      *status = fail(ctx, HM_ENOMEM, "device allocation failed");
      return nullptr;
    }
    p.push_back(q);
    return (T*)q;
  }
};

// Component labels (first-seen order) of the n-vertex CSR graph in device
// memory into out (device); the count into *ncomp. Synchronises `st`.
hm_status components_device(hm_ctx* ctx, hipStream_t st, const uint64_t* indptr,
                            const uint32_t* indices, int64_t n, uint32_t* out, uint64_t* ncomp) {
  hm_status s = HM_OK;
  DevBufs B(st);
  uint32_t* lab = B.get<uint32_t>(ctx, n, &s);
  uint32_t* root = lab ? B.get<uint32_t>(ctx, n, &s) : nullptr;
  uint32_t* rank = root ? B.get<uint32_t>(ctx, n, &s) : nullptr;
  uint32_t* flag = rank ? B.get<uint32_t>(ctx, 2, &s) : nullptr;
  uint32_t* tmp = flag ? B.get<uint32_t>(ctx, cdiv(n, kScanTile) + 2, &s) : nullptr;
  if (!tmp) return s;
  const unsigned g = grid_for(n, 256);
  hipLaunchKernelGGL(cc_init_kernel, dim3(g), dim3(256), 0, st, lab, n);
  HM_LAUNCH(ctx);
  // every step that changes a label lowers it; a step that changes none
  // leaves a fixed point (within diameter + 1 steps)
  for (int64_t it = 0; it <= n; ++it) {
    hipLaunchKernelGGL(fill_u32, dim3(1), dim3(256), 0, st, flag, (int64_t)1, 0u);
    hipLaunchKernelGGL(cc_step_kernel, dim3(g), dim3(256), 0, st, indptr, indices, n, lab, flag);
    hipLaunchKernelGGL(cc_jump_kernel, dim3(g), dim3(256), 0, st, lab, n);
    hipLaunchKernelGGL(cc_jump_kernel, dim3(g), dim3(256), 0, st, lab, n);
    HM_LAUNCH(ctx);
    uint32_t changed = 0;
    HM_HIP(ctx, hipMemcpyAsync(&changed, flag, 4, hipMemcpyDeviceToHost, st));
    HM_HIP(ctx, hipStreamSynchronize(st));
    if (!changed) break;
  }
  hipLaunchKernelGGL(cc_root_kernel, dim3(g), dim3(256), 0, st, (const uint32_t*)lab, n, root);
  HM_TRY(scan_u32_excl(ctx, st, root, n, rank, flag + 1, tmp));
  hipLaunchKernelGGL(cc_label_kernel, dim3(g), dim3(256), 0, st, (const uint32_t*)lab,
                     (const uint32_t*)rank, n, out);
  HM_LAUNCH(ctx);
  uint32_t nc = 0;
  HM_HIP(ctx, hipMemcpyAsync(&nc, flag + 1, 4, hipMemcpyDeviceToHost, st));
  HM_HIP(ctx, hipStreamSynchronize(st));
  *ncomp = nc;
  return HM_OK;
}

// src/sgd.rs:23-25: umap-learn's n_epochs default.
uint32_t default_n_epochs(uint64_t n) { return n <= 10000 ? 500u : 200u; }

// The edge schedule of the m-entry COO (rows, cols, w: device) into head,
// tail, eps (device, capacity m); the kept count into *count. Synchronises.
hm_status schedule_device(hm_ctx* ctx, hipStream_t st, const uint32_t* rows, const uint32_t* cols,
                          const float* w, int64_t m, uint32_t prune_epochs, uint32_t* head,
                          uint32_t* tail, double* eps, uint64_t* count) {
  if (m == 0) {
    *count = 0;
    return HM_OK;
  }
  hm_status s = HM_OK;
  DevBufs B(st);
  const unsigned gp = (unsigned)std::min<int64_t>(512, cdiv(m, 256));
  float* part = B.get<float>(ctx, gp + 1, &s);
  uint32_t* keep = part ? B.get<uint32_t>(ctx, m, &s) : nullptr;
  uint32_t* pos = keep ? B.get<uint32_t>(ctx, m, &s) : nullptr;
  uint32_t* tot = pos ? B.get<uint32_t>(ctx, cdiv(m, kScanTile) + 3, &s) : nullptr;
  if (!tot) return s;
  hipLaunchKernelGGL(sched_max_part_kernel, dim3(gp), dim3(256), 0, st, w, m, part);
  hipLaunchKernelGGL(sched_max_kernel, dim3(1), dim3(256), 0, st, (const float*)part, (int)gp,
                     part + gp);
  hipLaunchKernelGGL(sched_flag_kernel, dim3(grid_for(m, 256)), dim3(256), 0, st, w, m,
                     (const float*)(part + gp), prune_epochs, keep);
  HM_LAUNCH(ctx);
  HM_TRY(scan_u32_excl(ctx, st, keep, m, pos, tot, tot + 1));
  hipLaunchKernelGGL(sched_scatter_kernel, dim3(grid_for(m, 256)), dim3(256), 0, st, rows, cols, w,
                     m, (const uint32_t*)keep, (const uint32_t*)pos, (const float*)(part + gp),
                     head, tail, eps);
  HM_LAUNCH(ctx);
  uint32_t c = 0;
  HM_HIP(ctx, hipMemcpyAsync(&c, tot, 4, hipMemcpyDeviceToHost, st));
  HM_HIP(ctx, hipStreamSynchronize(st));
  *count = c;
  return HM_OK;
}

// A caller's CSR must index within its n vertices: indptr non-decreasing
// from 0, indices < n (the kernels follow them).
hm_status check_csr(hm_ctx* ctx, const uint64_t* indptr, const uint32_t* indices, uint64_t n) {
  if (indptr[0] != 0) return fail(ctx, HM_EINVAL, "indptr[0] must be 0");
  for (uint64_t i = 0; i < n; ++i)
    if (indptr[i + 1] < indptr[i]) return fail(ctx, HM_EINVAL, "indptr must be non-decreasing");
  for (uint64_t p = 0; p < indptr[n]; ++p)
    if (indices[p] >= n) return fail(ctx, HM_EINVAL, "column index out of range");
  return HM_OK;
}

}  // namespace

extern "C" {

hm_status hm_connected_components(hm_ctx* ctx, const uint64_t* indptr, const uint32_t* indices,
                                  uint64_t n, uint32_t* labels, uint64_t* n_components) {
  if (!ctx || !indptr || !labels || !n_components || (!indices && n && indptr[n]))
    return fail(ctx, HM_EINVAL, "null argument");
  if (n == 0) {
    *n_components = 0;
    return HM_OK;
  }
  if (n >= (1ull << 32)) return fail(ctx, HM_EUNSUPPORTED, "n >= 2^32");
  HM_TRY(check_csr(ctx, indptr, indices, n));
  HM_HIP(ctx, hipSetDevice(ctx->device));
  const uint64_t nnz = indptr[n];
  hipStream_t st = ctx->stream;
  hm_status s = HM_OK;
  DevBufs B(st);
  uint64_t* dptr = B.get<uint64_t>(ctx, n + 1, &s);
  uint32_t* didx = dptr ? B.get<uint32_t>(ctx, nnz, &s) : nullptr;
  uint32_t* dout = didx ? B.get<uint32_t>(ctx, n, &s) : nullptr;
  if (!dout) return s;
  HM_HIP(ctx, hipMemcpyAsync(dptr, indptr, (n + 1) * 8, hipMemcpyHostToDevice, st));
  if (nnz) HM_HIP(ctx, hipMemcpyAsync(didx, indices, nnz * 4, hipMemcpyHostToDevice, st));
  HM_TRY(components_device(ctx, st, dptr, didx, (int64_t)n, dout, n_components));
  HM_HIP(ctx, hipMemcpyAsync(labels, dout, n * 4, hipMemcpyDeviceToHost, st));
  HM_HIP(ctx, hipStreamSynchronize(st));
  return HM_OK;
}

hm_status hm_laplacian_new(hm_ctx* ctx, const uint64_t* indptr, const uint32_t* indices,
                           const float* vals, uint64_t n, double* svals, double* degrees) {
  if (!ctx || !indptr || (n && indptr[n] && (!indices || !vals)))
    return fail(ctx, HM_EINVAL, "null argument");
  if (n == 0) return HM_OK;
  HM_TRY(check_csr(ctx, indptr, indices, n));
  HM_HIP(ctx, hipSetDevice(ctx->device));
  const uint64_t nnz = indptr[n];
  hipStream_t st = ctx->stream;
  hm_status s = HM_OK;
  DevBufs B(st);
  uint64_t* dptr = B.get<uint64_t>(ctx, n + 1, &s);
  uint32_t* didx = dptr ? B.get<uint32_t>(ctx, nnz, &s) : nullptr;
  float* dval = didx ? B.get<float>(ctx, nnz, &s) : nullptr;
  double* ddeg = dval ? B.get<double>(ctx, n, &s) : nullptr;
  double* dinv = ddeg ? B.get<double>(ctx, n, &s) : nullptr;
  double* dsv = dinv ? B.get<double>(ctx, nnz, &s) : nullptr;
  if (!dsv) return s;
  HM_HIP(ctx, hipMemcpyAsync(dptr, indptr, (n + 1) * 8, hipMemcpyHostToDevice, st));
  if (nnz) {
    HM_HIP(ctx, hipMemcpyAsync(didx, indices, nnz * 4, hipMemcpyHostToDevice, st));
    HM_HIP(ctx, hipMemcpyAsync(dval, vals, nnz * 4, hipMemcpyHostToDevice, st));
  }
  const unsigned g = grid_for((int64_t)n, 256);
  hipLaunchKernelGGL(lap_degree_kernel, dim3(g), dim3(256), 0, st, (const uint64_t*)dptr,
                     (const float*)dval, (int64_t)n, ddeg, dinv);
  hipLaunchKernelGGL(lap_scale_kernel, dim3(g), dim3(256), 0, st, (const uint64_t*)dptr,
                     (const uint32_t*)didx, (const float*)dval, (const double*)dinv, (int64_t)n, dsv);
  HM_LAUNCH(ctx);
  if (degrees) HM_HIP(ctx, hipMemcpyAsync(degrees, ddeg, n * 8, hipMemcpyDeviceToHost, st));
  if (svals && nnz) HM_HIP(ctx, hipMemcpyAsync(svals, dsv, nnz * 8, hipMemcpyDeviceToHost, st));
  HM_HIP(ctx, hipStreamSynchronize(st));
  return HM_OK;
}

hm_status hm_laplacian_matvec(hm_ctx* ctx, const uint64_t* indptr, const uint32_t* indices,
                              const double* svals, uint64_t n, const double* x, double* y) {
  if (!ctx || !indptr || !x || !y || (n && indptr[n] && (!indices || !svals)))
    return fail(ctx, HM_EINVAL, "null argument");
  if (n == 0) return HM_OK;
  HM_TRY(check_csr(ctx, indptr, indices, n));
  HM_HIP(ctx, hipSetDevice(ctx->device));
  const uint64_t nnz = indptr[n];
  hipStream_t st = ctx->stream;
  hm_status s = HM_OK;
  DevBufs B(st);
  uint64_t* dptr = B.get<uint64_t>(ctx, n + 1, &s);
  uint32_t* didx = dptr ? B.get<uint32_t>(ctx, nnz, &s) : nullptr;
  double* dsv = didx ? B.get<double>(ctx, nnz, &s) : nullptr;
  double* dx = dsv ? B.get<double>(ctx, n, &s) : nullptr;
  double* dy = dx ? B.get<double>(ctx, n, &s) : nullptr;
  if (!dy) return s;
  HM_HIP(ctx, hipMemcpyAsync(dptr, indptr, (n + 1) * 8, hipMemcpyHostToDevice, st));
  if (nnz) {
    HM_HIP(ctx, hipMemcpyAsync(didx, indices, nnz * 4, hipMemcpyHostToDevice, st));
    HM_HIP(ctx, hipMemcpyAsync(dsv, svals, nnz * 8, hipMemcpyHostToDevice, st));
  }
  HM_HIP(ctx, hipMemcpyAsync(dx, x, n * 8, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(lap_matvec_kernel, dim3(grid_for((int64_t)n, 256)), dim3(256), 0, st,
                     (const uint64_t*)dptr, (const uint32_t*)didx, (const double*)dsv,
                     (const float*)nullptr, (const double*)nullptr, (int64_t)n, (const double*)dx,
                     dy);
  HM_LAUNCH(ctx);
  HM_HIP(ctx, hipMemcpyAsync(y, dy, n * 8, hipMemcpyDeviceToHost, st));
  HM_HIP(ctx, hipStreamSynchronize(st));
  return HM_OK;
}

hm_status hm_schedule_edges(hm_ctx* ctx, const uint32_t* rows, const uint32_t* cols,
                            const float* vals, uint64_t nnz, uint64_t n, uint32_t n_epochs,
                            uint32_t* head, uint32_t* tail, double* epochs_per_sample,
                            uint64_t* count) {
  if (!ctx || !count || (nnz && (!rows || !cols || !vals || !head || !tail || !epochs_per_sample)))
    return fail(ctx, HM_EINVAL, "null argument");
  if (nnz >= (1ull << 32)) return fail(ctx, HM_EUNSUPPORTED, "nnz >= 2^32");
  const uint32_t prune = n_epochs > 10 ? n_epochs : default_n_epochs(n);
  if (nnz == 0) {
    *count = 0;
    return HM_OK;
  }
  HM_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  hm_status s = HM_OK;
  DevBufs B(st);
  uint32_t* dr = B.get<uint32_t>(ctx, nnz, &s);
  uint32_t* dc = dr ? B.get<uint32_t>(ctx, nnz, &s) : nullptr;
  float* dw = dc ? B.get<float>(ctx, nnz, &s) : nullptr;
  uint32_t* dh = dw ? B.get<uint32_t>(ctx, nnz, &s) : nullptr;
  uint32_t* dt = dh ? B.get<uint32_t>(ctx, nnz, &s) : nullptr;
  double* de = dt ? B.get<double>(ctx, nnz, &s) : nullptr;
  if (!de) return s;
  HM_HIP(ctx, hipMemcpyAsync(dr, rows, nnz * 4, hipMemcpyHostToDevice, st));
  HM_HIP(ctx, hipMemcpyAsync(dc, cols, nnz * 4, hipMemcpyHostToDevice, st));
  HM_HIP(ctx, hipMemcpyAsync(dw, vals, nnz * 4, hipMemcpyHostToDevice, st));
  HM_TRY(schedule_device(ctx, st, dr, dc, dw, (int64_t)nnz, prune, dh, dt, de, count));
  const uint64_t c = *count;
  if (c) {
    HM_HIP(ctx, hipMemcpyAsync(head, dh, c * 4, hipMemcpyDeviceToHost, st));
    HM_HIP(ctx, hipMemcpyAsync(tail, dt, c * 4, hipMemcpyDeviceToHost, st));
    HM_HIP(ctx, hipMemcpyAsync(epochs_per_sample, de, c * 8, hipMemcpyDeviceToHost, st));
  }
  HM_HIP(ctx, hipStreamSynchronize(st));
  return HM_OK;
}

// Row pointer (local, n+1) and the start of dataset ds's COO in the job's
// graph arrays.
static hm_status job_csr(hm_job* J, uint32_t ds, const uint64_t** indptr, uint64_t* s0) {
  if (!(J->stages & HM_STAGE_FUZZY)) return fail(J->ctx, HM_EINVAL, "the job builds no graph");
  HM_TRY(sym_check(J));
  HM_HIP(J->ctx, hipMemcpy(s0, J->indptr_g + J->dsh[ds].row_off, 8, hipMemcpyDeviceToHost));
  *indptr = J->indptr_l + J->iptr_off_h[ds];
  return HM_OK;
}

hm_status hm_job_connected_components(hm_job* J, uint32_t ds, uint32_t* labels,
                                      uint64_t* n_components) {
  if (!J || ds >= (uint32_t)J->nds || !labels || !n_components) return HM_EINVAL;
  hm_ctx* ctx = J->ctx;
  HM_HIP(ctx, hipSetDevice(ctx->device));
  const uint64_t* iptr = nullptr;
  uint64_t s0 = 0;
  HM_TRY(job_csr(J, ds, &iptr, &s0));
  const int64_t n = J->dsh[ds].n;
  hm_status s = HM_OK;
  DevBufs B(J->stream);
  uint32_t* dout = B.get<uint32_t>(ctx, n, &s);
  if (!dout) return s;
  HM_TRY(components_device(ctx, J->stream, iptr, J->cols + s0, n, dout, n_components));
  HM_HIP(ctx, hipMemcpyAsync(labels, dout, n * 4, hipMemcpyDeviceToHost, J->stream));
  HM_HIP(ctx, hipStreamSynchronize(J->stream));
  return HM_OK;
}

hm_status hm_job_schedule_edges(hm_job* J, uint32_t ds, uint32_t n_epochs, uint32_t* head,
                                uint32_t* tail, double* epochs_per_sample, uint64_t* count) {
  if (!J || ds >= (uint32_t)J->nds || !head || !tail || !epochs_per_sample || !count)
    return HM_EINVAL;
  hm_ctx* ctx = J->ctx;
  HM_HIP(ctx, hipSetDevice(ctx->device));
  const uint64_t* iptr = nullptr;
  uint64_t s0 = 0, s1 = 0;
  HM_TRY(job_csr(J, ds, &iptr, &s0));
  HM_HIP(ctx, hipMemcpy(&s1, iptr + J->dsh[ds].n, 8, hipMemcpyDeviceToHost));
  const int64_t m = (int64_t)s1;  // the dataset's entries (local row pointer's end)
  const uint32_t prune = n_epochs > 10 ? n_epochs : default_n_epochs((uint64_t)J->dsh[ds].n);
  hm_status s = HM_OK;
  DevBufs B(J->stream);
  uint32_t* dh = B.get<uint32_t>(ctx, m, &s);
  uint32_t* dt = dh ? B.get<uint32_t>(ctx, m, &s) : nullptr;
  double* de = dt ? B.get<double>(ctx, m, &s) : nullptr;
  if (!de) return s;
  HM_TRY(schedule_device(ctx, J->stream, J->rows + s0, J->cols + s0, J->vals + s0, m, prune, dh,
                         dt, de, count));
  const uint64_t c = *count;
  if (c) {
    HM_HIP(ctx, hipMemcpyAsync(head, dh, c * 4, hipMemcpyDeviceToHost, J->stream));
    HM_HIP(ctx, hipMemcpyAsync(tail, dt, c * 4, hipMemcpyDeviceToHost, J->stream));
    HM_HIP(ctx, hipMemcpyAsync(epochs_per_sample, de, c * 8, hipMemcpyDeviceToHost, J->stream));
  }
  HM_HIP(ctx, hipStreamSynchronize(J->stream));
  return HM_OK;
}

hm_status hm_job_laplacian_matvec(hm_job* J, uint32_t ds, const double* x, double* y) {
  if (!J || ds >= (uint32_t)J->nds || !x || !y) return HM_EINVAL;
  hm_ctx* ctx = J->ctx;
  HM_HIP(ctx, hipSetDevice(ctx->device));
  const uint64_t* iptr = nullptr;
  uint64_t s0 = 0;
  HM_TRY(job_csr(J, ds, &iptr, &s0));
  const DsHost& h = J->dsh[ds];
  if (!J->lap_inv) HM_TRY(dalloc(J, &J->lap_inv, J->total_rows));
  if (!J->lap_valid) {  // degrees of every dataset's graph of the last run
    for (int i = 0; i < J->nds; ++i) {
      const uint64_t* ip = nullptr;
      uint64_t b0 = 0;
      HM_TRY(job_csr(J, (uint32_t)i, &ip, &b0));
      hipLaunchKernelGGL(lap_degree_kernel, dim3(grid_for(J->dsh[i].n, 256)), dim3(256), 0,
                         J->stream, ip, (const float*)(J->vals + b0), J->dsh[i].n,
                         (double*)nullptr, J->lap_inv + J->dsh[i].row_off);
    }
    HM_LAUNCH(ctx);
    J->lap_valid = true;
  }
  hipLaunchKernelGGL(lap_matvec_kernel, dim3(grid_for(h.n, 256)), dim3(256), 0, J->stream, iptr,
                     (const uint32_t*)(J->cols + s0), (const double*)nullptr,
                     (const float*)(J->vals + s0), (const double*)(J->lap_inv + h.row_off), h.n, x,
                     y);
  HM_LAUNCH(ctx);
  HM_HIP(ctx, hipStreamSynchronize(J->stream));
  return HM_OK;
}

hm_status hm_l2_normalise(hm_ctx* ctx, const float* data, uint64_t n, uint32_t d, float* out) {
  if (!ctx || !data || !out) return fail(ctx, HM_EINVAL, "null argument");
  if (n == 0 || d == 0) return HM_OK;
  HM_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const uint64_t m = n * d;
  float *dx = nullptr, *dy = nullptr;
  HM_HIP(ctx, hipMallocAsync((void**)&dx, m * 4, st));
  if (hipMallocAsync((void**)&dy, m * 4, st) != hipSuccess) {
    (void)hipFreeAsync(dx, st);
    return fail(ctx, HM_ENOMEM, "device allocation failed");
  }
  hm_status s = HM_OK;
  if (hipMemcpyAsync(dx, data, m * 4, hipMemcpyHostToDevice, st) != hipSuccess)
    s = fail(ctx, HM_EDEVICE, "upload failed");
  if (s == HM_OK) {
    hipLaunchKernelGGL(l2_normalise_kernel, dim3(grid_for((int64_t)n, 256)), dim3(256), 0, st,
                       (const float*)dx, (int64_t)n, (int)d, dy);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(out, dy, m * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      s = fail(ctx, HM_EDEVICE, "l2_normalise kernel failed");
  }
  (void)hipFreeAsync(dx, st);
  (void)hipFreeAsync(dy, st);
  return s;
}

hm_status hm_expf_array(hm_ctx* ctx, const float* x, float* y, uint64_t n) {
  if (!ctx || !x || !y) return fail(ctx, HM_EINVAL, "null argument");
  if (n == 0) return HM_OK;
  HM_HIP(ctx, hipSetDevice(ctx->device));
  float *dx = nullptr, *dy = nullptr;
  HM_HIP(ctx, hipMalloc(&dx, n * 4));
  if (hipMalloc(&dy, n * 4) != hipSuccess) {
    (void)hipFree(dx);
    return fail(ctx, HM_ENOMEM, "hipMalloc failed");
  }
  hm_status s = HM_OK;
  if (hipMemcpy(dx, x, n * 4, hipMemcpyHostToDevice) != hipSuccess) s = fail(ctx, HM_EDEVICE, "upload");
  if (s == HM_OK) {
    hipLaunchKernelGGL(expf_kernel, dim3(grid_for((int64_t)n, 256)), dim3(256), 0, ctx->stream,
                       (const float*)dx, dy, (int64_t)n);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess ||
        hipMemcpy(y, dy, n * 4, hipMemcpyDeviceToHost) != hipSuccess)
      s = fail(ctx, HM_EDEVICE, "expf kernel failed");
  }
  (void)hipFree(dx);
  (void)hipFree(dy);
  return s;
}

}  // extern "C"
