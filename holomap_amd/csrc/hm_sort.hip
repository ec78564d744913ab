// Deterministic, atomic-free building blocks: exclusive prefix scan
// (three-phase: tile sums, single-workgroup scan of the sums, tile scan +
// offset) for u32 and u64, flag compaction, fills and sums.
#include "hm_scan.hpp"

namespace hm {

#define SC_TILE 4096  // elements per scan tile (256 threads x 16)

template <typename T>
__global__ __launch_bounds__(256) void scan_tile_sums(const T* __restrict__ in, int64_t n,
                                                      T* __restrict__ sums) {
  __shared__ T red[256];
  const int64_t base = (int64_t)blockIdx.x * SC_TILE + threadIdx.x * 16;
  T s = 0;
#pragma unroll
  for (int e = 0; e < 16; ++e) s += (base + e < n) ? in[base + e] : (T)0;
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = red[0];
}

// Exclusive scan of the tile sums in place; also writes the grand total.
template <typename T>
__global__ __launch_bounds__(1024) void scan_sums(T* __restrict__ sums, int64_t ntiles,
                                                  T* __restrict__ total) {
  __shared__ T part[1024];
  __shared__ T carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t b = 0; b < ntiles; b += 1024) {
    const int64_t idx = b + threadIdx.x;
    const T v = idx < ntiles ? sums[idx] : (T)0;
    part[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {  // Hillis-Steele inclusive scan
      T add = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : (T)0;
      __syncthreads();
      part[threadIdx.x] += add;
      __syncthreads();
    }
    const T incl = part[threadIdx.x];
    const T c = carry;
    if (idx < ntiles) sums[idx] = c + incl - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry = c + incl;
    __syncthreads();
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

template <typename T>
__global__ __launch_bounds__(256) void scan_tiles(const T* __restrict__ in, int64_t n,
                                                  const T* __restrict__ offs, T* __restrict__ out) {
  __shared__ T part[256];
  const int64_t base = (int64_t)blockIdx.x * SC_TILE + threadIdx.x * 16;
  T v[16];
  T s = 0;
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    v[e] = (base + e < n) ? in[base + e] : (T)0;
    s += v[e];
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    T add = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : (T)0;
    __syncthreads();
    part[threadIdx.x] += add;
    __syncthreads();
  }
  T run = offs[blockIdx.x] + part[threadIdx.x] - s;
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    if (base + e < n) out[base + e] = run;
    run += v[e];
  }
}

// LIST: the flag-compaction form may be asked for (list non-null); without
// it the kernel is a plain scan and keeps no registers for the flags.
template <typename T, int NT, int MS, bool LIST>
__global__ __launch_bounds__(NT) void scan_small(const T* __restrict__ in, int64_t n,
                                                 T* __restrict__ out, T* __restrict__ total,
                                                 uint32_t* __restrict__ list,
                                                 int* __restrict__ count) {
  scan_small_body<T, NT, MS>(in, n, out, total, LIST ? list : nullptr, LIST ? count : nullptr);
}

template __global__ void scan_small<uint32_t, 512, 32, true>(const uint32_t*, int64_t, uint32_t*,
                                                             uint32_t*, uint32_t*, int*);
template __global__ void scan_small<uint32_t, 1024, 16, false>(const uint32_t*, int64_t,
                                                               uint32_t*, uint32_t*, uint32_t*,
                                                               int*);
template __global__ void scan_small<uint64_t, 1024, 8, false>(const uint64_t*, int64_t, uint64_t*,
                                                              uint64_t*, uint32_t*, int*);
template __global__ void scan_tile_sums<uint32_t>(const uint32_t*, int64_t, uint32_t*);
template __global__ void scan_tile_sums<uint64_t>(const uint64_t*, int64_t, uint64_t*);
template __global__ void scan_sums<uint32_t>(uint32_t*, int64_t, uint32_t*);
template __global__ void scan_sums<uint64_t>(uint64_t*, int64_t, uint64_t*);
template __global__ void scan_tiles<uint32_t>(const uint32_t*, int64_t, const uint32_t*, uint32_t*);
template __global__ void scan_tiles<uint64_t>(const uint64_t*, int64_t, const uint64_t*, uint64_t*);

// lower_bound of every value v in [0, nv) in a sorted key array
__global__ void lower_bounds(const uint32_t* __restrict__ sorted, int64_t n, int64_t nv,
                             uint32_t* __restrict__ out) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v > nv) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)sorted[mid] < v) lo = mid + 1; else hi = mid;
  }
  out[v] = (uint32_t)lo;
}

// list[pos[i]] = i for flagged rows (pos = exclusive scan of flag); count.
__global__ void compact_kernel(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                               int64_t n, uint32_t* __restrict__ list, int* __restrict__ count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (flag[i]) list[pos[i]] = (uint32_t)i;
  if (i == n - 1) *count = (int)(pos[i] + flag[i]);
}

__global__ void fill_u32(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Sum of a u32 array in two launches (a fixed order: the same total on every
// run, no atomics): per-workgroup partial sums, then one workgroup adds them.
__device__ __forceinline__ uint64_t block_sum_u64(uint64_t v) {
  __shared__ uint64_t red[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += shfl_xor_u64(v, o);
  if (lane == 0) red[wave] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}
__global__ __launch_bounds__(256) void sum_u32_part_kernel(const uint32_t* __restrict__ a,
                                                           int64_t n, uint64_t* __restrict__ part) {
  uint64_t s = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    s += a[i];
  const uint64_t t = block_sum_u64(s);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}
__global__ __launch_bounds__(256) void sum_u64_kernel(const uint64_t* __restrict__ part, int np,
                                                      uint64_t* __restrict__ out) {
  uint64_t s = 0;
  for (int i = threadIdx.x; i < np; i += 256) s += part[i];
  const uint64_t t = block_sum_u64(s);
  if (threadIdx.x == 0) *out = t;
}

}  // namespace hm
