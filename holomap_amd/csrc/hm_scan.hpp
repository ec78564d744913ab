// Wave scans and the single-workgroup scan body (device code shared by the
// scan kernels and the kernels that fuse a flag compaction with the work
// they feed).
#pragma once
#include "hm_internal.hpp"

namespace hm {

// Lane-order inclusive scan of a wave with DPP: row shifts by 1, 2, 4, 8
// scan each row of 16 lanes, then two row broadcasts (lane 15 into rows 1
// and 3, lane 31 into rows 2 and 3) join the rows. Six dependent VALU adds,
// no LDS round trips. Lanes whose DPP source is outside the row, or whose
// row is masked off, add the `old` operand, 0.
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t x) {
  return ((uint64_t)dpp_u32<CTRL, ROWS>((uint32_t)(x >> 32)) << 32) |
         dpp_u32<CTRL, ROWS>((uint32_t)x);
}
template <int CTRL, int ROWS, typename T>
__device__ __forceinline__ T dpp_t(T x) {
  if constexpr (sizeof(T) == 8) return dpp_u64<CTRL, ROWS>(x);
  else return dpp_u32<CTRL, ROWS>(x);
}
template <typename T>
__device__ __forceinline__ T wave_scan_incl(T x) {
  x += dpp_t<0x111, 0xf>(x);  // row_shr:1
  x += dpp_t<0x112, 0xf>(x);  // row_shr:2
  x += dpp_t<0x114, 0xf>(x);  // row_shr:4
  x += dpp_t<0x118, 0xf>(x);  // row_shr:8
  x += dpp_t<0x142, 0xa>(x);  // row_bcast:15 into rows 1, 3
  x += dpp_t<0x143, 0xc>(x);  // row_bcast:31 into rows 2, 3
  return x;
}
template <typename T>
__device__ __forceinline__ T wave_lane63(T x) {
  if constexpr (sizeof(T) == 8)
    return ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), 63) << 32) |
           (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, 63);
  else
    return (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}

// Single-workgroup exclusive scan for short arrays: NT threads, at most
// NT * 4 * MS elements (the runtime's kScanSmall). Wave w owns the contiguous
// segment [w*seg, w*seg + seg), seg a multiple of 256, and loads all of it
// at once into registers: step j of a lane is the quad of 4 consecutive
// elements at 4 * (64 j + lane) (at most MS steps), so the scan costs one
// memory round trip. A quad is scanned serially in its lane, the quad sums
// of a step by one DPP wave scan (the steps are independent, so their scans
// overlap); the step totals give each step's offset in the segment and the
// segment sums each wave's base. With `list` set it instead compacts:
// list[pos] = i for every i with in[i] != 0 (in is a 0/1 flag array), pos
// = the exclusive scan, and *count = the number of flags.
template <typename T>
struct Quad {
  T x[4];
};
template <typename T>
__device__ __forceinline__ Quad<T> load_quad(const T* __restrict__ p, int64_t i, int64_t e,
                                             bool vec) {
  Quad<T> q;
  if (vec && i + 3 < e) {
    if constexpr (sizeof(T) == 4) {
      const uint4 u = *(const uint4*)(p + i);
      q.x[0] = u.x; q.x[1] = u.y; q.x[2] = u.z; q.x[3] = u.w;
    } else {
      const ulonglong2 a = *(const ulonglong2*)(p + i), b = *(const ulonglong2*)(p + i + 2);
      q.x[0] = a.x; q.x[1] = a.y; q.x[2] = b.x; q.x[3] = b.y;
    }
  } else {
#pragma unroll
    for (int c = 0; c < 4; ++c) q.x[c] = i + c < e ? p[i + c] : (T)0;
  }
  return q;
}

template <typename T, int NT, int MS>
__device__ __forceinline__ void scan_small_body(const T* __restrict__ in, int64_t n,
                                                T* __restrict__ out, T* __restrict__ total,
                                                uint32_t* __restrict__ list,
                                                int* __restrict__ count) {
  constexpr int NW = NT / 64;  // waves
  __shared__ T wsum[NW];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t seg = (n + NW * 256 - 1) / (NW * 256) * 256;  // <= MS * 256 for n <= NT * 4 * MS
  const int64_t b = wave * seg < n ? wave * seg : n;
  const int64_t e = b + seg < n ? b + seg : n;
  const int nsteps = (int)(seg / 256);  // uniform
  const bool vec = (((uintptr_t)in | (list ? (uintptr_t)0 : (uintptr_t)out)) & 15) == 0;
  Quad<T> v[MS];
#pragma unroll
  for (int j = 0; j < MS; ++j)
    if (j < nsteps) v[j] = load_quad(in, b + 4 * (64 * j + lane), e, vec);
  // v[j] <- each element's exclusive scan plus the step's offset in the segment
  uint64_t flags[(MS + 15) / 16] = {};
  T run = 0;
#pragma unroll
  for (int j = 0; j < MS; ++j) {
    if (j < nsteps) {
      T pre[4];
      T t = 0;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        pre[c] = t;
        t += v[j].x[c];
        if (list) flags[(4 * j + c) >> 6] |= (uint64_t)(v[j].x[c] != 0) << ((4 * j + c) & 63);
      }
      const T x = wave_scan_incl(t);
      const T q0 = run + x - t;  // the quad's exclusive offset
#pragma unroll
      for (int c = 0; c < 4; ++c) v[j].x[c] = q0 + pre[c];
      run += wave_lane63(x);
    }
  }
  if (lane == 0) wsum[wave] = run;
  __syncthreads();
  T base = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    const T x = wsum[w];
    base += w < wave ? x : (T)0;
    tot += x;
  }
#pragma unroll
  for (int j = 0; j < MS; ++j) {
    if (j < nsteps) {
      const int64_t i = b + 4 * (64 * j + lane);
      if (list) {
#pragma unroll
        for (int c = 0; c < 4; ++c)
          if (i + c < e && ((flags[(4 * j + c) >> 6] >> ((4 * j + c) & 63)) & 1))
            list[base + v[j].x[c]] = (uint32_t)(i + c);
      } else if (vec && i + 3 < e) {
        if constexpr (sizeof(T) == 4) {
          *(uint4*)(out + i) =
              make_uint4(base + v[j].x[0], base + v[j].x[1], base + v[j].x[2], base + v[j].x[3]);
        } else {
          *(ulonglong2*)(out + i) = make_ulonglong2(base + v[j].x[0], base + v[j].x[1]);
          *(ulonglong2*)(out + i + 2) = make_ulonglong2(base + v[j].x[2], base + v[j].x[3]);
        }
      } else {
#pragma unroll
        for (int c = 0; c < 4; ++c)
          if (i + c < e) out[i + c] = base + v[j].x[c];
      }
    }
  }
  if (tid == 0) {
    if (list) *count = (int)tot;
    else if (total) *total = tot;
  }
}

}  // namespace hm
