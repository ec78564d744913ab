// Per-lane candidate-list compaction of the MFMA sweep (hm_gram.hip).
//
// Host-compilable: on the device HM_ANY(p) is a wave ballot used only to end
// loops early once every lane is done; on the host it is the lane's own
// predicate, so the CPU unit test (tests/cpu/select_test.cpp) runs the very
// same code one lane at a time.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define HM_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define HM_HD inline
#endif
#if defined(__HIP_DEVICE_COMPILE__)
#define HM_ANY(p) (__ballot(p) != 0ull)
#else
#define HM_ANY(p) (p)
#endif

namespace hm {

HM_HD uint32_t sel_ordered(float x) {
  uint32_t u;
  __builtin_memcpy(&u, &x, 4);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

HM_HD uint64_t sel_key(float v, uint32_t j) { return ((uint64_t)sel_ordered(v) << 32) | j; }

// A lane's list: entry e is (kf[e*stride], kj[e*stride]); J is the column
// id type (the sweep stores u16 offsets within its column split). Lanes with
// cnt > act are compacted to a count in [keep_lo, keep_hi] (act >= keep_hi);
// the kept entries are exactly the smallest ones in the u64 key order
// (ordered(value) << 32 | column), and tau becomes the largest kept key.
// Every dropped entry is larger than the new tau.
// The entry loops load 16 entries at a time unconditionally (the whole QL
// region is allocated) and mask by e < cnt afterwards, so the loads of a
// batch are in flight together instead of each waiting behind a guard.
#define HM_SEL_B 16

template <int QL, typename J>
HM_HD void lane_compact(float* kf, J* kj, int stride, int& cnt, uint64_t& tau, int act,
                        int keep_lo, int keep_hi) {
  static_assert(QL % HM_SEL_B == 0, "QL must be a multiple of the load batch");
  const bool active = cnt > act;
  // f(v) = #{entries with value <= v}
  float mx = -INFINITY, mn = INFINITY;
  int ninf = 0;
  for (int e0 = 0; e0 < QL; e0 += HM_SEL_B) {
    float v[HM_SEL_B];
#pragma unroll
    for (int u = 0; u < HM_SEL_B; ++u) v[u] = kf[(e0 + u) * stride];
#pragma unroll
    for (int u = 0; u < HM_SEL_B; ++u) {
      if (e0 + u < cnt) {
        mx = v[u] > mx ? v[u] : mx;
        if (v[u] == -INFINITY) ++ninf;
        else mn = v[u] < mn ? v[u] : mn;
      }
    }
  }
  uint64_t T = ~0ull;  // keep the entries with key <= T
  bool done = !active, tiep = false;
  float tie = INFINITY;
  if (active && ninf >= keep_lo) {  // -inf keys alone fill the window
    tiep = true;
    tie = -INFINITY;
    done = true;
  }
  // bisection on the value; invariant f(lo) < keep_lo, f(hi) > keep_hi
  float lo = mn - (fabsf(mn) * 0x1p-20f + 1e-30f);  // below every finite key
  float hi = mx;
  for (int it = 0; it < 64; ++it) {
    if (!HM_ANY(!done)) break;
    const float mid = lo * 0.5f + hi * 0.5f;
    int c = 0;
    for (int e0 = 0; e0 < QL; e0 += HM_SEL_B) {
      float v[HM_SEL_B];
#pragma unroll
      for (int u = 0; u < HM_SEL_B; ++u) v[u] = kf[(e0 + u) * stride];
#pragma unroll
      for (int u = 0; u < HM_SEL_B; ++u) c += (e0 + u < cnt && v[u] <= mid) ? 1 : 0;
    }
    if (!done) {
      if (!(mid > lo && mid < hi)) {
        // lo and hi are adjacent floats: the ties at value hi straddle the
        // window
        tiep = true;
        tie = hi;
        done = true;
      } else if (c < keep_lo) {
        lo = mid;
      } else if (c > keep_hi) {
        hi = mid;
      } else {
        T = ((uint64_t)sel_ordered(mid) << 32) | 0xffffffffu;
        done = true;
      }
    }
  }
  // ties: keep every value < tie and the smallest columns among value == tie
  if (HM_ANY(active && tiep)) {
    int c_lt = 0;
    for (int e0 = 0; e0 < QL; e0 += HM_SEL_B) {
      float v[HM_SEL_B];
#pragma unroll
      for (int u = 0; u < HM_SEL_B; ++u) v[u] = kf[(e0 + u) * stride];
#pragma unroll
      for (int u = 0; u < HM_SEL_B; ++u) c_lt += (e0 + u < cnt && v[u] < tie) ? 1 : 0;
    }
    const int need = keep_lo - c_lt;
    uint32_t jl = 0, jh = 0xffffffffu;
    for (int it = 0; it < 34; ++it) {
      const bool go = active && tiep && jl < jh;
      if (!HM_ANY(go)) break;
      const uint32_t jm = jl + (jh - jl) / 2;
      int c = 0;
      for (int e0 = 0; e0 < QL; e0 += HM_SEL_B) {
        float v[HM_SEL_B];
        uint32_t jv[HM_SEL_B];
#pragma unroll
        for (int u = 0; u < HM_SEL_B; ++u) {
          v[u] = kf[(e0 + u) * stride];
          jv[u] = (uint32_t)kj[(e0 + u) * stride];
        }
#pragma unroll
        for (int u = 0; u < HM_SEL_B; ++u) c += (e0 + u < cnt && v[u] == tie && jv[u] <= jm) ? 1 : 0;
      }
      if (go) {
        if (c >= need) jh = jm; else jl = jm + 1;
      }
    }
    if (active && tiep) T = ((uint64_t)sel_ordered(tie) << 32) | jl;
  }
  // no convergence (never expected): exact selection in the u64 key order
  const bool fb = active && !done;
  if (HM_ANY(fb)) {
    uint64_t l64 = 0, h64 = ~0ull;
    for (int it = 0; it < 66; ++it) {
      const bool go = fb && l64 < h64;
      if (!HM_ANY(go)) break;
      const uint64_t m64 = l64 + (h64 - l64) / 2;
      int c = 0;
      for (int e0 = 0; e0 < QL; e0 += HM_SEL_B) {
        float v[HM_SEL_B];
        uint32_t jv[HM_SEL_B];
#pragma unroll
        for (int u = 0; u < HM_SEL_B; ++u) {
          v[u] = kf[(e0 + u) * stride];
          jv[u] = (uint32_t)kj[(e0 + u) * stride];
        }
#pragma unroll
        for (int u = 0; u < HM_SEL_B; ++u) c += (e0 + u < cnt && sel_key(v[u], jv[u]) <= m64) ? 1 : 0;
      }
      if (go) {
        if (c >= keep_lo) h64 = m64; else l64 = m64 + 1;
      }
    }
    if (fb) T = l64;
  }
  if (active) {
    int w = 0;
    uint64_t tmax = 0;
    for (int e0 = 0; e0 < QL; e0 += HM_SEL_B) {
      float v[HM_SEL_B];
      J jv[HM_SEL_B];
#pragma unroll
      for (int u = 0; u < HM_SEL_B; ++u) {
        v[u] = kf[(e0 + u) * stride];
        jv[u] = kj[(e0 + u) * stride];
      }
#pragma unroll
      for (int u = 0; u < HM_SEL_B; ++u) {
        const uint64_t key = sel_key(v[u], (uint32_t)jv[u]);
        if (e0 + u < cnt && key <= T) {
          kf[w * stride] = v[u];
          kj[w * stride] = jv[u];
          ++w;
          tmax = key > tmax ? key : tmax;
        }
      }
    }
    cnt = w;
    tau = tmax;
  }
}

}  // namespace hm
