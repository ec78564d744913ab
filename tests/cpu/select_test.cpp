// CPU unit test of the sweep's per-lane list compaction and insertion rule
// (holomap_amd/csrc/hm_select.hpp compiled as host code, one lane at a time).
#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>

#include "../../holomap_amd/csrc/hm_select.hpp"

using hm::lane_compact;
using hm::sel_key;

static int fails = 0;
#define CHECK(c, ...)                         \
  do {                                        \
    if (!(c)) {                               \
      if (fails < 10) std::printf(__VA_ARGS__); \
      ++fails;                                \
    }                                         \
  } while (0)

template <int QL, typename J = uint32_t>
void check_compact(std::mt19937_64& rng, int mode, int keep_lo, int keep_hi) {
  const int stride = 3;  // any stride: the kernel uses 256 (lane-minor)
  std::vector<float> kf(QL * stride, 0.f);
  std::vector<J> kj(QL * stride, 0);
  std::uniform_int_distribution<int> cd(keep_hi + 1, QL);
  const int cnt0 = cd(rng);
  std::normal_distribution<float> nd(0.f, 1.f);
  std::vector<uint64_t> keys;
  for (int e = 0; e < cnt0; ++e) {
    float v;
    switch (mode) {
      case 0: v = nd(rng) * 1000.f - 5000.f; break;              // euclid-like
      case 1: v = -1.0f + 1e-4f * nd(rng); break;                // clustered cosine
      case 2: v = (float)(rng() % 3) * 0.5f - 7.0f; break;       // heavy ties
      case 3: v = (rng() % 4 == 0) ? -INFINITY : nd(rng); break; // -inf keys
      case 4: v = (rng() % 2) ? -INFINITY : 3.0f; break;         // only two values
      case 5: v = nd(rng) * 1e30f; break;                         // huge magnitudes
      default: v = std::ldexp(1.0f, -140) * (float)(rng() % 5); break;  // subnormals
    }
    const uint32_t jmod = sizeof(J) == 2 ? 65536u : 100000u;
    uint32_t j = (uint32_t)(rng() % jmod);
    // keep columns distinct, as in the sweep
    while (std::find_if(keys.begin(), keys.end(), [&](uint64_t k) { return (uint32_t)k == j; }) !=
           keys.end())
      j = (uint32_t)(rng() % jmod);
    kf[e * stride] = v;
    kj[e * stride] = (J)j;
    keys.push_back(sel_key(v, j));
  }
  int cnt = cnt0;
  uint64_t tau = ~0ull;
  lane_compact<QL>(kf.data(), kj.data(), stride, cnt, tau, keep_hi, keep_lo, keep_hi);
  std::sort(keys.begin(), keys.end());
  CHECK(cnt >= keep_lo && cnt <= keep_hi, "mode %d: cnt %d outside [%d,%d]\n", mode, cnt, keep_lo,
        keep_hi);
  std::vector<uint64_t> kept;
  for (int e = 0; e < cnt; ++e) kept.push_back(sel_key(kf[e * stride], (uint32_t)kj[e * stride]));
  std::sort(kept.begin(), kept.end());
  bool prefix = true;
  for (int e = 0; e < cnt; ++e) prefix &= kept[e] == keys[e];
  CHECK(prefix, "mode %d: kept entries are not the smallest\n", mode);
  CHECK(tau == keys[cnt - 1], "mode %d: tau is not the largest kept key\n", mode);
}

// The sweep's insertion rule over a stream of keys for one lane: append if
// key < tau; before every group of 16, compact when fewer than 16 slots are
// free; at the end compact to <= keep_hi. The final list must hold the
// stream's keep_lo smallest keys and every key not kept must exceed tau.
template <int QL>
void check_stream(std::mt19937_64& rng, int mode, int m, int keep_lo) {
  const int keep_hi = QL - 16;
  std::vector<float> kf(QL * 256);
  std::vector<uint32_t> kj(QL * 256);
  float* f = kf.data() + 17;  // lane 17 of a 256-lane layout
  uint32_t* jj = kj.data() + 17;
  std::normal_distribution<float> nd(0.f, 1.f);
  std::vector<uint64_t> all;
  int cnt = 0;
  uint64_t tau = ~0ull;
  for (int g = 0; g < m; g += 16) {
    if (cnt > QL - 16) lane_compact<QL>(f, jj, 256, cnt, tau, keep_lo + 4, keep_lo, keep_lo + 4);
    for (int t = 0; t < 16 && g + t < m; ++t) {
      const uint32_t j = (uint32_t)(g + t);
      float v = mode == 0 ? nd(rng) : (mode == 1 ? (float)(rng() % 7) : -(float)j * 0.001f);
      const uint64_t key = sel_key(v, j);
      all.push_back(key);
      if (key < tau) {
        CHECK(cnt < QL, "overflow\n");
        f[cnt * 256] = v;
        jj[cnt * 256] = j;
        ++cnt;
      }
    }
  }
  if (cnt > keep_hi) lane_compact<QL>(f, jj, 256, cnt, tau, keep_hi, keep_lo, keep_hi);
  std::sort(all.begin(), all.end());
  std::vector<uint64_t> kept;
  for (int e = 0; e < cnt; ++e) kept.push_back(sel_key(f[e * 256], jj[e * 256]));
  std::sort(kept.begin(), kept.end());
  const int need = std::min<int>(keep_lo, (int)all.size());
  bool ok = (int)kept.size() >= need;
  for (int e = 0; ok && e < need; ++e) ok = kept[e] == all[e];
  CHECK(ok, "stream mode %d m %d: smallest keys missing\n", mode, m);
  if (tau != ~0ull) {
    for (uint64_t k : all)
      if (!std::binary_search(kept.begin(), kept.end(), k)) CHECK(k > tau, "dropped key <= tau\n");
  }
}

int main() {
  std::mt19937_64 rng(1234);
  for (int rep = 0; rep < 3000; ++rep) {
    const int mode = rep % 7;
    check_compact<48>(rng, mode, 23, 32);
    check_compact<48>(rng, mode, 32, 32);
    check_compact<64>(rng, mode, 40, 48);
    check_compact<64>(rng, mode, 8, 48);
    check_compact<64, uint16_t>(rng, mode, 23, 35);
    check_compact<64, uint16_t>(rng, mode, 48, 48);
  }
  for (int rep = 0; rep < 60; ++rep) {
    check_stream<48>(rng, rep % 3, 50 + rep * 97, 23);
    check_stream<64>(rng, rep % 3, 50 + rep * 131, 45);
  }
  std::printf("%s (%d failures)\n", fails ? "FAIL" : "OK", fails);
  return fails ? 1 : 0;
}
