// scenes.cpp -- deterministic synthetic particle generators (host only).
// xoshiro256** (Blackman & Vigna) seeded through splitmix64, so a workload
// is fixed by its seed on every machine.
#include <stddef.h>
#include <stdint.h>

#include "../../include/aemh.h"

namespace {

struct Xoshiro256ss {
  uint64_t s[4];
  static uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  explicit Xoshiro256ss(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) s[i] = splitmix64(x);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  // uniform in [0, 1) with 24 random bits (exact in float)
  float uniform() { return (float)(next() >> 40) * (1.0f / 16777216.0f); }
};

}  // namespace

extern "C" void aemh_scene_uniform_box(uint64_t seed, uint32_t n, const float lo[3],
                                       const float hi[3], float* pos3) {
  if (!pos3 || !lo || !hi) return;
  Xoshiro256ss rng(seed);
  for (uint32_t i = 0; i < n; i++)
    for (int a = 0; a < 3; a++) pos3[3 * (size_t)i + a] = lo[a] + (hi[a] - lo[a]) * rng.uniform();
}
