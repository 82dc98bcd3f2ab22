// Test infrastructure (the oracle): the self-test's fixed-order reduction restated on the CPU,
// following crates/sherd-gpu/src/selftest.rs:405-418 (reduce_terms: the xorshift64 stream, 24 bits
// into [-1, 1), every 997th term times 1e5) and :423-460 (reduce_mirror: `groups` workgroups, each
// a lane-strided accumulate over its block of ceil(n / groups) terms and a 256 -> 1 halving tree,
// then one workgroup over the partials). Pinned by the reference's own figure for 1e7 terms over
// 256 workgroups, 0x49a7230c (README.md:1232).

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle.h"

namespace {

std::vector<float> pass(const std::vector<float> &x, size_t groups) {
  const size_t lanes = 256;
  size_t block = (x.size() + groups - 1) / groups;
  std::vector<float> out(groups);
  for (size_t g = 0; g < groups; ++g) {
    size_t start = g * block, end = std::min(start + block, x.size());
    float s[256];
    for (size_t l = 0; l < lanes; ++l) {
      float acc = 0.0f;
      for (size_t i = start + l; i < end; i += lanes) acc += x[i];
      s[l] = acc;
    }
    for (size_t stride = lanes / 2; stride > 0; stride /= 2)
      for (size_t l = 0; l < stride; ++l) s[l] += s[l + stride];
    out[g] = s[0];
  }
  return out;
}

}  // namespace

extern "C" uint32_t oracle_reduce_bits(size_t n, size_t groups) {
  std::vector<float> x(n);
  uint64_t st = 0x9e3779b97f4a7c15ull;
  for (size_t i = 0; i < n; ++i) {
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    float v = (float)(st >> 40) * (1.0f / 16777216.0f) * 2.0f - 1.0f;
    x[i] = i % 997 == 0 ? v * 1e5f : v;
  }
  float r = pass(pass(x, groups), 1)[0];
  uint32_t b;
  std::memcpy(&b, &r, 4);
  return b;
}
