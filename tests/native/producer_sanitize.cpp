// Drives the host-side producer of the C ABI (include/fcs/fcs_producer.h:
// volume recipes, streaming and materialised .csdb encoders, brick
// pyramids) under AddressSanitizer + UndefinedBehaviorSanitizer: every
// recipe at a small edge, random ragged u32/u64 volumes at every brick size,
// the streaming encoder against the materialised one, and invalid
// arguments. Prints "ok" when every result is as expected.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "fcs/fcs_producer.h"

static int fail(const char* what) {
  std::fprintf(stderr, "%s\n", what);
  return 1;
}

int main() {
  std::mt19937_64 rng(12345);
  // recipes 1-5 at a small edge: generate a box, stream-encode, compare with
  // the materialised encode of the same cube
  for (int cfg = 1; cfg <= 5; ++cfg) {
    // the regions recipe is made of 256^3 regions: a 64^3 corner box of it
    const uint32_t edge = cfg == 3 ? 256 : 64;
    fcs_gen_params p;
    if (fcs_gen_preset(cfg, edge, &p) != 0) return fail("preset");
    p.size = edge;
    fcs_gen* g = nullptr;
    if (fcs_gen_create(&p, &g) != 0) return fail("create");
    std::vector<uint32_t> box(64 * 64 * 64);
    if (fcs_gen_box(g, edge - 64, 0, edge - 64, 64, 64, 64, box.data(), 3) != 0) return fail("box");
    (void)fcs_gen_background_fraction(g);
    if (cfg == 3) {
      fcs_gen_destroy(g);
      continue;
    }
    for (uint32_t b : {8u, 16u, 32u}) {
      fcs_encode_opts o{b, 0, 3, 0};
      uint8_t *s1 = nullptr, *s2 = nullptr;
      size_t n1 = 0, n2 = 0;
      if (fcs_encode_generated(g, &o, &s1, &n1) != 0) return fail("encode_generated");
      if (fcs_encode_volume(box.data(), 32, 64, 64, 64, &o, &s2, &n2) != 0) return fail("encode_volume");
      if (n1 != n2 || std::memcmp(s1, s2, n1) != 0) return fail("streamed != materialised");
      fcs_free_buffer(s1);
      fcs_free_buffer(s2);
    }
    fcs_gen_destroy(g);
  }
  // random ragged volumes, u32 and u64, every brick size
  for (int trial = 0; trial < 40; ++trial) {
    const uint32_t b = 2u << (trial % 6);
    const uint32_t W = 1 + rng() % (2 * b), H = 1 + rng() % (2 * b), D = 1 + rng() % (b + 1);
    const uint32_t k = 1 + rng() % 7;
    std::vector<uint64_t> lab(k);
    for (auto& x : lab) x = (trial & 1) ? rng() : (rng() & 0xFFFFFFFFu);
    lab[0] = 0;
    const size_t n = (size_t)W * H * D;
    std::vector<uint64_t> v64(n);
    std::vector<uint32_t> v32(n);
    for (size_t i = 0; i < n; ++i) {
      v64[i] = lab[rng() % k];
      v32[i] = (uint32_t)v64[i];
    }
    fcs_encode_opts o{b, (uint32_t)(trial % 2), 2, 0};
    uint8_t* s = nullptr;
    size_t ns = 0;
    if (fcs_encode_volume(v64.data(), 64, W, H, D, &o, &s, &ns) != 0) return fail("encode u64");
    fcs_free_buffer(s);
    if (fcs_encode_volume(v32.data(), 32, W, H, D, &o, &s, &ns) != 0) return fail("encode u32");
    fcs_free_buffer(s);
    // pyramid of one brick
    std::vector<uint64_t> brick((size_t)b * b * b);
    for (auto& x : brick) x = lab[rng() % k];
    uint32_t L = 0;
    while ((1u << L) < b) ++L;
    std::vector<uint64_t> pyr(((1ull << (3 * (L + 1))) - 1) / 7);
    std::vector<uint8_t> cst(pyr.size());
    if (fcs_brick_pyramid(brick.data(), b, pyr.data(), cst.data()) != 0) return fail("pyramid");
    if (pyr[0] == 0 && k > 1) { /* a majority of zeros is allowed */ }
  }
  // invalid arguments are rejected, not crashed on
  fcs_encode_opts bad{3, 0, 1, 0};
  uint8_t* s = nullptr;
  size_t ns = 0;
  uint32_t one = 7;
  if (fcs_encode_volume(&one, 32, 1, 1, 1, &bad, &s, &ns) == 0) return fail("brick size 3 accepted");
  if (fcs_gen_create(nullptr, nullptr) == 0) return fail("null params accepted");
  fcs_gen_params z;
  std::memset(&z, 0, sizeof z);
  fcs_gen* g = nullptr;
  if (fcs_gen_create(&z, &g) == 0) return fail("empty params accepted");
  std::printf("ok\n");
  return 0;
}
