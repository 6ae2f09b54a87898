// ref_harness.cpp — C entry points over the reference encoder/decoder,
// compiled together with the reference sources under
// /root/reference/encoder/CSVEncoder/src by oracle/Makefile into
// oracle/_ref/libref.so. TEST INFRASTRUCTURE ONLY (the checker).
//
// Calls: splitGridIntoBricks<b> (brick.hpp:133-194), compressDataset
// (brickEncoding.hpp:466-531), saveDatasetToFile / loadDatasetFromFile
// (fileIO.cpp:15-183), decodeRansStream (brickDecoding.cpp:153-188) and
// decodeBrick<b> (brickDecoding.hpp:174-242).
#include "brickDecoding.hpp"
#include "brickEncoding.hpp"
#include "fileIO.hpp"

#include <atomic>
#include <cstring>
#include <exception>
#include <string>
#include <thread>
#include <vector>

namespace {

template <size_t b>
int encode_impl(const uint64_t* labels, uint32_t W, uint32_t H, uint32_t D, const char* path,
                uint64_t* pyr_out) {
  NpyArray arr;
  arr.shape = {W, H, D};
  arr.data.assign(labels, labels + (size_t)W * H * D);
  // the reference accumulates model counts into a process-global table
  // (brickEncoding.cpp:7); each call here encodes a fresh dataset
  Encoding::interiorFreqTable.fill(0);
  auto bricks = splitGridIntoBricks<b>(arr);
  auto ds = Encoding::compressDataset<b>(bricks);
  saveDatasetToFile(ds, path);
  if (pyr_out) {
    constexpr size_t levels = Brick<b>::Levels;  // L + 1
    size_t per = 0;
    for (size_t t = 0; t < levels; ++t) per += size_t(1) << (3 * t);
    for (size_t i = 0; i < bricks.size(); ++i) {
      uint64_t* o = pyr_out + per * i;
      for (size_t t = 0; t < levels; ++t) {
        const auto& lv = bricks[i].coarser(t);
        for (size_t k = 0; k < lv.size(); ++k) *o++ = lv[k].label;
      }
    }
  }
  return (int)bricks.size();
}

Encoding::CompressedDataset g_ds;
bool g_loaded = false;

template <size_t b>
int decode_impl(uint32_t idx, uint64_t* out) {
  auto brick = g_ds.bricks[idx];
  brick.encodedData = decodeRansStream(brick, g_ds.model);
  auto v = decodeBrick<b>(brick, Brick<b>::Levels - 1);
  std::memcpy(out, v.data(), v.size() * sizeof(uint64_t));
  return (int)v.size();
}

}  // namespace

extern "C" {

// Encode a flat x-fastest u64 volume with brick edge b; writes the .csbd file
// and (if pyr_out) every brick's pyramid, coarse-first, bricks in file order.
int ref_encode(const uint64_t* labels, uint32_t W, uint32_t H, uint32_t D, uint32_t b,
               const char* path, uint64_t* pyr_out) {
  try {
    switch (b) {
      case 2: return encode_impl<2>(labels, W, H, D, path, pyr_out);
      case 4: return encode_impl<4>(labels, W, H, D, path, pyr_out);
      case 8: return encode_impl<8>(labels, W, H, D, path, pyr_out);
      case 16: return encode_impl<16>(labels, W, H, D, path, pyr_out);
      case 32: return encode_impl<32>(labels, W, H, D, path, pyr_out);
      case 64: return encode_impl<64>(labels, W, H, D, path, pyr_out);
      default: return -1;
    }
  } catch (const std::exception&) {
    return -2;
  }
}

int ref_load(const char* path) {
  try {
    g_ds = loadDatasetFromFile(path);
    g_loaded = true;
    return (int)g_ds.bricks.size();
  } catch (const std::exception&) {
    g_loaded = false;
    return -2;
  }
}

// Brick `idx` in file order: rANS-decoded nibble bytes; returns the count.
int ref_rans(uint32_t idx, uint8_t* out) {
  if (!g_loaded || idx >= g_ds.bricks.size()) return -1;
  auto v = decodeRansStream(g_ds.bricks[idx], g_ds.model);
  std::memcpy(out, v.data(), v.size());
  return (int)v.size();
}

uint32_t ref_brick_id(uint32_t idx) { return g_loaded && idx < g_ds.bricks.size() ? g_ds.bricks[idx].ID : ~0u; }

// Finest-LOD decode of brick `idx` (file order); -2 if the reference throws
// (e.g. a constant brick: its root stop bit is ignored).
int ref_decode_finest(uint32_t idx, uint32_t b, uint64_t* out) {
  if (!g_loaded || idx >= g_ds.bricks.size()) return -1;
  try {
    switch (b) {
      case 2: return decode_impl<2>(idx, out);
      case 4: return decode_impl<4>(idx, out);
      case 8: return decode_impl<8>(idx, out);
      case 16: return decode_impl<16>(idx, out);
      case 32: return decode_impl<32>(idx, out);
      case 64: return decode_impl<64>(idx, out);
      default: return -1;
    }
  } catch (const std::exception&) {
    return -2;
  }
}

// CPU baseline: decodeRansStream + decodeBrick<b> at the finest LOD for the
// bricks idx[0..n) (file order) on `threads` std::threads, labels discarded
// after a checksum. Returns the number of bricks decoded; bricks the
// reference rejects (it throws on a set root stop bit) are counted in
// *rejected and skipped.
int64_t ref_decode_finest_batch(const uint32_t* idx, uint32_t n, uint32_t b, int threads, uint64_t* checksum,
                                uint32_t* rejected) {
  if (!g_loaded || threads < 1) return -1;
  std::atomic<uint32_t> next{0}, rej{0};
  std::atomic<int64_t> done{0};
  std::atomic<uint64_t> sum{0};
  auto work = [&] {
    std::vector<uint64_t> out((size_t)b * b * b);
    uint64_t local = 0;
    for (uint32_t k; (k = next.fetch_add(1)) < n;) {
      const int r = ref_decode_finest(idx[k], b, out.data());
      if (r < 0) { rej.fetch_add(1); continue; }
      for (uint64_t x : out) local += x;
      done.fetch_add(1);
    }
    sum.fetch_add(local);
  };
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) pool.emplace_back(work);
  for (auto& th : pool) th.join();
  if (checksum) *checksum = sum.load();
  if (rejected) *rejected = rej.load();
  return done.load();
}

}  // extern "C"
