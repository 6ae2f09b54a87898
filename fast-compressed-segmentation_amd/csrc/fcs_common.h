// Shared format constants and bit utilities for the .csdb brick codec.
// Used by host code (g++/hipcc host pass) and by device code (hipcc, gfx950).
//
// Format facts (reference: encoder/CSVEncoder/src/types.hpp:18-28,
// fileIO.cpp:15-99, utils.cpp:4-42):
//   * magic 0x43534244, version 1, little-endian u32 header fields
//   * one global 16-symbol rANS model: M = 2^12, L = 2^23, byte-wise renorm
//   * Morton order: x -> bit 0, y -> bit 1, z -> bit 2, 10 bits per axis
//   * op nibble = (op << 1) | stop; op 0..6 = Rp, Rx, Ry, Rz, Pa, P0, Pd;
//     Pd is followed by a delta nibble.
#pragma once
#include <stdint.h>

#include "fcs/fcs.h"

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define FCS_HD __host__ __device__ __forceinline__
#else
#define FCS_HD static inline
#endif

#define FCS_MAGIC 0x43534244u
#define FCS_VERSION 1u
#define FCS_RANS_L (1u << 23)
#define FCS_RANS_LOGM 12u
#define FCS_RANS_M (1u << 12)

enum fcs_op {
  FCS_OP_RP = 0,
  FCS_OP_RX = 1,
  FCS_OP_RY = 2,
  FCS_OP_RZ = 3,
  FCS_OP_PA = 4,
  FCS_OP_P0 = 5,
  FCS_OP_PD = 6,
};

FCS_HD uint32_t fcs_part1by2(uint32_t n) {
  n &= 0x3ffu;
  n = (n | (n << 16)) & 0x30000ffu;
  n = (n | (n << 8)) & 0x300f00fu;
  n = (n | (n << 4)) & 0x30c30c3u;
  n = (n | (n << 2)) & 0x9249249u;
  return n;
}

FCS_HD uint32_t fcs_compact1by2(uint32_t n) {
  n &= 0x9249249u;
  n = (n ^ (n >> 2)) & 0x30c30c3u;
  n = (n ^ (n >> 4)) & 0x300f00fu;
  n = (n ^ (n >> 8)) & 0x30000ffu;
  n = (n ^ (n >> 16)) & 0x3ffu;
  return n;
}

FCS_HD uint32_t fcs_morton3(uint32_t x, uint32_t y, uint32_t z) {
  return fcs_part1by2(x) | (fcs_part1by2(y) << 1) | (fcs_part1by2(z) << 2);
}

FCS_HD uint32_t fcs_unmorton_x(uint32_t m) { return fcs_compact1by2(m); }
FCS_HD uint32_t fcs_unmorton_y(uint32_t m) { return fcs_compact1by2(m >> 1); }
FCS_HD uint32_t fcs_unmorton_z(uint32_t m) { return fcs_compact1by2(m >> 2); }

// log2 of a power of two (brick edge b = 2^L).
FCS_HD uint32_t fcs_log2u(uint32_t v) {
  uint32_t r = 0;
  while ((1u << r) < v) ++r;
  return r;
}

// Offset (in voxels) of level t inside a brick's LOD pyramid laid out
// coarse-first: level 0 (1 voxel), level 1 (8), ..., level t at (8^t - 1) / 7.
FCS_HD uint64_t fcs_pyramid_level_offset(uint32_t t) {
  return ((1ull << (3u * t)) - 1ull) / 7ull;
}
