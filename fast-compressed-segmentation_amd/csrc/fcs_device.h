// Device-side data layout shared by the kernels and the host runtime.
#pragma once
#include <stdint.h>

// One record per brick, bricks in Morton (ID) order. 32 bytes.
struct DevBrick {
  uint64_t str_off;  // byte offset of the rANS stream in the stream blob (16-byte aligned)
  uint64_t nib_off;  // u32-word offset of the decoded nibble stream (4-word aligned)
  uint32_t E;        // rANS byte count
  uint32_t nsym;     // nibble count
  uint32_t pal_off;  // first palette entry
  uint32_t P;        // palette size
};

// Per-brick, per-op-level side table written by the rANS kernel:
// entry l (0 <= l <= L) = nibble index where level l's ops start and the
// palette cursor before them. Entry L = end of the last level.
struct LevelEntry {
  uint32_t nib_start;
  uint32_t ip_base;
};

// Per-brick rANS results: status bits and the number of complete op levels
// (a request at LOD t needs t <= levels_ok).
struct RansResult {
  uint32_t status;
  uint32_t levels_ok;
};

// One decode request. 16 bytes.
struct DecodeReq {
  uint32_t brick;
  uint32_t lod;      // bits 0..7: LOD t; bit 8: write every level 0..t (pyramid layout)
  uint64_t out_off;  // voxel offset in the brick cache pool
};

#define FCS_REQ_PYRAMID 0x100u

// Ops kernel launch geometry and LDS budget.
#define FCS_OPS_THREADS 1024
#define FCS_OPS_WAVES (FCS_OPS_THREADS / 64)
// Largest level held in LDS as 16-bit palette indices (32^3 nodes).
#define FCS_LDS_LEVEL_MAX 32768u
// 16-bit node value: bit 15 set = reference to another node of the same level.
#define FCS_REF16 0x8000u
#define FCS_REF32 0x80000000u
// Palette indices must fit the 15-bit payload.
#define FCS_PALETTE_MAX_LDS 32768u
