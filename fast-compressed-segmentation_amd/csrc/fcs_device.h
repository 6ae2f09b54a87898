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

// Per-brick rANS result: FCS_ST_* status bits.
struct RansResult {
  uint32_t status;
  uint32_t pad;
};

// One decode request. 16 bytes.
struct DecodeReq {
  uint32_t brick;
  uint32_t lod;      // bits 0..7: LOD t; bit 8: write every level 0..t (pyramid layout)
  uint64_t out_off;  // voxel offset in the brick cache pool
};

#define FCS_REQ_PYRAMID 0x100u

// Ops kernel launch geometry and LDS budget.
#define FCS_OPS_THREADS 256
#define FCS_OPS_WAVES (FCS_OPS_THREADS / 64)
// 16-bit node value: bit 15 set = reference to another node of the same level.
#define FCS_REF16 0x8000u
#define FCS_REF32 0x80000000u
// Palette indices must fit the 15-bit payload.
#define FCS_PALETTE_MAX_LDS 32768u
