// Kernel argument blocks and launchers (host-visible declarations).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fcs_device.h"

// rANS mapping of bricks to the machine
#define FCS_RANS_AUTO 0
#define FCS_RANS_LANE 1  // one lane per brick
#define FCS_RANS_WAVE 2  // one wave per brick, scalar-cache decode table
#define FCS_RANS_WAVE_BALLOT 3  // one wave per brick, ballot over a per-lane CDF
#define FCS_RANS_WAVE_LDS 4  // one wave per brick, decode table in LDS
#define FCS_RANS_WIN 5  // one wave per brick, scalar byte window, scalar-cache decode table
#define FCS_RANS_WIN_BALLOT 6  // one wave per brick, scalar byte window, ballot over a per-lane CDF

struct RansArgs {
  const DevBrick* bricks;
  const uint8_t* stream;
  const uint32_t* list;  // nullable: bricks 0..nlist-1
  uint32_t nlist;
  uint32_t L;
  uint32_t freq[16];
  uint32_t cum[17];
  uint32_t* nib;
  RansResult* res;
  const uint32_t* table;  // nullable: x -> (symbol << 25) | (freq << 12) | (x - cum)
};

template <typename LT>
struct OpsArgs {
  const DevBrick* bricks;
  const uint32_t* nib;
  const LT* palette;
  const DecodeReq* reqs;
  LT* out;
  uint32_t* status;
  uint32_t* counter;  // work counter (reset by the launcher)
  uint32_t nreq;      // requests (with nreq_dev: the capacity of the request list)
  uint32_t nbricks;
  uint32_t L;
  uint32_t head;      // set by the launcher: a head pass left levels 1..3 in the outputs
  const uint32_t* nreq_dev;  // nullable: the request count, read from device memory (a device-built queue)
};

// per-brick frame metadata for the marcher
#define FCS_META_LOD_MASK 0xFFu
#define FCS_META_EMPTY 0x100u
#define FCS_META_RESIDENT 0x200u

struct MarchCam {
  float position[3];
  float tan_half_fov;
  float forward[3];
  float ortho_half_height;
  float right[3];
  float pad1;
  float up[3];
  float pad2;
};

template <typename LT>
struct MarchArgs {
  MarchCam cam;
  const MarchCam* cam_dev;  // nullable: when set, the camera is read from device memory instead of `cam`
  const uint32_t* meta;   // per brick: lod | EMPTY | RESIDENT
  const uint64_t* offs;   // per brick: cache voxel offset of its labels at `lod`
  const LT* pal0;         // per brick: palette[0]
  const LT* cache;
  LT* out_label;
  uint64_t* out_voxel;
  float* out_t;
  uint32_t* visited;
  uint32_t* misses;
  uint32_t w, h;
  float wf, hf, aspect, grid;
  uint32_t brick_size, bricks_per_axis;
  // composited, shaded output (the RGBA form of the march; zero = first hit only)
  uint32_t* out_rgba;
  const uint64_t* tf_keys;  // transfer function hash table: label (0 = empty slot) ...
  const uint32_t* tf_rgba;  // ... -> RGBA8
  uint32_t tf_mask;         // slots - 1 (a power of two minus one)
  float light[3];           // light direction (towards the light, not normalised)
  float ambient, diffuse;
  const uint32_t* run_if;   // nullable: the launch does nothing when *run_if == 0
  // A march that re-traces only the rays that missed (all nullable):
  uint32_t* pix_miss;       // per pixel: 1 when the ray entered a non-resident brick
  const uint32_t* pix_only; // per pixel: trace only where this is non-zero (others keep their outputs)
  uint32_t* needed;         // per brick: entered by a ray after its first miss (`visited` then holds
                            // only bricks entered up to a ray's first miss, which a full march enters too)
  // nullable: what the rays read, summed -- samples, non-empty bricks entered,
  // empty bricks tested -- each over FCS_MARCH_STATS_SLOTS slots (cleared by
  // the caller), for the march's algorithmic bytes
  unsigned long long* stats;
};
#define FCS_MARCH_STATS_SLOTS 1024u

// Transfer function hash slot of a label (SplitMix64 finaliser).
__host__ __device__ __forceinline__ uint32_t fcs_tf_slot(uint64_t label, uint32_t mask) {
  uint64_t z = label;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (uint32_t)z & mask;
}

// Device residency word of a brick: cache voxel offset (bits 0-47), resident
// LOD (bits 48-55, 0xFF = not resident), pyramid layout (bit 56), empty (bit 57).
#define FCS_RES_OFF_MASK 0xFFFFFFFFFFFFull
#define FCS_RES_LOD_SHIFT 48
#define FCS_RES_NONE 0xFFull
#define FCS_RES_PYRAMID (1ull << 56)
#define FCS_RES_EMPTY (1ull << 57)

// Per-frame brick metadata on the device: the LOD rule of camera.js:103-137
// (float64, same operation order as the host rule) or a fixed LOD per brick,
// then meta/offs as the marcher reads them.
struct FrameSetupArgs {
  const uint64_t* res;     // per brick residency word
  const MarchCam* cam;     // device camera
  const uint8_t* lod_fixed;  // nullable: per-brick LOD instead of the rule
  uint32_t* meta;
  uint64_t* offs;
  uint32_t nb, b, L, w, h;
  // On-demand decode (a device-owned direct-mapped cache; all nullable):
  // every non-empty brick inside the view frustum that is not resident at its
  // LOD gets a request (brick, LOD, slot brick * stride) in `queue`, its
  // residency word in `res_w` is set to that LOD and it is marked resident
  // for the march that follows the decode.
  uint64_t* res_w;
  DecodeReq* queue;
  uint32_t* qcount;
  uint64_t stride;
};

// Requests for the bricks a march entered but found not resident at their
// LOD (visited && !resident, non-empty, LOD > 0), when *misses != 0: queued,
// marked resident in meta/offs and in the residency table (direct slots).
struct MissEnqueueArgs {
  const uint32_t* misses;
  const uint32_t* visited;
  const uint32_t* needed;  // nullable: bricks entered after a ray's first miss
  uint32_t* meta;
  uint64_t* offs;
  uint64_t* res_w;
  DecodeReq* queue;
  uint32_t* qcount;
  uint32_t nb;
  uint64_t stride;
};

// After a device-queued decode: requests that failed are OR-ed into *err and
// their bricks made non-resident again (residency table and meta).
struct QueueSettleArgs {
  const DecodeReq* queue;
  const uint32_t* qcount;
  const uint32_t* status;
  uint64_t* res_w;
  uint32_t* meta;     // nullable
  uint32_t* err;
  uint32_t cap;
};

// Prefetch candidates from the network's probabilities: non-empty bricks
// with p >= threshold whose LOD by the rule at the look-ahead camera is > 0
// and that are not resident at it; at most `cap` of them, the most probable
// first (probabilities bucketed to 1/1024), queued for a decode into their
// direct slots and marked resident.
struct PrefetchArgs {
  const float* probs;
  const MarchCam* cam;  // look-ahead camera (device)
  uint64_t* res_w;
  DecodeReq* queue;
  uint32_t* qcount;
  uint32_t* hist;       // 1024 + 4 counters, cleared by the launcher
  float threshold;
  uint32_t cap, nb, b, L, w, h;
  uint64_t stride;
};

hipError_t fcs_launch_frame_setup(const FrameSetupArgs& a, hipStream_t s, bool frustum = false);
hipError_t fcs_launch_miss_enqueue(const MissEnqueueArgs& a, hipStream_t s);
hipError_t fcs_launch_queue_settle(const QueueSettleArgs& a, hipStream_t s);
hipError_t fcs_launch_prefetch_enqueue(const PrefetchArgs& a, hipStream_t s);
// zero p[0..n) when *flag != 0
hipError_t fcs_launch_zero_if(uint32_t* p, uint64_t n, const uint32_t* flag, hipStream_t s);
hipError_t fcs_launch_rans(const RansArgs& a, int mode, hipStream_t s);
size_t fcs_ops_lds_bytes(uint32_t L);
// Op decoder LDS splits for bricks up to 32^3 (64^3 bricks have one).
#define FCS_OPS_3WG 0  // ~52 KB LDS, three workgroups per CU
#define FCS_OPS_4WG 1  // ~37 KB LDS, four workgroups per CU
// launches of at least this many requests decode their first levels in a head
// pass (head: FCS_OPS_HEAD_AUTO / _ON / _OFF)
#define FCS_OPS_HEAD_MIN 1024u
// Persistent launch: min(nreq, resident workgroups per CU x num_cu) workgroups.
template <typename LT>
hipError_t fcs_launch_ops(const OpsArgs<LT>& a, uint32_t num_cu, hipStream_t s, uint32_t* grid_out = nullptr,
                          int split = FCS_OPS_3WG, int head = 0);
// The launch configuration's one-time set-up (LDS attribute, occupancy
// query), done before a stream capture that launches the op decoder.
template <typename LT>
hipError_t fcs_ops_prepare(uint32_t L, int split);
// First-hit march; with a.out_rgba set, the composited shaded march that also
// writes the first-hit outputs.
template <typename LT>
hipError_t fcs_launch_march(const MarchArgs<LT>& a, hipStream_t s);
hipError_t fcs_launch_zero(uint32_t* p, uint64_t n, hipStream_t s);
template <typename LT>
hipError_t fcs_launch_frame_checksum(const LT* label, const uint64_t* voxel, uint64_t n, unsigned long long* out,
                                     hipStream_t s);
