/* fcs.h — C ABI of the MI355X brick decoder and DDA ray marcher.
 *
 * The reference exposes this path as a JavaScript module API plus WGSL bind
 * groups (web/app.js:9-12). Each entry point below replaces one of them:
 *
 *   fcs_load_csdb         <- loadDatasetFromArrayBuffer (web/scripts/datasetLoader.js:39-145)
 *                            and loadDatasetFromFile (encoder/CSVEncoder/src/fileIO.cpp:101-183)
 *   fcs_decode_rans       <- decodeAllRansBricks (web/scripts/decode.js:17-26),
 *                            decodeRansStream (encoder/CSVEncoder/src/brickDecoding.cpp:153-188)
 *   fcs_request_bricks    <- the decode dispatch of one frame: updateWorkQueue + runComputePass
 *                            (web/scripts/cacheProcessing.js:303-322, web/scripts/gpuPasses.js:15-30)
 *                            over the decodeBrick kernel (web/shaders/decode.wgsl:344-508)
 *   fcs_decode_all        <- initializeCache(forceAll) (web/scripts/renderLoop.js:134-181)
 *   fcs_decode_volume     <- decodeAllRansBricks followed by initializeCache(forceAll) (app.js:75-82 then
 *                            renderLoop.js:134-181), both stages in one call
 *   fcs_read_brick        <- decodeBrick<b> (encoder/CSVEncoder/src/brickDecoding.hpp:174-242)
 *   fcs_render            <- frame(): LOD update + decode + rayTrace pass
 *                            (web/scripts/renderLoop.js:212-359, web/shaders/rayTrace.wgsl:565-906)
 *   fcs_brick_lods        <- calculateAllBrickLODs / calculateBrickLOD (web/scripts/camera.js:103-201)
 *   fcs_camera_orbit      <- OrbitController.apply + Camera.getUniformData
 *                            (web/scripts/orbitController.js:100-123, web/scripts/camera.js:61-77)
 *
 * Conventions
 *   - Every function returns >= 0 on success and a negative FCS_E_* code on
 *     error; fcs_last_error() then holds a message. No exceptions cross.
 *   - The caller owns every host buffer; the library copies inputs.
 *   - Device memory is owned by the context / volume.
 *   - One context = one HIP stream; a context is not thread-safe. Separate
 *     contexts are independent (one per serving session).
 *   - Labels are returned as uint32 (label_bits 32) or uint64 (label_bits 64).
 *   - Per-request status words use the FCS_ST_* bits (stream overrun, palette
 *     out of range, nibble-count mismatch, cache full, rANS invariant).
 */
#ifndef FCS_H
#define FCS_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FCS_OK 0
#define FCS_E_INVALID -1    /* bad argument */
#define FCS_E_FORMAT -2     /* not a .csdb stream / truncated / inconsistent */
#define FCS_E_HIP -3        /* HIP runtime error */
#define FCS_E_NOMEM -4      /* host or device allocation failed */
#define FCS_E_NODEVICE -5   /* no HIP device */
#define FCS_E_CORRUPT -6    /* a brick stream failed decoding (see status words) */
#define FCS_E_CACHE_FULL -7 /* brick cache pool exhausted */

/* Per-brick / per-request status bits (OR-ed together). */
#define FCS_ST_OK 0u
#define FCS_ST_STREAM_OVERRUN 1u /* op stream ended before the requested level */
#define FCS_ST_PALETTE_RANGE 2u  /* palette cursor outside [0, P) */
#define FCS_ST_NSYM_MISMATCH 4u  /* nibble count disagrees with the op structure */
#define FCS_ST_CACHE_FULL 8u     /* brick cache pool exhausted */
#define FCS_ST_RANS_STATE 16u    /* rANS end-of-stream invariant violated */
#define FCS_ST_BAD_OP 32u        /* op code 7 (undefined by the format) */
#define FCS_ST_BAD_REQUEST 64u   /* brick index or LOD out of range */

typedef struct fcs_ctx fcs_ctx;
typedef struct fcs_volume fcs_volume;

typedef struct {
  uint32_t num_bricks;      /* = bricks_per_axis^3 */
  uint32_t brick_size;      /* b */
  uint32_t levels;          /* L = log2(b); LOD t in [0, L] */
  uint32_t bricks_per_axis;
  uint32_t grid_size;       /* bricks_per_axis * b */
  uint32_t label_bits;      /* 32 or 64 (64 when any palette entry >= 2^32) */
  uint32_t header_brick_size; /* value stored in the file header */
  uint32_t rans_ready;      /* 1 once fcs_decode_rans has run */
  uint64_t compressed_bytes;  /* sum of rANS stream bytes */
  uint64_t total_symbols;     /* sum of nSymbols */
  uint64_t palette_entries;   /* sum of palette sizes */
  uint64_t cache_capacity_voxels; /* brick cache pool size */
  uint32_t rans_freq[16];
  uint32_t rans_cum[16];
} fcs_volume_info;

/* Camera: the 64-byte uniform of camera.js:61-77 (position, tanHalfFov,
 * forward, right, up) plus an orthographic switch. ortho_half_height > 0
 * selects an orthographic camera: ray origin = position + right*ndc_x*aspect*h
 * + up*ndc_y*h, direction = forward. */
typedef struct {
  float position[3];
  float tan_half_fov;
  float forward[3];
  float ortho_half_height;
  float right[3];
  float pad1;
  float up[3];
  float pad2;
} fcs_camera;

/* Per-frame outputs (host pointers, each nullable). */
typedef struct {
  void* label;             /* w*h labels (uint32 or uint64 per label_bits); 0 = no hit */
  uint64_t* voxel;         /* w*h linear voxel index x + G*(y + G*z) of the hit
                              voxel's LOD-cell min corner; UINT64_MAX = no hit */
  float* t_hit;            /* w*h ray parameter at entry of the hit cell; -1 = no hit */
  uint32_t* brick_visited; /* num_bricks flags: 1 = a ray entered this non-empty brick */
  uint8_t* brick_lod_used; /* num_bricks LODs used this frame */
  /* filled by the library */
  uint32_t bricks_decoded; /* decode requests issued for this frame (cache misses) */
  uint32_t pad;
  float decode_ms;         /* device time of this frame's decode (hipEvent) */
  float march_ms;          /* device time of this frame's march (hipEvent) */
  /* caller-set, nullable: w*h RGBA8 pixels (R in bits 0-7, A in 24-31) of the
   * shaded front-to-back composite of rayTrace.wgsl:565-906 (transfer
   * function, LOD gradient, diffuse light, shadow rays, early end at alpha
   * 0.95). When set, the march runs in its compositing form; the first-hit
   * outputs are unchanged. */
  uint32_t* rgba;
} fcs_frame;

/* Lighting of shaded frames (webgpuSetup.js:21-31, 186-222): light direction
 * (sin(angle), 0.6, -cos(angle)), ambient term, diffuse strength. */
typedef struct {
  float light_angle_deg;  /* default 25 */
  float ambient;          /* default 0.5 */
  float diffuse_strength; /* default 0.8 */
  float pad;
} fcs_lighting;

typedef struct {
  float rans_ms;          /* last fcs_decode_rans */
  float decode_ms;        /* last fcs_request_bricks / fcs_decode_all */
  float march_ms;         /* last fcs_render march */
  uint32_t rans_mapping;  /* FCS_RANS_MAPPING_* used by the last fcs_decode_rans */
  uint64_t decoded_voxels;  /* labels written by the last decode */
  uint64_t decode_bytes;    /* algorithmic bytes of the last decode */
  uint64_t rans_symbols;    /* symbols decoded by the last fcs_decode_rans */
  uint64_t rans_bytes;      /* compressed bytes consumed by the last fcs_decode_rans */
  float tune_lane_ms;     /* autotuner: whole-volume rANS time, one lane per brick */
  float tune_wave_ms;     /* autotuner: whole-volume rANS time, one wave per brick (table) */
  float tune_ballot_ms;   /* autotuner: whole-volume rANS time, one wave per brick (ballot) */
  float tune_wave_lds_ms; /* autotuner: whole-volume rANS time, one wave per brick (LDS table) */
  uint32_t ops_split;     /* FCS_OPS_SPLIT_* the op decoder used last (bricks up to 32^3) */
  float tune_ops_3wg_ms;  /* autotuner: op decode time of the first large batch, 3 workgroups/CU */
  float tune_ops_4wg_ms;  /* autotuner: the same batch, 4 workgroups/CU */
  float tune_win_ms;      /* autotuner: whole-volume rANS time, one wave per brick, scalar byte window (table) */
  float tune_win_ballot_ms; /* autotuner: same, ballot over a per-lane CDF */
  float pipeline_ms;      /* last fcs_decode_volume: both stages */
  /* last fcs_render / fcs_render_visible march with FCS_OPT_MARCH_STATS set
   * (host-built frames): what its rays read -- label samples, non-empty
   * bricks entered (each reads its 48-byte records), empty bricks tested */
  uint64_t march_samples;
  uint64_t march_bricks_entered;
  uint64_t march_empty_tested;
} fcs_timing;

/* Options (fcs_set_option). */
#define FCS_OPT_RANS_MAPPING 1 /* FCS_RANS_MAPPING_*; default AUTO */
#define FCS_RANS_MAPPING_AUTO 0 /* measure every mapping on the volume's first rANS pass, keep the fastest */
#define FCS_RANS_MAPPING_LANE 1 /* one lane per brick (many bricks) */
#define FCS_RANS_MAPPING_WAVE 2 /* one wave per brick (few, long bricks), scalar decode table */
#define FCS_RANS_MAPPING_WAVE_BALLOT 3 /* one wave per brick, ballot over a per-lane CDF */
#define FCS_RANS_MAPPING_WAVE_LDS 4 /* one wave per brick, decode table in LDS */
#define FCS_RANS_MAPPING_WAVE_WIN 5 /* one wave per brick, byte window in scalar registers, scalar decode table */
#define FCS_RANS_MAPPING_WAVE_WIN_BALLOT 6 /* one wave per brick, byte window in scalar registers, ballot */
#define FCS_OPT_OPS_SPLIT 2 /* FCS_OPS_SPLIT_*; default AUTO */
#define FCS_OPS_SPLIT_AUTO 0 /* time both splits on a volume's first batch of >= 2048 requests, keep the faster */
#define FCS_OPS_SPLIT_3WG 1 /* op decoder LDS split for three workgroups per CU (bricks up to 32^3) */
#define FCS_OPS_SPLIT_4WG 2 /* op decoder LDS split for four workgroups per CU (bricks up to 32^3) */
#define FCS_OPT_OPS_HEAD 3 /* FCS_OPS_HEAD_*; default AUTO */
#define FCS_OPS_HEAD_AUTO 0 /* decode the levels up to 3 in a separate one-wave-per-request pass for launches of
                               >= 1024 requests */
#define FCS_OPS_HEAD_ON 1   /* always use that pass */
#define FCS_OPS_HEAD_OFF 2  /* never: the op decoder's first wave decodes those levels */
#define FCS_OPT_MARCH_STATS 4 /* 1: host-built frames count what their rays read (fcs_timing.march_*); 0: off */

int fcs_create(int hip_device, fcs_ctx** out);
void fcs_destroy(fcs_ctx* ctx);
const char* fcs_last_error(const fcs_ctx* ctx);
int fcs_device_count(void);

/* Parse a .csdb byte stream and upload it. brick_size_override != 0 replaces
 * the header brick size (the reference encoder always writes 64). */
int fcs_load_csdb(fcs_ctx* ctx, const uint8_t* bytes, size_t n, uint32_t brick_size_override,
                  fcs_volume** out);
void fcs_volume_free(fcs_volume* vol);
int fcs_volume_get_info(const fcs_volume* vol, fcs_volume_info* out);

/* Run the device rANS stage for every brick: nibble streams + per-level side
 * table. out_status (nullable) receives num_bricks status words. Returns
 * FCS_E_CORRUPT if any brick failed. */
int fcs_decode_rans(fcs_volume* vol, uint32_t* out_status);

/* Size the device brick cache pool (voxels). Drops resident bricks. */
int fcs_cache_reserve(fcs_volume* vol, uint64_t voxels);
/* Drop every resident brick (pool kept). */
int fcs_cache_clear(fcs_volume* vol);
/* The resident set, the brick cache's residency map (cacheManager.js:52-78,
 * brickCache.residency read by forceFullBrickRefresh / reDecodeCachedBricks,
 * renderLoop.js:502-595): resident bricks in ascending index order with the
 * LOD each holds (its finest level for a pyramid); writes min(cap, n) entries
 * (either pointer nullable) and returns n. */
int64_t fcs_cache_resident(fcs_volume* vol, uint32_t* bricks, uint8_t* lods, uint64_t cap);

/* Decode n (brick, LOD) requests into the brick cache (one batched launch).
 * Bricks already resident at that LOD are not decoded again.
 * out_cache_offset (nullable): voxel offset of each request's labels in the pool.
 * out_status (nullable): per-request FCS_ST_* bits.
 * A brick is resident at one LOD (or as a pyramid) at a time: when a call
 * requests a brick at several LODs, the last request wins and the regions of
 * the earlier ones stay valid until the next call. */
int fcs_request_bricks(fcs_volume* vol, const uint32_t* brick, const uint8_t* lod, uint32_t n,
                       uint64_t* out_cache_offset, uint32_t* out_status);

/* No reference counterpart: fcs_request_bricks for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops,
 * hosts that forbid allocation). Same requests, offsets, statuses and labels,
 * but the op decoder's request list, status words and claim counter live in a
 * fixed device array of the library (one per GPU, held until the process
 * ends) instead of buffers it allocates, and the call performs no device
 * allocation or free: the brick cache pool must already be reserved
 * (fcs_cache_reserve / fcs_cache_reserve_direct) and the rANS stage must have
 * run (fcs_decode_rans), else FCS_E_INVALID. At most FCS_STATIC_MAX_REQUESTS
 * requests per call (more: FCS_E_INVALID). Every call on one GPU shares that
 * array: calls that overlap on one GPU (other threads, other contexts) race
 * and must be serialised by the caller. */
#define FCS_STATIC_MAX_REQUESTS 262144u
int fcs_request_bricks_static(fcs_volume* vol, const uint32_t* brick, const uint8_t* lod, uint32_t n,
                              uint64_t* out_cache_offset, uint32_t* out_status);

/* Decode every brick at LOD `lod` (lod > levels: every brick at its finest LOD).
 * pyramid != 0 writes all levels 0..lod of each brick (coarse-first layout). */
int fcs_decode_all(fcs_volume* vol, uint32_t lod, int pyramid);

/* Every brick's rANS stage and then its decode at LOD `lod`, as fcs_decode_rans
 * followed by fcs_decode_all, in one call: both launches go onto the context
 * stream back to back, with no host synchronisation between them. Same labels
 * and statuses as the two calls. out_status (nullable, num_bricks entries): per
 * brick, the FCS_ST_* bits of both stages. fcs_get_timing: pipeline_ms. */
int fcs_decode_volume(fcs_volume* vol, uint32_t lod, int pyramid, uint32_t* out_status);

/* Copy the labels of `brick` at `lod` to the host (8^lod labels, Morton order),
 * decoding it first if it is not resident. */
int fcs_read_brick(fcs_volume* vol, uint32_t brick, uint32_t lod, void* host_labels);

/* Copy the whole brick cache pool region [offset, offset+count) to the host. */
int fcs_read_cache(fcs_volume* vol, uint64_t offset, uint64_t count, void* host_labels);
/* decodeAllRansBricks's per-brick result (web/scripts/decode.js:17-26): the
 * nibble stream of `brick` after fcs_decode_rans, nSymbols nibbles packed two
 * per byte with the first nibble in the high half ((nSymbols+1)/2 bytes; a
 * trailing odd nibble leaves the low half 0). Returns the byte count. */
int64_t fcs_read_nibbles(fcs_volume* vol, uint32_t brick, uint8_t* out, uint64_t cap);

/* Per-brick LOD by camera.js:103-137 (f64) for a w x h viewport. */
int fcs_brick_lods(const fcs_volume* vol, const fcs_camera* cam, uint32_t w, uint32_t h,
                   uint8_t* out_lods);

/* No reference counterpart (the reference's pool is a bump allocator with
 * per-LOD free stacks, cacheManager.js:52-78, as fcs_cache_reserve's is):
 * replace the pool with a direct-mapped one of bricks x 8^L voxels in which
 * brick i owns slot i, so regions never move or fragment. A brick then holds
 * one LOD at a time: when one fcs_request_bricks call asks for a brick more
 * than once, only its last request is decoded and the earlier ones report
 * that request's slot and status. */
int fcs_cache_reserve_direct(fcs_volume* vol);

/* Render one frame (replaces renderLoop.js:268-298, the frame's decode pass +
 * rayTrace pass): brick_lod (nullable: LOD rule) selects each brick's LOD;
 * non-empty bricks inside the view frustum (culling.js:87-171, exact
 * perspective planes) that are not resident at that LOD are decoded first,
 * on demand, then every pixel is marched. Decodes issued ahead with
 * fcs_prefetch_bricks finish before the frame's own work. Output arrays left
 * NULL in `out` are not copied back. */
int fcs_render(fcs_volume* vol, const fcs_camera* cam, const uint8_t* brick_lod, uint32_t w,
               uint32_t h, fcs_frame* out);

/* No reference counterpart: render one frame decoding only the bricks its rays
 * enter. March with what is resident; a ray entering a non-empty brick that is
 * not resident at its LOD is counted and passes through it; those bricks are
 * decoded and the frame is marched again until no ray misses, so the frame
 * equals fcs_render's. `out->bricks_decoded` counts the decodes. */
int fcs_render_visible(fcs_volume* vol, const fcs_camera* cam, const uint8_t* brick_lod, uint32_t w,
                       uint32_t h, fcs_frame* out);

/* No reference counterpart (the reference only decodes on demand): decode
 * (brick, LOD) pairs ahead of the frames that need them, asynchronously on the
 * context's second stream. Residency is updated at once; the next fcs_render
 * (or any cache call) orders itself after these decodes. A prefetch that fails
 * leaves its brick non-resident, so on-demand decode retries and reports it.
 * Returns the number of decodes issued (pairs already resident are skipped). */
int fcs_prefetch_bricks(fcs_volume* vol, const uint32_t* brick, const uint8_t* lod, uint32_t n);

/* No reference counterpart: order-independent 64-bit checksum of the last
 * rendered frame's first-hit labels and voxel indices (computed on the
 * device), for comparing frames without copying them back. */
int fcs_frame_checksum(fcs_volume* vol, uint64_t* out);

/* No reference counterpart: the 12 camera features the prefetch network reads,
 * position / grid size, forward, up, right. */
int fcs_camera_features(const fcs_volume* vol, const fcs_camera* cam, float* out12);

/* Orbit pose around the volume centre (orbitController.js:100-123) with the
 * basis of camera.js:61-77, fov in radians. */
/* Shading controls. updateLightingOptions (webgpuSetup.js:186-222). */
int fcs_set_lighting(fcs_volume* vol, const fcs_lighting* lighting);
/* applyLabelVisibility + applyLabelHover (webgpuSetup.js:827-882): labels in
 * `hidden` become transparent; `highlight` (nullable) is brightened x1.8
 * (clamped at 255); each call starts again from the default colours. */
int fcs_set_label_visibility(fcs_volume* vol, const uint64_t* hidden, uint32_t n_hidden, const uint64_t* highlight);
/* getDatasetLabels / buildTransferFunctionBuffers (webgpuSetup.js:57-59,
 * 389-458): the transfer function's labels (every distinct non-zero palette
 * label, ascending) and their current RGBA8 colours; writes min(cap, n)
 * entries (either pointer nullable) and returns n. */
int64_t fcs_transfer_function(fcs_volume* vol, uint64_t* labels, uint32_t* rgba, uint64_t cap);
/* Default RGBA8 of transfer function entry `index` (webgpuSetup.js:410-429,
 * hslToRgb 490-515, JavaScript double arithmetic): hue index*137.5 mod 360,
 * saturation 0.55 + (index mod 3)*0.08, lightness 0.35 + (index mod 4)*0.05,
 * opacity 0.25 + (index mod 5)*0.08. */
uint32_t fcs_tf_default_rgba(uint64_t index);

int fcs_camera_orbit(const fcs_volume* vol, double yaw, double pitch, double radius, double fov,
                     fcs_camera* out);

int fcs_get_timing(const fcs_ctx* ctx, fcs_timing* out);
/* No reference counterpart: runtime knobs (FCS_OPT_*). */
int fcs_set_option(fcs_ctx* ctx, int option, int64_t value);
int fcs_sync(fcs_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
