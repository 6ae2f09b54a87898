/*
 * Serving: many viewers over device-resident volumes. No counterpart in the
 * reference, which renders one frame at a time for one viewer
 * (web/scripts/renderLoop.js:212-359) and selects LODs and cache contents on
 * the host (web/scripts/camera.js:140-201, web/scripts/cacheProcessing.js:99-232).
 *
 * A view is a render target with its own HIP stream, device frame buffers up
 * to max_w x max_h and pinned host outputs. A view frame is set up on the
 * device: one kernel applies the LOD rule of fcs_brick_lods (camera.js:103-137,
 * float64, bit-identical to the host rule) to every brick and builds the
 * marcher's per-brick metadata from the volume's device residency table, so a
 * frame costs no per-brick host work. A frame of a view never decodes: a ray
 * that enters a brick not resident at its LOD is counted as a miss, and the
 * frame is then incomplete (the serving loop serves such a request again with
 * fcs_render_visible, which decodes what is missing).
 *
 * Views on a volume must not overlap with calls that change that volume's
 * cache (request, decode, prefetch, clear, reserve); the serving loop keeps
 * this rule.
 */
#ifndef FCS_SERVE_H
#define FCS_SERVE_H

#include <stddef.h>
#include <stdint.h>

#include "fcs/fcs.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fcs_view fcs_view;

#define FCS_VIEW_GRAPH 1u    /* replay each (volume, size) frame as one instantiated hipGraph */
#define FCS_VIEW_LABELS 2u   /* copy the first-hit label image to the view's pinned host buffer */
#define FCS_VIEW_CHECKSUM 4u /* compute the fcs_frame_checksum value of every frame on the device */
#define FCS_VIEW_RGBA 8u     /* shaded frames (the RGBA form of fcs_frame.rgba); with FCS_VIEW_LABELS the
                                RGBA image is copied to the pinned host buffer instead of the labels */

typedef struct {
  uint64_t checksum;   /* FCS_VIEW_CHECKSUM: equals fcs_frame_checksum of the same frame */
  uint32_t misses;     /* rays that entered a non-resident brick; 0 = the frame is complete */
  uint32_t w, h;
  uint32_t label_bits; /* 32 or 64: width of the labels in fcs_view_labels */
  float device_ms;     /* hipEvent time of the frame on the view's stream */
  uint32_t pad;
} fcs_view_result;

int fcs_view_create(fcs_ctx* ctx, uint32_t max_w, uint32_t max_h, uint32_t flags, fcs_view** out);
void fcs_view_destroy(fcs_view* view);
/* Enqueue a frame of `vol` seen from `cam` on the view's stream and return.
 * brick_lod (nullable, num_bricks entries): fixed LODs instead of the rule;
 * such frames are never graph-replayed. A view holds one frame: a call while
 * the previous frame runs waits for it first. */
int fcs_view_render_async(fcs_view* view, fcs_volume* vol, const fcs_camera* cam, const uint8_t* brick_lod,
                          uint32_t w, uint32_t h);
/* 1 when the view's last frame has finished (or none was issued), 0 while it runs. */
int fcs_view_query(fcs_view* view);
/* Wait for the view's last frame and describe it. */
int fcs_view_wait(fcs_view* view, fcs_view_result* out);
/* After fcs_view_wait, FCS_VIEW_LABELS views: the pinned host label image of
 * the last frame (w*h labels of label_bits bits; w*h RGBA8 words for
 * FCS_VIEW_RGBA views), valid until the next frame. */
const void* fcs_view_labels(const fcs_view* view);
/* After fcs_view_wait: copy the last frame's device outputs to host buffers,
 * each nullable (labels of label_bits, voxel index, t, the LOD used per brick). */
int fcs_view_read(fcs_view* view, void* label, uint64_t* voxel, float* t_hit, uint8_t* brick_lod_used);
/* After fcs_view_wait, FCS_VIEW_RGBA views: copy the last frame's RGBA8 image. */
int fcs_view_read_rgba(fcs_view* view, uint32_t* rgba);
/* hipGraph instantiations the view holds (one per volume and frame size seen). */
uint32_t fcs_view_graphs(const fcs_view* view);

/* Admission under an HBM budget of brick-cache bytes. full_bytes[i] = bytes
 * of volume i's full pyramid cache (every brick at every LOD), demand[i] =
 * its expected share of frames. Volumes are taken in order of demand per
 * full byte (ties: lower index first); each whose full pyramid fits in what is
 * left is admitted whole (plan 2: frames never miss). What remains is then
 * split among the other volumes in proportion to demand as on-demand pools
 * (plan 1, pool_bytes[i]); a volume with zero demand gets nothing (plan 0).
 * Pure host logic. */
int fcs_serve_plan(const uint64_t* full_bytes, const double* demand, uint32_t n, uint64_t budget_bytes,
                   uint32_t* plan, uint64_t* pool_bytes);
/* Apply fcs_serve_plan to volumes: plan 2 volumes get a pyramid cache with
 * every brick decoded at every LOD (rANS first if needed); plan 1 volumes an
 * empty on-demand pool of pool_bytes. plan/pool_bytes (nullable) receive the plan. */
int fcs_serve_admit(fcs_volume* const* vols, const double* demand, uint32_t n, uint64_t budget_bytes,
                    uint32_t* plan, uint64_t* pool_bytes);

typedef struct {
  double arrival_s;  /* seconds after the start of the run; requests sorted by arrival */
  uint32_t session;
  uint32_t volume;   /* index into the run's volumes */
  uint32_t w, h;
  fcs_camera cam;
} fcs_serve_request;

typedef struct {
  double issue_s;    /* when a view took the request */
  double done_s;     /* when the host saw the frame finished; latency = done_s - arrival_s */
  uint64_t checksum; /* frame checksum when the views compute it, else 0 */
  float device_ms;   /* device time of the view frame */
  uint32_t misses;   /* misses of the view frame (then served again on demand) */
  uint32_t view;
  uint32_t pad;
} fcs_serve_result;

/* Open-loop serving: requests are released at their arrival times (wall
 * clock from the call), wait in one FIFO queue and go to the first free view;
 * completions are seen by polling. A view frame with misses is served again
 * with fcs_render_visible once every view has drained. Views must all belong
 * to the volumes' context. Returns 0 or a negative FCS_E_* code. */
int fcs_serve_run(fcs_view* const* views, uint32_t n_views, fcs_volume* const* vols, uint32_t n_vols,
                  const fcs_serve_request* reqs, uint32_t n, fcs_serve_result* out);

#ifdef __cplusplus
}
#endif

#endif /* FCS_SERVE_H */
