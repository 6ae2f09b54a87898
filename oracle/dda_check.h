/* dda_check.h — an independent float64 check of the first-hit DDA.
 * TEST INFRASTRUCTURE ONLY: used by tests/ as a checker, never by the product.
 *
 * The FP32 marcher (rayTrace.wgsl:457-563, 597-619, 647-847; restated by
 * or_render and by the device kernel) walks a brick grid and, inside each
 * non-empty brick, a grid of LOD cells (stride b / 2^t) with entry/exit
 * epsilons and per-brick re-initialisation. This checker does not follow
 * that code: it derives each ray from the camera in float64
 * (dir = normalize(f + r*ndc_x*aspect*tan + u*ndc_y*tan); orthographic:
 * origin offset, dir = f), clips it against [0, G]^3 and enumerates, in
 * float64, every cell of the global LOD-cell grid the ray passes through with
 * its entry and exit parameter. The first cell with a non-zero label is the
 * exact first hit.
 *
 * Where the exact and the FP32 answers may legitimately differ (a "grazing
 * tie"): a cell the ray stays in for less than tau(t), a bound on the FP32
 * marcher's accumulated rounding (ray setup, tDelta sums) plus twice its
 * epsilon, may be skipped by the FP32 walk, and the cells around such a
 * short stay (the other crossing order) may be visited instead. On exact
 * float64 ties strictly inside a brick the reference's tie rule applies
 * (advanceDDA's strict comparisons: z steps first, then y, then x), with no
 * allowance: the intermediate cell it names is sampled and the other order's
 * cell is not. Every other difference is unexplained.
 */
#ifndef FCS_DDA_CHECK_H
#define FCS_DDA_CHECK_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Camera: the same 64-byte layout as or_camera / fcs_camera. */
typedef struct {
  float position[3], tan_half_fov, forward[3], ortho_half_height, right[3], pad1, up[3], pad2;
} dc_camera;

/* Per-pixel classes written to cls. */
#define DC_AGREE 0      /* same (label, voxel) as the exact first hit, t within tau */
#define DC_GRAZING 1    /* differs, explained by a grazing tie */
#define DC_UNEXPLAINED 2

typedef struct {
  uint64_t pixels, hits, agree, grazing, unexplained;
  double max_t_err; /* largest |t_fp32 - t_exact| over agreeing hits */
} dc_stats;

/* Check one frame. got_label / got_voxel / got_t: the FP32 marcher's
 * per-pixel first hit (label 0 / voxel UINT64_MAX / t -1 = no hit).
 * exp_label / exp_voxel / cls: nullable per-pixel outputs (the exact first
 * hit and the class). threads >= 1. Returns 0. */
int dc_check_frame(uint32_t bricks_per_axis, uint32_t brick_size, const uint8_t* brick_empty,
                   const uint8_t* brick_lod, const uint64_t* const* brick_labels,
                   const uint64_t* palette0, const dc_camera* cam, uint32_t w, uint32_t h,
                   const uint64_t* got_label, const uint64_t* got_voxel, const float* got_t,
                   int threads, uint64_t* exp_label, uint64_t* exp_voxel, uint8_t* cls,
                   dc_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
