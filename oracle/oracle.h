/* oracle.h — CPU restatement of the reference's brick decoder and DDA
 * marcher. TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py as the checker, never by the product.
 *
 * Each function cites the reference code it restates. Pinning (what checks
 * each function against something other than itself):
 *   or_parse_csdb      shipped fixture compressed_dataset.csbd (header,
 *                      model, 512 bricks), reference-encoder outputs.
 *   or_rans_decode     reference decodeRansStream (oracle/_ref) on every
 *                      brick of every golden .csdb; end-of-stream invariant
 *                      (state == 2^23, cursor == 0).
 *   or_decode_brick    finest LOD == source volume; every LOD == encoder
 *                      pyramid dumped from the reference encoder; reference
 *                      decodeBrick<b> at the finest LOD; LOD 0 == palette[0].
 *   or_render          orthographic +z frames == brute-force column scan;
 *                      perspective and orthographic frames == the float64
 *                      exact traversal of dda_check.c (every pixel agrees or
 *                      is a grazing tie within the FP32 rounding bound; the
 *                      tie rule on exact ties); planted mistakes (tie order,
 *                      entry epsilon, step sign) fail it.
 *   or_brick_lod       exact thresholds of camera.js:103-137 in tests.
 *   or_render_rgba     closed forms: a one-label volume seen along +z
 *                      composites to 1-(1-a)^n in the label's colour; a
 *                      planar interface gives the analytic normal and
 *                      diffuse term; hidden labels are transparent; a
 *                      planted mistake (opacity correction dropped) fails.
 */
#ifndef FCS_ORACLE_H
#define FCS_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  uint32_t id;        /* Morton code of the brick position */
  uint32_t P;         /* palette size */
  uint32_t E;         /* rANS byte count */
  uint32_t nsym;      /* nibble count */
  uint64_t pal_off;   /* index into or_dataset.palette */
  uint64_t str_off;   /* byte offset into the source buffer */
} or_brick;

typedef struct {
  uint32_t num_bricks, header_brick_size, brick_size, version;
  uint32_t freq[16], cum[17], total;
  or_brick* bricks;   /* sorted by id (datasetLoader.js:143) */
  uint64_t* palette;
  const uint8_t* src; /* caller's buffer (not owned) */
} or_dataset;

int or_parse_csdb(const uint8_t* bytes, size_t n, uint32_t brick_size_override, or_dataset* out);
void or_free(or_dataset* d);

/* decodeRansStream: returns the number of repacked bytes written to out
 * ((nsym+1)/2 bytes, high nibble first); *status gets FCS_ST bits
 * (16 = end-of-stream invariant violated). */
size_t or_rans_decode(const or_dataset* d, uint32_t brick, uint8_t* out, uint32_t* status);

/* decode.wgsl decodeBrick semantics at LOD t, emitting labels (u64) in
 * Morton order (8^t entries); the root stop bit fills the brick with
 * palette[0]. nib = repacked nibble bytes of length nbytes. */
int or_decode_brick(const uint8_t* nib, size_t nbytes, const uint64_t* palette, uint32_t P,
                    uint32_t t, uint64_t* out, uint32_t* status);

/* Whole chain for one brick of a dataset (rANS then op decode). */
int or_decode_dataset_brick(const or_dataset* d, uint32_t brick, uint32_t t, uint64_t* out,
                            uint32_t* status);

typedef struct {
  float position[3], tan_half_fov, forward[3], ortho_half_height, right[3], pad1, up[3], pad2;
} or_camera;

/* rayTrace.wgsl traversal up to the first hit (label != 0).
 * brick_labels[i] points at brick i's labels at LOD brick_lod[i] (Morton
 * order; unused for empty bricks or LOD 0, where palette0[i] is used);
 * NULL for a non-empty brick of LOD > 0 = not resident: rays pass through. */
int or_render(uint32_t bricks_per_axis, uint32_t brick_size, const uint8_t* brick_empty,
              const uint8_t* brick_lod, const uint64_t* const* brick_labels,
              const uint64_t* palette0, const or_camera* cam, uint32_t w, uint32_t h,
              uint64_t* out_label, uint64_t* out_voxel, float* out_t, uint32_t* visited);

/* The compositing form of rayTrace.wgsl:565-906: labels coloured by the
 * transfer function (tf_labels ascending, tf_rgba RGBA8; other non-zero
 * labels get decode.wgsl's hash colour), front-to-back compositing with the
 * LOD opacity correction, LOD-gradient diffuse shading and shadow rays,
 * early end at alpha 0.95. brick_labels[i] == NULL: brick i is not resident
 * (its samples are transparent). out_rgba: w*h RGBA8. */
int or_render_rgba(uint32_t bricks_per_axis, uint32_t brick_size, const uint8_t* brick_empty,
                   const uint8_t* brick_lod, const uint64_t* const* brick_labels, const uint64_t* palette0,
                   const or_camera* cam, uint32_t w, uint32_t h, const uint64_t* tf_labels,
                   const uint32_t* tf_rgba, uint64_t tf_n, const float light[3], float ambient,
                   float diffuse, uint32_t* out_rgba);

/* camera.js:103-137 (f64). */
uint32_t or_brick_lod(double cx, double cy, double cz, double brick_size, double px, double py,
                      double pz, double fov, double screen_width, uint32_t max_lod);

/* Multi-threaded full decode of every brick at every LOD 0..L into a
 * pyramid layout (per brick (8^(L+1)-1)/7 labels); used as the CPU baseline.
 * Returns the number of labels written. */
uint64_t or_decode_all_pyramid(const or_dataset* d, uint64_t* out, int threads,
                               uint32_t* status_or);

/* Threaded decode of n (brick, LOD) requests, request k's 8^lod[k] labels
 * at offset[k] of out32 (uint32, when non-NULL) or out64; status[k] gets
 * its status bits. The checker of batched device launches. */
int or_decode_requests(const or_dataset* d, const uint32_t* brick, const uint8_t* lod, uint32_t n,
                       const uint64_t* offset, uint32_t* out32, uint64_t* out64, int threads,
                       uint32_t* status);

#ifdef __cplusplus
}
#endif
#endif
