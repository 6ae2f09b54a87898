/* CPU producer side of the .csdb format: deterministic synthetic label
 * volumes and a streaming brick encoder.
 *
 * The encoder restates the reference producer pipeline
 *   splitGridIntoBricks<b>   (encoder/CSVEncoder/src/brick.hpp:133-194)
 *   Brick<b>::build          (brick.hpp:78-123, majority vote brick.hpp:24-37)
 *   encodeBrick / bestOperation (brickEncoding.hpp:188-447)
 *   packOperationsToNibbles  (brickEncoding.cpp:11-91)
 *   updateFrequencyTables / buildRansModel / ransEncodeSymbol
 *                            (brickEncoding.cpp:94-218, brickEncoding.hpp:466-531)
 *   saveDatasetToFile        (fileIO.cpp:15-99)
 * brick-at-a-time and multi-threaded, so volumes far larger than host RAM
 * can be encoded. Output bytes equal the reference encoder's, except header
 * bytes 12..15 (brickSize), which the reference always writes as 64
 * (fileIO.cpp:28) and this producer writes as the true brick edge unless
 * `header_brick_size_64` is set.
 *
 * All functions return 0 on success and a negative FCS_E_* code on error.
 */
#ifndef FCS_PRODUCER_H
#define FCS_PRODUCER_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Volume recipes. Every recipe is a pure function of (recipe, seed, x, y, z),
 * so any brick can be generated independently. Integer arithmetic only. */
typedef enum {
  FCS_GEN_VORONOI = 0, /* Voronoi cells, optional metric weights, noise, background */
  FCS_GEN_REGIONS = 1, /* 256^3 regions: background / sparse (~50 cells) / dense (~4000 noisy cells) */
} fcs_gen_kind;

typedef enum {
  FCS_BG_NONE = 0,
  FCS_BG_SEED_SLAB = 1, /* cells whose seed lies in a random-axis slab (30 % of the edge) get label 0 */
  FCS_BG_TWO_BLOBS = 2, /* two spheres (union ~25 % of voxels) of label 0 */
} fcs_gen_background;

typedef struct {
  uint32_t kind;        /* fcs_gen_kind */
  uint32_t size;        /* cube edge in voxels */
  uint64_t seed;
  uint32_t num_seeds;   /* Voronoi seeds (FCS_GEN_VORONOI) */
  uint32_t wx, wy, wz;  /* integer metric weights: d^2 = (wx dx)^2 + (wy dy)^2 + (wz dz)^2 */
  uint32_t noise_amp;   /* boundary warp amplitude in voxels (value noise, lattice 8) */
  uint32_t background;  /* fcs_gen_background */
} fcs_gen_params;

typedef struct fcs_gen fcs_gen;

/* Presets for the five benchmark configurations (1..5). `size` = 0 keeps the
 * preset edge; config 5 takes its edge from `size` (256/512/1024/2048). */
int fcs_gen_preset(int config, uint32_t size, fcs_gen_params* out);
int fcs_gen_create(const fcs_gen_params* p, fcs_gen** out);
void fcs_gen_destroy(fcs_gen* g);
/* Labels of the box [x0,x0+nx) x [y0,y0+ny) x [z0,z0+nz), flat x-fastest:
 * out[x + nx*(y + ny*z)]. */
int fcs_gen_box(const fcs_gen* g, uint32_t x0, uint32_t y0, uint32_t z0, uint32_t nx,
                uint32_t ny, uint32_t nz, uint32_t* out, int threads);
/* Source labels of `count` consecutive bricks of edge b, starting at brick
 * Morton ID `first` (the order bricks are loaded in, datasetLoader.js:143),
 * each brick's b^3 labels in Morton order (the decoder's finest-LOD layout,
 * decode.wgsl:239-245), brick after brick: out[b^3 * k + morton(dx, dy, dz)].
 * The finest LOD of a decoded brick equals this (closed form). */
int fcs_gen_bricks(const fcs_gen* g, uint32_t b, uint32_t first, uint32_t count, uint32_t* out,
                   int threads);
/* Fraction of background (label 0) voxels, estimated on a 64^3 lattice. */
double fcs_gen_background_fraction(const fcs_gen* g);

/* Encoder options. */
typedef struct {
  uint32_t brick_size;          /* b = 2^L, 2 <= b <= 64 */
  uint32_t header_brick_size_64;/* 1: write 64 in the header like the reference encoder */
  int threads;                  /* <= 0: hardware concurrency */
  uint32_t progress;            /* 1: print a line to stderr at every sixteenth of the bricks encoded */
} fcs_encode_opts;

/* Encode a flat label volume (x-fastest, data[x + W*(y + H*z)], W/H/D need
 * not be multiples of b; outside voxels pad as 0). label_bits 32 or 64.
 * *out_bytes is malloc'ed; release with fcs_free_buffer. */
int fcs_encode_volume(const void* labels, uint32_t label_bits, uint32_t W, uint32_t H,
                      uint32_t D, const fcs_encode_opts* opts, uint8_t** out_bytes,
                      size_t* out_len);
/* Encode a generated cube brick by brick (never materialises the volume). */
int fcs_encode_generated(const fcs_gen* g, const fcs_encode_opts* opts, uint8_t** out_bytes,
                         size_t* out_len);
void fcs_free_buffer(void* p);

/* Majority-vote LOD pyramid of one brick, coarse-first: level t (8^t labels,
 * Morton order) starts at out[(8^t-1)/7]; const_out (nullable) receives the
 * constantChildren flags in the same layout. brick_labels holds b^3 labels in
 * Morton order. Follows brick.hpp:78-123. */
int fcs_brick_pyramid(const uint64_t* brick_labels, uint32_t brick_size, uint64_t* out,
                      uint8_t* const_out);

#ifdef __cplusplus
}
#endif
#endif
