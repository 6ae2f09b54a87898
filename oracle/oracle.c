/* oracle.c — CPU restatement of the reference brick decoder and DDA marcher.
 * TEST INFRASTRUCTURE ONLY (the checker): tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg. The product never links or calls this file.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math), so that
 * FP32 ray setup rounds exactly like the device code.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* OR_MUTANT (tests only): build a deliberately wrong oracle to show that the
 * golden checks catch plausible mistakes. 0 = correct. */
#ifndef OR_MUTANT
#define OR_MUTANT 0
#endif

#define ST_OVERRUN 1u
#define ST_PALETTE 2u
#define ST_NSYM 4u
#define ST_RANS 16u
#define ST_BADOP 32u

/* ------------------------------------------------------------ Morton ---- */
/* utils.cpp:4-42 */
static uint32_t part1by2(uint32_t n) {
  n &= 0x3ff;
  n = (n | (n << 16)) & 0x30000ff;
  n = (n | (n << 8)) & 0x300f00f;
  n = (n | (n << 4)) & 0x30c30c3;
  n = (n | (n << 2)) & 0x9249249;
  return n;
}
static uint32_t compact1by2(uint32_t n) {
  n &= 0x9249249;
  n = (n ^ (n >> 2)) & 0x30c30c3;
  n = (n ^ (n >> 4)) & 0x300f00f;
  n = (n ^ (n >> 8)) & 0x30000ff;
  n = (n ^ (n >> 16)) & 0x3ff;
  return n;
}
static uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  return part1by2(x) | (part1by2(y) << 1) | (part1by2(z) << 2);
}

/* --------------------------------------------------------- .csdb parse -- */
/* fileIO.cpp:101-183 and datasetLoader.js:39-145 (sort by ID: line 143) */
static uint32_t rd32(const uint8_t* p) {
  return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
static int cmp_brick(const void* a, const void* b) {
  const or_brick* x = (const or_brick*)a;
  const or_brick* y = (const or_brick*)b;
  return x->id < y->id ? -1 : (x->id > y->id ? 1 : 0);
}

int or_parse_csdb(const uint8_t* bytes, size_t n, uint32_t brick_size_override, or_dataset* out) {
  memset(out, 0, sizeof *out);
  if (n < 16 + 8) return -2;
  if (rd32(bytes) != 0x43534244u) return -2;
  out->version = rd32(bytes + 4);
  out->num_bricks = rd32(bytes + 8);
  out->header_brick_size = rd32(bytes + 12);
  out->brick_size = brick_size_override ? brick_size_override : out->header_brick_size;
  size_t p = 16;
  const uint32_t symbols = rd32(bytes + p);
  out->total = rd32(bytes + p + 4);
  p += 8;
  if (symbols != 16 || p + 128 > n) return -2;
  for (int i = 0; i < 16; ++i) out->freq[i] = rd32(bytes + p + 4 * i);
  for (int i = 0; i < 16; ++i) out->cum[i] = rd32(bytes + p + 64 + 4 * i);
  out->cum[16] = out->total;
  p += 128;
  out->bricks = (or_brick*)calloc(out->num_bricks ? out->num_bricks : 1, sizeof(or_brick));
  /* first pass: count palette entries */
  uint64_t pal_total = 0;
  size_t q = p;
  for (uint32_t i = 0; i < out->num_bricks; ++i) {
    if (q + 8 > n) return -2;
    const uint32_t P = rd32(bytes + q + 4);
    q += 8;
    if (q + 8ull * P + 4 > n) return -2;
    q += 8ull * P;
    const uint32_t E = rd32(bytes + q);
    q += 4;
    if (q + (uint64_t)E + 4 > n) return -2;
    q += E + 4;
    pal_total += P;
  }
  out->palette = (uint64_t*)malloc((pal_total ? pal_total : 1) * sizeof(uint64_t));
  uint64_t po = 0;
  for (uint32_t i = 0; i < out->num_bricks; ++i) {
    or_brick* b = &out->bricks[i];
    b->id = rd32(bytes + p);
    b->P = rd32(bytes + p + 4);
    p += 8;
    memcpy(out->palette + po, bytes + p, 8ull * b->P);
    b->pal_off = po;
    po += b->P;
    p += 8ull * b->P;
    b->E = rd32(bytes + p);
    p += 4;
    b->str_off = p;
    p += b->E;
    b->nsym = rd32(bytes + p);
    p += 4;
  }
  qsort(out->bricks, out->num_bricks, sizeof(or_brick), cmp_brick);
  out->src = bytes;
  return 0;
}

void or_free(or_dataset* d) {
  free(d->bricks);
  free(d->palette);
  memset(d, 0, sizeof *d);
}

/* ---------------------------------------------------------------- rANS -- */
/* findSymbol / ransDecodeSymbol / decodeRansStream, brickDecoding.cpp:126-188 */
static uint32_t find_symbol(uint32_t x, const or_dataset* d) {
  uint32_t s = 0;
  while (s + 1 < 16 && x >= d->cum[s + 1]) ++s;
  return s;
}
static uint32_t rans_step(uint32_t* state, const uint8_t* in, size_t* pos, const or_dataset* d) {
  const uint32_t x = *state % d->total;
  const uint32_t s = find_symbol(x, d);
  *state = d->freq[s] * (*state / d->total) + (x - d->cum[s]);
  while (*state < (1u << 23) && *pos > 0) *state = (*state << 8) | in[--*pos];
  return s;
}

size_t or_rans_decode(const or_dataset* d, uint32_t brick, uint8_t* out, uint32_t* status) {
  *status = 0;
  if (brick >= d->num_bricks) { *status = ST_RANS; return 0; }
  const or_brick* b = &d->bricks[brick];
  const uint8_t* in = d->src + b->str_off;
  if (b->E < 4) { *status |= ST_RANS; return 0; }
  size_t pos = b->E;
  const uint32_t b0 = in[--pos], b1 = in[--pos], b2 = in[--pos], b3 = in[--pos];
  uint32_t state = (b0 << 24) | (b1 << 16) | (b2 << 8) | b3;
  size_t nout = 0;
  uint32_t done = 0;
  while (done < b->nsym) {
    const uint32_t hi = rans_step(&state, in, &pos, d) & 0xF;
    ++done;
    uint32_t lo = 0;
    if (done < b->nsym) { lo = rans_step(&state, in, &pos, d) & 0xF; ++done; }
    out[nout++] = (uint8_t)((hi << 4) | lo);
  }
  /* the encoder coded an even symbol count (a 0 pad after an odd count) and
   * started from state 2^23, so after the pad the state returns to 2^23 with
   * every byte consumed */
  if (b->nsym & 1) {
    const uint32_t pad = rans_step(&state, in, &pos, d);
    if (pad != 0) *status |= ST_RANS;
  }
  if (state != (1u << 23) || pos != 0) *status |= ST_RANS;
  return nout;
}

/* --------------------------------------------------------- op decoder --- */
/* decode.wgsl:239-508 with brickDecoding.cpp:7-124 semantics; explicit
 * `filled` flags replace the in-band EMPTY sentinel (decode.wgsl:46). */
static uint32_t node_to_out(uint32_t m, uint32_t l, uint32_t t) {
  /* coarseNodeToOutputIndex (brickDecoding.cpp:7-18) */
  const uint32_t s = 1u << (t - l);
  return morton3(compact1by2(m) * s, compact1by2(m >> 1) * s, compact1by2(m >> 2) * s);
}

typedef struct {
  const uint8_t* nib;
  size_t nbytes, pos;
  int high;
} reader;

static int read_nibble(reader* r, uint32_t* v) {
  /* readNibble (brickDecoding.cpp:20-41): low nibble of byte 0 first */
  if (r->pos >= r->nbytes) return 0;
  const uint8_t byte = r->nib[r->pos];
  if (r->high) { *v = byte >> 4; r->high = 0; } else { *v = byte & 0xF; r->high = 1; r->pos++; }
  return 1;
}

int or_decode_brick(const uint8_t* nib, size_t nbytes, const uint64_t* pal, uint32_t P, uint32_t t,
                    uint64_t* out, uint32_t* status) {
  *status = 0;
  const uint64_t n = 1ull << (3 * t);
  uint8_t* filled = (uint8_t*)calloc(n, 1);
  if (!filled) return -4;
  const int root_stop = (OR_MUTANT == 3) ? 0 : (nbytes > 0 ? ((nib[0] >> 4) & 1) : 0);
  if (P == 0) { *status |= ST_PALETTE; free(filled); return 0; }
  out[0] = pal[0];
  filled[0] = 1;
  if (root_stop) {
    /* set root stop bit: constant brick, every LOD = palette[0] */
    for (uint64_t i = 0; i < n; ++i) out[i] = pal[0];
    free(filled);
    return 0;
  }
  reader r = {nib, nbytes, 0, 0};
  int64_t ip = 0;
  for (uint32_t l = 0; l < t; ++l) {
    const uint32_t count = 1u << (3 * l);
    for (uint32_t m = 0; m < count; ++m) {
      const uint32_t i = node_to_out(m, l, t);
      const uint32_t last = node_to_out(8 * m + 7, l + 1, t);
      if (filled[last]) continue; /* alreadyFilled (decode.wgsl:433) */
      const uint64_t parent = out[i];
      for (uint32_t c = 0; c < 8; ++c) {
        const uint32_t child = 8 * m + c;
        const uint32_t j = node_to_out(child, l + 1, t);
        uint32_t nibv, delta = 0;
        if (!read_nibble(&r, &nibv)) { *status |= ST_OVERRUN; goto done; }
        const uint32_t op = nibv >> 1, stop = nibv & 1;
        if (op == 6 && !read_nibble(&r, &delta)) { *status |= ST_OVERRUN; goto done; }
        uint64_t val = 0;
        switch (op) {
          case 0: val = parent; break;
          case 1: case 2: case 3: {
            /* getNeighborValue (brickDecoding.cpp:77-124) */
            const int32_t cx = (int32_t)compact1by2(child), cy = (int32_t)compact1by2(child >> 1),
                          cz = (int32_t)compact1by2(child >> 2);
            int32_t nx = cx, ny = cy, nz = cz;
            if (op == 1) nx = (cx & 1) == 0 ? cx - 1 : cx + 1;
            if (op == 2) ny = (cy & 1) == 0 ? cy - 1 : cy + 1;
            if (op == 3) nz = (cz & 1) == 0 ? cz - 1 : cz + 1;
            const int32_t maxc = (int32_t)((1u << (l + 1)) - 1);
            if (nx < 0 || ny < 0 || nz < 0 || nx > maxc || ny > maxc || nz > maxc) {
              val = out[node_to_out(morton3(cx >> 1, cy >> 1, cz >> 1), l, t)];
            } else {
              const uint32_t nm = morton3(nx, ny, nz);
              if (nm < child || OR_MUTANT == 1) val = out[node_to_out(nm, l + 1, t)];
              else val = out[node_to_out(morton3(nx >> 1, ny >> 1, nz >> 1), l, t)];
            }
            break;
          }
          case 4:
            if (OR_MUTANT != 5) ++ip;
            if (ip >= P) { *status |= ST_PALETTE; val = pal[P - 1]; } else val = pal[ip];
            break;
          case 5:
            if (ip >= P) { *status |= ST_PALETTE; val = pal[P - 1]; } else val = pal[ip];
            break;
          case 6: {
            const int64_t k = ip - (int64_t)delta - (OR_MUTANT == 2 ? 0 : 1);
            if (k < 0 || k >= P) { *status |= ST_PALETTE; val = pal[P - 1]; } else val = pal[k];
            break;
          }
          default:
            *status |= ST_BADOP;
            val = 0;
        }
        out[j] = val;
        filled[j] = 1;
        if (stop) {
          /* fillSubBlock (brickDecoding.cpp:61-75) */
          const uint64_t bs = 1ull << (3 * (t - l - 1));
          for (uint64_t k = 0; k < bs; ++k) { out[j + k] = val; filled[j + k] = 1; }
        }
      }
    }
  }
done:
  free(filled);
  return 0;
}

int or_decode_dataset_brick(const or_dataset* d, uint32_t brick, uint32_t t, uint64_t* out,
                            uint32_t* status) {
  if (brick >= d->num_bricks) { *status = ST_OVERRUN; return -1; }
  const or_brick* b = &d->bricks[brick];
  uint8_t* nib = (uint8_t*)malloc((b->nsym + 2) / 2 + 1);
  uint32_t st1 = 0, st2 = 0;
  const size_t nb = or_rans_decode(d, brick, nib, &st1);
  const int rc = or_decode_brick(nib, nb, d->palette + b->pal_off, b->P, t, out, &st2);
  free(nib);
  *status = st1 | st2;
  return rc;
}

/* ------------------------------------------------------------- LOD rule - */
/* calculateBrickLOD (camera.js:103-137) */
uint32_t or_brick_lod(double cx, double cy, double cz, double brick_size, double px, double py,
                      double pz, double fov, double screen_width, uint32_t max_lod) {
  const double dx = cx - px, dy = cy - py, dz = cz - pz;
  double d = sqrt(dx * dx + dy * dy + dz * dz);
  if (d < 0.1) d = 0.1;
  const double focal = (screen_width / 2) / tan(fov / 2);
  uint32_t best = 0;
  for (uint32_t lod = 0; lod <= max_lod; ++lod) {
    const double voxel = brick_size / pow(2.0, (double)lod);
    const double pixels = (focal * voxel) / d;
    if (pixels >= 1.0) best = lod; else break;
  }
  return best;
}

/* ------------------------------------------------------------------ DDA - */
/* rayTrace.wgsl:457-563 and the traversal of 565-892, FP32 throughout,
 * with the zero-direction fix: a zero component has tMax = tDelta = +inf
 * and no slab constraint when the origin lies within that slab. */
typedef struct { float o[3], d[3], inv[3]; int zero[3]; } ray_t;

static float fmaxf_(float a, float b) { return a > b ? a : b; }
static float fminf_(float a, float b) { return a < b ? a : b; }

static int intersect_volume(const ray_t* r, float G, float* tenter, float* texit) {
  float tmin = -INFINITY, tmax = INFINITY;
  for (int a = 0; a < 3; ++a) {
    if (r->zero[a]) {
      if (r->o[a] < 0.0f || r->o[a] > G) return 0;
      continue;
    }
    const float t0 = (0.0f - r->o[a]) * r->inv[a];
    const float t1 = (G - r->o[a]) * r->inv[a];
    tmin = fmaxf_(tmin, fminf_(t0, t1));
    tmax = fminf_(tmax, fmaxf_(t0, t1));
  }
  if (tmax < 0.0f || tmin > tmax) return 0;
  *tenter = fmaxf_(tmin, 0.0f);
  *texit = tmax;
  return 1;
}

static float adv(int v[3], float tmax[3], const float tdelta[3], const int step[3]) {
  /* advanceDDA / advanceBrickDDA: ties prefer z, then y (strict <) */
  float t;
  if (tmax[0] < tmax[1] && tmax[0] < tmax[2]) { t = tmax[0]; tmax[0] += tdelta[0]; v[0] += step[0]; }
  else if (OR_MUTANT == 7 ? tmax[1] <= tmax[2] : tmax[1] < tmax[2]) { t = tmax[1]; tmax[1] += tdelta[1]; v[1] += step[1]; }
  else { t = tmax[2]; tmax[2] += tdelta[2]; v[2] += step[2]; }
  return t;
}

static float bias(float t) { return fmaxf_(t * 1e-6f, 1e-5f); }
/* the entry offset of initBrickDDA / initDDA (rayTrace.wgsl:522, 476) */
static float entry_bias(float t) { return OR_MUTANT == 8 ? 0.0f : bias(t); }

/* Scene of one frame; the compositing fields are used by or_render_rgba. */
typedef struct {
  uint32_t B, b;
  const uint8_t* empty;
  const uint8_t* lod;
  const uint64_t* const* labels;
  const uint64_t* pal0;
  int rgba;
  const uint64_t* tf_labels;
  const uint32_t* tf_rgba;
  uint64_t tf_n;
  float light[3], ambient, diffuse;
} scene_t;

/* ------------------------------------------------------------ shading --- */
/* decode.wgsl:135-172: transfer function entry, else a hash colour */
static uint32_t hash_colour(uint64_t label) {
  uint32_t lo = (uint32_t)label, hi = (uint32_t)(label >> 32);
  lo = lo ^ hi;
  lo = lo ^ (lo >> 16);
  lo = lo * 0x7feb352du;
  lo = lo ^ (lo >> 15);
  hi = hi ^ lo;
  hi = hi ^ (hi >> 16);
  hi = hi * 0x85ebca6bu;
  hi = hi ^ (hi >> 13);
  return (lo & 0xFFu) | (((lo >> 8) & 0xFFu) << 8) | (((hi >> 8) & 0xFFu) << 16) | (0x40u << 24);
}

static uint32_t tf_colour(const scene_t* sc, uint64_t label) {
  if (label == 0) return 0u;
  uint64_t lo = 0, hi = sc->tf_n;
  while (lo < hi) { /* binary search in the ascending labels */
    const uint64_t mid = lo + (hi - lo) / 2;
    if (sc->tf_labels[mid] < label) lo = mid + 1; else hi = mid;
  }
  if (lo < sc->tf_n && sc->tf_labels[lo] == label) return sc->tf_rgba[lo];
  return hash_colour(label);
}

static float alpha_of(uint32_t rgba) { return (float)((rgba >> 24) & 0xFFu) / 255.0f; }

static float powi(float x, uint32_t n) { /* repeated squaring */
  float r = 1.0f, p = x;
  while (n) {
    if (n & 1u) r = r * p;
    p = p * p;
    n >>= 1;
  }
  return r;
}

/* getVoxelColor's lookup: label of a voxel of brick bi at its frame LOD;
 * 0 when the brick is empty or has no labels (not resident) */
static int brick_label(const scene_t* sc, uint32_t bi, int lx, int ly, int lz, uint64_t* out) {
  if (sc->empty[bi]) return 0;
  const uint32_t lod = sc->lod[bi];
  if (lod == 0) { *out = sc->pal0[bi]; return 1; }
  if (!sc->labels[bi]) return 0;
  const int b = (int)sc->b, lodSize = 1 << lod;
  const float lodScale = (float)lodSize / (float)b;
  const int loc[3] = {lx, ly, lz};
  uint32_t cc[3];
  for (int k = 0; k < 3; ++k) {
    const int l = loc[k] < 0 ? 0 : (loc[k] > b - 1 ? b - 1 : loc[k]);
    float s = (float)l * lodScale;
    s = s < 0.0f ? 0.0f : (s > (float)(lodSize - 1) ? (float)(lodSize - 1) : s);
    cc[k] = (uint32_t)s;
  }
  *out = sc->labels[bi][morton3(cc[0], cc[1], cc[2])];
  return 1;
}

/* sampleAlpha (rayTrace.wgsl:262-298) */
static float sample_alpha(const scene_t* sc, int x, int y, int z) {
  const int G = (int)(sc->B * sc->b), b = (int)sc->b;
  if (x < 0 || y < 0 || z < 0 || x >= G || y >= G || z >= G) return 0.0f;
  const int bx = x / b, by = y / b, bz = z / b;
  uint64_t lab;
  if (!brick_label(sc, morton3((uint32_t)bx, (uint32_t)by, (uint32_t)bz), x - bx * b, y - by * b, z - bz * b, &lab))
    return 0.0f;
  return alpha_of(tf_colour(sc, lab));
}

/* sampleAlphaInBrick (rayTrace.wgsl:301-318) */
static float sample_alpha_in(const scene_t* sc, uint32_t bi, const int bmin[3], int x, int y, int z) {
  const int G = (int)(sc->B * sc->b), b = (int)sc->b;
  if (x < 0 || y < 0 || z < 0 || x >= G || y >= G || z >= G) return 0.0f;
  const int lx = x - bmin[0], ly = y - bmin[1], lz = z - bmin[2];
  if (lx < 0 || ly < 0 || lz < 0 || lx >= b || ly >= b || lz >= b) return sample_alpha(sc, x, y, z);
  uint64_t lab;
  if (!brick_label(sc, bi, lx, ly, lz, &lab)) return 0.0f;
  return alpha_of(tf_colour(sc, lab));
}

/* traceShadowRay (rayTrace.wgsl:329-395) */
static float shadow_ray(const scene_t* sc, const float start[3], const float normal[3], float accum,
                        float transmittance) {
  if (transmittance < 0.1f || accum < 0.2f) return 1.0f;
  const uint32_t s64 = (uint32_t)(64.0f * transmittance);
  const uint32_t steps = s64 > 24u ? s64 : 24u;
  float ll = sc->light[0] * sc->light[0];
  ll = ll + sc->light[1] * sc->light[1];
  ll = ll + sc->light[2] * sc->light[2];
  const float len = sqrtf(ll);
  float dir[3], rs[3];
  for (int k = 0; k < 3; ++k) {
    dir[k] = sc->light[k] / len;
    rs[k] = start[k] + normal[k] * 0.6f;
  }
  const float G = (float)(sc->B * sc->b);
  float t = 0.5f, occ = 0.0f;
  for (uint32_t i = 0; i < steps; ++i) {
    float p[3];
    for (int k = 0; k < 3; ++k) p[k] = rs[k] + dir[k] * t;
    if (p[0] < 0.0f || p[1] < 0.0f || p[2] < 0.0f || p[0] >= G || p[1] >= G || p[2] >= G) break;
    const float al = sample_alpha(sc, (int)floorf(p[0]), (int)floorf(p[1]), (int)floorf(p[2]));
    occ = occ + al * 1.0f;
    if (occ > 0.8f) return 0.0f;
    t = t + 1.0f;
  }
  const float vis = expf(-occ * 1.5f);
  return 1.0f * (1.0f - transmittance) + vis * transmittance;
}

/* rayTrace.wgsl:801-837: colour of one sample of alpha al at voxel vp */
static void shade_sample(const scene_t* sc, uint32_t bi, const int bmin[3], const int vp[3], int stride, float al,
                         float ca, float* sr, float* sg, float* sb) {
  float wp[3];
  for (int k = 0; k < 3; ++k) wp[k] = (float)vp[k] + 0.5f;
  const float xp = sample_alpha_in(sc, bi, bmin, vp[0] + stride, vp[1], vp[2]);
  const float xn = sample_alpha_in(sc, bi, bmin, vp[0] - stride, vp[1], vp[2]);
  const float yp = sample_alpha_in(sc, bi, bmin, vp[0], vp[1] + stride, vp[2]);
  const float yn = sample_alpha_in(sc, bi, bmin, vp[0], vp[1] - stride, vp[2]);
  const float zp = sample_alpha_in(sc, bi, bmin, vp[0], vp[1], vp[2] + stride);
  const float zn = sample_alpha_in(sc, bi, bmin, vp[0], vp[1], vp[2] - stride);
  const float g[3] = {((xp > 0.05f ? xp : al) - (xn > 0.05f ? xn : al)) * 0.5f,
                      ((yp > 0.05f ? yp : al) - (yn > 0.05f ? yn : al)) * 0.5f,
                      ((zp > 0.05f ? zp : al) - (zn > 0.05f ? zn : al)) * 0.5f};
  float gl = g[0] * g[0];
  gl = gl + g[1] * g[1];
  gl = gl + g[2] * g[2];
  const float glen = fmaxf_(sqrtf(gl), 0.001f);
  float nrm[3];
  for (int k = 0; k < 3; ++k) nrm[k] = -g[k] / glen;
  float sf = 1.0f;
  if (al > 0.4f) sf = shadow_ray(sc, wp, nrm, ca, 1.0f - ca);
  float ll = sc->light[0] * sc->light[0];
  ll = ll + sc->light[1] * sc->light[1];
  ll = ll + sc->light[2] * sc->light[2];
  const float llen = sqrtf(ll);
  float dl = nrm[0] * (sc->light[0] / llen);
  dl = dl + nrm[1] * (sc->light[1] / llen);
  dl = dl + nrm[2] * (sc->light[2] / llen);
  const float diff = fmaxf_(0.0f, dl);
  const float ds = sc->diffuse * sf;
  const float mina = diff > 0.01f ? 0.3f : 0.1f;
  const float shf = fminf_(fmaxf_(sc->ambient + diff * ds, mina), 1.0f);
  float lum = *sr * 0.299f;
  lum = lum + *sg * 0.587f;
  lum = lum + *sb * 0.114f;
  const float smix = fminf_(fmaxf_(1.0f - sf, 0.0f), 1.0f) * 0.5f;
  float* c3[3] = {sr, sg, sb};
  for (int k = 0; k < 3; ++k) {
    const float gray = lum * 0.8f + *c3[k] * 0.2f;
    const float base = *c3[k] * (1.0f - smix) + gray * smix;
    *c3[k] = base * shf;
  }
}

/* One pixel: the first hit (label != 0) and, in the compositing form, the
 * shaded front-to-back composite (rayTrace.wgsl:565-906). */
static void march_pixel(const scene_t* sc, const or_camera* cam, uint32_t w, uint32_t h, uint32_t px, uint32_t py,
                        uint64_t* out_label, uint64_t* out_voxel, float* out_t, uint32_t* out_rgba,
                        uint32_t* visited) {
  const uint32_t B = sc->B, b = sc->b;
  const float Wf = (float)w, Hf = (float)h;
  const float aspect = (float)((double)w / (double)h);
  const float G = (float)(B * b);
  const uint64_t Gi = (uint64_t)B * b;
  uint64_t hit_label = 0, hit_vox = UINT64_MAX;
  float hit_t = -1.0f;
  float cr = 0.0f, cg = 0.0f, cb = 0.0f, ca = 0.0f;
  /* rayTrace.wgsl:597-619 */
  float ndc_x = ((float)px + 0.5f) / Wf;
  ndc_x = ndc_x * 2.0f;
  ndc_x = ndc_x - 1.0f;
  float ndc_y = ((float)py + 0.5f) / Hf;
  ndc_y = ndc_y * 2.0f;
  ndc_y = 1.0f - ndc_y;
  ray_t r;
  if (cam->ortho_half_height > 0.0f) {
    for (int a = 0; a < 3; ++a) {
      float ox = cam->right[a] * ndc_x;
      ox = ox * aspect;
      ox = ox * cam->ortho_half_height;
      float oy = cam->up[a] * ndc_y;
      oy = oy * cam->ortho_half_height;
      r.o[a] = (cam->position[a] + ox) + oy;
      r.d[a] = cam->forward[a];
    }
  } else {
    float v[3];
    for (int a = 0; a < 3; ++a) {
      float rt = cam->right[a] * ndc_x;
      rt = rt * aspect;
      rt = rt * cam->tan_half_fov;
      float ut = cam->up[a] * ndc_y;
      ut = ut * cam->tan_half_fov;
      v[a] = (cam->forward[a] + rt) + ut;
      r.o[a] = cam->position[a];
    }
    float dd = v[0] * v[0];
    dd = dd + v[1] * v[1];
    dd = dd + v[2] * v[2];
    const float len = sqrtf(dd);
    for (int a = 0; a < 3; ++a) r.d[a] = v[a] / len;
  }
  for (int a = 0; a < 3; ++a) {
    r.zero[a] = r.d[a] == 0.0f;
    const float s = r.d[a] > 0.0f ? 1.0f : (r.d[a] < 0.0f ? -1.0f : 0.0f);
    r.inv[a] = r.zero[a] ? INFINITY : s / fmaxf_(fabsf(r.d[a]), 1e-6f);
  }
  float tEnter, tExit;
  if (!intersect_volume(&r, G, &tEnter, &tExit)) goto store;
  {
    /* initBrickDDA (rayTrace.wgsl:519-545) */
    int step[3], brick[3];
    float btmax[3], btdelta[3];
    {
      const float ts = tEnter + entry_bias(tEnter);
      for (int a = 0; a < 3; ++a) {
        const float pos = r.o[a] + r.d[a] * ts;
        brick[a] = (int)floorf(pos) / (int)b;
        step[a] = r.d[a] > 0.0f ? 1 : (r.d[a] < 0.0f ? -1 : 0);
        const float bf = (float)brick[a] * (float)b;
        const float nb = r.d[a] > 0.0f ? bf + (float)b : bf;
        btmax[a] = r.zero[a] ? INFINITY : (nb - r.o[a]) * r.inv[a];
        btdelta[a] = r.zero[a] ? INFINITY : fabsf((float)b * r.inv[a]);
      }
    }
    int be[3], bx[3];
    uint32_t maxb = 6;
    for (int a = 0; a < 3; ++a) {
      be[a] = (int)(r.o[a] + r.d[a] * tEnter) / (int)b;
      bx[a] = (int)(r.o[a] + r.d[a] * tExit) / (int)b;
      maxb += (uint32_t)abs(bx[a] - be[a]);
    }
    float brickEntryT = tEnter;
    for (uint32_t it = 0; it < maxb; ++it) {
      if (brick[0] < 0 || brick[1] < 0 || brick[2] < 0 || brick[0] >= (int)B || brick[1] >= (int)B ||
          brick[2] >= (int)B)
        break;
      const uint32_t bi = morton3((uint32_t)brick[0], (uint32_t)brick[1], (uint32_t)brick[2]);
      if (!sc->empty[bi]) {
        if (visited) visited[bi] = 1;
        const int bmin[3] = {brick[0] * (int)b, brick[1] * (int)b, brick[2] * (int)b};
        const float tEnterBrick = fmaxf_(brickEntryT, tEnter);
        float tExitBrick = fminf_(fminf_(btmax[0], btmax[1]), btmax[2]);
        tExitBrick = fmaxf_(tEnterBrick, tExitBrick - bias(tExitBrick));
        /* initDDA (rayTrace.wgsl:473-499) */
        int vox[3], vstep[3];
        float vtmax[3], vtdelta[3];
        const float ts = tEnterBrick + entry_bias(tEnterBrick);
        for (int a = 0; a < 3; ++a) {
          const float pos = r.o[a] + r.d[a] * ts;
          vox[a] = (int)floorf(pos);
          vstep[a] = (OR_MUTANT == 9 && a == 0) ? -step[a] : step[a];
          const float nb = r.d[a] > 0.0f ? (float)(vox[a] + 1) : (float)vox[a];
          vtmax[a] = r.zero[a] ? INFINITY : (nb - r.o[a]) * r.inv[a];
          vtdelta[a] = r.zero[a] ? INFINITY : fabsf(r.inv[a]);
        }
        const uint32_t lod = sc->lod[bi];
        const int lodSize = 1 << lod;
        const int stride = (int)b / lodSize;
        /* a brick without labels at its LOD is not resident: the ray passes
         * through it (the device counts a miss there) */
        const int resident = lod == 0 || sc->labels[bi] != NULL;
        for (int a = 0; a < 3; ++a) vox[a] -= bmin[a];
        if (stride > 1 && OR_MUTANT != 4) {
          for (int a = 0; a < 3; ++a) {
            int lv = (vox[a] / stride) * stride;
            lv = lv < 0 ? 0 : (lv > (int)b - 1 ? (int)b - 1 : lv);
            vox[a] = lv;
            /* the reference keeps a unit voxel step here (rayTrace.wgsl:721-744)
             * while tDelta covers `stride` voxels, so coarse cells are sampled
             * out of place; the cell walk steps by the stride */
            vstep[a] = vstep[a] * stride;
            vtdelta[a] = r.zero[a] ? INFINITY : fabsf(r.inv[a]) * (float)stride;
            const float lw = (float)bmin[a] + (float)vox[a];
            const float nb = r.d[a] > 0.0f ? lw + (float)stride : lw;
            vtmax[a] = r.zero[a] ? INFINITY : (nb - r.o[a]) * r.inv[a];
          }
        } else {
          for (int a = 0; a < 3; ++a) vox[a] = vox[a] < 0 ? 0 : (vox[a] > (int)b - 1 ? (int)b - 1 : vox[a]);
        }
        uint32_t maxs = 128;
        for (int a = 0; a < 3; ++a) {
          const float ve = r.o[a] + r.d[a] * tEnterBrick;
          const float vx = r.o[a] + r.d[a] * tExitBrick;
          const float le = (ve - (float)bmin[a]) / (float)stride;
          const float lx = (vx - (float)bmin[a]) / (float)stride;
          maxs += (uint32_t)abs((int)ceilf(lx) - (int)floorf(le));
        }
        float tcur = tEnterBrick;
        const float lodScale = (float)lodSize / (float)b;
        for (uint32_t s = 0; s < (resident ? maxs : 0u); ++s) {
          const int inb = vox[0] >= 0 && vox[1] >= 0 && vox[2] >= 0 && vox[0] < (int)b && vox[1] < (int)b &&
                          vox[2] < (int)b;
          if (inb) {
            /* getVoxelColor (rayTrace.wgsl:160-204); LOD 0 = palette[0] */
            uint64_t lab;
            if (lod == 0) {
              lab = sc->pal0[bi];
            } else {
              uint32_t cc[3];
              for (int a = 0; a < 3; ++a) {
                float scl = (float)vox[a] * lodScale;
                scl = scl < 0.0f ? 0.0f : (scl > (float)(lodSize - 1) ? (float)(lodSize - 1) : scl);
                cc[a] = (uint32_t)scl;
              }
              lab = sc->labels[bi][morton3(cc[0], cc[1], cc[2])];
            }
            if (lab != 0 && hit_t < 0.0f) {
              hit_label = lab;
              hit_vox = (uint64_t)(bmin[0] + vox[0]) +
                        Gi * ((uint64_t)(bmin[1] + vox[1]) + Gi * (uint64_t)(bmin[2] + vox[2]));
              hit_t = tcur;
              if (!sc->rgba) goto store;
            }
            if (sc->rgba) {
              const uint32_t rgba = tf_colour(sc, lab);
              const float al = alpha_of(rgba);
              if (al > 0.05f) {
                float sr = (float)(rgba & 0xFFu) / 255.0f, sg = (float)((rgba >> 8) & 0xFFu) / 255.0f,
                      sb = (float)((rgba >> 16) & 0xFFu) / 255.0f;
                if (al > 0.35f) {
                  const int vp[3] = {bmin[0] + vox[0], bmin[1] + vox[1], bmin[2] + vox[2]};
                  shade_sample(sc, bi, bmin, vp, stride, al, ca, &sr, &sg, &sb);
                }
                const float corr = fminf_(fmaxf_(1.0f - powi(1.0f - al, (uint32_t)(OR_MUTANT == 6 ? 1 : stride)), 0.0f), 1.0f);
                const float src = corr * (1.0f - ca);
                cr = cr + sr * src;
                cg = cg + sg * src;
                cb = cb + sb * src;
                ca = ca + src;
                if (ca > 0.95f) break;
              }
            }
          }
          const float t = adv(vox, vtmax, vtdelta, vstep);
          tcur = t;
          if (t > tExitBrick) break;
        }
        if (sc->rgba && ca > 0.95f) break;
      }
      const float t = adv(brick, btmax, btdelta, step);
      brickEntryT = t + bias(t);
      if (t > tExit) break;
    }
  }
store:
  if (out_label) *out_label = hit_label;
  if (out_voxel) *out_voxel = hit_vox;
  if (out_t) *out_t = hit_t;
  if (out_rgba)
    *out_rgba = (uint32_t)(cr * 255.0f) | ((uint32_t)(cg * 255.0f) << 8) | ((uint32_t)(cb * 255.0f) << 16) |
                ((uint32_t)(ca * 255.0f) << 24);
}

int or_render(uint32_t B, uint32_t b, const uint8_t* brick_empty, const uint8_t* brick_lod,
              const uint64_t* const* brick_labels, const uint64_t* palette0, const or_camera* cam,
              uint32_t w, uint32_t h, uint64_t* out_label, uint64_t* out_voxel, float* out_t,
              uint32_t* visited) {
  scene_t sc;
  memset(&sc, 0, sizeof sc);
  sc.B = B;
  sc.b = b;
  sc.empty = brick_empty;
  sc.lod = brick_lod;
  sc.labels = brick_labels;
  sc.pal0 = palette0;
  for (uint32_t py = 0; py < h; ++py)
    for (uint32_t px = 0; px < w; ++px) {
      const size_t pix = (size_t)py * w + px;
      march_pixel(&sc, cam, w, h, px, py, out_label ? out_label + pix : NULL, out_voxel ? out_voxel + pix : NULL,
                  out_t ? out_t + pix : NULL, NULL, visited);
    }
  return 0;
}

int or_render_rgba(uint32_t B, uint32_t b, const uint8_t* brick_empty, const uint8_t* brick_lod,
                   const uint64_t* const* brick_labels, const uint64_t* palette0, const or_camera* cam,
                   uint32_t w, uint32_t h, const uint64_t* tf_labels, const uint32_t* tf_rgba, uint64_t tf_n,
                   const float light[3], float ambient, float diffuse, uint32_t* out_rgba) {
  scene_t sc;
  memset(&sc, 0, sizeof sc);
  sc.B = B;
  sc.b = b;
  sc.empty = brick_empty;
  sc.lod = brick_lod;
  sc.labels = brick_labels;
  sc.pal0 = palette0;
  sc.rgba = 1;
  sc.tf_labels = tf_labels;
  sc.tf_rgba = tf_rgba;
  sc.tf_n = tf_n;
  for (int k = 0; k < 3; ++k) sc.light[k] = light[k];
  sc.ambient = ambient;
  sc.diffuse = diffuse;
  for (uint32_t py = 0; py < h; ++py)
    for (uint32_t px = 0; px < w; ++px)
      march_pixel(&sc, cam, w, h, px, py, NULL, NULL, NULL, out_rgba + (size_t)py * w + px, NULL);
  return 0;
}

/* ---------------------------------------------- threaded full decode ---- */
typedef struct {
  const or_dataset* d;
  uint64_t* out;
  uint32_t next;
  pthread_mutex_t mu;
  uint32_t status;
  uint32_t L;
} pool_t;

static void* pool_worker(void* arg) {
  pool_t* P = (pool_t*)arg;
  const uint64_t per = ((1ull << (3 * (P->L + 1))) - 1) / 7;
  uint8_t* nib = NULL;
  size_t cap = 0;
  uint32_t st_acc = 0;
  for (;;) {
    pthread_mutex_lock(&P->mu);
    const uint32_t i = P->next++;
    pthread_mutex_unlock(&P->mu);
    if (i >= P->d->num_bricks) break;
    const or_brick* b = &P->d->bricks[i];
    const size_t need = (b->nsym + 2) / 2 + 1;
    if (need > cap) { free(nib); nib = (uint8_t*)malloc(need); cap = need; }
    uint32_t st = 0;
    const size_t nb = or_rans_decode(P->d, i, nib, &st);
    st_acc |= st;
    for (uint32_t t = 0; t <= P->L; ++t) {
      uint32_t s2 = 0;
      or_decode_brick(nib, nb, P->d->palette + b->pal_off, b->P, t,
                      P->out + per * i + ((1ull << (3 * t)) - 1) / 7, &s2);
      st_acc |= s2;
    }
  }
  free(nib);
  pthread_mutex_lock(&P->mu);
  P->status |= st_acc;
  pthread_mutex_unlock(&P->mu);
  return NULL;
}

uint64_t or_decode_all_pyramid(const or_dataset* d, uint64_t* out, int threads, uint32_t* status_or) {
  uint32_t L = 0;
  while ((1u << L) < d->brick_size) ++L;
  pool_t P;
  P.d = d;
  P.out = out;
  P.next = 0;
  P.status = 0;
  P.L = L;
  pthread_mutex_init(&P.mu, NULL);
  if (threads < 1) threads = 1;
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * threads);
  for (int i = 0; i < threads; ++i) pthread_create(&th[i], NULL, pool_worker, &P);
  for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
  free(th);
  pthread_mutex_destroy(&P.mu);
  if (status_or) *status_or = P.status;
  return (uint64_t)d->num_bricks * (((1ull << (3 * (L + 1))) - 1) / 7);
}

/* ------------------------------------------- threaded request decode ---- */
/* The checker for batched device launches: request k is (brick[k], lod[k]),
 * its 8^lod labels written at out + offset[k] in the label width of `out32`
 * (uint32 when non-NULL, else uint64 into out64); status words per request.
 * Same restatement as or_decode_dataset_brick, run over a thread pool. */
typedef struct {
  const or_dataset* d;
  const uint32_t* brick;
  const uint8_t* lod;
  const uint64_t* offset;
  uint32_t* out32;
  uint64_t* out64;
  uint32_t* status;
  uint32_t n, next;
  pthread_mutex_t mu;
} req_pool_t;

static void* req_worker(void* arg) {
  req_pool_t* P = (req_pool_t*)arg;
  uint8_t* nib = NULL;
  size_t cap = 0;
  uint64_t* tmp = NULL;
  size_t tcap = 0;
  for (;;) {
    pthread_mutex_lock(&P->mu);
    const uint32_t k = P->next;
    P->next += 64;
    pthread_mutex_unlock(&P->mu);
    if (k >= P->n) break;
    const uint32_t end = k + 64 < P->n ? k + 64 : P->n;
    for (uint32_t q = k; q < end; ++q) {
      const uint32_t i = P->brick[q], t = P->lod[q];
      if (i >= P->d->num_bricks) { P->status[q] = ST_OVERRUN; continue; }
      const or_brick* b = &P->d->bricks[i];
      const size_t need = (b->nsym + 2) / 2 + 1;
      if (need > cap) { free(nib); nib = (uint8_t*)malloc(need); cap = need; }
      const size_t nl = (size_t)1 << (3 * t);
      if (nl > tcap) { free(tmp); tmp = (uint64_t*)malloc(nl * 8); tcap = nl; }
      uint32_t st1 = 0, st2 = 0;
      const size_t nb = or_rans_decode(P->d, i, nib, &st1);
      uint64_t* dst = P->out32 ? tmp : P->out64 + P->offset[q];
      or_decode_brick(nib, nb, P->d->palette + b->pal_off, b->P, t, dst, &st2);
      if (P->out32)
        for (size_t j = 0; j < nl; ++j) P->out32[P->offset[q] + j] = (uint32_t)tmp[j];
      P->status[q] = st1 | st2;
    }
  }
  free(nib);
  free(tmp);
  return NULL;
}

int or_decode_requests(const or_dataset* d, const uint32_t* brick, const uint8_t* lod, uint32_t n,
                       const uint64_t* offset, uint32_t* out32, uint64_t* out64, int threads,
                       uint32_t* status) {
  req_pool_t P;
  P.d = d;
  P.brick = brick;
  P.lod = lod;
  P.offset = offset;
  P.out32 = out32;
  P.out64 = out64;
  P.status = status;
  P.n = n;
  P.next = 0;
  pthread_mutex_init(&P.mu, NULL);
  if (threads < 1) threads = 1;
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * threads);
  for (int i = 0; i < threads; ++i) pthread_create(&th[i], NULL, req_worker, &P);
  for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
  free(th);
  pthread_mutex_destroy(&P.mu);
  return 0;
}
