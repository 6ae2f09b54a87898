/* fcs_napi.c — thin N-API addon over the C ABI (include/fcs/fcs.h), so the
 * JavaScript app path (web/app.js:9-12) can call the HIP decoder and marcher
 * from Node.js. Every function validates its arguments, calls exactly one
 * fcs_* entry point and throws a JS Error with fcs_last_error() on failure.
 */
#include <node_api.h>
#include <stdbool.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fcs/fcs.h"

#define NAPI_OK(env, call)                                       \
  do {                                                           \
    if ((call) != napi_ok) {                                     \
      napi_throw_error((env), NULL, "N-API call failed: " #call); \
      return NULL;                                               \
    }                                                            \
  } while (0)

static napi_value throw_fcs(napi_env env, int rc, fcs_ctx* ctx, const char* what) {
  char buf[512];
  snprintf(buf, sizeof buf, "%s failed (%d): %s", what, rc, ctx ? fcs_last_error(ctx) : "");
  napi_throw_error(env, NULL, buf);
  return NULL;
}

typedef struct {
  fcs_ctx* ctx;
} ctx_box;

typedef struct {
  fcs_volume* vol;
  fcs_ctx* ctx;
  fcs_volume_info info;
} vol_box;

static void ctx_finalize(napi_env env, void* data, void* hint) {
  (void)env; (void)hint;
  ctx_box* b = (ctx_box*)data;
  if (b->ctx) fcs_destroy(b->ctx);
  free(b);
}

static void vol_finalize(napi_env env, void* data, void* hint) {
  (void)env; (void)hint;
  vol_box* b = (vol_box*)data;
  if (b->vol) fcs_volume_free(b->vol);
  free(b);
}

static int get_args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, NULL, NULL) != napi_ok) return 0;
  for (size_t i = argc; i < want; ++i) napi_get_undefined(env, &argv[i]);
  return 1;
}

static int get_u32(napi_env env, napi_value v, uint32_t* out) { return napi_get_value_uint32(env, v, out) == napi_ok; }
static int get_f64(napi_env env, napi_value v, double* out) { return napi_get_value_double(env, v, out) == napi_ok; }

static int typed_data(napi_env env, napi_value v, napi_typedarray_type want, void** data, size_t* len) {
  bool is = false;
  if (napi_is_typedarray(env, v, &is) != napi_ok || !is) return 0;
  napi_typedarray_type t;
  napi_value ab;
  size_t off;
  if (napi_get_typedarray_info(env, v, &t, len, data, &ab, &off) != napi_ok) return 0;
  return t == want;
}

static napi_value new_typed(napi_env env, napi_typedarray_type t, size_t n, size_t esize, void** data) {
  napi_value ab, arr;
  if (napi_create_arraybuffer(env, n * esize, data, &ab) != napi_ok) return NULL;
  if (napi_create_typedarray(env, t, n, ab, 0, &arr) != napi_ok) return NULL;
  return arr;
}

/* createContext(device = 0) -> ctx (fcs_create) */
static napi_value js_create_context(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  uint32_t dev = 0;
  napi_valuetype vt;
  napi_typeof(env, argv[0], &vt);
  if (vt == napi_number && !get_u32(env, argv[0], &dev)) return NULL;
  fcs_ctx* c = NULL;
  int rc = fcs_create((int)dev, &c);
  if (rc < 0) return throw_fcs(env, rc, NULL, "fcs_create");
  ctx_box* b = (ctx_box*)calloc(1, sizeof *b);
  b->ctx = c;
  napi_value obj;
  NAPI_OK(env, napi_create_object(env, &obj));
  NAPI_OK(env, napi_wrap(env, obj, b, ctx_finalize, NULL, NULL));
  return obj;
}

static ctx_box* unwrap_ctx(napi_env env, napi_value v) {
  void* p = NULL;
  if (napi_unwrap(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, NULL, "expected a context from createContext()");
    return NULL;
  }
  return (ctx_box*)p;
}

static vol_box* unwrap_vol(napi_env env, napi_value v) {
  void* p = NULL;
  if (napi_unwrap(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, NULL, "expected a volume from loadCsdb()");
    return NULL;
  }
  return (vol_box*)p;
}

static void set_u32(napi_env env, napi_value obj, const char* k, uint32_t v) {
  napi_value x;
  napi_create_uint32(env, v, &x);
  napi_set_named_property(env, obj, k, x);
}
static void set_f64(napi_env env, napi_value obj, const char* k, double v) {
  napi_value x;
  napi_create_double(env, v, &x);
  napi_set_named_property(env, obj, k, x);
}

/* loadCsdb(ctx, Uint8Array bytes, brickSizeOverride = 0) -> volume
 * (fcs_load_csdb; loadDatasetFromArrayBuffer, datasetLoader.js:39-145) */
static napi_value js_load(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  ctx_box* cb = unwrap_ctx(env, argv[0]);
  if (!cb) return NULL;
  void* data;
  size_t len;
  if (!typed_data(env, argv[1], napi_uint8_array, &data, &len)) {
    napi_throw_type_error(env, NULL, "bytes must be a Uint8Array");
    return NULL;
  }
  uint32_t bso = 0;
  napi_valuetype vt;
  napi_typeof(env, argv[2], &vt);
  if (vt == napi_number) get_u32(env, argv[2], &bso);
  fcs_volume* v = NULL;
  int rc = fcs_load_csdb(cb->ctx, (const uint8_t*)data, len, bso, &v);
  if (rc < 0) return throw_fcs(env, rc, cb->ctx, "fcs_load_csdb");
  vol_box* vb = (vol_box*)calloc(1, sizeof *vb);
  vb->vol = v;
  vb->ctx = cb->ctx;
  fcs_volume_get_info(v, &vb->info);
  napi_value obj;
  NAPI_OK(env, napi_create_object(env, &obj));
  NAPI_OK(env, napi_wrap(env, obj, vb, vol_finalize, NULL, NULL));
  napi_value hdr;
  napi_create_object(env, &hdr);
  set_u32(env, hdr, "numBricks", vb->info.num_bricks);
  set_u32(env, hdr, "brickSize", vb->info.brick_size);
  set_u32(env, hdr, "levels", vb->info.levels);
  set_u32(env, hdr, "bricksPerAxis", vb->info.bricks_per_axis);
  set_u32(env, hdr, "gridSize", vb->info.grid_size);
  set_u32(env, hdr, "labelBits", vb->info.label_bits);
  set_f64(env, hdr, "compressedBytes", (double)vb->info.compressed_bytes);
  set_f64(env, hdr, "totalSymbols", (double)vb->info.total_symbols);
  napi_set_named_property(env, obj, "header", hdr);
  return obj;
}

/* decodeRans(volume) -> Uint32Array status per brick (fcs_decode_rans) */
static napi_value js_decode_rans(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  vol_box* vb = unwrap_vol(env, argv[0]);
  if (!vb) return NULL;
  void* st;
  napi_value arr = new_typed(env, napi_uint32_array, vb->info.num_bricks, 4, &st);
  if (!arr) return NULL;
  int rc = fcs_decode_rans(vb->vol, (uint32_t*)st);
  if (rc < 0 && rc != FCS_E_CORRUPT) return throw_fcs(env, rc, vb->ctx, "fcs_decode_rans");
  return arr;
}

/* readNibbles(volume, brick) -> Uint8Array: the brick's nibble stream after
 * the device rANS stage, two nibbles per byte, the first in the high half
 * (one element of decodeAllRansBricks' result, decode.js:17-26) */
static napi_value nibbles_of(napi_env env, vol_box* vb, uint32_t brick) {
  const int64_t n = fcs_read_nibbles(vb->vol, brick, NULL, 0);
  if (n < 0) return throw_fcs(env, (int)n, vb->ctx, "fcs_read_nibbles");
  void* p;
  napi_value arr = new_typed(env, napi_uint8_array, (size_t)n, 1, &p);
  if (!arr) return NULL;
  if (n > 0) {
    const int64_t rc = fcs_read_nibbles(vb->vol, brick, (uint8_t*)p, (uint64_t)n);
    if (rc < 0) return throw_fcs(env, (int)rc, vb->ctx, "fcs_read_nibbles");
  }
  return arr;
}

static napi_value js_read_nibbles(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  vol_box* vb = unwrap_vol(env, argv[0]);
  uint32_t brick;
  if (!vb || !get_u32(env, argv[1], &brick)) return NULL;
  return nibbles_of(env, vb, brick);
}

/* allNibbles(volume) -> Array<Uint8Array>, one nibble stream per brick in
 * brick (Morton ID) order */
static napi_value js_all_nibbles(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  vol_box* vb = unwrap_vol(env, argv[0]);
  if (!vb) return NULL;
  napi_value arr;
  if (napi_create_array_with_length(env, vb->info.num_bricks, &arr) != napi_ok) return NULL;
  for (uint32_t b = 0; b < vb->info.num_bricks; ++b) {
    napi_value e = nibbles_of(env, vb, b);
    if (!e) return NULL;
    napi_set_element(env, arr, b, e);
  }
  return arr;
}

/* cacheClear(volume): drop every resident brick (fcs_cache_clear) */
static napi_value js_cache_clear(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  vol_box* vb = unwrap_vol(env, argv[0]);
  if (!vb) return NULL;
  int rc = fcs_cache_clear(vb->vol);
  if (rc < 0) return throw_fcs(env, rc, vb->ctx, "fcs_cache_clear");
  return NULL;
}

/* cacheResident(volume) -> { bricks: Uint32Array, lods: Uint8Array } (fcs_cache_resident) */
static napi_value js_cache_resident(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  vol_box* vb = unwrap_vol(env, argv[0]);
  if (!vb) return NULL;
  int64_t n = fcs_cache_resident(vb->vol, NULL, NULL, 0);
  if (n < 0) return throw_fcs(env, (int)n, vb->ctx, "fcs_cache_resident");
  void *bp, *lp;
  napi_value jb = new_typed(env, napi_uint32_array, (size_t)n, 4, &bp);
  napi_value jl = new_typed(env, napi_uint8_array, (size_t)n, 1, &lp);
  if (!jb || !jl) return NULL;
  n = fcs_cache_resident(vb->vol, (uint32_t*)bp, (uint8_t*)lp, (uint64_t)n);
  if (n < 0) return throw_fcs(env, (int)n, vb->ctx, "fcs_cache_resident");
  napi_value obj;
  napi_create_object(env, &obj);
  napi_set_named_property(env, obj, "bricks", jb);
  napi_set_named_property(env, obj, "lods", jl);
  return obj;
}

/* requestBricks(volume, Uint32Array bricks, Uint8Array lods)
 *   -> { offsets: Float64Array (voxel offsets), status: Uint32Array } */
static napi_value js_request(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  vol_box* vb = unwrap_vol(env, argv[0]);
  if (!vb) return NULL;
  void *bp, *lp;
  size_t nb, nl;
  if (!typed_data(env, argv[1], napi_uint32_array, &bp, &nb) ||
      !typed_data(env, argv[2], napi_uint8_array, &lp, &nl) || nb != nl) {
    napi_throw_type_error(env, NULL, "bricks: Uint32Array, lods: Uint8Array of the same length");
    return NULL;
  }
  uint64_t* off = (uint64_t*)malloc((nb ? nb : 1) * 8);
  void *stp, *offp;
  napi_value st = new_typed(env, napi_uint32_array, nb, 4, &stp);
  napi_value offs = new_typed(env, napi_float64_array, nb, 8, &offp);
  int rc = fcs_request_bricks(vb->vol, (const uint32_t*)bp, (const uint8_t*)lp, (uint32_t)nb, off, (uint32_t*)stp);
  if (rc < 0 && rc != FCS_E_CORRUPT) { free(off); return throw_fcs(env, rc, vb->ctx, "fcs_request_bricks"); }
  for (size_t i = 0; i < nb; ++i) ((double*)offp)[i] = off[i] == UINT64_MAX ? -1.0 : (double)off[i];
  free(off);
  napi_value obj;
  napi_create_object(env, &obj);
  napi_set_named_property(env, obj, "offsets", offs);
  napi_set_named_property(env, obj, "status", st);
  return obj;
}

/* decodeAll(volume, lod = 255, pyramid = true) (fcs_decode_all) */
static napi_value js_decode_all(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  vol_box* vb = unwrap_vol(env, argv[0]);
  if (!vb) return NULL;
  uint32_t lod = 255;
  bool pyr = true;
  napi_valuetype vt;
  napi_typeof(env, argv[1], &vt);
  if (vt == napi_number) get_u32(env, argv[1], &lod);
  napi_typeof(env, argv[2], &vt);
  if (vt == napi_boolean) napi_get_value_bool(env, argv[2], &pyr);
  int rc = fcs_decode_all(vb->vol, lod, pyr ? 1 : 0);
  if (rc < 0) return throw_fcs(env, rc, vb->ctx, "fcs_decode_all");
  return NULL;
}

/* decodeVolume(volume, lod = 255, pyramid = true) -> Uint32Array status per
 * brick: decodeAllRansBricks then decodeAll in one call (fcs_decode_volume) */
static napi_value js_decode_volume(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  vol_box* vb = unwrap_vol(env, argv[0]);
  if (!vb) return NULL;
  uint32_t lod = 255;
  bool pyr = true;
  napi_valuetype vt;
  napi_typeof(env, argv[1], &vt);
  if (vt == napi_number) get_u32(env, argv[1], &lod);
  napi_typeof(env, argv[2], &vt);
  if (vt == napi_boolean) napi_get_value_bool(env, argv[2], &pyr);
  void* st;
  napi_value arr = new_typed(env, napi_uint32_array, vb->info.num_bricks, 4, &st);
  if (!arr) return NULL;
  int rc = fcs_decode_volume(vb->vol, lod, pyr ? 1 : 0, (uint32_t*)st);
  if (rc < 0 && rc != FCS_E_CORRUPT) return throw_fcs(env, rc, vb->ctx, "fcs_decode_volume");
  return arr;
}

/* readBrick(volume, brick, lod) -> Uint32Array | BigUint64Array (fcs_read_brick) */
static napi_value js_read_brick(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  vol_box* vb = unwrap_vol(env, argv[0]);
  if (!vb) return NULL;
  uint32_t brick, lod;
  if (!get_u32(env, argv[1], &brick) || !get_u32(env, argv[2], &lod) || lod > vb->info.levels) {
    napi_throw_range_error(env, NULL, "brick / lod out of range");
    return NULL;
  }
  const size_t n = (size_t)1 << (3 * lod);
  void* out;
  napi_value arr = vb->info.label_bits == 64 ? new_typed(env, napi_biguint64_array, n, 8, &out)
                                             : new_typed(env, napi_uint32_array, n, 4, &out);
  int rc = fcs_read_brick(vb->vol, brick, lod, out);
  if (rc < 0) return throw_fcs(env, rc, vb->ctx, "fcs_read_brick");
  return arr;
}

static int read_vec3(napi_env env, napi_value obj, const char* k, float* out) {
  napi_value v;
  if (napi_get_named_property(env, obj, k, &v) != napi_ok) return 0;
  for (uint32_t i = 0; i < 3; ++i) {
    napi_value e;
    double d;
    if (napi_get_element(env, v, i, &e) != napi_ok || !get_f64(env, e, &d)) return 0;
    out[i] = (float)d;
  }
  return 1;
}

static int read_camera(napi_env env, napi_value obj, fcs_camera* c) {
  memset(c, 0, sizeof *c);
  napi_value v;
  double d;
  if (!read_vec3(env, obj, "position", c->position) || !read_vec3(env, obj, "forward", c->forward) ||
      !read_vec3(env, obj, "right", c->right) || !read_vec3(env, obj, "up", c->up))
    return 0;
  if (napi_get_named_property(env, obj, "tanHalfFov", &v) == napi_ok && get_f64(env, v, &d)) c->tan_half_fov = (float)d;
  if (napi_get_named_property(env, obj, "orthoHalfHeight", &v) == napi_ok && get_f64(env, v, &d))
    c->ortho_half_height = (float)d;
  return 1;
}

static napi_value camera_object(napi_env env, const fcs_camera* c) {
  napi_value obj;
  napi_create_object(env, &obj);
  const char* keys[4] = {"position", "forward", "right", "up"};
  const float* vals[4] = {c->position, c->forward, c->right, c->up};
  for (int k = 0; k < 4; ++k) {
    napi_value a;
    napi_create_array_with_length(env, 3, &a);
    for (uint32_t i = 0; i < 3; ++i) {
      napi_value e;
      napi_create_double(env, vals[k][i], &e);
      napi_set_element(env, a, i, e);
    }
    napi_set_named_property(env, obj, keys[k], a);
  }
  set_f64(env, obj, "tanHalfFov", c->tan_half_fov);
  set_f64(env, obj, "orthoHalfHeight", c->ortho_half_height);
  return obj;
}

/* orbitCamera(volume, yaw, pitch, radius, fov) -> camera (fcs_camera_orbit) */
static napi_value js_orbit(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  if (!get_args(env, info, 5, argv)) return NULL;
  vol_box* vb = unwrap_vol(env, argv[0]);
  if (!vb) return NULL;
  double yaw, pitch, radius, fov;
  if (!get_f64(env, argv[1], &yaw) || !get_f64(env, argv[2], &pitch) || !get_f64(env, argv[3], &radius) ||
      !get_f64(env, argv[4], &fov)) {
    napi_throw_type_error(env, NULL, "orbitCamera(volume, yaw, pitch, radius, fov)");
    return NULL;
  }
  fcs_camera c;
  int rc = fcs_camera_orbit(vb->vol, yaw, pitch, radius, fov, &c);
  if (rc < 0) return throw_fcs(env, rc, vb->ctx, "fcs_camera_orbit");
  return camera_object(env, &c);
}

/* render(volume, camera, width, height, brickLods?: Uint8Array, shaded?: boolean)
 *   -> { label, voxel (Float64Array, -1 = no hit), tHit, visited, lods,
 *        bricksDecoded, decodeMs, marchMs[, rgba: Uint32Array] } (fcs_render) */
static napi_value js_render(napi_env env, napi_callback_info info) {
  napi_value argv[6];
  if (!get_args(env, info, 6, argv)) return NULL;
  vol_box* vb = unwrap_vol(env, argv[0]);
  if (!vb) return NULL;
  fcs_camera cam;
  uint32_t w, h;
  if (!read_camera(env, argv[1], &cam) || !get_u32(env, argv[2], &w) || !get_u32(env, argv[3], &h) || !w || !h) {
    napi_throw_type_error(env, NULL, "render(volume, camera, width, height[, brickLods])");
    return NULL;
  }
  const uint8_t* lods = NULL;
  bool is = false;
  napi_is_typedarray(env, argv[4], &is);
  if (is) {
    void* lp;
    size_t nl;
    if (!typed_data(env, argv[4], napi_uint8_array, &lp, &nl) || nl != vb->info.num_bricks) {
      napi_throw_type_error(env, NULL, "brickLods must be a Uint8Array with one LOD per brick");
      return NULL;
    }
    lods = (const uint8_t*)lp;
  }
  const size_t n = (size_t)w * h;
  void *lab, *tt, *vis, *used;
  napi_value jl = vb->info.label_bits == 64 ? new_typed(env, napi_biguint64_array, n, 8, &lab)
                                            : new_typed(env, napi_uint32_array, n, 4, &lab);
  napi_value jt = new_typed(env, napi_float32_array, n, 4, &tt);
  napi_value jv = new_typed(env, napi_uint32_array, vb->info.num_bricks, 4, &vis);
  napi_value ju = new_typed(env, napi_uint8_array, vb->info.num_bricks, 1, &used);
  void* voxd;
  napi_value jx = new_typed(env, napi_float64_array, n, 8, &voxd);
  bool shaded = false;
  napi_valuetype st;
  if (napi_typeof(env, argv[5], &st) == napi_ok && st == napi_boolean) napi_get_value_bool(env, argv[5], &shaded);
  void* rgbap = NULL;
  napi_value jr = NULL;
  if (shaded) jr = new_typed(env, napi_uint32_array, n, 4, &rgbap);
  uint64_t* vox = (uint64_t*)malloc(n * 8);
  fcs_frame f;
  memset(&f, 0, sizeof f);
  f.rgba = (uint32_t*)rgbap;
  f.label = lab;
  f.voxel = vox;
  f.t_hit = (float*)tt;
  f.brick_visited = (uint32_t*)vis;
  f.brick_lod_used = (uint8_t*)used;
  int rc = fcs_render(vb->vol, &cam, lods, w, h, &f);
  if (rc < 0) { free(vox); return throw_fcs(env, rc, vb->ctx, "fcs_render"); }
  for (size_t i = 0; i < n; ++i) ((double*)voxd)[i] = vox[i] == UINT64_MAX ? -1.0 : (double)vox[i];
  free(vox);
  napi_value obj;
  napi_create_object(env, &obj);
  napi_set_named_property(env, obj, "label", jl);
  napi_set_named_property(env, obj, "voxel", jx);
  napi_set_named_property(env, obj, "tHit", jt);
  napi_set_named_property(env, obj, "visited", jv);
  napi_set_named_property(env, obj, "lods", ju);
  set_u32(env, obj, "bricksDecoded", f.bricks_decoded);
  set_f64(env, obj, "decodeMs", f.decode_ms);
  set_f64(env, obj, "marchMs", f.march_ms);
  if (jr) napi_set_named_property(env, obj, "rgba", jr);
  return obj;
}

/* setLighting(volume, lightAngle, ambient, diffuseStrength) (fcs_set_lighting) */
static napi_value js_set_lighting(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv)) return NULL;
  vol_box* vb = unwrap_vol(env, argv[0]);
  if (!vb) return NULL;
  double a, amb, dif;
  if (!get_f64(env, argv[1], &a) || !get_f64(env, argv[2], &amb) || !get_f64(env, argv[3], &dif)) {
    napi_throw_type_error(env, NULL, "setLighting(volume, lightAngle, ambient, diffuseStrength)");
    return NULL;
  }
  fcs_lighting l = {(float)a, (float)amb, (float)dif, 0.0f};
  int rc = fcs_set_lighting(vb->vol, &l);
  if (rc < 0) return throw_fcs(env, rc, vb->ctx, "fcs_set_lighting");
  return NULL;
}

/* setLabelVisibility(volume, hidden: BigUint64Array, highlight?: bigint)
 * (fcs_set_label_visibility) */
static napi_value js_set_label_visibility(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  vol_box* vb = unwrap_vol(env, argv[0]);
  if (!vb) return NULL;
  void* hp = NULL;
  size_t nh = 0;
  if (!typed_data(env, argv[1], napi_biguint64_array, &hp, &nh)) {
    napi_throw_type_error(env, NULL, "hidden labels must be a BigUint64Array");
    return NULL;
  }
  uint64_t hl = 0;
  const uint64_t* hlp = NULL;
  napi_valuetype t;
  if (napi_typeof(env, argv[2], &t) == napi_ok && t == napi_bigint) {
    bool lossless = false;
    napi_get_value_bigint_uint64(env, argv[2], &hl, &lossless);
    hlp = &hl;
  }
  int rc = fcs_set_label_visibility(vb->vol, (const uint64_t*)hp, (uint32_t)nh, hlp);
  if (rc < 0) return throw_fcs(env, rc, vb->ctx, "fcs_set_label_visibility");
  return NULL;
}

/* transferFunction(volume) -> { labels: BigUint64Array, rgba: Uint32Array } (fcs_transfer_function) */
static napi_value js_transfer_function(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  vol_box* vb = unwrap_vol(env, argv[0]);
  if (!vb) return NULL;
  int64_t n = fcs_transfer_function(vb->vol, NULL, NULL, 0);
  if (n < 0) return throw_fcs(env, (int)n, vb->ctx, "fcs_transfer_function");
  void *lp, *rp;
  napi_value jl = new_typed(env, napi_biguint64_array, (size_t)n, 8, &lp);
  napi_value jr = new_typed(env, napi_uint32_array, (size_t)n, 4, &rp);
  n = fcs_transfer_function(vb->vol, (uint64_t*)lp, (uint32_t*)rp, (uint64_t)n);
  if (n < 0) return throw_fcs(env, (int)n, vb->ctx, "fcs_transfer_function");
  napi_value obj;
  napi_create_object(env, &obj);
  napi_set_named_property(env, obj, "labels", jl);
  napi_set_named_property(env, obj, "rgba", jr);
  return obj;
}

static napi_value js_device_count(napi_env env, napi_callback_info info) {
  (void)info;
  napi_value v;
  napi_create_int32(env, fcs_device_count(), &v);
  return v;
}

static napi_value init(napi_env env, napi_value exports) {
  napi_property_descriptor d[] = {
      {"deviceCount", NULL, js_device_count, NULL, NULL, NULL, napi_default, NULL},
      {"createContext", NULL, js_create_context, NULL, NULL, NULL, napi_default, NULL},
      {"loadCsdb", NULL, js_load, NULL, NULL, NULL, napi_default, NULL},
      {"decodeRans", NULL, js_decode_rans, NULL, NULL, NULL, napi_default, NULL},
      {"readNibbles", NULL, js_read_nibbles, NULL, NULL, NULL, napi_default, NULL},
      {"allNibbles", NULL, js_all_nibbles, NULL, NULL, NULL, napi_default, NULL},
      {"cacheClear", NULL, js_cache_clear, NULL, NULL, NULL, napi_default, NULL},
      {"cacheResident", NULL, js_cache_resident, NULL, NULL, NULL, napi_default, NULL},
      {"requestBricks", NULL, js_request, NULL, NULL, NULL, napi_default, NULL},
      {"decodeAll", NULL, js_decode_all, NULL, NULL, NULL, napi_default, NULL},
      {"decodeVolume", NULL, js_decode_volume, NULL, NULL, NULL, napi_default, NULL},
      {"readBrick", NULL, js_read_brick, NULL, NULL, NULL, napi_default, NULL},
      {"orbitCamera", NULL, js_orbit, NULL, NULL, NULL, napi_default, NULL},
      {"render", NULL, js_render, NULL, NULL, NULL, napi_default, NULL},
      {"setLighting", NULL, js_set_lighting, NULL, NULL, NULL, napi_default, NULL},
      {"setLabelVisibility", NULL, js_set_label_visibility, NULL, NULL, NULL, napi_default, NULL},
      {"transferFunction", NULL, js_transfer_function, NULL, NULL, NULL, napi_default, NULL},
  };
  napi_define_properties(env, exports, sizeof d / sizeof d[0], d);
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
