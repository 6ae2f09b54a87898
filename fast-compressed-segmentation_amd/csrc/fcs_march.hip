// DDA marcher for gfx950: one ray per lane over the brick cache.
#include <hip/hip_runtime.h>

#include "fcs_common.h"
#include "fcs_device.h"
#include "fcs_kernels.h"

namespace {

// ====================================================================== DDA
// One ray per lane, rayTrace.wgsl:565-892 up to the first hit, FP32 in the
// WGSL operation order (built with -ffp-contract=off, IEEE div/sqrt).
// A zero direction component has tMax = tDelta = +inf and no slab test.

__device__ __forceinline__ float fmx(float a, float b) { return a > b ? a : b; }
// x / 2^s with C's truncation toward zero (brick sizes and LOD strides are powers of two)
__device__ __forceinline__ int tdiv2(int x, int s) { return x >= 0 ? (x >> s) : -((-x) >> s); }
__device__ __forceinline__ float fmn(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float biasf(float t) { return fmx(t * 1e-6f, 1e-5f); }

// advanceDDA / advanceBrickDDA (rayTrace.wgsl:501-563): ties prefer z, then y.
// Written with selects so the axis choice never becomes an indexed access.
__device__ __forceinline__ float advance3(int& v0, int& v1, int& v2, float& m0, float& m1, float& m2,
                                          float d0, float d1, float d2, int s0, int s1, int s2) {
  const bool ax = m0 < m1 && m0 < m2;
  const bool ay = !ax && m1 < m2;
  const bool az = !ax && !ay;
  const float t = ax ? m0 : (ay ? m1 : m2);
  m0 = ax ? m0 + d0 : m0;
  m1 = ay ? m1 + d1 : m1;
  m2 = az ? m2 + d2 : m2;
  v0 += ax ? s0 : 0;
  v1 += ay ? s1 : 0;
  v2 += az ? s2 : 0;
  return t;
}

// advance3 that also steps the Morton code `mc` of the cell by one along the
// chosen axis (dilated-integer increment or decrement under the axis mask M0,
// M1 or M2; a cell stepping out of the brick wraps, and is never read)
__device__ __forceinline__ float advance3m(int& v0, int& v1, int& v2, float& m0, float& m1, float& m2, float d0,
                                           float d1, float d2, int s0, int s1, int s2, uint32_t& mc, uint32_t M0,
                                           uint32_t M1, uint32_t M2) {
  const bool ax = m0 < m1 && m0 < m2;
  const bool ay = !ax && m1 < m2;
  const float t = advance3(v0, v1, v2, m0, m1, m2, d0, d1, d2, s0, s1, s2);
  const uint32_t M = ax ? M0 : (ay ? M1 : M2);
  const int sg = ax ? s0 : (ay ? s1 : s2);
  const uint32_t u = sg > 0 ? ((mc | ~M) + 1u) : ((mc & M) - 1u);
  mc = (u & M) | (mc & ~M);
  return t;
}

// ------------------------------------------------------------ shading
// The RGBA form follows rayTrace.wgsl:565-906 past the first hit: labels are
// coloured by the transfer function (decode.wgsl:135-172 lookup, hash colour
// for labels outside it), samples with alpha > 0.05 are composited front to
// back with the LOD opacity correction 1 - (1 - a)^stride, samples with
// alpha > 0.35 are shaded from an LOD-stride central-difference gradient
// (computeGradientLOD, 428-455) with a diffuse term and, above alpha 0.4, a
// shadow ray (traceShadowRay, 329-395); the march ends at alpha > 0.95.
// Samples of bricks that are not resident at their frame LOD read as
// transparent. (1 - a)^stride is computed by repeated squaring (exact same
// order on host and device); the shadow falloff uses expf.

__device__ __forceinline__ uint32_t hash_colour(uint64_t label) {
  // decode.wgsl:149-171
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

template <typename LT>
__device__ __forceinline__ uint32_t tf_colour(const MarchArgs<LT>& a, LT label) {
  if (label == 0) return 0u;
  const uint64_t key = (uint64_t)label;
  uint32_t h = fcs_tf_slot(key, a.tf_mask);
  for (uint32_t probe = 0; probe <= a.tf_mask; ++probe) {
    const uint64_t k = a.tf_keys[h];
    if (k == key) return a.tf_rgba[h];
    if (k == 0) break;
    h = (h + 1) & a.tf_mask;
  }
  return hash_colour(key);
}

__device__ __forceinline__ float alpha_of(uint32_t rgba) { return (float)((rgba >> 24) & 0xFFu) / 255.0f; }

// x^n for an integer n >= 1 by repeated squaring
__device__ __forceinline__ float powi(float x, uint32_t n) {
  float r = 1.0f, p = x;
  while (n) {
    if (n & 1u) r = r * p;
    p = p * p;
    n >>= 1;
  }
  return r;
}

// Label at a voxel of brick `bi` quantised to the brick's frame LOD; false
// when the brick is empty or not resident (getVoxelColor, rayTrace.wgsl:160-204).
template <typename LT>
__device__ __forceinline__ bool brick_label(const MarchArgs<LT>& a, uint32_t bi, int lx, int ly, int lz, LT& out) {
  const uint32_t meta = a.meta[bi];
  if (meta & FCS_META_EMPTY) return false;
  const uint32_t lod = meta & FCS_META_LOD_MASK;
  if (lod == 0) {
    out = a.pal0[bi];
    return true;
  }
  if (!(meta & FCS_META_RESIDENT)) return false;
  const int b = (int)a.brick_size;
  const int lodSize = 1 << lod;
  const float lodScale = (float)lodSize / (float)b;
  const int loc[3] = {lx, ly, lz};
  uint32_t cc[3];
  _Pragma("unroll") for (int k = 0; k < 3; ++k) {
    const int l = loc[k] < 0 ? 0 : (loc[k] > b - 1 ? b - 1 : loc[k]);
    float sc = (float)l * lodScale;
    sc = sc < 0.0f ? 0.0f : (sc > (float)(lodSize - 1) ? (float)(lodSize - 1) : sc);
    cc[k] = (uint32_t)sc;
  }
  out = a.cache[a.offs[bi] + fcs_morton3(cc[0], cc[1], cc[2])];
  return true;
}

// sampleAlpha (rayTrace.wgsl:262-298): alpha at a world voxel
template <typename LT>
__device__ __forceinline__ float sample_alpha(const MarchArgs<LT>& a, int x, int y, int z) {
  const int G = (int)(a.bricks_per_axis * a.brick_size);
  if (x < 0 || y < 0 || z < 0 || x >= G || y >= G || z >= G) return 0.0f;
  const int b = (int)a.brick_size;
  const int bx = x / b, by = y / b, bz = z / b;
  const uint32_t bi = fcs_morton3((uint32_t)bx, (uint32_t)by, (uint32_t)bz);
  LT lab;
  if (!brick_label(a, bi, x - bx * b, y - by * b, z - bz * b, lab)) return 0.0f;
  return alpha_of(tf_colour(a, lab));
}

// sampleAlphaInBrick (rayTrace.wgsl:301-318)
template <typename LT>
__device__ __forceinline__ float sample_alpha_in(const MarchArgs<LT>& a, uint32_t bi, const int bmin[3], int x, int y,
                                                 int z) {
  const int G = (int)(a.bricks_per_axis * a.brick_size);
  if (x < 0 || y < 0 || z < 0 || x >= G || y >= G || z >= G) return 0.0f;
  const int b = (int)a.brick_size;
  const int lx = x - bmin[0], ly = y - bmin[1], lz = z - bmin[2];
  if (lx < 0 || ly < 0 || lz < 0 || lx >= b || ly >= b || lz >= b) return sample_alpha(a, x, y, z);
  LT lab;
  if (!brick_label(a, bi, lx, ly, lz, lab)) return 0.0f;
  return alpha_of(tf_colour(a, lab));
}

// traceShadowRay (rayTrace.wgsl:329-395)
template <typename LT>
__device__ float shadow_ray(const MarchArgs<LT>& a, const float start[3], const float normal[3], float accum,
                            float transmittance) {
  if (transmittance < 0.1f || accum < 0.2f) return 1.0f;
  const uint32_t steps = max(24u, (uint32_t)(64.0f * transmittance));
  float ll = a.light[0] * a.light[0];
  ll = ll + a.light[1] * a.light[1];
  ll = ll + a.light[2] * a.light[2];
  const float len = sqrtf(ll);
  float dir[3], rs[3];
  _Pragma("unroll") for (int k = 0; k < 3; ++k) {
    dir[k] = a.light[k] / len;
    rs[k] = start[k] + normal[k] * 0.6f;
  }
  const float G = a.grid;
  float t = 0.5f, occ = 0.0f;
  for (uint32_t i = 0; i < steps; ++i) {
    float p[3];
    _Pragma("unroll") for (int k = 0; k < 3; ++k) p[k] = rs[k] + dir[k] * t;
    if (p[0] < 0.0f || p[1] < 0.0f || p[2] < 0.0f || p[0] >= G || p[1] >= G || p[2] >= G) break;
    const float al = sample_alpha(a, (int)floorf(p[0]), (int)floorf(p[1]), (int)floorf(p[2]));
    occ = occ + al * 1.0f;
    if (occ > 0.8f) return 0.0f;
    t = t + 1.0f;
  }
  const float vis = expf(-occ * 1.5f);
  return 1.0f * (1.0f - transmittance) + vis * transmittance;
}

template <typename LT, bool RGBA, bool STATS>
__global__ __launch_bounds__(128) void march_kernel(MarchArgs<LT> a) {
  // a workgroup's 16x8 pixels as two 8x8 tiles side by side, one per wave: a
  // wave's rays start closer together than in a 16x4 strip and run more alike;
  // two-wave workgroups finish a frame's ragged tail sooner than four-wave ones
  const uint32_t wv = threadIdx.x >> 6, ln = threadIdx.x & 63u;
  const uint32_t px = blockIdx.x * 16 + wv * 8 + (ln & 7u);
  const uint32_t py = blockIdx.y * 8 + (ln >> 3);
  if (px >= a.w || py >= a.h) return;
  if (a.run_if && *a.run_if == 0u) return;
  const size_t pix = (size_t)py * a.w + px;
  if (a.pix_only && a.pix_only[pix] == 0u) return;
  if (a.cam_dev) a.cam = *a.cam_dev;
  bool missed = false;  // entered a non-resident brick
  uint32_t n_samples = 0, n_bricks = 0, n_empty = 0;  // STATS: what the ray read
  LT hit_label = 0;
  uint64_t hit_vox = ~0ull;
  float hit_t = -1.0f;
  float cr = 0.0f, cg = 0.0f, cb = 0.0f, ca = 0.0f;  // composited colour (RGBA form)
  const int b = (int)a.brick_size;
  const int lb = __builtin_ctz(a.brick_size);  // log2 of the brick size
  const int B = (int)a.bricks_per_axis;
  const float G = a.grid;

  float ndc_x = ((float)px + 0.5f) / a.wf;
  ndc_x = ndc_x * 2.0f;
  ndc_x = ndc_x - 1.0f;
  float ndc_y = ((float)py + 0.5f) / a.hf;
  ndc_y = ndc_y * 2.0f;
  ndc_y = 1.0f - ndc_y;
  float o[3], d[3], inv[3];
  bool zero[3];
  if (a.cam.ortho_half_height > 0.0f) {
    _Pragma("unroll") for (int k = 0; k < 3; ++k) {
      float ox = a.cam.right[k] * ndc_x;
      ox = ox * a.aspect;
      ox = ox * a.cam.ortho_half_height;
      float oy = a.cam.up[k] * ndc_y;
      oy = oy * a.cam.ortho_half_height;
      o[k] = (a.cam.position[k] + ox) + oy;
      d[k] = a.cam.forward[k];
    }
  } else {
    float v[3];
    _Pragma("unroll") for (int k = 0; k < 3; ++k) {
      float rt = a.cam.right[k] * ndc_x;
      rt = rt * a.aspect;
      rt = rt * a.cam.tan_half_fov;
      float ut = a.cam.up[k] * ndc_y;
      ut = ut * a.cam.tan_half_fov;
      v[k] = (a.cam.forward[k] + rt) + ut;
      o[k] = a.cam.position[k];
    }
    float dd = v[0] * v[0];
    dd = dd + v[1] * v[1];
    dd = dd + v[2] * v[2];
    const float len = sqrtf(dd);
    _Pragma("unroll") for (int k = 0; k < 3; ++k) d[k] = v[k] / len;
  }
  _Pragma("unroll") for (int k = 0; k < 3; ++k) {
    zero[k] = d[k] == 0.0f;
    const float s = d[k] > 0.0f ? 1.0f : (d[k] < 0.0f ? -1.0f : 0.0f);
    inv[k] = zero[k] ? __builtin_inff() : s / fmx(fabsf(d[k]), 1e-6f);
  }
  // slab test (rayTrace.wgsl:457-471)
  float tmin = -__builtin_inff(), tmax = __builtin_inff();
  bool miss = false;
  _Pragma("unroll") for (int k = 0; k < 3; ++k) {
    if (zero[k]) {
      if (o[k] < 0.0f || o[k] > G) miss = true;
      continue;
    }
    const float t0 = (0.0f - o[k]) * inv[k];
    const float t1 = (G - o[k]) * inv[k];
    tmin = fmx(tmin, fmn(t0, t1));
    tmax = fmn(tmax, fmx(t0, t1));
  }
  if (!miss && !(tmax < 0.0f || tmin > tmax)) {
    const float tEnter = fmx(tmin, 0.0f), tExit = tmax;
    int step[3], brick[3];
    float btmax[3], btdelta[3];
    {
      const float ts = tEnter + biasf(tEnter);
      _Pragma("unroll") for (int k = 0; k < 3; ++k) {
        const float pos = o[k] + d[k] * ts;
        brick[k] = tdiv2((int)floorf(pos), lb);
        step[k] = d[k] > 0.0f ? 1 : (d[k] < 0.0f ? -1 : 0);
        const float bf = (float)brick[k] * (float)b;
        const float nb = d[k] > 0.0f ? bf + (float)b : bf;
        btmax[k] = zero[k] ? __builtin_inff() : (nb - o[k]) * inv[k];
        btdelta[k] = zero[k] ? __builtin_inff() : fabsf((float)b * inv[k]);
      }
    }
    uint32_t maxb = 6;
    _Pragma("unroll") for (int k = 0; k < 3; ++k) {
      const int be = tdiv2((int)(o[k] + d[k] * tEnter), lb);
      const int bx = tdiv2((int)(o[k] + d[k] * tExit), lb);
      maxb += (uint32_t)abs(bx - be);
    }
    float brickEntryT = tEnter;
    for (uint32_t it = 0; it < maxb; ++it) {
      if (brick[0] < 0 || brick[1] < 0 || brick[2] < 0 || brick[0] >= B || brick[1] >= B || brick[2] >= B) break;
      const uint32_t bi = fcs_morton3((uint32_t)brick[0], (uint32_t)brick[1], (uint32_t)brick[2]);
      const uint32_t meta = a.meta[bi];
      if constexpr (STATS) {
        if (meta & FCS_META_EMPTY) ++n_empty;
        else ++n_bricks;
      }
      if (!(meta & FCS_META_EMPTY)) {
        if (a.visited) {
          if (missed && a.needed) a.needed[bi] = 1u;
          else a.visited[bi] = 1u;
        }
        const int bmin0 = brick[0] * b, bmin1 = brick[1] * b, bmin2 = brick[2] * b;
        const float tEnterBrick = fmx(brickEntryT, tEnter);
        float tExitBrick = fmn(fmn(btmax[0], btmax[1]), btmax[2]);
        tExitBrick = fmx(tEnterBrick, tExitBrick - biasf(tExitBrick));
        int vox[3];
        float vtmax[3], vtdelta[3];
        const float ts = tEnterBrick + biasf(tEnterBrick);
        const int bmin[3] = {bmin0, bmin1, bmin2};
        _Pragma("unroll") for (int k = 0; k < 3; ++k) {
          const float pos = o[k] + d[k] * ts;
          vox[k] = (int)floorf(pos);
          const float nb = d[k] > 0.0f ? (float)(vox[k] + 1) : (float)vox[k];
          vtmax[k] = zero[k] ? __builtin_inff() : (nb - o[k]) * inv[k];
          vtdelta[k] = zero[k] ? __builtin_inff() : fabsf(inv[k]);
        }
        const uint32_t lod = meta & FCS_META_LOD_MASK;
        const int lodSize = 1 << lod;
        const int ls = lb - (int)lod;  // log2 of the LOD stride (lod <= log2 b)
        const int stride = 1 << ls;
        const float inv_stride = 1.0f / (float)stride;  // exact: x / stride == x * inv_stride
        _Pragma("unroll") for (int k = 0; k < 3; ++k) vox[k] -= bmin[k];
        int vstep[3] = {step[0], step[1], step[2]};
        if (stride > 1) {
          _Pragma("unroll") for (int k = 0; k < 3; ++k) {
            int lvx = tdiv2(vox[k], ls) * stride;
            lvx = lvx < 0 ? 0 : (lvx > b - 1 ? b - 1 : lvx);
            vox[k] = lvx;
            vstep[k] = step[k] * stride;  // one LOD cell per step
            vtdelta[k] = zero[k] ? __builtin_inff() : fabsf(inv[k]) * (float)stride;
            const float lw = (float)bmin[k] + (float)vox[k];
            const float nb = d[k] > 0.0f ? lw + (float)stride : lw;
            vtmax[k] = zero[k] ? __builtin_inff() : (nb - o[k]) * inv[k];
          }
        } else {
          _Pragma("unroll") for (int k = 0; k < 3; ++k) vox[k] = vox[k] < 0 ? 0 : (vox[k] > b - 1 ? b - 1 : vox[k]);
        }
        uint32_t maxs = 128;
        _Pragma("unroll") for (int k = 0; k < 3; ++k) {
          const float ve = o[k] + d[k] * tEnterBrick;
          const float vx = o[k] + d[k] * tExitBrick;
          const float le = (ve - (float)bmin[k]) * inv_stride;
          const float lx = (vx - (float)bmin[k]) * inv_stride;
          maxs += (uint32_t)abs((int)ceilf(lx) - (int)floorf(le));
        }
        const bool resident = (meta & FCS_META_RESIDENT) != 0;
        if (lod != 0 && !resident) {
          if (a.misses) atomicAdd(a.misses, 1u);
          missed = true;
        } else {
          const LT* lab = a.cache + a.offs[bi];
          float tcur = tEnterBrick;
          const uint32_t cm = (1u << (3 * lod)) - 1u;  // Morton codes of the brick's cells
          const uint32_t M0 = 0x09249249u & cm, M1 = (0x09249249u << 1) & cm, M2 = (0x09249249u << 2) & cm;
          // the LOD cell of an in-brick voxel: vox * lodSize / b (exact in FP32,
          // within [0, lodSize)) is vox >> log2(stride); stepped along with the DDA
          uint32_t mcur = fcs_morton3((uint32_t)vox[0] >> ls, (uint32_t)vox[1] >> ls, (uint32_t)vox[2] >> ls);
          for (uint32_t s = 0; s < maxs; ++s) {
            if (vox[0] >= 0 && vox[1] >= 0 && vox[2] >= 0 && vox[0] < b && vox[1] < b && vox[2] < b) {
              if constexpr (STATS) ++n_samples;
              LT v;
              if (lod == 0) {
                v = a.pal0[bi];
              } else {
                v = lab[mcur];
              }
              if (v != 0 && hit_t < 0.0f) {
                hit_label = v;
                const uint64_t Gi = (uint64_t)B * (uint64_t)b;
                hit_vox = (uint64_t)(bmin[0] + vox[0]) +
                          Gi * ((uint64_t)(bmin[1] + vox[1]) + Gi * (uint64_t)(bmin[2] + vox[2]));
                hit_t = tcur;
                if (!RGBA) break;
              }
              if constexpr (RGBA) {
                const uint32_t rgba = tf_colour(a, v);
                const float al = alpha_of(rgba);
                if (al > 0.05f) {
                  float sr = (float)(rgba & 0xFFu) / 255.0f, sg = (float)((rgba >> 8) & 0xFFu) / 255.0f,
                        sb = (float)((rgba >> 16) & 0xFFu) / 255.0f;
                  if (al > 0.35f) {
                    // rayTrace.wgsl:807-837
                    const int vp[3] = {bmin[0] + vox[0], bmin[1] + vox[1], bmin[2] + vox[2]};
                    float wp[3];
                    _Pragma("unroll") for (int k = 0; k < 3; ++k) wp[k] = (float)vp[k] + 0.5f;
                    const float xp = sample_alpha_in(a, bi, bmin, vp[0] + stride, vp[1], vp[2]);
                    const float xn = sample_alpha_in(a, bi, bmin, vp[0] - stride, vp[1], vp[2]);
                    const float yp = sample_alpha_in(a, bi, bmin, vp[0], vp[1] + stride, vp[2]);
                    const float yn = sample_alpha_in(a, bi, bmin, vp[0], vp[1] - stride, vp[2]);
                    const float zp = sample_alpha_in(a, bi, bmin, vp[0], vp[1], vp[2] + stride);
                    const float zn = sample_alpha_in(a, bi, bmin, vp[0], vp[1], vp[2] - stride);
                    const float g[3] = {((xp > 0.05f ? xp : al) - (xn > 0.05f ? xn : al)) * 0.5f,
                                        ((yp > 0.05f ? yp : al) - (yn > 0.05f ? yn : al)) * 0.5f,
                                        ((zp > 0.05f ? zp : al) - (zn > 0.05f ? zn : al)) * 0.5f};
                    float gl = g[0] * g[0];
                    gl = gl + g[1] * g[1];
                    gl = gl + g[2] * g[2];
                    const float glen = fmx(sqrtf(gl), 0.001f);
                    float nrm[3];
                    _Pragma("unroll") for (int k = 0; k < 3; ++k) nrm[k] = -g[k] / glen;
                    float sf = 1.0f;
                    if (al > 0.4f) sf = shadow_ray(a, wp, nrm, ca, 1.0f - ca);
                    float ll = a.light[0] * a.light[0];
                    ll = ll + a.light[1] * a.light[1];
                    ll = ll + a.light[2] * a.light[2];
                    const float llen = sqrtf(ll);
                    float dl = nrm[0] * (a.light[0] / llen);
                    dl = dl + nrm[1] * (a.light[1] / llen);
                    dl = dl + nrm[2] * (a.light[2] / llen);
                    const float diff = fmx(0.0f, dl);
                    const float ds = a.diffuse * sf;
                    const float mina = diff > 0.01f ? 0.3f : 0.1f;
                    const float shf = fmn(fmx(a.ambient + diff * ds, mina), 1.0f);
                    float lum = sr * 0.299f;
                    lum = lum + sg * 0.587f;
                    lum = lum + sb * 0.114f;
                    const float smix = fmn(fmx(1.0f - sf, 0.0f), 1.0f) * 0.5f;
                    const float c3[3] = {sr, sg, sb};
                    float o3[3];
                    _Pragma("unroll") for (int k = 0; k < 3; ++k) {
                      const float gray = lum * 0.8f + c3[k] * 0.2f;
                      const float base = c3[k] * (1.0f - smix) + gray * smix;
                      o3[k] = base * shf;
                    }
                    sr = o3[0];
                    sg = o3[1];
                    sb = o3[2];
                  }
                  const float corr = fmn(fmx(1.0f - powi(1.0f - al, (uint32_t)stride), 0.0f), 1.0f);
                  const float src = corr * (1.0f - ca);
                  cr = cr + sr * src;
                  cg = cg + sg * src;
                  cb = cb + sb * src;
                  ca = ca + src;
                  if (ca > 0.95f) break;
                }
              }
            }
            const float tt = advance3m(vox[0], vox[1], vox[2], vtmax[0], vtmax[1], vtmax[2], vtdelta[0],
                                       vtdelta[1], vtdelta[2], vstep[0], vstep[1], vstep[2], mcur, M0, M1, M2);
            tcur = tt;
            if (tt > tExitBrick) break;
          }
          if (!RGBA && hit_t >= 0.0f) break;
          if (RGBA && ca > 0.95f) break;
        }
      }
      const float tt = advance3(brick[0], brick[1], brick[2], btmax[0], btmax[1], btmax[2], btdelta[0],
                                btdelta[1], btdelta[2], step[0], step[1], step[2]);
      brickEntryT = tt + biasf(tt);
      if (tt > tExit) break;
    }
  }
  if constexpr (STATS) {
    // spread over FCS_MARCH_STATS_SLOTS slots per counter (the host sums them)
    const uint32_t slot = (uint32_t)(pix % FCS_MARCH_STATS_SLOTS);
    if (n_samples) atomicAdd(&a.stats[slot], (unsigned long long)n_samples);
    if (n_bricks) atomicAdd(&a.stats[FCS_MARCH_STATS_SLOTS + slot], (unsigned long long)n_bricks);
    if (n_empty) atomicAdd(&a.stats[2 * FCS_MARCH_STATS_SLOTS + slot], (unsigned long long)n_empty);
  }
  if (a.pix_miss) a.pix_miss[pix] = missed ? 1u : 0u;
  if (a.out_label) a.out_label[pix] = hit_label;
  if (a.out_voxel) a.out_voxel[pix] = hit_vox;
  if (a.out_t) a.out_t[pix] = hit_t;
  if constexpr (RGBA) {
    // rayTrace.wgsl:899-905
    a.out_rgba[pix] = (uint32_t)(cr * 255.0f) | ((uint32_t)(cg * 255.0f) << 8) | ((uint32_t)(cb * 255.0f) << 16) |
                      ((uint32_t)(ca * 255.0f) << 24);
  }
}

}  // namespace

namespace {
// One thread per brick. The LOD rule is calculateBrickLOD (camera.js:103-137)
// in float64 with the host rule's exact operation order: focal =
// (min(w,h)/2)/tan(fov/2), d = |brick centre - camera| clamped to 0.1, the
// largest t with (focal * b/2^t)/d >= 1. Division and square root are the
// correctly rounded IEEE operations on both sides and nothing is contracted,
// so device and host pick the same LOD for every brick.
__device__ __forceinline__ uint32_t lod_rule(const MarchCam& c, uint32_t i, uint32_t b, uint32_t L, uint32_t w,
                                             uint32_t h) {
  const double sw = (double)(w < h ? w : h);
  const double focal = (sw / 2.0) / (double)c.tan_half_fov;
  const double half = (double)b / 2.0;
  const double cx = (double)fcs_unmorton_x(i) * (double)b + half;
  const double cy = (double)fcs_unmorton_y(i) * (double)b + half;
  const double cz = (double)fcs_unmorton_z(i) * (double)b + half;
  const double dx = cx - (double)c.position[0], dy = cy - (double)c.position[1], dz = cz - (double)c.position[2];
  double d = sqrt(dx * dx + dy * dy + dz * dz);
  if (d < 0.1) d = 0.1;
  uint32_t lod = 0;
  for (uint32_t l = 0; l <= L; ++l) {
    const double voxel = (double)b / (double)(1u << l);  // exact power of two
    if ((focal * voxel) / d >= 1.0) lod = l; else break;
  }
  return lod;
}

// View-frustum test of brick i (culling.js:87-171 with exact perspective
// planes): the brick's box grown by one voxel lies on the inner side of the
// four side planes and in front of the camera. Orthographic cameras keep
// every brick. float64, the same operations as the host test.
__device__ __forceinline__ bool in_frustum(const MarchCam& c, uint32_t i, uint32_t b, uint32_t w, uint32_t h) {
  if (c.ortho_half_height > 0.0f || !(c.tan_half_fov > 0.0f)) return true;
  const double A = (double)w / (double)h, T = c.tan_half_fov;
  double n[5][3];
  for (int k = 0; k < 3; ++k) {
    const double f = c.forward[k], r = c.right[k], u = c.up[k];
    n[0][k] = r - A * T * f;
    n[1][k] = -r - A * T * f;
    n[2][k] = u - T * f;
    n[3][k] = -u - T * f;
    n[4][k] = -f;
  }
  const double bd = b;
  const uint32_t bc[3] = {fcs_unmorton_x(i), fcs_unmorton_y(i), fcs_unmorton_z(i)};
  for (int q = 0; q < 5; ++q) {
    double d = 0;
    for (int k = 0; k < 3; ++k) {
      const double lo = bc[k] * bd - 1.0, hi = (bc[k] + 1) * bd + 1.0;
      d += n[q][k] * ((n[q][k] >= 0 ? lo : hi) - (double)c.position[k]);
    }
    if (d > 0) return false;
  }
  return true;
}

// Wave-aggregated slot in a device queue: one atomic per wave.
__device__ __forceinline__ uint32_t queue_slot(bool want, uint32_t* count) {
  const uint64_t m = __ballot(want);
  if (!m) return 0xFFFFFFFFu;
  const uint32_t lane = __lane_id();
  const uint32_t leader = (uint32_t)__ffsll((long long)m) - 1u;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
  base = __shfl(base, (int)leader);
  const uint64_t below = lane ? (m & ((1ull << lane) - 1ull)) : 0ull;
  return want ? base + (uint32_t)__popcll(below) : 0xFFFFFFFFu;
}

__device__ __forceinline__ uint64_t res_word(uint64_t off, uint32_t lod, uint64_t old) {
  return (off & FCS_RES_OFF_MASK) | ((uint64_t)lod << FCS_RES_LOD_SHIFT) | (old & FCS_RES_EMPTY);
}

template <bool kFrustum>
__global__ __launch_bounds__(256) void frame_setup_kernel(FrameSetupArgs a) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  const bool live = i < a.nb;
  uint32_t lod = 0;
  uint64_t r = 0;
  bool resident = false, pyr = false, need = false;
  if (live) {
    if (a.lod_fixed) lod = a.lod_fixed[i] < a.L ? a.lod_fixed[i] : a.L;
    else lod = lod_rule(*a.cam, i, a.b, a.L, a.w, a.h);
    r = a.res[i];
    const uint32_t rl = (uint32_t)((r >> FCS_RES_LOD_SHIFT) & 0xFF);
    pyr = (r & FCS_RES_PYRAMID) != 0;
    resident = rl != FCS_RES_NONE && (pyr ? rl >= lod : rl == lod);
    need = a.queue && !resident && lod > 0 && !(r & FCS_RES_EMPTY) &&
           (!kFrustum || in_frustum(*a.cam, i, a.b, a.w, a.h));
  }
  // every lane of the wave takes part in the ballot
  const uint32_t slot = a.queue ? queue_slot(need, a.qcount) : 0xFFFFFFFFu;
  if (!live) return;
  uint32_t m = lod;
  if (r & FCS_RES_EMPTY) m |= FCS_META_EMPTY;
  uint64_t off = 0;
  if (need) {
    off = (uint64_t)i * a.stride;
    a.queue[slot] = DecodeReq{i, lod, off};
    a.res_w[i] = res_word(off, lod, r);
    m |= FCS_META_RESIDENT;
  } else if (resident) {
    m |= FCS_META_RESIDENT;
    off = (r & FCS_RES_OFF_MASK) + (pyr ? fcs_pyramid_level_offset(lod) : 0);
  }
  a.meta[i] = m;
  a.offs[i] = off;
}

__global__ __launch_bounds__(256) void miss_enqueue_kernel(MissEnqueueArgs a) {
  if (*a.misses == 0u) return;
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  bool need = false;
  uint32_t m = 0;
  if (i < a.nb) {
    m = a.meta[i];
    need = (a.visited[i] || (a.needed && a.needed[i])) && !(m & (FCS_META_RESIDENT | FCS_META_EMPTY)) &&
           (m & FCS_META_LOD_MASK) > 0;
  }
  const uint32_t slot = queue_slot(need, a.qcount);
  if (!need) return;
  const uint32_t lod = m & FCS_META_LOD_MASK;
  const uint64_t off = (uint64_t)i * a.stride;
  a.queue[slot] = DecodeReq{i, lod, off};
  a.res_w[i] = res_word(off, lod, a.res_w[i]);
  a.meta[i] = m | FCS_META_RESIDENT;
  a.offs[i] = off;
}

__global__ __launch_bounds__(256) void queue_settle_kernel(QueueSettleArgs a) {
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  const uint32_t n = min(*a.qcount, a.cap);
  if (k >= n) return;
  const uint32_t st = a.status[k];
  if (!st) return;
  atomicOr(a.err, st);
  const uint32_t i = a.queue[k].brick;
  a.res_w[i] = (FCS_RES_NONE << FCS_RES_LOD_SHIFT) | (a.res_w[i] & FCS_RES_EMPTY);
  if (a.meta) a.meta[i] &= ~FCS_META_RESIDENT;
}

// Prefetch candidate of brick i: its LOD at the look-ahead camera and the
// probability bucket (1024 buckets), or false.
__device__ __forceinline__ bool prefetch_candidate(const PrefetchArgs& a, uint32_t i, uint32_t* lod,
                                                   uint32_t* bucket) {
  const float p = a.probs[i];
  if (!(p >= a.threshold)) return false;
  const uint64_t r = a.res_w[i];
  if (r & FCS_RES_EMPTY) return false;
  const uint32_t t = lod_rule(*a.cam, i, a.b, a.L, a.w, a.h);
  if (t == 0) return false;
  const uint32_t rl = (uint32_t)((r >> FCS_RES_LOD_SHIFT) & 0xFF);
  const bool pyr = (r & FCS_RES_PYRAMID) != 0;
  if (rl != FCS_RES_NONE && (pyr ? rl >= t : rl == t)) return false;
  *lod = t;
  *bucket = min((uint32_t)(p * 1024.0f), 1023u);
  return true;
}

__global__ __launch_bounds__(256) void prefetch_hist_kernel(PrefetchArgs a) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  uint32_t t, bk;
  if (i < a.nb && prefetch_candidate(a, i, &t, &bk)) atomicAdd(&a.hist[bk], 1u);
}

// One wave: the lowest bucket still taken whole (hist[1024]) and how many of
// the bucket below it fit (hist[1025]), from the most probable bucket down.
// Lane l holds buckets 1023 - 16 l - j (j = 0..15); a wave scan gives each
// lane the count above its first bucket, and the first lane whose buckets
// cross the cap writes the cut.
__global__ __launch_bounds__(64) void prefetch_cut_kernel(PrefetchArgs a) {
  const uint32_t lane = threadIdx.x;
  uint32_t c[16], tot = 0;
#pragma unroll
  for (uint32_t j = 0; j < 16; ++j) {
    c[j] = a.hist[1023u - (16u * lane + j)];
    tot += c[j];
  }
  uint32_t incl = tot;
#pragma unroll
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(incl, o);
    if (lane >= o) incl += t;
  }
  uint32_t acc = incl - tot;
  int found = -1;
  uint32_t part = 0;
#pragma unroll
  for (uint32_t j = 0; j < 16; ++j) {
    if (found < 0 && acc + c[j] > a.cap) {
      found = (int)j;
      part = a.cap - acc;
    }
    acc += c[j];
  }
  const uint64_t m = __ballot(found >= 0);
  if (m == 0) {
    if (lane == 0) {
      a.hist[1024] = 0u;
      a.hist[1025] = 0u;
    }
    return;
  }
  if (lane == (uint32_t)(__ffsll((unsigned long long)m) - 1)) {
    a.hist[1024] = 1024u - (16u * lane + (uint32_t)found);
    a.hist[1025] = part;
  }
}

__global__ __launch_bounds__(256) void prefetch_enqueue_kernel(PrefetchArgs a) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  uint32_t t = 0, bk = 0;
  bool want = i < a.nb && prefetch_candidate(a, i, &t, &bk);
  const uint32_t cut = a.hist[1024];
  if (want && bk < cut) {
    // the partly taken bucket: the first `part` claimants
    want = bk + 1u == cut && atomicAdd(&a.hist[1026], 1u) < a.hist[1025];
  }
  const uint32_t slot = queue_slot(want, a.qcount);
  if (!want) return;
  const uint64_t off = (uint64_t)i * a.stride;
  a.queue[slot] = DecodeReq{i, t, off};
  a.res_w[i] = res_word(off, t, a.res_w[i]);
}

__global__ void zero_u32_kernel_m(uint32_t* p, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

__global__ void zero_if_kernel(uint32_t* p, uint64_t n, const uint32_t* flag) {
  if (*flag == 0u) return;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}
}  // namespace

hipError_t fcs_launch_frame_setup(const FrameSetupArgs& a, hipStream_t s, bool frustum) {
  if (!a.nb) return hipSuccess;
  if (frustum)
    hipLaunchKernelGGL(frame_setup_kernel<true>, dim3((a.nb + 255) / 256), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(frame_setup_kernel<false>, dim3((a.nb + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t fcs_launch_miss_enqueue(const MissEnqueueArgs& a, hipStream_t s) {
  if (!a.nb) return hipSuccess;
  hipLaunchKernelGGL(miss_enqueue_kernel, dim3((a.nb + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t fcs_launch_queue_settle(const QueueSettleArgs& a, hipStream_t s) {
  if (!a.cap) return hipSuccess;
  hipLaunchKernelGGL(queue_settle_kernel, dim3((a.cap + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t fcs_launch_prefetch_enqueue(const PrefetchArgs& a, hipStream_t s) {
  if (!a.nb) return hipSuccess;
  const dim3 grid((a.nb + 255) / 256);
  hipLaunchKernelGGL(zero_u32_kernel_m, dim3(5), dim3(256), 0, s, a.hist, (uint64_t)1028);
  hipLaunchKernelGGL(prefetch_hist_kernel, grid, dim3(256), 0, s, a);
  hipLaunchKernelGGL(prefetch_cut_kernel, dim3(1), dim3(64), 0, s, a);
  hipLaunchKernelGGL(prefetch_enqueue_kernel, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t fcs_launch_zero_if(uint32_t* p, uint64_t n, const uint32_t* flag, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(zero_if_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, flag);
  return hipGetLastError();
}

template <typename LT>
hipError_t fcs_launch_march(const MarchArgs<LT>& a, hipStream_t s) {
  const dim3 grid((a.w + 15) / 16, (a.h + 7) / 8);
  if (a.out_rgba) {
    if (!a.tf_keys || !a.tf_rgba) return hipErrorInvalidValue;
    if (a.stats) hipLaunchKernelGGL((march_kernel<LT, true, true>), grid, dim3(128), 0, s, a);
    else hipLaunchKernelGGL((march_kernel<LT, true, false>), grid, dim3(128), 0, s, a);
  } else {
    if (a.stats) hipLaunchKernelGGL((march_kernel<LT, false, true>), grid, dim3(128), 0, s, a);
    else hipLaunchKernelGGL((march_kernel<LT, false, false>), grid, dim3(128), 0, s, a);
  }
  return hipGetLastError();
}

// Order-independent 64-bit checksum of a frame: the sum over pixels of a
// mixed (label, voxel index, pixel index) word, so equal frames give equal
// sums and any changed pixel changes it with overwhelming probability.
namespace {
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Order-independent 64-bit sum over the pixels. A fixed grid strides over
// the frame (four pixels per thread per round, loads issued together), waves
// reduce with lane shuffles and each workgroup adds its sum once: a few
// hundred atomics on the one output word instead of one per 256 pixels, which
// serialised at the cache holding it.
constexpr uint32_t CSUM_GRID = 512;

template <typename LT>
__global__ __launch_bounds__(256) void frame_checksum_kernel(const LT* __restrict__ label,
                                                             const uint64_t* __restrict__ voxel, uint64_t n,
                                                             unsigned long long* out) {
  __shared__ unsigned long long red[4];
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  unsigned long long h = 0;
  for (uint64_t i0 = (uint64_t)blockIdx.x * 256 + threadIdx.x; i0 < n; i0 += 4 * stride) {
    uint64_t lab[4], vox[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint64_t i = i0 + k * stride;
      lab[k] = i < n ? (uint64_t)label[i] : 0ull;
      vox[k] = i < n ? voxel[i] : 0ull;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint64_t i = i0 + k * stride;
      if (i < n) h += mix64(mix64(lab[k] + 0x9E3779B97F4A7C15ull * (i + 1)) ^ vox[k]);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = h;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(out, red[0] + red[1] + red[2] + red[3]);
}
}  // namespace

template <typename LT>
hipError_t fcs_launch_frame_checksum(const LT* label, const uint64_t* voxel, uint64_t n, unsigned long long* out,
                                     hipStream_t s) {
  if (!n) return hipSuccess;
  const uint64_t groups = (n + 255) / 256;
  hipLaunchKernelGGL(frame_checksum_kernel<LT>, dim3((unsigned)(groups < CSUM_GRID ? groups : CSUM_GRID)), dim3(256), 0, s,
                     label, voxel, n, out);
  return hipGetLastError();
}

template hipError_t fcs_launch_frame_checksum<uint32_t>(const uint32_t*, const uint64_t*, uint64_t,
                                                        unsigned long long*, hipStream_t);
template hipError_t fcs_launch_frame_checksum<uint64_t>(const uint64_t*, const uint64_t*, uint64_t,
                                                        unsigned long long*, hipStream_t);

template hipError_t fcs_launch_march<uint32_t>(const MarchArgs<uint32_t>&, hipStream_t);
template hipError_t fcs_launch_march<uint64_t>(const MarchArgs<uint64_t>&, hipStream_t);
