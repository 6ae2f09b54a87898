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

template <typename LT>
__global__ __launch_bounds__(256) void march_kernel(MarchArgs<LT> a) {
  const uint32_t px = blockIdx.x * 16 + (threadIdx.x & 15);
  const uint32_t py = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (px >= a.w || py >= a.h) return;
  if (a.cam_dev) a.cam = *a.cam_dev;
  const size_t pix = (size_t)py * a.w + px;
  LT hit_label = 0;
  uint64_t hit_vox = ~0ull;
  float hit_t = -1.0f;
  const int b = (int)a.brick_size;
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
        brick[k] = (int)floorf(pos) / b;
        step[k] = d[k] > 0.0f ? 1 : (d[k] < 0.0f ? -1 : 0);
        const float bf = (float)brick[k] * (float)b;
        const float nb = d[k] > 0.0f ? bf + (float)b : bf;
        btmax[k] = zero[k] ? __builtin_inff() : (nb - o[k]) * inv[k];
        btdelta[k] = zero[k] ? __builtin_inff() : fabsf((float)b * inv[k]);
      }
    }
    uint32_t maxb = 6;
    _Pragma("unroll") for (int k = 0; k < 3; ++k) {
      const int be = (int)(o[k] + d[k] * tEnter) / b;
      const int bx = (int)(o[k] + d[k] * tExit) / b;
      maxb += (uint32_t)abs(bx - be);
    }
    float brickEntryT = tEnter;
    for (uint32_t it = 0; it < maxb; ++it) {
      if (brick[0] < 0 || brick[1] < 0 || brick[2] < 0 || brick[0] >= B || brick[1] >= B || brick[2] >= B) break;
      const uint32_t bi = fcs_morton3((uint32_t)brick[0], (uint32_t)brick[1], (uint32_t)brick[2]);
      const uint32_t meta = a.meta[bi];
      if (!(meta & FCS_META_EMPTY)) {
        if (a.visited) a.visited[bi] = 1u;
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
        const int stride = b / lodSize;
        _Pragma("unroll") for (int k = 0; k < 3; ++k) vox[k] -= bmin[k];
        int vstep[3] = {step[0], step[1], step[2]};
        if (stride > 1) {
          _Pragma("unroll") for (int k = 0; k < 3; ++k) {
            int lvx = (vox[k] / stride) * stride;
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
          const float le = (ve - (float)bmin[k]) / (float)stride;
          const float lx = (vx - (float)bmin[k]) / (float)stride;
          maxs += (uint32_t)abs((int)ceilf(lx) - (int)floorf(le));
        }
        const bool resident = (meta & FCS_META_RESIDENT) != 0;
        if (lod != 0 && !resident) {
          if (a.misses) atomicAdd(a.misses, 1u);
        } else {
          const LT* lab = a.cache + a.offs[bi];
          const float lodScale = (float)lodSize / (float)b;
          float tcur = tEnterBrick;
          for (uint32_t s = 0; s < maxs; ++s) {
            if (vox[0] >= 0 && vox[1] >= 0 && vox[2] >= 0 && vox[0] < b && vox[1] < b && vox[2] < b) {
              LT v;
              if (lod == 0) {
                v = a.pal0[bi];
              } else {
                uint32_t cc[3];
                _Pragma("unroll") for (int k = 0; k < 3; ++k) {
                  float sc = (float)vox[k] * lodScale;
                  sc = sc < 0.0f ? 0.0f : (sc > (float)(lodSize - 1) ? (float)(lodSize - 1) : sc);
                  cc[k] = (uint32_t)sc;
                }
                v = lab[fcs_morton3(cc[0], cc[1], cc[2])];
              }
              if (v != 0) {
                hit_label = v;
                const uint64_t Gi = (uint64_t)B * (uint64_t)b;
                hit_vox = (uint64_t)(bmin[0] + vox[0]) +
                          Gi * ((uint64_t)(bmin[1] + vox[1]) + Gi * (uint64_t)(bmin[2] + vox[2]));
                hit_t = tcur;
                break;
              }
            }
            const float tt = advance3(vox[0], vox[1], vox[2], vtmax[0], vtmax[1], vtmax[2], vtdelta[0],
                                      vtdelta[1], vtdelta[2], vstep[0], vstep[1], vstep[2]);
            tcur = tt;
            if (tt > tExitBrick) break;
          }
          if (hit_t >= 0.0f) break;
        }
      }
      const float tt = advance3(brick[0], brick[1], brick[2], btmax[0], btmax[1], btmax[2], btdelta[0],
                                btdelta[1], btdelta[2], step[0], step[1], step[2]);
      brickEntryT = tt + biasf(tt);
      if (tt > tExit) break;
    }
  }
  if (a.out_label) a.out_label[pix] = hit_label;
  if (a.out_voxel) a.out_voxel[pix] = hit_vox;
  if (a.out_t) a.out_t[pix] = hit_t;
}

}  // namespace

namespace {
// One thread per brick. The LOD rule is calculateBrickLOD (camera.js:103-137)
// in float64 with the host rule's exact operation order: focal =
// (min(w,h)/2)/tan(fov/2), d = |brick centre - camera| clamped to 0.1, the
// largest t with (focal * b/2^t)/d >= 1. Division and square root are the
// correctly rounded IEEE operations on both sides and nothing is contracted,
// so device and host pick the same LOD for every brick.
__global__ __launch_bounds__(256) void frame_setup_kernel(FrameSetupArgs a) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= a.nb) return;
  uint32_t lod;
  if (a.lod_fixed) {
    lod = a.lod_fixed[i] < a.L ? a.lod_fixed[i] : a.L;
  } else {
    const MarchCam& c = *a.cam;
    const double sw = (double)(a.w < a.h ? a.w : a.h);
    const double focal = (sw / 2.0) / (double)c.tan_half_fov;
    const double half = (double)a.b / 2.0;
    const double cx = (double)fcs_unmorton_x(i) * (double)a.b + half;
    const double cy = (double)fcs_unmorton_y(i) * (double)a.b + half;
    const double cz = (double)fcs_unmorton_z(i) * (double)a.b + half;
    const double dx = cx - (double)c.position[0], dy = cy - (double)c.position[1], dz = cz - (double)c.position[2];
    double d = sqrt(dx * dx + dy * dy + dz * dz);
    if (d < 0.1) d = 0.1;
    lod = 0;
    for (uint32_t l = 0; l <= a.L; ++l) {
      const double voxel = (double)a.b / (double)(1u << l);  // exact power of two
      if ((focal * voxel) / d >= 1.0) lod = l; else break;
    }
  }
  const uint64_t r = a.res[i];
  const uint32_t rl = (uint32_t)((r >> FCS_RES_LOD_SHIFT) & 0xFF);
  const bool pyr = (r & FCS_RES_PYRAMID) != 0;
  const bool resident = rl != FCS_RES_NONE && (pyr ? rl >= lod : rl == lod);
  uint32_t m = lod;
  if (r & FCS_RES_EMPTY) m |= FCS_META_EMPTY;
  uint64_t off = 0;
  if (resident) {
    m |= FCS_META_RESIDENT;
    off = (r & FCS_RES_OFF_MASK) + (pyr ? fcs_pyramid_level_offset(lod) : 0);
  }
  a.meta[i] = m;
  a.offs[i] = off;
}
}  // namespace

hipError_t fcs_launch_frame_setup(const FrameSetupArgs& a, hipStream_t s) {
  if (!a.nb) return hipSuccess;
  hipLaunchKernelGGL(frame_setup_kernel, dim3((a.nb + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

template <typename LT>
hipError_t fcs_launch_march(const MarchArgs<LT>& a, hipStream_t s) {
  const dim3 grid((a.w + 15) / 16, (a.h + 15) / 16);
  hipLaunchKernelGGL(march_kernel<LT>, grid, dim3(256), 0, s, a);
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

template <typename LT>
__global__ void frame_checksum_kernel(const LT* __restrict__ label, const uint64_t* __restrict__ voxel, uint64_t n,
                                      unsigned long long* out) {
  __shared__ unsigned long long red[256];
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long h = 0;
  if (i < n) h = mix64(mix64((uint64_t)label[i] + 0x9E3779B97F4A7C15ull * (i + 1)) ^ voxel[i]);
  red[threadIdx.x] = h;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, red[0]);
}
}  // namespace

template <typename LT>
hipError_t fcs_launch_frame_checksum(const LT* label, const uint64_t* voxel, uint64_t n, unsigned long long* out,
                                     hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(frame_checksum_kernel<LT>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, label, voxel, n,
                     out);
  return hipGetLastError();
}

template hipError_t fcs_launch_frame_checksum<uint32_t>(const uint32_t*, const uint64_t*, uint64_t,
                                                        unsigned long long*, hipStream_t);
template hipError_t fcs_launch_frame_checksum<uint64_t>(const uint64_t*, const uint64_t*, uint64_t,
                                                        unsigned long long*, hipStream_t);

template hipError_t fcs_launch_march<uint32_t>(const MarchArgs<uint32_t>&, hipStream_t);
template hipError_t fcs_launch_march<uint64_t>(const MarchArgs<uint64_t>&, hipStream_t);
