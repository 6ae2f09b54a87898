// CPU producer of the .csdb format: synthetic volume recipes and a
// brick-at-a-time, multi-threaded encoder. See include/fcs/fcs_producer.h.
#include "fcs/fcs_producer.h"
#include "fcs/fcs.h"
#include "fcs_common.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <cstdio>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace {

// ---------------------------------------------------------------- RNG ----
// SplitMix64 finaliser; fcs_rng(seed, stream, index) is a counter-based
// generator, so every draw is addressable without sequential state.
inline uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t rng(uint64_t seed, uint64_t stream, uint64_t idx) {
  return splitmix64(splitmix64(splitmix64(seed) ^ stream) ^ idx);
}
inline uint64_t rng_below(uint64_t seed, uint64_t stream, uint64_t idx, uint64_t n) {
  return (uint64_t)(((unsigned __int128)rng(seed, stream, idx) * n) >> 64);
}

enum : uint64_t {
  ST_SEED_X = 1, ST_SEED_Y, ST_SEED_Z, ST_LABEL, ST_SLAB, ST_BLOB, ST_NOISE_X, ST_NOISE_Y,
  ST_NOISE_Z, ST_REGION, ST_REGION_SEEDS
};

constexpr int64_t FP = 256;  // fixed point: 1/256 voxel

// ------------------------------------------------------- value noise -----
// Per-axis value noise on a lattice of spacing 8 voxels, trilinear, values
// in [-256, 255] / 256. Displacement = amp * noise voxels, in 1/256 units.
struct Noise {
  uint64_t seed;
  int amp;
  int64_t lattice(uint64_t stream, int64_t i, int64_t j, int64_t k) const {
    uint64_t key = ((uint64_t)(i & 0x1fffff)) | ((uint64_t)(j & 0x1fffff) << 21) |
                   ((uint64_t)(k & 0x1fffff) << 42);
    return (int64_t)(rng(seed, stream, key) & 511u) - 256;
  }
  int64_t axis(uint64_t stream, int64_t x, int64_t y, int64_t z) const {
    int64_t i0 = x >> 3, j0 = y >> 3, k0 = z >> 3;
    int64_t fx = x & 7, fy = y & 7, fz = z & 7;
    int64_t acc = 0;
    for (int c = 0; c < 8; ++c) {
      int dx = c & 1, dy = (c >> 1) & 1, dz = (c >> 2) & 1;
      int64_t w = (dx ? fx : 8 - fx) * (dy ? fy : 8 - fy) * (dz ? fz : 8 - fz);
      acc += lattice(stream, i0 + dx, j0 + dy, k0 + dz) * w;
    }
    // acc in [-256*512, 255*512]; displacement (1/256 voxel) = amp * acc / 512
    return (acc * amp) >> 9;
  }
};

// ---------------------------------------------------- Voronoi field ------
struct Voronoi {
  // seeds in fixed point, region-local coordinates
  std::vector<int64_t> sx, sy, sz;
  std::vector<uint32_t> label;  // per seed (0 = background cell)
  int64_t wx = 1, wy = 1, wz = 1;
  int64_t ext = 0;  // region edge in voxels
  // grid in metric space
  int64_t cell = 1;
  int gx = 1, gy = 1, gz = 1;
  std::vector<uint32_t> cell_start, cell_items;

  void build_grid() {
    const double n = (double)sx.size();
    double mvol = (double)(wx * ext * FP) * (double)(wy * ext * FP) * (double)(wz * ext * FP);
    double c = std::cbrt(mvol / std::max(1.0, n / 2.0));
    cell = std::max<int64_t>(1, (int64_t)c);
    gx = (int)std::max<int64_t>(1, (wx * ext * FP + cell - 1) / cell);
    gy = (int)std::max<int64_t>(1, (wy * ext * FP + cell - 1) / cell);
    gz = (int)std::max<int64_t>(1, (wz * ext * FP + cell - 1) / cell);
    const size_t ncell = (size_t)gx * gy * gz;
    std::vector<uint32_t> count(ncell + 1, 0);
    std::vector<uint32_t> cid(sx.size());
    for (size_t i = 0; i < sx.size(); ++i) {
      int cx = clampi((int)(wx * sx[i] / cell), gx), cy = clampi((int)(wy * sy[i] / cell), gy),
          cz = clampi((int)(wz * sz[i] / cell), gz);
      cid[i] = (uint32_t)(cx + (size_t)gx * (cy + (size_t)gy * cz));
      count[cid[i] + 1]++;
    }
    for (size_t c2 = 0; c2 < ncell; ++c2) count[c2 + 1] += count[c2];
    cell_start = count;
    cell_items.assign(sx.size(), 0);
    std::vector<uint32_t> fill(count.begin(), count.end() - 1);
    for (size_t i = 0; i < sx.size(); ++i) cell_items[fill[cid[i]]++] = (uint32_t)i;  // ascending i
  }
  static int clampi(int v, int n) { return v < 0 ? 0 : (v >= n ? n - 1 : v); }

  // nearest seed to fixed-point position (px,py,pz); ties -> lowest index
  uint32_t nearest(int64_t px, int64_t py, int64_t pz) const {
    const int64_t mx = wx * px, my = wy * py, mz = wz * pz;
    const int c0x = clampi((int)floordiv(mx, cell), gx), c0y = clampi((int)floordiv(my, cell), gy),
              c0z = clampi((int)floordiv(mz, cell), gz);
    int64_t best = INT64_MAX;
    uint32_t bi = UINT32_MAX;
    const int rmax = std::max(gx, std::max(gy, gz));
    for (int r = 0; r <= rmax; ++r) {
      for (int cz = c0z - r; cz <= c0z + r; ++cz) {
        if (cz < 0 || cz >= gz) continue;
        for (int cy = c0y - r; cy <= c0y + r; ++cy) {
          if (cy < 0 || cy >= gy) continue;
          const bool inner_yz = std::abs(cz - c0z) < r && std::abs(cy - c0y) < r;
          for (int cx = c0x - r; cx <= c0x + r; ++cx) {
            if (cx < 0 || cx >= gx) continue;
            if (inner_yz && std::abs(cx - c0x) < r) continue;  // only the ring shell
            const size_t c = cx + (size_t)gx * (cy + (size_t)gy * cz);
            for (uint32_t k = cell_start[c]; k < cell_start[c + 1]; ++k) {
              const uint32_t i = cell_items[k];
              const int64_t dx = wx * (sx[i] - px), dy = wy * (sy[i] - py), dz = wz * (sz[i] - pz);
              const int64_t d = dx * dx + dy * dy + dz * dz;
              if (d < best || (d == best && i < bi)) { best = d; bi = i; }
            }
          }
        }
      }
      if (bi != UINT32_MAX) {
        // lower bound on the metric distance to any cell of ring r+1
        int64_t lb = INT64_MAX;
        lb = std::min(lb, mx - (int64_t)(c0x - r) * cell);
        lb = std::min(lb, (int64_t)(c0x + r + 1) * cell - mx);
        lb = std::min(lb, my - (int64_t)(c0y - r) * cell);
        lb = std::min(lb, (int64_t)(c0y + r + 1) * cell - my);
        lb = std::min(lb, mz - (int64_t)(c0z - r) * cell);
        lb = std::min(lb, (int64_t)(c0z + r + 1) * cell - mz);
        if (lb > 0 && (__int128)lb * lb > best) break;
      }
    }
    return bi;
  }
  static int64_t floordiv(int64_t a, int64_t b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }
};

}  // namespace

struct fcs_gen {
  fcs_gen_params p;
  Noise noise;
  // FCS_GEN_VORONOI
  Voronoi vor;
  int64_t blob_c[2][3] = {{0}};
  int64_t blob_r2 = -1;  // squared radius in (1/256 voxel)^2, <0 = no blobs
  // FCS_GEN_REGIONS
  uint32_t regions_per_axis = 0;
  std::vector<uint8_t> region_type;            // 0 bg, 1 sparse, 2 dense
  std::vector<std::unique_ptr<Voronoi>> region_field;

  uint32_t voronoi_label(const Voronoi& v, const Noise* nz, int64_t gx, int64_t gy, int64_t gz,
                         int64_t ox, int64_t oy, int64_t oz, bool blobs) const {
    int64_t px = (gx - ox) * FP + FP / 2, py = (gy - oy) * FP + FP / 2, pz = (gz - oz) * FP + FP / 2;
    if (nz) {
      px += nz->axis(ST_NOISE_X, gx, gy, gz);
      py += nz->axis(ST_NOISE_Y, gx, gy, gz);
      pz += nz->axis(ST_NOISE_Z, gx, gy, gz);
    }
    if (blobs) {
      for (int b = 0; b < 2; ++b) {
        const int64_t dx = px - blob_c[b][0], dy = py - blob_c[b][1], dz = pz - blob_c[b][2];
        if (dx * dx + dy * dy + dz * dz <= blob_r2) return 0;
      }
    }
    return v.label[v.nearest(px, py, pz)];
  }

  uint32_t at(uint32_t x, uint32_t y, uint32_t z) const {
    if (p.kind == FCS_GEN_VORONOI)
      return voronoi_label(vor, p.noise_amp ? &noise : nullptr, x, y, z, 0, 0, 0, blob_r2 >= 0);
    const uint32_t R = 256;
    const uint32_t rx = x / R, ry = y / R, rz = z / R;
    const size_t ri = rx + (size_t)regions_per_axis * (ry + (size_t)regions_per_axis * rz);
    if (region_type[ri] == 0) return 0;
    const Voronoi& v = *region_field[ri];
    return voronoi_label(v, region_type[ri] == 2 ? &noise : nullptr, x, y, z, (int64_t)rx * R,
                         (int64_t)ry * R, (int64_t)rz * R, false);
  }
};

namespace {

// Draw `count` seeds uniformly inside an `ext`^3 region, with labels drawn
// from the shared label stream (distinct, nonzero, via `used`).
void make_seeds(Voronoi& v, uint64_t seed, uint64_t seed_stream_base, uint32_t count, uint32_t ext,
                uint64_t& label_counter, std::unordered_set<uint32_t>& used) {
  v.ext = ext;
  v.sx.resize(count);
  v.sy.resize(count);
  v.sz.resize(count);
  v.label.resize(count);
  const uint64_t span = (uint64_t)ext * FP;
  for (uint32_t i = 0; i < count; ++i) {
    v.sx[i] = (int64_t)rng_below(seed, ST_SEED_X + seed_stream_base, i, span);
    v.sy[i] = (int64_t)rng_below(seed, ST_SEED_Y + seed_stream_base, i, span);
    v.sz[i] = (int64_t)rng_below(seed, ST_SEED_Z + seed_stream_base, i, span);
    uint32_t lab;
    do {
      lab = (uint32_t)(rng(seed, ST_LABEL, label_counter++) & 0xffffffffu);
    } while (lab == 0 || !used.insert(lab).second);
    v.label[i] = lab;
  }
  v.build_grid();
}

}  // namespace

extern "C" int fcs_gen_preset(int config, uint32_t size, fcs_gen_params* out) {
  if (!out) return FCS_E_INVALID;
  fcs_gen_params p;
  memset(&p, 0, sizeof p);
  p.wx = p.wy = p.wz = 1;
  switch (config) {
    case 1:  // 128^3, seed 0, 600 seeds, seed-slab background 30 %
      p.kind = FCS_GEN_VORONOI; p.size = 128; p.seed = 0; p.num_seeds = 600;
      p.background = FCS_BG_SEED_SLAB;
      break;
    case 2:  // 512^3, seed 1, 20k seeds, metric 1:1:2.5, +-2 voxel noise, two blobs 25 %
    case 4:  // 1024^3, seed 3, same recipe at the same seed density
    case 5:  // serving volumes, same recipe
      p.kind = FCS_GEN_VORONOI;
      p.size = config == 2 ? 512 : (config == 4 ? 1024 : (size ? size : 256));
      p.seed = config == 2 ? 1 : (config == 4 ? 3 : 5 + fcs_log2u(size ? size : 256));
      p.wx = 2; p.wy = 2; p.wz = 5;
      p.noise_amp = 2;
      p.background = FCS_BG_TWO_BLOBS;
      break;
    case 3:  // 2048^3, seed 2, 256^3 regions: background / sparse / dense
      p.kind = FCS_GEN_REGIONS; p.size = 2048; p.seed = 2; p.noise_amp = 2;
      break;
    default:
      return FCS_E_INVALID;
  }
  if (size && config != 5) p.size = size;
  if (p.kind == FCS_GEN_VORONOI && (config == 2 || config == 4 || config == 5)) {
    const double scale = (double)p.size / 512.0;
    p.num_seeds = (uint32_t)std::llround(20000.0 * scale * scale * scale);
    if (p.num_seeds < 1) p.num_seeds = 1;
  }
  *out = p;
  return 0;
}

extern "C" int fcs_gen_create(const fcs_gen_params* p, fcs_gen** out) {
  if (!p || !out || p->size == 0 || p->size > 65536) return FCS_E_INVALID;
  auto g = std::make_unique<fcs_gen>();
  g->p = *p;
  g->noise.seed = p->seed;
  g->noise.amp = (int)p->noise_amp;
  uint64_t label_counter = 0;
  std::unordered_set<uint32_t> used;
  if (p->kind == FCS_GEN_VORONOI) {
    if (p->num_seeds == 0) return FCS_E_INVALID;
    g->vor.wx = p->wx ? p->wx : 1;
    g->vor.wy = p->wy ? p->wy : 1;
    g->vor.wz = p->wz ? p->wz : 1;
    make_seeds(g->vor, p->seed, 0, p->num_seeds, p->size, label_counter, used);
    if (p->background == FCS_BG_SEED_SLAB) {
      const uint32_t axis = (uint32_t)rng_below(p->seed, ST_SLAB, 0, 3);
      const int64_t thick = (int64_t)std::llround(0.3 * p->size);
      const int64_t start = (int64_t)rng_below(p->seed, ST_SLAB, 1, p->size - thick + 1);
      for (uint32_t i = 0; i < p->num_seeds; ++i) {
        const int64_t c = (axis == 0 ? g->vor.sx[i] : axis == 1 ? g->vor.sy[i] : g->vor.sz[i]) / FP;
        if (c >= start && c < start + thick) g->vor.label[i] = 0;
      }
    } else if (p->background == FCS_BG_TWO_BLOBS) {
      // Centres uniform in the inner box; radius bisected so that the union
      // covers 25 % of a 64^3 sample lattice (noise-free positions).
      const double n = p->size;
      const double r0 = std::cbrt(0.125 * 3.0 / (4.0 * M_PI)) * n;
      for (int b = 0; b < 2; ++b)
        for (int a = 0; a < 3; ++a)
          g->blob_c[b][a] =
              (int64_t)((r0 * 0.5 + (n - r0) * (double)rng_below(p->seed, ST_BLOB, b * 3 + a, 1u << 20) /
                                    (double)(1u << 20)) * FP);
      double lo = 0.0, hi = n * FP;
      for (int it = 0; it < 40; ++it) {
        const double mid = 0.5 * (lo + hi);
        const double r2 = mid * mid;
        size_t inside = 0;
        for (int z = 0; z < 64; ++z)
          for (int y = 0; y < 64; ++y)
            for (int x = 0; x < 64; ++x) {
              const double px = (x + 0.5) * n / 64 * FP, py = (y + 0.5) * n / 64 * FP,
                           pz = (z + 0.5) * n / 64 * FP;
              for (int b = 0; b < 2; ++b) {
                const double dx = px - g->blob_c[b][0], dy = py - g->blob_c[b][1], dz = pz - g->blob_c[b][2];
                if (dx * dx + dy * dy + dz * dz <= r2) { ++inside; break; }
              }
            }
        if (inside < (size_t)(0.25 * 64 * 64 * 64)) lo = mid; else hi = mid;
      }
      const int64_t r = (int64_t)hi;
      g->blob_r2 = r * r;
    }
  } else if (p->kind == FCS_GEN_REGIONS) {
    if (p->size % 256) return FCS_E_INVALID;
    g->regions_per_axis = p->size / 256;
    const size_t nreg = (size_t)g->regions_per_axis * g->regions_per_axis * g->regions_per_axis;
    g->region_type.resize(nreg);
    g->region_field.resize(nreg);
    for (size_t r = 0; r < nreg; ++r) {
      g->region_type[r] = (uint8_t)rng_below(p->seed, ST_REGION, r, 3);
      if (g->region_type[r] == 0) continue;
      auto v = std::make_unique<Voronoi>();
      const uint32_t count = g->region_type[r] == 1 ? 50 : 4000;
      make_seeds(*v, p->seed, (ST_REGION_SEEDS + r) * 16, count, 256, label_counter, used);
      g->region_field[r] = std::move(v);
    }
  } else {
    return FCS_E_INVALID;
  }
  *out = g.release();
  return 0;
}

extern "C" void fcs_gen_destroy(fcs_gen* g) { delete g; }

static int resolve_threads(int t) {
  if (t > 0) return t;
  unsigned h = std::thread::hardware_concurrency();
  return h ? (int)h : 4;
}

template <class F>
static void parallel_for(size_t n, int threads, F&& f) {
  threads = std::max(1, std::min<int>(threads, (int)std::max<size_t>(1, n)));
  if (threads == 1) {
    for (size_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<size_t> next{0};
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&] {
      for (size_t i; (i = next.fetch_add(1)) < n;) f(i);
    });
  for (auto& th : pool) th.join();
}

extern "C" int fcs_gen_box(const fcs_gen* g, uint32_t x0, uint32_t y0, uint32_t z0, uint32_t nx,
                           uint32_t ny, uint32_t nz, uint32_t* out, int threads) {
  if (!g || !out) return FCS_E_INVALID;
  if ((uint64_t)x0 + nx > g->p.size || (uint64_t)y0 + ny > g->p.size || (uint64_t)z0 + nz > g->p.size)
    return FCS_E_INVALID;
  parallel_for((size_t)nz * ny, resolve_threads(threads), [&](size_t row) {
    const uint32_t y = (uint32_t)(row % ny), z = (uint32_t)(row / ny);
    uint32_t* o = out + (size_t)nx * row;
    for (uint32_t x = 0; x < nx; ++x) o[x] = g->at(x0 + x, y0 + y, z0 + z);
  });
  return 0;
}

extern "C" int fcs_gen_bricks(const fcs_gen* g, uint32_t b, uint32_t first, uint32_t count, uint32_t* out,
                              int threads) {
  if (!g || !out) return FCS_E_INVALID;
  const uint32_t L = fcs_log2u(b);
  if (b < 2 || (1u << L) != b || L > 6 || g->p.size % b) return FCS_E_INVALID;
  const uint32_t B = g->p.size / b;
  if (B > 1024 || (uint64_t)first + count > (uint64_t)B * B * B) return FCS_E_INVALID;
  const uint64_t nvox = 1ull << (3 * L);
  // one task per (brick, z-slice of the brick): Morton order within a brick
  // is filled through the per-axis dilated coordinates
  parallel_for((size_t)count * b, resolve_threads(threads), [&](size_t task) {
    const uint32_t k = (uint32_t)(task / b), dz = (uint32_t)(task % b);
    const uint32_t id = first + k;
    const uint32_t bx = fcs_unmorton_x(id), by = fcs_unmorton_y(id), bz = fcs_unmorton_z(id);
    uint32_t* o = out + nvox * k;
    const uint32_t mz = fcs_part1by2(dz) << 2;
    for (uint32_t dy = 0; dy < b; ++dy) {
      const uint32_t myz = mz | (fcs_part1by2(dy) << 1);
      for (uint32_t dx = 0; dx < b; ++dx) o[myz | fcs_part1by2(dx)] = g->at(bx * b + dx, by * b + dy, bz * b + dz);
    }
  });
  return 0;
}

extern "C" double fcs_gen_background_fraction(const fcs_gen* g) {
  if (!g) return -1.0;
  size_t zero = 0;
  const uint32_t n = g->p.size;
  for (int z = 0; z < 64; ++z)
    for (int y = 0; y < 64; ++y)
      for (int x = 0; x < 64; ++x)
        zero += g->at((uint32_t)((x * 2 + 1) * (uint64_t)n / 128), (uint32_t)((y * 2 + 1) * (uint64_t)n / 128),
                      (uint32_t)((z * 2 + 1) * (uint64_t)n / 128)) == 0;
  return (double)zero / (64.0 * 64.0 * 64.0);
}

// ================================================================ encoder ==
namespace {

// Majority of 8 labels. Ties break the way the reference does: the first
// maximum in std::unordered_map<uint64_t,int> iteration order
// (brick.hpp:24-37); the same container is used on ties to reproduce it.
inline uint64_t majority8(const uint64_t* v) {
  uint64_t lab[8];
  int cnt[8];
  int nd = 0;
  for (int i = 0; i < 8; ++i) {
    int j = 0;
    while (j < nd && lab[j] != v[i]) ++j;
    if (j == nd) { lab[nd] = v[i]; cnt[nd++] = 1; } else ++cnt[j];
  }
  if (nd == 1) return lab[0];
  int best = 0, arg = 0, nbest = 0;
  for (int j = 0; j < nd; ++j) {
    if (cnt[j] > best) { best = cnt[j]; arg = j; nbest = 1; } else if (cnt[j] == best) ++nbest;
  }
  if (nbest == 1) return lab[arg];
  std::unordered_map<uint64_t, int> freq;
  for (int i = 0; i < 8; ++i) freq[v[i]]++;
  return std::max_element(freq.begin(), freq.end(),
                          [](const auto& a, const auto& b) { return a.second < b.second; })
      ->first;
}

struct Pyramid {
  uint32_t L = 0;
  std::vector<uint64_t> lab;  // coarse-first pyramid layout
  std::vector<uint8_t> cst;
  uint64_t off(uint32_t t) const { return fcs_pyramid_level_offset(t); }
};

void build_pyramid(const uint64_t* finest, uint32_t L, Pyramid& P) {
  P.L = L;
  const uint64_t total = fcs_pyramid_level_offset(L + 1);
  P.lab.resize(total);
  P.cst.resize(total);
  const uint64_t nL = 1ull << (3 * L);
  std::memcpy(&P.lab[P.off(L)], finest, nL * sizeof(uint64_t));
  std::memset(&P.cst[P.off(L)], 1, nL);
  for (int t = (int)L - 1; t >= 0; --t) {
    const uint64_t n = 1ull << (3 * t);
    const uint64_t* cl = &P.lab[P.off(t + 1)];
    const uint8_t* cc = &P.cst[P.off(t + 1)];
    uint64_t* pl = &P.lab[P.off(t)];
    uint8_t* pc = &P.cst[P.off(t)];
    for (uint64_t m = 0; m < n; ++m) {
      const uint64_t maj = majority8(cl + 8 * m);
      bool same = true;
      for (int c = 0; c < 8; ++c) same &= (cl[8 * m + c] == maj) && cc[8 * m + c];
      pl[m] = maj;
      pc[m] = same ? 1 : 0;
    }
  }
}

struct EncodedBrick {
  uint32_t id = 0;
  std::vector<uint64_t> palette;
  std::vector<uint8_t> packed;  // nibble bytes, later replaced by rANS bytes
  uint32_t nsym = 0;
};

// Op selection and packing (brickEncoding.hpp:188-447, brickEncoding.cpp:11-91).
void encode_ops(const Pyramid& P, EncodedBrick& out) {
  const uint32_t L = P.L;
  std::vector<uint8_t> nib;
  nib.reserve(1024);
  std::unordered_map<uint64_t, uint32_t> last;  // label -> last palette index
  out.palette.clear();
  const uint64_t root = P.lab[0];
  out.palette.push_back(root);
  last[root] = 0;
  nib.push_back((uint8_t)((FCS_OP_PA << 1) | (P.cst[0] ? 1 : 0)));
  for (uint32_t t = 0; t < L; ++t) {
    const uint64_t n = 1ull << (3 * t);
    const uint64_t* pl = &P.lab[P.off(t)];
    const uint8_t* pc = &P.cst[P.off(t)];
    const uint64_t* cl = &P.lab[P.off(t + 1)];
    const uint8_t* cc = &P.cst[P.off(t + 1)];
    const int64_t side = 1ll << (t + 1);
    for (uint64_t m = 0; m < n; ++m) {
      if (pc[m]) continue;
      for (uint32_t c = 0; c < 8; ++c) {
        const uint64_t ci = 8 * m + c;
        const uint64_t v = cl[ci];
        int op = -1, delta = 0;
        if (pl[m] == v) op = FCS_OP_RP;
        if (op < 0) {
          const int64_t x = fcs_unmorton_x((uint32_t)ci), y = fcs_unmorton_y((uint32_t)ci),
                        z = fcs_unmorton_z((uint32_t)ci);
          for (int axis = 0; axis < 3 && op < 0; ++axis) {
            int64_t nx = x, ny = y, nz = z;
            int64_t& a = axis == 0 ? nx : (axis == 1 ? ny : nz);
            a = (a & 1) ? a + 1 : a - 1;
            if (nx < 0 || ny < 0 || nz < 0 || nx >= side || ny >= side || nz >= side) continue;
            const uint32_t nb = fcs_morton3((uint32_t)nx, (uint32_t)ny, (uint32_t)nz);
            const uint64_t ref = nb < ci ? cl[nb] : pl[nb >> 3];
            if (ref == v) op = FCS_OP_RX + axis;
          }
        }
        if (op < 0) {
          auto it = last.find(v);
          const uint32_t lastIdx = (uint32_t)out.palette.size() - 1;
          if (it != last.end() && it->second == lastIdx) {
            op = FCS_OP_P0;
          } else if (it != last.end() && lastIdx - it->second - 1 <= 15) {
            op = FCS_OP_PD;
            delta = (int)(lastIdx - it->second - 1);
          } else {
            op = FCS_OP_PA;
            out.palette.push_back(v);
            last[v] = (uint32_t)out.palette.size() - 1;
          }
        }
        nib.push_back((uint8_t)((op << 1) | (cc[ci] ? 1 : 0)));
        if (op == FCS_OP_PD) nib.push_back((uint8_t)delta);
      }
    }
  }
  out.nsym = (uint32_t)nib.size();
  out.packed.assign((nib.size() + 1) / 2, 0);
  for (size_t i = 0; i < nib.size(); ++i)
    out.packed[i >> 1] |= (uint8_t)((i & 1) ? nib[i] : (nib[i] << 4));
}

struct Model {
  uint32_t freq[16];
  uint32_t cum[17];
};

// buildRansModel (brickEncoding.cpp:146-200).
Model build_model(const uint64_t raw[16]) {
  Model m;
  uint64_t rawSum = 0;
  for (int i = 0; i < 16; ++i) {
    uint32_t f = (uint32_t)raw[i];
    if (f == 0) f = 1;
    m.freq[i] = f;
    rawSum += f;
  }
  uint64_t norm = 0;
  for (int i = 0; i < 16; ++i) {
    uint64_t s = ((uint64_t)m.freq[i] * FCS_RANS_M) / rawSum;
    if (s == 0) s = 1;
    m.freq[i] = (uint32_t)s;
    norm += s;
  }
  while (norm < FCS_RANS_M) { m.freq[0]++; norm++; }
  while (norm > FCS_RANS_M) {
    int largest = 0;
    for (int i = 1; i < 16; ++i)
      if (m.freq[i] > m.freq[largest]) largest = i;
    if (m.freq[largest] > 1) { m.freq[largest]--; norm--; } else break;
  }
  uint32_t acc = 0;
  for (int i = 0; i < 16; ++i) { m.cum[i] = acc; acc += m.freq[i]; }
  m.cum[16] = acc;
  return m;
}

// Reverse-order byte-wise rANS over every nibble of the packed bytes
// (pad included), then the 32-bit state LSB first (brickEncoding.hpp:490-526).
void rans_encode(const Model& M, EncodedBrick& b) {
  std::vector<uint8_t> out;
  out.reserve(b.packed.size() + 8);
  uint32_t x = FCS_RANS_L;
  for (size_t i = b.packed.size() * 2; i-- > 0;) {
    const uint32_t s = (i & 1) ? (b.packed[i >> 1] & 0xF) : (b.packed[i >> 1] >> 4);
    const uint32_t f = M.freq[s], c = M.cum[s];
    while (x >= ((FCS_RANS_L >> FCS_RANS_LOGM) << 8) * f) {
      out.push_back((uint8_t)(x & 0xFF));
      x >>= 8;
    }
    x = (x / f) * FCS_RANS_M + (x % f) + c;
  }
  for (int k = 0; k < 4; ++k) { out.push_back((uint8_t)(x & 0xFF)); x >>= 8; }
  b.packed.swap(out);
}

int encode_common(uint32_t BX, uint32_t BY, uint32_t BZ, uint32_t b, const fcs_encode_opts* o,
                  const std::function<void(uint32_t, uint32_t, uint32_t, uint64_t*)>& src,
                  uint8_t** out_bytes, size_t* out_len) {
  const uint32_t L = fcs_log2u(b);
  if ((1u << L) != b || L < 1 || L > 6) return FCS_E_INVALID;
  if (BX > 1024 || BY > 1024 || BZ > 1024) return FCS_E_INVALID;
  const size_t nb = (size_t)BX * BY * BZ;
  const int threads = resolve_threads(o->threads);
  const uint64_t nvox = 1ull << (3 * L);

  auto make_brick = [&](size_t i, EncodedBrick& eb) {
    const uint32_t bz = (uint32_t)(i % BZ), by = (uint32_t)((i / BZ) % BY), bx = (uint32_t)(i / ((size_t)BZ * BY));
    std::vector<uint64_t> fin(nvox);
    src(bx, by, bz, fin.data());
    Pyramid P;
    build_pyramid(fin.data(), L, P);
    encode_ops(P, eb);
    eb.id = fcs_morton3(bx, by, bz);
  };

  // Pass 1: the model samples bricks whose split-order index is a multiple of
  // 512 (brickEncoding.hpp:482), counting both nibbles of every byte.
  std::vector<size_t> sampled;
  for (size_t i = 0; i < nb; i += 512) sampled.push_back(i);
  std::vector<EncodedBrick> bricks(nb);
  parallel_for(sampled.size(), threads, [&](size_t k) { make_brick(sampled[k], bricks[sampled[k]]); });
  uint64_t raw[16] = {0};
  for (size_t i : sampled)
    for (uint8_t byte : bricks[i].packed) { raw[byte >> 4]++; raw[byte & 0xF]++; }
  const Model M = build_model(raw);

  // Pass 2: every brick, then rANS.
  std::atomic<size_t> finished{0};
  parallel_for(nb, threads, [&](size_t i) {
    if (i % 512 != 0) make_brick(i, bricks[i]);
    rans_encode(M, bricks[i]);
    if (o->progress) {
      const size_t k = finished.fetch_add(1) + 1;
      if (k * 16 / nb != (k - 1) * 16 / nb) fprintf(stderr, "[encode] %zu of %zu bricks\n", k, nb);
    }
  });

  size_t total = 16 + 8 + 64 + 64;
  for (auto& e : bricks) total += 4 + 4 + 8 * e.palette.size() + 4 + e.packed.size() + 4;
  uint8_t* buf = (uint8_t*)malloc(total);
  if (!buf) return FCS_E_NOMEM;
  size_t p = 0;
  auto w32 = [&](uint32_t v) { memcpy(buf + p, &v, 4); p += 4; };
  w32(FCS_MAGIC);
  w32(FCS_VERSION);
  w32((uint32_t)nb);
  w32(o->header_brick_size_64 ? 64u : b);
  w32(16);
  w32(FCS_RANS_M);
  for (int i = 0; i < 16; ++i) w32(M.freq[i]);
  for (int i = 0; i < 16; ++i) w32(M.cum[i]);
  for (auto& e : bricks) {
    w32(e.id);
    w32((uint32_t)e.palette.size());
    memcpy(buf + p, e.palette.data(), 8 * e.palette.size());
    p += 8 * e.palette.size();
    w32((uint32_t)e.packed.size());
    memcpy(buf + p, e.packed.data(), e.packed.size());
    p += e.packed.size();
    w32(e.nsym);
  }
  *out_bytes = buf;
  *out_len = p;
  return 0;
}

}  // namespace

extern "C" int fcs_encode_volume(const void* labels, uint32_t label_bits, uint32_t W, uint32_t H,
                                 uint32_t D, const fcs_encode_opts* opts, uint8_t** out_bytes,
                                 size_t* out_len) {
  if (!labels || !opts || !out_bytes || !out_len || (label_bits != 32 && label_bits != 64))
    return FCS_E_INVALID;
  const uint32_t b = opts->brick_size;
  if (b < 2 || W == 0 || H == 0 || D == 0) return FCS_E_INVALID;
  const uint32_t BX = (W + b - 1) / b, BY = (H + b - 1) / b, BZ = (D + b - 1) / b;
  auto src = [&](uint32_t bx, uint32_t by, uint32_t bz, uint64_t* o) {
    for (uint32_t dz = 0; dz < b; ++dz)
      for (uint32_t dy = 0; dy < b; ++dy)
        for (uint32_t dx = 0; dx < b; ++dx) {
          const uint64_t gx = (uint64_t)bx * b + dx, gy = (uint64_t)by * b + dy, gz = (uint64_t)bz * b + dz;
          uint64_t v = 0;
          if (gx < W && gy < H && gz < D) {
            const uint64_t idx = gx + (uint64_t)W * (gy + (uint64_t)H * gz);
            v = label_bits == 64 ? ((const uint64_t*)labels)[idx] : ((const uint32_t*)labels)[idx];
          }
          o[fcs_morton3(dx, dy, dz)] = v;
        }
  };
  return encode_common(BX, BY, BZ, b, opts, src, out_bytes, out_len);
}

extern "C" int fcs_encode_generated(const fcs_gen* g, const fcs_encode_opts* opts, uint8_t** out_bytes,
                                    size_t* out_len) {
  if (!g || !opts || !out_bytes || !out_len) return FCS_E_INVALID;
  const uint32_t b = opts->brick_size, n = g->p.size;
  if (b < 2 || n % b) return FCS_E_INVALID;
  const uint32_t B = n / b;
  auto src = [&](uint32_t bx, uint32_t by, uint32_t bz, uint64_t* o) {
    for (uint32_t dz = 0; dz < b; ++dz)
      for (uint32_t dy = 0; dy < b; ++dy)
        for (uint32_t dx = 0; dx < b; ++dx)
          o[fcs_morton3(dx, dy, dz)] = g->at(bx * b + dx, by * b + dy, bz * b + dz);
  };
  return encode_common(B, B, B, b, opts, src, out_bytes, out_len);
}

extern "C" void fcs_free_buffer(void* p) { free(p); }

extern "C" int fcs_brick_pyramid(const uint64_t* brick_labels, uint32_t brick_size, uint64_t* out,
                                 uint8_t* const_out) {
  if (!brick_labels || !out) return FCS_E_INVALID;
  const uint32_t L = fcs_log2u(brick_size);
  if ((1u << L) != brick_size || L > 6) return FCS_E_INVALID;
  Pyramid P;
  build_pyramid(brick_labels, L, P);
  memcpy(out, P.lab.data(), P.lab.size() * sizeof(uint64_t));
  if (const_out) memcpy(const_out, P.cst.data(), P.cst.size());
  return 0;
}
