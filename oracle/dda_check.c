/* dda_check.c — float64 exact-traversal check of first-hit frames.
 * TEST INFRASTRUCTURE ONLY (see dda_check.h). Independent of oracle.c: no
 * code or constants are shared with the FP32 restatement beyond the camera
 * layout and the scene description (brick grid, LODs, Morton-ordered labels).
 */
#include "dda_check.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

static uint32_t spread3(uint32_t n) {
  n &= 0x3ffu;
  n = (n | (n << 16)) & 0x30000ffu;
  n = (n | (n << 8)) & 0x300f00fu;
  n = (n | (n << 4)) & 0x30c30c3u;
  n = (n | (n << 2)) & 0x9249249u;
  return n;
}
static uint32_t mort(uint32_t x, uint32_t y, uint32_t z) { return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2); }

typedef struct {
  uint32_t B, b, G;
  const uint8_t* empty;
  const uint8_t* lod;
  const uint64_t* const* labels;
  const uint64_t* pal0;
  double rate; /* FP32 rounding allowance per unit of t (see tau) */
} scene_t;

typedef struct {
  uint64_t label, vox;
  double t;
} cellhit;

#define MAX_ALLOW 96
typedef struct {
  int any;       /* the ray only grazes the volume: every answer is a grazing tie */
  int has_exact; /* first non-zero cell in visit order */
  cellhit exact;
  int has_m;     /* first non-zero cell the FP32 walk cannot skip */
  cellhit m;
  int nallow;    /* non-zero cells it may visit instead, before m */
  cellhit allow[MAX_ALLOW];
  int overflow;
  double omax;
} rayres;

/* Allowance for the FP32 marcher at ray parameter t: its crossing times
 * carry the rounding of the ray setup (a few ulp of the direction, scaled by
 * t) and of up to 3b tDelta additions per brick, each half an ulp of a value
 * below t + |origin|; it also offsets brick entries and exits by its epsilon
 * max(t 1e-6, 1e-5). */
static double tau(const scene_t* S, const rayres* R, double t) {
  const double eps = fmax(t * 1e-6, 1e-5);
  return S->rate * (fabs(t) + R->omax) + 2.0 * eps;
}

static uint64_t label_of(const scene_t* S, uint32_t bi, uint32_t t, uint32_t cx, uint32_t cy, uint32_t cz) {
  if (S->empty[bi]) return 0;
  if (t == 0) return S->pal0[bi];
  if (!S->labels[bi]) return 0; /* not resident: the marcher passes through */
  return S->labels[bi][mort(cx, cy, cz)];
}

/* Cell of the global grid containing the point (x, y, z). */
static int cell_at(const scene_t* S, double x, double y, double z, uint64_t* label, uint64_t* vox) {
  const double G = (double)S->G;
  if (!(x >= 0.0 && y >= 0.0 && z >= 0.0 && x < G && y < G && z < G)) return 0;
  const uint32_t ix = (uint32_t)floor(x), iy = (uint32_t)floor(y), iz = (uint32_t)floor(z);
  const uint32_t b = S->b;
  const uint32_t bx = ix / b, by = iy / b, bz = iz / b;
  const uint32_t bi = mort(bx, by, bz);
  const uint32_t t = S->lod[bi];
  const uint32_t s = b >> t;
  const uint32_t cx = (ix - bx * b) / s, cy = (iy - by * b) / s, cz = (iz - bz * b) / s;
  *label = label_of(S, bi, t, cx, cy, cz);
  const uint64_t Gi = S->G;
  *vox = (uint64_t)(bx * b + cx * s) + Gi * ((uint64_t)(by * b + cy * s) + Gi * (uint64_t)(bz * b + cz * s));
  return 1;
}

static void allow(rayres* R, uint64_t label, uint64_t vox, double t) {
  if (label == 0) return;
  for (int i = 0; i < R->nallow; ++i)
    if (R->allow[i].vox == vox && R->allow[i].label == label) return;
  if (R->nallow == MAX_ALLOW) {
    R->overflow = 1;
    return;
  }
  R->allow[R->nallow].label = label;
  R->allow[R->nallow].vox = vox;
  R->allow[R->nallow].t = t;
  ++R->nallow;
}

/* One cell of the walk, entered at tin and left at tout. `rule` marks a
 * zero-length cell entered through an exact tie inside a brick, which the
 * tie rule makes the marcher sample. Returns 1 when the walk can stop. */
static int visit(const scene_t* S, rayres* R, const double o[3], const double d[3], uint64_t label, uint64_t vox,
                 double tin, double tout, int rule) {
  const double len = tout - tin;
  const double tu = tau(S, R, tin);
  if (label != 0 && !R->has_exact) {
    R->has_exact = 1;
    R->exact.label = label;
    R->exact.vox = vox;
    R->exact.t = tin;
  }
  if (len <= tu && !rule) {
    /* a grazing stay: the FP32 walk may skip this cell and pass through
     * any cell around it instead (the other crossing order) */
    allow(R, label, vox, tin);
    const double tm = 0.5 * (tin + tout), r = 0.5 * len + tu;
    const double off[3] = {-r, 0.0, r};
    for (int i = 0; i < 27; ++i) {
      uint64_t l2, v2;
      const double px = o[0] + d[0] * tm + off[i % 3], py = o[1] + d[1] * tm + off[(i / 3) % 3],
                   pz = o[2] + d[2] * tm + off[i / 9];
      if (cell_at(S, px, py, pz, &l2, &v2)) allow(R, l2, v2, tm);
    }
    return 0;
  }
  if (label != 0) {
    R->has_m = 1;
    R->m.label = label;
    R->m.vox = vox;
    R->m.t = tin;
    return 1;
  }
  return 0;
}

/* Index of the cell of edge `size` a ray moving with direction component d
 * enters at coordinate p (on a boundary it goes to the side d points to). */
static int enter_index(double p, double d, double size) {
  if (d < 0.0) return (int)ceil(p / size) - 1;
  return (int)floor(p / size);
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

static double min3(const double v[3]) { return fmin(v[0], fmin(v[1], v[2])); }

/* Among the axes whose crossing time equals tmin, the one the marcher steps
 * first on a tie (advanceDDA: x only when strictly smallest, then y only when
 * strictly below z): z, then y, then x. */
static int tie_axis(const double tc[3], double tmin) {
  if (tc[2] == tmin) return 2;
  if (tc[1] == tmin) return 1;
  return 0;
}

static void trace(const scene_t* S, const double o[3], const double d[3], rayres* R) {
  const double G = (double)S->G;
  double tmin = -INFINITY, tmax = INFINITY;
  R->omax = fmax(fabs(o[0]), fmax(fabs(o[1]), fabs(o[2])));
  for (int a = 0; a < 3; ++a) {
    if (d[a] == 0.0) {
      if (o[a] < 0.0 || o[a] > G) return;
      continue;
    }
    const double t0 = (0.0 - o[a]) / d[a], t1 = (G - o[a]) / d[a];
    tmin = fmax(tmin, fmin(t0, t1));
    tmax = fmin(tmax, fmax(t0, t1));
  }
  if (tmax < 0.0 || tmin > tmax) {
    if (tmin - tmax <= tau(S, R, tmin) && tmax >= -tau(S, R, 0.0)) R->any = 1;
    return;
  }
  const double tEnter = fmax(tmin, 0.0), tExit = tmax;
  if (tExit - tEnter <= tau(S, R, tEnter)) {
    R->any = 1; /* through an edge or corner region of the volume */
    return;
  }
  const int B = (int)S->B, b = (int)S->b;
  int bx[3];
  for (int a = 0; a < 3; ++a) bx[a] = clampi(enter_index(o[a] + d[a] * tEnter, d[a], b), 0, B - 1);
  double t = tEnter;
  for (;;) {
    double tb[3];
    for (int a = 0; a < 3; ++a) {
      if (d[a] > 0.0) tb[a] = ((double)(bx[a] + 1) * b - o[a]) / d[a];
      else if (d[a] < 0.0) tb[a] = ((double)bx[a] * b - o[a]) / d[a];
      else tb[a] = INFINITY;
    }
    double tbo = min3(tb);
    int last = 0;
    if (tbo >= tExit) {
      tbo = tExit;
      last = 1;
    }
    const uint32_t bi = mort((uint32_t)bx[0], (uint32_t)bx[1], (uint32_t)bx[2]);
    const uint32_t lt = S->lod[bi];
    const int walk = !S->empty[bi] && (lt == 0 || S->labels[bi] != NULL);
    if (walk) {
      const int s = b >> lt, n = 1 << lt;
      const double bmin[3] = {(double)bx[0] * b, (double)bx[1] * b, (double)bx[2] * b};
      int c[3];
      for (int a = 0; a < 3; ++a) c[a] = clampi(enter_index(o[a] + d[a] * t - bmin[a], d[a], s), 0, n - 1);
      int rule = 0;
      for (;;) {
        double tc[3];
        for (int a = 0; a < 3; ++a) {
          if (d[a] > 0.0) tc[a] = (bmin[a] + (double)(c[a] + 1) * s - o[a]) / d[a];
          else if (d[a] < 0.0) tc[a] = (bmin[a] + (double)c[a] * s - o[a]) / d[a];
          else tc[a] = INFINITY;
        }
        double tco = min3(tc);
        const int leaving = tco >= tbo;
        if (leaving) tco = tbo;
        const uint64_t label = label_of(S, bi, lt, (uint32_t)c[0], (uint32_t)c[1], (uint32_t)c[2]);
        const uint64_t Gi = S->G;
        const uint64_t vox = (uint64_t)(bmin[0] + c[0] * s) +
                             Gi * ((uint64_t)(bmin[1] + c[1] * s) + Gi * (uint64_t)(bmin[2] + c[2] * s));
        if (visit(S, R, o, d, label, vox, t, tco, rule)) return;
        if (leaving) break;
        const int a = tie_axis(tc, tco);
        const int tied = (tc[0] == tco) + (tc[1] == tco) + (tc[2] == tco);
        c[a] += d[a] > 0.0 ? 1 : -1;
        t = tco;
        rule = tied > 1;
      }
    }
    if (last) return;
    const int a = tie_axis(tb, tbo);
    bx[a] += d[a] > 0.0 ? 1 : -1;
    if (bx[a] < 0 || bx[a] >= B) return;
    t = tbo;
  }
}

typedef struct {
  const scene_t* S;
  const dc_camera* cam;
  uint32_t w, h;
  const uint64_t *got_label, *got_voxel;
  const float* got_t;
  uint64_t *exp_label, *exp_voxel;
  uint8_t* cls;
  uint32_t next_row;
  pthread_mutex_t mu;
  dc_stats st;
} job_t;

static void ray_of(const dc_camera* c, uint32_t w, uint32_t h, uint32_t px, uint32_t py, double o[3], double d[3]) {
  const double nx = ((double)px + 0.5) / (double)w * 2.0 - 1.0;
  const double ny = 1.0 - ((double)py + 0.5) / (double)h * 2.0;
  const double aspect = (double)w / (double)h;
  if (c->ortho_half_height > 0.0f) {
    const double hh = c->ortho_half_height;
    for (int a = 0; a < 3; ++a) {
      o[a] = (double)c->position[a] + (double)c->right[a] * nx * aspect * hh + (double)c->up[a] * ny * hh;
      d[a] = c->forward[a];
    }
    return;
  }
  const double th = c->tan_half_fov;
  double v[3], n2 = 0.0;
  for (int a = 0; a < 3; ++a) {
    v[a] = (double)c->forward[a] + (double)c->right[a] * nx * aspect * th + (double)c->up[a] * ny * th;
    n2 += v[a] * v[a];
    o[a] = c->position[a];
  }
  const double len = sqrt(n2);
  for (int a = 0; a < 3; ++a) d[a] = v[a] / len;
}

static void* worker(void* arg) {
  job_t* J = (job_t*)arg;
  dc_stats st;
  memset(&st, 0, sizeof st);
  rayres* R = (rayres*)malloc(sizeof(rayres));
  for (;;) {
    pthread_mutex_lock(&J->mu);
    const uint32_t py = J->next_row++;
    pthread_mutex_unlock(&J->mu);
    if (py >= J->h) break;
    for (uint32_t px = 0; px < J->w; ++px) {
      const size_t k = (size_t)py * J->w + px;
      double o[3], d[3];
      ray_of(J->cam, J->w, J->h, px, py, o, d);
      memset(R, 0, sizeof *R);
      trace(J->S, o, d, R);
      const uint64_t gl = J->got_label[k], gv = J->got_voxel[k];
      const int ghit = gl != 0;
      uint8_t c;
      const int same = R->has_exact ? (ghit && gl == R->exact.label && gv == R->exact.vox) : !ghit;
      if (same) {
        c = DC_AGREE;
        if (ghit) {
          const double err = fabs((double)J->got_t[k] - R->exact.t);
          if (err > tau(J->S, R, R->exact.t)) c = DC_UNEXPLAINED;
          else if (err > st.max_t_err) st.max_t_err = err;
        }
      } else if (R->any) {
        c = DC_GRAZING;
      } else if (!ghit) {
        c = R->has_m ? DC_UNEXPLAINED : DC_GRAZING;
      } else {
        int ok = R->has_m && gl == R->m.label && gv == R->m.vox;
        for (int i = 0; i < R->nallow && !ok; ++i) ok = gl == R->allow[i].label && gv == R->allow[i].vox;
        c = ok ? DC_GRAZING : DC_UNEXPLAINED;
      }
      if (J->cls) J->cls[k] = c;
      if (J->exp_label) J->exp_label[k] = R->has_exact ? R->exact.label : 0;
      if (J->exp_voxel) J->exp_voxel[k] = R->has_exact ? R->exact.vox : UINT64_MAX;
      st.pixels++;
      st.hits += R->has_exact;
      st.agree += c == DC_AGREE;
      st.grazing += c == DC_GRAZING;
      st.unexplained += c == DC_UNEXPLAINED;
    }
  }
  free(R);
  pthread_mutex_lock(&J->mu);
  J->st.pixels += st.pixels;
  J->st.hits += st.hits;
  J->st.agree += st.agree;
  J->st.grazing += st.grazing;
  J->st.unexplained += st.unexplained;
  if (st.max_t_err > J->st.max_t_err) J->st.max_t_err = st.max_t_err;
  pthread_mutex_unlock(&J->mu);
  return NULL;
}

int dc_check_frame(uint32_t B, uint32_t b, const uint8_t* brick_empty, const uint8_t* brick_lod,
                   const uint64_t* const* brick_labels, const uint64_t* palette0, const dc_camera* cam, uint32_t w,
                   uint32_t h, const uint64_t* got_label, const uint64_t* got_voxel, const float* got_t,
                   int threads, uint64_t* exp_label, uint64_t* exp_voxel, uint8_t* cls, dc_stats* stats) {
  scene_t S;
  S.B = B;
  S.b = b;
  S.G = B * b;
  S.empty = brick_empty;
  S.lod = brick_lod;
  S.labels = brick_labels;
  S.pal0 = palette0;
  /* 2^-24 per rounding, 3b + 16 roundings (tDelta sums within a brick, ray
   * setup, the brick-level sums) */
  S.rate = ldexp(1.0, -24) * (3.0 * b + 16.0);
  job_t J;
  memset(&J, 0, sizeof J);
  J.S = &S;
  J.cam = cam;
  J.w = w;
  J.h = h;
  J.got_label = got_label;
  J.got_voxel = got_voxel;
  J.got_t = got_t;
  J.exp_label = exp_label;
  J.exp_voxel = exp_voxel;
  J.cls = cls;
  pthread_mutex_init(&J.mu, NULL);
  if (threads < 1) threads = 1;
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * threads);
  for (int i = 0; i < threads; ++i) pthread_create(&th[i], NULL, worker, &J);
  for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
  free(th);
  pthread_mutex_destroy(&J.mu);
  if (stats) *stats = J.st;
  return 0;
}
