// Views (asynchronous frames with device-side setup, optionally replayed as
// hipGraphs), cache admission under an HBM budget and the open-loop serving
// loop. Declarations and contracts: include/fcs/fcs_serve.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <numeric>
#include <thread>
#include <vector>

#include "fcs/fcs.h"
#include "fcs/fcs_serve.h"
#include "fcs_common.h"
#include "fcs_internal.h"
#include "fcs_kernels.h"

namespace {

struct GraphEntry {
  uint64_t serial;
  const void* cache;
  uint64_t shade_epoch;
  uint32_t w, h;
  hipGraph_t graph;
  hipGraphExec_t exec;
};

}  // namespace

struct fcs_view {
  fcs_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_done = nullptr;
  uint32_t max_w = 0, max_h = 0, flags = 0;
  size_t max_px = 0;
  // device buffers
  uint32_t* d_meta = nullptr;
  uint64_t* d_offs = nullptr;
  uint8_t* d_lod = nullptr;
  uint32_t nb_cap = 0;
  void* d_label = nullptr;    // max_px x 8 bytes (either label width)
  uint64_t* d_voxel = nullptr;
  float* d_t = nullptr;
  uint32_t* d_rgba = nullptr;  // FCS_VIEW_RGBA: max_px RGBA8
  MarchCam* d_cam = nullptr;
  uint64_t* d_small = nullptr;  // [0] misses (u32), [1] checksum
  // pinned host buffers
  MarchCam* h_cam = nullptr;
  void* h_label = nullptr;
  uint64_t* h_small = nullptr;
  uint8_t* h_lod = nullptr;
  std::vector<GraphEntry> graphs;
  // the frame in flight / last finished
  bool pending = false, issued = false;
  fcs_view_result last{};
  uint32_t cur_nb = 0;
};

namespace {

int vfail(fcs_view* v, int code, const char* msg) { return fcs_fail(v ? v->ctx : nullptr, code, msg); }

#define VIEW_TRY(view, expr) FCS_HIP_TRY((view)->ctx, expr)

void drop_graphs(fcs_view* v) {
  for (auto& g : v->graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
  }
  v->graphs.clear();
}

int ensure_bricks(fcs_view* v, uint32_t nb) {
  if (nb <= v->nb_cap) return 0;
  // graphs hold the old buffer addresses
  VIEW_TRY(v, hipStreamSynchronize(v->stream));
  drop_graphs(v);
  if (v->d_meta) (void)hipFree(v->d_meta);
  if (v->d_offs) (void)hipFree(v->d_offs);
  if (v->d_lod) (void)hipFree(v->d_lod);
  if (v->h_lod) (void)hipHostFree(v->h_lod);
  v->d_meta = nullptr;
  v->d_offs = nullptr;
  v->d_lod = nullptr;
  v->h_lod = nullptr;
  v->nb_cap = 0;
  VIEW_TRY(v, hipMalloc((void**)&v->d_meta, (size_t)nb * 4));
  VIEW_TRY(v, hipMalloc((void**)&v->d_offs, (size_t)nb * 8));
  VIEW_TRY(v, hipMalloc((void**)&v->d_lod, (size_t)nb));
  VIEW_TRY(v, hipHostMalloc((void**)&v->h_lod, (size_t)nb, hipHostMallocDefault));
  v->nb_cap = nb;
  return 0;
}

// Shaded frames: the RGBA output, the volume's transfer function and lighting.
template <typename LT>
void shade(const fcs_view* v, const fcs_volume_desc& d, MarchArgs<LT>& a) {
  if (!(v->flags & FCS_VIEW_RGBA)) return;
  a.out_rgba = v->d_rgba;
  a.tf_keys = d.tf_keys;
  a.tf_rgba = d.tf_rgba;
  a.tf_mask = d.tf_mask;
  for (int k = 0; k < 3; ++k) a.light[k] = d.light[k];
  a.ambient = d.ambient;
  a.diffuse = d.diffuse;
}

// The frame's device work, in stream order: camera (and fixed LODs) in,
// per-brick setup, counters cleared by a kernel, march, checksum, outputs out.
// Identical whether issued directly or recorded into a graph.
int enqueue_frame(fcs_view* v, const fcs_volume_desc& d, bool fixed_lod, uint32_t w, uint32_t h) {
  hipStream_t s = v->stream;
  const size_t npx = (size_t)w * h;
  const size_t lb = d.label_bits / 8;
  VIEW_TRY(v, hipMemcpyAsync(v->d_cam, v->h_cam, sizeof(MarchCam), hipMemcpyHostToDevice, s));
  if (fixed_lod) VIEW_TRY(v, hipMemcpyAsync(v->d_lod, v->h_lod, d.nb, hipMemcpyHostToDevice, s));
  FrameSetupArgs fa{d.res, v->d_cam, fixed_lod ? v->d_lod : nullptr, v->d_meta, v->d_offs, d.nb, d.b, d.L, w, h};
  VIEW_TRY(v, fcs_launch_frame_setup(fa, s));
  VIEW_TRY(v, fcs_launch_zero(reinterpret_cast<uint32_t*>(v->d_small), 4, s));
  MarchCam zc;
  memset(&zc, 0, sizeof zc);
  const float aspect = (float)((double)w / (double)h);
  const float grid = (float)(d.B * d.b);
  if (d.label_bits == 32) {
    MarchArgs<uint32_t> a{zc, v->d_cam, v->d_meta, v->d_offs, (const uint32_t*)d.pal0, (const uint32_t*)d.cache,
                          (uint32_t*)v->d_label, v->d_voxel, v->d_t, nullptr,
                          reinterpret_cast<uint32_t*>(v->d_small), w, h, (float)w, (float)h, aspect, grid, d.b, d.B};
    shade(v, d, a);
    VIEW_TRY(v, fcs_launch_march<uint32_t>(a, s));
    if (v->flags & FCS_VIEW_CHECKSUM)
      VIEW_TRY(v, fcs_launch_frame_checksum<uint32_t>((const uint32_t*)v->d_label, v->d_voxel, npx,
                                                      (unsigned long long*)(v->d_small + 1), s));
  } else {
    MarchArgs<uint64_t> a{zc, v->d_cam, v->d_meta, v->d_offs, (const uint64_t*)d.pal0, (const uint64_t*)d.cache,
                          (uint64_t*)v->d_label, v->d_voxel, v->d_t, nullptr,
                          reinterpret_cast<uint32_t*>(v->d_small), w, h, (float)w, (float)h, aspect, grid, d.b, d.B};
    shade(v, d, a);
    VIEW_TRY(v, fcs_launch_march<uint64_t>(a, s));
    if (v->flags & FCS_VIEW_CHECKSUM)
      VIEW_TRY(v, fcs_launch_frame_checksum<uint64_t>((const uint64_t*)v->d_label, v->d_voxel, npx,
                                                      (unsigned long long*)(v->d_small + 1), s));
  }
  if (v->flags & FCS_VIEW_LABELS) {
    if (v->flags & FCS_VIEW_RGBA)
      VIEW_TRY(v, hipMemcpyAsync(v->h_label, v->d_rgba, npx * 4, hipMemcpyDeviceToHost, s));
    else
      VIEW_TRY(v, hipMemcpyAsync(v->h_label, v->d_label, npx * lb, hipMemcpyDeviceToHost, s));
  }
  VIEW_TRY(v, hipMemcpyAsync(v->h_small, v->d_small, 16, hipMemcpyDeviceToHost, s));
  return 0;
}

GraphEntry* find_graph(fcs_view* v, const fcs_volume_desc& d, uint32_t w, uint32_t h) {
  for (auto& g : v->graphs)
    if (g.serial == d.serial && g.cache == d.cache && g.shade_epoch == d.shade_epoch && g.w == w && g.h == h)
      return &g;
  return nullptr;
}

// Most graphs a view keeps; the oldest is destroyed to make room.
constexpr size_t kMaxViewGraphs = 64;

void destroy_graph(GraphEntry& g) {
  if (g.exec) (void)hipGraphExecDestroy(g.exec);
  if (g.graph) (void)hipGraphDestroy(g.graph);
  g.exec = nullptr;
  g.graph = nullptr;
}

// Called with the view's stream idle (its last frame waited for): graphs of
// the same volume and frame size whose cache or shading epoch is stale can
// never be replayed again and are destroyed; beyond kMaxViewGraphs the oldest
// goes too.
void evict_graphs(fcs_view* v, const fcs_volume_desc& d, uint32_t w, uint32_t h) {
  auto stale = [&](const GraphEntry& g) { return g.serial == d.serial && g.w == w && g.h == h; };
  for (auto& g : v->graphs)
    if (stale(g)) destroy_graph(g);
  v->graphs.erase(std::remove_if(v->graphs.begin(), v->graphs.end(), [](const GraphEntry& g) { return !g.exec; }),
                  v->graphs.end());
  while (v->graphs.size() >= kMaxViewGraphs) {
    destroy_graph(v->graphs.front());
    v->graphs.erase(v->graphs.begin());
  }
}

int capture_graph(fcs_view* v, const fcs_volume_desc& d, uint32_t w, uint32_t h, GraphEntry** out) {
  VIEW_TRY(v, hipStreamSynchronize(v->stream));
  evict_graphs(v, d, w, h);
  VIEW_TRY(v, hipStreamBeginCapture(v->stream, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue_frame(v, d, false, w, h);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(v->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  VIEW_TRY(v, e);
  hipGraphExec_t x = nullptr;
  const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    (void)hipGraphDestroy(g);
    VIEW_TRY(v, ei);
  }
  v->graphs.push_back(GraphEntry{d.serial, d.cache, d.shade_epoch, w, h, g, x});
  *out = &v->graphs.back();
  return 0;
}

double now_s(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

extern "C" {

int fcs_view_create(fcs_ctx* ctx, uint32_t max_w, uint32_t max_h, uint32_t flags, fcs_view** out) {
  if (!ctx || !out || max_w == 0 || max_h == 0 || max_w > 16384 || max_h > 16384 ||
      (flags & ~(FCS_VIEW_GRAPH | FCS_VIEW_LABELS | FCS_VIEW_CHECKSUM | FCS_VIEW_RGBA)))
    return FCS_E_INVALID;
  *out = nullptr;
  FCS_HIP_TRY(ctx, hipSetDevice(ctx->device));
  fcs_view* v = new fcs_view();
  v->ctx = ctx;
  v->max_w = max_w;
  v->max_h = max_h;
  v->flags = flags;
  v->max_px = (size_t)max_w * max_h;
  auto bad = [&](hipError_t e) {
    if (e == hipSuccess) return false;
    fcs_view_destroy(v);
    return true;
  };
  if (bad(hipStreamCreateWithFlags(&v->stream, hipStreamNonBlocking)) || bad(hipEventCreate(&v->ev_start)) ||
      bad(hipEventCreate(&v->ev_done)) || bad(hipMalloc(&v->d_label, v->max_px * 8)) ||
      bad(hipMalloc((void**)&v->d_voxel, v->max_px * 8)) || bad(hipMalloc((void**)&v->d_t, v->max_px * 4)) ||
      bad(hipMalloc((void**)&v->d_cam, sizeof(MarchCam))) || bad(hipMalloc((void**)&v->d_small, 64)) ||
      bad(hipHostMalloc((void**)&v->h_cam, sizeof(MarchCam), hipHostMallocDefault)) ||
      bad(hipHostMalloc((void**)&v->h_small, 64, hipHostMallocDefault)) ||
      ((flags & FCS_VIEW_RGBA) && bad(hipMalloc((void**)&v->d_rgba, v->max_px * 4))) ||
      ((flags & FCS_VIEW_LABELS) && bad(hipHostMalloc(&v->h_label, v->max_px * 8, hipHostMallocDefault))))
    return fcs_fail(ctx, FCS_E_NOMEM, "view allocation failed");
  *out = v;
  return 0;
}

void fcs_view_destroy(fcs_view* v) {
  if (!v) return;
  if (v->ctx) (void)hipSetDevice(v->ctx->device);
  if (v->stream) (void)hipStreamSynchronize(v->stream);
  drop_graphs(v);
  void* dev[] = {v->d_meta, v->d_offs, v->d_lod, v->d_label, v->d_voxel, v->d_t, v->d_cam, v->d_small, v->d_rgba};
  for (void* p : dev)
    if (p) (void)hipFree(p);
  void* host[] = {v->h_cam, v->h_label, v->h_small, v->h_lod};
  for (void* p : host)
    if (p) (void)hipHostFree(p);
  if (v->ev_start) (void)hipEventDestroy(v->ev_start);
  if (v->ev_done) (void)hipEventDestroy(v->ev_done);
  if (v->stream) (void)hipStreamDestroy(v->stream);
  delete v;
}

int fcs_view_render_async(fcs_view* v, fcs_volume* vol, const fcs_camera* cam, const uint8_t* brick_lod, uint32_t w,
                          uint32_t h) {
  if (!v || !vol || !cam || w == 0 || h == 0) return FCS_E_INVALID;
  if (fcs_volume_context(vol) != v->ctx) return vfail(v, FCS_E_INVALID, "volume and view belong to different contexts");
  if (w > v->max_w || h > v->max_h) return vfail(v, FCS_E_INVALID, "frame larger than the view");
  VIEW_TRY(v, hipSetDevice(v->ctx->device));
  if (v->pending) {
    fcs_view_result r;
    if (int rc = fcs_view_wait(v, &r)) return rc;
  }
  fcs_volume_desc d;
  if (int rc = fcs_volume_publish(vol, &d, (v->flags & FCS_VIEW_RGBA) != 0)) return rc;
  if (int rc = ensure_bricks(v, d.nb)) return rc;
  static_assert(sizeof(MarchCam) == sizeof(fcs_camera), "camera layouts");
  memcpy(v->h_cam, cam, sizeof(MarchCam));
  const bool fixed = brick_lod != nullptr;
  if (fixed) memcpy(v->h_lod, brick_lod, d.nb);
  if ((v->flags & FCS_VIEW_GRAPH) && !fixed) {
    GraphEntry* g = find_graph(v, d, w, h);
    if (!g)
      if (int rc = capture_graph(v, d, w, h, &g)) return rc;
    VIEW_TRY(v, hipEventRecord(v->ev_start, v->stream));
    VIEW_TRY(v, hipGraphLaunch(g->exec, v->stream));
  } else {
    VIEW_TRY(v, hipEventRecord(v->ev_start, v->stream));
    if (int rc = enqueue_frame(v, d, fixed, w, h)) return rc;
  }
  VIEW_TRY(v, hipEventRecord(v->ev_done, v->stream));
  v->pending = true;
  v->issued = true;
  v->cur_nb = d.nb;
  v->last = fcs_view_result{};
  v->last.w = w;
  v->last.h = h;
  v->last.label_bits = d.label_bits;
  return 0;
}

int fcs_view_query(fcs_view* v) {
  if (!v) return FCS_E_INVALID;
  if (!v->pending) return 1;
  const hipError_t e = hipEventQuery(v->ev_done);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  VIEW_TRY(v, e);
  return 0;
}

int fcs_view_wait(fcs_view* v, fcs_view_result* out) {
  if (!v || !out) return FCS_E_INVALID;
  if (!v->issued) return vfail(v, FCS_E_INVALID, "no frame issued on this view");
  if (v->pending) {
    VIEW_TRY(v, hipEventSynchronize(v->ev_done));
    float ms = 0;
    VIEW_TRY(v, hipEventElapsedTime(&ms, v->ev_start, v->ev_done));
    v->last.device_ms = ms;
    v->last.misses = (uint32_t)(v->h_small[0] & 0xFFFFFFFFu);
    v->last.checksum = (v->flags & FCS_VIEW_CHECKSUM) ? v->h_small[1] : 0;
    v->pending = false;
  }
  *out = v->last;
  return 0;
}

const void* fcs_view_labels(const fcs_view* v) {
  if (!v || !(v->flags & FCS_VIEW_LABELS) || v->pending || !v->issued) return nullptr;
  return v->h_label;
}

int fcs_view_read(fcs_view* v, void* label, uint64_t* voxel, float* t_hit, uint8_t* brick_lod_used) {
  if (!v) return FCS_E_INVALID;
  fcs_view_result r;
  if (int rc = fcs_view_wait(v, &r)) return rc;
  const size_t npx = (size_t)r.w * r.h;
  if (label) VIEW_TRY(v, hipMemcpyAsync(label, v->d_label, npx * (r.label_bits / 8), hipMemcpyDeviceToHost, v->stream));
  if (voxel) VIEW_TRY(v, hipMemcpyAsync(voxel, v->d_voxel, npx * 8, hipMemcpyDeviceToHost, v->stream));
  if (t_hit) VIEW_TRY(v, hipMemcpyAsync(t_hit, v->d_t, npx * 4, hipMemcpyDeviceToHost, v->stream));
  std::vector<uint32_t> meta;
  if (brick_lod_used) {
    meta.resize(v->cur_nb);
    VIEW_TRY(v, hipMemcpyAsync(meta.data(), v->d_meta, (size_t)v->cur_nb * 4, hipMemcpyDeviceToHost, v->stream));
  }
  VIEW_TRY(v, hipStreamSynchronize(v->stream));
  for (size_t i = 0; i < meta.size(); ++i) brick_lod_used[i] = (uint8_t)(meta[i] & FCS_META_LOD_MASK);
  return 0;
}

int fcs_view_read_rgba(fcs_view* v, uint32_t* rgba) {
  if (!v || !rgba) return FCS_E_INVALID;
  if (!(v->flags & FCS_VIEW_RGBA)) return vfail(v, FCS_E_INVALID, "not a shaded view");
  fcs_view_result r;
  if (int rc = fcs_view_wait(v, &r)) return rc;
  VIEW_TRY(v, hipMemcpyAsync(rgba, v->d_rgba, (size_t)r.w * r.h * 4, hipMemcpyDeviceToHost, v->stream));
  VIEW_TRY(v, hipStreamSynchronize(v->stream));
  return 0;
}

uint32_t fcs_view_graphs(const fcs_view* v) { return v ? (uint32_t)v->graphs.size() : 0; }

int fcs_serve_plan(const uint64_t* full_bytes, const double* demand, uint32_t n, uint64_t budget, uint32_t* plan,
                   uint64_t* pool_bytes) {
  if (n && (!full_bytes || !demand || !plan || !pool_bytes)) return FCS_E_INVALID;
  for (uint32_t i = 0; i < n; ++i)
    if (!(demand[i] >= 0.0)) return FCS_E_INVALID;
  std::vector<uint32_t> order(n);
  std::iota(order.begin(), order.end(), 0u);
  // demand per full byte, descending; compared as cross products (no division)
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return demand[a] * (double)full_bytes[b] > demand[b] * (double)full_bytes[a];
  });
  uint64_t left = budget;
  for (uint32_t i = 0; i < n; ++i) {
    plan[i] = 0;
    pool_bytes[i] = 0;
  }
  for (uint32_t i : order) {
    if (demand[i] > 0.0 && full_bytes[i] <= left) {
      plan[i] = 2;
      pool_bytes[i] = full_bytes[i];
      left -= full_bytes[i];
    }
  }
  double rest = 0.0;
  for (uint32_t i = 0; i < n; ++i)
    if (plan[i] == 0 && demand[i] > 0.0) rest += demand[i];
  if (rest > 0.0) {
    for (uint32_t i = 0; i < n; ++i) {
      if (plan[i] != 0 || demand[i] <= 0.0) continue;
      const uint64_t share = (uint64_t)((double)left * (demand[i] / rest));
      if (share > 0) {
        plan[i] = 1;
        pool_bytes[i] = share;
      }
    }
  }
  return 0;
}

int fcs_serve_admit(fcs_volume* const* vols, const double* demand, uint32_t n, uint64_t budget, uint32_t* plan_out,
                    uint64_t* pool_out) {
  if (n && (!vols || !demand)) return FCS_E_INVALID;
  std::vector<uint64_t> full(n), pool(n);
  std::vector<uint32_t> plan(n);
  std::vector<fcs_volume_info> info(n);
  for (uint32_t i = 0; i < n; ++i) {
    if (int rc = fcs_volume_get_info(vols[i], &info[i])) return rc;
    full[i] = (uint64_t)info[i].num_bricks * fcs_pyramid_level_offset(info[i].levels + 1) * (info[i].label_bits / 8);
  }
  if (int rc = fcs_serve_plan(full.data(), demand, n, budget, plan.data(), pool.data())) return rc;
  for (uint32_t i = 0; i < n; ++i) {
    if (plan[i] == 0) continue;
    if (!info[i].rans_ready)
      if (int rc = fcs_decode_rans(vols[i], nullptr); rc < 0) return rc;
    if (plan[i] == 2) {
      if (int rc = fcs_decode_all(vols[i], info[i].levels, 1)) return rc;
    } else {
      if (int rc = fcs_cache_reserve(vols[i], pool[i] / (info[i].label_bits / 8))) return rc;
    }
  }
  if (plan_out) std::copy(plan.begin(), plan.end(), plan_out);
  if (pool_out) std::copy(pool.begin(), pool.end(), pool_out);
  return 0;
}

int fcs_serve_run(fcs_view* const* views, uint32_t n_views, fcs_volume* const* vols, uint32_t n_vols,
                  const fcs_serve_request* reqs, uint32_t n, fcs_serve_result* out) {
  if (!views || n_views == 0 || !vols || n_vols == 0 || (n && (!reqs || !out))) return FCS_E_INVALID;
  fcs_ctx* c = fcs_volume_context(vols[0]);
  for (uint32_t k = 0; k < n_views; ++k)
    if (!views[k] || views[k]->ctx != c) return fcs_fail(c, FCS_E_INVALID, "views and volumes must share a context");
  for (uint32_t i = 0; i < n; ++i) {
    if (reqs[i].volume >= n_vols || !vols[reqs[i].volume]) return fcs_fail(c, FCS_E_INVALID, "bad volume index");
    if (i && reqs[i].arrival_s < reqs[i - 1].arrival_s) return fcs_fail(c, FCS_E_INVALID, "requests not sorted");
    if (reqs[i].w == 0 || reqs[i].h == 0) return fcs_fail(c, FCS_E_INVALID, "empty frame");
  }
  // every view's previous frame ends before the clock starts
  for (uint32_t k = 0; k < n_views; ++k) {
    fcs_view_result r;
    if (views[k]->pending)
      if (int rc = fcs_view_wait(views[k], &r)) return rc;
  }
  std::vector<int64_t> cur(n_views, -1);
  std::deque<uint32_t> queue;
  std::vector<uint32_t> redo;  // requests whose view frame missed
  uint32_t next = 0, done = 0, in_flight = 0;
  const auto t0 = std::chrono::steady_clock::now();

  auto complete = [&](uint32_t k) -> int {
    fcs_view_result r;
    if (int rc = fcs_view_wait(views[k], &r)) return rc;
    const uint32_t q = (uint32_t)cur[k];
    out[q].done_s = now_s(t0);
    out[q].checksum = r.checksum;
    out[q].device_ms = r.device_ms;
    out[q].misses = r.misses;
    out[q].view = k;
    cur[k] = -1;
    --in_flight;
    if (r.misses) redo.push_back(q);
    else ++done;
    return 0;
  };

  while (done < n) {
    const double now = now_s(t0);
    while (next < n && reqs[next].arrival_s <= now) queue.push_back(next++);
    for (uint32_t k = 0; k < n_views; ++k) {
      if (cur[k] < 0) continue;
      const int q = fcs_view_query(views[k]);
      if (q < 0) return q;
      if (q == 1)
        if (int rc = complete(k)) return rc;
    }
    if (!redo.empty()) {
      // a frame missed: drain the views, then decode what is missing and
      // render the request on the context's stream
      for (uint32_t k = 0; k < n_views; ++k)
        if (cur[k] >= 0)
          if (int rc = complete(k)) return rc;
      for (uint32_t q : redo) {
        const fcs_serve_request& rq = reqs[q];
        fcs_frame f;
        memset(&f, 0, sizeof f);
        if (int rc = fcs_render_visible(vols[rq.volume], &rq.cam, nullptr, rq.w, rq.h, &f)) return rc;
        uint64_t cs = 0;
        if (int rc = fcs_frame_checksum(vols[rq.volume], &cs)) return rc;
        out[q].checksum = cs;
        out[q].done_s = now_s(t0);
        ++done;
      }
      redo.clear();
      continue;
    }
    for (uint32_t k = 0; k < n_views && !queue.empty(); ++k) {
      if (cur[k] >= 0) continue;
      const uint32_t q = queue.front();
      queue.pop_front();
      const fcs_serve_request& rq = reqs[q];
      out[q] = fcs_serve_result{};
      out[q].issue_s = now_s(t0);
      if (int rc = fcs_view_render_async(views[k], vols[rq.volume], &rq.cam, nullptr, rq.w, rq.h)) return rc;
      cur[k] = q;
      ++in_flight;
    }
    if (in_flight == 0 && queue.empty() && next < n) {
      // idle until the next arrival: sleep, then spin the last 200 us
      const double wait = reqs[next].arrival_s - now_s(t0);
      if (wait > 4e-4) std::this_thread::sleep_for(std::chrono::duration<double>(wait - 2e-4));
    }
  }
  return 0;
}

}  // extern "C"
