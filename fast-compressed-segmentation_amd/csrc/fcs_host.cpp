// Host runtime of the C ABI (include/fcs/fcs.h): one HIP stream per context,
// device-resident volumes, a brick cache pool with per-size free lists, and
// batched decode / march launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fcs/fcs.h"
#include "fcs/fcs_mlp.h"
#include "fcs_common.h"
#include "fcs_device.h"
#include "fcs_internal.h"
#include "fcs_kernels.h"


namespace {

struct Alloc {
  uint64_t off = 0, size = 0;
};

// Device request / status buffers of one stream's decode launches, and the
// host copies of its last launch (requests in cost-binned order, the
// permutation back to the caller's order, statuses in device order).
struct DecodeLane {
  DecodeReq* d_req = nullptr;
  uint32_t* d_status = nullptr;
  uint32_t* d_counter = nullptr;
  size_t cap = 0;
  std::vector<DecodeReq> sorted;
  std::vector<uint32_t> perm;
  std::vector<uint32_t> status_sorted;
  std::vector<DecodeReq> reqs;  // caller order
  uint64_t gen = 0;             // set by every stage_decode into this lane
};

// Frames built on the device (direct-mapped pools): the per-brick LOD rule,
// frustum test and decode requests come from kernels, the op decoder drains
// a device-built queue, and the whole frame is replayed as a hipGraph.
struct DevGraph {
  int form;
  uint32_t w, h;
  bool rgba, fixed;
  const void *cache, *flabel, *frgba, *tfk, *tfr;
  uint64_t shade_epoch;
  hipGraph_t graph;
  hipGraphExec_t exec;
};

// Device-frame small words (d_small / h_small): queue count, misses of the
// first and second march, error bits, claim counter; the prefetch's queue
// count, claim counter and error bits.
enum { DF_QCOUNT = 0, DF_MISS1 = 1, DF_MISS2 = 2, DF_ERR = 3, DF_CLAIM = 4, DF_AQCOUNT = 8, DF_ACLAIM = 9,
       DF_AERR = 10, DF_WORDS = 16, DF_HIST = 32 };

struct DevFrames {
  DecodeReq* d_q = nullptr;     // frame queue (capacity nb)
  uint32_t* d_qst = nullptr;    // its status words
  DecodeReq* d_aq = nullptr;    // prefetch queue (capacity nb)
  uint32_t* d_aqst = nullptr;
  uint32_t* d_small = nullptr;  // DF_* words, then the prefetch histogram (1028 words)
  uint32_t* d_needed = nullptr;  // per brick: entered by a ray after its first miss
  uint32_t* d_pixmiss = nullptr; // per pixel: the ray missed (pix_cap pixels)
  size_t pix_cap = 0;
  void* d_cam = nullptr;        // MarchCam: camera, look-ahead camera
  void* d_cam_ahead = nullptr;
  uint8_t* d_lod = nullptr;     // fixed LODs
  void* h_cam = nullptr;        // pinned staging of the same
  void* h_cam_ahead = nullptr;
  uint8_t* h_lod = nullptr;
  uint32_t* h_small = nullptr;  // pinned copy of the DF_* words
  bool aux_pending = false;     // a prefetch decode is in flight on the aux stream
  std::vector<DevGraph> graphs;
};

}  // namespace

struct fcs_volume {
  fcs_ctx* ctx = nullptr;
  uint32_t nb = 0, b = 0, L = 0, B = 0, label_bits = 32, header_b = 0;
  uint64_t compressed = 0, symbols = 0, palette_entries = 0;
  uint32_t freq[16]{}, cum[17]{};
  std::vector<DevBrick> hb;
  std::vector<uint8_t> empty;     // all palette entries are 0
  std::vector<uint64_t> pal0;     // palette[0] per brick
  std::vector<uint32_t> rans_status;
  // device
  uint8_t* d_stream = nullptr;
  DevBrick* d_bricks = nullptr;
  void* d_pal = nullptr;
  void* d_pal0 = nullptr;
  uint32_t* d_nib = nullptr;
  RansResult* d_rres = nullptr;
  uint32_t* d_rtab = nullptr;  // x -> (symbol << 25) | (freq << 12) | (x - cum)
  uint32_t* d_rlist = nullptr; // bricks by nibble count, longest first (the rANS launch order)
  std::vector<uint32_t> rlist; // host copy
  uint32_t* d_counter = nullptr;
  bool rans_ready = false;
  int rans_tuned = 0;  // mapping picked by the autotuner for this volume (0 = not yet)
  int ops_tuned = 0;   // op decoder split picked by the autotuner (FCS_OPS_SPLIT_*, 0 = not yet)
  // brick cache pool (voxels of label_bits/8 bytes)
  void* d_cache = nullptr;
  uint64_t cap = 0, bump = 0;
  std::vector<std::vector<uint64_t>> free_lists;  // by size class
  std::vector<int16_t> res_lod;   // -1 = not resident
  std::vector<uint8_t> res_pyr;   // 1 = pyramid layout (levels 0..res_lod)
  std::vector<Alloc> res_alloc;
  uint64_t res_epoch = 1;     // bumped on every residency change
  uint64_t res_published = 0; // epoch of the device residency table
  uint64_t* d_res = nullptr;  // per brick residency word (FCS_RES_*), for device-side frame setup
  uint64_t serial = 0;
  // shaded frames: every palette entry (host copy), the transfer function
  // (ascending labels, default and current colours, device hash table) and
  // the lighting
  std::vector<uint64_t> pal_host;
  bool tf_built = false, tf_dirty = true;
  std::vector<uint64_t> tf_labels;
  std::vector<uint32_t> tf_base, tf_cur;
  uint64_t* d_tf_keys = nullptr;
  uint32_t* d_tf_rgba = nullptr;
  uint32_t tf_mask = 0;
  float light[3] = {0.0f, 0.0f, 0.0f};
  float ambient = 0.5f, diffuse = 0.8f;
  uint64_t shade_epoch = 0;     // bumped when the lighting or the table's buffers change
  uint32_t* d_frgba = nullptr;  // frame RGBA (frame_cap pixels)
  // direct-mapped cache: brick i owns voxels [i * stride, (i + 1) * stride)
  bool direct = false;
  uint64_t stride = 0;
  // decode launches: `lane_main` on the context stream (synchronous API
  // calls), `lane_aux` on the aux stream (prefetches, finished later)
  DecodeLane lane_main, lane_aux;
  uint64_t stage_gen = 0;  // source of DecodeLane::gen
  // pinned host copies of the per-brick results read back after a full
  // decode (num_bricks entries each; allocated at first use)
  RansResult* h_rres = nullptr;
  uint32_t* h_vstatus = nullptr;
  // fcs_decode_volume's binned requests as staged in lane_main: reused while
  // the lane still holds them (same generation) for the same LOD, layout and
  // cache slot size
  struct {
    uint64_t gen = 0;
    uint32_t t = 0;
    bool pyr = false;
    uint64_t slot = 0;
  } vol_plan;
  std::vector<uint8_t> aux_lod;  // LOD of each pending prefetch request
  uint32_t* d_meta = nullptr;
  uint64_t* d_offs = nullptr;
  void* d_flabel = nullptr;
  uint64_t* d_fvoxel = nullptr;
  float* d_ft = nullptr;
  uint32_t* d_visited = nullptr;
  uint32_t* d_misses = nullptr;
  unsigned long long* d_csum = nullptr;
  unsigned long long* d_mstats = nullptr;  // march statistics slots (3 x FCS_MARCH_STATS_SLOTS)
  size_t frame_cap = 0;
  size_t frame_px = 0;  // pixels of the last rendered frame
  // device frames: d_res is changed by the device and is newer than the
  // host copy (res_lod / res_pyr / res_alloc) until pull_residency
  DevFrames* df = nullptr;
  bool res_dev_newer = false;

  size_t lbytes() const { return label_bits / 8; }
};

namespace {

int fail(fcs_ctx* c, int code, const std::string& msg) { return fcs_fail(c, code, msg); }

#define HIP_TRY(ctx, expr) FCS_HIP_TRY(ctx, expr)

uint32_t rd32(const uint8_t* p) {
  uint32_t v;
  memcpy(&v, p, 4);
  return v;
}

int size_class(uint64_t voxels) {
  // sizes are 8^t or pyramid sizes (8^(t+1)-1)/7: class = (t, pyramid)
  for (int t = 0; t <= 6; ++t) {
    if (voxels == (1ull << (3 * t))) return t;
    if (voxels == fcs_pyramid_level_offset(t + 1)) return 8 + t;
  }
  return -1;
}

bool cache_alloc(fcs_volume* v, uint32_t brick, uint64_t voxels, uint64_t* off) {
  if (v->direct) {
    if (voxels > v->stride) return false;
    *off = v->stride * brick;
    return true;
  }
  const int cls = size_class(voxels);
  if (cls >= 0 && (size_t)cls < v->free_lists.size() && !v->free_lists[cls].empty()) {
    *off = v->free_lists[cls].back();
    v->free_lists[cls].pop_back();
    return true;
  }
  // keep every allocation 16-voxel aligned for vector stores
  const uint64_t start = (v->bump + 15) & ~15ull;
  if (start + voxels > v->cap) return false;
  *off = start;
  v->bump = start + voxels;
  return true;
}

bool resident_at(const fcs_volume* v, uint32_t i, uint32_t t) {
  const int r = v->res_lod[i];
  return r >= 0 && ((v->res_pyr[i] && (uint32_t)r >= t) || (!v->res_pyr[i] && (uint32_t)r == t));
}

// Releases are deferred to the end of a batch: a region dropped while the
// batch is being built may still be the target of an earlier request of the
// same launch, so it must not be handed to another request of that launch.
void cache_release(fcs_volume* v, uint32_t brick, std::vector<Alloc>* deferred) {
  if (v->res_lod[brick] < 0) return;
  const Alloc a = v->res_alloc[brick];
  ++v->res_epoch;
  if (v->direct) {
    v->res_lod[brick] = -1;
    return;
  }
  if (deferred) {
    deferred->push_back(a);
  } else {
    const int cls = size_class(a.size);
    if (cls >= 0) v->free_lists[cls].push_back(a.off);
  }
  v->res_lod[brick] = -1;
}

void cache_free_deferred(fcs_volume* v, const std::vector<Alloc>& deferred) {
  for (const Alloc& a : deferred) {
    const int cls = size_class(a.size);
    if (cls >= 0) v->free_lists[cls].push_back(a.off);
  }
}

int ensure_req(fcs_volume* v, DecodeLane& ln, size_t n) {
  fcs_ctx* c = v->ctx;
  if (!ln.d_counter) HIP_TRY(c, hipMalloc(&ln.d_counter, 64));
  if (n <= ln.cap) return 0;
  if (ln.d_req) { (void)hipFree(ln.d_req); ln.d_req = nullptr; }
  if (ln.d_status) { (void)hipFree(ln.d_status); ln.d_status = nullptr; }
  ln.cap = 0;
  size_t cap = std::max<size_t>(n, 1024);
  HIP_TRY(c, hipMalloc(&ln.d_req, cap * sizeof(DecodeReq)));
  HIP_TRY(c, hipMalloc(&ln.d_status, cap * sizeof(uint32_t)));
  ln.cap = cap;
  return 0;
}

void free_lane(DecodeLane& ln) {
  if (ln.d_req) (void)hipFree(ln.d_req);
  if (ln.d_status) (void)hipFree(ln.d_status);
  if (ln.d_counter) (void)hipFree(ln.d_counter);
  ln = DecodeLane{};
}

int ensure_cache(fcs_volume* v) {
  if (v->d_cache) return 0;
  return fcs_cache_reserve(v, (uint64_t)v->nb * fcs_pyramid_level_offset(v->L + 1));
}

// Launch the op decoder over `reqs`; statuses copied back.
// Launch the op decoder over `reqs` on `stream` through lane `ln`; the
// statuses are copied back asynchronously (read them after the stream is
// synchronised, with finish_decode). With `timed`, the launch is bracketed by
// the context's timing events.
// The op decoder split used for this volume: the option when forced, else
// the autotuner's pick (three workgroups per CU until it has run).
int ops_split_of(const fcs_volume* v) {
  const int forced = v->ctx->ops_split;
  if (forced != FCS_OPS_SPLIT_AUTO) return forced;
  return v->ops_tuned ? v->ops_tuned : FCS_OPS_SPLIT_3WG;
}

// Sort `reqs` into lane `ln` and copy them to the device on `stream`.
int stage_decode(fcs_volume* v, DecodeLane& ln, const std::vector<DecodeReq>& reqs, hipStream_t stream) {
  fcs_ctx* c = v->ctx;
  ln.reqs = reqs;
  ln.gen = ++v->stage_gen;
  const size_t n = reqs.size();
  if (n == 0) return 0;
  if (int rc = ensure_req(v, ln, n)) return rc;
  // Bin the requests by estimated cost, heaviest first (stable within a bin):
  // the persistent decoder deals the front three quarters round-robin and claims the
  // rest, so light requests fill the end. Cost ~ nibbles up to the LOD plus
  // labels written; bins are eighths of an octave, so requests of about the
  // same cost are claimed together (octave bins alone left each octave in
  // input order: the 2048^3 full decode's op kernel ran 9.5 ms in brick order
  // against 9.15 ms sorted by stream length).
  constexpr int kOct = 48, kSub = 8, kBins = kOct * kSub;
  std::vector<uint32_t> bin(n);
  ln.perm.resize(n);
  std::vector<size_t> count(kBins + 1, 0);
  for (size_t i = 0; i < n; ++i) {
    const DecodeReq& q = reqs[i];
    const uint32_t t = q.lod & 0xFF;
    const uint32_t nsym = q.brick < v->nb ? v->hb[q.brick].nsym : 0;
    const uint32_t drop = t < v->L ? 3 * (v->L - t) : 0;
    const uint64_t cost = (uint64_t)(nsym >> (drop < 32 ? drop : 31)) + ((1ull << (3 * (t < 20 ? t : 20))) >> 3) + 1;
    const int lg = 63 - __builtin_clzll(cost);
    const uint32_t sub = (uint32_t)((lg >= 3 ? cost >> (lg - 3) : cost << (3 - lg)) & 7u);
    const uint32_t key = (uint32_t)(lg < kOct ? lg : kOct - 1) * kSub + sub;  // grows with cost
    bin[i] = (uint32_t)(kBins - 1) - key;
    ++count[bin[i] + 1];
  }
  for (int b = 0; b < kBins; ++b) count[b + 1] += count[b];
  for (size_t i = 0; i < n; ++i) ln.perm[count[bin[i]]++] = (uint32_t)i;
  ln.sorted.resize(n);
  for (size_t i = 0; i < n; ++i) ln.sorted[i] = reqs[ln.perm[i]];
  HIP_TRY(c, hipMemcpyAsync(ln.d_req, ln.sorted.data(), n * sizeof(DecodeReq), hipMemcpyHostToDevice, stream));
  return 0;
}

// Launch the op decoder over lane `ln`'s staged requests on `stream`; with
// `copy_status` the statuses come back to the host in the same stream order
// (a copy to pageable memory, which holds the host until the launch is done).
int issue_decode(fcs_volume* v, DecodeLane& ln, hipStream_t stream, bool timed, int split, bool copy_status) {
  fcs_ctx* c = v->ctx;
  const size_t n = ln.reqs.size();
  if (n == 0) return 0;
  if (split == 0) split = ops_split_of(v);
  const int ksplit = split == FCS_OPS_SPLIT_4WG ? FCS_OPS_4WG : FCS_OPS_3WG;
  c->timing.ops_split = (uint32_t)split;
  if (timed) HIP_TRY(c, hipEventRecord(c->ev[0], stream));
  if (v->label_bits == 32) {
    OpsArgs<uint32_t> a{v->d_bricks, v->d_nib, (const uint32_t*)v->d_pal, ln.d_req,
                        (uint32_t*)v->d_cache, ln.d_status, ln.d_counter, (uint32_t)n, v->nb, v->L};
    HIP_TRY(c, fcs_launch_ops<uint32_t>(a, (uint32_t)c->num_cu, stream, nullptr, ksplit, c->ops_head));
  } else {
    OpsArgs<uint64_t> a{v->d_bricks, v->d_nib, (const uint64_t*)v->d_pal, ln.d_req,
                        (uint64_t*)v->d_cache, ln.d_status, ln.d_counter, (uint32_t)n, v->nb, v->L};
    HIP_TRY(c, fcs_launch_ops<uint64_t>(a, (uint32_t)c->num_cu, stream, nullptr, ksplit, c->ops_head));
  }
  if (timed) HIP_TRY(c, hipEventRecord(c->ev[1], stream));
  ln.status_sorted.resize(n);
  if (copy_status)
    HIP_TRY(c, hipMemcpyAsync(ln.status_sorted.data(), ln.d_status, n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                              stream));
  return 0;
}

int launch_decode(fcs_volume* v, DecodeLane& ln, const std::vector<DecodeReq>& reqs, hipStream_t stream,
                  bool timed, int split = 0) {
  if (int rc = stage_decode(v, ln, reqs, stream)) return rc;
  return issue_decode(v, ln, stream, timed, split, true);
}

// Statuses of lane `ln`'s last launch in the caller's order (the launch's
// stream must have been synchronised).
void finish_decode(DecodeLane& ln, std::vector<uint32_t>& status) {
  const size_t n = ln.reqs.size();
  status.assign(n, 0);
  for (size_t i = 0; i < n; ++i) status[ln.perm[i]] = ln.status_sorted[i];
}

// Labels and algorithmic bytes of a decode of `reqs` into the context's
// timing record; returns the labels.
uint64_t account_decode(fcs_volume* v, const std::vector<DecodeReq>& reqs) {
  uint64_t vox = 0, bytes = 0;
  for (const auto& q : reqs) {
    const uint32_t t = q.lod & 0xFF;
    const uint64_t n = (q.lod & FCS_REQ_PYRAMID) ? fcs_pyramid_level_offset(t + 1) : (1ull << (3 * t));
    vox += n;
    const DevBrick& bk = v->hb[q.brick < v->nb ? q.brick : 0];
    // algorithmic bytes: nibble stream consumed (0.5 B/nibble) + palette +
    // 48 B of brick records + labels written
    bytes += (uint64_t)(bk.nsym + 1) / 2 + 8ull * bk.P + 48 + n * v->lbytes();
  }
  v->ctx->timing.decoded_voxels = vox;
  v->ctx->timing.decode_bytes = bytes;
  return vox;
}

// Smallest batch the op decoder's split autotuner measures on.
constexpr size_t kOpsTuneMinRequests = 2048;

// Decode `reqs` on the context stream and wait; statuses in caller order.
int run_decode(fcs_volume* v, DecodeLane& ln, const std::vector<DecodeReq>& reqs, std::vector<uint32_t>& status,
               uint64_t* voxels_out) {
  fcs_ctx* c = v->ctx;
  status.assign(reqs.size(), 0);
  if (reqs.empty()) {
    c->timing.decode_ms = 0.0f;
    return 0;
  }
  float ms = 0;
  if (c->ops_split == FCS_OPS_SPLIT_AUTO && !v->ops_tuned && v->L <= 5 && reqs.size() >= kOpsTuneMinRequests) {
    // Autotune the LDS split on this batch: decoding is idempotent (every
    // request rewrites the same cache slots with the same labels), so the
    // batch runs three times -- a warm-up with three workgroups per CU, then
    // four, then three again, timed -- and the faster split is kept.
    float t[2] = {0, 0};
    const int order[3] = {FCS_OPS_SPLIT_3WG, FCS_OPS_SPLIT_4WG, FCS_OPS_SPLIT_3WG};
    for (int k = 0; k < 3; ++k) {
      if (int rc = launch_decode(v, ln, reqs, c->stream, true, order[k])) return rc;
      HIP_TRY(c, hipStreamSynchronize(c->stream));
      HIP_TRY(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
      if (k) t[order[k] - 1] = ms;
    }
    c->timing.tune_ops_3wg_ms = t[0];
    c->timing.tune_ops_4wg_ms = t[1];
    v->ops_tuned = t[1] < t[0] ? FCS_OPS_SPLIT_4WG : FCS_OPS_SPLIT_3WG;
    ms = t[v->ops_tuned - 1];
    c->timing.ops_split = (uint32_t)v->ops_tuned;
  } else {
    if (int rc = launch_decode(v, ln, reqs, c->stream, true)) return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  }
  finish_decode(ln, status);
  c->timing.decode_ms = ms;
  const uint64_t vox = account_decode(v, reqs);
  if (voxels_out) *voxels_out = vox;
  return 0;
}

// Wait for the prefetch launched on the aux stream (if any) and settle its
// statuses: a failed prefetch leaves its brick non-resident, so the next
// on-demand pass decodes it again and reports the failure.
int settle_prefetch(fcs_volume* v) {
  fcs_ctx* c = v->ctx;
  if (!c->aux_pending) return 0;
  HIP_TRY(c, hipEventSynchronize(c->aux_done));
  c->aux_pending = false;
  std::vector<uint32_t> st;
  finish_decode(v->lane_aux, st);
  for (size_t i = 0; i < st.size(); ++i)
    if (st[i]) cache_release(v, v->lane_aux.reqs[i].brick, nullptr);
  v->lane_aux.reqs.clear();
  return 0;
}

// Wait for a device-built prefetch decode on the aux stream (failed requests
// were made non-resident on the device by the settle kernel).
int settle_dev_prefetch(fcs_volume* v) {
  if (!v->df || !v->df->aux_pending) return 0;
  HIP_TRY(v->ctx, hipEventSynchronize(v->ctx->aux_done));
  v->df->aux_pending = false;
  return 0;
}

// Bring the host copy of the residency (res_lod / res_pyr / res_alloc) up to
// date after device frames changed the device table. Host residency
// operations call this first; afterwards host and device agree.
int pull_residency(fcs_volume* v) {
  if (int rc = settle_dev_prefetch(v)) return rc;
  if (!v->res_dev_newer) return 0;
  fcs_ctx* c = v->ctx;
  std::vector<uint64_t> w(v->nb);
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  HIP_TRY(c, hipMemcpy(w.data(), v->d_res, (size_t)v->nb * 8, hipMemcpyDeviceToHost));
  for (uint32_t i = 0; i < v->nb; ++i) {
    const uint32_t l = (uint32_t)((w[i] >> FCS_RES_LOD_SHIFT) & 0xFF);
    if (l == FCS_RES_NONE) {
      v->res_lod[i] = -1;
      continue;
    }
    v->res_lod[i] = (int16_t)l;
    v->res_pyr[i] = (w[i] & FCS_RES_PYRAMID) ? 1 : 0;
    v->res_alloc[i] = Alloc{w[i] & FCS_RES_OFF_MASK,
                            v->res_pyr[i] ? fcs_pyramid_level_offset(l + 1) : (1ull << (3 * l))};
  }
  v->res_dev_newer = false;
  v->res_published = v->res_epoch;
  return 0;
}

// Upload the host residency to the device table when it changed since the
// last upload. Frames in flight on any stream may read the old table, so the
// device is synchronised first.
int push_residency(fcs_volume* v) {
  fcs_ctx* c = v->ctx;
  if (!v->d_res) {
    HIP_TRY(c, hipMalloc((void**)&v->d_res, (size_t)v->nb * 8));
    v->res_published = 0;
  }
  if (v->res_dev_newer || v->res_published == v->res_epoch) return 0;
  std::vector<uint64_t> words(v->nb);
  for (uint32_t i = 0; i < v->nb; ++i) {
    uint64_t w = FCS_RES_NONE << FCS_RES_LOD_SHIFT;
    if (v->res_lod[i] >= 0) {
      w = (v->res_alloc[i].off & FCS_RES_OFF_MASK) | ((uint64_t)v->res_lod[i] << FCS_RES_LOD_SHIFT);
      if (v->res_pyr[i]) w |= FCS_RES_PYRAMID;
    }
    if (v->empty[i]) w |= FCS_RES_EMPTY;
    words[i] = w;
  }
  HIP_TRY(c, hipDeviceSynchronize());
  HIP_TRY(c, hipMemcpy(v->d_res, words.data(), words.size() * 8, hipMemcpyHostToDevice));
  v->res_published = v->res_epoch;
  return 0;
}

void drop_dev_graphs(fcs_volume* v) {
  if (!v->df) return;
  for (auto& g : v->df->graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
  }
  v->df->graphs.clear();
}

void free_dev_frames(fcs_volume* v) {
  if (!v->df) return;
  drop_dev_graphs(v);
  DevFrames* F = v->df;
  void* dev[] = {F->d_q, F->d_qst, F->d_aq, F->d_aqst, F->d_small, F->d_cam, F->d_cam_ahead, F->d_lod,
                 F->d_needed, F->d_pixmiss};
  for (void* q : dev)
    if (q) (void)hipFree(q);
  void* host[] = {F->h_cam, F->h_cam_ahead, F->h_lod, F->h_small};
  for (void* q : host)
    if (q) (void)hipHostFree(q);
  delete F;
  v->df = nullptr;
}

// Lighting uniform of webgpuSetup.js:205-219: the angle in degrees becomes
// the direction (sin, 0.6, -cos), computed in double like the JavaScript and
// rounded to float like its Float32Array upload.
void set_lighting(fcs_volume* v, float angle_deg, float ambient, float diffuse) {
  const double a = ((double)angle_deg * 3.141592653589793) / 180.0;
  const float l[3] = {(float)std::sin(a), 0.6f, (float)-std::cos(a)};
  // the values are baked into captured frames: a new epoch only when one changes
  if (l[0] == v->light[0] && l[1] == v->light[1] && l[2] == v->light[2] && ambient == v->ambient &&
      diffuse == v->diffuse && v->shade_epoch != 0)
    return;
  for (int k = 0; k < 3; ++k) v->light[k] = l[k];
  v->ambient = ambient;
  v->diffuse = diffuse;
  ++v->shade_epoch;
}

// Math.round for a non-negative double (halves go up)
double js_round(double x) {
  const double f = std::floor(x);
  return (x - f >= 0.5) ? f + 1.0 : f;
}

// Build the transfer function on first use (buildTransferFunctionBuffers,
// webgpuSetup.js:389-458) and upload its hash table when colours changed.
int ensure_tf(fcs_volume* v) {
  fcs_ctx* c = v->ctx;
  if (!v->tf_built) {
    std::vector<uint64_t> labs;
    labs.reserve(v->pal_host.size());
    for (uint64_t x : v->pal_host)
      if (x != 0) labs.push_back(x);
    std::sort(labs.begin(), labs.end());
    labs.erase(std::unique(labs.begin(), labs.end()), labs.end());
    v->tf_labels = std::move(labs);
    v->tf_base.resize(v->tf_labels.size());
    for (size_t i = 0; i < v->tf_labels.size(); ++i) v->tf_base[i] = fcs_tf_default_rgba(i);
    v->tf_cur = v->tf_base;
    v->tf_built = true;
    v->tf_dirty = true;
  }
  if (!v->tf_dirty) return 0;
  const size_t n = v->tf_labels.size();
  uint32_t cap = 16;
  while (cap < 2 * n) cap <<= 1;
  std::vector<uint64_t> keys(cap, 0);
  std::vector<uint32_t> cols(cap, 0);
  const uint32_t mask = cap - 1;
  for (size_t i = 0; i < n; ++i) {
    uint32_t h = fcs_tf_slot(v->tf_labels[i], mask);
    while (keys[h] != 0) h = (h + 1) & mask;
    keys[h] = v->tf_labels[i];
    cols[h] = v->tf_cur[i];
  }
  // Frames of any stream may be reading the table: views march on their own
  // streams and replay graphs that hold its address. Wait for the whole
  // device before the table is freed or overwritten.
  HIP_TRY(c, hipDeviceSynchronize());
  if (v->tf_mask != mask || !v->d_tf_keys) {
    if (v->d_tf_keys) (void)hipFree(v->d_tf_keys);
    if (v->d_tf_rgba) (void)hipFree(v->d_tf_rgba);
    v->d_tf_keys = nullptr;
    v->d_tf_rgba = nullptr;
    HIP_TRY(c, hipMalloc((void**)&v->d_tf_keys, (size_t)cap * 8));
    HIP_TRY(c, hipMalloc((void**)&v->d_tf_rgba, (size_t)cap * 4));
    v->tf_mask = mask;
    ++v->shade_epoch;
  }
  HIP_TRY(c, hipMemcpy(v->d_tf_keys, keys.data(), (size_t)cap * 8, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(v->d_tf_rgba, cols.data(), (size_t)cap * 4, hipMemcpyHostToDevice));
  v->tf_dirty = false;
  return 0;
}

// calculateBrickLOD (camera.js:103-137) for one brick: the largest LOD t
// with focal * (b / 2^t) / d >= 1, d = distance from the camera to the brick
// centre, focal from the camera's tan(fov/2) and screen width min(w, h)
// (webgpuSetup.js:587). float64 like the reference's JS numbers.
struct LodRule {
  const fcs_volume* v;
  double px, py, pz, focal, half;
  double voxel[32];
  LodRule(const fcs_volume* vol, const fcs_camera* cam, uint32_t w, uint32_t h) : v(vol) {
    const double sw = (double)std::min(w, h);
    focal = (sw / 2.0) / (double)cam->tan_half_fov;
    half = v->b / 2.0;
    px = cam->position[0];
    py = cam->position[1];
    pz = cam->position[2];
    for (uint32_t l = 0; l <= v->L; ++l) voxel[l] = (double)v->b / std::pow(2.0, (double)l);  // exact powers of two
  }
  uint8_t operator()(uint32_t i) const {
    const double cx = fcs_unmorton_x(i) * (double)v->b + half, cy = fcs_unmorton_y(i) * (double)v->b + half,
                 cz = fcs_unmorton_z(i) * (double)v->b + half;
    const double dx = cx - px, dy = cy - py, dz = cz - pz;
    double d = std::sqrt(dx * dx + dy * dy + dz * dz);
    if (d < 0.1) d = 0.1;
    uint32_t best = 0;
    for (uint32_t l = 0; l <= v->L; ++l) {
      if ((focal * voxel[l]) / d >= 1.0) best = l; else break;
    }
    return (uint8_t)best;
  }
};

// Bricks a ray of the frame can enter. Perspective: the brick box (grown by
// one voxel against rounding) is tested against the four side planes through
// the camera position and the plane behind it; a box wholly outside one plane
// is culled. Orthographic frames keep every brick.
std::vector<uint8_t> frustum_mask(const fcs_volume* v, const fcs_camera* cam, uint32_t w, uint32_t h) {
  std::vector<uint8_t> vis(v->nb, 1);
  if (cam->ortho_half_height > 0.0f || !(cam->tan_half_fov > 0.0f)) return vis;
  const double a = (double)w / (double)h, T = cam->tan_half_fov;
  double p[3], f[3], r[3], u[3];
  for (int k = 0; k < 3; ++k) {
    p[k] = cam->position[k];
    f[k] = cam->forward[k];
    r[k] = cam->right[k];
    u[k] = cam->up[k];
  }
  // inside: n . (x - p) <= 0 for every plane normal n
  double n[5][3];
  for (int k = 0; k < 3; ++k) {
    n[0][k] = r[k] - a * T * f[k];
    n[1][k] = -r[k] - a * T * f[k];
    n[2][k] = u[k] - T * f[k];
    n[3][k] = -u[k] - T * f[k];
    n[4][k] = -f[k];
  }
  const double b = v->b;
  for (uint32_t i = 0; i < v->nb; ++i) {
    const uint32_t bx = fcs_unmorton_x(i), by = fcs_unmorton_y(i), bz = fcs_unmorton_z(i);
    const double lo[3] = {bx * b - 1.0, by * b - 1.0, bz * b - 1.0};
    const double hi[3] = {(bx + 1) * b + 1.0, (by + 1) * b + 1.0, (bz + 1) * b + 1.0};
    for (int q = 0; q < 5; ++q) {
      double d = 0;
      for (int k = 0; k < 3; ++k) d += n[q][k] * ((n[q][k] >= 0 ? lo[k] : hi[k]) - p[k]);
      if (d > 0) {
        vis[i] = 0;
        break;
      }
    }
  }
  return vis;
}

template <typename T>
int dev_upload(fcs_ctx* c, T** dst, const void* src, size_t bytes) {
  HIP_TRY(c, hipMalloc((void**)dst, bytes ? bytes : 16));
  if (bytes) HIP_TRY(c, hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
  return 0;
}

}  // namespace

extern "C" {

int fcs_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int fcs_create(int hip_device, fcs_ctx** out) {
  if (!out) return FCS_E_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return FCS_E_NODEVICE;
  if (hip_device < 0 || hip_device >= n) return FCS_E_INVALID;
  fcs_ctx* c = new fcs_ctx();
  c->device = hip_device;
  if (hipSetDevice(hip_device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->aux_done, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->main_mark, hipEventDisableTiming) != hipSuccess) {
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->aux) (void)hipStreamDestroy(c->aux);
    delete c;
    return FCS_E_HIP;
  }
  if (fcs_scratch_setup(hip_device) != 0) {
    (void)hipEventDestroy(c->aux_done);
    (void)hipEventDestroy(c->main_mark);
    (void)hipStreamDestroy(c->stream);
    (void)hipStreamDestroy(c->aux);
    delete c;
    return FCS_E_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, hip_device) == hipSuccess) c->num_cu = prop.multiProcessorCount;
  for (auto& e : c->ev) (void)hipEventCreate(&e);
  *out = c;
  return 0;
}

void fcs_destroy(fcs_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  while (!c->volumes.empty()) fcs_volume_release_device(c->volumes.back());
  while (!c->mlps.empty()) fcs_mlp_release_device(c->mlps.back());
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->aux_done) (void)hipEventDestroy(c->aux_done);
  if (c->main_mark) (void)hipEventDestroy(c->main_mark);
  if (c->aux) (void)hipStreamDestroy(c->aux);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* fcs_last_error(const fcs_ctx* c) { return c ? c->err.c_str() : "no context"; }

int fcs_sync(fcs_ctx* c) {
  if (!c) return FCS_E_INVALID;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return 0;
}

int fcs_set_option(fcs_ctx* c, int option, int64_t value) {
  if (!c) return FCS_E_INVALID;
  if (option == FCS_OPT_RANS_MAPPING) {
    if (value < FCS_RANS_MAPPING_AUTO || value > FCS_RANS_MAPPING_WAVE_WIN_BALLOT)
      return fail(c, FCS_E_INVALID, "bad mapping");
    c->rans_mapping = (int)value;
    return 0;
  }
  if (option == FCS_OPT_OPS_SPLIT) {
    if (value < FCS_OPS_SPLIT_AUTO || value > FCS_OPS_SPLIT_4WG) return fail(c, FCS_E_INVALID, "bad ops split");
    c->ops_split = (int)value;
    return 0;
  }
  if (option == FCS_OPT_MARCH_STATS) {
    if (value < 0 || value > 1) return fail(c, FCS_E_INVALID, "bad march stats flag");
    c->march_stats = value != 0;
    return 0;
  }
  if (option == FCS_OPT_OPS_HEAD) {
    if (value < FCS_OPS_HEAD_AUTO || value > FCS_OPS_HEAD_OFF) return fail(c, FCS_E_INVALID, "bad head mode");
    c->ops_head = (int)value;
    return 0;
  }
  return fail(c, FCS_E_INVALID, "unknown option");
}

int fcs_get_timing(const fcs_ctx* c, fcs_timing* out) {
  if (!c || !out) return FCS_E_INVALID;
  *out = c->timing;
  return 0;
}

static int load_csdb_body(fcs_ctx* c, const uint8_t* bytes, size_t n, uint32_t bso, fcs_volume** out) {
  if (!c || !bytes || !out) return fail(c, FCS_E_INVALID, "null argument");
  *out = nullptr;
  // magic first, with the reference loaders' message (fileIO.cpp:118-119, datasetLoader.js:76-77)
  if (n < 4 || rd32(bytes) != FCS_MAGIC) return fail(c, FCS_E_FORMAT, "Invalid file format");
  if (n < 16 + 8 + 128) return fail(c, FCS_E_FORMAT, "stream too short for a .csdb header");
  HIP_TRY(c, hipSetDevice(c->device));
  auto v = new fcs_volume();
  v->ctx = c;
  v->nb = rd32(bytes + 8);
  v->header_b = rd32(bytes + 12);
  v->b = bso ? bso : v->header_b;
  const uint32_t symbols = rd32(bytes + 16), total = rd32(bytes + 20);
  auto bad = [&](const std::string& m) {
    fcs_volume_free(v);
    return fail(c, FCS_E_FORMAT, m);
  };
  v->L = fcs_log2u(v->b);
  if (v->b < 2 || v->b > 64 || (1u << v->L) != v->b) return bad("brick size must be a power of two in [2, 64]");
  if (symbols != 16 || total != FCS_RANS_M) return bad("rANS model must have 16 symbols and total 4096");
  uint32_t acc = 0;
  for (int i = 0; i < 16; ++i) {
    v->freq[i] = rd32(bytes + 24 + 4 * i);
    v->cum[i] = rd32(bytes + 24 + 64 + 4 * i);
    if (v->cum[i] != acc || v->freq[i] == 0) return bad("inconsistent rANS model");
    acc += v->freq[i];
  }
  if (acc != FCS_RANS_M) return bad("rANS frequencies do not sum to 4096");
  v->cum[16] = FCS_RANS_M;
  // brick grid: a full power-of-two cube (bricks are indexed by Morton code)
  uint32_t B = 1;
  while ((uint64_t)B * B * B < v->nb) B <<= 1;
  if ((uint64_t)B * B * B != v->nb || B > 1024) return bad("brick count must be a power-of-two cube");
  v->B = B;
  // every brick takes at least 24 bytes (ID, P, one palette entry, E,
  // nSymbols): reject a count the input cannot hold before sizing anything by it
  if (v->nb > (n - (16 + 8 + 128)) / 24) return bad("brick count exceeds the stream length");
  struct Raw { uint32_t id, P, E, nsym; size_t pal, str; };
  std::vector<Raw> raw(v->nb);
  size_t p = 16 + 8 + 128;
  for (uint32_t i = 0; i < v->nb; ++i) {
    if (p + 8 > n) return bad("truncated brick header");
    Raw r;
    r.id = rd32(bytes + p);
    r.P = rd32(bytes + p + 4);
    p += 8;
    if (r.P == 0 || p + 8ull * r.P + 4 > n) return bad("truncated palette");
    r.pal = p;
    p += 8ull * r.P;
    r.E = rd32(bytes + p);
    p += 4;
    if (p + (uint64_t)r.E + 4 > n) return bad("truncated stream");
    r.str = p;
    p += r.E;
    r.nsym = rd32(bytes + p);
    p += 4;
    if (r.id >= v->nb) return bad("brick ID outside the brick grid");
    raw[i] = r;
  }
  std::sort(raw.begin(), raw.end(), [](const Raw& a, const Raw& b) { return a.id < b.id; });
  for (uint32_t i = 0; i < v->nb; ++i)
    if (raw[i].id != i) return bad("brick IDs are not a permutation of the grid");
  // host staging
  uint64_t str_total = 0, nib_total = 0, pal_total = 0;
  v->hb.resize(v->nb);
  v->empty.resize(v->nb);
  v->pal0.resize(v->nb);
  for (uint32_t i = 0; i < v->nb; ++i) {
    DevBrick& d = v->hb[i];
    d.str_off = str_total;
    d.E = raw[i].E;
    d.nsym = raw[i].nsym;
    d.pal_off = (uint32_t)pal_total;
    d.P = raw[i].P;
    d.nib_off = nib_total;
    str_total += (raw[i].E + 15) & ~15ull;
    // whole 64-byte lines (the lane decoder stores line by line)
    nib_total += 16ull * std::max<uint64_t>(1, ((uint64_t)raw[i].nsym + 127) / 128);
    pal_total += raw[i].P;
    v->compressed += raw[i].E;
    v->symbols += raw[i].nsym;
  }
  v->palette_entries = pal_total;
  if (pal_total >= (1ull << 32)) return bad("palette too large");
  std::vector<uint8_t> hs(str_total + 16, 0);
  std::vector<uint64_t> hp(pal_total);
  bool wide = false;
  for (uint32_t i = 0; i < v->nb; ++i) {
    memcpy(hs.data() + v->hb[i].str_off, bytes + raw[i].str, raw[i].E);
    memcpy(hp.data() + v->hb[i].pal_off, bytes + raw[i].pal, 8ull * raw[i].P);
    bool all0 = true;
    for (uint32_t k = 0; k < raw[i].P; ++k) {
      const uint64_t x = hp[v->hb[i].pal_off + k];
      all0 &= (x == 0);
      wide |= (x >> 32) != 0;
    }
    v->empty[i] = all0;
    v->pal0[i] = hp[v->hb[i].pal_off];
  }
  v->label_bits = wide ? 64 : 32;
  set_lighting(v, 25.0f, 0.5f, 0.8f);
  int rc = 0;
  if ((rc = dev_upload(c, &v->d_stream, hs.data(), hs.size())) ||
      (rc = dev_upload(c, &v->d_bricks, v->hb.data(), v->hb.size() * sizeof(DevBrick)))) {
    fcs_volume_free(v);
    return rc;
  }
  if (wide) {
    rc = dev_upload(c, (uint64_t**)&v->d_pal, hp.data(), hp.size() * 8);
    if (!rc) rc = dev_upload(c, (uint64_t**)&v->d_pal0, v->pal0.data(), v->pal0.size() * 8);
  } else {
    std::vector<uint32_t> h32(hp.begin(), hp.end()), p32(v->pal0.begin(), v->pal0.end());
    rc = dev_upload(c, (uint32_t**)&v->d_pal, h32.data(), h32.size() * 4);
    if (!rc) rc = dev_upload(c, (uint32_t**)&v->d_pal0, p32.data(), p32.size() * 4);
  }
  if (rc) { fcs_volume_free(v); return rc; }
  if (hipMalloc(&v->d_nib, (nib_total + 8) * 4) != hipSuccess ||
      hipMalloc(&v->d_rres, (size_t)v->nb * sizeof(RansResult)) != hipSuccess ||
      hipMalloc(&v->d_counter, 64) != hipSuccess || hipMalloc(&v->d_meta, (size_t)v->nb * 4) != hipSuccess ||
      hipMalloc(&v->d_offs, (size_t)v->nb * 8) != hipSuccess) {
    fcs_volume_free(v);
    return fail(c, FCS_E_NOMEM, "device allocation failed");
  }
  (void)hipMemset(v->d_nib, 0, (nib_total + 8) * 4);
  {
    std::vector<uint32_t> tab(FCS_RANS_M);
    for (uint32_t x = 0; x < FCS_RANS_M; ++x) {
      uint32_t sym = 0;
      while (sym + 1 < 16 && x >= v->cum[sym + 1]) ++sym;
      tab[x] = (sym << 25) | (v->freq[sym] << 12) | (x - v->cum[sym]);
    }
    if (int rc2 = dev_upload(c, &v->d_rtab, tab.data(), tab.size() * 4)) { fcs_volume_free(v); return rc2; }
  }
  {
    // rANS launch order: longest streams first, so that a wave's lanes
    // decode streams of about the same length (a wave runs as long as its
    // longest lane) and the longest ones start first
    std::vector<uint32_t> order(v->nb);
    for (uint32_t i = 0; i < v->nb; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t x, uint32_t y) { return v->hb[x].nsym > v->hb[y].nsym; });
    if (int rc2 = dev_upload(c, &v->d_rlist, order.data(), order.size() * 4)) { fcs_volume_free(v); return rc2; }
    v->rlist = std::move(order);
  }
  v->pal_host = std::move(hp);
  {
    static std::atomic<uint64_t> next_serial{1};
    v->serial = next_serial.fetch_add(1);
  }
  v->res_lod.assign(v->nb, -1);
  v->res_pyr.assign(v->nb, 0);
  v->res_alloc.assign(v->nb, Alloc{});
  v->free_lists.assign(16, {});
  v->rans_status.assign(v->nb, 0);
  c->volumes.push_back(v);
  *out = v;
  return 0;
}

int fcs_load_csdb(fcs_ctx* c, const uint8_t* bytes, size_t n, uint32_t bso, fcs_volume** out) {
  return fcs_guard(c, [&] { return load_csdb_body(c, bytes, n, bso, out); });
}

}  // extern "C"

fcs_ctx* fcs_volume_context(const fcs_volume* v) { return v ? v->ctx : nullptr; }

int fcs_volume_publish(fcs_volume* v, fcs_volume_desc* o, bool shading) {
  if (!v || !v->ctx || !o) return FCS_E_INVALID;
  fcs_ctx* c = v->ctx;
  HIP_TRY(c, hipSetDevice(c->device));
  if (int rc = settle_prefetch(v)) return rc;
  if (int rc = pull_residency(v)) return rc;
  if (int rc = push_residency(v)) return rc;
  o->serial = v->serial;
  o->res = v->d_res;
  o->pal0 = v->d_pal0;
  o->cache = v->d_cache;
  o->nb = v->nb;
  o->b = v->b;
  o->L = v->L;
  o->B = v->B;
  o->label_bits = v->label_bits;
  o->tf_keys = nullptr;
  o->tf_rgba = nullptr;
  o->tf_mask = 0;
  if (shading) {
    if (int rc = ensure_tf(v)) return rc;
    o->tf_keys = v->d_tf_keys;
    o->tf_rgba = v->d_tf_rgba;
    o->tf_mask = v->tf_mask;
  }
  for (int k = 0; k < 3; ++k) o->light[k] = v->light[k];
  o->ambient = v->ambient;
  o->diffuse = v->diffuse;
  o->shade_epoch = v->shade_epoch;
  return 0;
}

void fcs_volume_release_device(fcs_volume* v) {
  fcs_ctx* c = v->ctx;
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)settle_prefetch(v);
  (void)settle_dev_prefetch(v);
  (void)hipStreamSynchronize(c->stream);
  free_dev_frames(v);
  void* ptrs[] = {v->d_stream, v->d_bricks, v->d_pal, v->d_pal0, v->d_nib, v->d_rres, v->d_counter, v->d_rtab, v->d_rlist,
                  v->d_cache, v->d_meta, v->d_offs, v->d_flabel, v->d_fvoxel, v->d_ft, v->d_visited,
                  v->d_misses, v->d_csum, v->d_res, v->d_tf_keys, v->d_tf_rgba, v->d_frgba, v->d_mstats};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  free_lane(v->lane_main);
  free_lane(v->lane_aux);
  if (v->h_rres) (void)hipHostFree(v->h_rres);
  if (v->h_vstatus) (void)hipHostFree(v->h_vstatus);
  v->h_rres = nullptr;
  v->h_vstatus = nullptr;
  auto& vs = c->volumes;
  vs.erase(std::remove(vs.begin(), vs.end(), v), vs.end());
  v->ctx = nullptr;
}

extern "C" {

void fcs_volume_free(fcs_volume* v) {
  if (!v) return;
  fcs_volume_release_device(v);
  delete v;
}

int fcs_volume_get_info(const fcs_volume* v, fcs_volume_info* o) {
  if (!v || !v->ctx || !o) return FCS_E_INVALID;
  memset(o, 0, sizeof *o);
  o->num_bricks = v->nb;
  o->brick_size = v->b;
  o->levels = v->L;
  o->bricks_per_axis = v->B;
  o->grid_size = v->B * v->b;
  o->label_bits = v->label_bits;
  o->header_brick_size = v->header_b;
  o->rans_ready = v->rans_ready;
  o->compressed_bytes = v->compressed;
  o->total_symbols = v->symbols;
  o->palette_entries = v->palette_entries;
  o->cache_capacity_voxels = v->cap;
  memcpy(o->rans_freq, v->freq, sizeof o->rans_freq);
  memcpy(o->rans_cum, v->cum, sizeof o->rans_cum);
  return 0;
}

// rANS launch over every brick, in launch order (longest first)
RansArgs rans_args(const fcs_volume* v) {
  RansArgs a{};
  a.bricks = v->d_bricks;
  a.stream = v->d_stream;
  a.list = v->d_rlist;
  a.nlist = v->nb;
  a.L = v->L;
  memcpy(a.freq, v->freq, sizeof a.freq);
  memcpy(a.cum, v->cum, sizeof a.cum);
  a.nib = v->d_nib;
  a.res = v->d_rres;
  a.table = v->d_rtab;
  return a;
}

// Per-brick rANS results of the last launch (on the context stream, which
// must have every rANS launch queued) into the volume; true if any failed.
int settle_rans(fcs_volume* v, uint32_t* out_status, bool* bad) {
  fcs_ctx* c = v->ctx;
  if (!v->h_rres)
    HIP_TRY(c, hipHostMalloc((void**)&v->h_rres, std::max<size_t>(v->nb, 1) * sizeof(RansResult), hipHostMallocDefault));
  const RansResult* rr = v->h_rres;
  HIP_TRY(c, hipMemcpyAsync(v->h_rres, v->d_rres, (size_t)v->nb * sizeof(RansResult), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  *bad = false;
  for (uint32_t i = 0; i < v->nb; ++i) {
    v->rans_status[i] = rr[i].status;
    if (out_status) out_status[i] = rr[i].status;
    *bad |= rr[i].status != 0;
  }
  v->rans_ready = true;
  return 0;
}

static int decode_rans_body(fcs_volume* v, uint32_t* out_status) {
  if (!v || !v->ctx) return FCS_E_INVALID;
  fcs_ctx* c = v->ctx;
  HIP_TRY(c, hipSetDevice(c->device));
  RansArgs a = rans_args(v);
  auto timed = [&](int mode, float* ms) -> int {
    HIP_TRY(c, hipEventRecord(c->ev[2], c->stream));
    HIP_TRY(c, fcs_launch_rans(a, mode, c->stream));
    HIP_TRY(c, hipEventRecord(c->ev[3], c->stream));
    HIP_TRY(c, hipEventSynchronize(c->ev[3]));
    HIP_TRY(c, hipEventElapsedTime(ms, c->ev[2], c->ev[3]));
    return 0;
  };
  int mode = c->rans_mapping;
  if (mode == FCS_RANS_MAPPING_AUTO) {
    if (!v->rans_tuned) {
      // autotuner: every mapping over the whole volume (identical outputs),
      // after a one-brick launch of each so code loading is not timed
      RansArgs one = a;
      one.list = v->d_rlist + (v->nb - 1);  // the shortest stream (the list runs longest first)
      one.nlist = 1;
      constexpr int kModes = 6;
      const int modes[kModes] = {FCS_RANS_WAVE, FCS_RANS_WAVE_BALLOT, FCS_RANS_WAVE_LDS, FCS_RANS_LANE,
                                 FCS_RANS_WIN, FCS_RANS_WIN_BALLOT};
      for (int md : modes) HIP_TRY(c, fcs_launch_rans(one, md, c->stream));
      float t[kModes] = {0, 0, 0, 0, 0, 0};
      for (int k = 0; k < kModes; ++k)
        if (int rc = timed(modes[k], &t[k])) return rc;
      c->timing.tune_wave_ms = t[0];
      c->timing.tune_ballot_ms = t[1];
      c->timing.tune_wave_lds_ms = t[2];
      c->timing.tune_lane_ms = t[3];
      c->timing.tune_win_ms = t[4];
      c->timing.tune_win_ballot_ms = t[5];
      int best = 0;
      for (int k = 1; k < kModes; ++k)
        if (t[k] < t[best]) best = k;
      v->rans_tuned = modes[best];
    }
    mode = v->rans_tuned;
  }
  float ms = 0;
  if (int rc = timed(mode, &ms)) return rc;
  c->timing.rans_mapping = (uint32_t)mode;
  bool bad = false;
  if (int rc = settle_rans(v, out_status, &bad)) return rc;
  c->timing.rans_ms = ms;
  c->timing.rans_symbols = v->symbols;
  c->timing.rans_bytes = v->compressed;
  if (bad) return fail(c, FCS_E_CORRUPT, "one or more brick streams failed rANS / structure checks");
  return 0;
}

int fcs_decode_rans(fcs_volume* v, uint32_t* out_status) {
  return fcs_guard(v ? v->ctx : nullptr, [&] { return decode_rans_body(v, out_status); });
}

int fcs_cache_reserve(fcs_volume* v, uint64_t voxels) {
  if (!v || !v->ctx || voxels == 0) return FCS_E_INVALID;
  fcs_ctx* c = v->ctx;
  HIP_TRY(c, hipSetDevice(c->device));
  if (int rc = settle_prefetch(v)) return rc;
  if (int rc = settle_dev_prefetch(v)) return rc;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  drop_dev_graphs(v);  // they hold the pool's address
  if (v->d_cache) { (void)hipFree(v->d_cache); v->d_cache = nullptr; }
  v->cap = 0;
  v->direct = false;
  v->stride = 0;
  if (hipMalloc(&v->d_cache, voxels * v->lbytes()) != hipSuccess)
    return fail(c, FCS_E_NOMEM, "brick cache pool allocation failed");
  v->cap = voxels;
  return fcs_cache_clear(v);
}

int fcs_cache_clear(fcs_volume* v) {
  if (!v || !v->ctx) return FCS_E_INVALID;
  if (int rc = settle_prefetch(v)) return rc;
  if (int rc = settle_dev_prefetch(v)) return rc;
  v->res_dev_newer = false;  // every brick dropped: the host copy is current again
  v->bump = 0;
  for (auto& f : v->free_lists) f.clear();
  std::fill(v->res_lod.begin(), v->res_lod.end(), (int16_t)-1);
  std::fill(v->res_pyr.begin(), v->res_pyr.end(), (uint8_t)0);
  ++v->res_epoch;
  return 0;
}

int64_t fcs_cache_resident(fcs_volume* v, uint32_t* bricks, uint8_t* lods, uint64_t cap) {
  if (!v || !v->ctx) return FCS_E_INVALID;
  if (int rc = pull_residency(v)) return rc;
  uint64_t n = 0;
  for (uint32_t i = 0; i < v->nb; ++i) {
    if (v->res_lod[i] < 0) continue;
    if (n < cap) {
      if (bricks) bricks[n] = i;
      if (lods) lods[n] = (uint8_t)v->res_lod[i];
    }
    ++n;
  }
  return (int64_t)n;
}

int fcs_cache_reserve_direct(fcs_volume* v) {
  if (!v || !v->ctx) return FCS_E_INVALID;
  const uint64_t finest = 1ull << (3 * v->L);
  if (int rc = fcs_cache_reserve(v, finest * v->nb)) return rc;
  v->direct = true;
  v->stride = finest;
  return 0;
}

// The requests of one call decoded through lane `ln`. With `no_alloc` the
// call allocates no device memory: the cache pool and the rANS stage must
// already exist (the lane's buffers are the caller's).
static int request_bricks_impl(fcs_volume* v, const uint32_t* brick, const uint8_t* lod, uint32_t n,
                               uint64_t* out_off, uint32_t* out_status, DecodeLane& ln, bool no_alloc) {
  fcs_ctx* c = v->ctx;
  HIP_TRY(c, hipSetDevice(c->device));
  if (int rc = settle_prefetch(v)) return rc;
  if (int rc = pull_residency(v)) return rc;
  if (no_alloc) {
    if (!v->rans_ready) return fail(c, FCS_E_INVALID, "fcs_decode_rans has not run");
    if (!v->d_cache) return fail(c, FCS_E_INVALID, "no brick cache pool reserved");
  }
  if (!v->rans_ready) {
    int rc = fcs_decode_rans(v, nullptr);
    if (rc && rc != FCS_E_CORRUPT) return rc;
  }
  if (int rc = ensure_cache(v)) return rc;
  std::vector<DecodeReq> reqs;
  std::vector<Alloc> deferred;
  std::vector<int64_t> req_of(n, -1);
  std::vector<uint32_t> st(n, 0);
  // direct-mapped pool: a brick holds one LOD at a time, so only its last
  // request of the call is decoded; earlier ones share that request's slot
  std::vector<int64_t> last_of;
  if (v->direct) {
    last_of.assign(v->nb, -1);
    for (uint32_t i = 0; i < n; ++i)
      if (brick[i] < v->nb && lod[i] <= v->L) last_of[brick[i]] = i;
  }
  std::vector<int64_t> alias(n, -1);
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t bi = brick[i], t = lod[i];
    if (bi >= v->nb || t > v->L) {
      st[i] = FCS_ST_BAD_REQUEST;
      if (out_off) out_off[i] = ~0ull;
      continue;
    }
    if (v->direct && last_of[bi] != (int64_t)i) {
      alias[i] = last_of[bi];
      continue;
    }
    if (v->res_lod[bi] >= 0 && ((v->res_pyr[bi] && (uint32_t)v->res_lod[bi] >= t) ||
                                (!v->res_pyr[bi] && (uint32_t)v->res_lod[bi] == t))) {
      if (out_off) out_off[i] = v->res_alloc[bi].off + (v->res_pyr[bi] ? fcs_pyramid_level_offset(t) : 0);
      continue;
    }
    cache_release(v, bi, &deferred);
    uint64_t off;
    const uint64_t sz = 1ull << (3 * t);
    if (!cache_alloc(v, bi, sz, &off)) {
      st[i] = FCS_ST_CACHE_FULL;
      if (out_off) out_off[i] = ~0ull;
      continue;
    }
    v->res_lod[bi] = (int16_t)t;
    v->res_pyr[bi] = 0;
    v->res_alloc[bi] = Alloc{off, sz};
    ++v->res_epoch;
    req_of[i] = (int64_t)reqs.size();
    reqs.push_back(DecodeReq{bi, t, off});
    if (out_off) out_off[i] = off;
  }
  std::vector<uint32_t> rs;
  const int rc_launch = run_decode(v, ln, reqs, rs, nullptr);
  cache_free_deferred(v, deferred);
  if (rc_launch) return rc_launch;
  bool bad = false;
  for (uint32_t i = 0; i < n; ++i) {
    if (req_of[i] >= 0) st[i] |= rs[req_of[i]];
  }
  for (uint32_t i = 0; i < n; ++i) {
    if (alias[i] >= 0) {
      st[i] = st[alias[i]];
      if (out_off) out_off[i] = out_off[alias[i]];
    }
    bad |= st[i] != 0;
    if (out_status) out_status[i] = st[i];
  }
  if (bad) return fail(c, FCS_E_CORRUPT, "one or more requests failed (see status words)");
  return 0;
}

int fcs_request_bricks(fcs_volume* v, const uint32_t* brick, const uint8_t* lod, uint32_t n,
                       uint64_t* out_off, uint32_t* out_status) {
  if (!v || !v->ctx || (n && (!brick || !lod))) return FCS_E_INVALID;
  return fcs_guard(v->ctx, [&] { return request_bricks_impl(v, brick, lod, n, out_off, out_status, v->lane_main, false); });
}

int fcs_request_bricks_static(fcs_volume* v, const uint32_t* brick, const uint8_t* lod, uint32_t n,
                              uint64_t* out_off, uint32_t* out_status) {
  if (!v || !v->ctx || (n && (!brick || !lod))) return FCS_E_INVALID;
  fcs_ctx* c = v->ctx;
  if (n > FCS_STATIC_MAX_REQUESTS)
    return fail(c, FCS_E_INVALID, "more than FCS_STATIC_MAX_REQUESTS requests in one call");
  return fcs_guard(c, [&] {
    // the lane's device buffers are the GPU's fixed scratch: ensure_req sees
    // them as already large enough and allocates nothing
    DecodeLane ln;
    if (int rc = fcs_scratch_lane(c->device, &ln.d_counter, &ln.d_req, &ln.d_status))
      return fail(c, rc, "no fixed scratch on this device");
    ln.cap = FCS_STATIC_MAX_REQUESTS;
    return request_bricks_impl(v, brick, lod, n, out_off, out_status, ln, true);
  });
}

// Cache slots and requests for every brick at `lod` (brick i at slot i).
int full_requests(fcs_volume* v, uint32_t lod, int pyramid, std::vector<DecodeReq>& reqs) {
  const uint32_t t = std::min(lod, v->L);
  const uint64_t per = pyramid ? fcs_pyramid_level_offset(t + 1) : (1ull << (3 * t));
  if (!v->d_cache || (v->direct ? v->stride < per : v->cap < per * v->nb)) {
    if (int rc = fcs_cache_reserve(v, per * v->nb)) return rc;
  }
  fcs_cache_clear(v);
  reqs.resize(v->nb);
  const uint64_t slot = v->direct ? v->stride : per;
  for (uint32_t i = 0; i < v->nb; ++i) {
    const uint64_t off = slot * i;
    reqs[i] = DecodeReq{i, t | (pyramid ? FCS_REQ_PYRAMID : 0u), off};
    v->res_lod[i] = (int16_t)t;
    v->res_pyr[i] = pyramid ? 1 : 0;
    v->res_alloc[i] = Alloc{off, per};
    ++v->res_epoch;
  }
  v->bump = per * v->nb;
  return 0;
}

static int decode_all_body(fcs_volume* v, uint32_t lod, int pyramid) {
  if (!v || !v->ctx) return FCS_E_INVALID;
  fcs_ctx* c = v->ctx;
  HIP_TRY(c, hipSetDevice(c->device));
  if (int rc = settle_prefetch(v)) return rc;
  if (!v->rans_ready) {
    int rc = fcs_decode_rans(v, nullptr);
    if (rc && rc != FCS_E_CORRUPT) return rc;
  }
  std::vector<DecodeReq> reqs;
  if (int rc = full_requests(v, lod, pyramid, reqs)) return rc;
  std::vector<uint32_t> rs;
  if (int rc = run_decode(v, v->lane_main, reqs, rs, nullptr)) return rc;
  uint32_t bad = 0, first = ~0u, any = 0;
  for (uint32_t i = 0; i < v->nb; ++i)
    if (rs[i]) {
      ++bad;
      any |= rs[i];
      if (first == ~0u) first = i;
    }
  if (bad)
    return fail(c, FCS_E_CORRUPT,
                std::to_string(bad) + " bricks failed decoding (status bits 0x" + [&] {
                  char b[16];
                  snprintf(b, sizeof b, "%x", any);
                  return std::string(b);
                }() + ", first brick " + std::to_string(first) + ")");
  return 0;
}

int fcs_decode_all(fcs_volume* v, uint32_t lod, int pyramid) {
  return fcs_guard(v ? v->ctx : nullptr, [&] { return decode_all_body(v, lod, pyramid); });
}

// Both stages for every brick in one call, on the context stream with no host
// synchronisation between them: the rANS launch over every stream (longest
// first), then the op decoder over every brick in the same order. Running the
// short streams' two stages on a second stream beside the long streams' rANS
// stage measured slower than this order once the rANS stage had become short,
// and one call in three much slower: the two stages compete for the same LDS.
static int decode_volume_body(fcs_volume* v, uint32_t lod, int pyramid, uint32_t* out_status) {
  if (!v || !v->ctx) return FCS_E_INVALID;
  fcs_ctx* c = v->ctx;
  HIP_TRY(c, hipSetDevice(c->device));
  if (int rc = settle_prefetch(v)) return rc;
  int mode = c->rans_mapping;
  if (mode == FCS_RANS_MAPPING_AUTO) {
    if (!v->rans_tuned) {
      // the volume's first rANS pass picks the mapping
      const int rc = fcs_decode_rans(v, nullptr);
      if (rc && rc != FCS_E_CORRUPT) return rc;
    }
    mode = v->rans_tuned;
  }
  if (!v->h_vstatus)
    HIP_TRY(c, hipHostMalloc((void**)&v->h_vstatus, std::max<size_t>(v->nb, 1) * sizeof(uint32_t), hipHostMallocDefault));
  // The rANS stage needs no requests, so it goes first and the host builds
  // the requests while it runs.
  HIP_TRY(c, hipEventRecord(c->ev[2], c->stream));
  HIP_TRY(c, fcs_launch_rans(rans_args(v), mode, c->stream));
  std::vector<DecodeReq> reqs;
  if (int rc = full_requests(v, lod, pyramid, reqs)) return rc;
  DecodeLane& ln = v->lane_main;
  const uint32_t t = std::min(lod, v->L);
  const uint64_t per = pyramid ? fcs_pyramid_level_offset(t + 1) : (1ull << (3 * t));
  const uint64_t slot = v->direct ? v->stride : per;
  auto& vp = v->vol_plan;
  // The requests (brick i at slot i, in rlist order) depend only on the LOD,
  // the layout and the slot size: while lane_main still holds the ones a
  // previous call binned and uploaded, they are launched again as they are.
  const bool reuse = vp.gen != 0 && vp.gen == ln.gen && vp.t == t && vp.pyr == (pyramid != 0) && vp.slot == slot &&
                     ln.reqs.size() == v->nb;
  if (!reuse) {
    std::vector<DecodeReq> rl;
    rl.reserve(v->nb);
    for (uint32_t k = 0; k < v->nb; ++k) rl.push_back(reqs[v->rlist[k]]);
    // uploaded on the aux stream, beside the rANS stage (the aux stream is
    // idle here: prefetches were settled above and by the cache clear)
    if (int rc = stage_decode(v, ln, rl, c->aux)) return rc;
    HIP_TRY(c, hipEventRecord(c->aux_done, c->aux));
    HIP_TRY(c, hipStreamWaitEvent(c->stream, c->aux_done, 0));
    vp.gen = ln.gen;
    vp.t = t;
    vp.pyr = pyramid != 0;
    vp.slot = slot;
  }
  if (int rc = issue_decode(v, ln, c->stream, false, 0, false)) return rc;
  HIP_TRY(c, hipEventRecord(c->ev[3], c->stream));
  account_decode(v, reqs);  // host bookkeeping while the device decodes
  if (!ln.reqs.empty())
    HIP_TRY(c, hipMemcpyAsync(v->h_vstatus, ln.d_status, ln.reqs.size() * sizeof(uint32_t), hipMemcpyDeviceToHost,
                              c->stream));
  bool rbad = false;
  if (int rc = settle_rans(v, nullptr, &rbad)) return rc;  // synchronises the context stream
  float ms = 0;
  HIP_TRY(c, hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
  c->timing.pipeline_ms = ms;
  c->timing.rans_mapping = (uint32_t)mode;
  // statuses in launch order -> bricks: launch position j holds request
  // perm[j], which is brick rlist[perm[j]]
  uint32_t nbad = 0, any = 0, first = ~0u;
  const size_t n = ln.reqs.size();
  for (size_t j = 0; j < n; ++j) {
    const uint32_t b = v->rlist[ln.perm[j]];
    const uint32_t st = v->rans_status[b] | v->h_vstatus[j];
    if (out_status) out_status[b] = st;
    if (st) {
      ++nbad;
      any |= st;
      first = std::min(first, b);
    }
  }
  if (nbad) {
    char bits[16];
    snprintf(bits, sizeof bits, "%x", any);
    return fail(c, FCS_E_CORRUPT, std::to_string(nbad) + " bricks failed decoding (status bits 0x" + bits +
                                      ", first brick " + std::to_string(first) + ")");
  }
  return 0;
}

int fcs_decode_volume(fcs_volume* v, uint32_t lod, int pyramid, uint32_t* out_status) {
  return fcs_guard(v ? v->ctx : nullptr, [&] { return decode_volume_body(v, lod, pyramid, out_status); });
}

int fcs_read_cache(fcs_volume* v, uint64_t off, uint64_t count, void* host) {
  if (!v || !v->ctx || !host) return FCS_E_INVALID;
  // written so that it cannot wrap: a failed request's offset (~0) or a huge
  // count is rejected, not turned into a pointer outside the pool
  if (off > v->cap || count > v->cap - off) return fail(v->ctx, FCS_E_INVALID, "region outside the brick cache pool");
  fcs_ctx* c = v->ctx;
  HIP_TRY(c, hipSetDevice(c->device));
  if (int rc = settle_prefetch(v)) return rc;
  HIP_TRY(c, hipMemcpyAsync(host, (uint8_t*)v->d_cache + off * v->lbytes(), count * v->lbytes(),
                            hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return 0;
}

int64_t fcs_read_nibbles(fcs_volume* v, uint32_t brick, uint8_t* out, uint64_t cap) {
  if (!v || !v->ctx || brick >= v->nb) return FCS_E_INVALID;
  fcs_ctx* c = v->ctx;
  if (!v->rans_ready) return fail(c, FCS_E_INVALID, "fcs_decode_rans has not run");
  const uint32_t nsym = v->hb[brick].nsym;
  const uint64_t nbytes = ((uint64_t)nsym + 1) / 2;
  if (!out) return (int64_t)nbytes;
  if (cap < nbytes) return fail(c, FCS_E_INVALID, "output buffer too small");
  HIP_TRY(c, hipSetDevice(c->device));
  if (int rc = settle_prefetch(v)) return rc;
  std::vector<uint32_t> w(((uint64_t)nsym + 7) / 8);
  if (!w.empty()) {
    HIP_TRY(c, hipMemcpyAsync(w.data(), v->d_nib + v->hb[brick].nib_off, w.size() * 4, hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
  }
  for (uint64_t j = 0; j < nbytes; ++j) {
    const uint64_t k0 = 2 * j, k1 = 2 * j + 1;
    const uint32_t hi = (w[k0 >> 3] >> (4 * (k0 & 7))) & 0xFu;
    const uint32_t lo = k1 < nsym ? (w[k1 >> 3] >> (4 * (k1 & 7))) & 0xFu : 0u;
    out[j] = (uint8_t)((hi << 4) | lo);
  }
  return (int64_t)nbytes;
}

int fcs_read_brick(fcs_volume* v, uint32_t brick, uint32_t lod, void* host) {
  if (!v || !v->ctx || !host || brick >= v->nb || lod > v->L) return FCS_E_INVALID;
  const uint8_t t = (uint8_t)lod;
  uint64_t off = 0;
  uint32_t st = 0;
  int rc = fcs_request_bricks(v, &brick, &t, 1, &off, &st);
  if (rc) return rc;
  return fcs_read_cache(v, off, 1ull << (3 * lod), host);
}

int fcs_brick_lods(const fcs_volume* v, const fcs_camera* cam, uint32_t w, uint32_t h, uint8_t* out) {
  if (!v || !v->ctx || !cam || !out || w == 0 || h == 0) return FCS_E_INVALID;
  // calculateBrickLOD (camera.js:103-137) with screen width min(w, h)
  // (webgpuSetup.js:587); focal length from the camera's tan(fov/2)
  const LodRule rule(v, cam, w, h);
  for (uint32_t i = 0; i < v->nb; ++i) out[i] = rule(i);
  return 0;
}

uint32_t fcs_tf_default_rgba(uint64_t index) {
  const double i = (double)index;
  const double hue = std::fmod(i * 137.5, 360.0);
  const double sat = 0.55 + (double)(index % 3) * 0.08;
  const double lig = 0.35 + (double)(index % 4) * 0.05;
  const double opa = 0.25 + (double)(index % 5) * 0.08;
  // hslToRgb (webgpuSetup.js:490-515)
  const double cc = (1.0 - std::fabs(2.0 * lig - 1.0)) * sat;
  const double x = cc * (1.0 - std::fabs(std::fmod(hue / 60.0, 2.0) - 1.0));
  const double m = lig - cc / 2.0;
  double r, g, b;
  if (hue < 60) { r = cc; g = x; b = 0; }
  else if (hue < 120) { r = x; g = cc; b = 0; }
  else if (hue < 180) { r = 0; g = cc; b = x; }
  else if (hue < 240) { r = 0; g = x; b = cc; }
  else if (hue < 300) { r = x; g = 0; b = cc; }
  else { r = cc; g = 0; b = x; }
  const double R = js_round((r + m) * 255.0), Gc = js_round((g + m) * 255.0), Bc = js_round((b + m) * 255.0);
  const uint32_t r8 = (uint32_t)std::min(std::floor(R * 0.92), 255.0);
  const uint32_t g8 = (uint32_t)std::min(std::floor(Gc * 0.92), 255.0);
  const uint32_t b8 = (uint32_t)std::min(std::floor(Bc * 0.92), 255.0);
  const uint32_t a8 = (uint32_t)std::floor(opa * 255.0);
  return r8 | (g8 << 8) | (b8 << 16) | (a8 << 24);
}

int fcs_set_lighting(fcs_volume* v, const fcs_lighting* l) {
  if (!v || !v->ctx || !l) return FCS_E_INVALID;
  set_lighting(v, l->light_angle_deg, l->ambient, l->diffuse_strength);
  return 0;
}

int fcs_set_label_visibility(fcs_volume* v, const uint64_t* hidden, uint32_t n_hidden, const uint64_t* highlight) {
  if (!v || !v->ctx || (n_hidden && !hidden)) return FCS_E_INVALID;
  if (int rc = ensure_tf(v)) return rc;
  std::vector<uint64_t> hs(hidden, hidden + n_hidden);
  std::sort(hs.begin(), hs.end());
  v->tf_cur = v->tf_base;
  for (size_t i = 0; i < v->tf_labels.size(); ++i) {
    const uint64_t lab = v->tf_labels[i];
    if (std::binary_search(hs.begin(), hs.end(), lab)) {
      v->tf_cur[i] = 0u;
    } else if (highlight && lab == *highlight) {
      // rebuildLabelColors (webgpuSetup.js:840-848)
      const uint32_t c0 = v->tf_base[i];
      const uint32_t r = (uint32_t)std::floor(std::min(255.0, (double)(c0 & 0xFFu) * 1.8));
      const uint32_t g = (uint32_t)std::floor(std::min(255.0, (double)((c0 >> 8) & 0xFFu) * 1.8));
      const uint32_t b = (uint32_t)std::floor(std::min(255.0, (double)((c0 >> 16) & 0xFFu) * 1.8));
      v->tf_cur[i] = r | (g << 8) | (b << 16) | (c0 & 0xFF000000u);
    }
  }
  v->tf_dirty = true;
  return ensure_tf(v);
}

int64_t fcs_transfer_function(fcs_volume* v, uint64_t* labels, uint32_t* rgba, uint64_t cap) {
  if (!v || !v->ctx) return FCS_E_INVALID;
  if (int rc = ensure_tf(v)) return rc;
  const uint64_t n = v->tf_labels.size();
  const uint64_t k = std::min<uint64_t>(n, cap);
  if (labels) std::copy(v->tf_labels.begin(), v->tf_labels.begin() + k, labels);
  if (rgba) std::copy(v->tf_cur.begin(), v->tf_cur.begin() + k, rgba);
  return (int64_t)n;
}

int fcs_camera_orbit(const fcs_volume* v, double yaw, double pitch, double radius, double fov, fcs_camera* o) {
  if (!v || !v->ctx || !o) return FCS_E_INVALID;
  // OrbitController.apply (orbitController.js:100-123) around the volume
  // centre; Camera.getUniformData basis (camera.js:61-77)
  const double G = (double)v->B * v->b;
  const double tx = G / 2, ty = G / 2, tz = G / 2;
  const double cp = std::cos(pitch), sp = std::sin(pitch), cy = std::cos(yaw), sy = std::sin(yaw);
  const double px = tx + radius * cp * sy, py = ty + radius * sp, pz = tz + radius * cp * cy;
  auto norm = [](double& x, double& y, double& z) {
    double l = std::hypot(x, std::hypot(y, z));
    if (l == 0) l = 1;
    x /= l; y /= l; z /= l;
  };
  double fx = tx - px, fy = ty - py, fz = tz - pz;
  norm(fx, fy, fz);
  norm(fx, fy, fz);
  double ux = 0, uy = 1, uz = 0;
  if (std::fabs(fy) > 0.999) { ux = 1; uy = 0; }
  double rx = fy * uz - fz * uy, ry = fz * ux - fx * uz, rz = fx * uy - fy * ux;
  norm(rx, ry, rz);
  double vx = ry * fz - rz * fy, vy = rz * fx - rx * fz, vz = rx * fy - ry * fx;
  norm(vx, vy, vz);
  memset(o, 0, sizeof *o);
  o->position[0] = (float)px; o->position[1] = (float)py; o->position[2] = (float)pz;
  o->tan_half_fov = (float)std::tan(fov * 0.5);
  o->forward[0] = (float)fx; o->forward[1] = (float)fy; o->forward[2] = (float)fz;
  o->right[0] = (float)rx; o->right[1] = (float)ry; o->right[2] = (float)rz;
  o->up[0] = (float)vx; o->up[1] = (float)vy; o->up[2] = (float)vz;
  return 0;
}

}  // extern "C"

namespace {

int frame_lods(fcs_volume* v, const fcs_camera* cam, const uint8_t* brick_lod, uint32_t w, uint32_t h,
               std::vector<uint8_t>& lods) {
  lods.resize(v->nb);
  if (brick_lod) {
    for (uint32_t i = 0; i < v->nb; ++i) lods[i] = std::min<uint8_t>(brick_lod[i], (uint8_t)v->L);
    return 0;
  }
  return fcs_brick_lods(v, cam, w, h, lods.data());
}

// Transfer function, lighting and RGBA output of a shaded march.
template <typename LT>
void shade_args(const fcs_volume* v, bool rgba, MarchArgs<LT>& a) {
  if (!rgba) return;
  a.out_rgba = v->d_frgba;
  a.tf_keys = v->d_tf_keys;
  a.tf_rgba = v->d_tf_rgba;
  a.tf_mask = v->tf_mask;
  for (int k = 0; k < 3; ++k) a.light[k] = v->light[k];
  a.ambient = v->ambient;
  a.diffuse = v->diffuse;
}

// Frame output buffers of at least npx pixels (RGBA when asked), the visited
// mask and the miss counter. Device-frame graphs hold their addresses: they
// are dropped when a buffer moves.
int ensure_frame_buffers(fcs_volume* v, size_t npx, bool rgba) {
  fcs_ctx* c = v->ctx;
  if (npx > v->frame_cap) {
    if (v->df && !v->df->graphs.empty()) {
      HIP_TRY(c, hipStreamSynchronize(c->stream));
      drop_dev_graphs(v);
    }
    void* ptrs[] = {v->d_flabel, v->d_fvoxel, v->d_ft, v->d_frgba};
    for (void* q : ptrs)
      if (q) (void)hipFree(q);
    v->d_flabel = nullptr;
    v->d_fvoxel = nullptr;
    v->d_ft = nullptr;
    v->d_frgba = nullptr;
    v->frame_cap = 0;
    HIP_TRY(c, hipMalloc(&v->d_flabel, npx * v->lbytes()));
    HIP_TRY(c, hipMalloc((void**)&v->d_fvoxel, npx * 8));
    HIP_TRY(c, hipMalloc((void**)&v->d_ft, npx * 4));
    v->frame_cap = npx;
  }
  if (rgba && !v->d_frgba) HIP_TRY(c, hipMalloc((void**)&v->d_frgba, v->frame_cap * 4));
  if (!v->d_visited) {
    HIP_TRY(c, hipMalloc((void**)&v->d_visited, (size_t)v->nb * 4));
    HIP_TRY(c, hipMalloc((void**)&v->d_misses, 64));
  }
  return 0;
}

// Upload the frame's per-brick metadata (LOD, empty, resident at that LOD,
// cache offset) and march every pixel on the context stream (not waited for).
int march_pass(fcs_volume* v, const fcs_camera* cam, const std::vector<uint8_t>& lods, uint32_t w, uint32_t h,
               bool rgba) {
  fcs_ctx* c = v->ctx;
  if (rgba)
    if (int rc = ensure_tf(v)) return rc;
  std::vector<uint32_t> meta(v->nb);
  std::vector<uint64_t> offs(v->nb, 0);
  for (uint32_t i = 0; i < v->nb; ++i) {
    uint32_t m = lods[i];
    if (v->empty[i]) m |= FCS_META_EMPTY;
    if (resident_at(v, i, lods[i])) {
      m |= FCS_META_RESIDENT;
      offs[i] = v->res_alloc[i].off + (v->res_pyr[i] ? fcs_pyramid_level_offset(lods[i]) : 0);
    }
    meta[i] = m;
  }
  if (int rc = ensure_frame_buffers(v, (size_t)w * h, rgba)) return rc;
  HIP_TRY(c, hipMemcpyAsync(v->d_meta, meta.data(), meta.size() * 4, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipMemcpyAsync(v->d_offs, offs.data(), offs.size() * 8, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, fcs_launch_zero(v->d_visited, v->nb, c->stream));
  HIP_TRY(c, fcs_launch_zero(v->d_misses, 1, c->stream));
  if (c->march_stats) {
    if (!v->d_mstats) HIP_TRY(c, hipMalloc((void**)&v->d_mstats, 3 * FCS_MARCH_STATS_SLOTS * 8));
    HIP_TRY(c, fcs_launch_zero(reinterpret_cast<uint32_t*>(v->d_mstats), 6 * FCS_MARCH_STATS_SLOTS, c->stream));
  }
  const float aspect = (float)((double)w / (double)h);
  MarchCam mc;
  memcpy(&mc, cam, sizeof mc);
  HIP_TRY(c, hipEventRecord(c->ev[0], c->stream));
  if (v->label_bits == 32) {
    MarchArgs<uint32_t> a{mc, nullptr, v->d_meta, v->d_offs, (const uint32_t*)v->d_pal0, (const uint32_t*)v->d_cache,
                          (uint32_t*)v->d_flabel, v->d_fvoxel, v->d_ft, v->d_visited, v->d_misses,
                          w, h, (float)w, (float)h, aspect, (float)(v->B * v->b), v->b, v->B};
    shade_args(v, rgba, a);
    a.stats = c->march_stats ? v->d_mstats : nullptr;
    HIP_TRY(c, fcs_launch_march<uint32_t>(a, c->stream));
  } else {
    MarchArgs<uint64_t> a{mc, nullptr, v->d_meta, v->d_offs, (const uint64_t*)v->d_pal0, (const uint64_t*)v->d_cache,
                          (uint64_t*)v->d_flabel, v->d_fvoxel, v->d_ft, v->d_visited, v->d_misses,
                          w, h, (float)w, (float)h, aspect, (float)(v->B * v->b), v->b, v->B};
    shade_args(v, rgba, a);
    a.stats = c->march_stats ? v->d_mstats : nullptr;
    HIP_TRY(c, fcs_launch_march<uint64_t>(a, c->stream));
  }
  HIP_TRY(c, hipEventRecord(c->ev[1], c->stream));
  return 0;
}

// Copy the requested outputs of the last march pass back, wait, and settle.
int finish_frame(fcs_volume* v, const std::vector<uint8_t>& lods, uint32_t w, uint32_t h, fcs_frame* out) {
  fcs_ctx* c = v->ctx;
  const size_t npx = (size_t)w * h;
  if (out->label) HIP_TRY(c, hipMemcpyAsync(out->label, v->d_flabel, npx * v->lbytes(), hipMemcpyDeviceToHost, c->stream));
  if (out->voxel) HIP_TRY(c, hipMemcpyAsync(out->voxel, v->d_fvoxel, npx * 8, hipMemcpyDeviceToHost, c->stream));
  if (out->t_hit) HIP_TRY(c, hipMemcpyAsync(out->t_hit, v->d_ft, npx * 4, hipMemcpyDeviceToHost, c->stream));
  if (out->rgba) HIP_TRY(c, hipMemcpyAsync(out->rgba, v->d_frgba, npx * 4, hipMemcpyDeviceToHost, c->stream));
  if (out->brick_visited)
    HIP_TRY(c, hipMemcpyAsync(out->brick_visited, v->d_visited, (size_t)v->nb * 4, hipMemcpyDeviceToHost, c->stream));
  if (out->brick_lod_used) memcpy(out->brick_lod_used, lods.data(), v->nb);
  std::vector<unsigned long long> mst;
  if (c->march_stats && v->d_mstats) {
    mst.resize(3 * FCS_MARCH_STATS_SLOTS);
    HIP_TRY(c, hipMemcpyAsync(mst.data(), v->d_mstats, mst.size() * 8, hipMemcpyDeviceToHost, c->stream));
  }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (!mst.empty()) {
    uint64_t sum[3] = {0, 0, 0};
    for (uint32_t k = 0; k < 3; ++k)
      for (uint32_t j = 0; j < FCS_MARCH_STATS_SLOTS; ++j) sum[k] += mst[k * FCS_MARCH_STATS_SLOTS + j];
    c->timing.march_samples = sum[0];
    c->timing.march_bricks_entered = sum[1];
    c->timing.march_empty_tested = sum[2];
  }
  v->frame_px = npx;
  if (int rc = settle_prefetch(v)) return rc;
  float ms = 0;
  HIP_TRY(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  out->march_ms += ms;
  c->timing.march_ms = out->march_ms;
  return 0;
}

// ---------------------------------------------------------------------------
// Frames built on the device (direct-mapped pools). Two forms:
//   plain    (fcs_render): frame setup (LOD rule, frustum test) queues every
//            non-empty brick in the frustum that is not resident at its LOD;
//            the op decoder drains the queue; the march follows.
//   visible  (fcs_render_visible): setup, march with what is resident, queue
//            the bricks rays entered but found missing, decode, and march
//            again only if there were misses (a second pass's rays enter a
//            subset of the bricks the first pass's rays entered, so two passes
//            see every brick they need).
// The sequence is captured once per (form, frame size, outputs) as a hipGraph;
// the camera (and fixed LODs) reach it through pinned staging buffers.
enum { FORM_PLAIN = 0, FORM_VISIBLE = 1 };

int ensure_dev_frames(fcs_volume* v) {
  if (v->df) return 0;
  fcs_ctx* c = v->ctx;
  DevFrames* F = new DevFrames();
  v->df = F;
  const size_t nb = v->nb;
  HIP_TRY(c, hipMalloc((void**)&F->d_q, nb * sizeof(DecodeReq)));
  HIP_TRY(c, hipMalloc((void**)&F->d_qst, nb * 4));
  HIP_TRY(c, hipMalloc((void**)&F->d_aq, nb * sizeof(DecodeReq)));
  HIP_TRY(c, hipMalloc((void**)&F->d_aqst, nb * 4));
  HIP_TRY(c, hipMalloc((void**)&F->d_small, (DF_HIST + 1028) * 4));
  HIP_TRY(c, hipMalloc(&F->d_cam, sizeof(MarchCam)));
  HIP_TRY(c, hipMalloc(&F->d_cam_ahead, sizeof(MarchCam)));
  HIP_TRY(c, hipMalloc((void**)&F->d_lod, nb));
  HIP_TRY(c, hipMalloc((void**)&F->d_needed, nb * 4));
  HIP_TRY(c, hipHostMalloc(&F->h_cam, sizeof(MarchCam), hipHostMallocDefault));
  HIP_TRY(c, hipHostMalloc(&F->h_cam_ahead, sizeof(MarchCam), hipHostMallocDefault));
  HIP_TRY(c, hipHostMalloc((void**)&F->h_lod, nb, hipHostMallocDefault));
  HIP_TRY(c, hipHostMalloc((void**)&F->h_small, DF_WORDS * 4, hipHostMallocDefault));
  return 0;
}

// The op decoder over a device-built queue (count in device memory).
int issue_queue_decode(fcs_volume* v, DecodeReq* q, uint32_t* status, uint32_t* qcount, uint32_t* claim,
                       int split, hipStream_t s) {
  fcs_ctx* c = v->ctx;
  const int ksplit = split == FCS_OPS_SPLIT_4WG ? FCS_OPS_4WG : FCS_OPS_3WG;
  if (v->label_bits == 32) {
    OpsArgs<uint32_t> a{v->d_bricks, v->d_nib, (const uint32_t*)v->d_pal, q, (uint32_t*)v->d_cache, status, claim,
                        v->nb, v->nb, v->L, 0u, qcount};
    HIP_TRY(c, fcs_launch_ops<uint32_t>(a, (uint32_t)c->num_cu, s, nullptr, ksplit, FCS_OPS_HEAD_OFF));
  } else {
    OpsArgs<uint64_t> a{v->d_bricks, v->d_nib, (const uint64_t*)v->d_pal, q, (uint64_t*)v->d_cache, status, claim,
                        v->nb, v->nb, v->L, 0u, qcount};
    HIP_TRY(c, fcs_launch_ops<uint64_t>(a, (uint32_t)c->num_cu, s, nullptr, ksplit, FCS_OPS_HEAD_OFF));
  }
  QueueSettleArgs sa{q, qcount, status, v->d_res, v->d_meta, v->df->d_small + DF_ERR, v->nb};
  HIP_TRY(c, fcs_launch_queue_settle(sa, s));
  return 0;
}

template <typename LT>
int dev_march(fcs_volume* v, uint32_t w, uint32_t h, bool rgba, uint32_t* misses, const uint32_t* run_if,
              uint32_t* pix_miss, const uint32_t* pix_only, uint32_t* needed) {
  fcs_ctx* c = v->ctx;
  MarchCam zc;
  memset(&zc, 0, sizeof zc);
  const float aspect = (float)((double)w / (double)h);
  MarchArgs<LT> a{zc, (const MarchCam*)v->df->d_cam, v->d_meta, v->d_offs, (const LT*)v->d_pal0,
                  (const LT*)v->d_cache, (LT*)v->d_flabel, v->d_fvoxel, v->d_ft, v->d_visited, misses,
                  w, h, (float)w, (float)h, aspect, (float)(v->B * v->b), v->b, v->B};
  shade_args(v, rgba, a);
  a.run_if = run_if;
  a.pix_miss = pix_miss;
  a.pix_only = pix_only;
  a.needed = needed;
  HIP_TRY(c, fcs_launch_march<LT>(a, c->stream));
  return 0;
}

// The device work of one frame, in stream order (captured into a graph).
int enqueue_dev_frame(fcs_volume* v, int form, uint32_t w, uint32_t h, bool rgba, bool fixed, int split) {
  fcs_ctx* c = v->ctx;
  DevFrames& F = *v->df;
  hipStream_t s = c->stream;
  uint32_t* sm = F.d_small;
  HIP_TRY(c, hipMemcpyAsync(F.d_cam, F.h_cam, sizeof(MarchCam), hipMemcpyHostToDevice, s));
  if (fixed) HIP_TRY(c, hipMemcpyAsync(F.d_lod, F.h_lod, v->nb, hipMemcpyHostToDevice, s));
  HIP_TRY(c, fcs_launch_zero(sm, 8, s));
  const bool plain = form == FORM_PLAIN;
  FrameSetupArgs fa{v->d_res, (const MarchCam*)F.d_cam, fixed ? F.d_lod : nullptr, v->d_meta, v->d_offs,
                    v->nb, v->b, v->L, w, h, plain ? v->d_res : nullptr, plain ? F.d_q : nullptr,
                    plain ? sm + DF_QCOUNT : nullptr, v->stride};
  HIP_TRY(c, fcs_launch_frame_setup(fa, s, plain));
  if (plain)
    if (int rc = issue_queue_decode(v, F.d_q, F.d_qst, sm + DF_QCOUNT, sm + DF_CLAIM, split, s)) return rc;
  HIP_TRY(c, fcs_launch_zero(v->d_visited, v->nb, s));
  auto march = [&](uint32_t* misses, const uint32_t* run_if, uint32_t* pix_miss, const uint32_t* pix_only,
                   uint32_t* needed) {
    return v->label_bits == 32 ? dev_march<uint32_t>(v, w, h, rgba, misses, run_if, pix_miss, pix_only, needed)
                               : dev_march<uint64_t>(v, w, h, rgba, misses, run_if, pix_miss, pix_only, needed);
  };
  if (plain) {
    if (int rc = march(sm + DF_MISS1, nullptr, nullptr, nullptr, nullptr)) return rc;
  } else {
    // the first march marks the bricks rays enter up to their first miss as
    // visited and the rest as needed; the second re-traces only the rays
    // that missed, adding the bricks they enter to the visited set
    HIP_TRY(c, fcs_launch_zero(F.d_needed, v->nb, s));
    if (int rc = march(sm + DF_MISS1, nullptr, F.d_pixmiss, nullptr, F.d_needed)) return rc;
    MissEnqueueArgs ma{sm + DF_MISS1, v->d_visited, F.d_needed, v->d_meta, v->d_offs, v->d_res, F.d_q,
                       sm + DF_QCOUNT, v->nb, v->stride};
    HIP_TRY(c, fcs_launch_miss_enqueue(ma, s));
    if (int rc = issue_queue_decode(v, F.d_q, F.d_qst, sm + DF_QCOUNT, sm + DF_CLAIM, split, s)) return rc;
    if (int rc = march(sm + DF_MISS2, sm + DF_MISS1, nullptr, F.d_pixmiss, nullptr)) return rc;
  }
  HIP_TRY(c, hipMemcpyAsync(F.h_small, sm, 8 * 4, hipMemcpyDeviceToHost, s));
  return 0;
}

int dev_frame_graph(fcs_volume* v, int form, uint32_t w, uint32_t h, bool rgba, bool fixed, DevGraph** out) {
  fcs_ctx* c = v->ctx;
  DevFrames& F = *v->df;
  const int split = ops_split_of(v);
  const void* tfk = rgba ? (const void*)v->d_tf_keys : nullptr;
  const void* tfr = rgba ? (const void*)v->d_tf_rgba : nullptr;
  const uint64_t se = rgba ? v->shade_epoch : 0;
  for (auto& g : F.graphs)
    if (g.form == form && g.w == w && g.h == h && g.rgba == rgba && g.fixed == fixed && g.cache == v->d_cache &&
        g.flabel == v->d_flabel && g.frgba == (rgba ? v->d_frgba : nullptr) && g.tfk == tfk && g.tfr == tfr &&
        g.shade_epoch == se) {
      *out = &g;
      return 0;
    }
  // graphs of this shape with stale addresses or shading can never be replayed
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  for (auto& g : F.graphs)
    if (g.form == form && g.w == w && g.h == h && g.rgba == rgba && g.fixed == fixed) {
      (void)hipGraphExecDestroy(g.exec);
      (void)hipGraphDestroy(g.graph);
      g.exec = nullptr;
    }
  F.graphs.erase(std::remove_if(F.graphs.begin(), F.graphs.end(), [](const DevGraph& g) { return !g.exec; }),
                 F.graphs.end());
  // the op decoder queries its occupancy on its first launch: not inside a capture
  HIP_TRY(c, v->label_bits == 32 ? fcs_ops_prepare<uint32_t>(v->L, split == FCS_OPS_SPLIT_4WG ? FCS_OPS_4WG : FCS_OPS_3WG)
                                 : fcs_ops_prepare<uint64_t>(v->L, split == FCS_OPS_SPLIT_4WG ? FCS_OPS_4WG : FCS_OPS_3WG));
  HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue_dev_frame(v, form, w, h, rgba, fixed, split);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  HIP_TRY(c, e);
  hipGraphExec_t x = nullptr;
  const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    (void)hipGraphDestroy(g);
    HIP_TRY(c, ei);
  }
  F.graphs.push_back(DevGraph{form, w, h, rgba, fixed, v->d_cache, v->d_flabel, rgba ? v->d_frgba : nullptr, tfk, tfr,
                              se, g, x});
  *out = &F.graphs.back();
  return 0;
}

// One device-built frame; outputs copied back as fcs_render's.
int dev_frame(fcs_volume* v, const fcs_camera* cam, const uint8_t* brick_lod, uint32_t w, uint32_t h, int form,
              fcs_frame* out) {
  fcs_ctx* c = v->ctx;
  if (int rc = settle_prefetch(v)) return rc;
  if (!v->rans_ready) {
    const int rc = fcs_decode_rans(v, nullptr);
    if (rc && rc != FCS_E_CORRUPT) return rc;
  }
  if (int rc = ensure_dev_frames(v)) return rc;
  DevFrames& F = *v->df;
  const bool rgba = out->rgba != nullptr;
  if (rgba)
    if (int rc = ensure_tf(v)) return rc;
  const size_t npx = (size_t)w * h;
  if (int rc = ensure_frame_buffers(v, npx, rgba)) return rc;
  if (npx > F.pix_cap) {
    if (!F.graphs.empty()) {
      HIP_TRY(c, hipStreamSynchronize(c->stream));
      drop_dev_graphs(v);
    }
    if (F.d_pixmiss) (void)hipFree(F.d_pixmiss);
    F.d_pixmiss = nullptr;
    F.pix_cap = 0;
    HIP_TRY(c, hipMalloc((void**)&F.d_pixmiss, npx * 4));
    F.pix_cap = npx;
  }
  if (int rc = push_residency(v)) return rc;
  memcpy(F.h_cam, cam, sizeof(MarchCam));
  if (brick_lod) memcpy(F.h_lod, brick_lod, v->nb);
  // a prefetch decode issued after the previous frame finishes first
  if (F.aux_pending) {
    HIP_TRY(c, hipStreamWaitEvent(c->stream, c->aux_done, 0));
    F.aux_pending = false;
  }
  DevGraph* g = nullptr;
  if (int rc = dev_frame_graph(v, form, w, h, rgba, brick_lod != nullptr, &g)) return rc;
  HIP_TRY(c, hipEventRecord(c->ev[0], c->stream));
  HIP_TRY(c, hipGraphLaunch(g->exec, c->stream));
  HIP_TRY(c, hipEventRecord(c->ev[1], c->stream));
  v->res_dev_newer = true;
  if (out->label) HIP_TRY(c, hipMemcpyAsync(out->label, v->d_flabel, npx * v->lbytes(), hipMemcpyDeviceToHost, c->stream));
  if (out->voxel) HIP_TRY(c, hipMemcpyAsync(out->voxel, v->d_fvoxel, npx * 8, hipMemcpyDeviceToHost, c->stream));
  if (out->t_hit) HIP_TRY(c, hipMemcpyAsync(out->t_hit, v->d_ft, npx * 4, hipMemcpyDeviceToHost, c->stream));
  if (out->rgba) HIP_TRY(c, hipMemcpyAsync(out->rgba, v->d_frgba, npx * 4, hipMemcpyDeviceToHost, c->stream));
  if (out->brick_visited)
    HIP_TRY(c, hipMemcpyAsync(out->brick_visited, v->d_visited, (size_t)v->nb * 4, hipMemcpyDeviceToHost, c->stream));
  std::vector<uint32_t> meta;
  if (out->brick_lod_used) {
    meta.resize(v->nb);
    HIP_TRY(c, hipMemcpyAsync(meta.data(), v->d_meta, (size_t)v->nb * 4, hipMemcpyDeviceToHost, c->stream));
  }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  for (size_t i = 0; i < meta.size(); ++i) out->brick_lod_used[i] = (uint8_t)(meta[i] & FCS_META_LOD_MASK);
  v->frame_px = npx;
  float ms = 0;
  HIP_TRY(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  out->bricks_decoded = F.h_small[DF_QCOUNT];
  out->decode_ms = 0.0f;
  out->march_ms = ms;
  c->timing.march_ms = ms;
  if (F.h_small[DF_ERR])
    return fail(c, FCS_E_CORRUPT, "a brick decoded for the frame failed (status bits in the request statuses)");
  const uint32_t left = form == FORM_PLAIN ? F.h_small[DF_MISS1] : (F.h_small[DF_MISS1] ? F.h_small[DF_MISS2] : 0u);
  if (left) return fail(c, FCS_E_CORRUPT, "rays entered bricks that could not be made resident");
  return 0;
}

// Prefetch decodes chosen on the device from the network's probabilities,
// issued on the aux stream after everything queued on the context stream;
// the next frame waits for them.
int dev_prefetch(fcs_volume* v, const float* d_probs, const fcs_camera* cam_ahead, uint32_t w, uint32_t h,
                 float threshold, uint32_t max_prefetch, uint32_t* out_prefetched, uint32_t* out_bricks) {
  fcs_ctx* c = v->ctx;
  DevFrames& F = *v->df;
  uint32_t* sm = F.d_small;
  const uint32_t cap = std::min<uint32_t>(max_prefetch, v->nb);
  memcpy(F.h_cam_ahead, cam_ahead, sizeof(MarchCam));
  HIP_TRY(c, hipEventRecord(c->main_mark, c->stream));
  HIP_TRY(c, hipStreamWaitEvent(c->aux, c->main_mark, 0));
  HIP_TRY(c, hipMemcpyAsync(F.d_cam_ahead, F.h_cam_ahead, sizeof(MarchCam), hipMemcpyHostToDevice, c->aux));
  HIP_TRY(c, fcs_launch_zero(sm + DF_AQCOUNT, 8, c->aux));
  PrefetchArgs pa{d_probs, (const MarchCam*)F.d_cam_ahead, v->d_res, F.d_aq, sm + DF_AQCOUNT, sm + DF_HIST,
                  threshold, cap, v->nb, v->b, v->L, w, h, v->stride};
  HIP_TRY(c, fcs_launch_prefetch_enqueue(pa, c->aux));
  uint32_t n = 0;
  HIP_TRY(c, hipMemcpyAsync(&n, sm + DF_AQCOUNT, 4, hipMemcpyDeviceToHost, c->aux));
  std::vector<DecodeReq> q;
  if (out_bricks && cap) {
    q.resize(cap);
    HIP_TRY(c, hipMemcpyAsync(q.data(), F.d_aq, cap * sizeof(DecodeReq), hipMemcpyDeviceToHost, c->aux));
  }
  HIP_TRY(c, hipStreamSynchronize(c->aux));
  n = std::min(n, cap);
  if (n) {
    const int split = ops_split_of(v);
    const int ksplit = split == FCS_OPS_SPLIT_4WG ? FCS_OPS_4WG : FCS_OPS_3WG;
    if (v->label_bits == 32) {
      OpsArgs<uint32_t> a{v->d_bricks, v->d_nib, (const uint32_t*)v->d_pal, F.d_aq, (uint32_t*)v->d_cache, F.d_aqst,
                          sm + DF_ACLAIM, n, v->nb, v->L, 0u, nullptr};
      HIP_TRY(c, fcs_launch_ops<uint32_t>(a, (uint32_t)c->num_cu, c->aux, nullptr, ksplit, FCS_OPS_HEAD_AUTO));
    } else {
      OpsArgs<uint64_t> a{v->d_bricks, v->d_nib, (const uint64_t*)v->d_pal, F.d_aq, (uint64_t*)v->d_cache, F.d_aqst,
                          sm + DF_ACLAIM, n, v->nb, v->L, 0u, nullptr};
      HIP_TRY(c, fcs_launch_ops<uint64_t>(a, (uint32_t)c->num_cu, c->aux, nullptr, ksplit, FCS_OPS_HEAD_AUTO));
    }
    QueueSettleArgs sa{F.d_aq, sm + DF_AQCOUNT, F.d_aqst, v->d_res, nullptr, sm + DF_AERR, n};
    HIP_TRY(c, fcs_launch_queue_settle(sa, c->aux));
    HIP_TRY(c, hipEventRecord(c->aux_done, c->aux));
    F.aux_pending = true;
    v->res_dev_newer = true;
  }
  if (out_prefetched) *out_prefetched = n;
  if (out_bricks)
    for (uint32_t i = 0; i < cap; ++i) out_bricks[i] = i < n ? q[i].brick : ~0u;
  return 0;
}

int decode_list(fcs_volume* v, const std::vector<uint32_t>& qb, const std::vector<uint8_t>& ql, fcs_frame* out) {
  if (qb.empty()) return 0;
  int rc = fcs_request_bricks(v, qb.data(), ql.data(), (uint32_t)qb.size(), nullptr, nullptr);
  if (rc && rc != FCS_E_CORRUPT) return rc;
  out->bricks_decoded += (uint32_t)qb.size();
  out->decode_ms += v->ctx->timing.decode_ms;
  return 0;
}

}  // namespace

extern "C" {

static int render_body(fcs_volume* v, const fcs_camera* cam, const uint8_t* brick_lod, uint32_t w, uint32_t h,
               fcs_frame* out) {
  if (!v || !v->ctx || !cam || !out || w == 0 || h == 0) return FCS_E_INVALID;
  fcs_ctx* c = v->ctx;
  HIP_TRY(c, hipSetDevice(c->device));
  if (v->direct) return dev_frame(v, cam, brick_lod, w, h, FORM_PLAIN, out);
  std::vector<uint8_t> lods;
  if (int rc = frame_lods(v, cam, brick_lod, w, h, lods)) return rc;
  // decodes issued ahead on the aux stream finish before this frame's work
  if (c->aux_pending) HIP_TRY(c, hipStreamWaitEvent(c->stream, c->aux_done, 0));
  // on-demand decode of every non-empty brick in the view frustum that is not
  // resident at its LOD (rays never enter a brick outside the frustum)
  const std::vector<uint8_t> vis = frustum_mask(v, cam, w, h);
  std::vector<uint32_t> qb;
  std::vector<uint8_t> ql;
  for (uint32_t i = 0; i < v->nb; ++i) {
    if (v->empty[i] || lods[i] == 0 || !vis[i] || resident_at(v, i, lods[i])) continue;
    qb.push_back(i);
    ql.push_back(lods[i]);
  }
  out->bricks_decoded = 0;
  out->decode_ms = 0.0f;
  out->march_ms = 0.0f;
  if (int rc = decode_list(v, qb, ql, out)) return rc;
  if (int rc = march_pass(v, cam, lods, w, h, out->rgba != nullptr)) return rc;
  return finish_frame(v, lods, w, h, out);
}

int fcs_render(fcs_volume* v, const fcs_camera* cam, const uint8_t* brick_lod, uint32_t w, uint32_t h,
               fcs_frame* out) {
  return fcs_guard(v ? v->ctx : nullptr, [&] { return render_body(v, cam, brick_lod, w, h, out); });
}

static int render_visible_body(fcs_volume* v, const fcs_camera* cam, const uint8_t* brick_lod, uint32_t w, uint32_t h,
                       fcs_frame* out) {
  if (!v || !v->ctx || !cam || !out || w == 0 || h == 0) return FCS_E_INVALID;
  fcs_ctx* c = v->ctx;
  HIP_TRY(c, hipSetDevice(c->device));
  if (v->direct) return dev_frame(v, cam, brick_lod, w, h, FORM_VISIBLE, out);
  std::vector<uint8_t> lods;
  if (int rc = frame_lods(v, cam, brick_lod, w, h, lods)) return rc;
  if (c->aux_pending) HIP_TRY(c, hipStreamWaitEvent(c->stream, c->aux_done, 0));
  out->bricks_decoded = 0;
  out->decode_ms = 0.0f;
  out->march_ms = 0.0f;
  std::vector<uint32_t> visited(v->nb);
  // march with what is resident; a ray that enters a non-resident brick is
  // counted as a miss and passes through it: decode the entered bricks that
  // were missing and march again until no ray misses. The last pass sees
  // every brick its rays enter, so the frame equals fcs_render's.
  for (int pass = 0;; ++pass) {
    if (int rc = march_pass(v, cam, lods, w, h, out->rgba != nullptr)) return rc;
    uint32_t misses = 0;
    HIP_TRY(c, hipMemcpyAsync(&misses, v->d_misses, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (misses == 0) break;
    float ms = 0;
    HIP_TRY(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    out->march_ms += ms;
    HIP_TRY(c, hipMemcpy(visited.data(), v->d_visited, (size_t)v->nb * 4, hipMemcpyDeviceToHost));
    std::vector<uint32_t> qb;
    std::vector<uint8_t> ql;
    for (uint32_t i = 0; i < v->nb; ++i) {
      if (!visited[i] || v->empty[i] || lods[i] == 0 || resident_at(v, i, lods[i])) continue;
      qb.push_back(i);
      ql.push_back(lods[i]);
    }
    if (qb.empty() || pass >= 64) return fail(c, FCS_E_CORRUPT, "misses without a decodable brick");
    if (int rc = decode_list(v, qb, ql, out)) return rc;
  }
  return finish_frame(v, lods, w, h, out);
}

int fcs_render_visible(fcs_volume* v, const fcs_camera* cam, const uint8_t* brick_lod, uint32_t w, uint32_t h,
                       fcs_frame* out) {
  return fcs_guard(v ? v->ctx : nullptr, [&] { return render_visible_body(v, cam, brick_lod, w, h, out); });
}

static int prefetch_bricks_body(fcs_volume* v, const uint32_t* brick, const uint8_t* lod, uint32_t n) {
  if (!v || !v->ctx || (n && (!brick || !lod))) return FCS_E_INVALID;
  fcs_ctx* c = v->ctx;
  HIP_TRY(c, hipSetDevice(c->device));
  if (!v->rans_ready) {
    int rc = fcs_decode_rans(v, nullptr);
    if (rc && rc != FCS_E_CORRUPT) return rc;
  }
  if (int rc = ensure_cache(v)) return rc;
  if (int rc = settle_prefetch(v)) return rc;  // one prefetch batch in flight
  if (int rc = pull_residency(v)) return rc;
  std::vector<DecodeReq> reqs;
  std::vector<Alloc> deferred;
  std::vector<uint8_t> seen(v->nb, 0);
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t bi = brick[i], t = lod[i];
    if (bi >= v->nb || t > v->L || t == 0 || seen[bi] || resident_at(v, bi, t)) continue;
    seen[bi] = 1;
    cache_release(v, bi, &deferred);
    uint64_t off;
    const uint64_t sz = 1ull << (3 * t);
    if (!cache_alloc(v, bi, sz, &off)) continue;  // best effort: a full pool skips the prefetch
    v->res_lod[bi] = (int16_t)t;
    v->res_pyr[bi] = 0;
    v->res_alloc[bi] = Alloc{off, sz};
    ++v->res_epoch;
    reqs.push_back(DecodeReq{bi, t, off});
  }
  if (reqs.empty()) {
    cache_free_deferred(v, deferred);
    return 0;
  }
  // the aux stream runs after everything already queued on the main stream
  HIP_TRY(c, hipEventRecord(c->main_mark, c->stream));
  HIP_TRY(c, hipStreamWaitEvent(c->aux, c->main_mark, 0));
  const int rc = launch_decode(v, v->lane_aux, reqs, c->aux, false);
  cache_free_deferred(v, deferred);
  if (rc) return rc;
  HIP_TRY(c, hipEventRecord(c->aux_done, c->aux));
  c->aux_pending = true;
  return (int)reqs.size();
}

int fcs_prefetch_bricks(fcs_volume* v, const uint32_t* brick, const uint8_t* lod, uint32_t n) {
  return fcs_guard(v ? v->ctx : nullptr, [&] { return prefetch_bricks_body(v, brick, lod, n); });
}

int fcs_frame_checksum(fcs_volume* v, uint64_t* out) {
  if (!v || !v->ctx || !out) return FCS_E_INVALID;
  fcs_ctx* c = v->ctx;
  if (!v->frame_px || !v->d_flabel) return fail(c, FCS_E_INVALID, "no frame rendered yet");
  HIP_TRY(c, hipSetDevice(c->device));
  if (!v->d_csum) HIP_TRY(c, hipMalloc((void**)&v->d_csum, 16));
  HIP_TRY(c, fcs_launch_zero(reinterpret_cast<uint32_t*>(v->d_csum), 2, c->stream));
  if (v->label_bits == 32)
    HIP_TRY(c, fcs_launch_frame_checksum<uint32_t>((const uint32_t*)v->d_flabel, v->d_fvoxel, v->frame_px, v->d_csum,
                                                   c->stream));
  else
    HIP_TRY(c, fcs_launch_frame_checksum<uint64_t>((const uint64_t*)v->d_flabel, v->d_fvoxel, v->frame_px, v->d_csum,
                                                   c->stream));
  unsigned long long h = 0;
  HIP_TRY(c, hipMemcpyAsync(&h, v->d_csum, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  *out = h;
  return 0;
}

int fcs_camera_features(const fcs_volume* v, const fcs_camera* cam, float* out12) {
  if (!v || !v->ctx || !cam || !out12) return FCS_E_INVALID;
  const float g = (float)(v->B * v->b);
  for (int k = 0; k < 3; ++k) {
    out12[k] = cam->position[k] / g;
    out12[3 + k] = cam->forward[k];
    out12[6 + k] = cam->up[k];
    out12[9 + k] = cam->right[k];
  }
  return 0;
}

static int render_prefetch_body(fcs_volume* v, fcs_mlp* m, const fcs_camera* cam, const fcs_camera* cam_ahead, uint32_t w,
                        uint32_t h, float threshold, uint32_t max_prefetch, fcs_frame* out, uint32_t* out_prefetched,
                        uint32_t* out_bricks) {
  if (!v || !v->ctx || !m || !cam || !out) return FCS_E_INVALID;
  if (fcs_mlp_inputs(m) == 0) return fail(v->ctx, FCS_E_INVALID, "prefetch network was released");
  if (out_prefetched) *out_prefetched = 0;
  fcs_ctx* c = v->ctx;
  if (fcs_mlp_outputs(m) != v->nb || fcs_mlp_inputs(m) != 12 + v->nb)
    return fail(c, FCS_E_INVALID, "prefetch network must have 12 + bricks inputs and bricks outputs");
  if (int rc = fcs_render_visible(v, cam, nullptr, w, h, out)) return rc;
  float feat[12];
  fcs_camera_features(v, cam, feat);
  const float* d_probs = nullptr;
  if (int rc = fcs_mlp_predict_dev(m, feat, 12, v->d_visited, &d_probs)) return rc;
  if (v->direct)
    return dev_prefetch(v, d_probs, cam_ahead ? cam_ahead : cam, w, h, threshold, max_prefetch, out_prefetched,
                        out_bricks);
  std::vector<float> probs(v->nb);
  HIP_TRY(c, hipMemcpyAsync(probs.data(), d_probs, (size_t)v->nb * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  const LodRule rule(v, cam_ahead ? cam_ahead : cam, w, h);
  std::vector<uint32_t> cand;
  std::vector<uint8_t> lod_of(v->nb, 0);
  for (uint32_t i = 0; i < v->nb; ++i) {
    if (!(probs[i] >= threshold) || v->empty[i]) continue;
    const uint8_t t = rule(i);
    if (t == 0 || resident_at(v, i, t)) continue;
    lod_of[i] = t;
    cand.push_back(i);
  }
  std::stable_sort(cand.begin(), cand.end(), [&](uint32_t a, uint32_t b) { return probs[a] > probs[b]; });
  if (cand.size() > max_prefetch) cand.resize(max_prefetch);
  std::vector<uint8_t> cl(cand.size());
  for (size_t i = 0; i < cand.size(); ++i) cl[i] = lod_of[cand[i]];
  const int n = fcs_prefetch_bricks(v, cand.data(), cl.data(), (uint32_t)cand.size());
  if (n < 0) return n;
  if (out_prefetched) *out_prefetched = (uint32_t)n;
  // candidates were neither resident nor repeated, so each was issued unless the pool was full
  if (out_bricks)
    for (size_t i = 0; i < cand.size(); ++i) out_bricks[i] = resident_at(v, cand[i], cl[i]) ? cand[i] : ~0u;
  return 0;
}

int fcs_render_prefetch(fcs_volume* v, fcs_mlp* m, const fcs_camera* cam, const fcs_camera* cam_ahead, uint32_t w,
                        uint32_t h, float threshold, uint32_t max_prefetch, fcs_frame* out, uint32_t* out_prefetched,
                        uint32_t* out_bricks) {
  return fcs_guard(v ? v->ctx : nullptr, [&] { return render_prefetch_body(v, m, cam, cam_ahead, w, h, threshold, max_prefetch, out, out_prefetched, out_bricks); });
}

}  // extern "C"
