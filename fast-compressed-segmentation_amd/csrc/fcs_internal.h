// Host-side internals shared by the runtime's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <exception>
#include <new>
#include <string>
#include <vector>

#include "fcs/fcs.h"

struct fcs_volume;
struct fcs_mlp;
struct DecodeReq;

// Fixed device scratch (fcs_scratch.hip) of fcs_request_bricks_static: its
// address on `device` is looked up once (fcs_create, device current);
// fcs_scratch_lane hands out the claim counter, request list and status words
// of FCS_STATIC_MAX_REQUESTS requests inside it.
int fcs_scratch_setup(int device);
int fcs_scratch_lane(int device, uint32_t** counter, DecodeReq** req, uint32_t** status);
size_t fcs_scratch_bytes();

// One HIP device and its streams (fcs_create). `stream` runs requests and
// frames; `aux` runs work issued ahead of them (prefetch decodes). The context
// tracks the volumes and networks created on it: fcs_destroy releases their
// device memory and detaches them, so freeing them afterwards is safe.
struct fcs_ctx {
  int device = 0;
  int num_cu = 256;
  hipStream_t stream = nullptr;
  hipStream_t aux = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t aux_done = nullptr;   // recorded on `aux` after the last prefetch launch
  hipEvent_t main_mark = nullptr;  // recorded on `stream` before a prefetch launch
  bool aux_pending = false;
  std::string err;
  fcs_timing timing{};
  int rans_mapping = FCS_RANS_MAPPING_AUTO;
  int ops_split = FCS_OPS_SPLIT_AUTO;
  int ops_head = FCS_OPS_HEAD_AUTO;
  bool march_stats = false;
  std::vector<fcs_volume*> volumes;
  std::vector<fcs_mlp*> mlps;
};

// Release a volume's / network's device memory and detach it from its
// context (called by fcs_destroy for the ones still alive).
void fcs_volume_release_device(fcs_volume* v);
void fcs_mlp_release_device(fcs_mlp* m);

inline int fcs_fail(fcs_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

// Runs an ABI body so that no C++ exception crosses the C boundary: a failed
// host allocation becomes FCS_E_NOMEM, anything else FCS_E_INVALID.
template <class F>
int fcs_guard(fcs_ctx* c, F&& body) {
  try {
    return body();
  } catch (const std::bad_alloc&) {
    return fcs_fail(c, FCS_E_NOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return fcs_fail(c, FCS_E_INVALID, std::string("internal error: ") + e.what());
  } catch (...) {
    return fcs_fail(c, FCS_E_INVALID, "internal error");
  }
}

#define FCS_HIP_TRY(ctx, expr)                                                            \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fcs_fail((ctx), FCS_E_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

// Learned prefetch network (fcs_mlp.hip): probabilities of one sample, from a
// pose and a device mask of one u32 per input position, in a device buffer
// owned by the network (valid until its next call).
struct fcs_mlp;
int fcs_mlp_predict_dev(fcs_mlp* m, const float* pose, uint32_t n_pose, const uint32_t* d_mask, const float** d_probs);
uint32_t fcs_mlp_outputs(const fcs_mlp* m);
uint32_t fcs_mlp_inputs(const fcs_mlp* m);

// Device-side description of a volume for frames launched outside the
// volume's own calls (views). `serial` identifies the volume for the lifetime
// of the process (addresses can be reused).
struct fcs_volume_desc {
  uint64_t serial;
  const uint64_t* res;  // per brick residency word (FCS_RES_* in fcs_kernels.h)
  const void* pal0;     // per brick palette[0]
  const void* cache;    // brick cache pool (may be NULL: nothing resident)
  uint32_t nb, b, L, B, label_bits;
  // shading (filled when requested): transfer function table and lighting;
  // shade_epoch changes whenever a value baked into a launch changes
  const uint64_t* tf_keys;
  const uint32_t* tf_rgba;
  uint32_t tf_mask;
  float light[3], ambient, diffuse;
  uint64_t shade_epoch;
};
// Bring the volume's device residency table (and, with `shading`, its
// transfer function table) up to date and describe the volume. When the
// residency changed, the device is synchronised first because frames in
// flight may read the old table; prefetches are settled first.
int fcs_volume_publish(fcs_volume* v, fcs_volume_desc* out, bool shading = false);
fcs_ctx* fcs_volume_context(const fcs_volume* v);
