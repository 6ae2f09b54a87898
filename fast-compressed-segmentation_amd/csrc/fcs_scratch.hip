// Fixed device scratch of the allocation-free request entry point
// (fcs_request_bricks_static): the op decoder's claim counter, request list
// and status words live in one file-scope device array instead of buffers the
// library allocates. Every GPU holds its own copy of the array for the life of
// the process; its address on each device is looked up once, at context
// creation with that device current, and only that address is handed to
// kernels and copies (the array's name in host code is not a device address).
#include <hip/hip_runtime.h>

#include <mutex>

#include "fcs_device.h"
#include "fcs_internal.h"

namespace {

constexpr size_t kCounterBytes = 256;
constexpr size_t kReqBytes = (size_t)FCS_STATIC_MAX_REQUESTS * sizeof(DecodeReq);
constexpr size_t kStatusBytes = (size_t)FCS_STATIC_MAX_REQUESTS * sizeof(uint32_t);
constexpr size_t kScratchBytes = kCounterBytes + kReqBytes + kStatusBytes;

__device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_addr[kMaxDevices] = {};

}  // namespace

int fcs_scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return FCS_E_INVALID;
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_addr[device]) return 0;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return FCS_E_HIP;
  g_addr[device] = p;
  return 0;
}

int fcs_scratch_lane(int device, uint32_t** counter, DecodeReq** req, uint32_t** status) {
  if (device < 0 || device >= kMaxDevices) return FCS_E_INVALID;
  std::lock_guard<std::mutex> lock(g_mu);
  uint8_t* p = static_cast<uint8_t*>(g_addr[device]);
  if (!p) return FCS_E_INVALID;
  *counter = reinterpret_cast<uint32_t*>(p);
  *req = reinterpret_cast<DecodeReq*>(p + kCounterBytes);
  *status = reinterpret_cast<uint32_t*>(p + kCounterBytes + kReqBytes);
  return 0;
}

size_t fcs_scratch_bytes() { return kScratchBytes; }
