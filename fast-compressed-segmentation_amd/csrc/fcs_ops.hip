// Op decoder for gfx950: decodes (brick, LOD) requests from the device
// nibble streams into the brick cache. See the comment block below.
#include <hip/hip_runtime.h>

#include "fcs_common.h"
#include "fcs_device.h"
#include "fcs_kernels.h"

namespace {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// ============================================================ op decoding
// One workgroup (16 waves) per request, persistent over a request queue.
// Level l -> l+1 for the 8^(l+1) child nodes, all in LDS as 16-bit palette
// indices (bit 15 = reference to an earlier node of the same level):
//   1. active parents (no stop bit) compacted in Morton order  -> ACT
//   2. children of stopped parents copy the parent; stop bits init
//   3. the level's nibble span is parsed 8 nibbles per lane: a nibble is a
//      delta iff the run of 12/13-nibbles before it (back to the level start)
//      has odd length; op index / palette cursor by block prefix scan
//   4. each op resolves to a palette index or a node reference
//   5. references are chased (always to a smaller node index)
//   6. labels (palette[idx]) stored coalesced to the brick cache
// This restates decode.wgsl:404-508 / brickDecoding.hpp:174-242 level by
// level: the value of node n at level l+1 only depends on level l and on
// nodes < n of level l+1, so each level is one parallel step.

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const T y = __shfl_up(v, d, 64);
    if (lane >= d) v += y;
  }
  return v;
}

// exclusive block scan over FCS_OPS_THREADS lanes; sh needs FCS_OPS_WAVES+1 words
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* sh, uint32_t& total) {
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  const uint32_t inc = wave_incl_scan(v);
  if (lane == 63) sh[wave] = inc;
  __syncthreads();
  if (threadIdx.x < 64) {
    const uint32_t w = threadIdx.x < FCS_OPS_WAVES ? sh[threadIdx.x] : 0u;
    const uint32_t wi = wave_incl_scan(w);
    if (threadIdx.x < FCS_OPS_WAVES) sh[threadIdx.x] = wi - w;
    if (threadIdx.x == FCS_OPS_WAVES - 1) sh[FCS_OPS_WAVES] = wi;
  }
  __syncthreads();
  const uint32_t r = sh[wave] + inc - v;
  total = sh[FCS_OPS_WAVES];
  __syncthreads();
  return r;
}

__device__ __forceinline__ uint32_t nib_word(const uint32_t* nibw, uint32_t wi) { return nibw[wi]; }

struct OpsShared {
  uint16_t* bigv;   // FCS_LDS_LEVEL_MAX
  uint16_t* smv;    // 4096
  uint16_t* act;    // act_cap
  uint32_t* sa;     // stop bits A (1024 words)
  uint32_t* sb;     // stop bits B (1024 words)
  uint32_t* scr;    // scan scratch (64 words)
};

template <typename LT>
__device__ void fill_const(LT* o, uint64_t count, LT v) {
  for (uint64_t i = threadIdx.x; i < count; i += FCS_OPS_THREADS) o[i] = v;
}

template <typename LT>
__device__ void decode_request(const OpsArgs<LT>& a, const OpsShared& S, uint32_t r) {
  const uint32_t tid = threadIdx.x;
  const DecodeReq q = a.reqs[r];
  const uint32_t b = q.brick, t = q.lod & 0xFFu;
  const bool pyr = (q.lod & FCS_REQ_PYRAMID) != 0;
  const uint32_t L = a.L;
  if (b >= a.nbricks || t > L) {
    if (tid == 0) a.status[r] = FCS_ST_BAD_REQUEST;
    return;
  }
  const DevBrick B = a.bricks[b];
  uint32_t status = 0;
  const uint32_t nsym = B.nsym;
  if (nsym == 0) status |= FCS_ST_STREAM_OVERRUN;
  const uint32_t P = B.P;
  const LT* pal = a.palette + B.pal_off;
  LT* o = a.out + q.out_off;
  if (P == 0 || P > FCS_PALETTE_MAX_LDS) {
    if (tid == 0) a.status[r] = status | (P == 0 ? FCS_ST_PALETTE_RANGE : FCS_ST_BAD_REQUEST);
    return;
  }
  const uint32_t* nibw = a.nib + B.nib_off;
  const uint32_t root = nibw[0] & 0xFu;
  const LT p0 = pal[0];
  if ((root & 1u) || t == 0 || (status & FCS_ST_STREAM_OVERRUN)) {
    // constant brick (root stop bit) or LOD 0: every voxel = palette[0]
    fill_const(o, pyr ? fcs_pyramid_level_offset(t + 1) : (1ull << (3 * t)), p0);
    if (tid == 0) a.status[r] = status;
    return;
  }
  if (pyr && tid == 0) o[0] = p0;
  uint16_t* vcur = S.smv;
  uint32_t* scur = S.sa;
  uint32_t* snext = S.sb;
  if (tid == 0) { vcur[0] = 0; scur[0] = 0; S.scr[60] = 0; }
  __syncthreads();
  // level l's ops start at nibble ns with palette cursor ip0 (nibble 0 is the
  // root op); each level's end is found while parsing it
  uint32_t ns = 1, ip0 = 0;
  const uint32_t w_stream_end = (nsym + 7) >> 3;

  for (uint32_t l = 0; l < t; ++l) {
    const uint32_t N = 1u << (3 * l), C = N << 3;
    const bool last = (l + 1 == t);
    const bool gnext = last && C > FCS_LDS_LEVEL_MAX;  // final level lives in the cache region
    LT* olev = o + (pyr ? fcs_pyramid_level_offset(l + 1) : 0ull);
    uint16_t* vnext = S.bigv;

    // 1. active-parent compaction (32 parents per stop word)
    const uint32_t nw = (N + 31) >> 5;
    uint32_t abits = 0;
    if (tid < nw) {
      const uint32_t valid = (N >= 32) ? 0xFFFFFFFFu : ((1u << N) - 1u);
      abits = ~scur[tid] & valid;
    }
    uint32_t A;
    const uint32_t abase = block_excl_scan(__popc(abits), S.scr, A);
    {
      uint32_t k = abase, bits = abits;
      while (bits) {
        const uint32_t j = __ffs(bits) - 1;
        bits &= bits - 1;
        S.act[k++] = (uint16_t)(tid * 32 + j);
      }
    }
    // 2. children of stopped parents; stop bits of the next level
    for (uint32_t n = tid; n < C; n += FCS_OPS_THREADS) {
      const uint32_t m = n >> 3;
      if ((scur[m >> 5] >> (m & 31)) & 1u) {
        const uint16_t v = vcur[m];
        if (gnext) olev[n] = (LT)v; else vnext[n] = v;
      }
    }
    if (!last) {
      for (uint32_t w = tid; w < ((C + 31) >> 5); w += FCS_OPS_THREADS) {
        const uint32_t pb = (scur[w >> 3] >> ((w & 7) * 4)) & 0xFu;  // 4 parents
        uint32_t m = 0;
        m |= (pb & 1u) ? 0x000000FFu : 0u;
        m |= (pb & 2u) ? 0x0000FF00u : 0u;
        m |= (pb & 4u) ? 0x00FF0000u : 0u;
        m |= (pb & 8u) ? 0xFF000000u : 0u;
        snext[w] = m;
      }
    }
    if (tid == 0) { S.scr[61] = 0; S.scr[62] = ns; S.scr[63] = ip0; }
    __syncthreads();

    // 3-4. parse this level's ops and resolve them to values / references
    const uint32_t K = 8 * A;
    uint32_t op_base = 0, pa_base = 0, carry = 0;  // carry: next nibble is a delta
    const uint32_t w_first = ns >> 3, w_end = w_stream_end;
    const uint32_t ne = nsym;  // parse bound; the level ends after its K-th op
    const uint32_t side = 1u << (l + 1);
    for (uint32_t wt = w_first; wt < w_end && op_base < K; wt += FCS_OPS_THREADS) {
      const uint32_t wi = wt + tid;
      const bool inw = wi < w_end;
      const uint32_t word = inw ? nib_word(nibw, wi) : 0u;
      const uint32_t g0 = wi * 8;
      // trailing run of 12/13 nibbles over the valid nibbles of this word
      uint32_t tr = 0;
      for (uint32_t j = 0; j < 8; ++j) {
        const uint32_t g = g0 + j;
        const uint32_t v = (word >> (4 * j)) & 0xFu;
        const bool is12 = inw && g >= ns && g < ne && (v >> 1) == FCS_OP_PD;
        tr = is12 ? tr + 1 : 0;
      }
      const bool full = (tr == 8);
      const uint64_t Pm = __ballot(tr & 1u);
      const uint64_t Nm = __ballot(!full);
      const uint32_t lane = lane_id(), wave = tid >> 6;
      // wave exit parity record
      if (lane == 0) {
        const uint32_t has = Nm != 0;
        const uint32_t hi = has ? 63 - __clzll(Nm) : 0;
        S.scr[32 + wave] = has ? (2u | ((uint32_t)(Pm >> hi) & 1u)) : 0u;
      }
      __syncthreads();
      uint32_t par;  // parity entering this lane
      {
        const uint64_t below = Nm & ((lane == 0) ? 0ull : (~0ull >> (64 - lane)));
        if (below) {
          par = (uint32_t)(Pm >> (63 - __clzll(below))) & 1u;
        } else {
          par = carry;
          for (int w = (int)wave - 1; w >= 0; --w) {
            const uint32_t rec = S.scr[32 + w];
            if (rec & 2u) { par = rec & 1u; break; }
          }
        }
      }
      uint32_t tile_exit = carry;
      for (int w = FCS_OPS_WAVES - 1; w >= 0; --w) {
        const uint32_t rec = S.scr[32 + w];
        if (rec & 2u) { tile_exit = rec & 1u; break; }
      }
      // count ops / palette advances of this lane
      uint32_t nops = 0, npa = 0;
      {
        uint32_t pd = par;
        for (uint32_t j = 0; j < 8; ++j) {
          const uint32_t g = g0 + j;
          if (!inw || g >= ne) break;
          if (g < ns) continue;
          if (g == ns) pd = 0;
          const uint32_t v = (word >> (4 * j)) & 0xFu;
          if (pd) { pd = 0; continue; }
          ++nops;
          npa += ((v >> 1) == FCS_OP_PA);
          pd = ((v >> 1) == FCS_OP_PD);
        }
      }
      uint32_t tot;
      const uint32_t ex = block_excl_scan(nops | (npa << 16), S.scr, tot);
      // process ops of this lane
      {
        uint32_t k = op_base + (ex & 0xFFFFu);
        uint32_t ipx = ip0 + pa_base + (ex >> 16);
        uint32_t pd = par;
        for (uint32_t j = 0; j < 8; ++j) {
          const uint32_t g = g0 + j;
          if (!inw || g >= ne) break;
          if (g < ns) continue;
          if (g == ns) pd = 0;
          const uint32_t v = (word >> (4 * j)) & 0xFu;
          if (pd) { pd = 0; continue; }
          const uint32_t op = v >> 1, stop = v & 1u;
          pd = (op == FCS_OP_PD);
          if (k >= K) break;  // the rest of the word belongs to the next level
          const uint32_t m = S.act[k >> 3], c = k & 7u, n = 8 * m + c;
          uint32_t val;  // palette index, or FCS_REF32 | node
          switch (op) {
            case FCS_OP_RP: val = vcur[m]; break;
            case FCS_OP_RX: case FCS_OP_RY: case FCS_OP_RZ: {
              uint32_t x = fcs_unmorton_x(n), y = fcs_unmorton_y(n), z = fcs_unmorton_z(n);
              uint32_t& qa = (op == FCS_OP_RX) ? x : ((op == FCS_OP_RY) ? y : z);
              const bool odd = qa & 1u;
              if ((!odd && qa == 0) || (odd && qa + 1 >= side)) {
                // outside the brick: the reference reads the parent's output
                // slot, which child 0 has already overwritten when c > 0
                val = (c == 0) ? (uint32_t)vcur[m] : (FCS_REF32 | (8 * m));
              } else {
                qa = odd ? qa + 1 : qa - 1;
                const uint32_t nb = fcs_morton3(x, y, z);
                val = (nb < n) ? (FCS_REF32 | nb) : (uint32_t)vcur[nb >> 3];
              }
              break;
            }
            case FCS_OP_PA:
              ++ipx;
              val = ipx;
              if (ipx >= P) { status |= FCS_ST_PALETTE_RANGE; val = P - 1; }
              break;
            case FCS_OP_P0:
              val = ipx;
              if (ipx >= P) { status |= FCS_ST_PALETTE_RANGE; val = P - 1; }
              break;
            case FCS_OP_PD: {
              if (g + 1 >= nsym) status |= FCS_ST_STREAM_OVERRUN;
              const uint32_t dn = (j < 7) ? ((word >> (4 * (j + 1))) & 0xFu) : (nib_word(nibw, wi + 1) & 0xFu);
              const int64_t idx = (int64_t)ipx - (int64_t)dn - 1;
              if (idx < 0 || idx >= (int64_t)P) { status |= FCS_ST_PALETTE_RANGE; val = P - 1; }
              else val = (uint32_t)idx;
              break;
            }
            default:
              status |= FCS_ST_BAD_OP;
              val = 0;
          }
          if (gnext) {
            olev[n] = (LT)val;
          } else {
            vnext[n] = (val & FCS_REF32) ? (uint16_t)(FCS_REF16 | (val & 0x7FFFu)) : (uint16_t)val;
            if (stop && !last) atomicOr(&snext[n >> 5], 1u << (n & 31));
          }
          if (k == K - 1) {
            // last op of the level: the next level starts after it (and its delta)
            S.scr[62] = g + 1 + (op == FCS_OP_PD ? 1u : 0u);
            S.scr[63] = ipx;
          }
          ++k;
        }
      }
      op_base += tot & 0xFFFFu;
      pa_base += tot >> 16;
      carry = tile_exit;
      __syncthreads();
    }
    if (op_base < K) status |= FCS_ST_STREAM_OVERRUN | FCS_ST_NSYM_MISMATCH;
    __syncthreads();
    ns = S.scr[62];
    ip0 = S.scr[63];
    if (l + 1 == L && ns != nsym) status |= FCS_ST_NSYM_MISMATCH;  // nibbles after the finest level

    // 5. chase references (targets are always smaller node indices)
    // A valid reference always points to a smaller node, so every chain
    // strictly decreases; anything else is corrupt input and stops the chase.
    if (gnext) {
      for (uint32_t n = tid; n < C; n += FCS_OPS_THREADS) {
        uint32_t v = (uint32_t)olev[n];
        if (v & FCS_REF32) {
          uint32_t cur = n;
          do {
            const uint32_t nx = v & ~FCS_REF32;
            if (nx >= cur) { status |= FCS_ST_NSYM_MISMATCH; v = 0; break; }
            cur = nx;
            v = (uint32_t)olev[nx];
          } while (v & FCS_REF32);
          olev[n] = (LT)v;
        }
      }
    } else {
      for (uint32_t n = tid; n < C; n += FCS_OPS_THREADS) {
        uint32_t v = vnext[n];
        if (v & FCS_REF16) {
          uint32_t cur = n;
          do {
            const uint32_t nx = v & 0x7FFFu;
            if (nx >= cur) { status |= FCS_ST_NSYM_MISMATCH; v = 0; break; }
            cur = nx;
            v = vnext[nx];
          } while (v & FCS_REF16);
          vnext[n] = (uint16_t)v;
        }
      }
    }
    __syncthreads();

    // 6. labels out
    if (gnext) {
      for (uint32_t n = tid; n < C; n += FCS_OPS_THREADS) {
        uint32_t v = (uint32_t)olev[n];
        olev[n] = pal[v < P ? v : P - 1];
      }
    } else if (last || pyr) {
      for (uint32_t n = tid; n < C; n += FCS_OPS_THREADS) {
        const uint32_t v = vnext[n];
        olev[n] = pal[v < P ? v : P - 1];
      }
    }
    if (!last) {
      // next level's parents
      if (C <= 4096) {
        for (uint32_t n = tid; n < C; n += FCS_OPS_THREADS) S.smv[n] = vnext[n];
        vcur = S.smv;
      } else {
        vcur = S.bigv;  // only feeds a final level held in the cache region
      }
      uint32_t* tmp = scur;
      scur = snext;
      snext = tmp;
    }
    __syncthreads();
  }
  if (status) atomicOr(&S.scr[60], status);
  __syncthreads();
  if (tid == 0) a.status[r] = S.scr[60];
}

template <typename LT>
__global__ __launch_bounds__(FCS_OPS_THREADS) void ops_kernel(OpsArgs<LT> a) {
  extern __shared__ __align__(16) uint8_t smem[];
  OpsShared S;
  S.bigv = reinterpret_cast<uint16_t*>(smem);
  S.smv = S.bigv + FCS_LDS_LEVEL_MAX;
  S.sa = reinterpret_cast<uint32_t*>(S.smv + 4096);
  S.sb = S.sa + 1024;
  S.scr = S.sb + 1024;
  S.act = reinterpret_cast<uint16_t*>(S.scr + 64);
  __shared__ uint32_t s_req;
  for (;;) {
    if (threadIdx.x == 0) s_req = atomicAdd(a.counter, 1u);
    __syncthreads();
    const uint32_t r = s_req;
    __syncthreads();
    if (r >= a.nreq) break;
    decode_request<LT>(a, S, r);
    __syncthreads();
  }
}

__global__ void zero_u32_kernel(uint32_t* p, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

}  // namespace

size_t fcs_ops_lds_bytes(uint32_t L) {
  const size_t act_cap = (L >= 6) ? 32768 : 4096;
  return FCS_LDS_LEVEL_MAX * 2 + 4096 * 2 + 1024 * 4 * 2 + 64 * 4 + act_cap * 2;
}

template <typename LT>
hipError_t fcs_launch_ops(const OpsArgs<LT>& a, uint32_t grid, hipStream_t s) {
  if (a.nreq == 0) return hipSuccess;
  hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, a.counter, (uint64_t)1);
  const size_t lds = fcs_ops_lds_bytes(a.L);
  static bool attr_set = false;
  if (!attr_set) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&ops_kernel<LT>),
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)fcs_ops_lds_bytes(6));
    if (e != hipSuccess) return e;
    attr_set = true;
  }
  hipLaunchKernelGGL(ops_kernel<LT>, dim3(grid), dim3(FCS_OPS_THREADS), lds, s, a);
  return hipGetLastError();
}

hipError_t fcs_launch_zero(uint32_t* p, uint64_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(zero_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

template hipError_t fcs_launch_ops<uint32_t>(const OpsArgs<uint32_t>&, uint32_t, hipStream_t);
template hipError_t fcs_launch_ops<uint64_t>(const OpsArgs<uint64_t>&, uint32_t, hipStream_t);
