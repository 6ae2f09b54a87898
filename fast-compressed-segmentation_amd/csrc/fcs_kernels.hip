// HIP kernels for gfx950 (MI355X): rANS stage, level-parallel op decoder,
// and the brick/voxel DDA marcher.
//
// Build flags (see the package build): -O3 --offload-arch=gfx950
// -ffp-contract=off (FP32 ray setup must round like the host restatement)
// -Xclang -target-feature -Xclang -packed-fp32-ops.
#include <hip/hip_runtime.h>

#include "fcs_common.h"
#include "fcs_device.h"
#include "fcs_kernels.h"

namespace {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// ===================================================================== rANS
// One lane per brick (bricks-per-wave): restates decodeRansStream
// (brickDecoding.cpp:126-188) and, in the same pass, tracks the op structure
// so that every op level's nibble span and palette cursor are known:
// level l has 8 * (#level-(l-1) ops without stop bit) ops, and the level
// boundaries follow from the op nibbles alone (a Pd op owns the next nibble).
//
// Decode table in LDS: x -> (symbol << 25) | (freq << 12) | cum.
// Compressed bytes are read backwards through a 16-byte block window with
// the next block prefetched one block ahead.

struct BackReader {
  const uint4* base;  // stream start (16-byte aligned)
  int64_t q;          // block index of `cur`
  uint4 cur, nxt;
  __device__ __forceinline__ void init(const uint8_t* s, uint32_t E) {
    base = reinterpret_cast<const uint4*>(s);
    q = (int64_t)((E - 1) >> 4);
    cur = base[q];
    if (q > 0) nxt = base[q - 1];
  }
  // byte at index p (p strictly decreasing by one per call)
  __device__ __forceinline__ uint32_t get(uint32_t p) {
    const int64_t pq = (int64_t)(p >> 4);
    if (pq != q) {
      cur = nxt;
      q = pq;
      if (q > 0) nxt = base[q - 1];
    }
    const uint32_t k = p & 15u;
    const uint32_t w = (k < 8) ? ((k < 4) ? cur.x : cur.y) : ((k < 12) ? cur.z : cur.w);
    return (w >> ((k & 3u) * 8u)) & 0xFFu;
  }
};

__global__ __launch_bounds__(256) void rans_kernel(RansArgs a) {
  __shared__ uint32_t table[FCS_RANS_M];
  for (uint32_t x = threadIdx.x; x < FCS_RANS_M; x += blockDim.x) {
    uint32_t s = 0;
    while (s + 1 < 16 && x >= a.cum[s + 1]) ++s;
    table[x] = (s << 25) | (a.freq[s] << 12) | a.cum[s];
  }
  __syncthreads();
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.nlist) return;
  const uint32_t b = a.list ? a.list[i] : i;
  const DevBrick B = a.bricks[b];
  const uint32_t L = a.L;
  LevelEntry* lv = a.lvl + (size_t)b * (L + 1);
  uint32_t* nibw = a.nib + B.nib_off;
  uint32_t status = 0;
  if (B.E < 4 || B.nsym == 0) {
    for (uint32_t l = 0; l <= L; ++l) lv[l] = LevelEntry{1u, 0u};
    a.res[b] = RansResult{FCS_ST_RANS_STATE | FCS_ST_STREAM_OVERRUN, 0u};
    return;
  }
  BackReader rd;
  rd.init(a.stream + B.str_off, B.E);
  uint32_t pos = B.E;
  uint32_t state = rd.get(--pos) << 24;
  state |= rd.get(--pos) << 16;
  state |= rd.get(--pos) << 8;
  state |= rd.get(--pos);

  // op-structure automaton
  uint32_t level = 0, K = 0, ops = 0, nonstop = 0, pa = 0;
  bool pend = false, done = false;
  uint32_t levels_ok = 0;
  // nibble packing: 8 per word, 4 words per 16-byte store
  uint32_t w0 = 0, w1 = 0, w2 = 0, w3 = 0, cw = 0;
  const uint32_t nsym = B.nsym;
  for (uint32_t g = 0; g < nsym; ++g) {
    const uint32_t x = state & (FCS_RANS_M - 1);
    const uint32_t e = table[x];
    const uint32_t s = e >> 25;
    state = ((e >> 12) & 0x1FFFu) * (state >> FCS_RANS_LOGM) + (x - (e & 0xFFFu));
    while (state < FCS_RANS_L && pos > 0) state = (state << 8) | rd.get(--pos);

    cw |= s << ((g & 7u) * 4u);
    if ((g & 7u) == 7u) {
      const uint32_t wi = (g >> 3) & 3u;
      if (wi == 0) w0 = cw; else if (wi == 1) w1 = cw; else if (wi == 2) w2 = cw; else w3 = cw;
      cw = 0;
      if (wi == 3) reinterpret_cast<uint4*>(nibw)[g >> 5] = make_uint4(w0, w1, w2, w3);
    }

    if (g == 0) {
      lv[0] = LevelEntry{1u, 0u};
      if (s & 1u) {  // root stop bit: constant brick, no op levels
        for (uint32_t l = 1; l <= L; ++l) lv[l] = LevelEntry{1u, 0u};
        done = true;
        levels_ok = L;
      } else {
        K = 8;
        if (L == 0) { done = true; levels_ok = 0; }
      }
    } else if (!done) {
      bool complete;
      if (pend) {
        pend = false;
        complete = true;
      } else {
        const uint32_t op = s >> 1;
        ++ops;
        pa += (op == FCS_OP_PA);
        nonstop += ((s & 1u) == 0u);
        if (op == 7) status |= FCS_ST_BAD_OP;
        pend = (op == FCS_OP_PD);
        complete = !pend;
      }
      if (complete && ops == K) {
        ++level;
        lv[level] = LevelEntry{g + 1, pa};
        K = 8 * nonstop;
        nonstop = 0;
        ops = 0;
        if (level == L) {
          done = true;
          levels_ok = L;
        } else if (K == 0) {
          for (uint32_t l = level + 1; l <= L; ++l) lv[l] = LevelEntry{g + 1, pa};
          done = true;
          levels_ok = L;
        }
      }
    } else {
      status |= FCS_ST_NSYM_MISMATCH;  // nibbles beyond the op structure
    }
  }
  // flush the partial 16-byte group
  if (nsym & 31u) {
    const uint32_t g = nsym;  // next nibble index
    const uint32_t wi = (g >> 3) & 3u;
    if (g & 7u) {
      if (wi == 0) w0 = cw; else if (wi == 1) w1 = cw; else if (wi == 2) w2 = cw; else w3 = cw;
    }
    // words beyond the last are don't-care
    reinterpret_cast<uint4*>(nibw)[g >> 5] = make_uint4(w0, w1, w2, w3);
  }
  if (!done) {
    levels_ok = level;
    for (uint32_t l = level + 1; l <= L; ++l) lv[l] = LevelEntry{nsym, pa};
    status |= FCS_ST_NSYM_MISMATCH;
  }
  // end-of-stream invariant: after the pad symbol of an odd count the state
  // is back at 2^23 with every byte consumed
  if (nsym & 1u) {
    const uint32_t x = state & (FCS_RANS_M - 1);
    const uint32_t e = table[x];
    state = ((e >> 12) & 0x1FFFu) * (state >> FCS_RANS_LOGM) + (x - (e & 0xFFFu));
    while (state < FCS_RANS_L && pos > 0) state = (state << 8) | rd.get(--pos);
    if ((e >> 25) != 0) status |= FCS_ST_RANS_STATE;
  }
  if (state != FCS_RANS_L || pos != 0) status |= FCS_ST_RANS_STATE;
  a.res[b] = RansResult{status, levels_ok};
}

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
  const RansResult rr = a.rres[b];
  uint32_t status = rr.status & FCS_ST_BAD_OP;
  if (t > rr.levels_ok) status |= FCS_ST_STREAM_OVERRUN | FCS_ST_NSYM_MISMATCH;
  const uint32_t P = B.P;
  const LT* pal = a.palette + B.pal_off;
  LT* o = a.out + q.out_off;
  if (P == 0 || P > FCS_PALETTE_MAX_LDS) {
    if (tid == 0) a.status[r] = status | (P == 0 ? FCS_ST_PALETTE_RANGE : FCS_ST_BAD_REQUEST);
    return;
  }
  const uint32_t* nibw = a.nib + B.nib_off;
  const LevelEntry* lv = a.lvl + (size_t)b * (L + 1);
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
      for (uint32_t w = tid; w < (C >> 5); w += FCS_OPS_THREADS) {
        const uint32_t pb = (scur[w >> 3] >> ((w & 7) * 4)) & 0xFu;  // 4 parents
        uint32_t m = 0;
        m |= (pb & 1u) ? 0x000000FFu : 0u;
        m |= (pb & 2u) ? 0x0000FF00u : 0u;
        m |= (pb & 4u) ? 0x00FF0000u : 0u;
        m |= (pb & 8u) ? 0xFF000000u : 0u;
        snext[w] = m;
      }
    }
    if (tid == 0) S.scr[61] = 0;
    __syncthreads();

    // 3-4. parse this level's ops and resolve them to values / references
    const uint32_t ns = lv[l].nib_start, ne = lv[l + 1].nib_start, ip0 = lv[l].ip_base;
    const uint32_t K = 8 * A;
    uint32_t op_base = 0, pa_base = 0, carry = 0;  // carry: next nibble is a delta
    const uint32_t w_first = ns >> 3, w_end = (ne + 7) >> 3;
    const uint32_t side = 1u << (l + 1);
    for (uint32_t wt = w_first; wt < w_end; wt += FCS_OPS_THREADS) {
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
        // lane holding the level's last nibble reports a dangling Pd op
        if (inw && ne > ns && ne - 1 >= g0 && ne - 1 < g0 + 8) S.scr[61] = pd;
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
          if (k >= K) { status |= FCS_ST_NSYM_MISMATCH; ++k; continue; }
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
          ++k;
        }
      }
      op_base += tot & 0xFFFFu;
      pa_base += tot >> 16;
      carry = tile_exit;
      __syncthreads();
    }
    if (op_base != K || S.scr[61] != 0) status |= FCS_ST_NSYM_MISMATCH;
    __syncthreads();

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

// ====================================================================== DDA
// One ray per lane, rayTrace.wgsl:565-892 up to the first hit, FP32 in the
// WGSL operation order (built with -ffp-contract=off, IEEE div/sqrt).
// A zero direction component has tMax = tDelta = +inf and no slab test.

__device__ __forceinline__ float fmx(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float fmn(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float biasf(float t) { return fmx(t * 1e-6f, 1e-5f); }

__device__ __forceinline__ float advance3(int& v0, int& v1, int& v2, float& m0, float& m1, float& m2,
                                          float d0, float d1, float d2, int s0, int s1, int s2) {
  float t;
  if (m0 < m1 && m0 < m2) { t = m0; m0 += d0; v0 += s0; }
  else if (m1 < m2) { t = m1; m1 += d1; v1 += s1; }
  else { t = m2; m2 += d2; v2 += s2; }
  return t;
}

template <typename LT>
__global__ __launch_bounds__(256) void march_kernel(MarchArgs<LT> a) {
  const uint32_t px = blockIdx.x * 16 + (threadIdx.x & 15);
  const uint32_t py = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (px >= a.w || py >= a.h) return;
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
    for (int k = 0; k < 3; ++k) {
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
    for (int k = 0; k < 3; ++k) {
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
    for (int k = 0; k < 3; ++k) d[k] = v[k] / len;
  }
  for (int k = 0; k < 3; ++k) {
    zero[k] = d[k] == 0.0f;
    const float s = d[k] > 0.0f ? 1.0f : (d[k] < 0.0f ? -1.0f : 0.0f);
    inv[k] = zero[k] ? __builtin_inff() : s / fmx(fabsf(d[k]), 1e-6f);
  }
  // slab test (rayTrace.wgsl:457-471)
  float tmin = -__builtin_inff(), tmax = __builtin_inff();
  bool miss = false;
  for (int k = 0; k < 3; ++k) {
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
      for (int k = 0; k < 3; ++k) {
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
    for (int k = 0; k < 3; ++k) {
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
        for (int k = 0; k < 3; ++k) {
          const float pos = o[k] + d[k] * ts;
          vox[k] = (int)floorf(pos);
          const float nb = d[k] > 0.0f ? (float)(vox[k] + 1) : (float)vox[k];
          vtmax[k] = zero[k] ? __builtin_inff() : (nb - o[k]) * inv[k];
          vtdelta[k] = zero[k] ? __builtin_inff() : fabsf(inv[k]);
        }
        const uint32_t lod = meta & FCS_META_LOD_MASK;
        const int lodSize = 1 << lod;
        const int stride = b / lodSize;
        for (int k = 0; k < 3; ++k) vox[k] -= bmin[k];
        int vstep[3] = {step[0], step[1], step[2]};
        if (stride > 1) {
          for (int k = 0; k < 3; ++k) {
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
          for (int k = 0; k < 3; ++k) vox[k] = vox[k] < 0 ? 0 : (vox[k] > b - 1 ? b - 1 : vox[k]);
        }
        uint32_t maxs = 128;
        for (int k = 0; k < 3; ++k) {
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
                for (int k = 0; k < 3; ++k) {
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

__global__ void zero_u32_kernel(uint32_t* p, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

}  // namespace

// ------------------------------------------------------------ launchers --
hipError_t fcs_launch_rans(const RansArgs& a, hipStream_t s) {
  if (a.nlist == 0) return hipSuccess;
  const uint32_t blocks = (a.nlist + 255) / 256;
  hipLaunchKernelGGL(rans_kernel, dim3(blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

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

template <typename LT>
hipError_t fcs_launch_march(const MarchArgs<LT>& a, hipStream_t s) {
  const dim3 grid((a.w + 15) / 16, (a.h + 15) / 16);
  hipLaunchKernelGGL(march_kernel<LT>, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t fcs_launch_zero(uint32_t* p, uint64_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(zero_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

template hipError_t fcs_launch_ops<uint32_t>(const OpsArgs<uint32_t>&, uint32_t, hipStream_t);
template hipError_t fcs_launch_ops<uint64_t>(const OpsArgs<uint64_t>&, uint32_t, hipStream_t);
template hipError_t fcs_launch_march<uint32_t>(const MarchArgs<uint32_t>&, hipStream_t);
template hipError_t fcs_launch_march<uint64_t>(const MarchArgs<uint64_t>&, hipStream_t);
