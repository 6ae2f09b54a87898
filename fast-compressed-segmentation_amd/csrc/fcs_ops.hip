// Op decoder for gfx950: decodes (brick, LOD) requests from the device
// nibble streams into the brick cache, bit-exactly like decode.wgsl:344-508 /
// brickDecoding.hpp:174-242 (with the root stop bit honoured).
//
// Decoding a brick at LOD t = computing the levels 1..t of its pyramid.
// Level l+1 depends only on level l and on *smaller* nodes of level l+1, so
// each level is one parallel step of a workgroup, persistent over a queue of
// requests. Per level, three phases separated by barriers:
//   P1  children of stopped parents copy the parent (one 16-byte LDS store
//       per parent); the level's nibble span is parsed 8 nibbles per lane: a
//       nibble is a Pd delta iff the run of 12/13-nibbles before it (back to
//       the level start) has odd length. Each wave publishes its op / palette
//       counts for both possible entering parities plus its exit parity, so
//       one barrier per 2048-nibble tile resolves every lane's op index and
//       palette cursor. Each op becomes a palette index (Pa/P0/Pd), the
//       parent's value (Rp) or a reference to a smaller node (Rx/Ry/Rz).
//   P2  references are chased; active nodes of the next level are counted.
//   P3  the next level's active list is compacted; labels are written.
// Intermediate levels live in LDS as 16-bit palette indices (bit 15 = node
// reference). The requested (final) level is never materialised in LDS: op
// results are kept per op (compact, in LDS) or per node in the cache region
// (when there are too many ops), and the labels are written per parent as
// 8 consecutive voxels (two or four 16-byte stores). Children of stopped
// parents are written straight from the parent's value. A level whose nodes
// are all stopped ends the brick: the rest is a direct expansion.
#include <hip/hip_runtime.h>

#include "fcs_common.h"
#include "fcs_device.h"
#include "fcs_kernels.h"

namespace {

template <uint32_t NT_, uint32_t VCAP_, uint32_t BIG_, uint32_t ACT_, uint32_t SW_, uint32_t KCAP_, uint32_t PCAP_>
struct OpsCfg {
  static constexpr uint32_t NT = NT_;      // threads per workgroup
  static constexpr uint32_t NW = NT_ / 64;  // waves
  static constexpr uint32_t VCAP = VCAP_;  // two ping-pong level buffers (u16)
  static constexpr uint32_t BIG = BIG_;    // level buffer beyond VCAP nodes (u16)
  static constexpr uint32_t ACT = ACT_;    // active-parent list (u16)
  static constexpr uint32_t SW = SW_;      // stop-bit words per level
  static constexpr uint32_t KCAP = KCAP_;  // compact final-level op results (u16)
  static constexpr uint32_t PCAP = PCAP_;  // palette entries staged in LDS
};
// bricks up to 32^3: four waves, ~46 KB of LDS (three workgroups per CU)
using CfgSmall = OpsCfg<256, 4096, 0, 4096, 128, 8192, 1024>;
// 64^3 bricks: sixteen waves, the 32768-node level in LDS (~155 KB)
using CfgLarge = OpsCfg<1024, 4096, 32768, 32768, 1024, 0, 0>;

constexpr uint32_t SCR_WORDS = 128;
constexpr uint32_t SCR_WAVESUM = 0;   // NW words: active-node counts per wave
constexpr uint32_t SCR_REC = 16;      // 2 x 48 words: parse records per wave
constexpr uint32_t SCR_STATUS = 120;
constexpr uint32_t SCR_NS = 122;      // nibble where the next level starts
constexpr uint32_t SCR_IP = 123;      // palette cursor at the next level's start
constexpr uint32_t SCR_REQ = 124;

template <class C, typename LT>
struct Layout {
  static constexpr size_t sA = 0;
  static constexpr size_t sB = sA + C::VCAP * 2;
  static constexpr size_t big = sB + C::VCAP * 2;
  static constexpr size_t act = big + C::BIG * 2;
  static constexpr size_t wpre = act + C::ACT * 2;
  static constexpr size_t opv = (wpre + C::SW * 2 + 15) / 16 * 16;
  static constexpr size_t st0 = (opv + C::KCAP * 2 + 15) / 16 * 16;
  static constexpr size_t st1 = st0 + C::SW * 4;
  static constexpr size_t scr = st1 + C::SW * 4;
  static constexpr size_t spal = (scr + SCR_WORDS * 4 + 15) / 16 * 16;
  static constexpr size_t bytes = spal + C::PCAP * sizeof(LT);
};

template <typename LT>
struct Lds {
  uint16_t* sA;
  uint16_t* sB;
  uint16_t* big;
  uint16_t* act;   // active parents of the level being expanded, Morton order
  uint16_t* wpre;  // active parents before each stop word (rank lookup)
  uint16_t* opv;   // compact final-level op results
  uint32_t* st0;
  uint32_t* st1;
  uint32_t* scr;
  LT* spal;
};

enum : uint32_t { MODE_LEVEL = 0, MODE_COMPACT = 1, MODE_GLOBAL = 2 };

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

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

__device__ __forceinline__ bool is_stopped(const uint32_t* st, uint32_t m) { return (st[m >> 5] >> (m & 31)) & 1u; }

// index of active node m among the active nodes of its level
__device__ __forceinline__ uint32_t rank_of(const uint32_t* st, const uint16_t* wpre, uint32_t m) {
  return wpre[m >> 5] + __popc(~st[m >> 5] & ((1u << (m & 31)) - 1u));
}

template <typename LT>
__device__ __forceinline__ uint4 splat16(LT v) {
  if constexpr (sizeof(LT) == 4) {
    const uint32_t u = (uint32_t)v;
    return make_uint4(u, u, u, u);
  } else {
    const uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    return make_uint4(lo, hi, lo, hi);
  }
}

template <typename LT>
__device__ __forceinline__ void store8(LT* p, const LT (&v)[8], bool al) {
  if (al) {
    uint4* q = reinterpret_cast<uint4*>(p);
    if constexpr (sizeof(LT) == 4) {
      q[0] = make_uint4(v[0], v[1], v[2], v[3]);
      q[1] = make_uint4(v[4], v[5], v[6], v[7]);
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i)
        q[i] = make_uint4((uint32_t)v[2 * i], (uint32_t)(v[2 * i] >> 32), (uint32_t)v[2 * i + 1],
                          (uint32_t)(v[2 * i + 1] >> 32));
    }
  } else {
#pragma unroll
    for (int i = 0; i < 8; ++i) p[i] = v[i];
  }
}

__device__ __forceinline__ bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

// store `count` labels of value v
template <typename LT, uint32_t NT>
__device__ __forceinline__ void fill_const(LT* o, uint64_t count, LT v) {
  const uint32_t tid = threadIdx.x;
  constexpr uint32_t PER = 16 / sizeof(LT);
  if (aligned16(o)) {
    const uint4 q = splat16(v);
    uint4* o4 = reinterpret_cast<uint4*>(o);
    const uint64_t n4 = count / PER;
    for (uint64_t i = tid; i < n4; i += NT) o4[i] = q;
    for (uint64_t i = n4 * PER + tid; i < count; i += NT) o[i] = v;
  } else {
    for (uint64_t i = tid; i < count; i += NT) o[i] = v;
  }
}

// level lt labels when every node of level lv is stopped: each node copies
// its level-lv ancestor
template <typename LT, uint32_t NT>
__device__ __forceinline__ void expand_level(LT* o, uint32_t lt, uint32_t lv, const uint16_t* vl, const LT* pal,
                                             uint32_t P) {
  const uint32_t tid = threadIdx.x;
  const uint32_t shift = 3 * (lt - lv);  // >= 3: 8 consecutive voxels share an ancestor
  const uint32_t n = 1u << (3 * lt);
  if (aligned16(o)) {
    constexpr uint32_t PER = 16 / sizeof(LT);
    uint4* o4 = reinterpret_cast<uint4*>(o);
    for (uint32_t i = tid; i < n / PER; i += NT) {
      const uint32_t v = vl[(i * PER) >> shift];
      o4[i] = splat16(pal[v < P ? v : P - 1]);
    }
  } else {
    for (uint32_t i = tid; i < n; i += NT) {
      const uint32_t v = vl[i >> shift];
      o[i] = pal[v < P ? v : P - 1];
    }
  }
}

// final level from compact op results: 8 labels per parent
template <typename LT, uint32_t NT>
__device__ __forceinline__ void out_compact(LT* olev, uint32_t N, const uint32_t* scur, const uint16_t* wpre,
                                            const uint16_t* vcur, const uint16_t* opv, const LT* pl, uint32_t P) {
  const bool al = aligned16(olev);
  for (uint32_t m = threadIdx.x; m < N; m += NT) {
    LT v8[8];
    if (is_stopped(scur, m)) {
      const uint32_t v = vcur[m];
      const LT u = pl[v < P ? v : P - 1];
#pragma unroll
      for (int c = 0; c < 8; ++c) v8[c] = u;
    } else {
      const uint4 q = *reinterpret_cast<const uint4*>(opv + 8 * rank_of(scur, wpre, m));
      const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const uint32_t v = (w[c >> 1] >> (16 * (c & 1))) & 0xFFFFu;
        v8[c] = pl[v < P ? v : P - 1];
      }
    }
    store8(olev + 8 * m, v8, al);
  }
}

// final level from per-node op results in the cache region
template <typename LT, uint32_t NT>
__device__ __forceinline__ void out_global(LT* olev, uint32_t N, const uint32_t* scur, const uint16_t* vcur,
                                           const LT* pl, uint32_t P) {
  const bool al = aligned16(olev);
  for (uint32_t m = threadIdx.x; m < N; m += NT) {
    LT v8[8];
    if (is_stopped(scur, m)) {
      const uint32_t v = vcur[m];
      const LT u = pl[v < P ? v : P - 1];
#pragma unroll
      for (int c = 0; c < 8; ++c) v8[c] = u;
    } else {
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const uint32_t v = (uint32_t)olev[8 * m + c];
        v8[c] = pl[v < P ? v : P - 1];
      }
    }
    store8(olev + 8 * m, v8, al);
  }
}

template <typename LT, class C>
__device__ void decode_request(const OpsArgs<LT>& a, const Lds<LT>& S, uint32_t r) {
  constexpr uint32_t NT = C::NT, NW = C::NW;
  const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
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
    fill_const<LT, NT>(o, pyr ? fcs_pyramid_level_offset(t + 1) : (1ull << (3 * t)), p0);
    if (tid == 0) a.status[r] = status;
    return;
  }
  const bool staged = C::PCAP != 0 && P <= C::PCAP;
  if (staged)
    for (uint32_t i = tid; i < P; i += NT) S.spal[i] = pal[i];
  const LT* lp = staged ? static_cast<const LT*>(S.spal) : pal;
  if (pyr && tid == 0) o[0] = p0;
  if (tid == 0) {
    // level 0: the root (palette index 0, active); level 1 not stopped yet
    S.sA[0] = 0;
    S.act[0] = 0;
    S.wpre[0] = 0;
    S.st0[0] = 0;
    S.st1[0] = 0;
    S.scr[SCR_STATUS] = 0;
    S.scr[SCR_NS] = 1;
    S.scr[SCR_IP] = 0;
  }
  __syncthreads();

  uint16_t* vcur = S.sA;
  uint32_t* scur = S.st0;
  uint32_t* snext = S.st1;
  // level l's ops start at nibble ns with palette cursor ip0 (nibble 0 is the
  // root op); each level's end is found while parsing it
  uint32_t A = 1, ns = 1, ip0 = 0, tile = 0;
  const uint32_t w_end = (nsym + 7) >> 3;
  const uint32_t ne = nsym;  // parse bound; a level ends after its K-th op

  for (uint32_t l = 0; l < t; ++l) {
    const uint32_t N = 1u << (3 * l), Cn = N << 3, K = 8 * A;
    const bool last = (l + 1 == t);
    const uint32_t mode = !last ? MODE_LEVEL : ((C::KCAP != 0 && K <= C::KCAP) ? MODE_COMPACT : MODE_GLOBAL);
    LT* olev = o + (pyr ? fcs_pyramid_level_offset(l + 1) : 0ull);
    uint16_t* vnext = (Cn <= C::VCAP) ? (vcur == S.sA ? S.sB : S.sA) : S.big;

    // ---- P1: stopped children; parse and resolve the level's ops
    if (mode == MODE_LEVEL) {
      for (uint32_t m = tid; m < N; m += NT)
        if (is_stopped(scur, m)) {
          const uint32_t v = vcur[m] | ((uint32_t)vcur[m] << 16);
          *reinterpret_cast<uint4*>(vnext + 8 * m) = make_uint4(v, v, v, v);
        }
    }
    uint32_t op_base = 0, pa_base = 0, carry = 0;  // carry: next nibble is a delta
    const uint32_t side = 1u << (l + 1);
    for (uint32_t wt = ns >> 3; wt < w_end && op_base < K; wt += NT, ++tile) {
      const uint32_t wi = wt + tid;
      const bool inw = wi < w_end;
      const uint32_t word = inw ? nibw[wi] : 0u;
      const uint32_t g0 = wi * 8;
      // trailing run of 12/13 nibbles over this lane's valid nibbles, and the
      // op / palette-advance counts for both entering parities
      uint32_t tr = 0, cnt0 = 0, cnt1 = 0;
      {
        uint32_t pd0 = 0, pd1 = 1;
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) {
          const uint32_t g = g0 + j;
          const uint32_t op = (word >> (4 * j + 1)) & 0x7u;
          const bool valid = inw && g >= ns && g < ne;
          tr = (valid && op == FCS_OP_PD) ? tr + 1 : 0;
          if (g == ns) { pd0 = 0; pd1 = 0; }
          const uint32_t inc = 1u | ((op == FCS_OP_PA) ? 0x10000u : 0u);
          if (valid) {
            if (pd0) pd0 = 0; else { cnt0 += inc; pd0 = (op == FCS_OP_PD); }
            if (pd1) pd1 = 0; else { cnt1 += inc; pd1 = (op == FCS_OP_PD); }
          }
        }
      }
      const uint64_t Pm = __ballot(tr & 1u);
      const uint64_t Nm = __ballot(tr != 8);
      const uint64_t below = Nm & ((lane == 0) ? 0ull : (~0ull >> (64 - lane)));
      const bool fixed = below != 0;  // entering parity set by a lane of this wave
      const uint32_t pfix = fixed ? (uint32_t)(Pm >> (63 - __clzll(below))) & 1u : 0u;
      const uint32_t cfix = pfix ? cnt1 : cnt0;
      const uint32_t v0 = fixed ? cfix : cnt0, v1 = fixed ? cfix : cnt1;
      const uint32_t i0 = wave_incl_scan(v0), i1 = wave_incl_scan(v1);
      uint32_t* R = S.scr + SCR_REC + (tile & 1u) * 48u;
      if (lane == 63) {
        R[3 * wave] = Nm ? (2u | ((uint32_t)(Pm >> (63 - __clzll(Nm))) & 1u)) : 0u;
        R[3 * wave + 1] = i0;
        R[3 * wave + 2] = i1;
      }
      __syncthreads();
      uint32_t X = carry, base = 0, myX = 0, mybase = 0;
#pragma unroll
      for (uint32_t w = 0; w < NW; ++w) {
        if (w == wave) { myX = X; mybase = base; }
        const uint32_t rec = R[3 * w];
        base += X ? R[3 * w + 2] : R[3 * w + 1];
        X = (rec & 2u) ? (rec & 1u) : X;
      }
      const uint32_t par = fixed ? pfix : myX;
      const uint32_t ex = mybase + (myX ? i1 - v1 : i0 - v0);
      // resolve this lane's ops
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
          if (k >= K) break;  // the rest belongs to the next level
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
              const uint32_t dn = (j < 7) ? ((word >> (4 * (j + 1))) & 0xFu) : (nibw[wi + 1] & 0xFu);
              const int64_t idx = (int64_t)ipx - (int64_t)dn - 1;
              if (idx < 0 || idx >= (int64_t)P) { status |= FCS_ST_PALETTE_RANGE; val = P - 1; }
              else val = (uint32_t)idx;
              break;
            }
            default:
              status |= FCS_ST_BAD_OP;
              val = 0;
          }
          const uint32_t v16 = (val & FCS_REF32) ? (FCS_REF16 | (val & 0x7FFFu)) : val;
          if (mode == MODE_LEVEL) {
            vnext[n] = (uint16_t)v16;
            if (stop) atomicOr(&snext[n >> 5], 1u << (n & 31));
          } else if (mode == MODE_COMPACT) {
            S.opv[k] = (uint16_t)v16;
          } else {
            olev[n] = (LT)val;
          }
          if (k == K - 1) {
            // last op of the level: the next level starts after it (and its delta)
            S.scr[SCR_NS] = g + 1 + (op == FCS_OP_PD ? 1u : 0u);
            S.scr[SCR_IP] = ipx;
          }
          ++k;
        }
      }
      op_base += base & 0xFFFFu;
      pa_base += base >> 16;
      carry = X;
    }
    if (op_base < K) status |= FCS_ST_STREAM_OVERRUN | FCS_ST_NSYM_MISMATCH;
    __syncthreads();
    ns = S.scr[SCR_NS];
    ip0 = S.scr[SCR_IP];
    if (l + 1 == L && ns != nsym) status |= FCS_ST_NSYM_MISMATCH;  // nibbles after the finest level

    // ---- P2: chase references. A valid reference points to a smaller node,
    // so a chain strictly decreases; anything else is corrupt input. Only op
    // children can hold references.
    uint32_t cnt = 0, inc = 0, w0 = 0, w1 = 0;
    const uint32_t nw = (Cn + 31) >> 5;
    if (mode == MODE_LEVEL) {
      for (uint32_t k = tid; k < K; k += NT) {
        const uint32_t n = 8 * S.act[k >> 3] + (k & 7u);
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
      // active nodes of level l+1, over consecutive stop words per thread
      const uint32_t wpt = (nw + NT - 1) / NT;
      w0 = min(nw, tid * wpt);
      w1 = min(nw, w0 + wpt);
      const uint32_t valid = (Cn >= 32) ? 0xFFFFFFFFu : ((1u << Cn) - 1u);
      for (uint32_t w = w0; w < w1; ++w) cnt += __popc(~snext[w] & valid);
      inc = wave_incl_scan(cnt);
      if (lane == 63) S.scr[SCR_WAVESUM + wave] = inc;
    } else if (mode == MODE_COMPACT) {
      for (uint32_t k = tid; k < K; k += NT) {
        uint32_t v = S.opv[k];
        if (v & FCS_REF16) {
          uint32_t cur = 8 * S.act[k >> 3] + (k & 7u);
          for (;;) {
            const uint32_t nb = v & 0x7FFFu;
            if (nb >= cur) { status |= FCS_ST_NSYM_MISMATCH; v = 0; break; }
            const uint32_t p = nb >> 3;
            if (is_stopped(scur, p)) { v = vcur[p]; break; }
            v = S.opv[8 * rank_of(scur, S.wpre, p) + (nb & 7u)];
            if (!(v & FCS_REF16)) break;
            cur = nb;
          }
          S.opv[k] = (uint16_t)v;
        }
      }
    } else {
      for (uint32_t k = tid; k < K; k += NT) {
        const uint32_t n = 8 * S.act[k >> 3] + (k & 7u);
        uint32_t v = (uint32_t)olev[n];
        if (v & FCS_REF32) {
          uint32_t cur = n;
          for (;;) {
            const uint32_t nb = v & ~FCS_REF32;
            if (nb >= cur) { status |= FCS_ST_NSYM_MISMATCH; v = 0; break; }
            const uint32_t p = nb >> 3;
            if (is_stopped(scur, p)) { v = vcur[p]; break; }
            v = (uint32_t)olev[nb];
            if (!(v & FCS_REF32)) break;
            cur = nb;
          }
          olev[n] = (LT)v;
        }
      }
    }
    __syncthreads();

    // ---- P3: next level's active list; labels out
    bool ended = false;
    if (mode == MODE_LEVEL) {
      uint32_t An = 0, wb = 0;
#pragma unroll
      for (uint32_t w = 0; w < NW; ++w) {
        const uint32_t s = S.scr[SCR_WAVESUM + w];
        wb += (w < wave) ? s : 0u;
        An += s;
      }
      if (pyr)
        for (uint32_t n = tid; n < Cn; n += NT) {
          const uint32_t v = vnext[n];
          olev[n] = lp[v < P ? v : P - 1];
        }
      if (An == 0) {
        // every node of level l+1 is stopped: the remaining levels are copies
        for (uint32_t lt = pyr ? l + 2 : t; lt <= t; ++lt) {
          LT* ol = o + (pyr ? fcs_pyramid_level_offset(lt) : 0ull);
          if (staged) expand_level<LT, NT>(ol, lt, l + 1, vnext, S.spal, P);
          else expand_level<LT, NT>(ol, lt, l + 1, vnext, pal, P);
        }
        if (t == L && ns != nsym) status |= FCS_ST_NSYM_MISMATCH;  // nibbles after the finest level
        ended = true;
      } else {
        uint32_t k0 = wb + inc - cnt;
        const uint32_t valid = (Cn >= 32) ? 0xFFFFFFFFu : ((1u << Cn) - 1u);
        for (uint32_t w = w0; w < w1; ++w) {
          S.wpre[w] = (uint16_t)k0;
          uint32_t bits = ~snext[w] & valid;
          while (bits) {
            const uint32_t j = __ffs(bits) - 1;
            bits &= bits - 1;
            S.act[k0++] = (uint16_t)(w * 32 + j);
          }
        }
        if (l + 2 < t) {
          // stop bits of level l+2: children of stopped nodes are stopped
          for (uint32_t w = tid; w < Cn / 4; w += NT) {
            const uint32_t pb = (snext[w >> 3] >> ((w & 7) * 4)) & 0xFu;  // 4 parents
            uint32_t mm = 0;
            mm |= (pb & 1u) ? 0x000000FFu : 0u;
            mm |= (pb & 2u) ? 0x0000FF00u : 0u;
            mm |= (pb & 4u) ? 0x00FF0000u : 0u;
            mm |= (pb & 8u) ? 0xFF000000u : 0u;
            scur[w] = mm;
          }
        }
        if (tid == 0) { S.scr[SCR_NS] = ns; S.scr[SCR_IP] = ip0; }
      }
      vcur = vnext;
      uint32_t* tmp = scur;
      scur = snext;
      snext = tmp;
      A = An;
    } else if (mode == MODE_COMPACT) {
      if (staged) out_compact<LT, NT>(olev, N, scur, S.wpre, vcur, S.opv, S.spal, P);
      else out_compact<LT, NT>(olev, N, scur, S.wpre, vcur, S.opv, pal, P);
    } else {
      if (staged) out_global<LT, NT>(olev, N, scur, vcur, S.spal, P);
      else out_global<LT, NT>(olev, N, scur, vcur, pal, P);
    }
    __syncthreads();
    if (ended) break;
  }
  if (status) atomicOr(&S.scr[SCR_STATUS], status);
  __syncthreads();
  if (tid == 0) a.status[r] = S.scr[SCR_STATUS];
}

template <typename LT, class C>
__global__ __launch_bounds__(C::NT) void ops_kernel(OpsArgs<LT> a) {
  extern __shared__ __align__(16) uint8_t smem[];
  using Y = Layout<C, LT>;
  Lds<LT> S;
  S.sA = reinterpret_cast<uint16_t*>(smem + Y::sA);
  S.sB = reinterpret_cast<uint16_t*>(smem + Y::sB);
  S.big = reinterpret_cast<uint16_t*>(smem + Y::big);
  S.act = reinterpret_cast<uint16_t*>(smem + Y::act);
  S.wpre = reinterpret_cast<uint16_t*>(smem + Y::wpre);
  S.opv = reinterpret_cast<uint16_t*>(smem + Y::opv);
  S.st0 = reinterpret_cast<uint32_t*>(smem + Y::st0);
  S.st1 = reinterpret_cast<uint32_t*>(smem + Y::st1);
  S.scr = reinterpret_cast<uint32_t*>(smem + Y::scr);
  S.spal = reinterpret_cast<LT*>(smem + Y::spal);
  for (;;) {
    if (threadIdx.x == 0) S.scr[SCR_REQ] = atomicAdd(a.counter, 1u);
    __syncthreads();
    const uint32_t r = S.scr[SCR_REQ];
    __syncthreads();
    if (r >= a.nreq) break;
    decode_request<LT, C>(a, S, r);
    __syncthreads();
  }
}

__global__ void zero_u32_kernel(uint32_t* p, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

template <typename LT, class C>
hipError_t launch_cfg(const OpsArgs<LT>& a, uint32_t num_cu, hipStream_t s, uint32_t* grid_out) {
  constexpr size_t lds = Layout<C, LT>::bytes;
  static_assert(lds <= 160 * 1024, "op decoder LDS budget");
  // resident workgroups per CU (queried once per instantiation)
  static int per_cu = 0;
  if (per_cu == 0) {
    const void* fn = reinterpret_cast<const void*>(&ops_kernel<LT, C>);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    int nb = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, (int)C::NT, lds);
    if (e != hipSuccess) return e;
    per_cu = nb > 0 ? nb : 1;
  }
  const uint64_t resident = (uint64_t)per_cu * (num_cu ? num_cu : 1);
  const uint32_t grid = (uint32_t)(a.nreq < resident ? a.nreq : resident);
  if (grid_out) *grid_out = grid;
  hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, a.counter, (uint64_t)1);
  hipLaunchKernelGGL((ops_kernel<LT, C>), dim3(grid), dim3(C::NT), lds, s, a);
  return hipGetLastError();
}

}  // namespace

size_t fcs_ops_lds_bytes(uint32_t L) {
  return L >= 6 ? Layout<CfgLarge, uint64_t>::bytes : Layout<CfgSmall, uint64_t>::bytes;
}

template <typename LT>
hipError_t fcs_launch_ops(const OpsArgs<LT>& a, uint32_t num_cu, hipStream_t s, uint32_t* grid_out) {
  if (a.nreq == 0) return hipSuccess;
  if (a.L > 6) return hipErrorInvalidValue;  // 16-bit node indices: bricks up to 64^3
  if (a.L == 6) return launch_cfg<LT, CfgLarge>(a, num_cu, s, grid_out);
  return launch_cfg<LT, CfgSmall>(a, num_cu, s, grid_out);
}

hipError_t fcs_launch_zero(uint32_t* p, uint64_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(zero_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

template hipError_t fcs_launch_ops<uint32_t>(const OpsArgs<uint32_t>&, uint32_t, hipStream_t, uint32_t*);
template hipError_t fcs_launch_ops<uint64_t>(const OpsArgs<uint64_t>&, uint32_t, hipStream_t, uint32_t*);
