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
//
// Large launches run in two kernels: ops_head_kernel decodes the non-final
// levels up to 3 of every request, one wave per request, and leaves the
// level state in the request's output; ops_kernel reads it back and decodes
// the remaining levels with the whole workgroup.
#include <hip/hip_runtime.h>

#include "fcs_common.h"
#include "fcs_device.h"
#include "fcs_kernels.h"

namespace {

template <uint32_t NT_, uint32_t VCAP_, uint32_t XCAP_, uint32_t BIG_, uint32_t ACT_, bool PACKED_, uint32_t SW_,
          uint32_t PBYTES_, uint32_t MINW_ = 1>
struct OpsCfg {
  static constexpr uint32_t MINW = MINW_;    // waves per SIMD the register budget must allow
  static constexpr uint32_t NT = NT_;        // threads per workgroup
  static constexpr uint32_t NW = NT_ / 64;    // waves
  static constexpr uint32_t VCAP = VCAP_;    // two ping-pong level buffers (u16)
  static constexpr uint32_t XCAP = XCAP_;    // u16 between them: the final level's compact op results use
                                             // it plus the ping-pong buffer not holding the parents
  static constexpr uint32_t KCAP = XCAP_ ? VCAP_ + XCAP_ : 0u;
  static constexpr uint32_t BIG = BIG_;      // level buffer beyond VCAP nodes (u16)
  static constexpr uint32_t ACT = ACT_;      // active-parent list entries
  static constexpr bool PACKED = PACKED_;    // active-parent entries carry the parent's value (u32)
  static constexpr uint32_t SW = SW_;        // stop-bit words per level
  static constexpr uint32_t PBYTES = PBYTES_;  // palette bytes staged in LDS
};
// bricks up to 32^3, two LDS splits picked per volume by the runtime's
// autotuner: ~52 KB (three workgroups per CU, parent values packed into the
// active list, 12,288 compact ops) or ~37 KB (four workgroups per CU within
// 128 VGPRs, 8,192 compact ops)
using CfgSmall = OpsCfg<256, 4096, 8192, 0, 4096, true, 128, 2048>;
using CfgSmall4 = OpsCfg<256, 4096, 4096, 0, 4096, false, 128, 1024, 4>;
// 64^3 bricks: sixteen waves, the 32768-node level in LDS (~155 KB)
using CfgLarge = OpsCfg<1024, 4096, 0, 32768, 32768, false, 1024, 0>;
// head pass: one wave decodes levels 1..3 of one request (~6.5 KB)
using CfgHead = OpsCfg<64, 512, 512, 0, 512, true, 128, 0>;

constexpr uint32_t SCR_WORDS = 256;
constexpr uint32_t SCR_WAVESUM = 0;   // NW words: active-node counts per wave
constexpr uint32_t SCR_REC = 16;      // 2 x 48 words: parse records per wave
constexpr uint32_t SCR_HAND = 112;    // 5 words: level state handed from the first wave to the workgroup
constexpr uint32_t SCR_STATUS = 120;
constexpr uint32_t SCR_NS = 122;      // nibble where the next level starts
constexpr uint32_t SCR_IP = 123;      // palette cursor at the next level's start
constexpr uint32_t SCR_DUM16 = 128;   // 64 u16: write target of lanes without an op
constexpr uint32_t SCR_DUM32 = 160;   // 64 u32: same, for stop bits
constexpr uint32_t SCR_IDX = 224;     // 4 words: request index ring
constexpr uint32_t SCR_HEAD = 232;    // 8 words: header of the head pass state

template <class C, typename LT>
struct Layout {
  static constexpr size_t sA = 0;
  static constexpr size_t opx = sA + C::VCAP * 2;
  static constexpr size_t sB = opx + C::XCAP * 2;
  static constexpr size_t big = sB + C::VCAP * 2;
  static constexpr size_t act = big + C::BIG * 2;
  static constexpr size_t wpre = act + C::ACT * (C::PACKED ? 4 : 2);
  static constexpr size_t st0 = (wpre + C::SW * 2 + 15) / 16 * 16;
  static constexpr size_t st1 = st0 + C::SW * 4;
  static constexpr size_t scr = st1 + C::SW * 4;
  static constexpr size_t ring = scr + SCR_WORDS * 4;  // 4 x (DecodeReq + DevBrick) = 192 bytes
  static constexpr size_t spal = (ring + 192 + 15) / 16 * 16;
  static constexpr size_t bytes = spal + C::PBYTES;
  static constexpr uint32_t PCAP = C::PBYTES / sizeof(LT);
  // kstop (XCAP / 32 words) ends the active list; a level parsed in compact
  // form has at most XCAP / 8 active parents, which must stay clear of it
  static_assert(C::XCAP / 8 * (C::PACKED ? 4 : 2) + C::XCAP / 8 <= C::ACT * (C::PACKED ? 4 : 2),
                "compact-level stop bits overlap the active list");
};

template <typename LT>
struct Lds {
  uint16_t* sA;
  uint16_t* sB;
  uint16_t* big;
  uint32_t* act32;  // active parents of the level being expanded (Morton order): node | value << 16
  uint16_t* act16;  // same, node only
  uint16_t* wpre;   // active parents before each stop word (rank lookup)
  uint32_t* kstop;  // stop bits in op order of a non-final level parsed in compact form (the
                    // active list's tail, unused while that level's parents are at most XCAP / 8)
  uint32_t* st0;
  uint32_t* st1;
  uint32_t* scr;
  DecodeReq* rq;    // request ring (4)
  DevBrick* rb;     // brick-record ring (4)
  LT* spal;
};

template <class C, typename LT>
__device__ __forceinline__ Lds<LT> make_lds(uint8_t* smem) {
  using Y = Layout<C, LT>;
  Lds<LT> S;
  S.sA = reinterpret_cast<uint16_t*>(smem + Y::sA);
  S.sB = reinterpret_cast<uint16_t*>(smem + Y::sB);
  S.big = reinterpret_cast<uint16_t*>(smem + Y::big);
  S.act32 = reinterpret_cast<uint32_t*>(smem + Y::act);
  S.act16 = reinterpret_cast<uint16_t*>(smem + Y::act);
  S.wpre = reinterpret_cast<uint16_t*>(smem + Y::wpre);
  S.kstop = reinterpret_cast<uint32_t*>(smem + Y::wpre) - C::XCAP / 32;
  S.st0 = reinterpret_cast<uint32_t*>(smem + Y::st0);
  S.st1 = reinterpret_cast<uint32_t*>(smem + Y::st1);
  S.scr = reinterpret_cast<uint32_t*>(smem + Y::scr);
  S.rq = reinterpret_cast<DecodeReq*>(smem + Y::ring);
  S.rb = reinterpret_cast<DevBrick*>(smem + Y::ring + 64);
  S.spal = reinterpret_cast<LT*>(smem + Y::spal);
  return S;
}

// MODE_LEVEL: ops resolved per nibble into the next level's node array;
// MODE_LCOMPACT: a non-final level parsed like MODE_COMPACT (op nibbles in op
// order), then resolved per active parent into the next level's node array;
// MODE_COMPACT / MODE_GLOBAL: the final level, op results per op in LDS or
// per node in the cache region.
enum : uint32_t { MODE_LEVEL = 0, MODE_COMPACT = 1, MODE_GLOBAL = 2, MODE_LCOMPACT = 3 };

// The same configuration run by one wave (the first of the workgroup): its
// LDS layout is the workgroup's, its thread count 64 and its syncs wave-level.
template <class C>
struct OneWave : C {
  static constexpr uint32_t NT = 64;
  static constexpr uint32_t NW = 1;
};

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Workgroup barrier that orders LDS only: global loads stay in flight across
// it (__syncthreads() would also wait for every outstanding global access,
// including prefetches meant for later). Used wherever threads exchange data
// through LDS alone.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// The sync of a (sub-)configuration: a workgroup barrier, or for one wave an
// ordering point only (a wave's LDS accesses complete in program order).
template <class C>
__device__ __forceinline__ void group_sync() {
  if constexpr (C::NW == 1) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
  } else {
    lds_barrier();
  }
}

// inclusive wave scan on the VALU data-parallel-primitive lanes: shifts by
// 1, 2, 4, 8 within each row of 16 lanes, then row 0's / row 1's last lane
// broadcast into the rows above (row_bcast:15, row_bcast:31). Lanes whose
// source is outside the row, and rows masked off, add 0.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);  // row_shr:1
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);  // row_shr:2
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);  // row_shr:4
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);  // row_shr:8
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast:15
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);  // row_bcast:31
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

// 16-byte store as one clang vector: a uint4 struct store built from
// separate values is emitted as four dword stores
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <typename LT>
__device__ __forceinline__ void store8(LT* p, const LT (&v)[8], bool al) {
  if (al) {
    u32x4* q = reinterpret_cast<u32x4*>(__builtin_assume_aligned(p, 16));
    if constexpr (sizeof(LT) == 4) {
      q[0] = u32x4{v[0], v[1], v[2], v[3]};
      q[1] = u32x4{v[4], v[5], v[6], v[7]};
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i)
        q[i] = u32x4{(uint32_t)v[2 * i], (uint32_t)(v[2 * i] >> 32), (uint32_t)v[2 * i + 1],
                     (uint32_t)(v[2 * i + 1] >> 32)};
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
      o4[i] = splat16(pal[min(v, P - 1)]);
    }
  } else {
    for (uint32_t i = tid; i < n; i += NT) {
      const uint32_t v = vl[i >> shift];
      o[i] = pal[min(v, P - 1)];
    }
  }
}

// final level from compact op results: 8 labels per parent. Branch-free so
// that two parents' LDS reads overlap: stopped parents still read a (valid)
// op group and then ignore it.
template <typename LT, uint32_t NT>
__device__ __forceinline__ void out_compact(LT* olev, uint32_t N, uint32_t A, const uint32_t* scur,
                                            const uint16_t* wpre, const uint16_t* vcur, const uint16_t* opv,
                                            const LT* pl, uint32_t P) {
  const bool al = aligned16(olev);
#pragma unroll 2
  for (uint32_t m = threadIdx.x; m < N; m += NT) {
    const uint32_t sw = scur[m >> 5];
    const bool stp = (sw >> (m & 31)) & 1u;
    const uint32_t pv = vcur[m];
    LT v8[8];
    if (__ballot(!stp) == 0) {
      // the wave's 64 parents are all stopped: one label per parent
      const LT u = pl[min(pv, P - 1)];
#pragma unroll
      for (int c = 0; c < 8; ++c) v8[c] = u;
      store8(olev + 8 * m, v8, al);
      continue;
    }
    const uint32_t rk = min(wpre[m >> 5] + __popc(~sw & ((1u << (m & 31)) - 1u)), A - 1);
    const uint4 q = *reinterpret_cast<const uint4*>(opv + 8 * rk);
    // a stopped parent's children all take its value: select per word, not per label
    const uint32_t pv2 = pv | (pv << 16);
    const uint32_t w[4] = {stp ? pv2 : q.x, stp ? pv2 : q.y, stp ? pv2 : q.z, stp ? pv2 : q.w};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const uint32_t v = (c & 1) ? (w[c >> 1] >> 16) : (w[c >> 1] & 0xFFFFu);
      v8[c] = pl[min(v, P - 1)];
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
      const LT u = pl[min(v, P - 1)];
#pragma unroll
      for (int c = 0; c < 8; ++c) v8[c] = u;
    } else {
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const uint32_t v = (uint32_t)olev[8 * m + c];
        v8[c] = pl[min(v, P - 1)];
      }
    }
    store8(olev + 8 * m, v8, al);
  }
}

// The brick's nibble stream is parsed in NT-word blocks; each thread keeps
// its word (and the word after it, for Pd deltas) of the current block and
// of the next one in registers, so a block's global loads are issued a full
// parse step before they are used.
struct Blocks {
  uint32_t bc;      // block held in w0 / n0
  uint32_t w0, n0;  // block bc
  uint32_t w1, n1;  // block bc + 1
};

template <uint32_t NT>
__device__ __forceinline__ void load_block(const uint32_t* nibw, uint32_t w_end, uint32_t b, uint32_t& w,
                                           uint32_t& n) {
  const uint32_t wi = b * NT + threadIdx.x;
  w = wi < w_end ? nibw[wi] : 0u;
  n = wi < w_end ? nibw[wi + 1] : 0u;  // streams carry padding words
}

// Per-level inputs of the op parser (uniform across the workgroup).
template <typename LT>
struct LevelIO {
  const uint32_t* nibw;
  uint32_t ns, ne, w_end, K, ip0, P, lmask;
  const uint16_t* vcur;
  uint16_t* vnext;   // MODE_LEVEL: next level values
  uint32_t* snext;   // MODE_LEVEL: next level stop bits
  uint16_t* opv;     // MODE_COMPACT / MODE_LCOMPACT: per-op results
  LT* olev;          // MODE_GLOBAL: per-node results in the cache region
  uint32_t* kstop;   // MODE_LCOMPACT: stop bits in op order (cleared before the parse)
};

// Parse and resolve the ops of one level (P1). Returns the number of ops
// found (< K means the stream ended early). The lane holding the level's last
// op publishes where the next level starts (SCR_NS / SCR_IP).
template <uint32_t MODE, typename LT, class C>
__device__ __forceinline__ uint32_t parse_level(const Lds<LT>& S, const LevelIO<LT>& io, Blocks& B, uint32_t& tile,
                                                uint32_t& status) {
  constexpr uint32_t NT = C::NT, NW = C::NW;
  const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  const uint32_t ns = io.ns, ne = io.ne, K = io.K, P = io.P;
  uint32_t op_base = 0, pa_base = 0, carry = 0;  // carry: next nibble is a delta
  uint32_t st = 0;
  // a level of K ops spans at most 2K nibbles (each op plus a Pd delta):
  // words past that carry no op of this level and are treated as invalid
  const uint32_t w_lim = min(io.w_end, ((ns + 2 * K + 7) >> 3) + 1);
  for (uint32_t b = (ns >> 3) / NT; b * NT < w_lim && op_base < K; ++b, ++tile) {
    if (b != B.bc) {
      // levels are contiguous: normally the next block, already in flight
      if (b == B.bc + 1) { B.w0 = B.w1; B.n0 = B.n1; }
      else load_block<NT>(io.nibw, io.w_end, b, B.w0, B.n0);
      B.bc = b;
      load_block<NT>(io.nibw, io.w_end, b + 1, B.w1, B.n1);
    }
    const uint32_t wi = b * NT + tid;
    const bool inw = wi < w_lim;
    const uint32_t word = B.w0;
    const uint32_t wnext = B.n0;
    const uint32_t g0 = wi * 8;
    // trailing run of 12/13 nibbles over this lane's valid nibbles, the op /
    // palette-advance counts for both entering parities, and which nibbles
    // are ops (flag at bit 4j) for both parities. Bit-parallel over the
    // word's 8 nibbles: a nibble is an op unless the one before it is a Pd
    // op (whose delta it then is); so within a run of Pd nibbles the ops
    // alternate from the run's first nibble, or from the one after it when
    // the run starts the word and the lane enters after a Pd op (parity 1).
    // A level's first nibble enters with parity 0.
    uint32_t tr, cnt0, cnt1, iso0, iso1;
    {
      constexpr uint32_t FL = 0x11111111u, EV = 0x01010101u, OD = 0x10101010u;
      const int lo = (int)ns - (int)g0, hi = (int)ne - (int)g0;
      const uint32_t mlo = (lo <= 0) ? ~0u : (lo >= 8) ? 0u : (~0u << (4 * lo));
      const uint32_t mhi = (hi >= 8) ? ~0u : (hi <= 0) ? 0u : ((1u << (4 * hi)) - 1u);
      const uint32_t V = inw ? (FL & mlo & mhi) : 0u;  // valid nibbles of this level
      const uint32_t b3 = word >> 3, b2 = word >> 2, b1 = word >> 1;
      const uint32_t Q = b3 & b2 & ~b1 & V;   // Pd (12/13)
      const uint32_t Pa = b3 & ~b2 & ~b1 & V;  // Pa (8/9)
      const uint32_t Qf = Q * 15u;             // runs of Pd nibbles as runs of bits
      const uint32_t S0 = Q & ~(Q << 4);       // run starts
      const uint32_t pin = (g0 > ns) ? 1u : 0u;
      const uint32_t S1 = S0 & ~pin;
      // runs that start on an even nibble: adding 1 at the start clears the run
      const uint32_t r0 = Qf & ~(Qf + (S0 & EV)), r1 = Qf & ~(Qf + (S1 & EV));
      const uint32_t d0 = Q & ((r0 & EV) | (~r0 & OD)), d1 = Q & ((r1 & EV) | (~r1 & OD));  // Pd ops
      iso0 = V & ~(d0 << 4);
      iso1 = V & ~((d1 << 4) | pin);
      cnt0 = (uint32_t)__popc(iso0) | ((uint32_t)__popc(iso0 & Pa) << 16);
      cnt1 = (uint32_t)__popc(iso1) | ((uint32_t)__popc(iso1 & Pa) << 16);
      tr = (uint32_t)__clz(~Qf) >> 2;  // 8 when every nibble is a valid Pd
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
    group_sync<C>();
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
    // resolve this lane's ops: straight-line code per nibble; lanes without a
    // live op write to a private dummy slot. Waves without a live op skip.
    const uint32_t my_ops = (myX ? v1 : v0) & 0xFFFFu;
    if (__ballot(my_ops != 0 && op_base + (ex & 0xFFFFu) < K) != 0) {
      const uint32_t k0 = op_base + (ex & 0xFFFFu), ip_start = io.ip0 + pa_base + (ex >> 16);
      uint32_t k = k0;
      const uint32_t iso = par ? iso1 : iso0;
      // node references: bit 15 of the LDS u16 forms, bit 31 of the labels
      constexpr uint32_t REFB = (MODE == MODE_GLOBAL) ? FCS_REF32 : FCS_REF16;
      uint16_t* dum16 = reinterpret_cast<uint16_t*>(S.scr + SCR_DUM16) + lane;
      uint32_t* dum32 = S.scr + SCR_DUM32 + lane;
      // an active parent owns 8 consecutive ops, so the lane's (at most 8)
      // ops belong to at most two parents: read both once (clamped to the
      // level's last parent; lanes past the level read a valid one and ignore it)
      const uint32_t qlast = (K - 1) >> 3, q0 = min(k0 >> 3, qlast), q1 = min(q0 + 1, qlast);
      uint32_t m0, pv0, m1, pv1;
      if constexpr (C::PACKED) {
        const uint32_t e0 = S.act32[q0], e1 = S.act32[q1];
        m0 = e0 & 0xFFFFu;
        pv0 = e0 >> 16;
        m1 = e1 & 0xFFFFu;
        pv1 = e1 >> 16;
      } else {
        m0 = S.act16[q0];
        m1 = S.act16[q1];
        pv0 = io.vcur[m0];
        pv1 = io.vcur[m1];
      }
      bool e_pal = false, e_ovr = false, e_op = false;
      if constexpr (MODE == MODE_COMPACT || MODE == MODE_LCOMPACT) {
        // the op nibble as it is, tagged with bit 15 (palette results, all
        // below 0x8000, overwrite their slots below); resolve_compact /
        // resolve_level turn it into a value or a node reference once the
        // level is parsed
        uint32_t sbits = 0;  // MODE_LCOMPACT: stop bits of the lane's ops, in op order
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) {
          const uint32_t v = (word >> (4 * j)) & 0xFu;
          const bool isop = (iso >> (4 * j)) & 1u;
          const bool live = isop && k < K;
          e_op = e_op || (live && (v >> 1) > FCS_OP_PD);
          *(live ? &io.opv[k] : dum16) = (uint16_t)(FCS_REF16 | v);
          if constexpr (MODE == MODE_LCOMPACT) sbits |= (live && (v & 1u)) ? (1u << (k - k0)) : 0u;
          k += isop ? 1u : 0u;
        }
        if constexpr (MODE == MODE_LCOMPACT) {
          // at most 8 consecutive op indices from k0: one or two words
          if (sbits != 0u) {
            const uint64_t sh = (uint64_t)sbits << (k0 & 31u);
            atomicOr(&io.kstop[k0 >> 5], (uint32_t)sh);
            if ((uint32_t)(sh >> 32)) atomicOr(&io.kstop[(k0 >> 5) + 1], (uint32_t)(sh >> 32));
          }
        }
      } else {
#pragma unroll
      for (uint32_t j = 0; j < 8; ++j) {
        const uint32_t v = (word >> (4 * j)) & 0xFu;
        const uint32_t op = v >> 1;
        const bool isop = (iso >> (4 * j)) & 1u;
        const bool live = isop && k < K;
        const bool second = (k >> 3) != q0;
        const uint32_t m = second ? m1 : m0, pv = second ? pv1 : pv0;
        const uint32_t c = k & 7u, n = (m << 3) | c;
        // Rx/Ry/Rz: neighbour along the axis by dilated-integer arithmetic on
        // the Morton code (x-1 for an even coordinate, x+1 for an odd one)
        const uint32_t ax = (op >= FCS_OP_RX && op <= FCS_OP_RZ) ? op - 1 : 0u;
        const uint32_t mA = (0x09249249u << ax) & io.lmask;
        const uint32_t xa = n & mA, rest = n & ~mA;
        const bool odd = (n >> ax) & 1u;
        const bool outside = odd ? (xa == mA) : (xa == 0u);
        const uint32_t nb = (odd ? (((xa | ~mA) + 1u) & mA) : ((xa - 1u) & mA)) | rest;
        const uint32_t nv = io.vcur[nb >> 3];
        // outside the brick the reference reads the parent's output slot,
        // which child 0 has already overwritten when c > 0. Bitwise selects
        // (v_bfi): the read stays unconditional and the code straight-line.
        const uint32_t mo = 0u - (uint32_t)outside, mc = 0u - (uint32_t)(c == 0), mb = 0u - (uint32_t)(nb < n);
        const uint32_t oval = (mc & pv) | (~mc & (REFB | (8 * m)));
        const uint32_t ival = (mb & (REFB | nb)) | (~mb & nv);
        const uint32_t rval = (mo & oval) | (~mo & ival);
        // palette ops (about 1 % of ops) are written below, in a loop over
        // only those nibbles; here they take a placeholder
        const uint32_t val = (op == FCS_OP_RP) ? pv : (op <= FCS_OP_RZ) ? rval : 0u;
        // status as a lane mask (scalar ors), folded in once per word
        e_op = e_op || (live && op > FCS_OP_PD);
        if constexpr (MODE == MODE_LEVEL) {
          *(live ? &io.vnext[n] : dum16) = (uint16_t)val;
          atomicOr(live ? &io.snext[n >> 5] : dum32, (live && (v & 1u)) ? (1u << (n & 31)) : 0u);
        } else {
          if (live) io.olev[n] = (LT)val;
        }
        k += isop ? 1u : 0u;
      }
      }
      {
        // Pa / P0 / Pd. Nibble masks with the flag in bit 4j: op nibbles,
        // Pa ops (nibble 8|9), all palette ops (nibbles 8..13)
        const uint32_t isoN = iso;
        const uint32_t b3 = word >> 3, b2 = word >> 2, b1 = word >> 1;
        const uint32_t paN = b3 & ~b2 & ~b1 & isoN;
        uint32_t pm = b3 & ~(b2 & b1) & isoN;
        while (pm != 0u) {
          const uint32_t j4 = __builtin_ctz(pm);  // 4 * nibble
          pm &= pm - 1u;
          const uint32_t below = (1u << j4) - 1u;
          const uint32_t kj = k0 + __popc(isoN & below);
          if (kj >= K) break;  // past the level's last op: so are the later ones
          const uint32_t op = (word >> (j4 + 1)) & 0x7u;
          // palette cursor: Pa ops up to this one; Pd: the delta nibble dn
          // after it selects cursor - dn - 1
          const uint32_t ipj = ip_start + __popc(paN & (below | (1u << j4)));
          const bool isd = (op == FCS_OP_PD);
          const uint32_t dn = ((j4 < 28u) ? (word >> (j4 + 4u)) : wnext) & 0xFu;
          const uint32_t pidx = ipj - (isd ? dn + 1u : 0u);  // wraps past P when the delta reaches before the palette
          e_pal = e_pal || pidx >= P;
          e_ovr = e_ovr || (isd && g0 + (j4 >> 2) + 1u >= ne);
          const uint32_t pval = min(pidx, P - 1);
          const uint32_t m = ((kj >> 3) != q0) ? m1 : m0;
          const uint32_t n = (m << 3) | (kj & 7u);
          if constexpr (MODE == MODE_LEVEL) {
            io.vnext[n] = (uint16_t)pval;
          } else if constexpr (MODE == MODE_COMPACT || MODE == MODE_LCOMPACT) {
            io.opv[kj] = (uint16_t)pval;
          } else {
            io.olev[n] = (LT)pval;
          }
        }
      }
      st |= (e_pal ? FCS_ST_PALETTE_RANGE : 0u) | (e_ovr ? FCS_ST_STREAM_OVERRUN : 0u) | (e_op ? FCS_ST_BAD_OP : 0u);
      if (k0 < K && k >= K) {
        // this lane holds the level's last op: the next level starts after it
        // (and its delta); the palette cursor there
        uint32_t kq = k0, ipq = ip_start;
        for (uint32_t j = 0; j < 8; ++j) {
          if (!((iso >> (4 * j)) & 1u)) continue;
          const uint32_t op = (word >> (4 * j + 1)) & 0x7u;
          ipq += (op == FCS_OP_PA) ? 1u : 0u;
          if (kq == K - 1) {
            S.scr[SCR_NS] = g0 + j + 1 + ((op == FCS_OP_PD) ? 1u : 0u);
            S.scr[SCR_IP] = ipq;
            break;
          }
          ++kq;
        }
      }
    }
    op_base += base & 0xFFFFu;
    pa_base += base >> 16;
    carry = X;
  }
  status |= st;
  return op_base;
}


// One active parent's 8 op results (8 consecutive ops = its children in
// order, packed two per word in w): the op nibbles left by parse_level (bit 15
// set) become values or node references, palette results stay. Rx/Ry/Rz name
// the neighbour across the face of the child's own 2x2x2 block: children with
// the axis bit clear reference the matching child of the parent's previous
// neighbour along that axis, children with it set take the value of the
// parent's next neighbour (that level is not decoded yet). Outside the brick
// the reference reads the parent's output slot, which child 0 has already
// overwritten when c > 0. The neighbours are worked out once per parent; a
// reference into a stopped parent takes that parent's value at once (its
// children all copy it), and one to child 0 takes child 0's result.
__device__ __forceinline__ void resolve_group(uint32_t (&w)[4], uint32_t m, uint32_t pv, const uint16_t* vcur,
                                              const uint32_t* scur, uint32_t pmask) {
  uint32_t prev[3], next[3];
  bool first[3], last[3];
#pragma unroll
  for (uint32_t ax = 0; ax < 3; ++ax) {
    // (the neighbours wrap inside the level at the faces, where they are unused)
    const uint32_t mA = (0x09249249u << ax) & pmask;
    const uint32_t xa = m & mA, rest = m & ~mA;
    const uint32_t mp = ((xa - 1u) & mA) | rest;
    first[ax] = xa == 0u;
    last[ax] = xa == mA;
    const bool pstop = (scur[mp >> 5] >> (mp & 31)) & 1u;
    const uint32_t pval = vcur[mp];
    prev[ax] = pstop ? pval : (FCS_REF16 | (8u * mp));
    next[ax] = vcur[(((xa | ~mA) + 1u) & mA) | rest];
  }
  uint32_t v0 = pv;  // child 0's result
#pragma unroll
  for (uint32_t c = 0; c < 8; ++c) {
    const uint32_t sh = 16 * (c & 1);
    const uint32_t x = (w[c >> 1] >> sh) & 0xFFFFu;
    const uint32_t oval = (c == 0) ? pv : v0;
    uint32_t cand[3];
#pragma unroll
    for (uint32_t ax = 0; ax < 3; ++ax) {
      const uint32_t pb = (prev[ax] & FCS_REF16) ? (prev[ax] | (c | (1u << ax))) : prev[ax];
      cand[ax] = ((c >> ax) & 1u) ? (last[ax] ? oval : next[ax]) : (first[ax] ? oval : pb);
    }
    const uint32_t op = (x >> 1) & 0x7u;
    const uint32_t lo = (op & 1u) ? cand[0] : pv;       // Rx : Rp
    const uint32_t hi = (op & 1u) ? cand[2] : cand[1];  // Rz : Ry
    const uint32_t r = (op & 4u) ? 0u : ((op & 2u) ? hi : lo);  // corrupt op nibbles: 0
    const uint32_t val = (x & FCS_REF16) ? r : x;
    if (c == 0) v0 = val;
    w[c >> 1] = (w[c >> 1] & ~(0xFFFFu << sh)) | (val << sh);
  }
}

template <typename LT, class C>
__device__ __forceinline__ void act_entry(const Lds<LT>& S, const uint16_t* vcur, uint32_t gp, uint32_t& m,
                                          uint32_t& pv) {
  if constexpr (C::PACKED) {
    const uint32_t e = S.act32[gp];
    m = e & 0xFFFFu;
    pv = e >> 16;
  } else {
    m = S.act16[gp];
    pv = vcur[m];
  }
}

// Final level: each active parent's group of 8 op results is resolved in place.
template <typename LT, class C>
__device__ __forceinline__ void resolve_compact(const Lds<LT>& S, uint16_t* opv, const uint16_t* vcur,
                                                const uint32_t* scur, uint32_t A, uint32_t l) {
  const uint32_t pmask = (1u << (3 * l)) - 1u;  // Morton codes of the parent level
  for (uint32_t gp = threadIdx.x; gp < A; gp += C::NT) {
    const uint4 q = *reinterpret_cast<const uint4*>(opv + 8 * gp);
    if (!((q.x | q.y | q.z | q.w) & 0x80008000u)) continue;  // palette results only
    uint32_t m, pv;
    act_entry<LT, C>(S, vcur, gp, m, pv);
    uint32_t w[4] = {q.x, q.y, q.z, q.w};
    resolve_group(w, m, pv, vcur, scur, pmask);
    *reinterpret_cast<uint4*>(opv + 8 * gp) = make_uint4(w[0], w[1], w[2], w[3]);
  }
}

// Non-final level parsed in compact form: each active parent's 8 resolved
// results go to its children in the next level's node array (one 16-byte
// store), and their stop bits (op order = child order) to the children's
// byte of the next level's stop words, which the previous level preset to 0
// for children of active parents.
template <typename LT, class C>
__device__ __forceinline__ void resolve_level(const Lds<LT>& S, const uint16_t* opv, const uint16_t* vcur,
                                              const uint32_t* scur, uint16_t* vnext, uint32_t* snext, uint32_t A,
                                              uint32_t l) {
  const uint32_t pmask = (1u << (3 * l)) - 1u;
  for (uint32_t gp = threadIdx.x; gp < A; gp += C::NT) {
    const uint4 q = *reinterpret_cast<const uint4*>(opv + 8 * gp);
    uint32_t m, pv;
    act_entry<LT, C>(S, vcur, gp, m, pv);
    uint32_t w[4] = {q.x, q.y, q.z, q.w};
    if ((q.x | q.y | q.z | q.w) & 0x80008000u) resolve_group(w, m, pv, vcur, scur, pmask);
    *reinterpret_cast<uint4*>(vnext + 8 * m) = make_uint4(w[0], w[1], w[2], w[3]);
    reinterpret_cast<uint8_t*>(snext)[m] = (uint8_t)(S.kstop[gp >> 2] >> (8 * (gp & 3u)));
  }
}

// Per-request state carried from level to level.
template <typename LT>
struct LevelVars {
  uint16_t* vcur;    // values of the current level (u16 palette indices)
  uint32_t* scur;    // stop bits of the current level
  uint32_t* snext;   // stop bits of the next level
  uint32_t A;        // active (not stopped) nodes of the current level
  uint32_t ns;       // nibble where the current level's ops start
  uint32_t ip0;      // palette cursor at that nibble
  uint32_t tile;     // parse tiles so far (record double-buffering)
};

// Per-request constants of the level loop.
template <typename LT>
struct LevelConst {
  const uint32_t* nibw;
  uint32_t nsym, w_end, t, L, P;
  const LT* pal;   // palette in global memory
  const LT* lp;    // palette as read by the output (LDS when staged)
  LT* o;           // request output
  bool pyr, staged;
};

// Every level from lv+1 to the target (or only the target) copies level lv.
template <typename LT, class C>
__device__ __forceinline__ void expand_rest(const Lds<LT>& S, const LevelConst<LT>& c, const uint16_t* vl,
                                            uint32_t lv) {
  for (uint32_t lt = c.pyr ? lv + 1 : c.t; lt <= c.t; ++lt) {
    LT* ol = c.o + (c.pyr ? fcs_pyramid_level_offset(lt) : 0ull);
    if (c.staged) expand_level<LT, C::NT>(ol, lt, lv, vl, S.spal, c.P);
    else expand_level<LT, C::NT>(ol, lt, lv, vl, c.pal, c.P);
  }
}

// One level step (level l -> l+1) run by the threads of CS: the workgroup
// (CS = C) or its first wave (CS = OneWave<C>, non-final levels only, the
// end-of-brick expansion left to the caller). Three phases: stopped
// children and the parse/resolve of the level's ops (P1), reference chasing
// (P2), the next level's active list or the labels (P3). Returns true when
// the brick ends at this step (every child stopped).
template <typename LT, class C, class CS>
__device__ __forceinline__ bool level_step(const OpsArgs<LT>& a, const Lds<LT>& S, LevelVars<LT>& v,
                                           const LevelConst<LT>& c, Blocks& blk, uint32_t l, uint32_t& status) {
  constexpr uint32_t NT = CS::NT, NW = CS::NW;
  constexpr bool ONE = NW == 1;
  const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  const uint32_t t = c.t, P = c.P, nsym = c.nsym;
  // Wave priorities among the CU's workgroups rise toward the label stores:
  // 0 for the parse, 1 from the reference chase on, 2 while labels are
  // written, so the stores keep flowing and a workgroup near its labels goes
  // ahead of one that is parsing. On the 2048^3 decode call: 10.88-10.91 ms
  // with one priority, 10.78-10.80 with 0 / 2, 10.68 with 0 / 1 / 2; the
  // reverse (level steps above labels) measured 11.02.
  if constexpr (!ONE) __builtin_amdgcn_s_setprio(0);
  const uint32_t N = 1u << (3 * l), Cn = N << 3, K = 8 * v.A;
  const bool last = (l + 1 == t);
  uint32_t mode = MODE_LEVEL;
  if constexpr (!ONE) mode = !last ? MODE_LEVEL : ((C::KCAP != 0 && K <= C::KCAP) ? MODE_COMPACT : MODE_GLOBAL);
  // A level decoded by one wave whose ops fit the area between the ping-pong
  // buffers is parsed in compact form and resolved per active parent
  // (MODE_LCOMPACT): fewer instructions than the per-nibble resolve, and one
  // wave has no barrier to pay for the extra phase. (A workgroup's level
  // step measured the same either way and stays per nibble.)
  bool lcomp = false;
  if constexpr (ONE && C::XCAP != 0) lcomp = !last && K <= C::XCAP && Cn <= C::VCAP;
  LT* olev = c.o + (c.pyr ? fcs_pyramid_level_offset(l + 1) : 0ull);
  uint16_t* vcur = v.vcur;
  uint16_t* vnext = (Cn <= C::VCAP) ? (vcur == S.sA ? S.sB : S.sA) : S.big;
  // compact op results: the free ping-pong buffer and the area next to it
  // (for a non-final level, the area alone: both buffers hold levels)
  uint16_t* opv = (lcomp || vcur == S.sA) ? S.sA + C::VCAP : S.sA;
  uint32_t* scur = v.scur;
  uint32_t* snext = v.snext;
  const uint32_t A = v.A;

  // ---- P1: stopped children; parse and resolve the level's ops
  if (mode == MODE_LEVEL) {
    for (uint32_t m = tid; m < N; m += NT)
      if (is_stopped(scur, m)) {
        const uint32_t x = vcur[m] | ((uint32_t)vcur[m] << 16);
        *reinterpret_cast<uint4*>(vnext + 8 * m) = make_uint4(x, x, x, x);
      }
    if (lcomp)
      for (uint32_t i = tid; i < (K + 31) / 32; i += NT) S.kstop[i] = 0u;  // ordered by the parse's first sync
  }
  LevelIO<LT> io{c.nibw, v.ns, nsym, c.w_end, K, v.ip0, P, (1u << (3 * (l + 1))) - 1u, vcur, vnext, snext, opv, olev,
                 S.kstop};
  uint32_t found;
  if constexpr (ONE) {
    if (lcomp) found = parse_level<MODE_LCOMPACT, LT, CS>(S, io, blk, v.tile, status);
    else found = parse_level<MODE_LEVEL, LT, CS>(S, io, blk, v.tile, status);
  } else {
    if (mode == MODE_LEVEL) found = parse_level<MODE_LEVEL, LT, CS>(S, io, blk, v.tile, status);
    else if (mode == MODE_COMPACT) found = parse_level<MODE_COMPACT, LT, CS>(S, io, blk, v.tile, status);
    else found = parse_level<MODE_GLOBAL, LT, CS>(S, io, blk, v.tile, status);
  }
  if (found < K) status |= FCS_ST_STREAM_OVERRUN | FCS_ST_NSYM_MISMATCH;
  if (mode == MODE_GLOBAL) __syncthreads(); else group_sync<CS>();  // per-node results in the cache region
  v.ns = S.scr[SCR_NS];
  v.ip0 = S.scr[SCR_IP];
  if (l + 1 == c.L && v.ns != nsym) status |= FCS_ST_NSYM_MISMATCH;  // nibbles after the finest level
  if constexpr (!ONE) {
    if (mode == MODE_COMPACT) {
      resolve_compact<LT, C>(S, opv, vcur, scur, A, l);
      group_sync<CS>();
    }
  } else {
    if (lcomp) {
      resolve_level<LT, CS>(S, opv, vcur, scur, vnext, snext, A, l);
      group_sync<CS>();
    }
  }

  if constexpr (!ONE) __builtin_amdgcn_s_setprio(1);
  // ---- P2: chase references. A valid reference points to a smaller node,
  // so a chain strictly decreases; anything else is corrupt input. Only op
  // children can hold references.
  auto act_node = [&](uint32_t k) -> uint32_t {
    if constexpr (C::PACKED) return S.act32[k >> 3] & 0xFFFFu; else return S.act16[k >> 3];
  };
  uint32_t cnt = 0, inc = 0, w0 = 0, w1 = 0, mine = 0;
  const uint32_t nw = (Cn + 31) >> 5;
  if (mode == MODE_LEVEL) {
    // one active parent's 8 children per step: one 16-byte LDS read, the
    // references chased from registers, one 16-byte write back (a chase
    // reading a group another lane rewrites sees either of two values that
    // lead to the same label)
    for (uint32_t gp = tid; gp < A; gp += NT) {
      const uint32_t n0 = 8 * act_node(8 * gp);
      const uint4 q = *reinterpret_cast<const uint4*>(vnext + n0);
      if ((q.x | q.y | q.z | q.w) & 0x80008000u) {
        uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (uint32_t ch = 0; ch < 8; ++ch) {
          const uint32_t sh = 16 * (ch & 1);
          uint32_t x = (w[ch >> 1] >> sh) & 0xFFFFu;
          if (x & FCS_REF16) {
            uint32_t cur = n0 + ch;
            do {
              const uint32_t nx = x & 0x7FFFu;
              if (nx >= cur) { status |= FCS_ST_NSYM_MISMATCH; x = 0; break; }
              cur = nx;
              x = vnext[nx];
            } while (x & FCS_REF16);
            w[ch >> 1] = (w[ch >> 1] & ~(0xFFFFu << sh)) | (x << sh);
          }
        }
        *reinterpret_cast<uint4*>(vnext + n0) = make_uint4(w[0], w[1], w[2], w[3]);
      }
    }
    // active nodes of level l+1: each thread takes a span of Cn / NT
    // consecutive nodes (within one stop word; whole words per thread when a
    // level has more than 32 nodes per thread)
    const uint32_t valid = (Cn >= 32) ? 0xFFFFFFFFu : ((1u << Cn) - 1u);
    const uint32_t span = Cn >= NT ? Cn / NT : 1u;
    if (span <= 32u) {
      const uint32_t bit0 = tid * span;
      if (bit0 < Cn) {
        const uint32_t sw = ~snext[bit0 >> 5] & valid;
        mine = span == 32u ? sw : (sw >> (bit0 & 31u)) & ((1u << span) - 1u);
      }
      cnt = __popc(mine);
    } else {
      const uint32_t wpt = (nw + NT - 1) / NT;
      w0 = min(nw, tid * wpt);
      w1 = min(nw, w0 + wpt);
      for (uint32_t w = w0; w < w1; ++w) cnt += __popc(~snext[w] & valid);
    }
    inc = wave_incl_scan(cnt);
    if (lane == 63) S.scr[SCR_WAVESUM + wave] = inc;
  } else if (mode == MODE_COMPACT) {
    if constexpr (!ONE) {
      // one active parent's 8 op results per step: one 16-byte LDS read, the
      // references among them chased from registers, one 16-byte write back.
      // A chase may read a group another lane is rewriting: either value
      // (the reference or what it resolves to) leads to the same label.
      for (uint32_t gp = tid; gp < A; gp += NT) {
        const uint4 q = *reinterpret_cast<const uint4*>(opv + 8 * gp);
        if ((q.x | q.y | q.z | q.w) & 0x80008000u) {
          const uint32_t n0 = 8 * act_node(8 * gp);
          uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
          for (uint32_t ch = 0; ch < 8; ++ch) {
            const uint32_t sh = 16 * (ch & 1);
            uint32_t x = (w[ch >> 1] >> sh) & 0xFFFFu;
            if (x & FCS_REF16) {
              uint32_t cur = n0 + ch;
              for (;;) {
                const uint32_t nb = x & 0x7FFFu;
                if (nb >= cur) { status |= FCS_ST_NSYM_MISMATCH; x = 0; break; }
                const uint32_t p = nb >> 3;
                const uint32_t sw = scur[p >> 5];
                const uint32_t wp = S.wpre[p >> 5];
                if ((sw >> (p & 31)) & 1u) { x = vcur[p]; break; }
                const uint32_t rk = wp + __popc(~sw & ((1u << (p & 31)) - 1u));
                x = opv[8 * rk + (nb & 7u)];
                if (!(x & FCS_REF16)) break;
                cur = nb;
              }
              w[ch >> 1] = (w[ch >> 1] & ~(0xFFFFu << sh)) | (x << sh);
            }
          }
          *reinterpret_cast<uint4*>(opv + 8 * gp) = make_uint4(w[0], w[1], w[2], w[3]);
        }
      }
    }
  } else {
    if constexpr (!ONE) {
      for (uint32_t k = tid; k < K; k += NT) {
        const uint32_t n = 8 * act_node(k) + (k & 7u);
        uint32_t x = (uint32_t)olev[n];
        if (x & FCS_REF32) {
          uint32_t cur = n;
          for (;;) {
            const uint32_t nb = x & ~FCS_REF32;
            if (nb >= cur) { status |= FCS_ST_NSYM_MISMATCH; x = 0; break; }
            const uint32_t p = nb >> 3;
            if (is_stopped(scur, p)) { x = vcur[p]; break; }
            x = (uint32_t)olev[nb];
            if (!(x & FCS_REF32)) break;
            cur = nb;
          }
          olev[n] = (LT)x;
        }
      }
    }
  }
  if (mode == MODE_GLOBAL) __syncthreads(); else group_sync<CS>();

  // ---- P3: next level's active list; labels out
  bool ended = false;
  if (mode == MODE_LEVEL) {
    uint32_t An = 0, wb = 0;
#pragma unroll
    for (uint32_t w = 0; w < NW; ++w) {
      const uint32_t sw = S.scr[SCR_WAVESUM + w];
      wb += (w < wave) ? sw : 0u;
      An += sw;
    }
    if (c.pyr)
      for (uint32_t n = tid; n < Cn; n += NT) {
        const uint32_t x = vnext[n];
        olev[n] = c.lp[min(x, P - 1)];
      }
    if (An == 0) {
      // every node of level l+1 is stopped: the remaining levels are copies
      // (the first wave alone leaves them to the whole workgroup)
      if constexpr (!ONE) {
        expand_rest<LT, CS>(S, c, vnext, l + 1);
        if (t == c.L && v.ns != nsym) status |= FCS_ST_NSYM_MISMATCH;  // nibbles after the finest level
      }
      ended = true;
    } else {
      uint32_t k0 = wb + inc - cnt;
      const uint32_t valid = (Cn >= 32) ? 0xFFFFFFFFu : ((1u << Cn) - 1u);
      const uint32_t span = Cn >= NT ? Cn / NT : 1u;
      if (span <= 32u) {
        // the thread whose span opens a stop word records the count before it
        const uint32_t bit0 = tid * span;
        if (bit0 < Cn && (bit0 & 31u) == 0u) S.wpre[bit0 >> 5] = (uint16_t)k0;
        uint32_t bits = mine;
        while (bits) {
          const uint32_t j = __ffs(bits) - 1;
          bits &= bits - 1;
          const uint32_t m = bit0 + j;
          if constexpr (C::PACKED) S.act32[k0++] = m | ((uint32_t)vnext[m] << 16);
          else S.act16[k0++] = (uint16_t)m;
        }
      } else {
        for (uint32_t w = w0; w < w1; ++w) {
          S.wpre[w] = (uint16_t)k0;
          uint32_t bits = ~snext[w] & valid;
          while (bits) {
            const uint32_t j = __ffs(bits) - 1;
            bits &= bits - 1;
            const uint32_t m = w * 32 + j;
            if constexpr (C::PACKED) S.act32[k0++] = m | ((uint32_t)vnext[m] << 16);
            else S.act16[k0++] = (uint16_t)m;
          }
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
      if (tid == 0) { S.scr[SCR_NS] = v.ns; S.scr[SCR_IP] = v.ip0; }
    }
    v.vcur = vnext;
    v.scur = snext;
    v.snext = scur;
    v.A = An;
  } else if (mode == MODE_COMPACT) {
    if constexpr (!ONE) {
      __builtin_amdgcn_s_setprio(2);
      if (c.staged) out_compact<LT, NT>(olev, N, A, scur, S.wpre, vcur, opv, S.spal, P);
      else out_compact<LT, NT>(olev, N, A, scur, S.wpre, vcur, opv, c.pal, P);
    }
  } else {
    if constexpr (!ONE) {
      __builtin_amdgcn_s_setprio(2);
      if (c.staged) out_global<LT, NT>(olev, N, scur, vcur, S.spal, P);
      else out_global<LT, NT>(olev, N, scur, vcur, c.pal, P);
    }
  }
  group_sync<CS>();
  (void)a;
  return ended;
}

// Level state after the first levels of a request, as the head pass leaves it
// in the request's output region (the finest level's slot of a pyramid, which
// only the request's last level writes). u32 words:
//   0: HEAD_MAGIC | level l | ended << 4   1: active nodes A   2: next nibble
//   3: palette cursor   4: status bits   5-7: unused
//   then level l's values (u16 pairs), its stop bits, the next level's stop
//   bits preset from stopped parents (when that level is not the last), the
//   active counts before each stop word (u16 pairs), the active list
//   (node | value << 16). At most 928 words (level 3 of a 32^3 request),
//   within every output the head pass writes to (8^t labels, t >= 2).
constexpr uint32_t HEAD_MAGIC = 0xC5D00000u;
constexpr uint32_t HEAD_MAX_LEVEL = 3;
constexpr uint32_t HEAD_WORDS = 928;

struct HeadLayout {
  uint32_t N, nvw, nsw, npw, nww;  // nodes; words of values, stop bits, preset stop bits, rank prefixes
  uint32_t v, st, ps, wp, act;     // word offsets
};

__device__ __forceinline__ HeadLayout head_layout(uint32_t l, uint32_t t) {
  HeadLayout h;
  h.N = 1u << (3 * l);
  h.nvw = (h.N + 1) / 2;
  h.nsw = (h.N + 31) / 32;
  h.npw = (l + 1 < t) ? (8 * h.N + 31) / 32 : 0u;
  h.nww = (h.nsw + 1) / 2;
  h.v = 8;
  h.st = h.v + h.nvw;
  h.ps = h.st + h.nsw;
  h.wp = h.ps + h.npw;
  h.act = h.wp + h.nww;
  return h;
}

template <typename LT>
__device__ __forceinline__ uint32_t* head_state(LT* o, uint32_t t, bool pyr) {
  return reinterpret_cast<uint32_t*>(o + (pyr ? fcs_pyramid_level_offset(t) : 0ull));
}

// `commit` is called once on every path, as soon as the request's first
// global loads have been waited for: the caller stores its prefetched ring
// entries there, so that their wait does not queue behind this request's
// label stores (global loads and stores complete in issue order).
template <typename LT, class C, class Commit>
__device__ void decode_request(const OpsArgs<LT>& a, const Lds<LT>& S, uint32_t r, const DecodeReq& q,
                               const DevBrick& B, Commit&& commit) {
  constexpr uint32_t NT = C::NT, NW = C::NW;
  constexpr uint32_t PCAP = Layout<C, LT>::PCAP;
  constexpr uint32_t PPT = PCAP ? (PCAP + NT - 1) / NT : 1;  // staged palette entries per thread
  constexpr uint32_t HW = (HEAD_WORDS + NT - 1) / NT;        // head state words per thread
  const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  const uint32_t b = q.brick, t = q.lod & 0xFFu;
  __builtin_amdgcn_s_setprio(2);  // start-up loads and constant fills: label-writing priority
  const bool pyr = (q.lod & FCS_REQ_PYRAMID) != 0;
  const uint32_t L = a.L;
  if (b >= a.nbricks || t > L) {
    commit();
    if (tid == 0) a.status[r] = FCS_ST_BAD_REQUEST;
    return;
  }
  uint32_t status = 0;
  const uint32_t nsym = B.nsym;
  if (nsym == 0) status |= FCS_ST_STREAM_OVERRUN;
  const uint32_t P = B.P;
  const LT* pal = a.palette + B.pal_off;
  LT* o = a.out + q.out_off;
  if (P == 0 || P > FCS_PALETTE_MAX_LDS) {
    commit();
    if (tid == 0) a.status[r] = status | (P == 0 ? FCS_ST_PALETTE_RANGE : FCS_ST_BAD_REQUEST);
    return;
  }
  // Every start-up load is issued here, before any wait: the nibble blocks,
  // the root, the staged palette and (for streams longer than the root) the
  // head pass state, read whole from the request's output.
  const uint32_t* nibw = a.nib + B.nib_off;
  const uint32_t w_end = (nsym + 7) >> 3;
  Blocks blk;
  blk.bc = 0;
  load_block<NT>(nibw, w_end, 0, blk.w0, blk.n0);
  load_block<NT>(nibw, w_end, 1, blk.w1, blk.n1);
  const uint32_t root = nibw[0] & 0xFu;
  const LT p0 = pal[0];
  const bool staged = PCAP != 0 && P <= PCAP;
  LT pv[PPT];
#pragma unroll
  for (uint32_t k = 0; k < PPT; ++k) {
    const uint32_t j = tid + k * NT;
    pv[k] = (staged && j < P) ? pal[j] : LT(0);
  }
  const bool use_head = a.head && t >= 2;
  const uint32_t* hs = head_state(o, t, pyr);
  const uint32_t hn = min(HEAD_WORDS, (uint32_t)((sizeof(LT) / 4) << (3 * t)));
  uint32_t hw[HW];
  const bool pre = use_head && nsym > 1;
#pragma unroll
  for (uint32_t k = 0; k < HW; ++k) {
    const uint32_t j = tid + k * NT;
    hw[k] = (pre && j < hn) ? hs[j] : 0u;
  }
  commit();
  if ((root & 1u) || t == 0 || (status & FCS_ST_STREAM_OVERRUN)) {
    // constant brick (root stop bit) or LOD 0: every voxel = palette[0]
    fill_const<LT, NT>(o, pyr ? fcs_pyramid_level_offset(t + 1) : (1ull << (3 * t)), p0);
    if (tid == 0) a.status[r] = status;
    return;
  }
  if (use_head && !pre) {
    // (a one-nibble stream whose root is not stopped: corrupt, rare)
#pragma unroll
    for (uint32_t k = 0; k < HW; ++k) {
      const uint32_t j = tid + k * NT;
      hw[k] = j < hn ? hs[j] : 0u;
    }
  }
  if (staged) {
#pragma unroll
    for (uint32_t k = 0; k < PPT; ++k) {
      const uint32_t j = tid + k * NT;
      if (j < P) S.spal[j] = pv[k];
    }
  }
  const LT* lp = staged ? static_cast<const LT*>(S.spal) : pal;
  if (pyr && tid == 0) o[0] = p0;
  if (tid == 0) {
    // level 0: the root (palette index 0, active); level 1 not stopped yet
    S.sA[0] = 0;
    if constexpr (C::PACKED) S.act32[0] = 0; else S.act16[0] = 0;
    S.wpre[0] = 0;
    S.st0[0] = 0;
    S.st1[0] = 0;
    S.scr[SCR_STATUS] = 0;
    S.scr[SCR_NS] = 1;
    S.scr[SCR_IP] = 0;
  }
  if (use_head && tid < 8) S.scr[SCR_HEAD + tid] = hw[0];
  lds_barrier();

  // level 0 state: the root (palette index 0) is the one active node
  LevelVars<LT> lv{S.sA, S.st0, S.st1, 1u, 1u, 0u, 0u};
  LevelConst<LT> lc{nibw, nsym, w_end, t, L, P, pal, lp, o, pyr, staged};
  uint32_t l = 0, ended_at = 0;
  if (use_head) {
    // the head pass decoded the first levels: its state, held by the
    // threads in word order, goes to the LDS arrays
    const uint32_t h0 = S.scr[SCR_HEAD];
    if ((h0 & 0xFFFF0000u) != HEAD_MAGIC || (h0 & 0xFu) > HEAD_MAX_LEVEL || (h0 & 0xFu) >= t ||
        ((h0 >> 4) & 0xFu) > (h0 & 0xFu)) {
      if (tid == 0) a.status[r] = status | FCS_ST_BAD_REQUEST;
      return;
    }
    l = h0 & 0xFu;
    ended_at = (h0 >> 4) & 0xFu;
    const HeadLayout h = head_layout(l, t);
    lv.A = min(S.scr[SCR_HEAD + 1], h.N);
    lv.ns = S.scr[SCR_HEAD + 2];
    lv.ip0 = S.scr[SCR_HEAD + 3];
    status |= S.scr[SCR_HEAD + 4];
    uint32_t* vc = reinterpret_cast<uint32_t*>((l & 1u) ? S.sB : S.sA);
    uint32_t* sc = (l & 1u) ? S.st1 : S.st0;
    uint32_t* sn = (l & 1u) ? S.st0 : S.st1;
    uint32_t* wp = reinterpret_cast<uint32_t*>(S.wpre);
    const uint32_t end = min(h.act + lv.A, hn);
#pragma unroll
    for (uint32_t k = 0; k < HW; ++k) {
      const uint32_t j = tid + k * NT, w = hw[k];
      if (j < h.v || j >= end) continue;
      if (j < h.st) vc[j - h.v] = w;
      else if (j < h.ps) sc[j - h.st] = w;
      else if (j < h.wp) sn[j - h.ps] = w;
      else if (j < h.act) wp[j - h.wp] = w;
      else if constexpr (C::PACKED) S.act32[j - h.act] = w;
      else S.act16[j - h.act] = (uint16_t)w;
    }
    if (tid == 0) {
      S.scr[SCR_NS] = lv.ns;
      S.scr[SCR_IP] = lv.ip0;
    }
    lds_barrier();
  } else {
    // Non-final levels whose children fit one wave's tiles (8, 64, 512) are
    // decoded by the first wave alone, with wave-level syncs; the other waves
    // wait at the handoff barrier. A level whose children are all stopped ends
    // the brick there; its expansion then runs with the whole workgroup.
    if (wave == 0) {
      Blocks sb;
      sb.bc = 0;
      load_block<64>(nibw, w_end, 0, sb.w0, sb.n0);
      load_block<64>(nibw, w_end, 1, sb.w1, sb.n1);
      for (; l + 1 < t && l < 3; ++l) {
        if (level_step<LT, C, OneWave<C>>(a, S, lv, lc, sb, l, status)) {
          ended_at = l + 1;
          ++l;
          break;
        }
      }
      if (tid == 0) {
        S.scr[SCR_HAND + 0] = lv.A;
        S.scr[SCR_HAND + 1] = lv.ns;
        S.scr[SCR_HAND + 2] = lv.ip0;
        S.scr[SCR_HAND + 3] = l;
        S.scr[SCR_HAND + 4] = ended_at;
      }
    }
    lds_barrier();
    l = S.scr[SCR_HAND + 3];
    ended_at = S.scr[SCR_HAND + 4];
    lv.A = S.scr[SCR_HAND + 0];
    lv.ns = S.scr[SCR_HAND + 1];
    lv.ip0 = S.scr[SCR_HAND + 2];
  }
  lv.tile = 0;
  // after l levels the ping-pong buffers have swapped l times
  lv.vcur = (l & 1u) ? S.sB : S.sA;
  lv.scur = (l & 1u) ? S.st1 : S.st0;
  lv.snext = (l & 1u) ? S.st0 : S.st1;
  if (ended_at) {
    // every node of level ended_at is stopped: the remaining levels are copies
    expand_rest<LT, C>(S, lc, lv.vcur, ended_at);
    if (t == L && lv.ns != nsym) status |= FCS_ST_NSYM_MISMATCH;  // nibbles after the finest level
  } else {
    for (; l < t; ++l)
      if (level_step<LT, C, C>(a, S, lv, lc, blk, l, status)) break;
  }
  if (status) atomicOr(&S.scr[SCR_STATUS], status);
  lds_barrier();
  if (tid == 0) a.status[r] = S.scr[SCR_STATUS];
}

// Head pass: one wave per request decodes the non-final levels up to level 3
// (8, 64 and 512 nodes) and leaves the level state in the request's output
// (head_state). These levels are short chains of dependent steps; run here
// at a wave's LDS footprint, many requests overlap on every CU, where inside
// the op kernel they would hold a whole workgroup's LDS. Requests the op
// kernel rejects or finishes without levels (constant bricks, t < 2) are
// skipped with the same tests.
template <typename LT>
__global__ __launch_bounds__(64) void ops_head_kernel(OpsArgs<LT> a) {
  using C = CfgHead;
  extern __shared__ __align__(16) uint8_t smem[];
  const Lds<LT> S = make_lds<C, LT>(smem);
  const uint32_t tid = threadIdx.x;
  const DecodeReq q = a.reqs[blockIdx.x];
  const uint32_t b = q.brick, t = q.lod & 0xFFu;
  const bool pyr = (q.lod & FCS_REQ_PYRAMID) != 0;
  if (b >= a.nbricks || t > a.L || t < 2) return;
  const DevBrick B = a.bricks[b];
  const uint32_t nsym = B.nsym, P = B.P;
  if (nsym == 0 || P == 0 || P > FCS_PALETTE_MAX_LDS) return;
  const uint32_t* nibw = a.nib + B.nib_off;
  const uint32_t w_end = (nsym + 7) >> 3;
  Blocks sb;
  sb.bc = 0;
  load_block<64>(nibw, w_end, 0, sb.w0, sb.n0);
  load_block<64>(nibw, w_end, 1, sb.w1, sb.n1);
  if (nibw[0] & 1u) return;  // constant brick
  const LT* pal = a.palette + B.pal_off;
  LT* o = a.out + q.out_off;
  if (tid == 0) {
    S.sA[0] = 0;
    S.act32[0] = 0;
    S.wpre[0] = 0;
    S.st0[0] = 0;
    S.st1[0] = 0;
    S.scr[SCR_STATUS] = 0;
    S.scr[SCR_NS] = 1;
    S.scr[SCR_IP] = 0;
  }
  group_sync<C>();
  LevelVars<LT> lv{S.sA, S.st0, S.st1, 1u, 1u, 0u, 0u};
  LevelConst<LT> lc{nibw, nsym, w_end, t, a.L, P, pal, pal, o, pyr, false};
  uint32_t status = 0, l = 0, ended_at = 0;
  for (; l + 1 < t && l < HEAD_MAX_LEVEL; ++l) {
    if (level_step<LT, C, C>(a, S, lv, lc, sb, l, status)) {
      ended_at = l + 1;
      ++l;
      break;
    }
  }
  if (status) atomicOr(&S.scr[SCR_STATUS], status);
  group_sync<C>();
  uint32_t* hs = head_state(o, t, pyr);
  const HeadLayout h = head_layout(l, t);
  if (tid == 0) {
    hs[0] = HEAD_MAGIC | l | (ended_at << 4);
    hs[1] = lv.A;
    hs[2] = lv.ns;
    hs[3] = lv.ip0;
    hs[4] = S.scr[SCR_STATUS];
  }
  const uint32_t* vc = reinterpret_cast<const uint32_t*>(lv.vcur);
  const uint32_t* wp = reinterpret_cast<const uint32_t*>(S.wpre);
  for (uint32_t i = tid; i < h.nvw; i += 64) hs[h.v + i] = vc[i];
  for (uint32_t i = tid; i < h.nsw; i += 64) hs[h.st + i] = lv.scur[i];
  for (uint32_t i = tid; i < h.npw; i += 64) hs[h.ps + i] = lv.snext[i];
  for (uint32_t i = tid; i < h.nww; i += 64) hs[h.wp + i] = wp[i];
  for (uint32_t i = tid; i < lv.A; i += 64) hs[h.act + i] = S.act32[i];
}

// Persistent workgroups drain the request list, which the host orders by
// estimated cost, heaviest first. Claims are numbered: the first three quarters are dealt
// round-robin (equal shares, since neighbouring requests cost about the same);
// the rest is claimed through one atomic counter to even out the finishing
// times. (One counter shared by every workgroup serialises: claiming every
// request through it caps a launch of cheap requests at the counter's atomic
// rate.) A launch of at most one request per workgroup -- a frame's few
// decodes, or none -- is dealt entirely, so its workgroups never touch the
// counter; one of fewer than two per workgroup deals the first round. Claims map to requests alternately from the heavy front and the light
// back of the list, one grid-sized chunk at a time, so every workgroup mixes
// compute-bound requests with cheap ones -- constant bricks are pure label
// stores, and all of them together at the end would leave the launch's tail
// bound by the write bandwidth with nothing to overlap. Request metadata is
// fetched ahead through 4-entry rings in LDS so that its latency overlaps the
// previous request's decode: while request i decodes, thread 0 takes the
// index of request i+3, loads the DecodeReq of request i+2 and the brick
// record of request i+1; they are stored to LDS as soon as the first loads of
// request i have arrived.
template <typename LT, class C>
__global__ __launch_bounds__(C::NT, C::MINW) void ops_kernel(OpsArgs<LT> a) {
  extern __shared__ __align__(16) uint8_t smem[];
  const Lds<LT> S = make_lds<C, LT>(smem);
  // a device-built queue carries its count in device memory (at most the
  // list's capacity, a.nreq)
  const uint32_t nreq = a.nreq_dev ? min(*a.nreq_dev, a.nreq) : a.nreq;
  const DecodeReq none{0xFFFFFFFFu, 0u, 0ull};
  const DevBrick nobrick{0ull, 0ull, 0u, 0u, 0u, 0u};
  const uint32_t G = gridDim.x;
  // the first three quarters of the claims are dealt round-robin (dealing a
  // half measured 1 % slower, dealing all 5 % slower: the claimed tail
  // evens out the workgroups' ends)
  const uint32_t n_static = nreq <= G ? nreq : max((nreq / 4 * 3) / G * G, G);
  uint32_t sj = 0;  // thread 0: round-robin requests taken so far
  // claim c -> request: chunk c / G even takes the next G requests from the
  // front, odd from the back (a bijection on [0, nreq))
  auto order = [&](uint32_t c) -> uint32_t {
    if (c >= nreq) return c;
    const uint32_t qc = c / G, pos = (qc >> 1) * G + (c - qc * G);
    return (qc & 1u) ? nreq - 1u - pos : pos;
  };
  auto next_index = [&]() -> uint32_t {
    const uint32_t rr = blockIdx.x + sj * G;
    if (rr < n_static) { ++sj; return order(rr); }
    if (n_static >= nreq) return nreq;  // everything was dealt
    return order(n_static + atomicAdd(a.counter, 1u));
  };
  auto fetch_req = [&](uint32_t i) -> DecodeReq { return i < nreq ? a.reqs[i] : none; };
  auto fetch_brick = [&](const DecodeReq& q) -> DevBrick { return q.brick < a.nbricks ? a.bricks[q.brick] : nobrick; };
  if (threadIdx.x == 0) {
    const uint32_t r0 = next_index();
    const uint32_t r1 = next_index();
    const uint32_t r2 = next_index();
    const DecodeReq q0 = fetch_req(r0), q1 = fetch_req(r1);
    const DevBrick b0 = fetch_brick(q0);
    S.scr[SCR_IDX + 0] = r0;
    S.scr[SCR_IDX + 1] = r1;
    S.scr[SCR_IDX + 2] = r2;
    S.rq[0] = q0;
    S.rq[1] = q1;
    S.rb[0] = b0;
  }
  lds_barrier();
  for (uint32_t i = 0;; ++i) {
    const uint32_t s0 = i & 3u;
    const uint32_t r = S.scr[SCR_IDX + s0];
    if (r >= nreq) break;
    const DecodeReq q = S.rq[s0];
    const DevBrick B = S.rb[s0];
    // prefetch (thread 0): queue index for i+3, request i+2, brick record i+1
    uint32_t r3 = 0;
    DecodeReq q2 = none;
    DevBrick b1 = nobrick;
    if (threadIdx.x == 0) {
      r3 = next_index();
      q2 = fetch_req(S.scr[SCR_IDX + ((i + 2) & 3u)]);
      b1 = fetch_brick(S.rq[(i + 1) & 3u]);
    }
    // the ring slots written here are not read before the next iteration
    auto commit = [&]() {
      if (threadIdx.x == 0) {
        S.scr[SCR_IDX + ((i + 3) & 3u)] = r3;
        S.rq[(i + 2) & 3u] = q2;
        S.rb[(i + 1) & 3u] = b1;
      }
    };
    decode_request<LT, C>(a, S, r, q, B, commit);
    lds_barrier();
  }
}

__global__ void zero_u32_kernel(uint32_t* p, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

template <typename LT, class C>
hipError_t occupancy(int* out) {
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
  *out = per_cu;
  return hipSuccess;
}

template <typename LT, class C>
hipError_t launch_cfg(const OpsArgs<LT>& a, uint32_t num_cu, hipStream_t s, uint32_t* grid_out, int head) {
  constexpr size_t lds = Layout<C, LT>::bytes;
  int per_cu = 0;
  if (hipError_t e = occupancy<LT, C>(&per_cu); e != hipSuccess) return e;
  const uint64_t resident = (uint64_t)per_cu * (num_cu ? num_cu : 1);
  const uint32_t grid = (uint32_t)(a.nreq < resident ? a.nreq : resident);
  if (grid_out) *grid_out = grid;
  OpsArgs<LT> b = a;
  // the head pass needs the request count at launch: never for a device-built queue
  b.head = !a.nreq_dev && (head == FCS_OPS_HEAD_ON || (head == FCS_OPS_HEAD_AUTO && a.nreq >= FCS_OPS_HEAD_MIN)) ? 1u : 0u;
  if (b.head) {
    constexpr size_t hl = Layout<CfgHead, LT>::bytes;
    hipLaunchKernelGGL(ops_head_kernel<LT>, dim3(a.nreq), dim3(64), hl, s, b);
  }
  hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, a.counter, (uint64_t)1);
  hipLaunchKernelGGL((ops_kernel<LT, C>), dim3(grid), dim3(C::NT), lds, s, b);
  return hipGetLastError();
}

}  // namespace

size_t fcs_ops_lds_bytes(uint32_t L) {
  return L >= 6 ? Layout<CfgLarge, uint64_t>::bytes : Layout<CfgSmall, uint64_t>::bytes;
}

template <typename LT>
hipError_t fcs_launch_ops(const OpsArgs<LT>& a, uint32_t num_cu, hipStream_t s, uint32_t* grid_out, int split,
                          int head) {
  if (a.nreq == 0) return hipSuccess;
  if (a.L > 6) return hipErrorInvalidValue;  // 16-bit node indices: bricks up to 64^3
  if (a.L == 6) return launch_cfg<LT, CfgLarge>(a, num_cu, s, grid_out, head);
  if (split == FCS_OPS_4WG) return launch_cfg<LT, CfgSmall4>(a, num_cu, s, grid_out, head);
  return launch_cfg<LT, CfgSmall>(a, num_cu, s, grid_out, head);
}

hipError_t fcs_launch_zero(uint32_t* p, uint64_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(zero_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

template <typename LT>
hipError_t fcs_ops_prepare(uint32_t L, int split) {
  int per_cu = 0;
  if (L == 6) return occupancy<LT, CfgLarge>(&per_cu);
  if (split == FCS_OPS_4WG) return occupancy<LT, CfgSmall4>(&per_cu);
  return occupancy<LT, CfgSmall>(&per_cu);
}

template hipError_t fcs_ops_prepare<uint32_t>(uint32_t, int);
template hipError_t fcs_ops_prepare<uint64_t>(uint32_t, int);
template hipError_t fcs_launch_ops<uint32_t>(const OpsArgs<uint32_t>&, uint32_t, hipStream_t, uint32_t*, int, int);
template hipError_t fcs_launch_ops<uint64_t>(const OpsArgs<uint64_t>&, uint32_t, hipStream_t, uint32_t*, int, int);
