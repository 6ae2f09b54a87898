// rANS stage for gfx950: every brick's rANS-coded op stream -> its nibble
// stream in HBM (8 nibbles per u32, nibble k of a brick at bits 4*(k&7) of
// word k>>3). Restates decodeRansStream (brickDecoding.cpp:126-188):
//   x = s & 4095; symbol = CDF^-1(x); s = f*(s>>12) + x - c;
//   while (s < 2^23 && bytes left) s = (s << 8) | next byte (read backwards)
// The initial state is the last 4 bytes (byte E-1 most significant).
// After the pad symbol of an odd count, a well-formed stream is back at
// state 2^23 with every byte consumed; violations set FCS_ST_RANS_STATE.
//
// Two mappings of bricks to the machine (picked by the runtime):
//   rans_wave_kernel: one wave per brick. The stream is staged in an LDS ring
//     (coalesced 64-lane loads, next half-ring prefetched in VGPRs), the rANS
//     state is wave-uniform (SGPRs, scalar ALU), the symbol search is one
//     vector compare of x against a per-lane CDF + ballot + popcount, and
//     nibble word k of a 64-word group is kept in lane k, stored 256 B at a time.
//     Lowest latency per brick: for volumes with few, long bricks.
//     The decode table is read through the scalar cache, or copied into LDS.
//   rans_win_kernel: one wave per brick, the byte source in scalar registers.
//   rans_lane_kernel: one lane per brick, 64 bricks per wave (launched longest
//     first), symbol table in LDS, each lane's stream through a ring in LDS
//     refilled at wave-wide checkpoints. Highest throughput: for volumes with
//     many bricks.
#include <hip/hip_runtime.h>

#include "fcs_common.h"
#include "fcs_device.h"
#include "fcs_kernels.h"

namespace {

// ---------------------------------------------------------- wave per brick
#define RING_WORDS 4096u
#define RING_HALF 2048u
#define RING_PEND (RING_HALF / 64u)

// TM: 0 = decode table through the scalar cache, 1 = ballot over a per-lane
// CDF (no table), 2 = decode table copied into LDS (one LDS round trip per
// symbol instead of a scalar-cache lookup that can miss to L2)
template <int TM>
__global__ __launch_bounds__(64) void rans_wave_kernel(RansArgs a) {
  __shared__ uint32_t ring[RING_WORDS];
  __shared__ uint32_t ltab[TM == 2 ? FCS_RANS_M : 1];
  const uint32_t lane = threadIdx.x;
  const uint32_t b = a.list ? a.list[blockIdx.x] : blockIdx.x;
  const DevBrick B = a.bricks[b];
  const uint32_t E = B.E, nsym = B.nsym;
  if (E < 4 || nsym == 0) {
    if (lane == 0) a.res[b] = RansResult{FCS_ST_RANS_STATE | FCS_ST_STREAM_OVERRUN, 0u};
    return;
  }
  uint32_t* nibw = a.nib + B.nib_off;
  const uint32_t* words = reinterpret_cast<const uint32_t*>(a.stream + B.str_off);
  // per-lane CDF: lane k < 15 holds cum[k+1] (so the symbol is the number of
  // lanes with cum[k+1] <= x), lane k < 16 holds freq[k] and cum[k]
  const uint32_t my_next = lane < 15 ? a.cum[lane + 1] : 0xFFFFFFFFu;
  const uint32_t my_f = lane < 16 ? a.freq[lane] : 0u;
  const uint32_t my_c = lane < 16 ? a.cum[lane] : 0u;

  const int32_t nwords = (int32_t)((E + 3) >> 2);
  int32_t lo = nwords > (int32_t)RING_WORDS ? nwords - (int32_t)RING_WORDS : 0;
  for (int32_t w = lo + (int32_t)lane; w < nwords; w += 64) ring[w & (RING_WORDS - 1)] = words[w];
  if (TM == 2) {
    const uint4* t4 = reinterpret_cast<const uint4*>(a.table);
    uint4* l4 = reinterpret_cast<uint4*>(ltab);
#pragma unroll
    for (uint32_t i = 0; i < FCS_RANS_M / 4 / 64; ++i) l4[lane + 64 * i] = t4[lane + 64 * i];
  }
  uint32_t pend[RING_PEND];
#pragma unroll
  for (uint32_t i = 0; i < RING_PEND; ++i) {
    const int32_t w = lo - (int32_t)RING_HALF + (int32_t)(lane + 64 * i);
    pend[i] = words[w < 0 ? 0 : w];
  }
  __syncthreads();

  // wave-uniform byte source: cw holds the current word, next byte on top
  uint32_t pos = E;  // bytes left to read
  int32_t wnext = (int32_t)((E - 1) >> 2);  // next word to fetch from the ring
  uint32_t cw, nbb;
  auto fetch_word = [&]() {
    if (lo > 0 && wnext < lo + (int32_t)RING_HALF) {
      // words [lo + HALF, lo + RING) are consumed: their slots take the
      // prefetched half [lo - HALF, lo); then prefetch the half below
      __syncthreads();
#pragma unroll
      for (uint32_t i = 0; i < RING_PEND; ++i) {
        const int32_t w = lo - (int32_t)RING_HALF + (int32_t)(lane + 64 * i);
        ring[(w < 0 ? 0 : w) & (RING_WORDS - 1)] = w < 0 ? ring[0] : pend[i];
      }
      lo = lo > (int32_t)RING_HALF ? lo - (int32_t)RING_HALF : 0;
#pragma unroll
      for (uint32_t i = 0; i < RING_PEND; ++i) {
        const int32_t w = lo - (int32_t)RING_HALF + (int32_t)(lane + 64 * i);
        pend[i] = words[w < 0 ? 0 : w];
      }
      __syncthreads();
    }
    cw = __builtin_amdgcn_readfirstlane(ring[wnext & (RING_WORDS - 1)]);
    --wnext;
    nbb = 4;
  };
  fetch_word();
  {
    const uint32_t skip = 3u - ((E - 1) & 3u);  // bytes of the last word beyond E
    cw <<= 8u * skip;
    nbb -= skip;
  }
  auto next_byte = [&]() -> uint32_t {
    if (nbb == 0) fetch_word();
    const uint32_t r = cw >> 24;
    cw <<= 8;
    --nbb;
    --pos;
    return r;
  };
  uint32_t state = next_byte() << 24;
  state |= next_byte() << 16;
  state |= next_byte() << 8;
  state |= next_byte();

  // symbol step. With a decode table (x -> symbol | freq | x - cum) the
  // chain stays on the scalar unit (one scalar-cache load per symbol);
  // otherwise the symbol is the ballot count of per-lane CDF compares.
  // the table is read through the constant address space so that the
  // uniform-index loads go through the scalar cache (s_load)
  typedef __attribute__((address_space(4))) const uint32_t ctab_t;
  ctab_t* tab = (ctab_t*)a.table;
  auto decode_one = [&]() -> uint32_t {
    const uint32_t x = state & (FCS_RANS_M - 1);
    uint32_t sym;
    if (TM == 2) {
      const uint32_t e = (uint32_t)__builtin_amdgcn_readfirstlane((int)ltab[x]);
      sym = e >> 25;
      state = ((e >> 12) & 0x1FFFu) * (state >> FCS_RANS_LOGM) + (e & 0xFFFu);
    } else if (TM == 0) {
      const uint32_t e = tab[x];
      sym = e >> 25;
      state = ((e >> 12) & 0x1FFFu) * (state >> FCS_RANS_LOGM) + (e & 0xFFFu);
    } else {
      sym = (uint32_t)__popcll(__ballot(x >= my_next));
      const uint32_t f = (uint32_t)__builtin_amdgcn_readlane((int)my_f, (int)sym);
      const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)my_c, (int)sym);
      state = f * (state >> FCS_RANS_LOGM) + x - c;
    }
    // the empty asm keeps the byte in a scalar register: (state << 8) | byte
    // is otherwise matched as a funnel shift, which exists only on the vector
    // unit, and each renormalisation step would round-trip through a vector
    // register on the serial chain
    auto renorm = [&]() {
      while (state < FCS_RANS_L && pos > 0) {
        uint32_t byte = next_byte();
        __asm__ volatile("" : "+s"(byte));
        state = (state << 8) | byte;
      }
    };
    if constexpr (TM == 0) {
      // most symbols take no byte: that path falls through
      if (__builtin_expect(state < FCS_RANS_L, 0)) renorm();
    } else {
      renorm();
    }
    return sym;
  };

  // 8 symbols -> one nibble word; word k of a 64-word group lives in lane k
  uint32_t outv = 0;
  const uint32_t nfull = nsym >> 3;
  for (uint32_t wd = 0; wd < nfull; ++wd) {
    uint32_t word = 0;
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) word |= decode_one() << (4u * k);
    outv = lane == (wd & 63u) ? word : outv;
    if ((wd & 63u) == 63u) nibw[(wd >> 6) * 64u + lane] = outv;
  }
  if (nsym & 7u) {
    uint32_t word = 0;
    for (uint32_t k = 0; k < (nsym & 7u); ++k) word |= decode_one() << (4u * k);
    outv = lane == (nfull & 63u) ? word : outv;
  }
  const uint32_t total_words = (nsym + 7) >> 3;
  const uint32_t stored = (nsym >> 9) * 64u;
  if (lane < total_words - stored) nibw[stored + lane] = outv;

  uint32_t status = 0;
  if (nsym & 1u) {
    if (decode_one() != 0) status |= FCS_ST_RANS_STATE;
  }
  if (state != FCS_RANS_L || pos != 0) status |= FCS_ST_RANS_STATE;
  if (lane == 0) a.res[b] = RansResult{status, 0u};
}

// ------------------------------------------------- wave per brick, window
// Same mapping, with the byte source in scalar registers: a 64-bit window of
// upcoming stream bytes (next byte in bits 63..56) refilled 4 bytes at a time
// from a 64-word block held one word per lane (v_readlane), the next block
// loading while the current one is consumed; no LDS. A group of 8 symbols
// whose start has the state >= 2^23 and >= 16 bytes left needs at most 2
// bytes per symbol: its symbol step is straight scalar code and the
// renormalisation runs out of line. Other groups (stream end, malformed
// input) take the byte-by-byte loop of the reference.
// SYM: 0 = decode table through the scalar cache, 1 = ballot over a per-lane CDF
template <int SYM>
__global__ __launch_bounds__(64) void rans_win_kernel(RansArgs a) {
  const uint32_t lane = threadIdx.x;
  const uint32_t b = a.list ? a.list[blockIdx.x] : blockIdx.x;
  const DevBrick B = a.bricks[b];
  const uint32_t E = B.E, nsym = B.nsym;
  if (E < 4 || nsym == 0) {
    if (lane == 0) a.res[b] = RansResult{FCS_ST_RANS_STATE | FCS_ST_STREAM_OVERRUN, 0u};
    return;
  }
  uint32_t* nibw = a.nib + B.nib_off;
  const uint32_t* words = reinterpret_cast<const uint32_t*>(a.stream + B.str_off);
  const uint32_t my_next = lane < 15 ? a.cum[lane + 1] : 0xFFFFFFFFu;
  const uint32_t my_f = lane < 16 ? a.freq[lane] : 0u;
  const uint32_t my_c = lane < 16 ? a.cum[lane] : 0u;
  typedef __attribute__((address_space(4))) const uint32_t ctab_t;
  ctab_t* tab = (ctab_t*)a.table;

  // lane i of a block holds word (base - i); words before the stream read 0
  int32_t base = (int32_t)((E + 3) >> 2) - 1;
  auto ldw = [&](int32_t top) -> uint32_t {
    const int32_t w = top - (int32_t)lane;
    return w >= 0 ? words[w] : 0u;
  };
  uint32_t blk = ldw(base), nblk = ldw(base - 64);
  uint32_t widx = 0;
  auto next_word = [&]() -> uint32_t {
    const uint32_t w = (uint32_t)__builtin_amdgcn_readlane((int)blk, (int)widx);
    if (++widx == 64u) {
      blk = nblk;
      base -= 64;
      nblk = ldw(base - 64);
      widx = 0;
    }
    return w;
  };
  uint64_t win;
  uint32_t avail;
  {
    const uint32_t skip = 3u - ((E - 1) & 3u);  // bytes of the last word beyond E
    win = (uint64_t)(next_word() << (8u * skip)) << 32;
    avail = 32u - 8u * skip;
  }
  // avail <= 32 before every refill
  auto refill = [&]() {
    win |= (uint64_t)next_word() << (32u - avail);
    avail += 32u;
  };
  refill();
  uint32_t pos = E - 4;  // bytes left after the initial state
  uint32_t state = (uint32_t)(win >> 32);
  win <<= 32;
  avail -= 32u;

  auto lookup = [&](uint32_t x, uint32_t& f, uint32_t& r) -> uint32_t {
    if (SYM == 0) {
      const uint32_t e = tab[x];
      f = (e >> 12) & 0x1FFFu;
      r = e & 0xFFFu;
      return e >> 25;
    } else {
      const uint32_t sym = (uint32_t)__popcll(__ballot(x >= my_next));
      f = (uint32_t)__builtin_amdgcn_readlane((int)my_f, (int)sym);
      r = x - (uint32_t)__builtin_amdgcn_readlane((int)my_c, (int)sym);
      return sym;
    }
  };
  auto fast_one = [&]() -> uint32_t {
    uint32_t f, r;
    const uint32_t sym = lookup(state & (FCS_RANS_M - 1), f, r);
    uint32_t s1 = f * (state >> FCS_RANS_LOGM) + r;
    if (__builtin_expect(s1 < FCS_RANS_L, 0)) {
      if (avail < 16u) refill();
      const uint32_t two = (uint32_t)(win >> 48);
      if (s1 < (1u << 15)) {
        s1 = (s1 << 16) | two;
        win <<= 16;
        avail -= 16u;
        pos -= 2;
      } else {
        s1 = (s1 << 8) | (two >> 8);
        win <<= 8;
        avail -= 8u;
        pos -= 1;
      }
    }
    state = s1;
    return sym;
  };
  auto exact_one = [&]() -> uint32_t {
    uint32_t f, r;
    const uint32_t sym = lookup(state & (FCS_RANS_M - 1), f, r);
    state = f * (state >> FCS_RANS_LOGM) + r;
    while (state < FCS_RANS_L && pos > 0) {
      if (avail < 8u) refill();
      uint32_t byte = (uint32_t)(win >> 56);
      __asm__ volatile("" : "+s"(byte));  // scalar: no vector funnel shift on the chain
      state = (state << 8) | byte;
      win <<= 8;
      avail -= 8u;
      --pos;
    }
    return sym;
  };

  uint32_t outv = 0;
  const uint32_t nfull = nsym >> 3;
  for (uint32_t wd = 0; wd < nfull; ++wd) {
    uint32_t word = 0;
    if (pos >= 16u && state >= FCS_RANS_L) {
#pragma unroll
      for (uint32_t k = 0; k < 8; ++k) word |= fast_one() << (4u * k);
    } else {
      for (uint32_t k = 0; k < 8; ++k) word |= exact_one() << (4u * k);
    }
    outv = lane == (wd & 63u) ? word : outv;
    if ((wd & 63u) == 63u) nibw[(wd >> 6) * 64u + lane] = outv;
  }
  if (nsym & 7u) {
    uint32_t word = 0;
    for (uint32_t k = 0; k < (nsym & 7u); ++k) word |= exact_one() << (4u * k);
    outv = lane == (nfull & 63u) ? word : outv;
  }
  const uint32_t total_words = (nsym + 7) >> 3;
  const uint32_t stored = (nsym >> 9) * 64u;
  if (lane < total_words - stored) nibw[stored + lane] = outv;

  uint32_t status = 0;
  if (nsym & 1u) {
    if (exact_one() != 0) status |= FCS_ST_RANS_STATE;
  }
  if (state != FCS_RANS_L || pos != 0) status |= FCS_ST_RANS_STATE;
  if (lane == 0) a.res[b] = RansResult{status, 0u};
}

// ---------------------------------------------------------- lane per brick
// Each lane decodes one brick. Its stream is read backwards through a ring of
// LR_BLOCKS 16-byte blocks in LDS (kept in consumption order: ring row d % 16
// holds the d-th word read) and a 64-bit bit buffer in registers, topped up
// from a ring word read ahead of use, once per pair of symbols. Global memory
// is touched only every LR_CHECK symbols, an iteration all lanes of the wave
// share: the blocks loaded at the previous checkpoint go into the ring, the
// next ones are requested, and finished nibble words are stored. Waits for
// global memory (which complete in issue order, per wave) therefore only ever
// cover operations issued a whole interval earlier; a lane refilling its
// buffer at its own moment waits for LDS only. From an initial state in
// [2^23, 2^31), every symbol leaves the state in [2^11, 2^31) and
// renormalising brings it back to [2^23, 2^31) with at most 2 bytes, so a
// lane reads at most 8 ring words per interval and a ring holding two
// intervals never runs dry; the bytes left need no per-symbol count, since
// the sequential rule's "and bytes are left" only differs once the stream has
// been read past its start, which the bits consumed show at the end. A lane
// whose initial state is outside that range, whose stream was read past its
// start, or whose read index was found past the write index (corrupt input
// only) decodes its brick again by the sequential rule, byte by byte, so
// every stream decodes exactly as that rule reads it.
constexpr uint32_t LANE_NT = 256;
constexpr uint32_t LR_BLOCKS = 4;             // 64 bytes per lane
constexpr uint32_t LR_WORDS = LR_BLOCKS * 4;
constexpr uint32_t LR_CHECK = 16;             // symbols between checkpoints (<= 32 bytes)
constexpr uint32_t LR_ROW = LANE_NT * 4u;     // bytes between ring rows
constexpr uint32_t LR_MASK = (LR_WORDS - 1u) * LR_ROW;

struct LaneStream {
  const uint4* base;  // the brick's stream as 16-byte blocks
  char* ring;         // LDS, LR_WORDS rows of LANE_NT words, this lane's column
  int32_t q;          // next block to request (decreasing; below 0 reads as zeros)
  // ring words read / written so far, times LR_ROW (so that the low bits
  // masked by LR_MASK are the row's byte offset); on a lane not marked bad,
  // rd never passes wr and they differ by at most LR_WORDS words, so
  // comparisons survive wrapping (a bad lane's values are discarded)
  uint32_t rdx, wrx;
  uint32_t rdx0, bits0;  // read index and stream bits in bb after the initial fill
  uint64_t bb;        // upcoming stream bytes, the next one in bits 63..56
  uint32_t nbits;     // valid bits in bb (a multiple of 8)
  uint32_t pre;       // ring word rd, read at the end of the previous symbol
  bool staged;
  bool bad;           // the fast rule left the sequential one: redo the brick
  uint4 s0, s1;       // blocks requested at the last checkpoint

  __device__ __forceinline__ uint4 blk(int32_t k) const {
    return k >= 0 ? base[k] : make_uint4(0u, 0u, 0u, 0u);
  }
  __device__ __forceinline__ uint32_t at(uint32_t dx) const {
    return *reinterpret_cast<const uint32_t*>(ring + (dx & LR_MASK));
  }
  // a block's words are read last to first
  __device__ __forceinline__ void put(const uint4& v) {
    char* r = ring + (wrx & LR_MASK);
    *reinterpret_cast<uint32_t*>(r) = v.w;
    *reinterpret_cast<uint32_t*>(r + LR_ROW) = v.z;
    *reinterpret_cast<uint32_t*>(r + 2u * LR_ROW) = v.y;
    *reinterpret_cast<uint32_t*>(r + 3u * LR_ROW) = v.x;
    wrx += 4u * LR_ROW;
  }
  // issued once per symbol, by every lane: the word a refill would take next
  // (read again unchanged when nothing was taken); used one symbol later
  __device__ __forceinline__ void prefetch() { pre = at(rdx); }
  // fewer than 2 bytes left after the initial state: append the prefetched
  // word below them
  __device__ __forceinline__ void refill() {
    bad = bad || rdx == wrx;
    bb |= (uint64_t)pre << (32u - nbits);
    nbits += 32u;
    rdx += LR_ROW;
  }
  __device__ __forceinline__ void init(const uint8_t* s, uint32_t E) {
    base = reinterpret_cast<const uint4*>(s);
    const int32_t q0 = (int32_t)((E - 1) >> 4);
    wrx = 0;
    for (uint32_t k = 0; k < LR_BLOCKS; ++k) put(blk(q0 - (int32_t)k));
    q = q0 - (int32_t)LR_BLOCKS;
    // the first interval's blocks are requested now, so that every
    // checkpoint finds its blocks requested one interval earlier
    s0 = blk(q);
    s1 = blk(q - 1);
    q -= 2;
    staged = true;
    bad = false;
    const uint32_t j = ((E - 1) >> 2) & 3u;     // word of byte E-1 in its block
    const uint32_t skip = 3u - ((E - 1) & 3u);  // bytes of that word beyond E
    rdx = (3u - j) * LR_ROW;
    bb = (uint64_t)(at(rdx) << (8u * skip)) << 32;
    rdx += LR_ROW;
    nbits = 32u - 8u * skip;
    bb |= (uint64_t)at(rdx) << (32u - nbits);
    rdx += LR_ROW;
    nbits += 32u;  // 40..64 bits: the initial state and more
    rdx0 = rdx;
    bits0 = nbits;
    prefetch();
  }
  // every lane of the wave, the same symbol: stage the blocks requested one
  // interval ago, request the next ones. A read index past the write index
  // (a refill took a ring word not yet written) marks the lane for the
  // sequential redo.
  __device__ __forceinline__ void checkpoint() {
    bad = bad || (int32_t)(wrx - rdx) < 0;
    if (staged && wrx - rdx + 8u * LR_ROW <= LR_WORDS * LR_ROW) {
      put(s0);
      put(s1);
      staged = false;
    }
    if (!staged) {
      s0 = blk(q);
      s1 = blk(q - 1);
      q -= 2;
      staged = true;
    }
  }
  // renormalise after a symbol: while state < 2^23, shift the next byte in.
  // A state in [2^11, 2^31) (every state of a lane whose initial state is in
  // [2^23, 2^31)) takes (clz(state) - 1) rounded down to a multiple of 8
  // bits: 0, 1 or 2 bytes, shifted in from bb at once (bb holds at least 16
  // bits here). Other lanes compute garbage here and are redone.
  __device__ __forceinline__ void shift(uint32_t& state) {
    const uint32_t sh = ((uint32_t)__clz((int)state) - 1u) & 24u;  // 0, 8 or 16 here (24 only on a redone lane)
    state = (uint32_t)(((((uint64_t)state) << 32) | (bb >> 32)) >> (32u - sh));
    bb <<= sh;
    nbits -= sh;
  }
  // stream bits consumed so far (the bits taken into bb less those still in it)
  __device__ __forceinline__ uint32_t consumed() const { return bits0 + 32u * ((rdx - rdx0) / LR_ROW) - nbits; }
  // fewer than T bits left (T <= 32): append the word prefetched from the
  // ring below them. With at most 2 bytes per symbol and a refill of one
  // word, a lane not marked reads at most 8 ring words per 16 symbols, which
  // the checkpoints keep staged ahead of it: its ring never runs dry.
  template <uint32_t T>
  __device__ __forceinline__ void top_up() {
    const bool low = nbits < T;
    bb |= (uint64_t)(low ? pre : 0u) << ((32u - nbits) & 63u);
    const uint32_t r = low ? 32u : 0u;
    nbits += r;
    rdx += r * (LR_ROW / 32u);
  }
};

// decode table entry for x: symbol (bits 0..3), x - cum[symbol] (bits 4..15),
// freq[symbol] (bits 16..28)
__device__ __forceinline__ uint32_t lane_step(uint32_t& state, const uint32_t* table) {
  const uint32_t e = table[state & (FCS_RANS_M - 1)];
  state = (e >> 16) * (state >> FCS_RANS_LOGM) + ((e >> 4) & 0xFFFu);
  return e;
}

// A corrupt stream's brick decoded again by the sequential rule (a byte at a
// time while the state is below 2^23 and bytes are left), over the nibble
// words the fast pass wrote; returns the stream's status bits.
__device__ __noinline__ uint32_t lane_redo(const uint8_t* in, uint32_t E, uint32_t nsym, uint32_t* nibw,
                                           const uint32_t* table) {
  uint32_t pos = E - 4;
  uint32_t state = ((uint32_t)in[E - 1] << 24) | ((uint32_t)in[E - 2] << 16) | ((uint32_t)in[E - 3] << 8) | in[E - 4];
  auto step = [&]() -> uint32_t {
    const uint32_t e = lane_step(state, table);
    while (state < FCS_RANS_L && pos > 0) state = (state << 8) | in[--pos];
    return e & 15u;
  };
  for (uint32_t wd = 0; 8 * wd < nsym; ++wd) {
    const uint32_t cnt = min(8u, nsym - 8 * wd);
    uint32_t word = 0;
    for (uint32_t k = 0; k < cnt; ++k) word = (word >> 4) | (step() << 28);
    nibw[wd] = word >> (32u - 4u * cnt);
  }
  uint32_t status = 0;
  if ((nsym & 1u) && step() != 0) status |= FCS_ST_RANS_STATE;
  if (state != FCS_RANS_L || pos != 0) status |= FCS_ST_RANS_STATE;
  return status;
}

__global__ __launch_bounds__(LANE_NT) void rans_lane_kernel(RansArgs a) {
  __shared__ uint32_t table[FCS_RANS_M];
  __shared__ uint32_t rings[LR_WORDS * LANE_NT];
  for (uint32_t x = threadIdx.x; x < FCS_RANS_M; x += blockDim.x) {
    uint32_t s = 0;
    while (s + 1 < 16 && x >= a.cum[s + 1]) ++s;
    table[x] = s | ((x - a.cum[s]) << 4) | (a.freq[s] << 16);
  }
  __syncthreads();
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.nlist) return;
  const uint32_t b = a.list ? a.list[i] : i;
  const DevBrick B = a.bricks[b];
  const uint32_t E = B.E, nsym = B.nsym;
  if (E < 4 || nsym == 0) {
    a.res[b] = RansResult{FCS_ST_RANS_STATE | FCS_ST_STREAM_OVERRUN, 0u};
    return;
  }
  uint32_t* nibw = a.nib + B.nib_off;
  LaneStream rd;
  rd.ring = reinterpret_cast<char*>(rings + threadIdx.x);
  rd.init(a.stream + B.str_off, E);
  // the initial state: the stream's last 4 bytes
  uint32_t state = (uint32_t)(rd.bb >> 32);
  rd.bb <<= 32;
  rd.nbits -= 32u;
  // at least 32 bits for the first pair of symbols
  if (rd.nbits < 32u) rd.refill();
  rd.prefetch();
  // a state in [2^23, 2^31) before every symbol keeps every later one there
  rd.bad = rd.bad || state < FCS_RANS_L || state >= 0x80000000u;
  // one symbol; at least 16 bits left for the next
  auto step = [&]() -> uint32_t {
    const uint32_t e = lane_step(state, table);
    rd.shift(state);
    rd.top_up<16u>();
    rd.prefetch();
    return e;
  };
  // two symbols (at most 32 bits) from a buffer of at least 32; one refill
  // to at least 32 bits for the next pair
  auto step2 = [&](uint32_t& word) {
    const uint32_t e0 = lane_step(state, table);
    rd.shift(state);
    const uint32_t e1 = lane_step(state, table);
    rd.shift(state);
    rd.top_up<32u>();
    rd.prefetch();
    word = (word >> 4) | (e0 << 28);
    word = (word >> 4) | (e1 << 28);
  };
  // 8 symbols -> one nibble word (each symbol shifted in at the top, so the
  // first ends in bits 0..3); 16 words -> one 64-byte line, stored whole. The
  // word counter is the same for every active lane (checkpoints every 16
  // symbols and stores happen at iterations the whole wave shares), so the
  // word array is indexed uniformly.
  uint32_t w[16];
#pragma unroll
  for (uint32_t k = 0; k < 16; ++k) w[k] = 0;
  auto flush = [&](uint32_t line) {
    uint4* o = reinterpret_cast<uint4*>(nibw) + line * 4u;
    o[0] = make_uint4(w[0], w[1], w[2], w[3]);
    o[1] = make_uint4(w[4], w[5], w[6], w[7]);
    o[2] = make_uint4(w[8], w[9], w[10], w[11]);
    o[3] = make_uint4(w[12], w[13], w[14], w[15]);
  };
  const uint32_t nfull = nsym >> 3;
  for (uint32_t wd = 0; wd < nfull; ++wd) {
    if ((wd & 1u) == 0 && wd != 0) {
      rd.checkpoint();
      if ((wd & 15u) == 0) flush((wd >> 4) - 1);
    }
    uint32_t word = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) step2(word);
    w[wd & 15u] = word;
  }
  const uint32_t r = nsym & 7u;
  if (r) {
    if ((nfull & 1u) == 0 && nfull != 0) {
      rd.checkpoint();
      // the partial word opens a new line: the previous one is complete
      if ((nfull & 15u) == 0) flush((nfull >> 4) - 1);
    }
    uint32_t word = 0;
    for (uint32_t k = 0; k < r; ++k) word = (word >> 4) | (step() << 28);
    w[nfull & 15u] = word >> (32u - 4u * r);
  }
  // the last line (nibble regions are whole 64-byte lines)
  flush((nsym - 1u) >> 7);
  uint32_t status = 0;
  if (nsym & 1u) {
    const uint32_t e = lane_step(state, table);
    rd.shift(state);
    if ((e & 15u) != 0) status |= FCS_ST_RANS_STATE;
  }
  // a stream that ran out before its symbols did was read past its start
  const uint32_t used = rd.consumed(), have = 8u * E;
  if (state != FCS_RANS_L || used != have) status |= FCS_ST_RANS_STATE;
  rd.bad = rd.bad || used > have || (int32_t)(rd.wrx - rd.rdx) < 0;
  if (rd.bad) status = lane_redo(a.stream + B.str_off, E, nsym, nibw, table);
  a.res[b] = RansResult{status, 0u};
}

}  // namespace

hipError_t fcs_launch_rans(const RansArgs& a, int mode, hipStream_t s) {
  if (a.nlist == 0) return hipSuccess;
  if (mode == FCS_RANS_WAVE) {
    hipLaunchKernelGGL(rans_wave_kernel<0>, dim3(a.nlist), dim3(64), 0, s, a);
  } else if (mode == FCS_RANS_WAVE_BALLOT) {
    hipLaunchKernelGGL(rans_wave_kernel<1>, dim3(a.nlist), dim3(64), 0, s, a);
  } else if (mode == FCS_RANS_WAVE_LDS) {
    hipLaunchKernelGGL(rans_wave_kernel<2>, dim3(a.nlist), dim3(64), 0, s, a);
  } else if (mode == FCS_RANS_WIN) {
    hipLaunchKernelGGL(rans_win_kernel<0>, dim3(a.nlist), dim3(64), 0, s, a);
  } else if (mode == FCS_RANS_WIN_BALLOT) {
    hipLaunchKernelGGL(rans_win_kernel<1>, dim3(a.nlist), dim3(64), 0, s, a);
  } else {
    hipLaunchKernelGGL(rans_lane_kernel, dim3((a.nlist + 255) / 256), dim3(256), 0, s, a);
  }
  return hipGetLastError();
}
