// Segmented stable LSD radix sort of edge records on gfx950.
//
// The reference sorts with load_balance_radix_sort (core/graph.hpp:465-584):
// a stable LSD sort whose result depends only on stability, so any stable sort
// by the same key gives the identical order.  A sort here covers nseg
// independent contiguous segments (one per rank slice, or one) and takes
// ceil(key_bits/8) balanced digit passes, or ceil(key_bits/9) with 9-bit
// digits where that is fewer.  Two pass schemes, each with the tile size as a
// template parameter (NT threads x 16 records):
//
//   three-launch   upsweep: per-tile digit histogram (LDS atomics) into a
//                  (segment, digit, tile)-major table; device scan of the
//                  table; downsweep: re-read, rank, scatter.
//   one-pass       a histogram of all passes up front, then per pass one
//                  launch: tiles taken by ticket publish their digit counts
//                  and look back over earlier tiles of their segment
//                  (decoupled look-back, bounded wait) before scattering.
//
// Ranking inside a tile is stable: wave w owns a contiguous run of the tile,
// handled in rounds of 64 records; the lanes holding the same digit find each
// other through a per-wave LDS table of lane masks (those sharing the round's
// first digit through one ballot), and per-wave running counters in LDS order
// the rounds.  The tile is then staged in LDS in digit order, so each digit's
// run is written contiguously (bigger tiles => longer runs => fewer partially
// written lines).  The variant (scheme, threads, digit width) is process-wide
// per record size: sort_variant(), set by eg_set_sort_variant or
// eg_autotune_sort.
//
// A sort's input, intermediate lists and result are each either records or
// field arrays (src[], dst[], w[]; RxStorage): where a list is only an
// intermediate, its later upsweeps read the key's array alone.
#pragma once

#include <type_traits>
#include <vector>

#include "common.hpp"

namespace eg {

constexpr int kRxIPT = 16;
constexpr int kRxMaxBits = 9;  // widest digit any variant uses
constexpr int kRxMaxPasses = 8;
constexpr uint32_t kRxSpinLimit = 1u << 22;
constexpr uint64_t kFlagAgg = 1, kFlagInc = 2;

enum SortScheme { kThreeLaunch = 0, kOnePass = 1, kThreeLaunchXcd = 2 };

struct RxPlan {
  int passes = 0;
  int shift[kRxMaxPasses] = {0};
  int bits[kRxMaxPasses] = {0};
};

// Per-wave digit counters of a tile (after the scan, digit starts in the
// tile) stay below 2^16, a tile holding at most 16384 records: 9-bit digits
// keep them as 16-bit words so that two 512-thread tiles share a CU's LDS.
template <int BM>
using RxWc = std::conditional_t<(BM > 8), uint16_t, uint32_t>;

inline RxPlan rx_plan(int key_bits, int max_bits) {
  RxPlan p;
  if (key_bits <= 0) key_bits = 1;
  p.passes = (key_bits + max_bits - 1) / max_bits;
  const int d = (key_bits + p.passes - 1) / p.passes;
  for (int i = 0; i < p.passes; ++i) {
    p.shift[i] = i * d;
    p.bits[i] = d < key_bits - i * d ? d : key_bits - i * d;
  }
  return p;
}

// Key of every launch after a sort's first upsweep: a key functor with a
// one-time side effect (a range check) overloads this for its own type.
template <class KeyFn>
KeyFn rx_quiet(const KeyFn& k) {
  return k;
}

// Input of a sort pass.  Plain: record i of the sort order is p[i].  Block
// view (first pass of a sort that regroups contiguous blocks on the fly):
// record i is p[i + shift[t]] for tile t holding i, or, in a tile that
// straddles a block boundary (shift[t] == kRxStraddle), p[bsrc[b] + i - bdst[b]]
// for the last block b with bdst[b] <= i.
// Records kept as field arrays (structure of arrays): src[i], dst[i] and,
// for 12-byte records, w[i].  A sort whose intermediate lists are field
// arrays reads only the key's array in its upsweeps (half the bytes of
// 8-byte records); the record layout is kept where a list is a result.
template <class R>
struct RxSoa {
  uint32_t* s = nullptr;
  uint32_t* d = nullptr;
  float* w = nullptr;  // 12-byte records only
  RxSoa at(uint64_t off) const {
    RxSoa o;
    o.s = s + off;
    o.d = d + off;
    o.w = w ? w + off : nullptr;
    return o;
  }
};

// Field arrays of n records in one buffer, each array starting on a 256-byte
// boundary: rx_soa_bytes<R>(n) bytes.
inline uint64_t rx_soa_stride(uint64_t n) { return (n + 63) & ~uint64_t(63); }
template <class R>
uint64_t rx_soa_bytes(uint64_t n) {
  return rx_soa_stride(n) * sizeof(R);
}
template <class R>
RxSoa<R> rx_soa_in(void* buf, uint64_t n) {
  const uint64_t st = rx_soa_stride(n);
  RxSoa<R> o;
  o.s = static_cast<uint32_t*>(buf);
  o.d = o.s + st;
  if constexpr (sizeof(R) == 12) o.w = reinterpret_cast<float*>(o.d + st);
  return o;
}

template <class R>
__device__ __forceinline__ R rx_soa_load(const RxSoa<R>& a, uint64_t j) {
  R r;
  r.src = a.s[j];
  r.dst = a.d[j];
  if constexpr (sizeof(R) == 12) r.w = a.w[j];
  return r;
}

template <class R>
__device__ __forceinline__ void rx_soa_store(const RxSoa<R>& a, uint64_t j, const R& r) {
  a.s[j] = r.src;
  a.d[j] = r.dst;
  if constexpr (sizeof(R) == 12) a.w[j] = r.w;
}

// Field a key reads (0 src, 1 dst) when the key functor names one
// (static constexpr int kField, with of_field(v) the key of field value v),
// else -1: an upsweep over field arrays then reads that array only.
template <class K, class = void>
struct RxKeyField {
  static constexpr int value = -1;
};
template <class K>
struct RxKeyField<K, std::void_t<decltype(K::kField)>> {
  static constexpr int value = K::kField;
};

template <class R>
struct RxIn {
  const R* p;            // records (record layout), or null for field arrays
  const int64_t* shift;  // per tile, or null for plain input
  const uint64_t* bdst;  // nblk + 1 block starts in sort order
  const uint64_t* bsrc;  // nblk block starts in p
  int nblk;
  RxSoa<R> f;            // field arrays (kernels with SIN = true)
};

// Output of a pass: records, or field arrays (SOUT = true).
template <class R>
struct RxOut {
  R* p = nullptr;
  RxSoa<R> f;
};

constexpr int64_t kRxStraddle = INT64_MIN;

template <class R>
__device__ __forceinline__ int64_t rx_tile_shift(const RxIn<R>& in, uint32_t t) {
  return in.shift ? in.shift[t] : 0;
}

// position in the input's storage of record i of the sort order
template <class R>
__device__ __forceinline__ uint64_t rx_src_index(const RxIn<R>& in, int64_t sh, uint64_t i) {
  if (sh != kRxStraddle) return uint64_t(int64_t(i) + sh);
  int lo = 0, hi = in.nblk;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (in.bdst[mid] <= i)
      lo = mid;
    else
      hi = mid;
  }
  return in.bsrc[lo] + (i - in.bdst[lo]);
}

template <bool SIN, class R>
__device__ __forceinline__ R rx_load(const RxIn<R>& in, int64_t sh, uint64_t i) {
  const uint64_t j = rx_src_index(in, sh, i);
  if constexpr (SIN)
    return rx_soa_load(in.f, j);
  else
    return in.p[j];
}

template <bool SOUT, class R>
__device__ __forceinline__ void rx_store(const RxOut<R>& out, uint64_t j, const R& r) {
  if constexpr (SOUT)
    rx_soa_store(out.f, j, r);
  else
    out.p[j] = r;
}

struct RxPassArgs {
  int npass;
  int shift[kRxMaxPasses];
  int bits[kRxMaxPasses];
};

__device__ __forceinline__ uint64_t rx_pack(uint32_t epoch, uint64_t flag, uint32_t v) {
  return (uint64_t(epoch) << 34) | (flag << 32) | v;
}

// segment of global tile t: the last s with tile_start[s] <= t (skips empty
// segments, whose tile ranges are empty)
__device__ __forceinline__ int rx_seg_of(const uint32_t* tile_start, int nseg, uint32_t t) {
  int lo = 0, hi = nseg;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (tile_start[mid] <= t)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

struct RxTile {
  int s;          // segment
  uint32_t lt;    // tile index inside the segment
  uint32_t nt;    // tiles of the segment
  uint64_t base;  // first record
  uint64_t end;   // segment end
  uint32_t count;
};

template <int TILE>
__device__ __forceinline__ RxTile rx_tile(const uint64_t* seg, const uint32_t* ts, int nseg, uint32_t t) {
  RxTile x;
  x.s = rx_seg_of(ts, nseg, t);
  x.lt = t - ts[x.s];
  x.nt = ts[x.s + 1] - ts[x.s];
  x.base = seg[x.s] + uint64_t(x.lt) * TILE;
  x.end = seg[x.s + 1];
  x.count = (x.end - x.base) < uint64_t(TILE) ? uint32_t(x.end - x.base) : uint32_t(TILE);
  return x;
}

// Tile of block b in the three-launch kernels with XCD order (scheme 2).
// Blocks go to the 8 XCDs in turn (block b to XCD b % 8); numbering tiles so
// that each XCD takes one contiguous range puts neighbouring tiles in the
// same L2, where the partial lines they share (table entries of adjacent
// tiles, the ends of a digit's runs) are combined before they are written
// back.  Scheme 0 keeps tile = block.
constexpr uint32_t kXcds = 8;

__device__ __forceinline__ uint32_t rx_block_tile(uint32_t b, uint32_t tiles) {
  const uint32_t q = tiles / kXcds, r = tiles % kXcds, x = b % kXcds;
  return x * q + (x < r ? x : r) + b / kXcds;
}

// s_waitcnt immediate of gfx9-family ISAs: lgkmcnt(0), vmcnt and expcnt at
// their maxima (no wait on vector memory)
constexpr unsigned kLgkmZero = 0xC07F;

// Stable in-tile ranking shared by both schemes.  On return: item/dig/rank per
// record of this thread, wcnt[w][d] = records of digit d in waves < w, *tot
// (threads d < nb) = tile count of digit d.  The 16 loads are issued before
// any of them is used (invalid lanes re-read the tile's first record), so a
// wave has all its loads in flight at once.  peer_tab: (NT/64) << BM u64 of
// LDS, free for the duration of the call.

template <int NT, int BM, bool SIN, class R, class KeyFn>
__device__ __forceinline__ void rx_rank(const RxIn<R>& in, int64_t sh, const RxTile& x, KeyFn key, int shift, int bits,
                                        R (&item)[kRxIPT], uint32_t (&dig)[kRxIPT], uint32_t (&rank)[kRxIPT],
                                        RxWc<BM> (*wcnt)[1 << BM], uint32_t* tot, uint64_t* peer_tab) {
  constexpr int W = NT / kWave;
  const int nb = 1 << bits;
  const uint32_t mask = uint32_t(nb - 1);
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  const uint32_t first = uint32_t(w) * (kRxIPT * kWave) + lane;
#pragma unroll
  for (int k = 0; k < kRxIPT; ++k) {
    const uint32_t li = first + uint32_t(k) * kWave;
    item[k] = rx_load<SIN>(in, sh, x.base + (li < x.count ? li : 0u));
  }
#pragma unroll
  for (int k = 0; k < kRxIPT; ++k) dig[k] = uint32_t(key(item[k]) >> shift) & mask;
  // Peers (lanes of this wave with the same digit) through a per-wave LDS
  // table of lane masks: every lane ORs its bit into its digit's entry, reads
  // the entry back and clears it (one LDS round trip per access instead of a
  // ballot per digit bit).
  uint64_t* tab = peer_tab + w * (1 << BM);
  for (int i = lane; i < nb; i += kWave) __hip_atomic_store(tab + i, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
  __builtin_amdgcn_s_waitcnt(kLgkmZero);
  __asm__ volatile("" ::: "memory");
#pragma unroll
  for (int k = 0; k < kRxIPT; ++k) {
    const bool valid = first + uint32_t(k) * kWave < x.count;
    const uint32_t d = dig[k];
    // The compiler sees one lane: to it the accesses of record k and those
    // of record k+1 touch unrelated addresses, while across lanes they alias.
    // Three waits (s_waitcnt lgkmcnt(0)) order them without relying on the
    // LDS completing a wave's accesses in issue order: the read sees every
    // lane's OR, the clear follows every lane's read, and the clear has
    // landed before record k+1's OR of the same entry (the wait after the
    // count read; the compiler emitted it only inside the branch of the
    // count write).  The empty asm statements keep the compiler from moving
    // memory accesses across those points.
    // The lanes sharing the first lane's digit (in a later pass, often a
    // run of one vertex of high degree, whose records the earlier passes
    // brought together) take their peers from one ballot and stay off the
    // table, whose same-entry ORs would serialise; a wave with no other
    // digit skips the table.
    const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
    const bool lead = valid && d == d0;
    const uint64_t m0 = __ballot(lead);
    const bool use_tab = valid && !lead;
    uint64_t peers = lead ? m0 : 0ull;
    if (__ballot(use_tab)) {
      if (use_tab) __hip_atomic_fetch_or(tab + d, uint64_t(1) << lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
      __builtin_amdgcn_s_waitcnt(kLgkmZero);
      __asm__ volatile("" ::: "memory");
      if (use_tab) peers = __hip_atomic_load(tab + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
      __builtin_amdgcn_s_waitcnt(kLgkmZero);
      __asm__ volatile("" ::: "memory");
      if (use_tab) __hip_atomic_store(tab + d, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    }
    const uint32_t prior = wcnt[w][d];
    __builtin_amdgcn_s_waitcnt(kLgkmZero);
    __asm__ volatile("" ::: "memory");
    const uint32_t below = __builtin_amdgcn_mbcnt_hi(uint32_t(peers >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(peers), 0u));
    const uint32_t total = uint32_t(__popcll(peers));
    rank[k] = prior + below;
    if (valid && below + 1 == total) wcnt[w][d] = prior + total;  // highest peer updates
  }
  __syncthreads();
  uint32_t t = 0;
  for (int d = threadIdx.x; d < nb; d += NT) {
#pragma unroll
    for (int ww = 0; ww < W; ++ww) {
      const uint32_t c = wcnt[ww][d];
      wcnt[ww][d] = t;
      t += c;
    }
  }
  *tot = t;
}

// Stage in digit order and write every digit's run contiguously (LDS reads,
// then digits, then stores, so the key lookups of a thread overlap).
template <int NT, int BM, bool SOUT, class R, class KeyFn>
__device__ __forceinline__ void rx_scatter(const RxOut<R>& out, const RxTile& x, KeyFn key, int shift, int bits,
                                           const R (&item)[kRxIPT], const uint32_t (&dig)[kRxIPT],
                                           const uint32_t (&rank)[kRxIPT], RxWc<BM> (*wcnt)[1 << BM],
                                           uint32_t* goff, const uint32_t* dstart, R* stage) {
  const uint32_t mask = uint32_t((1 << bits) - 1);
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  const uint32_t first = uint32_t(w) * (kRxIPT * kWave) + lane;
  // fold the tile's digit starts into the per-wave bases and the global
  // offsets, so each record needs one table read per step
  const int nb = 1 << bits;
  for (int d = threadIdx.x; d < nb; d += NT) {
    const uint32_t ds = dstart[d];
#pragma unroll
    for (int ww = 0; ww < NT / kWave; ++ww) wcnt[ww][d] += ds;
    goff[d] -= ds;  // no wrap: the global prefix of d counts this tile's smaller digits
  }
  __syncthreads();
  // Every LDS read of a step is issued (unconditionally: the digits of lanes
  // past the tile's end are those of a real record) before the first use,
  // and the branches hold only the stores, so a thread's round trips overlap
  // instead of one per record.
  uint32_t at[kRxIPT];
#pragma unroll
  for (int k = 0; k < kRxIPT; ++k) at[k] = wcnt[w][dig[k]] + rank[k];
#pragma unroll
  for (int k = 0; k < kRxIPT; ++k)
    if (first + uint32_t(k) * kWave < x.count) stage[at[k]] = item[k];
  __syncthreads();
  R r[kRxIPT];
  uint32_t d[kRxIPT];
#pragma unroll
  for (int k = 0; k < kRxIPT; ++k) {
    const uint32_t lp = uint32_t(k) * NT + threadIdx.x;
    r[k] = stage[lp < x.count ? lp : 0u];
  }
#pragma unroll
  for (int k = 0; k < kRxIPT; ++k) d[k] = uint32_t(key(r[k]) >> shift) & mask;
  uint32_t go[kRxIPT];
#pragma unroll
  for (int k = 0; k < kRxIPT; ++k) go[k] = goff[d[k]];
#pragma unroll
  for (int k = 0; k < kRxIPT; ++k) {
    const uint32_t lp = uint32_t(k) * NT + threadIdx.x;
    const uint64_t pos = uint64_t(go[k]) + lp;
    // pos < end always holds unless a look-back gave up
    if (lp < x.count && pos < x.end) rx_store<SOUT>(out, pos, r[k]);
  }
}

// dstart[d] = records of digits < d in the tile (tot: thread d's count for
// d < nb; a thread may own digits d and d + NT when nb > NT).
template <int NT>
__device__ __forceinline__ void rx_dstart(int nb, uint32_t tot, uint32_t* wt, uint32_t* dstart) {
  uint32_t btot;
  const uint32_t ex = block_exclusive_scan<NT>(int(threadIdx.x) < nb ? tot : 0u, 0u, OpSum{}, wt, &btot);
  if (int(threadIdx.x) < nb) dstart[threadIdx.x] = ex;
  __syncthreads();
}

// ---------------------------------------------------------- three-launch --
// table index of (tile t of segment s, digit d): ts[s]*nb + d*nt + lt
// Histogram order is free, so a full tile of 8-byte records at an even
// record address is read as 16-byte pairs; each wave counts into its own
// LDS histogram (fewer lanes of a skewed digit on one counter), and the
// wave histograms are added at the end.  With a second key (Key2 other than
// RxNoKey) the same read also counts the first digit of a second sort of
// the same records into hist2 (two sorts of one list share their first
// upsweep).
struct RxNoKey {};

template <int NT, int BM, bool SIN, class R, class KeyFn, class Key2 = RxNoKey>
__global__ __launch_bounds__(NT) void k_rx_up(RxIn<R> in, const uint64_t* __restrict__ seg,
                                              const uint32_t* __restrict__ ts, int nseg, KeyFn key, int shift,
                                              int bits, uint32_t* __restrict__ hist, int xcd_order, Key2 key2 = {},
                                              uint32_t* __restrict__ hist2 = nullptr) {
  constexpr int TILE = NT * kRxIPT;
  constexpr int W = NT / kWave;
  constexpr bool kTwo = !std::is_same<Key2, RxNoKey>::value;
  constexpr int kField = RxKeyField<KeyFn>::value;
  // Key arrays of a later pass hold the keys in the previous digit's order,
  // where a wave's lanes often share a digit (a vertex of high degree
  // repeats): each 16-lane group counts into its own histogram, rows one
  // bank apart, so the lanes of a group contend only among themselves.
  constexpr bool kKeys = SIN && kField >= 0 && !kTwo;
  constexpr int G = kKeys ? 4 : 1;         // histograms per wave
  constexpr int HS = (1 << BM) + (G > 1);  // row stride in words
  __shared__ uint32_t h[W * G * HS];
  __shared__ uint32_t h2[kTwo ? W : 1][kTwo ? (1 << BM) : 1];
  const int nb = 1 << bits;
  for (int i = threadIdx.x; i < W * G * HS; i += NT) h[i] = 0;
  if constexpr (kTwo)
    for (int i = threadIdx.x; i < W * nb; i += NT) h2[i / nb][i % nb] = 0;
  __syncthreads();
  const uint32_t t = xcd_order ? rx_block_tile(blockIdx.x, gridDim.x) : blockIdx.x;
  const RxTile x = rx_tile<TILE>(seg, ts, nseg, t);
  const int64_t sh = rx_tile_shift(in, t);
  const uint32_t mask = uint32_t(nb - 1);
  uint32_t* hw = h + ((threadIdx.x / kWave) * G + (threadIdx.x & (kWave - 1)) / (kWave / G)) * HS;
  if constexpr (kKeys) {
    // field arrays: read the key's array only, four keys per 16-byte load
    // in full tiles
    const uint32_t* fa = kField == 0 ? in.f.s : in.f.d;
    uint32_t v[kRxIPT];
    const uint32_t* p0 = fa + uint64_t(int64_t(x.base) + (sh == kRxStraddle ? 0 : sh));
    const bool quads =
        x.count == uint32_t(TILE) && sh != kRxStraddle && (reinterpret_cast<uintptr_t>(p0) & 15) == 0;
    if (quads) {
      const uint4* q4 = reinterpret_cast<const uint4*>(p0);
#pragma unroll
      for (int k = 0; k < kRxIPT / 4; ++k) {
        const uint4 q = q4[uint32_t(k) * NT + threadIdx.x];
        v[4 * k] = q.x;
        v[4 * k + 1] = q.y;
        v[4 * k + 2] = q.z;
        v[4 * k + 3] = q.w;
      }
    } else {
#pragma unroll
      for (int k = 0; k < kRxIPT; ++k) {
        const uint32_t li = uint32_t(k) * NT + threadIdx.x;
        v[k] = fa[rx_src_index(in, sh, x.base + (li < x.count ? li : 0u))];
      }
    }
    uint32_t dg[kRxIPT];
#pragma unroll
    for (int k = 0; k < kRxIPT; ++k) dg[k] = uint32_t(key.of_field(v[k]) >> shift) & mask;
    if (quads) {
      // a load's four keys are consecutive records: a run of one digit
      // inside it (a vertex of high degree) is one atomic of its length
#pragma unroll
      for (int q = 0; q < kRxIPT / 4; ++q) {
        uint32_t run = 1;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int k = 4 * q + j;
          if (j < 3 && dg[k + 1] == dg[k]) {
            ++run;
          } else {
            atomicAdd(&hw[dg[k]], run);
            run = 1;
          }
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < kRxIPT; ++k)
        if (uint32_t(k) * NT + threadIdx.x < x.count) atomicAdd(&hw[dg[k]], 1u);
    }
  } else {
  R r[kRxIPT];
  bool pairs = false;
  if constexpr (sizeof(R) == 8 && !SIN) {
    pairs = x.count == uint32_t(TILE) && sh != kRxStraddle && ((int64_t(x.base) + sh) & 1) == 0;
    if (pairs) {
      const uint4* p = reinterpret_cast<const uint4*>(in.p + uint64_t(int64_t(x.base) + sh));
#pragma unroll
      for (int k = 0; k < kRxIPT / 2; ++k) {
        const uint4 v = p[uint32_t(k) * NT + threadIdx.x];
        __builtin_memcpy(&r[2 * k], &v, 16);
      }
    }
  }
  if (!pairs) {
#pragma unroll
    for (int k = 0; k < kRxIPT; ++k) {
      const uint32_t li = uint32_t(k) * NT + threadIdx.x;
      r[k] = rx_load<SIN>(in, sh, x.base + (li < x.count ? li : 0u));
    }
  }
#pragma unroll
  for (int k = 0; k < kRxIPT; ++k) {
    const uint32_t d = uint32_t(key(r[k]) >> shift) & mask;
    // full tiles count every record; otherwise record k of this thread is li
    const bool live = pairs || uint32_t(k) * NT + threadIdx.x < x.count;
    if (live) atomicAdd(&hw[d], 1u);
    if constexpr (kTwo)
      if (live) atomicAdd(&h2[threadIdx.x / kWave][uint32_t(key2(r[k]) >> shift) & mask], 1u);
  }
  }
  __syncthreads();
  for (int d = threadIdx.x; d < nb; d += NT) {
    uint32_t c = 0, c2 = 0;
#pragma unroll
    for (int ww = 0; ww < W * G; ++ww) c += h[ww * HS + d];
    if constexpr (kTwo) {
#pragma unroll
      for (int ww = 0; ww < W; ++ww) c2 += h2[ww][d];
    }
    const uint64_t at = uint64_t(ts[x.s]) * nb + uint64_t(d) * x.nt + x.lt;
    hist[at] = c;
    if constexpr (kTwo) hist2[at] = c2;
  }
}

template <int NT, int BM, bool SIN, bool SOUT, class R, class KeyFn>
__global__ __launch_bounds__(NT) void k_rx_down(RxIn<R> in, RxOut<R> out,
                                                const uint64_t* __restrict__ seg, const uint32_t* __restrict__ ts,
                                                int nseg, KeyFn key, int shift, int bits,
                                                const uint32_t* __restrict__ offs, int xcd_order) {
  constexpr int TILE = NT * kRxIPT;
  constexpr int W = NT / kWave;
  static_assert(TILE < 65536, "16-bit digit counters");
  __shared__ RxWc<BM> wcnt[W][1 << BM];
  __shared__ uint32_t goff[1 << BM];
  __shared__ uint32_t dstart[1 << BM];
  __shared__ uint32_t wt[W];
  __shared__ __align__(16) R stage[TILE];
  const int nb = 1 << bits;
  for (int i = threadIdx.x; i < W * nb; i += NT) wcnt[i / nb][i % nb] = 0;
  const uint32_t t = xcd_order ? rx_block_tile(blockIdx.x, gridDim.x) : blockIdx.x;
  const RxTile x = rx_tile<TILE>(seg, ts, nseg, t);
  for (int d = threadIdx.x; d < nb; d += NT) goff[d] = offs[uint64_t(ts[x.s]) * nb + uint64_t(d) * x.nt + x.lt];
  __syncthreads();
  R item[kRxIPT];
  uint32_t dig[kRxIPT], rank[kRxIPT];
  uint32_t tot;
  // the peer table of the ranking lives in the staging buffer, which is
  // first written after the ranking's last barrier
  rx_rank<NT, BM, SIN>(in, rx_tile_shift(in, t), x, key, shift, bits, item, dig, rank, wcnt, &tot,
                       reinterpret_cast<uint64_t*>(stage));
  rx_dstart<NT>(nb, tot, wt, dstart);
  rx_scatter<NT, BM, SOUT>(out, x, key, shift, bits, item, dig, rank, wcnt, goff, dstart, stage);
}

// -------------------------------------------------------------- one-pass --
template <int NT, int BM, bool SIN, class R, class KeyFn>
__global__ __launch_bounds__(NT) void k_rx_hist_all(RxIn<R> in, const uint64_t* __restrict__ seg,
                                                    const uint32_t* __restrict__ ts, int nseg, KeyFn key,
                                                    RxPassArgs pa, uint32_t* __restrict__ ghist) {
  constexpr int TILE = NT * kRxIPT;
  __shared__ uint32_t h[kRxMaxPasses][1 << BM];
  for (int i = threadIdx.x; i < kRxMaxPasses * (1 << BM); i += NT) (&h[0][0])[i] = 0;
  __syncthreads();
  const RxTile x = rx_tile<TILE>(seg, ts, nseg, blockIdx.x);
  const int64_t sh = rx_tile_shift(in, blockIdx.x);
  R r[kRxIPT];
#pragma unroll
  for (int k = 0; k < kRxIPT; ++k) {
    const uint32_t li = uint32_t(k) * NT + threadIdx.x;
    r[k] = rx_load<SIN>(in, sh, x.base + (li < x.count ? li : 0u));
  }
#pragma unroll
  for (int k = 0; k < kRxIPT; ++k) {
    if (uint32_t(k) * NT + threadIdx.x < x.count) {
      const uint64_t kv = key(r[k]);
      for (int p = 0; p < pa.npass; ++p) atomicAdd(&h[p][uint32_t(kv >> pa.shift[p]) & ((1u << pa.bits[p]) - 1)], 1u);
    }
  }
  __syncthreads();
  for (int p = 0; p < pa.npass; ++p) {
    const int nb = 1 << pa.bits[p];
    uint32_t* g = ghist + (uint64_t(x.s) * kRxMaxPasses + p) * (1 << kRxMaxBits);
    for (int d = threadIdx.x; d < nb; d += NT)
      if (h[p][d]) atomicAdd(&g[d], h[p][d]);
  }
}

template <int NT, int BM, bool SIN, bool SOUT, class R, class KeyFn>
__global__ __launch_bounds__(NT) void k_rx_onepass(RxIn<R> in, RxOut<R> out,
                                                   const uint64_t* __restrict__ seg, const uint32_t* __restrict__ ts,
                                                   int nseg, KeyFn key, int shift, int bits,
                                                   const uint32_t* __restrict__ dbase, uint64_t* __restrict__ status,
                                                   uint32_t* __restrict__ ticket, uint32_t epoch,
                                                   uint32_t* __restrict__ err) {
  constexpr int TILE = NT * kRxIPT;
  constexpr int W = NT / kWave;
  static_assert(TILE < 65536, "16-bit digit counters");
  __shared__ RxWc<BM> wcnt[W][1 << BM];
  __shared__ uint32_t goff[1 << BM];
  __shared__ uint32_t dstart[1 << BM];
  __shared__ uint32_t wt[W];
  __shared__ uint32_t s_tile;
  __shared__ __align__(16) R stage[TILE];
  const int nb = 1 << bits;
  if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
  for (int i = threadIdx.x; i < W * nb; i += NT) wcnt[i / nb][i % nb] = 0;
  __syncthreads();
  const uint32_t t = s_tile;
  const RxTile x = rx_tile<TILE>(seg, ts, nseg, t);
  const uint32_t* db = dbase + uint64_t(x.s) * kRxMaxPasses * (1 << kRxMaxBits);
  R item[kRxIPT];
  uint32_t dig[kRxIPT], rank[kRxIPT];
  uint32_t tot;
  rx_rank<NT, BM, SIN>(in, rx_tile_shift(in, t), x, key, shift, bits, item, dig, rank, wcnt, &tot,
                       reinterpret_cast<uint64_t*>(stage));
  // A status word is its own payload (epoch | flag | count): relaxed
  // device-scope atomics suffice; release/acquire would write back or
  // invalidate the L2 on every publish and poll.
  for (int d = threadIdx.x; d < nb; d += NT) {
    uint64_t* my = status + uint64_t(t) * (1 << kRxMaxBits) + d;
    if (x.lt == 0) {
      __hip_atomic_store(my, rx_pack(epoch, kFlagInc, tot), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      goff[d] = db[d];
    } else {
      __hip_atomic_store(my, rx_pack(epoch, kFlagAgg, tot), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      uint32_t excl = 0, spins = 0;
      uint32_t j = t - 1;
      const uint32_t first = t - x.lt;
      while (true) {
        const uint64_t sw =
            __hip_atomic_load(status + uint64_t(j) * (1 << kRxMaxBits) + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t fl = (sw >> 32) & 3ull;
        if (uint32_t(sw >> 34) != epoch || fl == 0) {
          if (++spins > kRxSpinLimit) {
            atomicOr(err, 1u);
            break;
          }
          continue;
        }
        excl += uint32_t(sw);
        if (fl == kFlagInc || j == first) break;
        --j;
      }
      __hip_atomic_store(my, rx_pack(epoch, kFlagInc, excl + tot), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      goff[d] = db[d] + excl;
    }
  }
  rx_dstart<NT>(nb, tot, wt, dstart);
  rx_scatter<NT, BM, SOUT>(out, x, key, shift, bits, item, dig, rank, wcnt, goff, dstart, stage);
}

namespace {  // internal linkage: this header is included by several sources

__global__ __attribute__((unused)) void k_rx_base(uint32_t* __restrict__ ghist, const uint64_t* __restrict__ seg, int nseg, int npass) {
  const int s = blockIdx.x;
  if (s >= nseg || int(threadIdx.x) >= npass) return;
  uint32_t* g = ghist + (uint64_t(s) * kRxMaxPasses + threadIdx.x) * (1 << kRxMaxBits);
  uint32_t run = uint32_t(seg[s]);
  for (int d = 0; d < (1 << kRxMaxBits); ++d) {
    const uint32_t c = g[d];
    g[d] = run;
    run += c;
  }
}

// per-tile shift of a block view (kRxStraddle where a tile spans blocks)
template <int TILE>
__global__ void k_rx_tile_shift(const uint64_t* __restrict__ seg, const uint32_t* __restrict__ ts, int nseg,
                                uint32_t tiles, const uint64_t* __restrict__ bdst, const uint64_t* __restrict__ bsrc,
                                int nblk, int64_t* __restrict__ out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= tiles) return;
  const RxTile x = rx_tile<TILE>(seg, ts, nseg, t);
  int lo = 0, hi = nblk;  // last block starting at or before the tile
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (bdst[mid] <= x.base)
      lo = mid;
    else
      hi = mid;
  }
  out[t] = bdst[lo + 1] >= x.base + x.count ? int64_t(bsrc[lo]) - int64_t(bdst[lo]) : kRxStraddle;
}

__global__ __attribute__((unused)) void k_rx_zero(uint32_t* __restrict__ p, uint64_t n) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = 0;
}

}  // namespace

// ------------------------------------------------------------------ host --
// Sort variant: scheme (three-launch in block order, one-pass, three-launch
// in XCD order) x tile threads.  LDS per block ~ NT*16*rec bytes, so
// 12-byte records stop at 512 threads (96 KB), 8-byte ones at 1024 (128 KB).
struct SortVariant {
  int scheme = kThreeLaunch;
  int nt = 256;
  int bits = 8;  // widest digit: 8, or 9 with 512 threads (used where it saves a pass)
};

// process-wide choice per record size (eg_set_sort_variant / autotuner)
SortVariant& sort_variant(int rec_bytes);

// Device workspace (grow-only, owned by the caller).
struct RxWork {
  DevBuf table, table2, ghist, status, ticket, segs, err, scan, tshift;
  uint32_t epoch = 0;
  int resize_segs(int nseg) {
    EG_TRY(segs.resize((uint64_t(nseg) + 1) * 12 + 16));
    EG_TRY(ticket.resize(kRxMaxPasses * 4));
    EG_TRY(err.resize(4));
    return EG_OK;
  }
};

// First pass of a three-launch sort: its own upsweep, its own upsweep that
// also counts a second sort's first digit into w.table2 (key2), or no
// upsweep, the counts being in w.table2 already.
enum RxFirst { kFirstOwn = 0, kFirstAlsoSecond = 1, kFirstCounted = 2 };

// Storage of a sort's lists: records throughout (kRecords); the input in
// records and every pass's output in field arrays (kToFields: the result is
// an intermediate list of the build); the input and every intermediate in
// field arrays and the result in records (kFromFields: passes alternate
// between tmp and tmp2, both field arrays, the last writes out); the input
// and the result in records, the intermediates in field arrays
// (kThroughFields, tmp and tmp2 as in kFromFields).
enum RxStorage { kRecords = 0, kToFields = 1, kFromFields = 2, kThroughFields = 3 };

template <class R>
RxIn<R> rx_in_of(const RxOut<R>& o) {
  RxIn<R> in{o.p, nullptr, nullptr, nullptr, 0, o.f};
  return in;
}

template <int NT, int BM, bool SIN, class R, class KeyFn, class Key2>
int rx_launch_up(const RxIn<R>& src, uint32_t tiles, const uint64_t* d_seg, const uint32_t* d_ts, int nseg,
                 KeyFn key, int shift, int bits, uint32_t* tab, int xcd, Key2 key2, uint32_t* tab2, uint64_t n,
                 hipStream_t s) {
  constexpr bool kKeyOnly = SIN && RxKeyField<KeyFn>::value >= 0 && std::is_same<Key2, RxNoKey>::value;
  prof_begin(s);
  hipLaunchKernelGGL((k_rx_up<NT, BM, SIN, R, KeyFn, Key2>), dim3(tiles), dim3(NT), 0, s, src, d_seg, d_ts, nseg, key,
                     shift, bits, tab, xcd, key2, tab2);
  prof_end(PK_RADIX_UPSWEEP, n * (kKeyOnly ? 4 : sizeof(R)), s);
  EG_HIP(hipGetLastError());
  return EG_OK;
}

template <int NT, int BM, bool SIN, bool SOUT, class R, class KeyFn>
int rx_launch_down(const RxIn<R>& src, const RxOut<R>& dst, uint32_t tiles, const uint64_t* d_seg,
                   const uint32_t* d_ts, int nseg, KeyFn key, int shift, int bits, const uint32_t* tab, int xcd,
                   uint64_t n, hipStream_t s) {
  prof_begin(s);
  hipLaunchKernelGGL((k_rx_down<NT, BM, SIN, SOUT, R, KeyFn>), dim3(tiles), dim3(NT), 0, s, src, dst, d_seg, d_ts,
                     nseg, key, shift, bits, tab, xcd);
  prof_end(PK_RADIX_DOWNSWEEP, 2 * n * sizeof(R), s);
  EG_HIP(hipGetLastError());
  return EG_OK;
}

template <int NT, int BM, bool SIN, bool SOUT, class R, class KeyFn>
int rx_launch_onepass(const RxIn<R>& src, const RxOut<R>& dst, uint32_t tiles, const uint64_t* d_seg,
                      const uint32_t* d_ts, int nseg, KeyFn key, int shift, int bits, const uint32_t* dbase,
                      RxWork& w, int p, uint32_t epoch, uint64_t n, hipStream_t s) {
  prof_begin(s);
  hipLaunchKernelGGL((k_rx_onepass<NT, BM, SIN, SOUT, R, KeyFn>), dim3(tiles), dim3(NT), 0, s, src, dst, d_seg, d_ts,
                     nseg, key, shift, bits, dbase, w.status.as<uint64_t>(), w.ticket.as<uint32_t>() + p, epoch,
                     w.err.as<uint32_t>());
  prof_end(PK_RADIX_DOWNSWEEP, 2 * n * sizeof(R), s);
  EG_HIP(hipGetLastError());
  return EG_OK;
}

// pass p of `passes` (storage ST): the input / output is field arrays
template <int ST>
constexpr bool rx_pass_in_fields(int p) {
  return ST == kFromFields || ((ST == kToFields || ST == kThroughFields) && p > 0);
}
template <int ST>
constexpr bool rx_pass_out_fields(int p, int passes) {
  return ST == kToFields || ((ST == kFromFields || ST == kThroughFields) && p != passes - 1);
}

template <int NT, int BM, int ST = kRecords, class R, class KeyFn, class Key2 = RxNoKey>
int rx_sort_nt(RxIn<R> in, RxOut<R> out, RxOut<R> tmp, RxOut<R> tmp2, const std::vector<uint64_t>& seg, KeyFn key,
               int key_bits, int scheme, RxWork& w, hipStream_t s, RxFirst first = kFirstOwn, Key2 key2 = {}) {
  static_assert((1 << BM) <= NT, "one digit per thread in the count/scan steps");
  static_assert(NT * kRxIPT * sizeof(R) + (NT / kWave) * (4 << BM) + (8 << BM) <= 160 * 1024, "LDS of one block");
  static_assert((NT / kWave) * (8 << BM) <= NT * kRxIPT * sizeof(R), "peer tables fit the staging buffer");
  constexpr int TILE = NT * kRxIPT;
  const int nseg = int(seg.size()) - 1;
  std::vector<uint32_t> ts(nseg + 1, 0);
  for (int k = 0; k < nseg; ++k) ts[k + 1] = ts[k] + uint32_t(ceil_div(seg[k + 1] - seg[k], TILE));
  const uint32_t tiles = ts[nseg];
  if (tiles == 0) return EG_OK;
  EG_TRY(w.resize_segs(nseg));
  uint64_t* d_seg = w.segs.as<uint64_t>();
  uint32_t* d_ts = reinterpret_cast<uint32_t*>(d_seg + nseg + 1);
  EG_HIP(hipMemcpyAsync(d_seg, seg.data(), (nseg + 1) * 8, hipMemcpyHostToDevice, s));
  EG_HIP(hipMemcpyAsync(d_ts, ts.data(), (nseg + 1) * 4, hipMemcpyHostToDevice, s));
  const uint64_t n = seg[nseg] - seg[0];
  const RxPlan plan = rx_plan(key_bits, BM);
  const KeyFn quiet = rx_quiet(key);
  if (in.nblk > 0) {
    EG_TRY(w.tshift.resize(uint64_t(tiles) * 8));
    hipLaunchKernelGGL((k_rx_tile_shift<TILE>), dim3(ceil_div(tiles, 256u)), dim3(256), 0, s, (const uint64_t*)d_seg,
                       (const uint32_t*)d_ts, nseg, tiles, in.bdst, in.bsrc, in.nblk, w.tshift.as<int64_t>());
    EG_HIP(hipGetLastError());
    in.shift = w.tshift.as<int64_t>();
  }
  // destination of pass p: kFromFields alternates tmp, tmp2 and ends in out;
  // the others alternate out and tmp so that the last pass writes out
  auto dest = [&](int p) -> const RxOut<R>& {
    if (p == plan.passes - 1) return out;
    if (ST == kFromFields || ST == kThroughFields) return p % 2 == 0 ? tmp : tmp2;
    return (plan.passes - 1 - p) % 2 == 0 ? out : tmp;
  };
  RxIn<R> src = in;
  if (scheme == kThreeLaunch || scheme == kThreeLaunchXcd) {
    const int xcd = scheme == kThreeLaunchXcd;
    const uint64_t tn = uint64_t(tiles) << kRxMaxBits;
    EG_TRY(w.table.resize(tn * 4));
    if (first == kFirstAlsoSecond) EG_TRY(w.table2.resize(tn * 4));
    if (first == kFirstCounted && w.table2.bytes < tn * 4)
      return fail(EG_ERR_STATE, "internal: the first digit's counts were not made for this sort");
    if (w.scan.bytes < scan_tmp_elems(tn) * 4) EG_TRY(w.scan.resize(scan_tmp_elems(tn) * 4));
    for (int p = 0; p < plan.passes; ++p) {
      const RxOut<R>& dst = dest(p);
      uint32_t* tab = p == 0 && first == kFirstCounted ? w.table2.as<uint32_t>() : w.table.as<uint32_t>();
      const bool fin = rx_pass_in_fields<ST>(p), fout = rx_pass_out_fields<ST>(p, plan.passes);
      if (p == 0 && first == kFirstAlsoSecond) {
        EG_TRY((rx_launch_up<NT, BM, ST == kFromFields>(src, tiles, d_seg, d_ts, nseg, key, plan.shift[p],
                                                        plan.bits[p], tab, xcd, key2, w.table2.as<uint32_t>(), n, s)));
      } else if (!(p == 0 && first == kFirstCounted)) {
        const KeyFn kp = p == 0 ? key : quiet;
        if constexpr (ST == kRecords) {
          EG_TRY((rx_launch_up<NT, BM, false>(src, tiles, d_seg, d_ts, nseg, kp, plan.shift[p], plan.bits[p], tab, xcd,
                                              RxNoKey{}, nullptr, n, s)));
        } else {
          if (fin)
            EG_TRY((rx_launch_up<NT, BM, true>(src, tiles, d_seg, d_ts, nseg, kp, plan.shift[p], plan.bits[p], tab,
                                               xcd, RxNoKey{}, nullptr, n, s)));
          else
            EG_TRY((rx_launch_up<NT, BM, false>(src, tiles, d_seg, d_ts, nseg, kp, plan.shift[p], plan.bits[p], tab,
                                                xcd, RxNoKey{}, nullptr, n, s)));
        }
      }
      EG_TRY(exclusive_scan_u32(tab, tab, uint64_t(tiles) << plan.bits[p], w.scan.as<uint32_t>(), s));
      if constexpr (ST == kRecords) {
        EG_TRY((rx_launch_down<NT, BM, false, false>(src, dst, tiles, d_seg, d_ts, nseg, quiet, plan.shift[p],
                                                     plan.bits[p], tab, xcd, n, s)));
      } else if constexpr (ST == kToFields) {
        if (fin)
          EG_TRY((rx_launch_down<NT, BM, true, true>(src, dst, tiles, d_seg, d_ts, nseg, quiet, plan.shift[p],
                                                     plan.bits[p], tab, xcd, n, s)));
        else
          EG_TRY((rx_launch_down<NT, BM, false, true>(src, dst, tiles, d_seg, d_ts, nseg, quiet, plan.shift[p],
                                                      plan.bits[p], tab, xcd, n, s)));
      } else if constexpr (ST == kFromFields) {
        if (fout)
          EG_TRY((rx_launch_down<NT, BM, true, true>(src, dst, tiles, d_seg, d_ts, nseg, quiet, plan.shift[p],
                                                     plan.bits[p], tab, xcd, n, s)));
        else
          EG_TRY((rx_launch_down<NT, BM, true, false>(src, dst, tiles, d_seg, d_ts, nseg, quiet, plan.shift[p],
                                                      plan.bits[p], tab, xcd, n, s)));
      } else {  // kThroughFields
        if (fin && fout)
          EG_TRY((rx_launch_down<NT, BM, true, true>(src, dst, tiles, d_seg, d_ts, nseg, quiet, plan.shift[p],
                                                     plan.bits[p], tab, xcd, n, s)));
        else if (fin)
          EG_TRY((rx_launch_down<NT, BM, true, false>(src, dst, tiles, d_seg, d_ts, nseg, quiet, plan.shift[p],
                                                      plan.bits[p], tab, xcd, n, s)));
        else if (fout)
          EG_TRY((rx_launch_down<NT, BM, false, true>(src, dst, tiles, d_seg, d_ts, nseg, quiet, plan.shift[p],
                                                      plan.bits[p], tab, xcd, n, s)));
        else
          EG_TRY((rx_launch_down<NT, BM, false, false>(src, dst, tiles, d_seg, d_ts, nseg, quiet, plan.shift[p],
                                                       plan.bits[p], tab, xcd, n, s)));
      }
      src = rx_in_of(dst);
    }
    return EG_OK;
  }
  // one-pass
  EG_TRY(w.ghist.resize(uint64_t(nseg) * kRxMaxPasses * (1 << kRxMaxBits) * 4));
  const uint64_t sb = uint64_t(tiles) * (1 << kRxMaxBits) * 8;
  if (w.status.bytes < sb) {
    EG_TRY(w.status.resize(sb));
    EG_HIP(hipMemsetAsync(w.status.ptr, 0, w.status.bytes, s));  // epoch 0 is never used
  }
  RxPassArgs pa;
  pa.npass = plan.passes;
  for (int p = 0; p < kRxMaxPasses; ++p) {
    pa.shift[p] = plan.shift[p];
    pa.bits[p] = plan.bits[p];
  }
  const uint64_t nh = uint64_t(nseg) * kRxMaxPasses * (1 << kRxMaxBits);
  hipLaunchKernelGGL(k_rx_zero, dim3(unsigned(std::min<uint64_t>(ceil_div(nh, 256), 4096))), dim3(256), 0, s,
                     w.ghist.as<uint32_t>(), nh);
  hipLaunchKernelGGL(k_rx_zero, dim3(1), dim3(64), 0, s, w.ticket.as<uint32_t>(), uint64_t(kRxMaxPasses));
  prof_begin(s);
  hipLaunchKernelGGL((k_rx_hist_all<NT, BM, ST == kFromFields, R, KeyFn>), dim3(tiles), dim3(NT), 0, s, in,
                     (const uint64_t*)d_seg, (const uint32_t*)d_ts, nseg, key, pa, w.ghist.as<uint32_t>());
  prof_end(PK_RADIX_UPSWEEP, n * sizeof(R), s);
  EG_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_rx_base, dim3(nseg), dim3(64), 0, s, w.ghist.as<uint32_t>(), (const uint64_t*)d_seg, nseg,
                     plan.passes);
  EG_HIP(hipGetLastError());
  for (int p = 0; p < plan.passes; ++p) {
    const RxOut<R>& dst = dest(p);
    const uint32_t epoch = ++w.epoch;
    const uint32_t* db = (const uint32_t*)w.ghist.as<uint32_t>() + uint64_t(p) * (1 << kRxMaxBits);
    const bool fin = rx_pass_in_fields<ST>(p), fout = rx_pass_out_fields<ST>(p, plan.passes);
    if constexpr (ST == kRecords) {
      EG_TRY((rx_launch_onepass<NT, BM, false, false>(src, dst, tiles, d_seg, d_ts, nseg, quiet, plan.shift[p],
                                                      plan.bits[p], db, w, p, epoch, n, s)));
    } else if constexpr (ST == kToFields) {
      if (fin)
        EG_TRY((rx_launch_onepass<NT, BM, true, true>(src, dst, tiles, d_seg, d_ts, nseg, quiet, plan.shift[p],
                                                      plan.bits[p], db, w, p, epoch, n, s)));
      else
        EG_TRY((rx_launch_onepass<NT, BM, false, true>(src, dst, tiles, d_seg, d_ts, nseg, quiet, plan.shift[p],
                                                       plan.bits[p], db, w, p, epoch, n, s)));
    } else if constexpr (ST == kFromFields) {
      if (fout)
        EG_TRY((rx_launch_onepass<NT, BM, true, true>(src, dst, tiles, d_seg, d_ts, nseg, quiet, plan.shift[p],
                                                      plan.bits[p], db, w, p, epoch, n, s)));
      else
        EG_TRY((rx_launch_onepass<NT, BM, true, false>(src, dst, tiles, d_seg, d_ts, nseg, quiet, plan.shift[p],
                                                       plan.bits[p], db, w, p, epoch, n, s)));
    } else {  // kThroughFields
      if (fin && fout)
        EG_TRY((rx_launch_onepass<NT, BM, true, true>(src, dst, tiles, d_seg, d_ts, nseg, quiet, plan.shift[p],
                                                      plan.bits[p], db, w, p, epoch, n, s)));
      else if (fin)
        EG_TRY((rx_launch_onepass<NT, BM, true, false>(src, dst, tiles, d_seg, d_ts, nseg, quiet, plan.shift[p],
                                                       plan.bits[p], db, w, p, epoch, n, s)));
      else if (fout)
        EG_TRY((rx_launch_onepass<NT, BM, false, true>(src, dst, tiles, d_seg, d_ts, nseg, quiet, plan.shift[p],
                                                       plan.bits[p], db, w, p, epoch, n, s)));
      else
        EG_TRY((rx_launch_onepass<NT, BM, false, false>(src, dst, tiles, d_seg, d_ts, nseg, quiet, plan.shift[p],
                                                        plan.bits[p], db, w, p, epoch, n, s)));
    }
    src = rx_in_of(dst);
  }
  return EG_OK;
}

// Sort every segment [seg[k], seg[k+1]) of `in` stably by key into the same
// positions of `out` (tmp = ping-pong; in, out, tmp distinct; in unchanged;
// tmp2: the second field-array temporary of kFromFields, which may be the
// input's storage).  The caller zeroes w.err before and checks it after
// synchronising.
template <int ST = kRecords, class R, class KeyFn>
int rx_sort_with(const SortVariant& v, RxIn<R> in, RxOut<R> out, RxOut<R> tmp, RxOut<R> tmp2,
                 const std::vector<uint64_t>& seg, KeyFn key, int key_bits, RxWork& w, hipStream_t s) {
  const int nseg = int(seg.size()) - 1;
  if (nseg < 1 || seg.back() == seg.front()) return EG_OK;
  if (seg.front() != 0) return fail(EG_ERR_ARG, "internal: sort segments must start at 0");
  if (seg.back() >= (uint64_t(1) << 32)) return fail(EG_ERR_ARG, "sort of more than 2^32-1 records");
  if (v.nt == 256) return rx_sort_nt<256, 8, ST>(in, out, tmp, tmp2, seg, key, key_bits, v.scheme, w, s);
  if constexpr (sizeof(R) <= 8) {
    if (v.nt >= 1024) return rx_sort_nt<1024, 8, ST>(in, out, tmp, tmp2, seg, key, key_bits, v.scheme, w, s);
  }
  // 9-bit digits only where they save a pass (25-27-bit keys: 3 passes
  // instead of 4); otherwise the 8-bit kernel, whose scatter writes longer runs
  if (v.bits >= 9 && rx_plan(key_bits, 9).passes < rx_plan(key_bits, 8).passes)
    return rx_sort_nt<512, 9, ST>(in, out, tmp, tmp2, seg, key, key_bits, v.scheme, w, s);
  return rx_sort_nt<512, 8, ST>(in, out, tmp, tmp2, seg, key, key_bits, v.scheme, w, s);
}

template <class R, class KeyFn>
int rx_sort_with(const SortVariant& v, RxIn<R> in, R* out, R* tmp, const std::vector<uint64_t>& seg, KeyFn key,
                 int key_bits, RxWork& w, hipStream_t s) {
  RxOut<R> o, t;
  o.p = out;
  t.p = tmp;
  return rx_sort_with<kRecords>(v, in, o, t, RxOut<R>{}, seg, key, key_bits, w, s);
}

template <class R, class KeyFn>
int rx_sort(RxIn<R> in, R* out, R* tmp, const std::vector<uint64_t>& seg, KeyFn key, int key_bits, RxWork& w,
            hipStream_t s) {
  return rx_sort_with(sort_variant(int(sizeof(R))), in, out, tmp, seg, key, key_bits, w, s);
}

// Two stable sorts of the same records (plain input) by key_a into out_a and
// by key_b into out_b, each as rx_sort would do it.  Three-launch variants
// read the input once for both first digits (the first key's first upsweep
// also counts the second's); the one-pass variant runs the two sorts.
// ST: kToFields (the results as field arrays) or kThroughFields (records
// in and out, the intermediate passes as field arrays).
template <int NT, int BM, int ST, class R, class KeyA, class KeyB>
int rx_sort_pair_nt(const R* in, RxOut<R> out_a, RxOut<R> out_b, RxOut<R> tmp, RxOut<R> tmp2,
                    const std::vector<uint64_t>& seg, KeyA key_a, KeyB key_b, int key_bits, int scheme, RxWork& w,
                    hipStream_t s) {
  const RxIn<R> src{in, nullptr, nullptr, nullptr, 0, {}};
  if (scheme == kOnePass) {
    EG_TRY((rx_sort_nt<NT, BM, ST>(src, out_a, tmp, tmp2, seg, key_a, key_bits, scheme, w, s)));
    return rx_sort_nt<NT, BM, ST>(src, out_b, tmp, tmp2, seg, key_b, key_bits, scheme, w, s);
  }
  EG_TRY((rx_sort_nt<NT, BM, ST>(src, out_a, tmp, tmp2, seg, key_a, key_bits, scheme, w, s, kFirstAlsoSecond,
                                rx_quiet(key_b))));
  return rx_sort_nt<NT, BM, ST>(src, out_b, tmp, tmp2, seg, key_b, key_bits, scheme, w, s, kFirstCounted);
}

template <int ST, class R, class KeyA, class KeyB>
int rx_sort_pair_st(const R* in, RxOut<R> out_a, RxOut<R> out_b, RxOut<R> tmp, RxOut<R> tmp2,
                    const std::vector<uint64_t>& seg, KeyA key_a, KeyB key_b, int key_bits, RxWork& w,
                    hipStream_t s) {
  const SortVariant v = sort_variant(int(sizeof(R)));
  const int nseg = int(seg.size()) - 1;
  if (nseg < 1 || seg.back() == seg.front()) return EG_OK;
  if (seg.front() != 0) return fail(EG_ERR_ARG, "internal: sort segments must start at 0");
  if (seg.back() >= (uint64_t(1) << 32)) return fail(EG_ERR_ARG, "sort of more than 2^32-1 records");
  if (v.nt == 256)
    return rx_sort_pair_nt<256, 8, ST>(in, out_a, out_b, tmp, tmp2, seg, key_a, key_b, key_bits, v.scheme, w, s);
  if constexpr (sizeof(R) <= 8) {
    if (v.nt >= 1024)
      return rx_sort_pair_nt<1024, 8, ST>(in, out_a, out_b, tmp, tmp2, seg, key_a, key_b, key_bits, v.scheme, w, s);
  }
  if (v.bits >= 9 && rx_plan(key_bits, 9).passes < rx_plan(key_bits, 8).passes)
    return rx_sort_pair_nt<512, 9, ST>(in, out_a, out_b, tmp, tmp2, seg, key_a, key_b, key_bits, v.scheme, w, s);
  return rx_sort_pair_nt<512, 8, ST>(in, out_a, out_b, tmp, tmp2, seg, key_a, key_b, key_bits, v.scheme, w, s);
}

// Two sorts of the same records as above with the two results (and the
// ping-pong) as field arrays.
template <class R, class KeyA, class KeyB>
int rx_sort_pair_fields(const R* in, RxSoa<R> out_a, RxSoa<R> out_b, RxSoa<R> tmp, const std::vector<uint64_t>& seg,
                        KeyA key_a, KeyB key_b, int key_bits, RxWork& w, hipStream_t s) {
  RxOut<R> a, b, t;
  a.f = out_a;
  b.f = out_b;
  t.f = tmp;
  return rx_sort_pair_st<kToFields>(in, a, b, t, RxOut<R>{}, seg, key_a, key_b, key_bits, w, s);
}

// The same with records in and out and the intermediate passes in the field
// arrays tmp and tmp2 (which must not overlap in, out_a or out_b).
template <class R, class KeyA, class KeyB>
int rx_sort_pair_through(const R* in, R* out_a, R* out_b, RxSoa<R> tmp, RxSoa<R> tmp2,
                         const std::vector<uint64_t>& seg, KeyA key_a, KeyB key_b, int key_bits, RxWork& w,
                         hipStream_t s) {
  RxOut<R> a, b, t, t2;
  a.p = out_a;
  b.p = out_b;
  t.f = tmp;
  t2.f = tmp2;
  return rx_sort_pair_st<kThroughFields>(in, a, b, t, t2, seg, key_a, key_b, key_bits, w, s);
}

// Device workspace a three-launch sort of these segments takes with tiles of
// `tile` records: the digit table (every tile, widest digit) and its scan
// temporaries, in bytes; the segment tables are RxWork::resize_segs's.
inline uint64_t rx_three_launch_table_bytes(const std::vector<uint64_t>& seg, uint64_t tile) {
  uint64_t tiles = 0;
  for (size_t k = 0; k + 1 < seg.size(); ++k) tiles += ceil_div(seg[k + 1] - seg[k], tile);
  return (tiles << kRxMaxBits) * 4;
}
inline uint64_t rx_three_launch_scan_bytes(const std::vector<uint64_t>& seg, uint64_t tile) {
  return scan_tmp_elems(rx_three_launch_table_bytes(seg, tile) / 4) * 4;
}

// plain input
template <class R, class KeyFn>
int rx_sort(const R* in, R* out, R* tmp, const std::vector<uint64_t>& seg, KeyFn key, int key_bits, RxWork& w,
            hipStream_t s) {
  return rx_sort(RxIn<R>{in, nullptr, nullptr, nullptr, 0, {}}, out, tmp, seg, key, key_bits, w, s);
}

// Sort of records into field arrays: every pass writes field arrays (out or
// tmp, alternating so that the last writes out), the later ones read only
// the key's array in their upsweeps.
template <class R, class KeyFn>
int rx_sort_to_fields(const R* in, RxSoa<R> out, RxSoa<R> tmp, const std::vector<uint64_t>& seg, KeyFn key,
                      int key_bits, RxWork& w, hipStream_t s) {
  RxOut<R> o, t;
  o.f = out;
  t.f = tmp;
  return rx_sort_with<kToFields>(sort_variant(int(sizeof(R))), RxIn<R>{in, nullptr, nullptr, nullptr, 0, {}}, o, t,
                                 RxOut<R>{}, seg, key, key_bits, w, s);
}

// Sort of a block view: the input in sort order is nblk contiguous blocks,
// block b = in[bsrc[b], bsrc[b] + bdst[b+1] - bdst[b]) placed at bdst[b]
// (device tables; bdst[0] == 0, bdst[nblk] == the records).  The first pass
// reads the blocks in place, so no regrouping copy is made; tmp must not
// overlap in.
template <class R, class KeyFn>
int rx_sort_blocks(const R* in, const uint64_t* bdst, const uint64_t* bsrc, int nblk, R* out, R* tmp,
                   const std::vector<uint64_t>& seg, KeyFn key, int key_bits, RxWork& w, hipStream_t s) {
  if (nblk < 1) return fail(EG_ERR_ARG, "internal: a block view needs at least one block");
  return rx_sort(RxIn<R>{in, nullptr, bdst, bsrc, nblk, {}}, out, tmp, seg, key, key_bits, w, s);
}

// The same over field arrays: the blocks and the intermediate passes are
// field arrays (tmp, and tmp2, which may be in's storage, since the first
// pass has read in before the second writes tmp2); the result is records.
template <class R, class KeyFn>
int rx_sort_blocks_fields(RxSoa<R> in, const uint64_t* bdst, const uint64_t* bsrc, int nblk, R* out, RxSoa<R> tmp,
                          RxSoa<R> tmp2, const std::vector<uint64_t>& seg, KeyFn key, int key_bits, RxWork& w,
                          hipStream_t s) {
  if (nblk < 1) return fail(EG_ERR_ARG, "internal: a block view needs at least one block");
  RxOut<R> o, t, t2;
  o.p = out;
  t.f = tmp;
  t2.f = tmp2;
  return rx_sort_with<kFromFields>(sort_variant(int(sizeof(R))), RxIn<R>{nullptr, nullptr, bdst, bsrc, nblk, in}, o, t,
                                   t2, seg, key, key_bits, w, s);
}

}  // namespace eg
