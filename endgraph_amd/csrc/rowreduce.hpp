// Segmented row reduction over edge lists grouped into rows (the engine's
// dense / pull kernels).  Edges are walked 64 per step: a wave takes the
// words [w0, w0 + K) of the row-start bitmap (bit e set iff edge e starts a
// row; wordrank[w] = rows started before word w), loads the row bits,
// sources and gathered values of its whole range first (every load in
// flight before the first use), reduces rows with a segmented wave scan
// carried across words, and hands each finished row's partial to
// emit(row, value, shared); shared is true for the rows cut by the range
// ends, which need atomics.
#pragma once

#include <type_traits>

#include "common.hpp"

namespace eg {
namespace {

constexpr int kRowWordsPerWave = 16;  // 1024 edges per wave: 8 and 32 measured slower / no better

// first word of the calling wave's range in a one-range-per-wave grid
template <int K = kRowWordsPerWave>
__device__ __forceinline__ uint64_t first_word() {
  return (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave * K;
}

// Lane value moved by a DPP control (gfx9 wave64: row_shr:n = 0x110 + n
// within rows of 16 lanes, row_bcast:15 = 0x142 and row_bcast:31 = 0x143
// across rows).  Lanes whose source is outside the row, or whose row is not
// in row_mask, get `old`.  T is 4 or 8 bytes (moved as 32-bit halves).
template <int kCtrl, int kRowMask, class T>
__device__ __forceinline__ T dpp_move(T old, T v) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "4- or 8-byte lane values");
  int o[sizeof(T) / 4], x[sizeof(T) / 4], r[sizeof(T) / 4];
  __builtin_memcpy(o, &old, sizeof(T));
  __builtin_memcpy(x, &v, sizeof(T));
#pragma unroll
  for (int i = 0; i < int(sizeof(T) / 4); ++i) r[i] = __builtin_amdgcn_update_dpp(o[i], x[i], kCtrl, kRowMask, 0xF, false);
  T out;
  __builtin_memcpy(&out, r, sizeof(T));
  return out;
}

// Value of one lane (wave-uniform lane index) in every lane, through the
// scalar unit.
template <class T>
__device__ __forceinline__ T lane_value(T v, int lane) {
  static_assert(sizeof(T) % 4 == 0, "32-bit multiple");
  int x[sizeof(T) / 4];
  __builtin_memcpy(x, &v, sizeof(T));
#pragma unroll
  for (int i = 0; i < int(sizeof(T) / 4); ++i) x[i] = __builtin_amdgcn_readlane(x[i], lane);
  T out;
  __builtin_memcpy(&out, x, sizeof(T));
  return out;
}

// Inclusive segmented scan over the wave: start = the lane of the segment's
// head at or below this lane (-1: the segment began before lane 0).  Shifts
// by 1, 2, 4, 8 inside rows of 16 lanes, then the last lane of row 0 / 2 is
// broadcast to rows 1 / 3 and the last lane of row 1 to rows 2 and 3; a lane
// takes a moved value only when its source lies in its own segment.  DPP
// moves stay in the VALU (no LDS round trip), and the takes are selects, so
// the scan is branch-free.
template <class T, class Op>
__device__ __forceinline__ T seg_scan(T v, int start, T identity, Op op) {
  const int lane = threadIdx.x & (kWave - 1);
  T t, r;
  t = dpp_move<0x111, 0xF>(identity, v);
  r = op(t, v);
  v = lane - 1 >= start ? r : v;
  t = dpp_move<0x112, 0xF>(identity, v);
  r = op(t, v);
  v = lane - 2 >= start ? r : v;
  t = dpp_move<0x114, 0xF>(identity, v);
  r = op(t, v);
  v = lane - 4 >= start ? r : v;
  t = dpp_move<0x118, 0xF>(identity, v);
  r = op(t, v);
  v = lane - 8 >= start ? r : v;
  t = dpp_move<0x142, 0xA>(identity, v);
  r = op(t, v);
  v = (lane & 16) && start < (lane & ~15) ? r : v;
  t = dpp_move<0x143, 0xC>(identity, v);
  r = op(t, v);
  v = (lane & 32) && start < 32 ? r : v;
  return v;
}

// No second per-edge stream: val(e, source).
struct NoEdgeValue {};

// ev (optional): a second per-edge stream (weights), loaded like the sources
// (16-byte vectors when aligned) and handed to val(e, source, ev[e]).
// AllLanes: every load is unconditional (addresses of lanes past the end
// clamped to the range's first word and the list's first edges, their
// values then dropped) and val is called for every lane, those past the end
// with source 0, so that the loads and the arithmetic on them are one
// straight-line block whose loads all issue before the first use.  Needs src
// 16-byte aligned and readable up to E rounded up to a multiple of 4, and
// val accepting source 0 (every graph with an edge has vertex 0).
template <class T, int K = kRowWordsPerWave, class EV = NoEdgeValue, bool AllLanes = false, class Op, class Val,
          class Emit>
__device__ __forceinline__ void row_reduce_wave(uint64_t w0, uint64_t E, const uint32_t* __restrict__ src,
                                                const uint64_t* __restrict__ rowflag,
                                                const uint32_t* __restrict__ wordrank, T identity, Op op, Val val,
                                                Emit emit, const EV* __restrict__ ev = nullptr) {
  // The wave's K words are taken in rounds of 4 words (256 edges); in a
  // round lane l holds the 4 consecutive edges 4l..4l+3 (one 16-byte load
  // of sources), reduces them serially, hands the open end of its edges to
  // one segmented scan across the lanes, and emits the rows that end in it.
  constexpr int IPL = 4;                 // edges per lane per round
  constexpr int RW = kWave * IPL / 64;   // words per round
  constexpr int RR = K / RW;             // rounds
  static_assert(K % RW == 0, "whole rounds");
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t nwords = ceil_div(E, 64);
  if (w0 >= nwords) return;
  const int nwd = w0 + K < nwords ? K : int(nwords - w0);
  const int lw = lane / (64 / IPL);       // word of this lane's edges within a round
  const int pos0 = (lane % (64 / IPL)) * IPL;  // bit of its first edge in that word
  const bool vec = (reinterpret_cast<uintptr_t>(src) & 15) == 0;
  // every load of the range first, so all are in flight before the first use
  uint64_t f[RR];
  uint32_t base[RR], sv[RR][IPL];
  T v[RR][IPL];
#pragma unroll
  for (int r = 0; r < RR; ++r) {
    const int k = r * RW + lw;
    const uint64_t wd = w0 + k;
    const bool ok = k < nwd;
    const uint64_t e0 = wd * 64 + pos0;
    if constexpr (AllLanes) {
      const uint64_t wdc = ok ? wd : w0;
      const uint64_t fl = rowflag[wdc];
      const uint32_t bs = wordrank[wdc];
      f[r] = ok ? fl : 0ull;
      base[r] = ok ? bs : 0u;
      const bool in = ok && e0 < E;
      const uint4 q = *reinterpret_cast<const uint4*>(src + (in ? e0 : 0));
      sv[r][0] = in ? q.x : 0u;
      sv[r][1] = in && e0 + 1 < E ? q.y : 0u;
      sv[r][2] = in && e0 + 2 < E ? q.z : 0u;
      sv[r][3] = in && e0 + 3 < E ? q.w : 0u;
      continue;
    }
    f[r] = ok ? rowflag[wd] : 0ull;
    base[r] = ok ? wordrank[wd] : 0u;
    if (ok && vec && e0 + IPL <= E) {
      const uint4 q = *reinterpret_cast<const uint4*>(src + e0);
      sv[r][0] = q.x;
      sv[r][1] = q.y;
      sv[r][2] = q.z;
      sv[r][3] = q.w;
    } else {
#pragma unroll
      for (int i = 0; i < IPL; ++i) sv[r][i] = ok && e0 + i < E ? src[e0 + i] : 0u;
    }
  }
  constexpr bool kEV = !std::is_same<EV, NoEdgeValue>::value;
  constexpr int kEvSize = kEV ? int(sizeof(EV)) : 4;
  EV evv[kEV ? RR : 1][kEV ? IPL : 1];
  if constexpr (kEV) {
    static_assert((kEvSize * IPL) % 16 == 0, "a lane's edge values are whole 16-byte vectors");
    const bool evec = (reinterpret_cast<uintptr_t>(ev) & 15) == 0;
#pragma unroll
    for (int r = 0; r < RR; ++r) {
      const uint64_t e0 = (w0 + r * RW + lw) * 64 + pos0;
      const bool ok = r * RW + lw < nwd;
      if (ok && evec && e0 + IPL <= E) {
        uint4 q[kEvSize * IPL / 16];
#pragma unroll
        for (int h = 0; h < kEvSize * IPL / 16; ++h) q[h] = reinterpret_cast<const uint4*>(ev + e0)[h];
        __builtin_memcpy(evv[r], q, sizeof(q));
      } else {
#pragma unroll
        for (int i = 0; i < IPL; ++i) evv[r][i] = ok && e0 + i < E ? ev[e0 + i] : EV{};
      }
    }
  }
#pragma unroll
  for (int r = 0; r < RR; ++r) {
    const uint64_t e0 = (w0 + r * RW + lw) * 64 + pos0;
    const bool ok = r * RW + lw < nwd;
#pragma unroll
    for (int i = 0; i < IPL; ++i) {
      if constexpr (AllLanes) {
        static_assert(!kEV, "AllLanes with a per-edge value stream is not needed");
        const T t = val(e0 + i, sv[r][i]);  // sv is 0 past the end
        v[r][i] = ok && e0 + i < E ? t : identity;
      } else if constexpr (kEV) {
        v[r][i] = ok && e0 + i < E ? val(e0 + i, sv[r][i], evv[r][i]) : identity;
      } else {
        v[r][i] = ok && e0 + i < E ? val(e0 + i, sv[r][i]) : identity;
      }
    }
  }
  const uint64_t lt = (1ull << lane) - 1;
  T carry = identity;
  bool carry_live = false;   // a row is open from earlier edges of the range
  bool carry_shared = true;  // ... and it started before the range
  // the row open at the range start (rows started before word w0, minus one;
  // word 0 starts with a head, so this is never used as -1)
  uint32_t carry_row = uint32_t(__builtin_amdgcn_readlane(int(base[0]), 0)) - 1u;
#pragma unroll
  for (int r = 0; r < RR; ++r) {
    if (r * RW >= nwd) break;  // wave-uniform
    const uint32_t bits = uint32_t(f[r] >> pos0) & ((1u << IPL) - 1);
    const uint32_t below = uint32_t(__popcll(f[r] & ((1ull << pos0) - 1)));  // heads before the lane in its word
    // serial pass: pre = edges before the lane's first head (they close the
    // open row), rows between two heads of the lane are emitted here, post
    // = edges from the last head on (the whole lane when it has no head)
    T pre = identity, acc = identity;
    uint32_t nh = 0;
#pragma unroll
    for (int i = 0; i < IPL; ++i) {
      if ((bits >> i) & 1u) {
        if (nh == 0)
          pre = acc;
        else
          emit(base[r] + below + nh - 1, acc, false);
        acc = identity;
        ++nh;
      }
      acc = op(acc, v[r][i]);
    }
    const bool has = nh != 0;
    if (lane == 0) {  // the open row goes on in lane 0
      if (has)
        pre = op(carry, pre);
      else
        acc = op(carry, acc);
    }
    const uint64_t H = __ballot(has);
    const uint64_t upto = lane == 63 ? H : (H & ((2ull << lane) - 1));
    const int start = upto ? 63 - __clzll((long long)upto) : -1;
    const T S = seg_scan(acc, start, identity, op);
    // the open row at the end of the previous lane (lane 0: folded above)
    const T prev = dpp_move<0x138, 0xF>(identity, S);  // wave_shr:1
    if (has) {
      const T close = lane == 0 ? pre : op(prev, pre);
      const bool first_in_round = (H & lt) == 0;
      const bool open_before = lane > 0 || carry_live || (bits & 1u) == 0;
      if (open_before) emit(base[r] + below - 1, close, first_in_round ? carry_shared : false);
    }
    // the row open after this round: lane 63's scan, started at the round's
    // last head (or carried on)
    carry = lane_value(S, 63);
    carry_live = true;
    if (H) {
      const int L = 63 - __clzll((long long)H);
      const uint32_t my_last = base[r] + below + nh - 1;
      carry_row = uint32_t(__builtin_amdgcn_readlane(int(my_last), L));
      carry_shared = false;
    }
  }
  if (lane == 0 && carry_live) emit(carry_row, carry, true);
}

// grid of 256-thread blocks with one K-word range per wave
template <int K = kRowWordsPerWave>
inline unsigned row_grid(uint64_t E) {
  const uint64_t waves = ceil_div(ceil_div(E, 64), K);
  const uint64_t b = ceil_div(waves, 4);
  return unsigned(b < 1 ? 1 : (b > (1u << 30) ? (1u << 30) : b));
}

}  // namespace
}  // namespace eg
