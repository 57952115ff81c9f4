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

#include "common.hpp"

namespace eg {
namespace {

constexpr int kRowWordsPerWave = 8;  // 4 and 16 measured slower on the pull

// first word of the calling wave's range in a one-range-per-wave grid
__device__ __forceinline__ uint64_t first_word() {
  return (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave * kRowWordsPerWave;
}

template <class T, class Op>
__device__ __forceinline__ T seg_scan(T v, bool head, Op op) {
  const int lane = threadIdx.x & (kWave - 1);
  bool h = head;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const T ov = shfl_up_t(v, d);
    const bool oh = __shfl_up(int(h), d, kWave) != 0;
    if (lane >= d && !h) {
      v = op(ov, v);
      h = oh;
    }
  }
  return v;
}

template <class T, class Op, class Val, class Emit>
__device__ __forceinline__ void row_reduce_wave(uint64_t w0, uint64_t E, const uint32_t* __restrict__ src,
                                                const uint64_t* __restrict__ rowflag,
                                                const uint32_t* __restrict__ wordrank, T identity, Op op, Val val,
                                                Emit emit) {
  constexpr int K = kRowWordsPerWave;
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t nwords = ceil_div(E, 64);
  if (w0 >= nwords) return;
  const int nwd = w0 + K < nwords ? K : int(nwords - w0);
  // all row bits, sources and gathered values of the wave's range first, so
  // every load of the range is in flight before the first use
  uint64_t f[K];
  uint32_t base[K], sv[K];
  T v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint64_t wd = w0 + k;
    const uint64_t e = wd * 64 + lane;
    const bool ok = k < nwd;
    f[k] = ok ? rowflag[wd] : 0ull;
    base[k] = ok ? wordrank[wd] : 0u;
    sv[k] = ok && e < E ? src[e] : 0u;
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint64_t e = (w0 + k) * 64 + lane;
    v[k] = k < nwd && e < E ? val(e, sv[k]) : identity;
  }
  const uint64_t lt = (1ull << lane) - 1;
  T carry = identity;
  bool carry_live = false;     // a row is open from earlier edges
  bool carry_shared = true;    // ... and it started before this wave's range
  uint32_t carry_row = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (k >= nwd) break;  // wave-uniform
    const uint64_t e = (w0 + k) * 64 + lane;
    T x = v[k];
    const bool head = (f[k] >> lane) & 1ull;
    if (lane == 0 && !head) x = op(carry, x);
    if ((f[k] & 1ull) && carry_live && lane == 0) emit(carry_row, carry, carry_shared);
    const T inc = seg_scan(x, head, op);
    // tails: lanes whose successor starts a row (lane 63 stays open)
    const bool tail = lane < 63 ? ((f[k] >> (lane + 1)) & 1ull) : false;
    const uint64_t heads_le = __popcll(f[k] & (lt | (1ull << lane)));
    const uint32_t row = base[k] + uint32_t(heads_le) - 1;  // rows started before + own
    const bool first_seg_open = heads_le == 0;              // segment continues the carried row
    if (tail && e < E) emit(row, inc, first_seg_open ? carry_shared : false);
    // new carry = lane 63's segment
    const T c63 = shfl_t(inc, 63);
    const uint64_t h63 = __popcll(f[k]);
    carry = c63;
    carry_live = true;
    if (h63 == 0) {
      carry_row = base[k] - 1;  // no head in this word: the open row continues
    } else {
      carry_shared = false;
      carry_row = base[k] + uint32_t(h63) - 1;
    }
  }
  if (lane == 0 && carry_live) emit(carry_row, carry, true);
}

// grid of 256-thread blocks with one K-word range per wave
inline unsigned row_grid(uint64_t E) {
  const uint64_t waves = ceil_div(ceil_div(E, 64), kRowWordsPerWave);
  const uint64_t b = ceil_div(waves, 4);
  return unsigned(b < 1 ? 1 : (b > (1u << 30) ? (1u << 30) : b));
}

}  // namespace
}  // namespace eg
