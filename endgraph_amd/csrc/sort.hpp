// Stable LSD radix sort of edge records on gfx950.
//
// The reference sorts with load_balance_radix_sort (core/graph.hpp:465-584):
// a stable LSD sort whose result only depends on stability, so any stable
// sort by the same key yields the identical order.  Here one pass is three
// launches:
//   upsweep   per-tile digit histogram (LDS atomics), digit-major table
//   scan      exclusive scan of the (digit, tile) table = global offsets
//   downsweep re-reads the tile, ranks every record stably inside its tile
//             (64-lane ballot match per round + per-wave running counters),
//             stages the tile in LDS in digit order and writes each digit
//             run contiguously.
// Digits are at most 8 bits; a key of B bits takes ceil(B/8) balanced passes.
#pragma once

#include "common.hpp"

namespace eg {

constexpr int kSortNT = 256;                 // threads per block (4 waves)
constexpr int kSortIPT = 16;                 // records per thread
constexpr int kSortTile = kSortNT * kSortIPT;  // 4096 records per tile
constexpr int kSortWaves = kSortNT / kWave;
constexpr int kMaxDigitBits = 8;

template <class R, class KeyFn>
__global__ __launch_bounds__(kSortNT) void k_radix_upsweep(const R* __restrict__ in, uint64_t n, KeyFn key,
                                                           int shift, int bits, uint32_t* __restrict__ hist,
                                                           uint32_t ntiles) {
  __shared__ uint32_t h[1 << kMaxDigitBits];
  const int nb = 1 << bits;
  for (int i = threadIdx.x; i < nb; i += kSortNT) h[i] = 0;
  __syncthreads();
  const uint64_t base = uint64_t(blockIdx.x) * kSortTile;
  const uint32_t mask = uint32_t(nb - 1);
#pragma unroll 4
  for (int k = 0; k < kSortIPT; ++k) {
    uint64_t i = base + uint64_t(k) * kSortNT + threadIdx.x;
    if (i < n) {
      uint32_t d = uint32_t(key(in[i]) >> shift) & mask;
      atomicAdd(&h[d], 1u);
    }
  }
  __syncthreads();
  for (int d = threadIdx.x; d < nb; d += kSortNT) hist[uint64_t(d) * ntiles + blockIdx.x] = h[d];
}

template <class R, class KeyFn>
__global__ __launch_bounds__(kSortNT) void k_radix_downsweep(const R* __restrict__ in, R* __restrict__ out,
                                                             uint64_t n, KeyFn key, int shift, int bits,
                                                             const uint32_t* __restrict__ offs, uint32_t ntiles) {
  __shared__ uint32_t wcnt[kSortWaves][1 << kMaxDigitBits];
  __shared__ uint32_t goff[1 << kMaxDigitBits];
  __shared__ uint32_t dstart[1 << kMaxDigitBits];
  __shared__ uint32_t wt[kSortWaves];
  __shared__ R stage[kSortTile];

  const int nb = 1 << bits;
  const uint32_t mask = uint32_t(nb - 1);
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  const uint64_t base = uint64_t(blockIdx.x) * kSortTile;
  const uint32_t count = (n - base) < uint64_t(kSortTile) ? uint32_t(n - base) : uint32_t(kSortTile);

  for (int i = threadIdx.x; i < kSortWaves * nb; i += kSortNT) wcnt[i / nb][i % nb] = 0;
  for (int d = threadIdx.x; d < nb; d += kSortNT) goff[d] = offs[uint64_t(d) * ntiles + blockIdx.x];
  __syncthreads();

  // Wave w owns records [w*1024, (w+1)*1024) of the tile, 16 rounds of 64.
  R item[kSortIPT];
  uint32_t rank[kSortIPT];
  uint32_t dig[kSortIPT];
  const uint64_t lt_mask = (uint64_t(1) << lane) - 1;
#pragma unroll
  for (int k = 0; k < kSortIPT; ++k) {
    const uint32_t li = uint32_t(w) * (kSortIPT * kWave) + uint32_t(k) * kWave + lane;
    const bool valid = li < count;
    if (valid) item[k] = in[base + li];
    const uint32_t d = valid ? (uint32_t(key(item[k]) >> shift) & mask) : 0u;
    dig[k] = d;
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < bits; ++b) {
      const bool set = (d >> b) & 1u;
      const uint64_t bal = __ballot(set);
      peers &= set ? bal : ~bal;
    }
    const uint32_t below = uint32_t(__popcll(peers & lt_mask));
    const uint32_t total = uint32_t(__popcll(peers));
    uint32_t prior = 0;
    if (valid) prior = wcnt[w][d];
    rank[k] = prior + below;
    if (valid && below + 1 == total) wcnt[w][d] = prior + total;  // highest peer updates
  }
  __syncthreads();

  // Per-digit: wave-exclusive offsets, tile totals, tile-local digit starts.
  uint32_t tot = 0;
  if (threadIdx.x < nb) {
    const int d = threadIdx.x;
#pragma unroll
    for (int ww = 0; ww < kSortWaves; ++ww) {
      uint32_t c = wcnt[ww][d];
      wcnt[ww][d] = tot;
      tot += c;
    }
  }
  uint32_t btot;
  uint32_t ex = block_exclusive_scan<kSortNT>(threadIdx.x < nb ? tot : 0u, 0u, OpSum{}, wt, &btot);
  if (threadIdx.x < nb) dstart[threadIdx.x] = ex;
  __syncthreads();

#pragma unroll
  for (int k = 0; k < kSortIPT; ++k) {
    const uint32_t li = uint32_t(w) * (kSortIPT * kWave) + uint32_t(k) * kWave + lane;
    if (li < count) stage[dstart[dig[k]] + wcnt[w][dig[k]] + rank[k]] = item[k];
  }
  __syncthreads();

#pragma unroll 4
  for (int k = 0; k < kSortIPT; ++k) {
    const uint32_t lp = uint32_t(k) * kSortNT + threadIdx.x;
    if (lp < count) {
      R r = stage[lp];
      const uint32_t d = uint32_t(key(r) >> shift) & mask;
      out[uint64_t(goff[d]) + (lp - dstart[d])] = r;
    }
  }
}

// Workspace for one sort of up to n records.
inline uint64_t radix_hist_elems(uint64_t n) {
  return (uint64_t(1) << kMaxDigitBits) * ceil_div(n, kSortTile);
}

struct RadixPlan {
  int passes = 0;
  int shift[8] = {0};
  int bits[8] = {0};
};

inline RadixPlan radix_plan(int key_bits) {
  RadixPlan p;
  if (key_bits <= 0) key_bits = 1;
  p.passes = (key_bits + kMaxDigitBits - 1) / kMaxDigitBits;
  const int dbits = (key_bits + p.passes - 1) / p.passes;
  for (int i = 0; i < p.passes; ++i) {
    p.shift[i] = i * dbits;
    p.bits[i] = dbits < key_bits - i * dbits ? dbits : key_bits - i * dbits;
  }
  return p;
}

// Stable sort of n records from `in` into `out` by key (key_bits wide), using
// `tmp` as the ping-pong buffer.  in, out and tmp must be distinct; `in` is
// not modified.  hist: radix_hist_elems(n) u32; scan_tmp: scan_tmp_elems of
// that count.
template <class R, class KeyFn>
int radix_sort(const R* in, R* out, R* tmp, uint64_t n, KeyFn key, int key_bits, uint32_t* hist,
               uint32_t* scan_tmp, hipStream_t s) {
  if (n == 0) return EG_OK;
  if (n >= (uint64_t(1) << 32))
    return fail(EG_ERR_ARG, "radix_sort: %llu records exceed the 2^32 limit of one sort", (unsigned long long)n);
  const RadixPlan plan = radix_plan(key_bits);
  const uint32_t ntiles = uint32_t(ceil_div(n, kSortTile));
  const R* src = in;
  for (int p = 0; p < plan.passes; ++p) {
    R* dst = ((plan.passes - 1 - p) % 2 == 0) ? out : tmp;
    const uint64_t nh = (uint64_t(1) << plan.bits[p]) * ntiles;
    prof_begin(s);
    hipLaunchKernelGGL((k_radix_upsweep<R, KeyFn>), dim3(ntiles), dim3(kSortNT), 0, s, src, n, key, plan.shift[p],
                       plan.bits[p], hist, ntiles);
    prof_end(PK_RADIX_UPSWEEP, n * sizeof(R), s);
    EG_HIP(hipGetLastError());
    prof_begin(s);
    EG_TRY(exclusive_scan_u32(hist, hist, nh, scan_tmp, s));
    prof_end(PK_SCAN, nh * 8, s);
    prof_begin(s);
    hipLaunchKernelGGL((k_radix_downsweep<R, KeyFn>), dim3(ntiles), dim3(kSortNT), 0, s, src, dst, n, key,
                       plan.shift[p], plan.bits[p], (const uint32_t*)hist, ntiles);
    // algorithmic traffic: every record read once and written once
    prof_end(PK_RADIX_DOWNSWEEP, 2 * n * sizeof(R), s);
    EG_HIP(hipGetLastError());
    src = dst;
  }
  return EG_OK;
}

}  // namespace eg
