// Device-wide prefix scans (reduce-then-scan).  Used for the page-weight
// prefix, the radix-sort digit offsets and the compressed-index compaction.
#include "common.hpp"

namespace eg {
namespace {

constexpr int kNT = 256;        // threads per block
constexpr int kIPT = 16;        // items per thread
constexpr int kTile = kNT * kIPT;

template <class T, class Op>
__global__ __launch_bounds__(kNT) void k_tile_reduce(const T* __restrict__ in, uint64_t n, T identity, Op op,
                                                     T* __restrict__ partial) {
  __shared__ T wt[kNT / kWave];
  const uint64_t base = uint64_t(blockIdx.x) * kTile;
  T acc = identity;
#pragma unroll
  for (int k = 0; k < kIPT; ++k) {
    uint64_t i = base + uint64_t(k) * kNT + threadIdx.x;  // striped: coalesced
    if (i < n) acc = op(acc, in[i]);
  }
  acc = wave_reduce(acc, op);
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  if (lane == 0) wt[w] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    T t = identity;
    for (int i = 0; i < kNT / kWave; ++i) t = op(t, wt[i]);
    partial[blockIdx.x] = t;
  }
}

// Exclusive scan of `n` partials in place by one block.
template <class T, class Op>
__global__ __launch_bounds__(kNT) void k_scan_partials(T* __restrict__ partial, uint64_t n, T identity, Op op) {
  __shared__ T wt[kNT / kWave];
  T carry = identity;
  for (uint64_t base = 0; base < n; base += kNT) {
    uint64_t i = base + threadIdx.x;
    T v = i < n ? partial[i] : identity;
    T tot;
    T exc = block_exclusive_scan<kNT>(v, identity, op, wt, &tot);
    if (i < n) partial[i] = op(carry, exc);
    carry = op(carry, tot);
    __syncthreads();
  }
}

// Each block scans its tile (staged through LDS so that global reads and
// writes stay coalesced) and adds the block's prefix.
template <class T, class Op, bool kInclusive>
__global__ __launch_bounds__(kNT) void k_tile_scan(const T* __restrict__ in, T* __restrict__ out, uint64_t n,
                                                   const T* __restrict__ partial, T identity, Op op) {
  __shared__ T buf[kTile + kTile / 32];
  __shared__ T wt[kNT / kWave];
  const uint64_t base = uint64_t(blockIdx.x) * kTile;
  auto pad = [](int i) { return i + (i >> 5); };
#pragma unroll
  for (int k = 0; k < kIPT; ++k) {
    int li = k * kNT + threadIdx.x;
    uint64_t i = base + li;
    buf[pad(li)] = i < n ? in[i] : identity;
  }
  __syncthreads();
  T v[kIPT];
  T acc = identity;
#pragma unroll
  for (int k = 0; k < kIPT; ++k) {
    v[k] = buf[pad(threadIdx.x * kIPT + k)];
    acc = op(acc, v[k]);
  }
  T tot;
  T exc = block_exclusive_scan<kNT>(acc, identity, op, wt, &tot);
  T run = op(partial[blockIdx.x], exc);
#pragma unroll
  for (int k = 0; k < kIPT; ++k) {
    if (kInclusive) {
      run = op(run, v[k]);
      buf[pad(threadIdx.x * kIPT + k)] = run;
    } else {
      buf[pad(threadIdx.x * kIPT + k)] = run;
      run = op(run, v[k]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kIPT; ++k) {
    int li = k * kNT + threadIdx.x;
    uint64_t i = base + li;
    if (i < n) out[i] = buf[pad(li)];
  }
}

template <class T, class Op, bool kInclusive>
int run_scan(const T* in, T* out, uint64_t n, T* tmp, T identity, Op op, hipStream_t s) {
  if (n == 0) return EG_OK;
  const uint64_t tiles = ceil_div(n, kTile);
  if (tiles > 0x7fffffffull) return fail(EG_ERR_ARG, "scan of %llu elements is too large", (unsigned long long)n);
  hipLaunchKernelGGL((k_tile_reduce<T, Op>), dim3(unsigned(tiles)), dim3(kNT), 0, s, in, n, identity, op, tmp);
  EG_HIP(hipGetLastError());
  hipLaunchKernelGGL((k_scan_partials<T, Op>), dim3(1), dim3(kNT), 0, s, tmp, tiles, identity, op);
  EG_HIP(hipGetLastError());
  hipLaunchKernelGGL((k_tile_scan<T, Op, kInclusive>), dim3(unsigned(tiles)), dim3(kNT), 0, s, in, out, n,
                     (const T*)tmp, identity, op);
  EG_HIP(hipGetLastError());
  return EG_OK;
}

}  // namespace

uint64_t scan_tmp_elems(uint64_t n) { return ceil_div(n, kTile) + 1; }

int exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* tmp, hipStream_t s) {
  return run_scan<uint32_t, OpSum, false>(in, out, n, tmp, 0u, OpSum{}, s);
}
int exclusive_scan_u64(const uint64_t* in, uint64_t* out, uint64_t n, uint64_t* tmp, hipStream_t s) {
  return run_scan<uint64_t, OpSum, false>(in, out, n, tmp, uint64_t(0), OpSum{}, s);
}
int inclusive_max_scan_u64(const uint64_t* in, uint64_t* out, uint64_t n, uint64_t* tmp, hipStream_t s) {
  return run_scan<uint64_t, OpMax, true>(in, out, n, tmp, uint64_t(0), OpMax{}, s);
}

}  // namespace eg
