// Shared definitions for the gfx950 kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include "../../include/endgraph.h"

namespace eg {

// ---------------------------------------------------------------- errors --
// The C-ABI never aborts: every entry point returns an eg_status and leaves a
// message for eg_last_error() (thread-local).
void set_error(const char* fmt, ...);
int fail(int status, const char* fmt, ...);

#define EG_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                         \
      return ::eg::fail(EG_ERR_HIP, "%s failed at %s:%d: %s", #call, __FILE__,    \
                        __LINE__, hipGetErrorString(e_));                         \
  } while (0)

#define EG_TRY(expr)              \
  do {                            \
    int s_ = (expr);              \
    if (s_ != EG_OK) return s_;   \
  } while (0)

// --------------------------------------------------------------- records --
// Record layouts of core/type.hpp:26-40 (EdgeType<float>, EdgeType<Empty>).
struct Rec8 {
  uint32_t src, dst;
};
struct Rec12 {
  uint32_t src, dst;
  float w;
};
static_assert(sizeof(Rec8) == 8, "EdgeType<Empty> is 8 bytes");
static_assert(sizeof(Rec12) == 12, "EdgeType<float> is 12 bytes");

// core/type.hpp:54-59 CompressedAdjIndexType: {u64 index; u32 vertex} packed.
#pragma pack(push, 1)
struct CIdx {
  uint64_t index;
  uint32_t vertex;
};
#pragma pack(pop)
static_assert(sizeof(CIdx) == 12, "CompressedAdjIndexType is 12 bytes");

// Wave-level helpers (CDNA wavefronts are 64 lanes).
constexpr int kWave = 64;

template <class T, class Op>
__device__ __forceinline__ T wave_inclusive_scan(T v, Op op) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    T o = __shfl_up(v, d, kWave);
    if (lane >= d) v = op(o, v);
  }
  return v;
}

template <class T, class Op>
__device__ __forceinline__ T wave_reduce(T v, Op op) {
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) v = op(v, __shfl_xor(v, d, kWave));
  return v;
}

// Exclusive scan across a block of NT threads; `wave_tot` is LDS scratch of
// NT/64 elements.  Returns the exclusive prefix; *total gets the block sum.
template <int NT, class T, class Op>
__device__ __forceinline__ T block_exclusive_scan(T v, T identity, Op op, T* wave_tot, T* total) {
  constexpr int NW = NT / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  T inc = wave_inclusive_scan(v, op);
  if (lane == kWave - 1) wave_tot[w] = inc;
  __syncthreads();
  T wave_prefix = identity, tot = identity;
#pragma unroll
  for (int i = 0; i < NW; ++i) {
    if (i < w) wave_prefix = op(wave_prefix, wave_tot[i]);
    tot = op(tot, wave_tot[i]);
  }
  __syncthreads();
  *total = tot;
  T exc = __shfl_up(inc, 1, kWave);
  if (lane == 0) exc = identity;
  return op(wave_prefix, exc);
}

struct OpSum {
  template <class T>
  __device__ __forceinline__ T operator()(T a, T b) const { return a + b; }
};
struct OpMax {
  template <class T>
  __device__ __forceinline__ T operator()(T a, T b) const { return a > b ? a : b; }
};
struct OpMin {
  template <class T>
  __device__ __forceinline__ T operator()(T a, T b) const { return b < a ? b : a; }
};
struct OpOr {
  template <class T>
  __device__ __forceinline__ T operator()(T a, T b) const { return a | b; }
};

// Shuffles through one name so that record types can add specialisations.
template <class T>
__device__ __forceinline__ T shfl_up_t(T v, unsigned d) { return __shfl_up(v, d, kWave); }
template <class T>
__device__ __forceinline__ T shfl_t(T v, int lane) { return __shfl(v, lane, kWave); }

constexpr uint32_t kPageSize = 4096;  // core/constants.hpp:21 PAGESIZE
constexpr int kPageShift = 12;

__host__ __device__ inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

inline int counting_bits(uint64_t n) {  // core/graph.hpp:1087-1092
  int i = 0;
  for (; n > 0; n >>= 1) i++;
  return i;
}

// Device memory owned by the library.  hipMalloc'd buffers are 256-byte
// aligned; resize() only grows.  borrow() turns the buffer into a view of
// memory someone else keeps (the allocation-free entry point's static
// scratch): resize() then never allocates and fails when the view is too
// small, and release() frees nothing.
struct DevBuf {
  void* ptr = nullptr;
  uint64_t bytes = 0;
  bool owned = true;
  int resize(uint64_t want);
  void release();
  void borrow(void* p, uint64_t n) {
    release();
    ptr = p;
    bytes = n;
    owned = false;
  }
  ~DevBuf() { release(); }
  template <class T>
  T* as() const { return static_cast<T*>(ptr); }
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  void swap(DevBuf& o) {
    void* p = ptr;
    uint64_t b = bytes;
    bool w = owned;
    ptr = o.ptr;
    bytes = o.bytes;
    owned = o.owned;
    o.ptr = p;
    o.bytes = b;
    o.owned = w;
  }
};

// ------------------------------------------------------------ profiling --
// Optional HIP-event timing of individual kernel launches (eg_profile_*):
// prof_begin/prof_end bracket one launch on its own stream; bytes are the
// launch's algorithmic traffic.
enum ProfKernel {
  PK_RADIX_UPSWEEP = 0,
  PK_RADIX_DOWNSWEEP,
  PK_SCAN,
  PK_DEGREE,
  PK_CINDEX,
  PK_CSR_INDEX,
  PK_PR_PULL,
  PK_BFS,
  PK_COUNT
};
bool prof_on();
void prof_begin(hipStream_t s);
void prof_end(int kernel, uint64_t bytes, hipStream_t s);

// ------------------------------------------------------------- scans -----
// Device-wide exclusive prefix sums (reduce-then-scan, three launches).
// `tmp` must hold scan_tmp_elems(n) elements of the value type.
uint64_t scan_tmp_elems(uint64_t n);
int exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* tmp, hipStream_t s);
int exclusive_scan_u64(const uint64_t* in, uint64_t* out, uint64_t n, uint64_t* tmp, hipStream_t s);
int inclusive_max_scan_u64(const uint64_t* in, uint64_t* out, uint64_t n, uint64_t* tmp, hipStream_t s);

// ---------------------------------------------------------- file reader --
// Host file bytes -> device, range by range (io.hip): after the copies of
// range k are queued, stream s waits for them and on_range(k) queues the
// work that reads them, while the next range is being read.
int upload_file_ranges(int fd, const std::vector<uint64_t>& bounds, char* d, hipStream_t s,
                       const std::function<int(int)>& on_range);

}  // namespace eg
