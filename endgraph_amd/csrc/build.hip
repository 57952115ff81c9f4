// Graph construction: EndGraph's preprocessing (core/graph.hpp:424-1373)
// recast for one MI355X with the P partitions side by side in HBM.
//
//   symmetrise   per-rank reversed copies            (:1057-1078)
//   degree       out/in-degree histogram, id check   (:671-716, R5)
//   page weights + CPBT boundary search             (:586-626, :718-780)
//   CSC order    (q(src), dst, r, src, j); P=1: (q(src), dst, j)
//   CSR order    (q(dst), src, r, dst, j); P=1: (q(dst), src, j)
//                r = source rank, j = position in the rank's list,
//                q = p*S + s the sub-partition (partition, socket) of the
//                owning endpoint.  The reference reaches these orders by the
//                shuffle's stable sort per rank (:815), concatenation by
//                source rank (:957) and two stable sorts per partition
//                (:1107/:1113, :1262/:1268).  Here: for P>1 one stable sort
//                of every rank slice by the owner id, then one stable sort of
//                everything by the composite key (q, key).
//   indexes      compressed CSC index (:1147-1188), dense CSR index and
//                bitmap (:1312-1353).
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cstring>

#include "graph_state.hpp"

namespace eg {
namespace {

constexpr int kNT = 256;

template <class T>
__global__ void k_fill(T* __restrict__ p, uint64_t n, T v) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

template <class T>
int fill(T* p, uint64_t n, T v, hipStream_t s) {
  if (n == 0) return EG_OK;
  unsigned grid = unsigned(std::min<uint64_t>(ceil_div(n, kNT), 8192));
  hipLaunchKernelGGL((k_fill<T>), dim3(grid), dim3(kNT), 0, s, p, n, v);
  EG_HIP(hipGetLastError());
  return EG_OK;
}

// Per-rank symmetrisation (core/graph.hpp:1057-1078): rank r's list is its
// slice followed by the reversed copy of that slice.
template <class R>
__global__ void k_symmetrize(const R* __restrict__ in, R* __restrict__ out, uint64_t E0, int P, uint64_t o_begin,
                             uint64_t o_end) {
  const uint64_t base = E0 / P;
  for (uint64_t o = o_begin + uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; o < o_end;
       o += uint64_t(gridDim.x) * blockDim.x) {
    uint64_t r = base ? o / (2 * base) : uint64_t(P - 1);
    if (r > uint64_t(P - 1)) r = P - 1;
    const uint64_t n_r = base + (r == uint64_t(P - 1) ? E0 % P : 0);
    const uint64_t local = o - 2 * base * r;
    R e;
    if (local < n_r) {
      e = in[base * r + local];
    } else {
      e = in[base * r + local - n_r];
      uint32_t t = e.src;
      e.src = e.dst;
      e.dst = t;
    }
    out[o] = e;
  }
}

// Degree histogram (core/graph.hpp:689-707, R5) from an array sorted by
// the counted endpoint (within rank slices: a key that continues across a
// slice boundary only splits its count in two).  Each wave owns 512
// consecutive records and adds, per run inside them, the run's length: one
// atomic per run and chunk, at ascending addresses.
template <bool kSrc>
struct KeyEnd {
  template <class R>
  __device__ __forceinline__ uint64_t operator()(const R& r) const { return kSrc ? r.src : r.dst; }
  // a list kept as field arrays: the owner's array itself
  __device__ __forceinline__ uint64_t operator()(uint32_t v) const { return v; }
};

// page_amount (core/graph.hpp:718-741): one block per page.  The weight
// (1+(P-1)/16)*deg + eta*(P-1 + m*deg) keeps the reference's integer types:
// the first product and the bracket are 32-bit unsigned, the rest 64-bit.
// Also writes each page's out- and in-degree sums (edges owned per page).
__global__ __launch_bounds__(kNT) void k_page_weights(const uint32_t* __restrict__ deg,
                                                       const uint32_t* __restrict__ in_deg, uint32_t V, uint32_t f1,
                                                       uint32_t pm1, uint32_t m, uint64_t eta,
                                                       uint64_t* __restrict__ page_w, uint32_t pages,
                                                       uint64_t* __restrict__ page_out, uint64_t* __restrict__ page_in) {
  __shared__ uint64_t wt[3][kNT / kWave];
  const uint32_t pg = blockIdx.x;
  uint64_t acc = 0, so = 0, si = 0;
  for (uint32_t k = threadIdx.x; k < kPageSize; k += kNT) {
    const uint32_t v = pg * kPageSize + k;
    if (v < V) {
      const uint32_t d = deg[v];
      const uint32_t a = f1 * d;
      const uint32_t b = pm1 + m * d;
      acc += uint64_t(a) + eta * uint64_t(b);
      so += d;
      si += in_deg[v];
    }
  }
  acc = wave_reduce(acc, OpSum{});
  so = wave_reduce(so, OpSum{});
  si = wave_reduce(si, OpSum{});
  if ((threadIdx.x & (kWave - 1)) == 0) {
    wt[0][threadIdx.x / kWave] = acc;
    wt[1][threadIdx.x / kWave] = so;
    wt[2][threadIdx.x / kWave] = si;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0, to = 0, ti = 0;
    for (int i = 0; i < kNT / kWave; ++i) {
      t += wt[0][i];
      to += wt[1][i];
      ti += wt[2][i];
    }
    page_w[pg] = t;
    page_out[pg] = to;
    page_in[pg] = ti;
    if (pg == pages - 1) page_w[pages] = 0;  // slot that becomes the total after the scan
  }
}

// index of the last prefix entry <= x among prefix[lo..hi] (upper_bound - 1);
// prefix is non-decreasing and prefix[lo] <= x holds for every caller.
__device__ uint32_t last_le(const uint64_t* prefix, uint32_t lo, uint32_t hi, uint64_t x) {
  uint32_t a = lo, b = hi + 1;  // answer in [a, b)
  while (b - a > 1) {
    uint32_t mid = a + (b - a) / 2;
    if (prefix[mid] <= x)
      a = mid;
    else
      b = mid;
  }
  return a;
}

// judge (core/graph.hpp:586-594) on the pages [lo, lo+n] of the global prefix.
__device__ bool cpbt_judge(const uint64_t* prefix, uint32_t lo, uint32_t n, uint32_t k, uint64_t s) {
  uint64_t sum = prefix[lo];
  for (uint32_t i = 0; i < k; ++i) sum = prefix[last_le(prefix, lo, lo + n, sum + s)];
  return sum == prefix[lo + n];
}

// optimal_offset (core/graph.hpp:597-626) by one wave: the smallest feasible
// bottleneck in (lb, ub] found by a 64-way search (judge is monotone in s, so
// this is the bisection's answer), then the greedy boundaries.  Writes k+1
// absolute page ids to out.
__device__ void cpbt_wave(const uint64_t* prefix, uint32_t lo, uint32_t n, uint32_t k, uint32_t* out) {
  const int lane = threadIdx.x & (kWave - 1);
  uint64_t lb = 0, ub = prefix[lo + n] - prefix[lo];
  while (lb + 1 < ub) {
    const uint64_t range = ub - lb - 1;  // candidates lb+1 .. ub-1
    const uint64_t c = lb + 1 + (range / 64) * uint64_t(lane) + ((range % 64) * uint64_t(lane)) / 64;
    const bool ok = cpbt_judge(prefix, lo, n, k, c);
    const uint64_t mask = __ballot(ok);
    if (mask) {
      const int first = __ffsll((unsigned long long)mask) - 1;
      const uint64_t cf = __shfl(c, first, kWave);
      const uint64_t cp = __shfl(c, first > 0 ? first - 1 : 0, kWave);
      ub = cf;
      if (first > 0) lb = cp;
    } else {
      lb = __shfl(c, kWave - 1, kWave);
    }
  }
  if (lane == 0) {
    out[0] = lo;
    out[k] = lo + n;
    uint64_t sum = prefix[lo];
    for (uint32_t i = 0; i + 1 < k; ++i) {
      uint32_t at = last_le(prefix, lo, lo + n, sum + ub);
      out[i + 1] = at;
      sum = prefix[at];
    }
  }
}

// Machine level (k = P) then socket level (k = S) for every partition; one
// wave.  Page boundaries become vertex boundaries clamped to V (R9).
__global__ void k_cpbt(const uint64_t* __restrict__ prefix, uint32_t pages, uint32_t V, int P, int S,
                       eg_graph::Meta* __restrict__ meta) {
  cpbt_wave(prefix, 0, pages, uint32_t(P), meta->page_offset);
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i <= P; ++i)
      meta->partition_offset[i] = uint32_t(min(uint64_t(meta->page_offset[i]) * kPageSize, uint64_t(V)));
    meta->partition_offset[0] = 0;
    meta->partition_offset[P] = V;
  }
  __syncthreads();
  for (int p = 0; p < P; ++p) {
    const uint32_t pb = meta->page_offset[p], pe = meta->page_offset[p + 1];
    uint32_t* lpo = meta->local_page_offset + p * (S + 1);
    cpbt_wave(prefix, pb, pe - pb, uint32_t(S), lpo);
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t* lo = meta->local_partition_offset + p * (S + 1);
      const uint32_t vb = meta->partition_offset[p], ve = meta->partition_offset[p + 1];
      for (int i = 0; i <= S; ++i)
        lo[i] = uint32_t(min(uint64_t(lpo[i]) * kPageSize, uint64_t(ve)));
      lo[0] = vb;
      lo[S] = ve;
    }
    __syncthreads();
  }
}

// qmap[page] = p*S + s of the page (vertex ids in one page share it).
__global__ void k_qmap(const eg_graph::Meta* __restrict__ meta, uint32_t pages, int P, int S,
                       uint16_t* __restrict__ qmap) {
  const uint32_t pg = blockIdx.x * blockDim.x + threadIdx.x;
  if (pg >= pages) return;
  int p = 0;
  while (p + 1 < P && meta->page_offset[p + 1] <= pg) ++p;
  const uint32_t* lpo = meta->local_page_offset + p * (S + 1);
  int s = 0;
  while (s + 1 < S && lpo[s + 1] <= pg) ++s;
  qmap[pg] = uint16_t(p * S + s);
}

// Edges per sub-partition (CSC owner = src: out-degree; CSR owner = dst:
// in-degree) from the per-page sums; one block, LDS counters.
__global__ __launch_bounds__(kNT) void k_q_counts(const uint64_t* __restrict__ page_out,
                                                   const uint64_t* __restrict__ page_in, uint32_t pages,
                                                   const uint16_t* __restrict__ qmap, int Q, eg_graph::Meta* meta) {
  __shared__ unsigned long long co[256], ci[256];
  for (int q = threadIdx.x; q < 256; q += kNT) co[q] = ci[q] = 0;
  __syncthreads();
  for (uint32_t pg = threadIdx.x; pg < pages; pg += kNT) {
    const int q = qmap[pg];
    atomicAdd(&co[q], (unsigned long long)page_out[pg]);
    atomicAdd(&ci[q], (unsigned long long)page_in[pg]);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < Q; q += kNT) {
    meta->q_cnt_csc[q] = co[q];
    meta->q_cnt_csr[q] = ci[q];
  }
}

__global__ void k_q_offsets(eg_graph::Meta* meta, int Q) {
  if (threadIdx.x != 0) return;
  uint64_t a = 0, b = 0;
  for (int q = 0; q < Q; ++q) {
    meta->q_off_csc[q] = a;
    meta->q_off_csr[q] = b;
    a += meta->q_cnt_csc[q];
    b += meta->q_cnt_csr[q];
  }
  meta->q_off_csc[Q] = a;
  meta->q_off_csr[Q] = b;
}

// ----------------------------------------------------------- sort keys --
// First sort of a build: key = one endpoint; a record with an id >= V sets
// *err and sorts as V-1, so every access stays in bounds until the check.
template <bool kSrc>
struct KeyChecked {
  static constexpr int kField = kSrc ? 0 : 1;
  uint32_t V;
  uint32_t* err;  // null: clamp only (the ids were checked in an earlier launch)
  template <class R>
  __device__ __forceinline__ uint64_t operator()(const R& r) const {
    if (err && (r.src >= V || r.dst >= V)) atomicOr(err, 1u);
    const uint32_t k = kSrc ? r.src : r.dst;
    return k < V ? k : V - 1;
  }
  // the key of a field value alone (passes after the check only)
  __device__ __forceinline__ uint64_t of_field(uint32_t k) const { return k < V ? k : V - 1; }
};
// the range check runs once per record: in the first upsweep of the first
// rank sort only
template <bool kSrc>
KeyChecked<kSrc> rx_quiet(const KeyChecked<kSrc>& k) {
  return KeyChecked<kSrc>{k.V, nullptr};
}
// Stable pass on the sub-partition of the owner only.
template <bool kPull>
struct KeyQOnly {
  const uint16_t* qmap;
  template <class R>
  __device__ __forceinline__ uint64_t operator()(const R& r) const {
    return qmap[(kPull ? r.src : r.dst) >> kPageShift];
  }
};
struct KeySrc {
  static constexpr int kField = 0;
  template <class R>
  __device__ __forceinline__ uint64_t operator()(const R& r) const { return r.src; }
  __device__ __forceinline__ uint64_t of_field(uint32_t k) const { return k; }
};
struct KeyDst {
  static constexpr int kField = 1;
  template <class R>
  __device__ __forceinline__ uint64_t operator()(const R& r) const { return r.dst; }
  __device__ __forceinline__ uint64_t of_field(uint32_t k) const { return k; }
};
// (q(owner) << vbits) | other endpoint of a record of the CSC (kPull: owner
// = src, key = dst) or CSR list.  The list is sorted by sub-partition, so q
// is the sub-partition of the record's position, [q_off[q], q_off[q+1]):
// taken from the offsets (a search per chunk, scalar loads) instead of a
// qmap lookup of the owner, a dependent load per record.
template <bool kPull>
struct KeyQ {
  const uint64_t* q_off;  // Q + 1 sub-partition offsets of the list
  int Q;
  int vbits;
  // sub-partition of position i: the last q with q_off[q] <= i (empty
  // sub-partitions before a non-empty one share its offset)
  __device__ __forceinline__ uint32_t q_at(uint64_t i) const {
    int lo = 0, hi = Q;
    while (hi - lo > 1) {
      const int mid = (lo + hi) / 2;
      if (q_off[mid] <= i)
        lo = mid;
      else
        hi = mid;
    }
    return uint32_t(lo);
  }
  // ... from the sub-partition q of an earlier position on
  __device__ __forceinline__ uint32_t q_from(uint32_t q, uint64_t i) const {
    while (int(q) + 1 < Q && q_off[q + 1] <= i) ++q;
    return q;
  }
  template <class R>
  __device__ __forceinline__ uint64_t of(const R& r, uint32_t q) const {
    return (uint64_t(q) << vbits) | (kPull ? r.dst : r.src);
  }
};

// Key of record i of list a: the functor's own, or KeyQ's from the position.
template <class R, class KeyFn>
__device__ __forceinline__ uint64_t key_at(const R* a, uint64_t i, KeyFn key) {
  return key(a[i]);
}
template <class R, bool kPull>
__device__ __forceinline__ uint64_t key_at(const R* a, uint64_t i, KeyQ<kPull> key) {
  return key.of(a[i], key.q_at(i));
}

// ------------------------------------------------------------- indexes --
// Run-boundary kernels over a sorted list: wave g owns the 512 consecutive
// records [g*512, g*512+512) in 8 rounds of 64.  All loads are issued before
// use, neighbours come from wave shuffles (or one extra load at each chunk
// end), and the sub-partition and vertex are the composite key's two fields.
constexpr int kCR = 8;                 // rounds per wave
constexpr int kChunk = kCR * kWave;    // records per wave
constexpr uint64_t kNoKey = ~0ull;     // never a real key

struct ChunkKeys {
  uint64_t k[kCR];
  uint64_t prev, next;  // keys just before / after the chunk (kNoKey at the ends)
};

template <class R, class KeyFn>
__device__ __forceinline__ void chunk_keys(const R* __restrict__ a, uint64_t n, uint64_t c0, KeyFn key,
                                           ChunkKeys& c) {
  const int lane = threadIdx.x & (kWave - 1);
  R r[kCR];
#pragma unroll
  for (int j = 0; j < kCR; ++j) {
    const uint64_t i = c0 + uint64_t(j) * kWave + lane;
    r[j] = a[i < n ? i : n - 1];
  }
  const R rp = a[c0 > 0 ? c0 - 1 : 0];
  const uint64_t nx = c0 + kChunk;
  const R rn = a[nx < n ? nx : n - 1];
#pragma unroll
  for (int j = 0; j < kCR; ++j) c.k[j] = key(r[j]);
  c.prev = c0 > 0 ? key(rp) : kNoKey;
  c.next = nx < n ? key(rn) : kNoKey;
}

// The same for KeyQ: the sub-partition of the chunk's previous record by a
// search, then of each record from there on.
template <class R, bool kPull>
__device__ __forceinline__ void chunk_keys(const R* __restrict__ a, uint64_t n, uint64_t c0, KeyQ<kPull> key,
                                           ChunkKeys& c) {
  const int lane = threadIdx.x & (kWave - 1);
  R r[kCR];
#pragma unroll
  for (int j = 0; j < kCR; ++j) {
    const uint64_t i = c0 + uint64_t(j) * kWave + lane;
    r[j] = a[i < n ? i : n - 1];
  }
  const R rp = a[c0 > 0 ? c0 - 1 : 0];
  const uint64_t nx = c0 + kChunk;
  const R rn = a[nx < n ? nx : n - 1];
  const uint32_t q0 = key.q_at(c0 > 0 ? c0 - 1 : 0);
  // the next sub-partition's start; a chunk usually ends before it
  const uint64_t nb = int(q0) + 1 < key.Q ? key.q_off[q0 + 1] : ~0ull;
#pragma unroll
  for (int j = 0; j < kCR; ++j) {
    const uint64_t i = c0 + uint64_t(j) * kWave + lane;
    const uint64_t ic = i < n ? i : n - 1;
    c.k[j] = key.of(r[j], ic < nb ? q0 : key.q_from(q0, ic));
  }
  c.prev = c0 > 0 ? key.of(rp, q0) : kNoKey;
  c.next = nx < n ? key.of(rn, nx < nb ? q0 : key.q_from(q0, nx)) : kNoKey;
}

// Neighbour lanes through DPP wave shifts, with no LDS round trip: lane l
// gets lane l-1's value (wave_shr:1; lane 0 gets `fill`), or lane l+1's
// (wave_shl:1; lane 63 gets `fill`).  With bound_ctrl off, a lane whose
// source is outside the wave keeps the `old` operand, and so does a lane
// whose source is disabled in EXEC: call these with all 64 lanes active, at
// the top of a function, never under a condition (a short-circuit `&&`
// included) that can switch a neighbour off.
template <int kCtrl>
__device__ __forceinline__ uint64_t dpp_shift_u64(uint64_t v, uint64_t fill) {
  const int lo = __builtin_amdgcn_update_dpp(int(uint32_t(fill)), int(uint32_t(v)), kCtrl, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(int(uint32_t(fill >> 32)), int(uint32_t(v >> 32)), kCtrl, 0xf, 0xf, false);
  return (uint64_t(uint32_t(hi)) << 32) | uint32_t(lo);
}
__device__ __forceinline__ uint64_t lane_prev(uint64_t v, uint64_t fill) { return dpp_shift_u64<0x138>(v, fill); }
__device__ __forceinline__ uint64_t lane_next(uint64_t v, uint64_t fill) { return dpp_shift_u64<0x130>(v, fill); }

// head: first record of a run (key differs from the previous record)
__device__ __forceinline__ bool chunk_head(const ChunkKeys& c, int j, uint64_t i, uint64_t n) {
  const uint64_t last_prev = j > 0 ? shfl_t(c.k[j > 0 ? j - 1 : 0], kWave - 1) : c.prev;
  const uint64_t pk = lane_prev(c.k[j], last_prev);  // every lane, before the branch
  return i < n && pk != c.k[j];
}

// tail: last record of a run (key differs from the next record)
__device__ __forceinline__ bool chunk_tail(const ChunkKeys& c, int j, uint64_t i, uint64_t n) {
  const uint64_t first_next = j < kCR - 1 ? shfl_t(c.k[j < kCR - 1 ? j + 1 : j], 0) : c.next;
  const uint64_t nk = lane_next(c.k[j], first_next);  // every lane, before the branch
  return i < n && (i == n - 1 || nk != c.k[j]);
}

#define EG_FOR_CHUNKS(n)                                                                        \
  const uint64_t n_chunks_ = ceil_div((n), kChunk);                                             \
  const uint64_t gw_ = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;                \
  const uint64_t nw_ = uint64_t(gridDim.x) * blockDim.x / kWave;                                \
  for (uint64_t chunk = gw_; chunk < n_chunks_; chunk += nw_)

// Run heads per wave chunk (for the compressed index).
template <class R, class KeyFn>
__global__ __launch_bounds__(kNT) void k_run_count(const R* __restrict__ a, uint64_t n, KeyFn key,
                                                   uint32_t* __restrict__ cnt, uint64_t nchunks) {
  const int lane = threadIdx.x & (kWave - 1);
  EG_FOR_CHUNKS(n) {
    const uint64_t c0 = chunk * kChunk;
    ChunkKeys c;
    chunk_keys(a, n, c0, key, c);
    uint32_t h = 0;
#pragma unroll
    for (int j = 0; j < kCR; ++j) h += uint32_t(__popcll(__ballot(chunk_head(c, j, c0 + uint64_t(j) * kWave + lane, n))));
    if (lane == 0) cnt[chunk] = h;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) cnt[nchunks] = 0;
}

// Compressed CSC index entries: the r-th run head overall, lying in
// sub-partition q, lands at r + q (one sentinel slot per earlier q).
template <class R>
__global__ __launch_bounds__(kNT) void k_cindex_write(const R* __restrict__ a, uint64_t n, KeyQ<true> key,
                                                      const uint32_t* __restrict__ off,
                                                      const eg_graph::Meta* __restrict__ meta,
                                                      CIdx* __restrict__ out) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t lt = (1ull << lane) - 1;
  const uint64_t vmask = (1ull << key.vbits) - 1;
  EG_FOR_CHUNKS(n) {
    const uint64_t c0 = chunk * kChunk;
    ChunkKeys c;
    chunk_keys(a, n, c0, key, c);
    uint64_t run = off[chunk];
    // the chunk's first sub-partition start, loaded once (wave-uniform)
    const uint32_t qc = uint32_t(shfl_t(c.k[0], 0) >> key.vbits);
    const uint64_t qb0 = meta->q_off_csc[qc < uint32_t(key.Q) ? qc : 0];
#pragma unroll
    for (int j = 0; j < kCR; ++j) {
      const uint64_t i = c0 + uint64_t(j) * kWave + lane;
      const bool h = chunk_head(c, j, i, n);
      const uint64_t H = __ballot(h);
      if (h) {
        const uint32_t q = uint32_t(c.k[j] >> key.vbits);
        uint64_t qbase = qb0;
        if (q != qc) qbase = meta->q_off_csc[q];
        CIdx e;
        e.index = i - qbase;
        e.vertex = uint32_t(c.k[j] & vmask);
        out[run + __popcll(H & lt) + q] = e;
      }
      run += __popcll(H);
    }
  }
}

// Number of run heads in [0, x): chunk prefix + the partial chunk (one block).
template <class R, class KeyFn>
__device__ uint64_t runs_before(const R* a, uint64_t n, uint64_t x, KeyFn key, const uint32_t* off, uint32_t* wt) {
  const uint64_t ch = x / kChunk;
  const uint64_t b = ch * kChunk;
  uint32_t c = 0;
  for (uint64_t i = b + threadIdx.x; i < x; i += kNT) c += (i == 0 || key_at(a, i, key) != key_at(a, i - 1, key)) ? 1u : 0u;
  c = wave_reduce(c, OpSum{});
  if ((threadIdx.x & (kWave - 1)) == 0) wt[threadIdx.x / kWave] = c;
  __syncthreads();
  uint64_t tot = off[ch];
  for (int i = 0; i < kNT / kWave; ++i) tot += wt[i];
  __syncthreads();
  return tot;
}

// One block per sub-partition: where its index starts, its entry count and
// the sentinel {n_edges, V} (core/graph.hpp:1186-1187; R8 for empty ones).
template <class R>
__global__ __launch_bounds__(kNT) void k_cindex_sentinel(const R* __restrict__ a, uint64_t n, KeyQ<true> key,
                                                         const uint32_t* __restrict__ off, uint32_t V,
                                                         eg_graph::Meta* __restrict__ meta, CIdx* __restrict__ out) {
  __shared__ uint32_t wt[kNT / kWave];
  const int q = blockIdx.x;
  const uint64_t b = meta->q_off_csc[q], e = meta->q_off_csc[q + 1];
  const uint64_t rb = runs_before(a, n, b, key, off, wt);
  const uint64_t re = runs_before(a, n, e, key, off, wt);
  if (threadIdx.x == 0) {
    CIdx c;
    c.index = e - b;
    c.vertex = V;
    out[re + q] = c;
    meta->cidx_off[q] = rb + q;
    meta->cidx_cnt[q] = re - rb + 1;
  }
}

// Dense CSR index (core/graph.hpp:1312-1353) and its bitmap, every entry
// defined: index[q][u] = local position of the first record of sub-partition
// q whose source is >= u (u in [0, V]), so index[q][u]..index[q][u+1] is u's
// run and empty u get an empty range.  The head of the run of source u, whose
// previous record in q has source u' (none: u' = -1), writes entries
// u'+1..u; the last record of q writes u+1..V.  Gaps longer than kShortGap
// are cut into pieces of at most kGapPiece entries that k_csr_long_gaps
// writes, one wave per piece.  Bit u of the bitmap is set by u's run head;
// the lanes of a round holding the same bitmap word combine their bits
// first (one atomic per word and round).
constexpr uint32_t kShortGap = 64;
constexpr uint32_t kGapPiece = 1u << 16;

struct GapPiece {
  uint64_t start;  // first index entry
  uint32_t count;
  uint32_t value;
};

__device__ __forceinline__ void put_gap(uint64_t* __restrict__ index, uint64_t start, uint64_t cnt, uint32_t value,
                                        GapPiece* __restrict__ gaps, uint32_t* __restrict__ ngaps, uint32_t max_gaps) {
  if (cnt <= kShortGap) {
    for (uint64_t c = 0; c < cnt; ++c) index[start + c] = value;
    return;
  }
  const uint32_t pieces = uint32_t(ceil_div(cnt, uint64_t(kGapPiece)));
  const uint32_t slot = atomicAdd(ngaps, pieces);
  if (uint64_t(slot) + pieces <= max_gaps) {
    for (uint32_t k = 0; k < pieces; ++k) {
      const uint64_t off = uint64_t(k) * kGapPiece;
      gaps[slot + k] = GapPiece{start + off, uint32_t(cnt - off < kGapPiece ? cnt - off : kGapPiece), value};
    }
    return;
  }
  // list full: empty the slots taken and write the gap here
  for (uint64_t k = slot; k < max_gaps && k < uint64_t(slot) + pieces; ++k) gaps[k] = GapPiece{0, 0, 0};
  for (uint64_t c = 0; c < cnt; ++c) index[start + c] = value;
}

template <class R>
__global__ __launch_bounds__(kNT) void k_csr_gaps(const R* __restrict__ a, uint64_t n, KeyQ<false> key, uint32_t V,
                                                  const eg_graph::Meta* __restrict__ meta, uint64_t* __restrict__ index,
                                                  unsigned long long* __restrict__ bitmap,
                                                  GapPiece* __restrict__ gaps, uint32_t* __restrict__ ngaps,
                                                  uint32_t max_gaps) {
  // per wave: bits of the 64 bitmap words from the chunk's first word on
  __shared__ unsigned long long bm[kNT / kWave][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  unsigned long long* my = bm[threadIdx.x / kWave];
  const uint64_t vmask = (1ull << key.vbits) - 1;
  const uint64_t row = uint64_t(V) + 1;
  const uint64_t words = (V >> 6) + 1;
  __hip_atomic_store(my + lane, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
  __asm__ volatile("" ::: "memory");
  EG_FOR_CHUNKS(n) {
    const uint64_t c0 = chunk * kChunk;
    ChunkKeys c;
    chunk_keys(a, n, c0, key, c);
    const uint64_t k0 = shfl_t(c.k[0], 0);
    const uint64_t wbase = (k0 >> key.vbits) * words + ((k0 & vmask) >> 6);
#pragma unroll
    for (int j = 0; j < kCR; ++j) {
      const uint64_t i = c0 + uint64_t(j) * kWave + lane;
      const uint64_t k = c.k[j];
      const uint32_t q = uint32_t(k >> key.vbits);
      const uint32_t u = uint32_t(k & vmask);
      const uint64_t last_prev = j > 0 ? shfl_t(c.k[j > 0 ? j - 1 : 0], kWave - 1) : c.prev;
      const uint64_t pk = lane_prev(k, last_prev);
      const uint64_t first_next = j < kCR - 1 ? shfl_t(c.k[j < kCR - 1 ? j + 1 : j], 0) : c.next;
      const uint64_t nk = lane_next(k, first_next);
      const bool valid = i < n;
      if (valid && pk != k) {  // run head
        const uint64_t qbase = meta->q_off_csr[q];
        const uint64_t lo = (pk != kNoKey && (pk >> key.vbits) == q) ? (pk & vmask) + 1 : 0;
        put_gap(index, q * row + lo, u - lo + 1, uint32_t(i - qbase), gaps, ngaps, max_gaps);
        const uint64_t wd = q * words + (u >> 6);
        const unsigned long long bit = 1ull << (u & 63);
        if (wd - wbase < kWave)
          __hip_atomic_fetch_or(my + (wd - wbase), bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        else
          atomicOr(&bitmap[wd], bit);
      }
      if (valid && (i == n - 1 || (nk >> key.vbits) != q))  // last record of q
        put_gap(index, q * row + u + 1, V - u, uint32_t(i + 1 - meta->q_off_csr[q]), gaps, ngaps, max_gaps);
    }
    // Flush the chunk's words.  Every OR of the wave has completed
    // (lgkmcnt(0)) before the read, and the read before the clear; the empty
    // asm statements keep the compiler from moving LDS accesses across.
    __builtin_amdgcn_s_waitcnt(kLgkmZero);
    __asm__ volatile("" ::: "memory");
    const unsigned long long b = __hip_atomic_load(my + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    __builtin_amdgcn_s_waitcnt(kLgkmZero);
    __asm__ volatile("" ::: "memory");
    if (b) {
      atomicOr(&bitmap[wbase + lane], b);
      __hip_atomic_store(my + lane, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    }
    __asm__ volatile("" ::: "memory");
  }
}

// Pieces of long gaps: one wave per piece.
__global__ __launch_bounds__(kNT) void k_csr_long_gaps(uint64_t* __restrict__ index, const GapPiece* __restrict__ gaps,
                                                       const uint32_t* __restrict__ ngaps, uint32_t max_gaps) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint32_t n = *ngaps < max_gaps ? *ngaps : max_gaps;
  const uint64_t gw = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  const uint64_t nw = uint64_t(gridDim.x) * blockDim.x / kWave;
  for (uint64_t p = gw; p < n; p += nw) {
    const GapPiece g = gaps[p];
    for (uint32_t e = lane; e < g.count; e += kWave) index[g.start + e] = g.value;
  }
}

// ---- one sub-partition (P = S = 1): both indexes from the degrees ----
// The CSC holds the records by destination and the CSR by source, each in
// one run per vertex, so a vertex's run starts at the exclusive prefix of
// its degree: the compressed CSC index is {prefix of in-degree, v} for the
// v with in-edges, then {E, V}; the dense CSR index is the prefix of the
// out-degree over V + 1 entries (the run-scanning kernels give the same
// arrays for Q > 1 one sub-partition at a time).  deg as u64 (entry V = 0)
// and flags deg > 0 as u32 (entry V = 0), for the scans.
__global__ void k_deg_u64(const uint32_t* __restrict__ deg, uint32_t V, uint64_t* __restrict__ out) {
  for (uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; v <= V; v += uint64_t(gridDim.x) * blockDim.x)
    out[v] = v < V ? deg[v] : 0ull;
}
__global__ void k_deg_flags(const uint32_t* __restrict__ deg, uint32_t V, uint32_t* __restrict__ out) {
  for (uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; v <= V; v += uint64_t(gridDim.x) * blockDim.x)
    out[v] = v < V && deg[v] > 0 ? 1u : 0u;
}
// pos, slot: the exclusive scans of k_deg_u64 and k_deg_flags over in_deg
__global__ void k_cindex_from_deg(const uint32_t* __restrict__ in_deg, const uint64_t* __restrict__ pos,
                                  const uint32_t* __restrict__ slot, uint32_t V, uint64_t E,
                                  eg_graph::Meta* __restrict__ meta, CIdx* __restrict__ out) {
  for (uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += uint64_t(gridDim.x) * blockDim.x)
    if (in_deg[v]) {
      CIdx c;
      c.index = pos[v];
      c.vertex = uint32_t(v);
      out[slot[v]] = c;
    }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    CIdx c;
    c.index = E;
    c.vertex = V;
    out[slot[V]] = c;
    meta->cidx_off[0] = 0;
    meta->cidx_cnt[0] = uint64_t(slot[V]) + 1;
  }
}
// bit u of the CSR bitmap: u has out-edges; one thread per word
__global__ __launch_bounds__(kNT) void k_bitmap_from_deg(const uint32_t* __restrict__ deg, uint32_t V,
                                                          uint64_t words, uint64_t* __restrict__ bitmap) {
  // one wave per word: lane i reads the degree of vertex 64w + i (coalesced)
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t nw = uint64_t(gridDim.x) * blockDim.x / kWave;
  for (uint64_t w = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave; w < words; w += nw) {
    const uint64_t u = w * 64 + lane;
    const uint64_t b = __ballot(u < V && deg[u] > 0);
    if (lane == 0) bitmap[w] = b;
  }
}

// Rows of empty sub-partitions (no records, so no run writes them): all 0.
__global__ __launch_bounds__(kNT) void k_csr_empty_rows(uint64_t* __restrict__ index, uint32_t V,
                                                        const eg_graph::Meta* __restrict__ meta) {
  const uint32_t q = blockIdx.y;
  if (meta->q_off_csr[q] != meta->q_off_csr[q + 1]) return;
  const uint64_t row = uint64_t(V) + 1;
  for (uint64_t u = uint64_t(blockIdx.x) * kNT + threadIdx.x; u < row; u += uint64_t(gridDim.x) * kNT)
    index[q * row + u] = 0;
}

template <class R, bool kSrc>
__global__ __launch_bounds__(kNT) void k_run_degree(const R* __restrict__ a, uint64_t n, uint32_t* __restrict__ deg) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t le = lane == 63 ? ~0ull : ((2ull << lane) - 1);
  EG_FOR_CHUNKS(n) {
    const uint64_t c0 = chunk * kChunk;
    ChunkKeys c;
    chunk_keys(a, n, c0, KeyEnd<kSrc>{}, c);
    uint64_t H[kCR];
#pragma unroll
    for (int j = 0; j < kCR; ++j) H[j] = __ballot(chunk_head(c, j, c0 + uint64_t(j) * kWave + lane, n));
#pragma unroll
    for (int j = 0; j < kCR; ++j) {
      const uint64_t i = c0 + uint64_t(j) * kWave + lane;
      const bool tail = chunk_tail(c, j, i, n) || (lane == 63 && j == kCR - 1 && i < n);
      if (!tail) continue;
      // latest head at or before (j, lane) inside the chunk; chunk start if none
      int sj = -1, sl = 0;
      if (H[j] & le) {
        sj = j;
        sl = 63 - __clzll((long long)(H[j] & le));
      } else {
        for (int jj = j - 1; jj >= 0; --jj)
          if (H[jj]) {
            sj = jj;
            sl = 63 - __clzll((long long)H[jj]);
            break;
          }
      }
      const uint32_t start = sj < 0 ? 0u : uint32_t(sj * kWave + sl);
      atomicAdd(&deg[uint32_t(c.k[j])], uint32_t(j * kWave + lane) - start + 1u);
    }
  }
}


// --- P > 1: regroup the rank slices by sub-partition ---------------------
// A rank slice sorted by its owner id is already grouped by sub-partition
// (q is monotone in the vertex id).  Block (r, q) = the records of rank r
// owned by q.  Copying the blocks in (q, r) order and then sorting each q
// stably by the other endpoint gives the order (q, key, r, owner, j): the
// composite (q, key) sort with the q digit done by a copy.
template <class R, bool kSrc>
__global__ void k_block_bounds(const R* __restrict__ a, const uint64_t* __restrict__ rank_start, int P, int Q, int S,
                               const eg_graph::Meta* __restrict__ meta, uint64_t* __restrict__ bpos) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P * (Q + 1)) return;
  const int r = t / (Q + 1), q = t % (Q + 1);
  const uint64_t lo0 = rank_start[r], hi0 = rank_start[r + 1];
  if (q == 0 || q == Q) {
    bpos[t] = q == 0 ? lo0 : hi0;
    return;
  }
  const int p = q / S, sk = q % S;
  const uint32_t vb = meta->local_partition_offset[p * (S + 1) + sk];  // first vertex of q
  uint64_t lo = lo0, hi = hi0;  // first record whose owner >= vb
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (KeyEnd<kSrc>{}(a[mid]) < vb)
      lo = mid + 1;
    else
      hi = mid;
  }
  bpos[t] = lo;
}

// (destination start, source start) of every block in (q, r) order, plus the
// end sentinel; one thread (P*Q is small).
__global__ void k_block_table(const uint64_t* __restrict__ bpos, int P, int Q, uint64_t* __restrict__ tab_dst,
                              uint64_t* __restrict__ tab_src) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  uint64_t d = 0;
  int b = 0;
  for (int q = 0; q < Q; ++q)
    for (int r = 0; r < P; ++r) {
      const uint64_t s0 = bpos[r * (Q + 1) + q], s1 = bpos[r * (Q + 1) + q + 1];
      tab_dst[b] = d;
      tab_src[b] = s0;
      d += s1 - s0;
      ++b;
    }
  tab_dst[b] = d;
}


unsigned grid_for(uint64_t n, uint64_t per_block = kNT, uint64_t cap = 65536) {
  return unsigned(std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(n, per_block), cap)));
}

// First record of rank slice r in the logical list (mult = 2 when each slice
// is followed by its reversed copy); r = P gives the list's end.
inline uint64_t slice_start(uint64_t E0, int P, uint64_t mult, int r) {
  return r >= P ? mult * E0 : mult * (E0 / P) * uint64_t(r);
}

// Stable sorts of the rank slices r0..r1-1 (one segment each) by source
// into a and by destination into b; the two sorts share their first read of
// L (the first upsweep counts both first digits).  A, B and the ping-pong
// are field arrays: with P > 1 A and B are intermediate lists, read by key
// only in the degree and upsweep passes.
template <class R, class KeyA, class KeyB>
int rank_sorts_fields(eg_graph* g, const R* L, RxSoa<R> a, RxSoa<R> b, RxSoa<R> tmp, uint64_t E0, int P,
                      uint64_t mult, int r0, int r1, KeyA key_a, KeyB key_b, int bits, hipStream_t s) {
  const uint64_t b0 = slice_start(E0, P, mult, r0);
  std::vector<uint64_t> seg(r1 - r0 + 1);
  for (int r = r0; r <= r1; ++r) seg[r - r0] = slice_start(E0, P, mult, r) - b0;
  return rx_sort_pair_fields(L + b0, a.at(b0), b.at(b0), tmp.at(b0), seg, key_a, key_b, bits, g->sortw, s);
}

template <class R, class KeyFn>
int whole_sort(eg_graph* g, const R* in, R* out, R* tmp, uint64_t n, KeyFn key, int bits, hipStream_t s) {
  return rx_sort(in, out, tmp, std::vector<uint64_t>{0, n}, key, bits, g->sortw, s);
}

unsigned chunk_grid(uint64_t n) { return grid_for(ceil_div(n, kChunk), kNT / kWave, 65536); }

template <class R, bool kSrc>
int run_degree(const R* a, uint64_t n, uint32_t* deg, hipStream_t s) {
  if (!n) return EG_OK;
  prof_begin(s);
  hipLaunchKernelGGL((k_run_degree<R, kSrc>), dim3(chunk_grid(n)), dim3(kNT), 0, s, a, n, deg);
  prof_end(PK_DEGREE, n * sizeof(R), s);
  EG_HIP(hipGetLastError());
  return EG_OK;
}

// Build order (every sort is stable; see the file comment for the target
// orders):
//   A = rank slices sorted by src   (P = 1: the whole list)   -> out-degree
//   B = rank slices sorted by dst   (P = 1: the whole list)   -> in-degree
//   page weights, CPBT, sub-partition map
//   CSC = A sorted by (q(src), dst)  (P = 1: B sorted by q(src))
//   CSR = B sorted by (q(dst), src)  (P = 1: A sorted by q(dst))
// At P = 1 and S = 1 the q pass is the identity and A, B are the results.
// The first sorts are per rank slice, so a caller that receives the list
// slice by slice (the file reader) can run them for each slice as it lands:
// build_setup, build_slices for every slice, build_finish.
template <class R>
int build_setup(eg_graph* g, uint64_t E0, int sym, hipStream_t s) {
  const uint32_t V = g->V;
  const int P = g->P, Q = g->Q;
  const uint64_t E = sym ? 2 * E0 : E0;
  const uint64_t rb = sizeof(R);
  g->built = false;
  g->eng_ready = false;
  g->pv_ready = false;
  g->tv_ready = false;
  g->ad_ready = false;
  g->E0 = E0;
  g->E = E;
  g->symmetrized = sym;
  g->vbits = std::max(1, counting_bits(V - 1));
  g->qbits = counting_bits(uint64_t(Q - 1));
  const uint64_t pages = g->pages;
  const uint64_t Emax = std::max<uint64_t>(E, 1);

  // --- buffers (grow-only; reused by repeated builds) ---
  EG_TRY(g->out_deg.resize(uint64_t(V) * 4));
  EG_TRY(g->in_deg.resize(uint64_t(V) * 4));
  EG_TRY(g->page_prefix.resize((pages + 1) * 8 + 2 * pages * 8));
  EG_TRY(g->meta.resize(sizeof(eg_graph::Meta)));
  EG_TRY(g->qmap.resize(pages * 2));
  // P = 1: the two lists' buffers are also the first sorts' field-array
  // ping-pong (they are free until the sorts within sub-partitions)
  const uint64_t lb = P == 1 ? std::max(Emax * rb, rx_soa_bytes<R>(Emax)) : Emax * rb;
  EG_TRY(g->csc_list.resize(lb));
  EG_TRY(g->csr_list.resize(lb));
  const uint64_t cidx_cap = std::min<uint64_t>(E, uint64_t(Q) * V) + Q;
  EG_TRY(g->csc_cidx.resize(cidx_cap * sizeof(CIdx)));
  const uint64_t idx_n = uint64_t(Q) * (V + 1);
  const uint64_t words = (V >> 6) + 1;
  EG_TRY(g->csr_index.resize(idx_n * 8));
  EG_TRY(g->csr_bitmap.resize(uint64_t(Q) * words * 8));
  // P > 1: the rank-sorted lists and the sorts' ping-pong are field arrays
  const uint64_t tb = P > 1 ? rx_soa_bytes<R>(Emax) : Emax * rb;
  EG_TRY(g->t1.resize(tb));
  EG_TRY(g->t2.resize(tb));
  if (P > 1 || Q > 1) EG_TRY(g->t3.resize(tb));
  const uint64_t hist_n = 1;
  const uint64_t scan_n = std::max<uint64_t>({scan_tmp_elems(hist_n), scan_tmp_elems(pages + 1),
                                              scan_tmp_elems(ceil_div(E, kChunk) + 1),
                                              Q == 1 ? scan_tmp_elems(uint64_t(V) + 1) : 0});
  EG_TRY(g->scan_tmp.resize(scan_n * 8));
  EG_TRY(g->tile_cnt.resize((ceil_div(E, kChunk) + 2) * 4));
  EG_TRY(g->err.resize(4));
  if (sym) EG_TRY(g->sym.resize(Emax * rb));
  // the first sorts check every id once (only mid-build read-back: build_finish)
  EG_TRY(fill<uint32_t>(g->err.as<uint32_t>(), 1, 0u, s));
  EG_TRY(g->sortw.resize_segs(1));
  EG_TRY(fill<uint32_t>(g->sortw.err.as<uint32_t>(), 1, 0u, s));
  return EG_OK;
}

// Rank slices r0..r1-1 of the input: their reversed copies (symmetrised
// builds) and their two first sorts.
template <class R>
int build_slices(eg_graph* g, const R* d_in, uint64_t E0, int sym, int r0, int r1, hipStream_t s) {
  const int P = g->P;
  const uint64_t mult = sym ? 2 : 1;
  const R* L = d_in;
  if (sym) {
    const uint64_t o0 = slice_start(E0, P, 2, r0), o1 = slice_start(E0, P, 2, r1);
    if (o1 > o0)
      hipLaunchKernelGGL((k_symmetrize<R>), dim3(grid_for(o1 - o0, kNT * 4, 65536)), dim3(kNT), 0, s, d_in,
                         g->sym.as<R>(), E0, P, o0, o1);
    EG_HIP(hipGetLastError());
    L = g->sym.as<R>();
  }
  // the ids are range-checked once, by the source sort's first upsweep
  if (P > 1) {
    const uint64_t Emax = std::max<uint64_t>(g->E, 1);
    EG_TRY(rank_sorts_fields(g, L, rx_soa_in<R>(g->t1.ptr, Emax), rx_soa_in<R>(g->t2.ptr, Emax),
                             rx_soa_in<R>(g->t3.ptr, Emax), E0, P, mult, r0, r1,
                             KeyChecked<true>{g->V, g->err.as<uint32_t>()}, KeyChecked<false>{g->V, nullptr},
                             g->vbits, s));
    return EG_OK;
  }
  // P = 1: the whole list, records in and out, intermediate passes as field
  // arrays in the CSC and CSR buffers
  const uint64_t Emax = std::max<uint64_t>(g->E, 1);
  const uint64_t b0 = slice_start(E0, P, mult, r0);
  std::vector<uint64_t> seg(r1 - r0 + 1);
  for (int r = r0; r <= r1; ++r) seg[r - r0] = slice_start(E0, P, mult, r) - b0;
  EG_TRY(rx_sort_pair_through(L + b0, g->t1.as<R>() + b0, g->t2.as<R>() + b0,
                              rx_soa_in<R>(g->csc_list.ptr, Emax).at(b0), rx_soa_in<R>(g->csr_list.ptr, Emax).at(b0),
                              seg, KeyChecked<true>{g->V, g->err.as<uint32_t>()}, KeyChecked<false>{g->V, nullptr},
                              g->vbits, g->sortw, s));
  return EG_OK;
}

template <class R>
int build_finish(eg_graph* g, hipStream_t s) {
  const uint32_t V = g->V;
  const int P = g->P, S = g->S, Q = g->Q;
  const uint64_t E0 = g->E0, E = g->E;
  const int sym = g->symmetrized;
  const uint64_t rb = sizeof(R);
  const uint64_t pages = g->pages;
  const uint64_t idx_n = uint64_t(Q) * (V + 1);
  const uint64_t words = (V >> 6) + 1;
  eg_graph::Meta* meta = g->dmeta();
  uint32_t* err = g->err.as<uint32_t>();
  uint32_t* out_deg = g->out_deg.as<uint32_t>();
  uint32_t* in_deg = g->in_deg.as<uint32_t>();
  uint16_t* qmap = g->qmap.as<uint16_t>();
  uint32_t* stmp32 = g->scan_tmp.as<uint32_t>();
  const uint64_t mult = sym ? 2 : 1;
  R* A = g->t1.as<R>();
  R* B = g->t2.as<R>();
  uint32_t h_err = 0;
  EG_HIP(hipMemcpyAsync(&h_err, err, 4, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  if (h_err) return fail(EG_ERR_RANGE, "edge list holds a vertex id >= vertices_num (%u)", V);

  // --- degrees from runs ---
  EG_TRY(fill<uint32_t>(out_deg, V, 0u, s));
  EG_TRY(fill<uint32_t>(in_deg, V, 0u, s));
  const uint64_t Emax = std::max<uint64_t>(E, 1);
  const RxSoa<R> fa = rx_soa_in<R>(g->t1.ptr, Emax), fb = rx_soa_in<R>(g->t2.ptr, Emax);
  if (P > 1) {  // the owner's field array of each rank-sorted list
    EG_TRY((run_degree<uint32_t, true>(fa.s, E, out_deg, s)));
    EG_TRY((run_degree<uint32_t, false>(fb.d, E, in_deg, s)));
  } else {
    EG_TRY((run_degree<R, true>(A, E, out_deg, s)));
    EG_TRY((run_degree<R, false>(B, E, in_deg, s)));
  }

  // --- CPBT ---
  uint64_t* prefix = g->page_prefix.as<uint64_t>();
  uint64_t* page_out = prefix + pages + 1;
  uint64_t* page_in = page_out + pages;
  hipLaunchKernelGGL(k_page_weights, dim3(unsigned(pages)), dim3(kNT), 0, s, (const uint32_t*)out_deg,
                     (const uint32_t*)in_deg, V, uint32_t(1 + (P - 1) / 16), uint32_t(P - 1), g->m, g->eta, prefix,
                     uint32_t(pages), page_out, page_in);
  EG_HIP(hipGetLastError());
  EG_TRY(exclusive_scan_u64(prefix, prefix, pages + 1, g->scan_tmp.as<uint64_t>(), s));
  EG_HIP(hipMemsetAsync(meta, 0, sizeof(eg_graph::Meta), s));
  hipLaunchKernelGGL(k_cpbt, dim3(1), dim3(kWave), 0, s, prefix, uint32_t(pages), V, P, S, meta);
  EG_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_qmap, dim3(grid_for(pages)), dim3(kNT), 0, s, meta, uint32_t(pages), P, S, qmap);
  EG_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_q_counts, dim3(1), dim3(kNT), 0, s, (const uint64_t*)page_out, (const uint64_t*)page_in,
                     uint32_t(pages), (const uint16_t*)qmap, Q, meta);
  EG_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_q_offsets, dim3(1), dim3(kWave), 0, s, meta, Q);
  EG_HIP(hipGetLastError());

  // --- second sorts ---
  if (P > 1) {
    // sub-partition sizes to the host: they delimit the segmented sorts
    std::vector<uint64_t> qoff(2 * (Q + 1));
    EG_HIP(hipMemcpyAsync(qoff.data(), meta->q_off_csc, (Q + 1) * 8, hipMemcpyDeviceToHost, s));
    EG_HIP(hipMemcpyAsync(qoff.data() + Q + 1, meta->q_off_csr, (Q + 1) * 8, hipMemcpyDeviceToHost, s));
    std::vector<uint64_t> rank_start(P + 1);
    for (int r = 0; r < P; ++r) rank_start[r] = mult * (E0 / P) * r;
    rank_start[P] = E;
    EG_TRY(g->blk.resize((uint64_t(P) + 1) * 8 + 2 * uint64_t(P) * (Q + 1) * 8 + 2 * (uint64_t(P) * Q + 1) * 8 + 64));
    uint64_t* d_rank = g->blk.as<uint64_t>();
    uint64_t* bpos_dir = d_rank + P + 1;  // block bounds of each shuffle: by source, then by destination
    uint64_t* tab_dst = bpos_dir + 2 * uint64_t(P) * (Q + 1);
    uint64_t* tab_src = tab_dst + uint64_t(P) * Q + 1;
    EG_HIP(hipMemcpyAsync(d_rank, rank_start.data(), (P + 1) * 8, hipMemcpyHostToDevice, s));
    EG_HIP(hipStreamSynchronize(s));
    const std::vector<uint64_t> seg_csc(qoff.begin(), qoff.begin() + Q + 1);
    const std::vector<uint64_t> seg_csr(qoff.begin() + Q + 1, qoff.end());
    const RxSoa<R> X = rx_soa_in<R>(g->t3.ptr, Emax);
    for (int dir = 0; dir < 2; ++dir) {
      const bool pull = dir == 0;
      const RxSoa<R> in = pull ? fa : fb;
      uint64_t* bpos = bpos_dir + uint64_t(dir) * P * (Q + 1);
      const unsigned bb = unsigned(ceil_div(uint64_t(P) * (Q + 1), kNT));
      if (pull)
        hipLaunchKernelGGL((k_block_bounds<uint32_t, true>), dim3(bb), dim3(kNT), 0, s, (const uint32_t*)in.s,
                           (const uint64_t*)d_rank, P, Q, S, (const eg_graph::Meta*)meta, bpos);
      else
        hipLaunchKernelGGL((k_block_bounds<uint32_t, false>), dim3(bb), dim3(kNT), 0, s, (const uint32_t*)in.d,
                           (const uint64_t*)d_rank, P, Q, S, (const eg_graph::Meta*)meta, bpos);
      EG_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_block_table, dim3(1), dim3(kWave), 0, s, (const uint64_t*)bpos, P, Q, tab_dst, tab_src);
      EG_HIP(hipGetLastError());
      // the first pass of the segmented sort reads the (q, r) blocks in
      // place; its later passes ping-pong between X and the rank-sorted
      // list it has consumed
      if (pull)
        EG_TRY(rx_sort_blocks_fields(in, tab_dst, tab_src, P * Q, g->csc_list.as<R>(), X, in, seg_csc, KeyDst{},
                                     g->vbits, g->sortw, s));
      else
        EG_TRY(rx_sort_blocks_fields(in, tab_dst, tab_src, P * Q, g->csr_list.as<R>(), X, in, seg_csr, KeySrc{},
                                     g->vbits, g->sortw, s));
    }
  } else if (Q == 1) {
    g->csr_list.swap(g->t1);  // A is the CSR, B the CSC
    g->csc_list.swap(g->t2);
  } else {
    const int qb = std::max(1, g->qbits);
    EG_TRY(whole_sort(g, (const R*)B, g->csc_list.as<R>(), g->t3.as<R>(), E, KeyQOnly<true>{qmap}, qb, s));
    EG_TRY(whole_sort(g, (const R*)A, g->csr_list.as<R>(), g->t3.as<R>(), E, KeyQOnly<false>{qmap}, qb, s));
  }

  if (Q == 1) {
    // --- both indexes from the degrees (one sub-partition) ---
    CIdx* cidx = g->csc_cidx.as<CIdx>();
    EG_TRY(g->idx_tmp.resize((uint64_t(V) + 1) * 12 + 64));
    uint64_t* pos = g->idx_tmp.as<uint64_t>();
    uint32_t* slot = reinterpret_cast<uint32_t*>(pos + uint64_t(V) + 1);
    const unsigned gv = grid_for(uint64_t(V) + 1);
    hipLaunchKernelGGL(k_deg_u64, dim3(gv), dim3(kNT), 0, s, (const uint32_t*)in_deg, V, pos);
    EG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_deg_flags, dim3(gv), dim3(kNT), 0, s, (const uint32_t*)in_deg, V, slot);
    EG_HIP(hipGetLastError());
    EG_TRY(exclusive_scan_u64(pos, pos, uint64_t(V) + 1, g->scan_tmp.as<uint64_t>(), s));
    EG_TRY(exclusive_scan_u32(slot, slot, uint64_t(V) + 1, stmp32, s));
    prof_begin(s);
    hipLaunchKernelGGL(k_cindex_from_deg, dim3(grid_for(V)), dim3(kNT), 0, s, (const uint32_t*)in_deg,
                       (const uint64_t*)pos, (const uint32_t*)slot, V, E, meta, cidx);
    prof_end(PK_CINDEX, uint64_t(V) * 16, s);
    EG_HIP(hipGetLastError());
    uint64_t* index = g->csr_index.as<uint64_t>();
    hipLaunchKernelGGL(k_deg_u64, dim3(gv), dim3(kNT), 0, s, (const uint32_t*)out_deg, V, index);
    EG_HIP(hipGetLastError());
    prof_begin(s);
    EG_TRY(exclusive_scan_u64(index, index, uint64_t(V) + 1, g->scan_tmp.as<uint64_t>(), s));
    prof_end(PK_CSR_INDEX, (uint64_t(V) + 1) * 16, s);
    hipLaunchKernelGGL(k_bitmap_from_deg, dim3(grid_for(words, kNT / kWave, 65536)), dim3(kNT), 0, s,
                       (const uint32_t*)out_deg, V, words, g->csr_bitmap.as<uint64_t>());
    EG_HIP(hipGetLastError());
  } else {
    // --- compressed CSC index ---
    {
      const R* a = g->csc_list.as<R>();
      const KeyQ<true> key{meta->q_off_csc, Q, g->vbits};
      const uint64_t nch = ceil_div(E, kChunk);
      uint32_t* tc = g->tile_cnt.as<uint32_t>();
      CIdx* cidx = g->csc_cidx.as<CIdx>();
      if (E) {
        hipLaunchKernelGGL((k_run_count<R, KeyQ<true>>), dim3(chunk_grid(E)), dim3(kNT), 0, s, a, E, key, tc, nch);
        EG_HIP(hipGetLastError());
        EG_TRY(exclusive_scan_u32(tc, tc, nch + 1, stmp32, s));
        prof_begin(s);
        hipLaunchKernelGGL((k_cindex_write<R>), dim3(chunk_grid(E)), dim3(kNT), 0, s, a, E, key, (const uint32_t*)tc,
                           meta, cidx);
        prof_end(PK_CINDEX, E * rb, s);
        EG_HIP(hipGetLastError());
      } else {
        EG_TRY(fill<uint32_t>(tc, 2, 0u, s));
      }
      hipLaunchKernelGGL((k_cindex_sentinel<R>), dim3(Q), dim3(kNT), 0, s, a, E, key, (const uint32_t*)tc, V, meta,
                         cidx);
      EG_HIP(hipGetLastError());
    }

    // --- dense CSR index + bitmap: runs write their gaps, long gaps go to a
    // list written by a second launch, empty sub-partitions by a third ---
    {
      uint64_t* index = g->csr_index.as<uint64_t>();
      uint64_t* bitmap = g->csr_bitmap.as<uint64_t>();
      // a gap of L > kShortGap entries takes ceil(L / kGapPiece) <= L / 65 + 1
      // pieces and the gaps of a row add up to at most V + 1 entries, so the
      // list cannot fill up (the fallback in put_gap is never taken)
      const uint64_t max_pieces = 2 * idx_n / (kShortGap + 1) + 1;
      if (max_pieces >= (uint64_t(1) << 32)) return fail(EG_ERR_ARG, "CSR index of %llu entries is too large",
                                                         (unsigned long long)idx_n);
      const uint32_t max_gaps = uint32_t(max_pieces);
      EG_TRY(g->idx_tmp.resize(uint64_t(max_gaps) * sizeof(GapPiece) + 64));
      GapPiece* gaps = g->idx_tmp.as<GapPiece>();
      uint32_t* ngaps = reinterpret_cast<uint32_t*>(gaps + max_gaps);
      EG_TRY(fill<uint64_t>(bitmap, uint64_t(Q) * words, uint64_t(0), s));
      EG_TRY(fill<uint32_t>(ngaps, 1, 0u, s));
      if (E) {
        prof_begin(s);
        hipLaunchKernelGGL((k_csr_gaps<R>), dim3(chunk_grid(E)), dim3(kNT), 0, s, g->csr_list.as<R>(), E,
                           KeyQ<false>{meta->q_off_csr, Q, g->vbits}, V, (const eg_graph::Meta*)meta, index,
                           reinterpret_cast<unsigned long long*>(bitmap), gaps, ngaps, max_gaps);
        prof_end(PK_CSR_INDEX, E * rb, s);
        EG_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_csr_long_gaps, dim3(1024), dim3(kNT), 0, s, index, (const GapPiece*)gaps,
                           (const uint32_t*)ngaps, max_gaps);
        EG_HIP(hipGetLastError());
      }
      hipLaunchKernelGGL(k_csr_empty_rows, dim3(64, unsigned(Q)), dim3(kNT), 0, s, index, V,
                         (const eg_graph::Meta*)meta);
      EG_HIP(hipGetLastError());
    }
  }

  // --- small tables back to the host ---
  eg_graph::Meta h;
  uint32_t sort_err = 0;
  std::vector<uint64_t> bpos_h(P > 1 ? 2 * uint64_t(P) * (Q + 1) : 0);
  EG_HIP(hipMemcpyAsync(&h, meta, sizeof(h), hipMemcpyDeviceToHost, s));
  EG_HIP(hipMemcpyAsync(&sort_err, g->sortw.err.ptr, 4, hipMemcpyDeviceToHost, s));
  if (P > 1)
    EG_HIP(hipMemcpyAsync(bpos_h.data(), g->blk.as<uint64_t>() + P + 1, bpos_h.size() * 8, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  if (sort_err) return fail(EG_ERR_HIP, "radix sort look-back exceeded its spin limit");
  // prep_data: every rank's allreduce of each degree array (V x 4 B; the
  // build keeps out- and in-degrees, core/graph.hpp:716 counts the one it
  // computes), and per shuffle the records of rank slice r owned by another
  // partition (core/graph.hpp:906-908; block (r, q) holds rank r's records
  // owned by sub-partition q, of partition q / S)
  g->prep_data = 2 * uint64_t(P) * V * 4;
  for (int dir = 0; dir < 2 && P > 1; ++dir)
    for (int r = 0; r < P; ++r)
      for (int q = 0; q < Q; ++q) {
        const uint64_t* b = bpos_h.data() + uint64_t(dir) * P * (Q + 1) + uint64_t(r) * (Q + 1);
        if (q / S != r) g->prep_data += (b[q + 1] - b[q]) * rb;
      }
  g->pr_cross_rows = ~0ull;
  g->page_offset.assign(h.page_offset, h.page_offset + P + 1);
  g->partition_offset.assign(h.partition_offset, h.partition_offset + P + 1);
  g->local_partition_offset.assign(h.local_partition_offset, h.local_partition_offset + P * (S + 1));
  g->q_off_csc.assign(h.q_off_csc, h.q_off_csc + Q + 1);
  g->q_off_csr.assign(h.q_off_csr, h.q_off_csr + Q + 1);
  g->cidx_off.assign(h.cidx_off, h.cidx_off + Q);
  g->cidx_cnt.assign(h.cidx_cnt, h.cidx_cnt + Q);
  if (g->q_off_csc[Q] != E || g->q_off_csr[Q] != E)
    return fail(EG_ERR_STATE, "internal: sub-partition edge counts do not add up to E");
  g->built = true;
  return EG_OK;
}

template <class R>
int build_impl(eg_graph* g, const R* d_in, uint64_t E0, int sym, hipStream_t s) {
  EG_TRY(build_setup<R>(g, E0, sym, s));
  EG_TRY(build_slices<R>(g, d_in, E0, sym, 0, g->P, s));
  return build_finish<R>(g, s);
}

}  // namespace
}  // namespace eg

using namespace eg;

extern "C" {

int eg_graph_create(uint32_t V, int P, int S, uint64_t eta, uint32_t m, int rec_bytes, eg_graph** out) {
  if (!out) return fail(EG_ERR_ARG, "eg_graph_create: out is NULL");
  *out = nullptr;
  if (V == 0) return fail(EG_ERR_ARG, "vertices_num must be positive");
  if (P < 1 || S < 1 || S > 8 || P * S > 256) return fail(EG_ERR_ARG, "need 1 <= P, 1 <= S <= 8, P*S <= 256");
  if (rec_bytes != 8 && rec_bytes != 12) return fail(EG_ERR_ARG, "rec_bytes must be 8 or 12");
  eg_graph* g = new (std::nothrow) eg_graph();
  if (!g) return fail(EG_ERR_NOMEM, "host allocation failed");
  g->V = V;
  g->P = P;
  g->S = S;
  g->Q = P * S;
  g->eta = eta;
  g->m = m;
  g->rec = rec_bytes;
  g->pages = uint32_t(ceil_div(V, kPageSize));
  *out = g;
  return EG_OK;
}

int eg_graph_destroy(eg_graph* g) {
  delete g;
  return EG_OK;
}

int eg_graph_build(eg_graph* g, const void* d_edges, uint64_t E0, int symmetrize, void* stream) {
  if (!g) return fail(EG_ERR_ARG, "eg_graph_build: graph is NULL");
  if (!d_edges && E0) return fail(EG_ERR_ARG, "eg_graph_build: edge list is NULL");
  const uint64_t E = symmetrize ? 2 * E0 : E0;
  if (E >= (uint64_t(1) << 32)) return fail(EG_ERR_ARG, "edge count %llu exceeds 2^32-1", (unsigned long long)E);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (g->rec == 8) return build_impl<Rec8>(g, static_cast<const Rec8*>(d_edges), E0, symmetrize ? 1 : 0, s);
  return build_impl<Rec12>(g, static_cast<const Rec12*>(d_edges), E0, symmetrize ? 1 : 0, s);
}

int eg_graph_build_file(eg_graph* g, const char* path, void* d_edges, uint64_t capacity_records, int symmetrize,
                        uint64_t* h_records, void* stream) {
  if (!g) return fail(EG_ERR_ARG, "eg_graph_build_file: graph is NULL");
  uint64_t E0 = 0;
  EG_TRY(eg_edge_file_records(path, g->rec, &E0));
  if (E0 > capacity_records)
    return fail(EG_ERR_ARG, "%s holds %llu records, the buffer %llu", path, (unsigned long long)E0,
                (unsigned long long)capacity_records);
  if (!d_edges && E0) return fail(EG_ERR_ARG, "eg_graph_build_file: d_edges is NULL");
  const int sym = symmetrize ? 1 : 0;
  if ((sym ? 2 * E0 : E0) >= (uint64_t(1) << 32))
    return fail(EG_ERR_ARG, "edge count %llu exceeds 2^32-1", (unsigned long long)(sym ? 2 * E0 : E0));
  const int fd = ::open(path, O_RDONLY);
  if (fd < 0) return fail(EG_ERR_ARG, "cannot open %s", path);
  hipStream_t s = static_cast<hipStream_t>(stream);
  std::vector<uint64_t> bounds(g->P + 1);
  for (int r = 0; r <= g->P; ++r) bounds[r] = slice_start(E0, g->P, 1, r) * uint64_t(g->rec);
  int rc;
  if (g->rec == 8) {
    rc = build_setup<Rec8>(g, E0, sym, s);
    if (rc == EG_OK)
      rc = upload_file_ranges(fd, bounds, static_cast<char*>(d_edges), s, [&](int r) {
        return build_slices<Rec8>(g, static_cast<const Rec8*>(d_edges), E0, sym, r, r + 1, s);
      });
    if (rc == EG_OK) rc = build_finish<Rec8>(g, s);
  } else {
    rc = build_setup<Rec12>(g, E0, sym, s);
    if (rc == EG_OK)
      rc = upload_file_ranges(fd, bounds, static_cast<char*>(d_edges), s, [&](int r) {
        return build_slices<Rec12>(g, static_cast<const Rec12*>(d_edges), E0, sym, r, r + 1, s);
      });
    if (rc == EG_OK) rc = build_finish<Rec12>(g, s);
  }
  ::close(fd);
  if (rc != EG_OK) {
    (void)hipStreamSynchronize(s);  // nothing queued may outlive a failed call's buffers
    return rc;
  }
  if (h_records) *h_records = E0;
  return EG_OK;
}

int eg_autotune_sort(eg_graph* g, const void* d_edges, uint64_t E0, int symmetrize, int* h_scheme, int* h_threads,
                     int* h_bits, double* h_ms, void* stream) {
  if (!g) return fail(EG_ERR_ARG, "eg_autotune_sort: graph is NULL");
  struct Cand {
    int scheme, nt, bits;
  };
  std::vector<Cand> cands;
  // 9-bit digits change nothing unless they save a pass on the vertex keys
  const int vb = std::max(1, counting_bits(g->V - 1));
  const bool nine = rx_plan(vb, 9).passes < rx_plan(vb, 8).passes;
  for (int sc : {kThreeLaunch, kThreeLaunchXcd, kOnePass}) {
    cands.push_back({sc, 512, 8});
    if (nine) cands.push_back({sc, 512, 9});
    if (sc != kOnePass) cands.push_back({sc, 256, 8});
    if (g->rec == 8) cands.push_back({sc, 1024, 8});
  }
  const SortVariant keep = sort_variant(g->rec);
  double best = 1e300;
  SortVariant pick = keep;
  for (const Cand& c : cands) {
    sort_variant(g->rec) = SortVariant{c.scheme, c.nt, c.bits};
    double t = 0;
    for (int rep = 0; rep < 2; ++rep) {  // the first build may allocate
      const auto t0 = std::chrono::steady_clock::now();
      const int rc = eg_graph_build(g, d_edges, E0, symmetrize, stream);
      if (rc != EG_OK) {
        sort_variant(g->rec) = keep;
        return rc;
      }
      t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    if (t < best) {
      best = t;
      pick = SortVariant{c.scheme, c.nt, c.bits};
    }
  }
  sort_variant(g->rec) = pick;
  const int rc = eg_graph_build(g, d_edges, E0, symmetrize, stream);  // leave the graph built
  if (h_scheme) *h_scheme = pick.scheme;
  if (h_threads) *h_threads = pick.nt;
  if (h_bits) *h_bits = pick.bits;
  if (h_ms) *h_ms = best;
  return rc;
}

int eg_graph_get_info(const eg_graph* g, eg_graph_info* out) {
  if (!g || !out) return fail(EG_ERR_ARG, "eg_graph_get_info: NULL argument");
  out->vertices_num = g->V;
  out->partitions = g->P;
  out->sockets = g->S;
  out->record_bytes = g->rec;
  out->edges_num = g->E;
  out->pages_num = g->pages;
  out->symmetrized = g->symmetrized;
  return EG_OK;
}

static int need_built(const eg_graph* g) {
  if (!g) return fail(EG_ERR_ARG, "graph is NULL");
  if (!g->built) return fail(EG_ERR_STATE, "graph has not been built");
  return EG_OK;
}

int eg_graph_partition_offset(const eg_graph* g, uint32_t* h_out) {
  EG_TRY(need_built(g));
  if (!h_out) return fail(EG_ERR_ARG, "h_out is NULL");
  std::memcpy(h_out, g->partition_offset.data(), sizeof(uint32_t) * (g->P + 1));
  return EG_OK;
}

int eg_graph_local_partition_offset(const eg_graph* g, int p, uint32_t* h_out) {
  EG_TRY(need_built(g));
  if (!h_out || p < 0 || p >= g->P) return fail(EG_ERR_ARG, "bad partition %d", p);
  std::memcpy(h_out, g->local_partition_offset.data() + p * (g->S + 1), sizeof(uint32_t) * (g->S + 1));
  return EG_OK;
}

int eg_graph_degree(const eg_graph* g, const uint32_t** d_out_degree, const uint32_t** d_in_degree) {
  EG_TRY(need_built(g));
  if (d_out_degree) *d_out_degree = g->out_deg.as<uint32_t>();
  if (d_in_degree) *d_in_degree = g->in_deg.as<uint32_t>();
  return EG_OK;
}

int eg_graph_page_prefix(const eg_graph* g, const uint64_t** d_prefix) {
  EG_TRY(need_built(g));
  if (!d_prefix) return fail(EG_ERR_ARG, "d_prefix is NULL");
  *d_prefix = g->page_prefix.as<uint64_t>();
  return EG_OK;
}

int eg_graph_csc(const eg_graph* g, int p, int s, const void** d_list, uint64_t* n_edges, const void** d_cindex,
                 uint64_t* n_cindex) {
  EG_TRY(need_built(g));
  if (p < 0 || p >= g->P || s < 0 || s >= g->S) return fail(EG_ERR_ARG, "bad (partition, socket) (%d, %d)", p, s);
  const int q = p * g->S + s;
  if (d_list) *d_list = static_cast<const char*>(g->csc_list.ptr) + g->q_off_csc[q] * g->rec;
  if (n_edges) *n_edges = g->q_off_csc[q + 1] - g->q_off_csc[q];
  if (d_cindex) *d_cindex = g->csc_cidx.as<CIdx>() + g->cidx_off[q];
  if (n_cindex) *n_cindex = g->cidx_cnt[q];
  return EG_OK;
}

int eg_graph_csr(const eg_graph* g, int p, int s, const void** d_list, uint64_t* n_edges, const uint64_t** d_index,
                 const uint64_t** d_bitmap) {
  EG_TRY(need_built(g));
  if (p < 0 || p >= g->P || s < 0 || s >= g->S) return fail(EG_ERR_ARG, "bad (partition, socket) (%d, %d)", p, s);
  const int q = p * g->S + s;
  if (d_list) *d_list = static_cast<const char*>(g->csr_list.ptr) + g->q_off_csr[q] * g->rec;
  if (n_edges) *n_edges = g->q_off_csr[q + 1] - g->q_off_csr[q];
  if (d_index) *d_index = g->csr_index.as<uint64_t>() + uint64_t(q) * (g->V + 1);
  if (d_bitmap) *d_bitmap = g->csr_bitmap.as<uint64_t>() + uint64_t(q) * ((g->V >> 6) + 1);
  return EG_OK;
}

}  // extern "C"
