// Batched traversals: BFS and SSSP from up to 64 roots at once (no
// reference counterpart: toolkits/bfs.cpp and toolkits/sssp.cpp run one
// root per process).  The 64 roots are the 64 lanes of a CDNA wavefront:
// per-vertex state is one u64 mask (bit r = root r) and per-(vertex, root)
// results are stored vertex-major, x[v*64 + r], so that a wave serving one
// vertex reads or writes its 64 roots as one 256-byte row.
//
// A traversal runs as steps of three launches each, queued four steps at a
// time between host checks (the device decides the direction of each step,
// so no step waits for the host):
//   push       sparse steps expand work items (a frontier vertex and a
//              4096-edge chunk of its out-edges, so hubs spread over waves);
//   pull       dense steps: a wave per destination walks its in-edges 64 at
//              a time (rows sorted by source) and, per root, takes the first
//              frontier source it meets;
//   end        new bits -> visited / next frontier, results, work items and
//              the next step's counters.
// Only one of push / pull works in a step (the rule of
// core/graph.hpp:1483-1498 with the policies of the single-root programs,
// evaluated on the device from the counters); a step queued after the
// frontier emptied returns at once.
// Results equal those of the single-root programs bit for bit: BFS parents
// follow the same min-id rule (the smallest frontier in-neighbour of the
// level that reaches v), depths are levels, SSSP distances are the unique
// fixpoint of dist[v] = min(dist[v], dist[u] + w) with fp32 adds.
#include <algorithm>
#include <chrono>
#include <vector>

#include "graph_state.hpp"
#include "rowreduce.hpp"
#include "stepgraph.hpp"

namespace eg {
namespace {

constexpr int kNT = 256;
constexpr int kEndStrides = 4;         // grid strides a level-end wave takes per step
constexpr uint32_t kPullEdges = kTvPullEdges;  // in-edges per pull work item
constexpr uint32_t kUnreached = 0xffffffffu;
constexpr float kInfDist = 1e9f;       // toolkits/sssp.cpp INF

unsigned grid_for(uint64_t n, uint64_t per = kNT, uint64_t cap = 65536) {
  uint64_t b = ceil_div(n, per);
  return unsigned(b < 1 ? 1 : (b > cap ? cap : b));
}

// Counters of one level: vertices with new bits, their out-degree sum, the
// roots that still have a frontier and the number of push work items.
struct MCount {
  unsigned long long nv, ne, roots;
  unsigned int items, pad;
};

struct OpOr {
  __device__ __forceinline__ uint64_t operator()(uint64_t a, uint64_t b) const { return a | b; }
};

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int lane) {
  return (uint64_t(__shfl(uint32_t(v >> 32), lane, kWave)) << 32) | __shfl(uint32_t(v), lane, kWave);
}
__device__ __forceinline__ uint64_t wave_or(uint64_t v) {
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) {
    const uint32_t hi = __shfl_xor(uint32_t(v >> 32), d, kWave), lo = __shfl_xor(uint32_t(v), d, kWave);
    v |= (uint64_t(hi) << 32) | lo;
  }
  return v;
}

// Inclusive OR over lanes 0..l (DPP row shifts, then row broadcasts).
__device__ __forceinline__ uint64_t or_scan(uint64_t v) {
  v |= dpp_move<0x111, 0xF>(uint64_t(0), v);
  v |= dpp_move<0x112, 0xF>(uint64_t(0), v);
  v |= dpp_move<0x114, 0xF>(uint64_t(0), v);
  v |= dpp_move<0x118, 0xF>(uint64_t(0), v);
  v |= dpp_move<0x142, 0xA>(uint64_t(0), v);
  v |= dpp_move<0x143, 0xC>(uint64_t(0), v);
  return v;
}

__device__ __forceinline__ bool sparse_level(int policy, unsigned long long active_edges, uint64_t E,
                                             uint64_t dense_div = 20) {
  if (policy == EG_POLICY_PUSH) return true;
  if (policy == EG_POLICY_PULL) return false;
  if (policy == EG_POLICY_REFERENCE) return true;  // rooted programs
  return active_edges < E / dense_div;
}

// auto policy of the batched SSSP: dense once the active out-edges reach
// E / tune.msssp_div.  Its pull skips the sources no root has active and
// gathers 64-root distance rows, so pushes pay until half of E (default 2;
// 64 R-MAT 22 roots: E/20 47.7 ms, E/5 46.1, E/3 45.4, E/2 44.4, E/1 106,
// push only 109).  The batched BFS goes dense at E / tune.mbfs_div (default
// 20; 64 R-MAT 22 roots: E/5..E/20 within 1%, E/40 7% slower).

template <class T>
__global__ void k_fill_t(T* __restrict__ p, uint64_t n, T v) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

template <class T>
int fill_t(T* p, uint64_t n, T v, hipStream_t s) {
  if (!n) return EG_OK;
  hipLaunchKernelGGL((k_fill_t<T>), dim3(grid_for(n, kNT, 65536)), dim3(kNT), 0, s, p, n, v);
  EG_HIP(hipGetLastError());
  return EG_OK;
}

// ------------------------------------------------------ traversal view --
// The built CSR is grouped by the sub-partition of the destination, and inside
// one by (source, rank, destination, position); a stable sort of it by
// destination alone (the sub-partition follows from the destination) leaves
// every destination's in-edges in ascending source order: one sort of the
// vertex-id width instead of one of twice that width over the CSC.
struct KeyDstOnly {
  static constexpr int kField = 1;
  template <class R>
  __device__ __forceinline__ uint64_t operator()(const R& r) const { return r.dst; }
  __device__ __forceinline__ uint64_t of_field(uint32_t k) const { return k; }
};

__global__ void k_copy_u32(const uint32_t* __restrict__ a, uint32_t* __restrict__ b, uint64_t n) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    b[i] = a[i];
}

// Pull work items: in-edge chunks [c0, c1) of at most kPullEdges edges per
// vertex, of the vertices with more than min_deg in-edges; chunks c0.. up
// to the last (c1 = 0) or to c1.
__global__ void k_tv_item_count(const uint32_t* __restrict__ in_deg, uint32_t V, uint32_t min_deg, uint32_t c0,
                                uint32_t c1, uint32_t* __restrict__ cnt) {
  for (uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += uint64_t(gridDim.x) * blockDim.x) {
    uint32_t n = in_deg[v] > min_deg ? uint32_t(ceil_div(in_deg[v], kPullEdges)) : 0u;
    if (c1 && n > c1) n = c1;
    cnt[v] = n > c0 ? n - c0 : 0u;
  }
}

__global__ void k_tv_items(const uint32_t* __restrict__ off, uint32_t V, uint32_t c0, uint64_t* __restrict__ items) {
  for (uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += uint64_t(gridDim.x) * blockDim.x)
    for (uint32_t j = 0; j < off[v + 1] - off[v]; ++j) items[off[v] + j] = (v << 32) | (c0 + j);
}

template <class R>
int tview_impl(eg_graph* g, hipStream_t s) {
  const uint64_t E = g->E;
  const uint32_t V = g->V;
  EG_TRY(g->tv_ptr.resize((uint64_t(V) + 1) * 4));
  EG_TRY(g->tv_src.resize((std::max<uint64_t>(E, 1) + 4) * 4));  // 4 words of padding for 16-byte group loads
  if (sizeof(R) == 12) EG_TRY(g->tv_w.resize(std::max<uint64_t>(E, 1) * 4));
  uint32_t* ptr = g->tv_ptr.as<uint32_t>();
  // rows sorted by destination: row v starts at the prefix of the in-degree
  hipLaunchKernelGGL(k_copy_u32, dim3(grid_for(V)), dim3(kNT), 0, s, (const uint32_t*)g->in_deg.as<uint32_t>(), ptr,
                     uint64_t(V));
  EG_HIP(hipGetLastError());
  EG_TRY(fill_t<uint32_t>(ptr + V, 1, 0u, s));
  EG_TRY(exclusive_scan_u32(ptr, ptr, uint64_t(V) + 1, g->scan_tmp.as<uint32_t>(), s));
  // items of every row (batched programs); of the rows above kTvSmallRow
  // in-edges (single-root BFS pull; smaller rows go one per lane): their
  // first chunks, then the others
  struct List {
    uint32_t min_deg, c0, c1;
    DevBuf* items;
    uint64_t* n;
  };
  const List lists[3] = {{0u, 0u, 0u, &g->tv_items, &g->tv_nitems},
                         {kTvSmallRow, 0u, 1u, &g->tv_big, &g->tv_nbig},
                         {kTvSmallRow, 1u, 0u, &g->tv_big_rest, &g->tv_nbig_rest}};
  // item offsets in the engine's scratch (V * 4 + 4096 bytes or more, idle here)
  if (g->e_scratch.bytes < (uint64_t(V) + 1) * 4) return fail(EG_ERR_STATE, "internal: engine scratch too small");
  uint32_t* o = g->e_scratch.as<uint32_t>();
  for (const List& l : lists) {
    hipLaunchKernelGGL(k_tv_item_count, dim3(grid_for(V)), dim3(kNT), 0, s, (const uint32_t*)g->in_deg.as<uint32_t>(),
                       V, l.min_deg, l.c0, l.c1, o);
    EG_HIP(hipGetLastError());
    EG_TRY(fill_t<uint32_t>(o + V, 1, 0u, s));
    EG_TRY(exclusive_scan_u32(o, o, uint64_t(V) + 1, g->scan_tmp.as<uint32_t>(), s));
    uint32_t n = 0;
    EG_HIP(hipMemcpyAsync(&n, o + V, 4, hipMemcpyDeviceToHost, s));
    EG_HIP(hipStreamSynchronize(s));
    *l.n = n;
    EG_TRY(l.items->resize(std::max<uint64_t>(n, 1) * 8));
    hipLaunchKernelGGL(k_tv_items, dim3(grid_for(V)), dim3(kNT), 0, s, (const uint32_t*)o, V, l.c0,
                       static_cast<uint64_t*>(l.items->ptr));
    EG_HIP(hipGetLastError());
  }
  if (E) {
    // The CSR sorted stably by destination, written as field arrays: the
    // sources straight into tv_src (and the weights into tv_w), the
    // destinations (implied by tv_ptr) into the build's sort temporary t1;
    // t2 holds the intermediate passes.  Both are free after the build.
    EG_TRY(g->sortw.resize_segs(1));
    EG_TRY(fill_t<uint32_t>(g->sortw.err.as<uint32_t>(), 1, 0u, s));
    EG_TRY(g->t1.resize(E * 4));
    EG_TRY(g->t2.resize(rx_soa_bytes<R>(E)));
    RxSoa<R> out;
    out.s = g->tv_src.as<uint32_t>();
    out.d = g->t1.as<uint32_t>();
    out.w = sizeof(R) == 12 ? g->tv_w.as<float>() : nullptr;
    EG_TRY(rx_sort_to_fields(g->csr_list.as<R>(), out, rx_soa_in<R>(g->t2.ptr, E), std::vector<uint64_t>{0, E},
                             KeyDstOnly{}, g->vbits, g->sortw, s));
  }
  if (E) {
    uint32_t err = 0;
    EG_HIP(hipMemcpyAsync(&err, g->sortw.err.ptr, 4, hipMemcpyDeviceToHost, s));
    EG_HIP(hipStreamSynchronize(s));
    if (err) return fail(EG_ERR_HIP, "radix sort look-back exceeded its spin limit");
  }
  EG_HIP(hipStreamSynchronize(s));
  return EG_OK;
}

}  // namespace

int ensure_tview(eg_graph* g, hipStream_t s) {
  EG_TRY(engine_ensure(g, s));
  if (g->tv_ready) return EG_OK;
  const auto t0 = std::chrono::steady_clock::now();
  EG_TRY(g->rec == 8 ? tview_impl<Rec8>(g, s) : tview_impl<Rec12>(g, s));
  g->tv_prepare_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  g->tv_ready = true;
  return EG_OK;
}

namespace {

// ------------------------------------------------------------ shared state --
struct MultiState {
  uint64_t *vis, *fr, *nx;  // u64[V] masks (SSSP uses fr as the active set, vis unused)
  uint64_t* items[2];       // push work items (v << 32 | chunk), ping-pong by step parity
  MCount* cnt;              // ring of 3 by step: read at s, written at s-1, cleared at s-2
  uint32_t* steps;          // steps that had a nonempty frontier
  uint32_t* base;           // first step of the queued batch (step = *base + j)
  uint32_t item_edges;      // out-edges per push work item (the graph's tune.m_item)
};

uint64_t item_cap(const eg_graph* g) { return uint64_t(g->V) + g->E / g->tune.m_item + 1; }

int multi_state(eg_graph* g, MultiState* st) {
  const uint64_t V = g->V;
  const uint64_t ic = item_cap(g);
  const uint64_t bytes = 3 * V * 8 + 2 * ic * 8 + 3 * sizeof(MCount) + 256;
  EG_TRY(g->e_multi.resize(bytes));
  char* p = g->e_multi.as<char>();
  st->vis = reinterpret_cast<uint64_t*>(p);
  st->fr = st->vis + V;
  st->nx = st->fr + V;
  st->items[0] = st->nx + V;
  st->items[1] = st->items[0] + ic;
  st->cnt = reinterpret_cast<MCount*>(st->items[1] + ic);
  st->steps = reinterpret_cast<uint32_t*>(st->cnt + 3);
  st->base = st->steps + 1;
  st->item_edges = g->tune.m_item;
  return EG_OK;
}

// Roots: lane r < n sets bit r of its root; work items and counters of the
// first step come from the distinct roots.  One wave.
__global__ void k_multi_roots(const uint32_t* __restrict__ roots, int n, uint32_t item_edges,
                              const uint32_t* __restrict__ deg,
                              uint64_t* __restrict__ vis, uint64_t* __restrict__ fr, uint64_t* __restrict__ items,
                              MCount* __restrict__ cnt, uint32_t* __restrict__ parent, uint32_t* __restrict__ depth,
                              float* __restrict__ dist) {
  const int r = threadIdx.x;
  if (r >= n) return;
  const uint32_t v = roots[r];
  if (vis) atomicOr((unsigned long long*)&vis[v], 1ull << r);
  atomicOr((unsigned long long*)&fr[v], 1ull << r);
  if (parent) parent[uint64_t(v) * 64 + r] = v;
  if (depth) depth[uint64_t(v) * 64 + r] = 0;
  if (dist) dist[uint64_t(v) * 64 + r] = 0.0f;
  bool first = true;  // the lowest lane holding this root adds its work items
  for (int q = 0; q < r; ++q) first = first && roots[q] != v;
  if (!first) return;
  atomicOr(&cnt[0].roots, 1ull << r);
  const uint32_t d = deg[v];
  atomicAdd(&cnt[0].nv, 1ull);
  atomicAdd(&cnt[0].ne, (unsigned long long)d);
  const uint32_t k = uint32_t(ceil_div(d, item_edges));
  const uint32_t at = atomicAdd(&cnt[0].items, k);
  for (uint32_t j = 0; j < k; ++j) items[at + j] = (uint64_t(v) << 32) | j;
}

// Level end shared by both programs.  A wave owns the vertex groups
// v0 = (gw + i*nw) * 64 ..; lane = vertex.  Pass 1 sums the work items and
// counters of the wave's vertices, one atomic per wave reserves the items,
// pass 2 writes them and hands each vertex with new bits to `visit`
// (wave-uniform call per group: lanes without new bits pass nb = 0).
// Counters reach memory once per block.
template <class Visit>
__device__ __forceinline__ void level_end(uint32_t V, const uint32_t* __restrict__ deg, uint64_t* __restrict__ nx,
                                          uint64_t* __restrict__ fr, MCount* nc, uint64_t* __restrict__ items,
                                          uint32_t item_edges, Visit visit) {
  // A wave takes kEndStrides grid strides of 64 vertices per step, with the
  // mask loads of all of them, then the degree loads, in flight together;
  // the block reserves its work items with one atomic (a block scan).
  constexpr int R = kEndStrides;
  __shared__ unsigned long long s_nv[kNT / kWave], s_ne[kNT / kWave], s_roots[kNT / kWave];
  __shared__ uint32_t s_wt[kNT / kWave];
  __shared__ uint32_t s_base;
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t gw = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  const uint64_t nw = uint64_t(gridDim.x) * blockDim.x / kWave;
  const uint64_t stride = nw * kWave;
  unsigned long long nv = 0, ne = 0, roots = 0;
  uint32_t ksum = 0;
  for (uint64_t v0 = gw * kWave; v0 < V; v0 += stride * R) {
    uint64_t nb[R];
    uint32_t d[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint64_t v = v0 + r * stride + lane;
      nb[r] = v < V ? nx[v] : 0ull;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) d[r] = nb[r] ? deg[v0 + r * stride + lane] : 0u;
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (nb[r]) {
        roots |= nb[r];
        ++nv;
        ne += d[r];
        ksum += uint32_t(ceil_div(d[r], item_edges));
      }
  }
  uint32_t tot = 0;
  const uint32_t excl = block_exclusive_scan<kNT>(ksum, 0u, OpSum{}, s_wt, &tot);
  if (threadIdx.x == 0) s_base = tot ? atomicAdd(&nc->items, tot) : 0u;
  __syncthreads();
  uint32_t at = s_base + excl;
  for (uint64_t v0 = gw * kWave; v0 < V; v0 += stride * R) {
    uint64_t nb[R], f[R];
    uint32_t d[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint64_t v = v0 + r * stride + lane;
      nb[r] = v < V ? nx[v] : 0ull;
      f[r] = v < V ? fr[v] : 0ull;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) d[r] = nb[r] ? deg[v0 + r * stride + lane] : 0u;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint64_t v = v0 + r * stride + lane;
      if (v < V) {
        if (nb[r]) {
          nx[v] = 0;
          const uint32_t k = uint32_t(ceil_div(d[r], item_edges));
          for (uint32_t q = 0; q < k; ++q) items[at + q] = (v << 32) | q;
          at += k;
        }
        if (f[r] != nb[r]) fr[v] = nb[r];
      }
      if (v0 + r * stride < V) visit(v0 + r * stride, v, nb[r]);  // wave-uniform
    }
  }
  nv = wave_reduce(nv, OpSum{});
  ne = wave_reduce(ne, OpSum{});
  roots = wave_or(roots);
  if (lane == 0) {
    s_nv[threadIdx.x / kWave] = nv;
    s_ne[threadIdx.x / kWave] = ne;
    s_roots[threadIdx.x / kWave] = roots;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long tv = 0, te = 0, tr = 0;
    for (int i = 0; i < kNT / kWave; ++i) {
      tv += s_nv[i];
      te += s_ne[i];
      tr |= s_roots[i];
    }
    if (tv) {
      atomicAdd(&nc->nv, tv);
      atomicAdd(&nc->ne, te);
      atomicOr(&nc->roots, tr);
    }
  }
}

// Step bookkeeping shared by the level-end kernels: clears the counters of
// step s+2 (last read at step s-1) and records step s as run.
__device__ __forceinline__ void step_bookkeeping(const MultiState& st, uint32_t step, bool live) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st.cnt[(step + 2) % 3] = MCount{0, 0, 0, 0, 0};
    if (live) *st.steps = step + 1;
  }
}

// ------------------------------------------------------------------ BFS --
struct BfsArgs {
  static constexpr uint64_t kPullItemsPerWave = 64;  // the pull filters 64 items per step
  uint32_t V;
  uint64_t E;
  int policy;
  uint32_t dense_div;  // the graph's tune.mbfs_div
  uint64_t lanes;  // bits of the live roots
  const uint32_t *out_ptr, *out_dst, *deg, *in_ptr, *in_src;
  const uint64_t* pull_items;  // in-edge chunks (v << 32 | chunk)
  uint64_t n_pull_items;
  MultiState st;
  uint32_t *parent, *depth;
};

// Sparse step (bfs.cpp:52-58 per root): unvisited out-neighbours of the
// frontier take the smallest frontier id as parent.
__global__ __launch_bounds__(kNT) void k_mbfs_push(BfsArgs a, uint32_t j) {
  const uint32_t step = *a.st.base + j;
  const MCount c = a.st.cnt[step % 3];
  if (!c.nv || !sparse_level(a.policy, c.ne, a.E, a.dense_div)) return;
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t gw = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  const uint64_t nw = uint64_t(gridDim.x) * blockDim.x / kWave;
  const uint64_t* items = a.st.items[step & 1];
  for (uint64_t i = gw; i < c.items; i += nw) {
    const uint64_t it = items[i];
    const uint32_t u = uint32_t(it >> 32), ch = uint32_t(it);
    const uint64_t fm = a.st.fr[u];
    const uint32_t e0 = a.out_ptr[u] + ch * a.st.item_edges;
    const uint32_t e1 = min(e0 + a.st.item_edges, a.out_ptr[u + 1]);
    for (uint32_t e = e0 + lane; e < e1; e += kWave) {
      const uint32_t v = a.out_dst[e];
      uint64_t bits = fm & ~a.st.vis[v];
      if (!bits) continue;
      atomicOr((unsigned long long*)&a.st.nx[v], bits);
      while (bits) {
        const int r = __ffsll((unsigned long long)bits) - 1;
        atomicMin(&a.parent[uint64_t(v) * 64 + r], u);
        bits &= bits - 1;
      }
    }
  }
}

// Dense step (bfs.cpp:60-68 per root): a wave takes an in-edge chunk (at
// most 512 edges, in source order) of a destination; the first frontier
// source of a root in the chunk is the chunk's smallest, write_min over the
// chunks gives the smallest of the row.  Only roots that still have a
// frontier are looked for, and a chunk stops once all of them are found.
__global__ __launch_bounds__(kNT) void k_mbfs_pull(BfsArgs a, uint32_t j) {
  const uint32_t step = *a.st.base + j;
  const MCount c = a.st.cnt[step % 3];
  if (!c.nv || sparse_level(a.policy, c.ne, a.E, a.dense_div)) return;
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t gw = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  const uint64_t nw = uint64_t(gridDim.x) * blockDim.x / kWave;
  const uint64_t live_roots = a.lanes & c.roots;
  __shared__ uint32_t hand[kNT / kWave][kWave];
  uint32_t* pslot = hand[threadIdx.x / kWave];
  // 64 items per step: every lane checks one (all their loads in flight at
  // once), then the wave works through the items that still need a root
  for (uint64_t i0 = gw * kWave; i0 < a.n_pull_items; i0 += nw * kWave) {
    const bool mine = i0 + lane < a.n_pull_items;
    const uint64_t my_it = mine ? a.pull_items[i0 + lane] : 0ull;
    const uint64_t my_need = mine ? live_roots & ~a.st.vis[my_it >> 32] : 0ull;
    uint64_t pending = __ballot(my_need != 0);
    while (pending) {  // wave-uniform
      const int jj = __ffsll((unsigned long long)pending) - 1;
      pending &= pending - 1;
      const uint64_t it = shfl64(my_it, jj);
      const uint64_t v = it >> 32;
      const uint64_t need = shfl64(my_need, jj);
      const uint32_t r0 = a.in_ptr[v], r1 = a.in_ptr[v + 1];
      const uint32_t e0 = r0 + uint32_t(it) * kPullEdges;
      const uint32_t b1 = min(e0 + kPullEdges, r1);
      // an item that is its row's only one (at most kPullEdges in-edges) is
      // the only writer of the row's parents and next mask in this step:
      // plain stores instead of atomics
      const bool whole = e0 == r0 && b1 == r1;
      uint64_t found = 0;
      for (uint32_t e = e0; e < b1 && found != need; e += kWave) {
        const bool ok = e + lane < b1;
        const uint32_t u = ok ? a.in_src[e + lane] : 0u;
        const uint64_t f = ok ? (a.st.fr[u] & need & ~found) : 0ull;
        // root r's parent is the source of the first lane with bit r (the
        // smallest: sources ascend); that lane hands it to lane r in LDS
        const uint64_t incl = or_scan(f);
        const uint64_t below = dpp_move<0x138, 0xF>(uint64_t(0), incl);  // wave_shr:1: lanes below
        const uint64_t newly = lane_value(incl, 63);
        if (!newly) continue;  // wave-uniform
        for (uint64_t m = f & ~below; m; m &= m - 1) pslot[__ffsll((unsigned long long)m) - 1] = u;
        __builtin_amdgcn_s_waitcnt(kLgkmZero);  // every hand-over written before lane r reads it
        __asm__ volatile("" ::: "memory");
        if ((newly >> lane) & 1ull) {
          if (whole)
            a.parent[v * 64 + lane] = pslot[lane];
          else
            atomicMin(&a.parent[v * 64 + lane], pslot[lane]);
        }
        __asm__ volatile("" ::: "memory");
        found |= newly;
      }
      if (found && lane == 0) {
        if (whole)
          a.st.nx[v] = found;
        else
          atomicOr((unsigned long long*)&a.st.nx[v], found);
      }
    }
  }
}

// Level end (bfs.cpp:70-90 per root): visited |= next, depth = level + 1,
// next becomes the frontier; counters and work items of the next step.
__global__ __launch_bounds__(kNT) void k_mbfs_end(BfsArgs a, uint32_t j) {
  const uint32_t step = *a.st.base + j;
  const bool live = a.st.cnt[step % 3].nv != 0;
  step_bookkeeping(a.st, step, live);
  if (!live) return;  // block-uniform
  const int lane = threadIdx.x & (kWave - 1);
  level_end(a.V, a.deg, a.st.nx, a.st.fr, &a.st.cnt[(step + 1) % 3], a.st.items[(step + 1) & 1], a.st.item_edges,
            [&](uint64_t v0, uint64_t v, uint64_t nb) {
              if (nb) a.st.vis[v] |= nb;
              // depth rows: lane = root, one 256-byte row per vertex
              uint64_t rows = __ballot(nb != 0);
              while (rows) {
                const int j = __ffsll((unsigned long long)rows) - 1;
                rows &= rows - 1;
                const uint64_t bj = shfl64(nb, j);
                if ((bj >> lane) & 1ull) a.depth[(v0 + j) * 64 + lane] = step + 1;
              }
            });
}

// ----------------------------------------------------------------- SSSP --
struct SsspArgs {
  static constexpr uint64_t kPullItemsPerWave = 1;
  uint32_t V;
  uint64_t E;
  int policy;
  uint32_t dense_div;  // the graph's tune.msssp_div
  uint64_t lanes;
  const uint32_t *out_ptr, *out_dst, *deg, *in_ptr, *in_src;
  const float *out_w, *in_w;
  const uint64_t* pull_items;  // in-edge chunks (v << 32 | chunk)
  uint64_t n_pull_items;
  MultiState st;
  float* dist;
};

// Sparse round (sssp.cpp:44-58 per root): relax dist[u] + w into the
// out-neighbours; write_min on the bits of non-negative floats.
__global__ __launch_bounds__(kNT) void k_msssp_push(SsspArgs a, uint32_t j) {
  const uint32_t step = *a.st.base + j;
  const MCount c = a.st.cnt[step % 3];
  if (!c.nv || !sparse_level(a.policy, c.ne, a.E, a.dense_div)) return;
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t gw = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  const uint64_t nw = uint64_t(gridDim.x) * blockDim.x / kWave;
  const uint64_t* items = a.st.items[step & 1];
  for (uint64_t i = gw; i < c.items; i += nw) {
    const uint64_t it = items[i];
    const uint32_t u = uint32_t(it >> 32), ch = uint32_t(it);
    const uint64_t am = a.st.fr[u];
    const uint32_t e0 = a.out_ptr[u] + ch * a.st.item_edges;
    const uint32_t e1 = min(e0 + a.st.item_edges, a.out_ptr[u + 1]);
    for (uint32_t e = e0 + lane; e < e1; e += kWave) {
      const uint32_t v = a.out_dst[e];
      const float w = a.out_w[e];
      for (uint64_t b = am; b; b &= b - 1) {
        const int r = __ffsll((unsigned long long)b) - 1;
        const float nd = __fadd_rn(a.dist[uint64_t(u) * 64 + r], w);
        float* p = &a.dist[uint64_t(v) * 64 + r];
        if (nd < *p) {
          const int old = atomicMin(reinterpret_cast<int*>(p), __float_as_int(nd));
          if (nd < __int_as_float(old)) atomicOr((unsigned long long*)&a.st.nx[v], 1ull << r);
        }
      }
    }
  }
}

// Dense round (sssp.cpp:60-73 per root): min over active in-neighbours of
// dist[u] + w; lane r is root r, so each in-edge reads its source's 64
// distances as one row.  A wave takes an in-edge chunk of at most 512 edges
// (hub rows spread over many waves) and lowers dist[v] with write_min.
__global__ __launch_bounds__(kNT) void k_msssp_pull(SsspArgs a, uint32_t j) {
  const uint32_t step = *a.st.base + j;
  const MCount c = a.st.cnt[step % 3];
  if (!c.nv || sparse_level(a.policy, c.ne, a.E, a.dense_div)) return;
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t gw = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  const uint64_t nw = uint64_t(gridDim.x) * blockDim.x / kWave;
  for (uint64_t i = gw; i < a.n_pull_items; i += nw) {
    const uint64_t it = a.pull_items[i];
    const uint64_t v = it >> 32;
    const uint32_t r0 = a.in_ptr[v], r1 = a.in_ptr[v + 1];
    const uint32_t e0 = r0 + uint32_t(it) * kPullEdges;
    const uint32_t b1 = min(e0 + kPullEdges, r1);
    const bool whole = e0 == r0 && b1 == r1;  // the row's only item: its only writer in this step
    float best = kInfDist;
    for (uint32_t e = e0; e < b1; e += kWave) {
      const bool ok = e + lane < b1;
      const uint32_t u = ok ? a.in_src[e + lane] : 0u;
      const float w = ok ? a.in_w[e + lane] : 0.0f;
      const uint64_t am = ok ? a.st.fr[u] : 0ull;
      uint64_t edges = __ballot(am != 0);  // in-edges of the chunk with an active source
      while (edges) {
        // eight active in-edges at a time: all distance rows in flight
        // before the first use
        constexpr int kG = 8;
        uint32_t uj[kG];
        float wj[kG], dj[kG];
        bool take[kG];
#pragma unroll
        for (int k = 0; k < kG; ++k) {
          const int j = edges ? __ffsll((unsigned long long)edges) - 1 : 0;
          const bool live = edges != 0;
          edges &= edges - 1;
          uj[k] = __shfl(u, j, kWave);
          wj[k] = __shfl(w, j, kWave);
          take[k] = live && ((shfl64(am, j) >> lane) & 1ull);
        }
#pragma unroll
        for (int k = 0; k < kG; ++k) dj[k] = take[k] ? a.dist[uint64_t(uj[k]) * 64 + lane] : kInfDist;
#pragma unroll
        for (int k = 0; k < kG; ++k)
          if (take[k]) best = fminf(best, __fadd_rn(dj[k], wj[k]));
      }
    }
    float* p = &a.dist[v * 64 + lane];
    bool lower = ((a.lanes >> lane) & 1ull) && best < *p;
    if (lower) {
      if (whole)
        *p = best;
      else
        lower = best < __int_as_float(atomicMin(reinterpret_cast<int*>(p), __float_as_int(best)));
    }
    const uint64_t m = __ballot(lower);
    if (m && lane == 0) {
      if (whole)
        a.st.nx[v] = m;
      else
        atomicOr((unsigned long long*)&a.st.nx[v], m);
    }
  }
}

// Round end: the lowered vertices are the next active set.
__global__ __launch_bounds__(kNT) void k_msssp_end(SsspArgs a, uint32_t j) {
  const uint32_t step = *a.st.base + j;
  const bool live = a.st.cnt[step % 3].nv != 0;
  step_bookkeeping(a.st, step, live);
  if (!live) return;  // block-uniform
  level_end(a.V, a.deg, a.st.nx, a.st.fr, &a.st.cnt[(step + 1) % 3], a.st.items[(step + 1) & 1], a.st.item_edges,
            [](uint64_t, uint64_t, uint64_t) {});
}

// Per-root statistics of a finished BFS: levels (max depth + 1) and edges
// traversed (out-degree sum of the reached vertices), as eg_bfs reports.
__global__ __launch_bounds__(kNT) void k_bfs_multi_stats(const uint32_t* __restrict__ depth, uint32_t V,
                                                         const uint32_t* __restrict__ deg, uint32_t* __restrict__ maxd,
                                                         unsigned long long* __restrict__ edges) {
  // a wave reads 8 vertices' depth rows per step (all in flight), the block
  // combines its waves in LDS, then one atomic pair per lane per block
  constexpr int U = 8;
  __shared__ uint32_t s_md[kNT / kWave][kWave];
  __shared__ unsigned long long s_ed[kNT / kWave][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t gw = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  const uint64_t nw = uint64_t(gridDim.x) * blockDim.x / kWave;
  uint32_t md = 0;
  unsigned long long ed = 0;
  for (uint64_t v0 = gw * U; v0 < V; v0 += nw * U) {
    uint32_t d[U], dg[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const uint64_t v = v0 + k;
      d[k] = v < V ? depth[v * 64 + lane] : kUnreached;
      dg[k] = v < V ? deg[v] : 0u;
    }
#pragma unroll
    for (int k = 0; k < U; ++k)
      if (d[k] != kUnreached) {
        md = max(md, d[k]);
        ed += dg[k];
      }
  }
  s_md[threadIdx.x / kWave][lane] = md;
  s_ed[threadIdx.x / kWave][lane] = ed;
  __syncthreads();
  if (threadIdx.x < kWave) {
    for (int w = 1; w < kNT / kWave; ++w) {
      md = max(md, s_md[w][lane]);
      ed += s_ed[w][lane];
    }
    if (md) atomicMax(&maxd[lane], md);
    if (ed) atomicAdd(&edges[lane], ed);
  }
}

int check_roots(const eg_graph* g, const uint32_t* h_roots, int n) {
  if (n < 1 || n > 64 || !h_roots) return fail(EG_ERR_ARG, "need 1..64 roots");
  for (int r = 0; r < n; ++r)
    if (h_roots[r] >= g->V) return fail(EG_ERR_ARG, "root %u >= vertices_num", h_roots[r]);
  return EG_OK;
}

int start_multi(eg_graph* g, const uint32_t* h_roots, int n, MultiState* st, uint32_t* parent, uint32_t* depth,
                float* dist, hipStream_t s) {
  EG_TRY(multi_state(g, st));
  const uint64_t V = g->V;
  EG_TRY(fill_t<uint64_t>(st->vis, 3 * V, 0ull, s));  // vis, fr, nx
  EG_TRY(fill_t<unsigned long long>(reinterpret_cast<unsigned long long*>(st->cnt), 3 * sizeof(MCount) / 8, 0ull, s));
  EG_TRY(fill_t<uint32_t>(st->steps, 2, 0u, s));  // steps, base
  if (parent) EG_TRY(fill_t<uint32_t>(parent, V * 64, g->V, s));
  if (depth) EG_TRY(fill_t<uint32_t>(depth, V * 64, kUnreached, s));
  if (dist) EG_TRY(fill_t<float>(dist, V * 64, kInfDist, s));
  EG_TRY(g->e_multi_roots.resize(64 * 4 + 64 * 4 + 64 * 8 + 256));
  EG_HIP(hipMemcpyAsync(g->e_multi_roots.ptr, h_roots, size_t(n) * 4, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_multi_roots, dim3(1), dim3(64), 0, s, (const uint32_t*)g->e_multi_roots.as<uint32_t>(), n,
                     st->item_edges, (const uint32_t*)g->out_deg.as<uint32_t>(), parent ? st->vis : nullptr, st->fr, st->items[0],
                     st->cnt, parent, depth, dist);
  EG_HIP(hipGetLastError());
  return EG_OK;
}

// Runs steps until a step has an empty frontier (push, pull, level end per
// step, as a captured step graph).
template <class Args>
int run_steps(eg_graph* g, const Args& a, void (*push)(Args, uint32_t), void (*pull)(Args, uint32_t),
              void (*end)(Args, uint32_t), uint32_t* h_steps, hipStream_t s) {
  const StepLaunch<Args> ls[3] = {{push, 8192},
                                  {pull, grid_for(g->tv_nitems, Args::kPullItemsPerWave * (kNT / kWave), 65536)},
                                  {end, grid_for(ceil_div(g->V, kWave), kNT / kWave, 1024)}};
  const unsigned long long* const nv[3] = {&a.st.cnt[0].nv, &a.st.cnt[1].nv, &a.st.cnt[2].nv};
  return run_step_graph(g->multi_graph, a, ls, a.st.base, nv, a.st.steps, g->V + 1, h_steps, s);
}

}  // namespace
}  // namespace eg

using namespace eg;

extern "C" {

int eg_bfs_multi(eg_graph* g, const uint32_t* h_roots, int n_roots, int policy, uint32_t* d_parent, uint32_t* d_depth,
                 uint32_t* h_levels, uint64_t* h_edges, uint32_t* h_steps, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!g) return fail(EG_ERR_ARG, "graph is NULL");
  EG_TRY(check_roots(g, h_roots, n_roots));
  if (!d_parent || !d_depth) return fail(EG_ERR_ARG, "d_parent and d_depth (V*64 each) are required");
  if (policy < EG_POLICY_REFERENCE || policy > EG_POLICY_PULL) return fail(EG_ERR_ARG, "unknown policy %d", policy);
  EG_TRY(ensure_tview(g, s));
  MultiState st;
  EG_TRY(start_multi(g, h_roots, n_roots, &st, d_parent, d_depth, nullptr, s));
  BfsArgs a{};  // zeroed padding: the bytes key the captured step graph
  a.V = g->V;
  a.E = g->E;
  a.policy = policy;
  a.dense_div = g->tune.mbfs_div;
  a.lanes = n_roots == 64 ? ~0ull : ((1ull << n_roots) - 1);
  a.out_ptr = g->e_out_ptr.as<uint32_t>();
  a.out_dst = g->e_out_dst.as<uint32_t>();
  a.deg = g->out_deg.as<uint32_t>();
  a.in_ptr = g->tv_ptr.as<uint32_t>();
  a.in_src = g->tv_src.as<uint32_t>();
  a.pull_items = g->tv_items.as<uint64_t>();
  a.n_pull_items = g->tv_nitems;
  a.st = st;
  a.parent = d_parent;
  a.depth = d_depth;
  uint32_t steps = 0;
  prof_begin(s);
  EG_TRY(run_steps<BfsArgs>(g, a, k_mbfs_push, k_mbfs_pull, k_mbfs_end, &steps, s));
  prof_end(PK_BFS, 0, s);
  // per-root levels and traversed edges
  uint32_t* maxd = g->e_multi_roots.as<uint32_t>() + 64;
  unsigned long long* edges = reinterpret_cast<unsigned long long*>(g->e_multi_roots.as<char>() + 512);
  EG_TRY(fill_t<uint32_t>(maxd, 64, 0u, s));
  EG_TRY(fill_t<unsigned long long>(edges, 64, 0ull, s));
  hipLaunchKernelGGL(k_bfs_multi_stats, dim3(grid_for(ceil_div(uint64_t(g->V), uint64_t(8)), kNT / kWave, 512)),
                     dim3(kNT), 0, s, (const uint32_t*)d_depth, g->V, (const uint32_t*)g->out_deg.as<uint32_t>(), maxd, edges);
  EG_HIP(hipGetLastError());
  uint32_t h_maxd[64];
  unsigned long long h_ed[64];
  EG_HIP(hipMemcpyAsync(h_maxd, maxd, sizeof(h_maxd), hipMemcpyDeviceToHost, s));
  EG_HIP(hipMemcpyAsync(h_ed, edges, sizeof(h_ed), hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  for (int r = 0; r < n_roots; ++r) {
    if (h_levels) h_levels[r] = h_maxd[r] + 1;
    if (h_edges) h_edges[r] = h_ed[r];
  }
  if (h_steps) *h_steps = steps;
  return EG_OK;
}

int eg_sssp_multi(eg_graph* g, const uint32_t* h_roots, int n_roots, int policy, float* d_dist, uint32_t* h_steps,
                  void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!g) return fail(EG_ERR_ARG, "graph is NULL");
  if (g->rec != 12) return fail(EG_ERR_STATE, "SSSP needs a graph of EdgeType<float> (12-byte) records");
  EG_TRY(check_roots(g, h_roots, n_roots));
  if (!d_dist) return fail(EG_ERR_ARG, "d_dist (V*64 floats) is required");
  if (policy < EG_POLICY_REFERENCE || policy > EG_POLICY_PULL) return fail(EG_ERR_ARG, "unknown policy %d", policy);
  EG_TRY(ensure_tview(g, s));
  MultiState st;
  EG_TRY(start_multi(g, h_roots, n_roots, &st, nullptr, nullptr, d_dist, s));
  SsspArgs a{};  // zeroed padding: the bytes key the captured step graph
  a.V = g->V;
  a.E = g->E;
  a.policy = policy;
  a.dense_div = g->tune.msssp_div;
  a.lanes = n_roots == 64 ? ~0ull : ((1ull << n_roots) - 1);
  a.out_ptr = g->e_out_ptr.as<uint32_t>();
  a.out_dst = g->e_out_dst.as<uint32_t>();
  a.out_w = g->e_out_w.as<float>();
  a.deg = g->out_deg.as<uint32_t>();
  a.in_ptr = g->tv_ptr.as<uint32_t>();
  a.in_src = g->tv_src.as<uint32_t>();
  a.in_w = g->tv_w.as<float>();
  a.pull_items = g->tv_items.as<uint64_t>();
  a.n_pull_items = g->tv_nitems;
  a.st = st;
  a.dist = d_dist;
  uint32_t steps = 0;
  EG_TRY(run_steps<SsspArgs>(g, a, k_msssp_push, k_msssp_pull, k_msssp_end, &steps, s));
  if (h_steps) *h_steps = steps;
  return EG_OK;
}

}  // extern "C"
