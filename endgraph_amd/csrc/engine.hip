// Gemini's edge-processing engine (core/graph.hpp:1376-1817) on one GPU,
// specialised per toolkit signal/slot pair (toolkits/*.cpp): the
// std::function lambdas of the reference cannot cross a C ABI, so each
// process_edges / process_vertices pair used by a toolkit is a kernel here.
//
// Engine view (built once per graph build by eg_engine_prepare):
//   pull side  the CSC lists of all sub-partitions back to back as a source
//              array (4 B/edge), a row-start bit per edge and the destination
//              of each row (a row = one compressed-index entry).  Dense-mode
//              kernels walk edges 64 at a time per wave and reduce rows with
//              a segmented wave scan; only rows cut by a wave boundary (or
//              present in several sub-partitions) use atomics.
//   push side  one CSR merged over the sub-partitions (u32 offsets).
// Mode switch (core/graph.hpp:1483-1498, R3/R4): iterative => dense (pull);
// rooted => sparse (push) under the reference policy; the auto policy and
// undirected programs use sparse iff sum of out-degrees of the active set
// < E/20.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <type_traits>

#include "graph_state.hpp"
#include "rowreduce.hpp"
#include "stepgraph.hpp"

namespace eg {
namespace {

constexpr int kNT = 256;

__device__ __forceinline__ bool bm_get(const uint64_t* bm, uint32_t v) { return (bm[v >> 6] >> (v & 63)) & 1ull; }
__device__ __forceinline__ void bm_set(uint64_t* bm, uint32_t v) {
  atomicOr((unsigned long long*)&bm[v >> 6], 1ull << (v & 63));
}

unsigned grid_for(uint64_t n, uint64_t per_block = kNT, uint64_t cap = 65536) {
  uint64_t b = ceil_div(n, per_block);
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return unsigned(b);
}

template <class T>
__global__ void k_fill(T* __restrict__ p, uint64_t n, T v) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}
__global__ void k_iota(uint32_t* __restrict__ p, uint32_t n) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = uint32_t(i);
}

template <class T>
int fill(T* p, uint64_t n, T v, hipStream_t s) {
  if (!n) return EG_OK;
  hipLaunchKernelGGL((k_fill<T>), dim3(grid_for(n, kNT, 8192)), dim3(kNT), 0, s, p, n, v);
  EG_HIP(hipGetLastError());
  return EG_OK;
}

// ------------------------------------------------------------- prepare --
// SoA source (and weight) of every CSC edge, and the row-start bit per edge
// (a row = one compressed-index entry: a destination within a sub-partition).
// Here a row starts where the destination changes; the starts of the
// sub-partitions, where it may not, are added by k_q_heads.
// A wave takes kPrepU consecutive words per step, their loads issued before
// any is used (one word per step left a wave with a single load in flight:
// 1.9 TB/s on the twitter-shaped graph's 11.7 GB list).
constexpr int kPrepU = 4;
template <class R>
__global__ __launch_bounds__(kNT) void k_prep_csc(const R* __restrict__ lst, uint64_t E, uint32_t* __restrict__ src,
                                                  float* __restrict__ w, uint64_t* __restrict__ rowflag) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t nwords = ceil_div(E, 64);
  const uint64_t gw = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  const uint64_t nw = uint64_t(gridDim.x) * blockDim.x / kWave;
  for (uint64_t w0 = gw * kPrepU; w0 < nwords; w0 += nw * kPrepU) {
    R r[kPrepU], p[kPrepU];
#pragma unroll
    for (int k = 0; k < kPrepU; ++k) {
      const uint64_t e = (w0 + k) * 64 + lane;
      const uint64_t ec = e < E ? e : 0;  // past the end: record 0, unused
      r[k] = lst[ec];
      p[k] = lst[ec > 0 ? ec - 1 : 0];
    }
#pragma unroll
    for (int k = 0; k < kPrepU; ++k) {
      const uint64_t e = (w0 + k) * 64 + lane;
      bool head = false;
      if (e < E) {
        src[e] = r[k].src;
        if constexpr (sizeof(R) == 12) w[e] = reinterpret_cast<const Rec12&>(r[k]).w;
        head = e == 0 || p[k].dst != r[k].dst;
      }
      const uint64_t m = __ballot(head);
      if (lane == 0 && w0 + k < nwords) rowflag[w0 + k] = m;
    }
  }
}

// The first edge of every non-empty sub-partition starts a row (the CSC is
// sorted by sub-partition, then destination).
__global__ void k_q_heads(uint64_t* __restrict__ rowflag, const uint64_t* __restrict__ q_off, int Q, uint64_t E) {
  for (int q = int(blockIdx.x * blockDim.x + threadIdx.x); q < Q; q += int(gridDim.x * blockDim.x)) {
    const uint64_t e = q_off[q];
    if (e < E) atomicOr(reinterpret_cast<unsigned long long*>(rowflag + (e >> 6)), 1ull << (e & 63));
  }
}

__global__ void k_popc_words(const uint64_t* __restrict__ words, uint64_t n, uint32_t* __restrict__ cnt) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    cnt[i] = uint32_t(__popcll(words[i]));
}

// Row r starts at edge e: row_start[r] = e, row_dst[r] = dst of e.
template <class R>
__global__ __launch_bounds__(kNT) void k_prep_rows(const R* __restrict__ lst, uint64_t E,
                                                   const uint64_t* __restrict__ rowflag,
                                                   const uint32_t* __restrict__ wordrank,
                                                   uint32_t* __restrict__ row_start, uint32_t* __restrict__ row_dst) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t nwords = ceil_div(E, 64);
  const uint64_t gw = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  const uint64_t nw = uint64_t(gridDim.x) * blockDim.x / kWave;
  for (uint64_t w0 = gw * kPrepU; w0 < nwords; w0 += nw * kPrepU) {
    uint64_t f[kPrepU];
    uint32_t base[kPrepU], d[kPrepU];
#pragma unroll
    for (int k = 0; k < kPrepU; ++k) {
      const bool ok = w0 + k < nwords;
      f[k] = ok ? rowflag[w0 + k] : 0ull;
      base[k] = ok ? wordrank[w0 + k] : 0u;
      const uint64_t e = (w0 + k) * 64 + lane;
      d[k] = ((f[k] >> lane) & 1ull) ? lst[e].dst : 0u;  // a head bit is set only below E
    }
#pragma unroll
    for (int k = 0; k < kPrepU; ++k)
      if ((f[k] >> lane) & 1ull) {
        const uint64_t e = (w0 + k) * 64 + lane;
        const uint32_t r = base[k] + uint32_t(__popcll(f[k] & ((1ull << lane) - 1)));
        if (row_start) row_start[r] = uint32_t(e);
        row_dst[r] = d[k];
      }
  }
}

// Merged CSR (every source's out-edges from sub-partition 0, then 1, ...)
// in two steps.  k_prep_base: per (sub-partition q, source u) the write base
// out_ptr[u] + (u's edges in sub-partitions before q) - index[q][u], modulo
// 2^32 (a position is below 2^32), so that edge i of sub-partition q (local
// index li) goes to base[q][u] + li.  k_prep_csr: one read of that base per
// edge.
__global__ void k_prep_base(const uint64_t* __restrict__ index, uint32_t V, int Q,
                            const uint32_t* __restrict__ out_ptr, uint32_t* __restrict__ base) {
  for (uint64_t u = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; u < V; u += uint64_t(gridDim.x) * blockDim.x) {
    uint32_t at = out_ptr[u];
    for (int q = 0; q < Q; ++q) {
      const uint64_t* ix = index + uint64_t(q) * (V + 1);
      const uint64_t lo = ix[u], hi = ix[u + 1];
      base[uint64_t(q) * V + u] = at - uint32_t(lo);
      at += uint32_t(hi - lo);
    }
  }
}

template <class R>
__global__ void k_prep_csr(const R* __restrict__ lst, uint64_t E, uint32_t V, int Q, const uint64_t* __restrict__ q_off,
                           const uint32_t* __restrict__ base, uint32_t* __restrict__ dst, float* __restrict__ w) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < E; i += uint64_t(gridDim.x) * blockDim.x) {
    int q = 0;
    while (q + 1 < Q && q_off[q + 1] <= i) ++q;
    const R r = lst[i];
    const uint32_t pos = base[uint64_t(q) * V + r.src] + uint32_t(i - q_off[q]);
    dst[pos] = r.dst;
    if constexpr (sizeof(R) == 12) w[pos] = reinterpret_cast<const Rec12&>(r).w;
  }
}

__global__ void k_u32_from_u64(const uint64_t* __restrict__ a, uint32_t* __restrict__ b, uint64_t n) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    b[i] = uint32_t(a[i]);
}

template <class R>
__global__ void k_soa_csr(const R* __restrict__ lst, uint64_t E, uint32_t* __restrict__ dst, float* __restrict__ w) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < E; i += uint64_t(gridDim.x) * blockDim.x) {
    const R r = lst[i];
    dst[i] = r.dst;
    if constexpr (sizeof(R) == 12) w[i] = reinterpret_cast<const Rec12&>(r).w;
  }
}

// Single-root BFS / SSSP push: vertices of out-degree above the graph's hub
// threshold (tune.hub, 4096 by default) are cut into work items of
// tune.hub_chunk out-edges, listed once per graph.
__global__ void k_hub_item_count(const uint32_t* __restrict__ deg, uint32_t V, uint32_t hub, uint32_t chunk,
                                 uint32_t* __restrict__ cnt) {
  for (uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += uint64_t(gridDim.x) * blockDim.x)
    cnt[v] = deg[v] > hub ? uint32_t(ceil_div(deg[v], chunk)) : 0u;
}

__global__ void k_hub_items(const uint32_t* __restrict__ off, uint32_t V, uint64_t* __restrict__ items) {
  for (uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += uint64_t(gridDim.x) * blockDim.x)
    for (uint32_t j = 0; j < off[v + 1] - off[v]; ++j) items[off[v] + j] = (v << 32) | j;
}

// out-degree in one byte, saturated at 255 (the exact value is read only
// for the saturated ones)
__global__ void k_deg8(const uint32_t* __restrict__ deg, uint32_t V, uint8_t* __restrict__ deg8) {
  for (uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += uint64_t(gridDim.x) * blockDim.x)
    deg8[v] = uint8_t(deg[v] < 255u ? deg[v] : 255u);
}

template <class R>
int prepare_impl(eg_graph* g, hipStream_t s) {
  const uint64_t E = g->E;
  const uint32_t V = g->V;
  const int Q = g->Q;
  const uint64_t nwords = ceil_div(E, 64);
  const uint64_t words_v = (V >> 6) + 1;
  EG_TRY(g->e_csc_src.resize(std::max<uint64_t>(E, 1) * 4));
  EG_TRY(g->e_rowflag.resize((nwords + 1) * 8));
  EG_TRY(g->e_wordrank.resize((nwords + 2) * 4));
  EG_TRY(g->e_out_ptr.resize((uint64_t(V) + 1) * 4));
  EG_TRY(g->e_out_dst.resize(std::max<uint64_t>(E, 1) * 4));
  if (sizeof(R) == 12) {
    EG_TRY(g->e_csc_w.resize(std::max<uint64_t>(E, 1) * 4));
    EG_TRY(g->e_out_w.resize(std::max<uint64_t>(E, 1) * 4));
  }
  EG_TRY(g->e_scratch.resize(uint64_t(V) * 4 + 4096));
  for (auto& b : g->e_bm) EG_TRY(b.resize(words_v * 8));
  uint64_t scan_n = std::max<uint64_t>(scan_tmp_elems(nwords + 1), scan_tmp_elems(uint64_t(V) + 1));
  if (g->scan_tmp.bytes < scan_n * 8) EG_TRY(g->scan_tmp.resize(scan_n * 8));

  // Pull side: the sub-partitions' CSC lists as built, one row per
  // (sub-partition, destination).  Keeping the sub-partitions apart keeps
  // each row's gathers inside one source range (better cache reuse than one
  // merged row per destination, measured); rows of one destination in
  // several sub-partitions combine with atomics.
  const R* csc = g->csc_list.as<R>();
  const R* csr = g->csr_list.as<R>();
  uint32_t* wordrank = g->e_wordrank.as<uint32_t>();
  uint64_t* rowflag = g->e_rowflag.as<uint64_t>();
  if (E) {
    hipLaunchKernelGGL((k_prep_csc<R>), dim3(grid_for(ceil_div(nwords, kPrepU), kNT / kWave, 65536)), dim3(kNT), 0, s,
                       csc, E,
                       g->e_csc_src.as<uint32_t>(), g->e_csc_w.as<float>(), rowflag);
    EG_HIP(hipGetLastError());
    if (Q > 1) {
      hipLaunchKernelGGL(k_q_heads, dim3(1), dim3(kNT), 0, s, rowflag, (const uint64_t*)g->dmeta()->q_off_csc, Q, E);
      EG_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_popc_words, dim3(grid_for(nwords, kNT, 65536)), dim3(kNT), 0, s, rowflag, nwords, wordrank);
    EG_HIP(hipGetLastError());
  }
  EG_TRY(fill<uint32_t>(wordrank + nwords, 1, 0u, s));
  EG_TRY(exclusive_scan_u32(wordrank, wordrank, nwords + 1, g->scan_tmp.as<uint32_t>(), s));
  uint32_t R_h = 0;
  EG_HIP(hipMemcpyAsync(&R_h, wordrank + nwords, 4, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  g->R = R_h;
  EG_TRY(g->e_row_start.resize((uint64_t(R_h) + 1) * 4));
  EG_TRY(g->e_row_dst.resize((uint64_t(R_h) + 1) * 4));
  if (E) {
    hipLaunchKernelGGL((k_prep_rows<R>), dim3(grid_for(ceil_div(nwords, kPrepU), kNT / kWave, 65536)), dim3(kNT), 0,
                       s, csc, E, rowflag,
                       (const uint32_t*)wordrank, g->e_row_start.as<uint32_t>(), g->e_row_dst.as<uint32_t>());
    EG_HIP(hipGetLastError());
  }
  EG_TRY(fill<uint32_t>(g->e_row_start.as<uint32_t>() + R_h, 1, uint32_t(E), s));

  uint32_t* out_ptr = g->e_out_ptr.as<uint32_t>();
  if (Q == 1) {
    hipLaunchKernelGGL(k_u32_from_u64, dim3(grid_for(uint64_t(V) + 1, kNT, 65536)), dim3(kNT), 0, s,
                       (const uint64_t*)g->csr_index.as<uint64_t>(), out_ptr, uint64_t(V) + 1);
    EG_HIP(hipGetLastError());
    if (E) {
      hipLaunchKernelGGL((k_soa_csr<R>), dim3(grid_for(E, kNT, 65536)), dim3(kNT), 0, s, csr, E,
                         g->e_out_dst.as<uint32_t>(), g->e_out_w.as<float>());
      EG_HIP(hipGetLastError());
    }
  } else {
    // out_ptr = exclusive prefix of the out-degree
    EG_HIP(hipMemcpyAsync(out_ptr, g->out_deg.ptr, uint64_t(V) * 4, hipMemcpyDeviceToDevice, s));
    EG_TRY(fill<uint32_t>(out_ptr + V, 1, 0u, s));
    EG_TRY(exclusive_scan_u32(out_ptr, out_ptr, uint64_t(V) + 1, g->scan_tmp.as<uint32_t>(), s));
    if (E) {
      // the write bases in the build's sort temporary (free after the build)
      EG_TRY(g->t3.resize(std::max<uint64_t>(g->t3.bytes, uint64_t(Q) * V * 4)));
      uint32_t* base = g->t3.as<uint32_t>();
      hipLaunchKernelGGL(k_prep_base, dim3(grid_for(V, kNT, 65536)), dim3(kNT), 0, s,
                         (const uint64_t*)g->csr_index.as<uint64_t>(), V, Q, (const uint32_t*)out_ptr, base);
      EG_HIP(hipGetLastError());
      hipLaunchKernelGGL((k_prep_csr<R>), dim3(grid_for(E, kNT, 65536)), dim3(kNT), 0, s, csr, E, V, Q,
                         g->dmeta()->q_off_csr, (const uint32_t*)base, g->e_out_dst.as<uint32_t>(),
                         g->e_out_w.as<float>());
      EG_HIP(hipGetLastError());
    }
  }
  EG_TRY(g->e_deg8.resize(V ? V : 1));
  hipLaunchKernelGGL(k_deg8, dim3(grid_for(V, kNT, 65536)), dim3(kNT), 0, s, (const uint32_t*)g->out_deg.as<uint32_t>(),
                     V, g->e_deg8.as<uint8_t>());
  EG_HIP(hipGetLastError());
  g->e_hub_for[0] = g->e_hub_for[1] = 0;
  EG_TRY(ensure_hub_items(g, s));
  return EG_OK;
}

}  // namespace

// The hubs' push items for the graph's tune: a count per vertex, its
// exclusive prefix (in the engine scratch, V * 4 + 4096 bytes or more), the
// items; rebuilt only when the tune changed.
int ensure_hub_items(eg_graph* g, hipStream_t s) {
  const uint32_t V = g->V, hub = g->tune.hub, chunk = g->tune.hub_chunk;
  if (g->e_hub_for[0] == hub && g->e_hub_for[1] == chunk) return EG_OK;
  if (chunk == 0) return fail(EG_ERR_ARG, "hub chunk size must be positive");
  if (g->e_scratch.bytes < (uint64_t(V) + 1) * 4) EG_TRY(g->e_scratch.resize((uint64_t(V) + 1) * 4 + 4096));
  uint32_t* o = g->e_scratch.as<uint32_t>();
  hipLaunchKernelGGL(k_hub_item_count, dim3(grid_for(V, kNT, 65536)), dim3(kNT), 0, s,
                     (const uint32_t*)g->out_deg.as<uint32_t>(), V, hub, chunk, o);
  EG_HIP(hipGetLastError());
  EG_TRY(fill<uint32_t>(o + V, 1, 0u, s));
  EG_TRY(exclusive_scan_u32(o, o, uint64_t(V) + 1, g->scan_tmp.as<uint32_t>(), s));
  uint32_t n = 0;
  EG_HIP(hipMemcpyAsync(&n, o + V, 4, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  EG_TRY(g->e_hub_items.resize(std::max<uint64_t>(n, 1) * 8));
  g->e_nhub_items = n;
  hipLaunchKernelGGL(k_hub_items, dim3(grid_for(V, kNT, 65536)), dim3(kNT), 0, s, (const uint32_t*)o, V,
                     g->e_hub_items.as<uint64_t>());
  EG_HIP(hipGetLastError());
  g->e_hub_for[0] = hub;
  g->e_hub_for[1] = chunk;
  return EG_OK;
}

namespace {

int ensure_engine(eg_graph* g, hipStream_t s) {
  if (!g) return fail(EG_ERR_ARG, "graph is NULL");
  if (!g->built) return fail(EG_ERR_STATE, "graph has not been built");
  if (g->eng_ready) return EG_OK;
  if (g->E >= (uint64_t(1) << 32)) return fail(EG_ERR_ARG, "the engine's edge offsets are 32-bit: need E < 2^32");
  auto t0 = std::chrono::steady_clock::now();
  int rc = g->rec == 8 ? prepare_impl<Rec8>(g, s) : prepare_impl<Rec12>(g, s);
  if (rc != EG_OK) return rc;
  g->eng_prepare_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  g->eng_ready = true;
  return EG_OK;
}

// ------------------------------------------ degree-ordered PageRank view --
// The pull gathers x[src] once per in-edge, so a source is read out-degree
// times.  Renumbering vertices by descending out-degree (ties by id) packs
// the most-read values into the fewest cache lines: on a twitter-shaped graph
// the top 2^20 sources hold 85% of the edges, 4 MiB of FP32 values, where
// the original ids spread them over one cache line each.  The view keeps one
// row per destination, ids renumbered, in one of two forms:
//   relabelled CSC (P = S = 1)  the built CSC's rows with every id replaced
//       by its new id: one gather per edge to build, the CSC's terms in the
//       CSC's order;
//   sorted rows  the renumbered graph's rows with sources ascending by new
//       id (a row copy and one sort by destination to build): the lanes of a
//       gather then share cache lines, and the vector memory unit takes one
//       line per cycle (twitter-shaped FP32C pull 8.2 -> 7.5 ms, R-MAT 24
//       P = 8 1.67 -> 1.62 ms; the build of the view 17.5 -> 50 ms and 13 ->
//       9 ms against the relabelled CSC, resp. the relabelled traversal view).
// The default is the relabelled CSC with one sub-partition and sorted rows
// otherwise (the CSC's rows there are one per sub-partition and destination,
// which would need an atomic add per row); EG_PR_SORTED asks for sorted rows.
// Both hold the CSC rows' terms; a different order moves an FP64 row sum by
// a few ulps and a compensated FP32 row sum not at all at the FP32 rounding.
struct KeyDstField {
  static constexpr int kField = 1;
  template <class R>
  __device__ __forceinline__ uint64_t operator()(const R& r) const { return r.dst; }
  __device__ __forceinline__ uint64_t of_field(uint32_t k) const { return k; }
};

// record v = (v, E - deg[v]): ascending key = descending degree
__global__ void k_deg_keys(const uint32_t* __restrict__ deg, uint32_t V, uint64_t E, Rec8* __restrict__ rec) {
  for (uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += uint64_t(gridDim.x) * blockDim.x)
    rec[v] = Rec8{uint32_t(v), uint32_t(E - deg[v])};
}

__global__ void k_view_perm(const Rec8* __restrict__ sorted, uint32_t V, const uint32_t* __restrict__ deg,
                            uint32_t* __restrict__ perm, uint32_t* __restrict__ inv, uint32_t* __restrict__ vdeg) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < V; i += uint64_t(gridDim.x) * blockDim.x) {
    const uint32_t u = sorted[i].src;
    inv[i] = u;
    perm[u] = uint32_t(i);
    vdeg[i] = deg[u];
  }
}

// out[i] = perm[in[i]], four ids per thread step with 16-byte accesses
__global__ void k_relabel(const uint32_t* __restrict__ in, uint64_t n, const uint32_t* __restrict__ perm,
                          uint32_t* __restrict__ out) {
  const uint64_t n4 = n / 4;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n4; i += uint64_t(gridDim.x) * blockDim.x) {
    const uint4 v = reinterpret_cast<const uint4*>(in)[i];
    reinterpret_cast<uint4*>(out)[i] = uint4{perm[v.x], perm[v.y], perm[v.z], perm[v.w]};
  }
  const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < n - n4 * 4) out[n4 * 4 + t] = perm[in[n4 * 4 + t]];
}

int sorted_rows(eg_graph* g, hipStream_t s);

// The view in form kViewCsc (P = S = 1 only: the built CSC's rows, every id
// relabelled) or kViewSorted (rows sorted by new source id).
enum ViewForm { kViewNone = -1, kViewCsc = 0, kViewSorted = 1 };
constexpr int kHotNT = 1024;  // threads per block of the FP32C pull with LDS-staged hot values

int view_impl(eg_graph* g, int form, hipStream_t s) {
  const uint64_t E = g->E;
  const uint32_t V = g->V;
  EG_TRY(g->pv_perm.resize(uint64_t(V) * 4));
  EG_TRY(g->pv_inv.resize(uint64_t(V) * 4));
  EG_TRY(g->pv_deg.resize(uint64_t(V) * 4));
  EG_TRY(g->pv_src.resize((std::max<uint64_t>(E, 1) + 4) * 4));
  // new ids: a stable sort of (v, E - deg) by the second field, in the
  // build's sort temporaries (free after the build; growing them only, they
  // keep the E records the traversal view's sort uses)
  const uint64_t vb = std::max<uint64_t>(V, 1) * 8;
  EG_TRY(g->t1.resize(vb));
  EG_TRY(g->t2.resize(vb));
  EG_TRY(g->t3.resize(vb));
  Rec8* ra = g->t1.as<Rec8>();
  Rec8* rb = g->t2.as<Rec8>();
  EG_TRY(g->sortw.resize_segs(1));
  EG_TRY(fill<uint32_t>(g->sortw.err.as<uint32_t>(), 1, 0u, s));
  hipLaunchKernelGGL(k_deg_keys, dim3(grid_for(V, kNT, 65536)), dim3(kNT), 0, s,
                     (const uint32_t*)g->out_deg.as<uint32_t>(), V, E, ra);
  EG_HIP(hipGetLastError());
  EG_TRY(rx_sort((const Rec8*)ra, rb, g->t3.as<Rec8>(), std::vector<uint64_t>{0, V}, KeyDstField{},
                 std::max(1, counting_bits(E)), g->sortw, s));
  uint32_t* perm = g->pv_perm.as<uint32_t>();
  hipLaunchKernelGGL(k_view_perm, dim3(grid_for(V, kNT, 65536)), dim3(kNT), 0, s, (const Rec8*)rb, V,
                     (const uint32_t*)g->out_deg.as<uint32_t>(), perm, g->pv_inv.as<uint32_t>(),
                     g->pv_deg.as<uint32_t>());
  EG_HIP(hipGetLastError());
  if (form == kViewSorted) {
    EG_TRY(sorted_rows(g, s));
  } else {
    if (g->Q != 1) return fail(EG_ERR_STATE, "internal: the relabelled-CSC view needs one sub-partition");
    const uint64_t rows = g->R;
    EG_TRY(g->pv_row_dst.resize((rows + 1) * 4));
    if (E) {
      hipLaunchKernelGGL(k_relabel, dim3(grid_for(ceil_div(E, 4), kNT, 65536)), dim3(kNT), 0, s,
                         (const uint32_t*)g->e_csc_src.as<uint32_t>(), E, (const uint32_t*)perm,
                         g->pv_src.as<uint32_t>());
      EG_HIP(hipGetLastError());
    }
    if (rows) {
      hipLaunchKernelGGL(k_relabel, dim3(grid_for(ceil_div(rows, 4), kNT, 65536)), dim3(kNT), 0, s,
                         (const uint32_t*)g->e_row_dst.as<uint32_t>(), rows, (const uint32_t*)perm,
                         g->pv_row_dst.as<uint32_t>());
      EG_HIP(hipGetLastError());
    }
    g->pv_R = uint32_t(rows);
  }
  uint32_t sort_err = 0;
  EG_HIP(hipMemcpyAsync(&sort_err, g->sortw.err.ptr, 4, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  if (sort_err) return fail(EG_ERR_HIP, "radix sort look-back exceeded its spin limit");
  return EG_OK;
}

// The ordered view of eg_pagerank's layout flags: EG_PR_SORTED the sorted-row
// form, EG_PR_ORDERED (or the FP32C default) the graph's default form.
int view_form_of(const eg_graph* g, int layout) {
  if (layout == EG_PR_SORTED || g->Q > 1) return kViewSorted;
  return kViewCsc;
}

int ensure_view(eg_graph* g, int form, hipStream_t s) {
  EG_TRY(ensure_engine(g, s));
  if (g->pv_ready && g->pv_form == form) return EG_OK;
  g->pv_ready = false;
  auto t0 = std::chrono::steady_clock::now();
  EG_TRY(view_impl(g, form, s));
  g->pv_prepare_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  g->pv_form = form;
  // The FP32C pull's LDS staging: as many hot values as one block's LDS
  // holds (160 KB on gfx950: 40,960 sources), one block per CU (its 1024
  // threads at the kernel's register count leave no room for a second).
  int dev = 0, cus = 0, lds = 0;
  EG_HIP(hipGetDevice(&dev));
  EG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  EG_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  const uint64_t waves = ceil_div(ceil_div(g->E, 64), kRowWordsPerWave);
  g->pv_hot = uint32_t(std::min<uint64_t>(g->V, uint64_t(std::max(lds, 0)) / 4));
  g->pv_hot_grid = unsigned(std::min<uint64_t>(ceil_div(waves, kHotNT / kWave), uint64_t(std::max(cus, 1))));
  if (!g->E) g->pv_hot = 0;
  g->pv_ready = true;
  return EG_OK;
}

__global__ void k_gather_by(const uint32_t* __restrict__ idx, uint32_t V, const double* __restrict__ from,
                            double* __restrict__ to) {
  for (uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += uint64_t(gridDim.x) * blockDim.x)
    to[v] = from[idx[v]];
}

// Compensated (Neumaier) pair for the FP32 pull sums (row reduction: see
// rowreduce.hpp).
struct F2 {
  float s, c;
};
__device__ __forceinline__ F2 f2_add(F2 a, F2 b) {
  const float s = __fadd_rn(a.s, b.s);
  const float bp = __fsub_rn(s, a.s);
  const float err = __fadd_rn(__fsub_rn(a.s, __fsub_rn(s, bp)), __fsub_rn(b.s, bp));
  return F2{s, __fadd_rn(__fadd_rn(a.c, b.c), err)};
}
struct OpF2 {
  __device__ __forceinline__ F2 operator()(F2 a, F2 b) const { return f2_add(a, b); }
};


// --------------------------------------------------------------- PageRank
// dense_signal/dense_slot of toolkits/pagerank.cpp:64-75: next[dst] += sum
// of curr[src] over the in-edges, partials of several sub-partitions added
// with atomics.
__global__ __launch_bounds__(kNT) void k_pr_pull_f64(uint64_t E, const uint32_t* __restrict__ src,
                                                      const uint64_t* __restrict__ rowflag,
                                                      const uint32_t* __restrict__ wordrank,
                                                      const uint32_t* __restrict__ row_dst,
                                                      const double* __restrict__ curr, double* __restrict__ next,
                                                      int all_atomic) {
  row_reduce_wave<double>(
      first_word(), E, src, rowflag, wordrank, 0.0, OpSum{}, [&](uint64_t, uint32_t s) { return curr[s]; },
      [&](uint32_t row, double v, bool shared) {
        double* p = &next[row_dst[row]];
        if (shared || all_atomic)
          atomicAdd(p, v);
        else
          *p = v;
      });
}

// FP32 gathers with compensated row sums; the row total is handed over as
// FP64 (s + c) for the FP64 residual step.
__global__ __launch_bounds__(kNT) void k_pr_pull_f32c(uint64_t E, const uint32_t* __restrict__ src,
                                                       const uint64_t* __restrict__ rowflag,
                                                       const uint32_t* __restrict__ wordrank,
                                                       const uint32_t* __restrict__ row_dst,
                                                       const float* __restrict__ curr, double* __restrict__ next,
                                                       int all_atomic) {
  row_reduce_wave<F2>(
      first_word(), E, src, rowflag, wordrank, F2{0.f, 0.f}, OpF2{}, [&](uint64_t, uint32_t s) { return F2{curr[s], 0.f}; },
      [&](uint32_t row, F2 v, bool shared) {
        double* p = &next[row_dst[row]];
        const double t = double(v.s) + double(v.c);
        if (shared || all_atomic)
          atomicAdd(p, t);
        else
          *p = t;
      });
}

// The same pull on the degree-ordered view, where new ids [0, nhot) are the
// highest out-degrees (on the twitter-shaped graph the 40,960 hottest of
// 41.6 M sources carry about 70% of the edges): their values are staged once
// per block in LDS, and a gather of one of them is an LDS read instead of a
// cache-line request to the L2.  A persistent grid (one kHotNT-thread block
// per CU, as many hot values as the LDS holds) whose waves take the K-word
// ranges in turn; each range is reduced exactly as k_pr_pull_f32c does.
__global__ __launch_bounds__(kHotNT) void k_pr_pull_f32c_hot(uint64_t E, const uint32_t* __restrict__ src,
                                                             const uint64_t* __restrict__ rowflag,
                                                             const uint32_t* __restrict__ wordrank,
                                                             const uint32_t* __restrict__ row_dst,
                                                             const float* __restrict__ curr,
                                                             double* __restrict__ next, int all_atomic,
                                                             uint32_t nhot, uint64_t nranges) {
  extern __shared__ float hot[];
  for (uint32_t i = threadIdx.x; i < nhot; i += kHotNT) hot[i] = curr[i];
  __syncthreads();
  constexpr uint64_t kWpb = kHotNT / kWave;
  for (uint64_t rg = uint64_t(blockIdx.x) * kWpb + threadIdx.x / kWave; rg < nranges; rg += uint64_t(gridDim.x) * kWpb)
    row_reduce_wave<F2, kRowWordsPerWave, NoEdgeValue, true>(
        rg * kRowWordsPerWave, E, src, rowflag, wordrank, F2{0.f, 0.f}, OpF2{},
        [&](uint64_t, uint32_t s) {
          // Both loads in every lane, without branches: hot lanes read
          // curr[0] (one line per wave), the others hot[0].  The value is
          // a sum of the two, one of them +0 (exact): a select of two loads
          // would be folded into one flat load through the selected address.
          const bool h = s < nhot;
          const float a = hot[h ? s : 0u];
          const float b = curr[h ? 0u : s];
          return F2{(h ? a : 0.f) + (h ? 0.f : b), 0.f};
        },
        [&](uint32_t row, F2 v, bool shared) {
          double* p = &next[row_dst[row]];
          const double t = double(v.s) + double(v.c);
          if (shared || all_atomic)
            atomicAdd(p, t);
          else
            *p = t;
        });
}

// process_vertices of toolkits/pagerank.cpp:36-47 (init) and :79-98 (apply).
__global__ __launch_bounds__(kNT) void k_pr_init(const uint32_t* __restrict__ deg, uint32_t V, double* __restrict__ x,
                                                  float* __restrict__ x32) {
  for (uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += uint64_t(gridDim.x) * blockDim.x) {
    double c = 1.0;
    if (deg[v] > 0) c = c / double(deg[v]);
    x[v] = c;
    if (x32) x32[v] = float(c);
  }
}

__global__ __launch_bounds__(kNT) void k_pr_apply(const uint32_t* __restrict__ deg, uint32_t V,
                                                   const double* __restrict__ curr, double* __restrict__ next,
                                                   float* __restrict__ next32, int last, double* __restrict__ delta) {
  __shared__ double wt[kNT / kWave];
  const double d = 0.85;
  double acc = 0;
  for (uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += uint64_t(gridDim.x) * blockDim.x) {
    const double y = __dadd_rn(1.0 - d, __dmul_rn(d, next[v]));
    double x = y;
    if (!last) {
      const uint32_t dg = deg[v];
      if (dg > 0) {
        x = y / double(dg);
        acc += fabs(x - curr[v]) * double(dg);
      } else {
        acc += fabs(x - curr[v]);
      }
    }
    next[v] = x;
    if (next32) next32[v] = float(x);
  }
  acc = wave_reduce(acc, OpSum{});
  if ((threadIdx.x & (kWave - 1)) == 0) wt[threadIdx.x / kWave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int i = 0; i < kNT / kWave; ++i) t += wt[i];
    atomicAdd(delta, t);
  }
}

// ----------------------------------------------------- frontier helpers --
// Frontier list from a bitmap (ascending ids) via per-word counts + scan.
__global__ void k_bm_count(const uint64_t* __restrict__ bm, uint64_t words, uint32_t* __restrict__ cnt) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < words; i += uint64_t(gridDim.x) * blockDim.x)
    cnt[i] = uint32_t(__popcll(bm[i]));
}
__global__ void k_bm_list(const uint64_t* __restrict__ bm, uint64_t words, const uint32_t* __restrict__ off,
                          uint32_t* __restrict__ list) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < words; i += uint64_t(gridDim.x) * blockDim.x) {
    uint64_t w = bm[i];
    uint32_t o = off[i];
    while (w) {
      const int b = __ffsll((unsigned long long)w) - 1;
      list[o++] = uint32_t(i * 64 + b);
      w &= w - 1;
    }
  }
}

__global__ void k_bm_or(uint64_t* __restrict__ dst, const uint64_t* __restrict__ src, uint64_t words) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < words; i += uint64_t(gridDim.x) * blockDim.x)
    dst[i] |= src[i];
}

// Level bookkeeping: visited |= out; depth of new vertices; counts of new
// vertices and of their out-edges (for the next mode decision).
// A wave loads 64 words at once (one per lane: the visited update and the
// count are per lane), then walks the non-empty ones: lane b holds vertex
// 64*word + b, so the depth stores and degree loads of a word are one
// coalesced access each, and empty words cost nothing more.
__global__ __launch_bounds__(kNT) void k_level_end(const uint64_t* __restrict__ out, uint64_t* __restrict__ visited,
                                                    uint64_t words, uint32_t* __restrict__ depth, uint32_t level,
                                                    const uint32_t* __restrict__ deg,
                                                    unsigned long long* __restrict__ counters) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t gw = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  const uint64_t nw = uint64_t(gridDim.x) * blockDim.x / kWave;
  unsigned long long nv = 0, ne = 0;
  for (uint64_t base = gw * kWave; base < words; base += nw * kWave) {
    const uint64_t wi = base + lane;
    const uint64_t mine = wi < words ? out[wi] : 0ull;
    if (mine) {
      if (visited) visited[wi] |= mine;
      nv += __popcll(mine);
    }
    if (!depth && !deg) continue;
    uint64_t nz = __ballot(mine != 0);
    while (nz) {  // wave-uniform
      const int j = __ffsll((unsigned long long)nz) - 1;
      nz &= nz - 1;
      const uint64_t w = shfl_t(mine, j);
      if ((w >> lane) & 1ull) {
        const uint64_t v = (base + j) * 64 + lane;
        if (depth) depth[v] = level;
        if (deg) ne += deg[v];
      }
    }
  }
  // one atomic pair per block (every wave adding its own serialised the
  // counters' cache line)
  __shared__ unsigned long long part[2][kNT / kWave];
  nv = wave_reduce(nv, OpSum{});
  ne = wave_reduce(ne, OpSum{});
  if (lane == 0) {
    part[0][threadIdx.x / kWave] = nv;
    part[1][threadIdx.x / kWave] = ne;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long a = 0, b = 0;
    for (int k = 0; k < kNT / kWave; ++k) {
      a += part[0][k];
      b += part[1][k];
    }
    if (a) atomicAdd(&counters[0], a);
    if (b) atomicAdd(&counters[1], b);
  }
}

// ------------------------------------------------------ sparse (push) ----
// Edge-balanced frontier expansion: block b takes the 2048 consecutive slots
// [b*2048, (b+1)*2048) of the frontier's concatenated out-adjacency
// (foff = exclusive prefix of the frontier degrees, foff[fn] = total); each
// slot finds its frontier vertex by binary search, in an LDS copy of the
// block's window of foff when it fits.  A hub's edges spread over many
// blocks instead of one wave.
constexpr int kExpSlots = kNT * 8;
constexpr int kExpWin = kExpSlots + 2;

__device__ __forceinline__ uint32_t last_le_u32(const uint32_t* a, uint32_t lo, uint32_t hi, uint32_t x) {
  // largest i in [lo, hi) with a[i] <= x (a non-decreasing, a[lo] <= x)
  while (hi - lo > 1) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (a[mid] <= x)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// An Op with `kSlot` takes (list index, vertex, edge, slot) instead of
// (vertex, edge).
template <class Op, class = void>
struct ExpandSlot : std::false_type {};
template <class Op>
struct ExpandSlot<Op, std::void_t<decltype(Op::kSlot)>> : std::true_type {};

template <class Op>
__global__ __launch_bounds__(kNT) void k_expand(const uint32_t* __restrict__ flist, const uint32_t* __restrict__ foff,
                                                uint32_t fn, const uint32_t* __restrict__ out_ptr, Op op) {
  __shared__ uint32_t win[kExpWin];
  __shared__ uint32_t s_lo, s_hi;
  const uint32_t total = foff[fn];
  const uint32_t t0 = blockIdx.x * uint32_t(kExpSlots);
  if (t0 >= total) return;  // block-uniform
  const uint32_t t1 = uint32_t(uint64_t(t0) + kExpSlots < total ? uint64_t(t0) + kExpSlots : total);
  if (threadIdx.x == 0) {
    s_lo = last_le_u32(foff, 0, fn, t0);
    s_hi = last_le_u32(foff, s_lo, fn, t1 - 1);
  }
  __syncthreads();
  const uint32_t lo = s_lo, hi = s_hi;  // vertices holding the block's slots
  const bool in_lds = hi - lo + 2 <= uint32_t(kExpWin);
  if (in_lds)
    for (uint32_t i = threadIdx.x; i < hi - lo + 2; i += kNT) win[i] = foff[lo + i];
  __syncthreads();
  uint32_t u[8], e[8], ix[8];
  bool ok[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const uint64_t t64 = uint64_t(t0) + uint32_t(k) * kNT + threadIdx.x;
    ok[k] = t64 < t1;
    const uint32_t t = uint32_t(t64);
    uint32_t i, start;
    if (in_lds) {
      i = lo + last_le_u32(win, 0, hi - lo + 1, ok[k] ? t : t0);
      start = win[i - lo];
    } else {
      i = last_le_u32(foff, lo, hi + 1, ok[k] ? t : t0);
      start = foff[i];
    }
    ix[k] = i;
    u[k] = flist[i];
    e[k] = out_ptr[u[k]] + ((ok[k] ? t : t0) - start);
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (!ok[k]) continue;
    if constexpr (ExpandSlot<Op>::value)
      op(ix[k], u[k], e[k], t0 + uint32_t(k) * kNT + threadIdx.x);
    else
      op(u[k], e[k]);
  }
}

// Sorted-row form: the renumbered graph's rows, sources ascending by new id
// inside each row, so the lanes of one gather touch neighbouring lines (the
// vector memory unit takes one cache line per cycle).  The merged CSR's rows
// are copied in new-id order as (new source, destination) records (the slot
// of an out-edge is its new source's offset plus its rank), then sorted
// stably by destination.
struct RowCopyOp {
  static constexpr bool kSlot = true;
  const uint32_t* out_dst;
  Rec8* rec;
  __device__ __forceinline__ void operator()(uint32_t i, uint32_t, uint32_t e, uint32_t t) const {
    rec[t] = Rec8{i, out_dst[e]};
  }
};

__global__ void k_copy_u32_v(const uint32_t* __restrict__ a, uint32_t* __restrict__ b, uint64_t n) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    b[i] = a[i];
}

// Row starts of the sorted list from its destination array: bit e of the
// row bits is set where the destination changes; cnt = heads per word.
__global__ __launch_bounds__(kNT) void k_view_flags(const uint32_t* __restrict__ dst, uint64_t E,
                                                     uint64_t* __restrict__ rowflag, uint32_t* __restrict__ cnt) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t nwords = ceil_div(E, 64);
  const uint64_t nw = uint64_t(gridDim.x) * blockDim.x / kWave;
  for (uint64_t w0 = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave * kPrepU; w0 < nwords;
       w0 += nw * kPrepU) {
    uint32_t cur[kPrepU], prev[kPrepU];
#pragma unroll
    for (int k = 0; k < kPrepU; ++k) {
      const uint64_t e = (w0 + k) * 64 + lane;
      const uint64_t ec = e < E ? e : 0;
      cur[k] = dst[ec];
      prev[k] = dst[ec > 0 ? ec - 1 : 0];
    }
#pragma unroll
    for (int k = 0; k < kPrepU; ++k) {
      const uint64_t e = (w0 + k) * 64 + lane;
      const uint64_t m = __ballot(e < E && (e == 0 || prev[k] != cur[k]));
      if (lane == 0 && w0 + k < nwords) {
        rowflag[w0 + k] = m;
        cnt[w0 + k] = uint32_t(__popcll(m));
      }
    }
  }
}

__global__ void k_view_row_dst(const uint32_t* __restrict__ dst, uint64_t E, const uint64_t* __restrict__ rowflag,
                               const uint32_t* __restrict__ wordrank, const uint32_t* __restrict__ perm,
                               uint32_t* __restrict__ row_dst) {
  for (uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < E; e += uint64_t(gridDim.x) * blockDim.x) {
    const uint64_t f = rowflag[e >> 6];
    if ((f >> (e & 63)) & 1ull)
      row_dst[wordrank[e >> 6] + uint32_t(__popcll(f & ((1ull << (e & 63)) - 1)))] = perm[dst[e]];
  }
}

// Sorted-row form from the new ids (perm, inv, pv_deg made by view_impl).
int sorted_rows(eg_graph* g, hipStream_t s) {
  const uint64_t E = g->E;
  const uint32_t V = g->V;
  if (!E) return EG_OK;
  const uint64_t nwords = ceil_div(E, 64);
  uint32_t* noff = g->e_scratch.as<uint32_t>();  // (V + 1) * 4 bytes or more: out-degree prefix in new order
  hipLaunchKernelGGL(k_copy_u32_v, dim3(grid_for(V, kNT, 65536)), dim3(kNT), 0, s,
                     (const uint32_t*)g->pv_deg.as<uint32_t>(), noff, uint64_t(V));
  EG_HIP(hipGetLastError());
  EG_TRY(fill<uint32_t>(noff + V, 1, 0u, s));
  EG_TRY(exclusive_scan_u32(noff, noff, uint64_t(V) + 1, g->scan_tmp.as<uint32_t>(), s));
  EG_TRY(g->t1.resize(E * 8));
  EG_TRY(g->t2.resize(E * 4));
  EG_TRY(g->t3.resize(rx_soa_bytes<Rec8>(E)));
  Rec8* rec = g->t1.as<Rec8>();
  uint32_t* sdst = g->t2.as<uint32_t>();  // destinations of the sorted rows
  hipLaunchKernelGGL((k_expand<RowCopyOp>), dim3(unsigned(ceil_div(E, kExpSlots))), dim3(kNT), 0, s,
                     (const uint32_t*)g->pv_inv.as<uint32_t>(), (const uint32_t*)noff, V,
                     (const uint32_t*)g->e_out_ptr.as<uint32_t>(), RowCopyOp{g->e_out_dst.as<uint32_t>(), rec});
  EG_HIP(hipGetLastError());
  EG_TRY(g->sortw.resize_segs(1));
  EG_TRY(fill<uint32_t>(g->sortw.err.as<uint32_t>(), 1, 0u, s));
  // sorted by destination straight into field arrays: the sources are the
  // view's, the destinations give the row starts
  RxSoa<Rec8> out;
  out.s = g->pv_src.as<uint32_t>();
  out.d = sdst;
  EG_TRY(rx_sort_to_fields((const Rec8*)rec, out, rx_soa_in<Rec8>(g->t3.ptr, E), std::vector<uint64_t>{0, E},
                           KeyDstField{}, g->vbits, g->sortw, s));
  EG_TRY(g->pv_rowflag.resize((nwords + 1) * 8));
  EG_TRY(g->pv_wordrank.resize((nwords + 2) * 4));
  uint32_t* wr = g->pv_wordrank.as<uint32_t>();
  hipLaunchKernelGGL(k_view_flags, dim3(grid_for(ceil_div(nwords, kPrepU), kNT / kWave, 65536)), dim3(kNT), 0, s, (const uint32_t*)sdst,
                     E, g->pv_rowflag.as<uint64_t>(), wr);
  EG_HIP(hipGetLastError());
  EG_TRY(fill<uint32_t>(wr + nwords, 1, 0u, s));
  EG_TRY(exclusive_scan_u32(wr, wr, nwords + 1, g->scan_tmp.as<uint32_t>(), s));
  uint32_t R = 0;
  EG_HIP(hipMemcpyAsync(&R, wr + nwords, 4, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  EG_TRY(g->pv_row_dst.resize((uint64_t(R) + 1) * 4));
  hipLaunchKernelGGL(k_view_row_dst, dim3(grid_for(E, kNT, 65536)), dim3(kNT), 0, s, (const uint32_t*)sdst, E,
                     (const uint64_t*)g->pv_rowflag.as<uint64_t>(), (const uint32_t*)wr,
                     (const uint32_t*)g->pv_perm.as<uint32_t>(), g->pv_row_dst.as<uint32_t>());
  EG_HIP(hipGetLastError());
  g->pv_R = R;
  return EG_OK;
}

// Sparse slot of toolkits/bfs.cpp:52-58 with R12: an unvisited out-neighbour
// takes the smallest frontier id as parent (atomicMin; visited is not
// modified during a level, so the rule is order-independent).
struct BfsPushOp {
  const uint32_t* out_dst;
  const uint64_t* visited;
  uint32_t* parent;
  uint64_t* next;
  __device__ __forceinline__ void operator()(uint32_t u, uint32_t e) const {
    const uint32_t v = out_dst[e];
    if (!bm_get(visited, v) && parent[v] > u) {
      atomicMin(&parent[v], u);
      if (!bm_get(next, v)) bm_set(next, v);
    }
  }
};

// toolkits/cc.cpp:47-58: a smaller label flows to out-neighbours.
struct CcPushOp {
  const uint32_t* out_dst;
  uint32_t* label;
  uint64_t* next;
  __device__ __forceinline__ void operator()(uint32_t u, uint32_t e) const {
    const uint32_t v = out_dst[e];
    const uint32_t msg = label[u];
    if (msg < label[v]) {
      atomicMin(&label[v], msg);
      bm_set(next, v);
    }
  }
};

// toolkits/sssp.cpp:44-58: relax = dist[src] + w (one fp32 add, round to
// nearest); distances are non-negative, so the float order is the order of
// their bit patterns and atomicMin on the bits is write_min.
struct SsspPushOp {
  const uint32_t* out_dst;
  const float* out_w;
  float* dist;
  uint64_t* next;
  __device__ __forceinline__ void operator()(uint32_t u, uint32_t e) const {
    const uint32_t v = out_dst[e];
    const float nd = __fadd_rn(dist[u], out_w[e]);
    if (nd < dist[v]) {
      const int old = atomicMin(reinterpret_cast<int*>(&dist[v]), __float_as_int(nd));
      if (nd < __int_as_float(old)) bm_set(next, v);
    }
  }
};

// frontier degrees for the expansion's prefix
__global__ void k_front_deg(const uint32_t* __restrict__ flist, uint32_t fn, const uint32_t* __restrict__ out_ptr,
                            uint32_t* __restrict__ fdeg) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i <= fn; i += uint64_t(gridDim.x) * blockDim.x)
    fdeg[i] = i < fn ? out_ptr[flist[i] + 1] - out_ptr[flist[i]] : 0u;
}

// Dense signal of toolkits/bfs.cpp:60-68 with R12 on the traversal view
// (in-edges by destination, sources ascending): the first frontier source
// of an unvisited vertex is its smallest, so a scan stops there.  Rows of at
// most kTvSmallRow in-edges take one lane each (4 sources in flight per
// step); larger rows are cut into work items of kTvPullEdges edges, one wave
// each, and write_min over the items gives the row's smallest.  A row's
// first item runs in an earlier launch than its others, which skip when the
// first found a source (it is smaller than all of theirs).
constexpr int kPullRowWords = 4;  // bitmap words a wave takes per step, their loads batched

__device__ __forceinline__ void bfs_pull_rows(uint32_t V, const uint32_t* __restrict__ in_ptr,
                                              const uint32_t* __restrict__ in_src,
                                              const uint64_t* __restrict__ frontier,
                                              const uint64_t* __restrict__ visited, uint32_t* __restrict__ parent,
                                              uint64_t* __restrict__ active_out) {
  // a wave's lanes are the 64 vertices of one bitmap word, so the word's
  // found rows are one plain store (the item launches, the word's other
  // writers, run after this one); a wave takes kPullRowWords words per step
  // and issues each stage's loads for all of them before the next stage
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  constexpr int R = kPullRowWords;
  for (uint64_t v0 = uint64_t(blockIdx.x) * blockDim.x + (threadIdx.x & ~(kWave - 1)); v0 < V; v0 += stride * R) {
    uint64_t vis[R];
#pragma unroll
    for (int r = 0; r < R; ++r) vis[r] = v0 + r * stride < V ? visited[(v0 + r * stride) >> 6] : ~0ull;
    bool any = false;
#pragma unroll
    for (int r = 0; r < R; ++r) any |= vis[r] != ~0ull;
    if (!any) continue;  // wave-uniform
    uint32_t b[R], e[R];
    bool act[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint64_t v = v0 + r * stride + lane;
      act[r] = v < V && !((vis[r] >> lane) & 1ull);
      b[r] = act[r] ? in_ptr[v] : 0u;
      e[r] = act[r] ? in_ptr[v + 1] : 0u;
    }
    // each row's first source alone first (its smallest: most rows of a
    // sparse level's pull end there, on one probe instead of four)
    uint32_t u0[R];
#pragma unroll
    for (int r = 0; r < R; ++r) u0[r] = b[r] < e[r] && e[r] - b[r] <= kTvSmallRow ? in_src[b[r]] : 0u;
    bool found[R];
#pragma unroll
    for (int r = 0; r < R; ++r) found[r] = b[r] < e[r] && e[r] - b[r] <= kTvSmallRow && bm_get(frontier, u0[r]);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint64_t v = v0 + r * stride + lane;
      if (found[r]) {
        parent[v] = u0[r];  // the row's only writer in this step
      } else if (e[r] - b[r] > 1 && e[r] - b[r] <= kTvSmallRow) {  // larger rows are wave items
        // 16-byte aligned groups of 4 sources, one vector load each (the
        // view holds 4 words of padding past its end); sources outside the
        // row are masked
        for (uint32_t k = (b[r] + 1) & ~3u; k < e[r]; k += 4) {
          const uint4 q = *reinterpret_cast<const uint4*>(in_src + k);
          const uint32_t u[4] = {q.x, q.y, q.z, q.w};
          bool f[4];
#pragma unroll
          for (int jj = 0; jj < 4; ++jj) f[jj] = k + jj > b[r] && k + jj < e[r] && bm_get(frontier, u[jj]);
          int hit = -1;
#pragma unroll
          for (int jj = 3; jj >= 0; --jj)
            if (f[jj]) hit = jj;
          if (hit >= 0) {
            parent[v] = u[hit];
            found[r] = true;
            break;
          }
        }
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint64_t m = __ballot(found[r]);
      if (lane == 0 && m) active_out[(v0 + r * stride) >> 6] = m;
    }
  }
}

// kFirst: the items are the first chunks of their rows, one per row, so a
// found parent is a plain store; other chunks of a row race and take the
// smallest.  The found rows' next bits are set with one atomic per bitmap
// word among the wave's items.
template <bool kFirst>
__device__ __forceinline__ void bfs_pull_items(const uint64_t* __restrict__ items, uint64_t n_items,
                                               const uint32_t* __restrict__ in_ptr,
                                               const uint32_t* __restrict__ in_src,
                                               const uint64_t* __restrict__ frontier,
                                               const uint64_t* __restrict__ visited, uint32_t* __restrict__ parent,
                                               uint64_t* __restrict__ active_out) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t gw = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  const uint64_t nw = uint64_t(gridDim.x) * blockDim.x / kWave;
  // 64 items per step: every lane checks one (visited, or an earlier item
  // of the row already holds a smaller parent), then the wave scans the
  // items that are left
  for (uint64_t i0 = gw * kWave; i0 < n_items; i0 += nw * kWave) {
    bool live = false;
    uint32_t my_v = 0, my_e0 = 0, my_e1 = 0;
    if (i0 + lane < n_items) {
      const uint64_t it = items[i0 + lane];
      my_v = uint32_t(it >> 32);
      if (!bm_get(visited, my_v)) {
        my_e0 = in_ptr[my_v] + uint32_t(it) * kTvPullEdges;
        my_e1 = min(my_e0 + kTvPullEdges, in_ptr[my_v + 1]);
        live = !(parent[my_v] < in_src[my_e0]);
      }
    }
    uint32_t my_p = ~0u;  // parent found for this lane's item
    // every lane probes its item's first source (the item's smallest)
    // before the wave scans the items that are left
    if (live) {
      const uint32_t u0 = in_src[my_e0];
      if (bm_get(frontier, u0)) {
        my_p = u0;
        live = false;
      }
    }
    uint64_t pending = __ballot(live);
    // four items at a time, one per group of 16 lanes; a group scans its
    // item in 16-byte aligned groups of 4 sources per lane (64 per round;
    // the view holds 4 words of padding past its end), and its first hit
    // in address order is the smallest source
    const int grp = lane >> 4;
    const uint64_t gmask = 0xffffull << (grp * 16);
    while (pending) {  // wave-uniform
      int jj[4];
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        jj[g] = pending ? __ffsll((unsigned long long)pending) - 1 : -1;
        pending &= pending - 1;
      }
      const int mj = grp == 0 ? jj[0] : grp == 1 ? jj[1] : grp == 2 ? jj[2] : jj[3];
      const uint32_t e0 = __shfl(my_e0, mj < 0 ? 0 : mj, kWave), e1 = __shfl(my_e1, mj < 0 ? 0 : mj, kWave);
      bool done = mj < 0;
      uint32_t found = ~0u;  // the group's parent
      for (uint64_t k = (e0 & ~3u) + uint64_t(lane & 15) * 4; __ballot(!done); k += 64) {
        int hit = -1;
        uint32_t u[4] = {0u, 0u, 0u, 0u};
        if (!done && k < e1) {
          const uint4 q = *reinterpret_cast<const uint4*>(in_src + k);
          u[0] = q.x;
          u[1] = q.y;
          u[2] = q.z;
          u[3] = q.w;
#pragma unroll
          for (int j = 3; j >= 0; --j)
            if (k + j >= e0 && k + j < e1 && bm_get(frontier, u[j])) hit = j;
        }
        const uint32_t cand = hit == 0 ? u[0] : hit == 1 ? u[1] : hit == 2 ? u[2] : u[3];
        const uint64_t hm = __ballot(hit >= 0) & gmask;
        if (hm) {
          const uint32_t p = __shfl(cand, __ffsll((unsigned long long)hm) - 1, kWave);
          if (!done) found = p;
          done = true;
        }
        if (k - uint64_t(lane & 15) * 4 + 64 >= e1) done = true;  // the group's last round
      }
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const uint32_t pg = __shfl(found, g * 16, kWave);
        if (jj[g] >= 0 && lane == jj[g]) my_p = pg;
      }
    }
    const bool found = my_p != ~0u;
    if (found) {
      if constexpr (kFirst)
        parent[my_v] = my_p;
      else
        atomicMin(&parent[my_v], my_p);
    }
    uint64_t fm = __ballot(found);
    while (fm) {  // wave-uniform: one word of found rows at a time
      const uint32_t w = __shfl(my_v, __ffsll((unsigned long long)fm) - 1, kWave) >> 6;
      const bool mine = found && (my_v >> 6) == w;
      const uint64_t bits = wave_reduce(mine ? 1ull << (my_v & 63) : 0ull, OpOr{});
      if (lane == 0) atomicOr(reinterpret_cast<unsigned long long*>(&active_out[w]), (unsigned long long)bits);
      fm &= ~__ballot(mine);
    }
  }
}

__global__ __launch_bounds__(kNT) void k_bfs_pull_rows(uint32_t V, const uint32_t* __restrict__ in_ptr,
                                                       const uint32_t* __restrict__ in_src,
                                                       const uint64_t* __restrict__ frontier,
                                                       const uint64_t* __restrict__ visited,
                                                       uint32_t* __restrict__ parent,
                                                       uint64_t* __restrict__ active_out) {
  bfs_pull_rows(V, in_ptr, in_src, frontier, visited, parent, active_out);
}

template <bool kFirst>
__global__ __launch_bounds__(kNT) void k_bfs_pull_items(const uint64_t* __restrict__ items, uint64_t n_items,
                                                        const uint32_t* __restrict__ in_ptr,
                                                        const uint32_t* __restrict__ in_src,
                                                        const uint64_t* __restrict__ frontier,
                                                        const uint64_t* __restrict__ visited,
                                                        uint32_t* __restrict__ parent,
                                                        uint64_t* __restrict__ active_out) {
  bfs_pull_items<kFirst>(items, n_items, in_ptr, in_src, frontier, visited, parent, active_out);
}

// ------------------------------------------- single-root BFS, device-driven
// eg_bfs's levels as steps of a captured graph (no host round trip per
// level): the level end counts the next frontier and its out-edges (which
// pick the next step's direction on the device); a push step walks the
// frontier bitmap for the vertices up to the hub threshold and the graph's
// list of hub items for the others.  The frontier and next bitmaps
// alternate by step parity, and the level end clears the bitmap the next
// step writes into.

// Counters of one step's frontier.  The frontier's out-edges are added by
// every block of the level end, so they are spread over kNeSlots cache
// lines (atomics on one address serialise); the vertex count is only ever
// tested for zero, so it is a flag set by plain stores.
constexpr int kNeSlots = 16;

struct SCount {
  unsigned long long nv;  // nonzero iff the frontier has a vertex
  unsigned long long pad[7];
  unsigned long long ne[kNeSlots][8];  // out-edges: the sum of ne[i][0]
};

struct SCountView {
  unsigned long long nv, ne;
};

struct SbfsState {
  uint64_t* bm[2];     // frontier of step s is bm[s & 1], its next bm[(s + 1) & 1]
  uint64_t* vis;
  SCount* cnt;         // ring of 3 by step
  uint32_t* steps;     // steps that had a frontier
  uint32_t* base;      // first step of the queued batch
  unsigned long long* total;  // out-edges of every reached vertex
};

struct SbfsArgs {
  uint32_t V;
  uint64_t E;
  int policy;
  const uint32_t *out_ptr, *out_dst, *deg, *in_ptr, *in_src;
  const uint8_t* deg8;
  const uint64_t *big, *big_rest;
  uint64_t n_big, n_big_rest;
  const uint64_t* hub_items;
  uint64_t n_hub_items;
  uint32_t dense_div, hub, hub_chunk;  // the graph's tune
  SbfsState st;
  uint32_t *parent, *depth;
};

template <class A>
__device__ __forceinline__ SCountView count_of(const A& a, uint32_t step) {
  const SCount& c = a.st.cnt[step % 3];
  SCountView r{c.nv, 0};
#pragma unroll
  for (int i = 0; i < kNeSlots; ++i) r.ne += c.ne[i][0];
  return r;
}

// auto policy: dense once the frontier's out-edges reach E / div (the
// graph's tune; default 20: twitter-shaped BFS E/10 4.73 ms, E/20 3.97,
// E/40 6.38; RMAT-24 within 1%)
__device__ __forceinline__ bool sbfs_sparse(int policy, unsigned long long ne, uint64_t E, uint32_t div) {
  if (policy == EG_POLICY_PULL) return false;
  if (policy == EG_POLICY_AUTO) return ne < E / div;
  return true;  // reference (rooted programs push, R11) and push
}

// A push step whose frontier has at least V/4 out-edges leaves the next
// bitmap to the level end, which reads it off parent (V until reached) with
// one coalesced pass: cheaper than an atomic per reached vertex.
__device__ __forceinline__ bool sbfs_derive(unsigned long long ne, uint32_t V) { return ne >= V / 4; }

constexpr int kRelaxBatch = 16;  // out-edges per lane whose loads are all in flight at once

// Push relaxation (bfs.cpp:52-58 with R12) of the out-edges
// e = first + k * stride, k < n <= kRelaxBatch, of frontier vertex u: the
// targets, their visited words and their parents are each loaded for the
// whole batch before the first is used, then the atomics of the batch go
// out together.  parent is V until a vertex is reached, so the one atomic
// that sees V sets the vertex's next bit (kDerive: the level end does).
static_assert(kSbfsLaneEdges <= uint32_t(kRelaxBatch), "a lane's vertex is one batch");

template <bool kDerive>
__device__ __forceinline__ void sbfs_relax(const SbfsArgs& a, uint64_t* __restrict__ nx, uint32_t u, uint32_t first,
                                           uint32_t stride, uint32_t n) {
  uint32_t v[kRelaxBatch], p[kRelaxBatch], old[kRelaxBatch];
  uint64_t w[kRelaxBatch];
#pragma unroll
  for (int k = 0; k < kRelaxBatch; ++k) v[k] = uint32_t(k) < n ? a.out_dst[first + uint32_t(k) * stride] : 0u;
#pragma unroll
  for (int k = 0; k < kRelaxBatch; ++k) w[k] = uint32_t(k) < n ? a.st.vis[v[k] >> 6] : ~0ull;
#pragma unroll
  for (int k = 0; k < kRelaxBatch; ++k) p[k] = ((w[k] >> (v[k] & 63)) & 1ull) ? 0u : a.parent[v[k]];
  if constexpr (kDerive) {
#pragma unroll
    for (int k = 0; k < kRelaxBatch; ++k)
      if (p[k] > u) atomicMin(&a.parent[v[k]], u);
  } else {
#pragma unroll
    for (int k = 0; k < kRelaxBatch; ++k) old[k] = p[k] > u ? atomicMin(&a.parent[v[k]], u) : 0u;
#pragma unroll
    for (int k = 0; k < kRelaxBatch; ++k)
      if (old[k] == a.V) bm_set(nx, v[k]);
  }
}

// the edges [e0, e1) of u over the wave, kRelaxBatch per lane at a time
template <bool kDerive>
__device__ __forceinline__ void sbfs_relax_wave(const SbfsArgs& a, uint64_t* __restrict__ nx, uint32_t u, uint32_t e0,
                                                uint32_t e1, int lane) {
  for (uint32_t b = e0; b < e1; b += kWave * kRelaxBatch) {
    const uint32_t f = b + uint32_t(lane);
    const uint32_t n = f < e1 ? ceil_div(e1 - f, uint32_t(kWave)) : 0u;
    sbfs_relax<kDerive>(a, nx, u, f, kWave, n < uint32_t(kRelaxBatch) ? n : uint32_t(kRelaxBatch));
  }
}

// Sparse step: frontier vertices of out-degree up to kSbfsLaneEdges run in
// their lane, up to the hub threshold with the whole wave; hubs' items are strided
// over the waves.
template <bool kDerive>
__device__ __forceinline__ void sbfs_push(const SbfsArgs& a, uint32_t step) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t gw = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  const uint64_t nw = uint64_t(gridDim.x) * blockDim.x / kWave;
  const uint64_t* fr = a.st.bm[step & 1];
  uint64_t* nx = a.st.bm[(step + 1) & 1];
  // the frontier bitmap, 64 words per wave step: each lane first runs the
  // small vertices of its own word, then the wave walks the words that hold
  // larger ones
  const uint64_t words = (uint64_t(a.V) >> 6) + 1;
  for (uint64_t b0 = gw * kWave; b0 < words; b0 += nw * kWave) {
    const uint64_t wi = b0 + lane;
    const uint64_t mine = wi < words ? fr[wi] : 0ull;
    bool has_mid = false;
    for (uint64_t m = mine; m; m &= m - 1) {
      const uint32_t u = uint32_t(wi * 64 + __ffsll((unsigned long long)m) - 1);
      const uint32_t e0 = a.out_ptr[u], e1 = a.out_ptr[u + 1];
      if (e1 - e0 <= kSbfsLaneEdges)
        sbfs_relax<kDerive>(a, nx, u, e0, 1, e1 - e0);
      else if (e1 - e0 <= a.hub)
        has_mid = true;
    }
    uint64_t wm = __ballot(has_mid);
    while (wm) {  // wave-uniform
      const int jj = __ffsll((unsigned long long)wm) - 1;
      wm &= wm - 1;
      const bool set = (shfl_t(mine, jj) >> lane) & 1ull;
      const uint32_t u = uint32_t((b0 + jj) * 64 + lane);
      const uint32_t e0 = set ? a.out_ptr[u] : 0u, e1 = set ? a.out_ptr[u + 1] : 0u;
      uint64_t mid = __ballot(e1 - e0 > kSbfsLaneEdges && e1 - e0 <= a.hub);
      while (mid) {  // wave-uniform
        const int l = __ffsll((unsigned long long)mid) - 1;
        mid &= mid - 1;
        sbfs_relax_wave<kDerive>(a, nx, __shfl(u, l, kWave), __shfl(e0, l, kWave), __shfl(e1, l, kWave), lane);
      }
    }
  }
  // hubs: the graph's list of their items, 64 per wave step with item
  // (r + lane) * nw + gw in lane `lane` of wave gw (a hub's consecutive
  // items go to different waves); the lanes whose hub is in the frontier
  // are run by the wave one after another
  for (uint64_t r = 0; r * nw < a.n_hub_items; r += kWave) {
    const uint64_t i = (r + lane) * nw + gw;
    uint64_t it = 0;
    bool live = false;
    if (i < a.n_hub_items) {
      it = a.hub_items[i];
      live = bm_get(fr, uint32_t(it >> 32));
    }
    uint64_t lm = __ballot(live);
    while (lm) {  // wave-uniform
      const int l = __ffsll((unsigned long long)lm) - 1;
      lm &= lm - 1;
      const uint64_t t = shfl_t(it, l);
      const uint32_t u = uint32_t(t >> 32);
      const uint32_t e0 = a.out_ptr[u] + uint32_t(t) * a.hub_chunk;
      sbfs_relax_wave<kDerive>(a, nx, u, e0, min(e0 + a.hub_chunk, a.out_ptr[u + 1]), lane);
    }
  }
}

__global__ __launch_bounds__(kNT) void k_sbfs_push(SbfsArgs a, uint32_t j) {
  const uint32_t step = *a.st.base + j;
  const SCountView c = count_of(a, step);
  if (!c.nv || !sbfs_sparse(a.policy, c.ne, a.E, a.dense_div)) return;
  if (sbfs_derive(c.ne, a.V))
    sbfs_push<true>(a, step);
  else
    sbfs_push<false>(a, step);
}

// Dense step: the pulls above, on this step's bitmaps.
__global__ __launch_bounds__(kNT) void k_sbfs_pull_rows(SbfsArgs a, uint32_t j) {
  const uint32_t step = *a.st.base + j;
  const SCountView c = count_of(a, step);
  if (!c.nv || sbfs_sparse(a.policy, c.ne, a.E, a.dense_div)) return;
  bfs_pull_rows(a.V, a.in_ptr, a.in_src, a.st.bm[step & 1], a.st.vis, a.parent, a.st.bm[(step + 1) & 1]);
}

template <int kRest>
__global__ __launch_bounds__(kNT) void k_sbfs_pull_items(SbfsArgs a, uint32_t j) {
  const uint32_t step = *a.st.base + j;
  const SCountView c = count_of(a, step);
  if (!c.nv || sbfs_sparse(a.policy, c.ne, a.E, a.dense_div)) return;
  bfs_pull_items<!kRest>(kRest ? a.big_rest : a.big, kRest ? a.n_big_rest : a.n_big, a.in_ptr, a.in_src,
                 a.st.bm[step & 1], a.st.vis, a.parent, a.st.bm[(step + 1) & 1]);
}

// Level end (bfs.cpp:70-90): visited |= next, depth = level + 1, counters
// of the next step; clears this step's frontier bitmap (the next step's
// next).  The next bitmap's non-empty words are taken kEndWords at a time
// with a lane per vertex of each word; the depth stores and byte-degree
// loads of all of them are issued before the first degree is used.  After
// a push step that left the next bitmap to it, the level end reads it off
// parent (V until reached), kEndWords words at a time.
constexpr int kEndWords = 16;

__global__ __launch_bounds__(kNT) void k_sbfs_end(SbfsArgs a, uint32_t j) {
  const uint32_t step = *a.st.base + j;
  const SCountView c = count_of(a, step);
  const bool live = c.nv != 0;
  if (blockIdx.x == 0) {
    // the slot after next, last read at step - 1
    unsigned long long* z = reinterpret_cast<unsigned long long*>(&a.st.cnt[(step + 2) % 3]);
    for (uint32_t i = threadIdx.x; i < sizeof(SCount) / 8; i += blockDim.x) z[i] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (live) {
      *a.st.steps = step + 1;
      *a.st.total += c.ne;  // this step's frontier edges (its only writer)
    }
  }
  if (!live) return;  // block-uniform
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t gw = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  const uint64_t nw = uint64_t(gridDim.x) * blockDim.x / kWave;
  const uint64_t words = (uint64_t(a.V) >> 6) + 1;
  uint64_t* fr = a.st.bm[step & 1];
  uint64_t* nx = a.st.bm[(step + 1) & 1];
  SCount* nc = &a.st.cnt[(step + 1) % 3];
  const bool derive = sbfs_sparse(a.policy, c.ne, a.E, a.dense_div) && sbfs_derive(c.ne, a.V);
  unsigned long long nv = 0, ne = 0;
  // the vertices of up to kEndWords words b0 + idx[k] (idx[k] < 0: none),
  // whose bits lane idx[k] holds in `mine`: depth stores and degree loads
  // all issued before the first degree is used
  auto vertices = [&](uint64_t b0, const int (&idx)[kEndWords], uint64_t mine) {
    uint32_t d[kEndWords];
#pragma unroll
    for (int k = 0; k < kEndWords; ++k) {
      const uint64_t wk = idx[k] >= 0 ? lane_value(mine, idx[k]) : 0ull;
      d[k] = 0;
      if ((wk >> lane) & 1ull) {
        const uint32_t v = uint32_t((b0 + uint64_t(idx[k])) * 64 + lane);
        a.depth[v] = step + 1;
        d[k] = a.deg8[v];
      }
    }
#pragma unroll
    for (int k = 0; k < kEndWords; ++k)
      if (d[k] == 255u) d[k] = a.deg[uint32_t((b0 + uint64_t(idx[k])) * 64 + lane)];  // saturated: the exact degree
#pragma unroll
    for (int k = 0; k < kEndWords; ++k) ne += d[k];
  };
  if (derive) {
    // next = reached (parent set) and not visited, kEndWords words at a
    // time (a big push level: most words have vertices); lanes below
    // kEndWords move the words
    for (uint64_t b0 = gw * kEndWords; b0 < words; b0 += nw * kEndWords) {
      const bool wl = lane < kEndWords && b0 + lane < words;
      uint64_t mine = 0;
      const uint64_t vw = wl ? a.st.vis[b0 + lane] : ~0ull;
      uint32_t pv[kEndWords];
#pragma unroll
      for (int k = 0; k < kEndWords; ++k) {
        const uint64_t v = (b0 + k) * 64 + lane;
        pv[k] = v < a.V ? a.parent[v] : a.V;
      }
#pragma unroll
      for (int k = 0; k < kEndWords; ++k) {
        const uint64_t m = __ballot(pv[k] != a.V && !((lane_value(vw, k) >> lane) & 1ull));
        if (lane == k) mine = m;
      }
      if (wl) {
        fr[b0 + lane] = 0;
        if (mine) {
          nx[b0 + lane] = mine;
          a.st.vis[b0 + lane] = vw | mine;
        }
      }
      nv += __popcll(mine);
      int idx[kEndWords];
#pragma unroll
      for (int k = 0; k < kEndWords; ++k) idx[k] = k;
      vertices(b0, idx, mine);
    }
  } else {
    // 64 words per wave step, a lane each (visited |= next, frontier = 0),
    // then the non-empty ones kEndWords at a time
    for (uint64_t b0 = gw * kWave; b0 < words; b0 += nw * kWave) {
      const uint64_t wi = b0 + lane;
      uint64_t mine = 0;
      if (wi < words) {
        mine = nx[wi];
        fr[wi] = 0;
        if (mine) a.st.vis[wi] |= mine;
      }
      nv += __popcll(mine);
      uint64_t nz = __ballot(mine != 0);
      while (nz) {  // wave-uniform
        int idx[kEndWords];
#pragma unroll
        for (int k = 0; k < kEndWords; ++k) {
          idx[k] = nz ? __ffsll((unsigned long long)nz) - 1 : -1;
          nz &= nz - 1;
        }
        vertices(b0, idx, mine);
      }
    }
  }
  __shared__ unsigned long long part[2][kNT / kWave];
  nv = wave_reduce(nv, OpSum{});
  ne = wave_reduce(ne, OpSum{});
  if (lane == 0) {
    part[0][threadIdx.x / kWave] = nv;
    part[1][threadIdx.x / kWave] = ne;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long x = 0, y = 0;
    for (int k = 0; k < kNT / kWave; ++k) {
      x += part[0][k];
      y += part[1][k];
    }
    if (x) nc->nv = 1;
    if (y) atomicAdd(&nc->ne[blockIdx.x % kNeSlots][0], y);
  }
}

// The start state in one launch: parent V and depth ~0 but the root's,
// the frontier, next and visited bitmaps empty but the root's bit in the
// frontier and visited, the counter ring and scalars (the frontier is the
// root and its out-edges).
__global__ void k_sbfs_init(SbfsArgs a, uint32_t root) {
  const uint64_t words = (uint64_t(a.V) >> 6) + 1;
  const uint64_t n = a.V > 3 * words ? a.V : 3 * words;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
    if (i < a.V) {
      a.parent[i] = i == root ? root : a.V;
      a.depth[i] = i == root ? 0u : 0xffffffffu;
    }
    if (i < 3 * words) {  // bm[0], bm[1], vis are contiguous
      const bool rw = i == (root >> 6) || i == 2 * words + (root >> 6);
      a.st.bm[0][i] = rw ? 1ull << (root & 63) : 0ull;
    }
  }
  if (blockIdx.x != 0) return;
  unsigned long long* z = reinterpret_cast<unsigned long long*>(a.st.cnt);
  for (uint32_t i = threadIdx.x; i < 3 * sizeof(SCount) / 8; i += blockDim.x) z[i] = 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    a.st.cnt[0].nv = 1;
    a.st.cnt[0].ne[0][0] = a.out_ptr[root + 1] - a.out_ptr[root];
    *a.st.steps = 0;
    *a.st.base = 0;
    *a.st.total = 0;
  }
}

// ------------------------------------------ single-root SSSP, device-driven
// eg_sssp's iterations (sssp.cpp:38-85) as steps of a captured graph, like
// eg_bfs: a push step walks the active bitmap (vertices up to the hub threshold
// out-edges) and the hub items, a dense step is the pull above, and the
// step end counts the next active set and its out-edges (the mode switch
// input) and clears the bitmap the next step writes into.
// Mode switch of the auto policy for SSSP: dense once the active
// vertices' out-edges reach E/80.  A push step's improving relaxations are
// atomics and most of its edges improve, while the dense pass streams;
// 64 R-MAT 22 runs: E/2 594 ms, E/4 344, E/8 244, E/20 198, E/40 189,
// E/80 187, E/160 190 (the default; the graph's tune sets it).
__device__ __forceinline__ bool dsssp_sparse(int policy, unsigned long long ne, uint64_t E, uint32_t div) {
  return sbfs_sparse(policy, ne, E, div);
}

struct DssspArgs {
  uint32_t V;
  uint64_t E;
  int policy;
  const uint32_t *out_ptr, *out_dst, *deg;
  const float* out_w;
  const uint8_t* deg8;
  const uint64_t* hub_items;
  uint64_t n_hub_items;
  uint32_t dense_div, hub, hub_chunk;  // the graph's tune
  const uint32_t *csc_src, *wordrank, *row_dst;
  const float* csc_w;
  const uint64_t* rowflag;
  SbfsState st;  // bm[2], cnt, steps, base (vis and total unused)
  float* dist;
  uint64_t* sum[2];  // active summary of step s in sum[s & 1]: bit w set iff word w of its bitmap is not empty
  uint32_t nsum;     // summary words, ((V - 1) >> 12) + 1: bit w of word w >> 6 for bitmap words w < V/64 rounded up
};

// Summary words a dense step stages in LDS (32 KB: V up to 2^24); larger
// graphs run the dense step without the summary, and reserve no LDS for it.
constexpr uint32_t kSumMax = 4096;
inline uint32_t summary_words(uint32_t V) { return ((V ? V - 1 : 0) >> 12) + 1; }

// Sparse slot (sssp.cpp:44-58) for the out-edges e = first + k * stride,
// k < n <= kRelaxBatch, of active vertex u at distance du: targets, weights
// and target distances loaded for the batch before the first use.
__device__ __forceinline__ void dsssp_relax(const DssspArgs& a, uint64_t* __restrict__ nx, float du, uint32_t first,
                                            uint32_t stride, uint32_t n) {
  uint32_t v[kRelaxBatch];
  float w[kRelaxBatch], dv[kRelaxBatch];
  int old[kRelaxBatch];
#pragma unroll
  for (int k = 0; k < kRelaxBatch; ++k) {
    const bool ok = uint32_t(k) < n;
    v[k] = ok ? a.out_dst[first + uint32_t(k) * stride] : 0u;
    w[k] = ok ? a.out_w[first + uint32_t(k) * stride] : 0.0f;
  }
#pragma unroll
  for (int k = 0; k < kRelaxBatch; ++k) dv[k] = uint32_t(k) < n ? a.dist[v[k]] : 0.0f;
#pragma unroll
  for (int k = 0; k < kRelaxBatch; ++k) {
    const float nd = __fadd_rn(du, w[k]);
    old[k] = uint32_t(k) < n && nd < dv[k] ? atomicMin(reinterpret_cast<int*>(&a.dist[v[k]]), __float_as_int(nd))
                                           : __float_as_int(0.0f);
  }
#pragma unroll
  for (int k = 0; k < kRelaxBatch; ++k)
    if (uint32_t(k) < n && __fadd_rn(du, w[k]) < __int_as_float(old[k])) bm_set(nx, v[k]);
}

__device__ __forceinline__ void dsssp_relax_wave(const DssspArgs& a, uint64_t* __restrict__ nx, float du, uint32_t e0,
                                                 uint32_t e1, int lane) {
  for (uint32_t b = e0; b < e1; b += kWave * kRelaxBatch) {
    const uint32_t f = b + uint32_t(lane);
    const uint32_t n = f < e1 ? ceil_div(e1 - f, uint32_t(kWave)) : 0u;
    dsssp_relax(a, nx, du, f, kWave, n < uint32_t(kRelaxBatch) ? n : uint32_t(kRelaxBatch));
  }
}

__global__ __launch_bounds__(kNT) void k_dsssp_push(DssspArgs a, uint32_t j) {
  const uint32_t step = *a.st.base + j;
  const SCountView c = count_of(a, step);
  if (!c.nv || !dsssp_sparse(a.policy, c.ne, a.E, a.dense_div)) return;
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t gw = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  const uint64_t nw = uint64_t(gridDim.x) * blockDim.x / kWave;
  const uint64_t* fr = a.st.bm[step & 1];
  uint64_t* nx = a.st.bm[(step + 1) & 1];
  const uint64_t words = (uint64_t(a.V) >> 6) + 1;
  for (uint64_t b0 = gw * kWave; b0 < words; b0 += nw * kWave) {
    const uint64_t wi = b0 + lane;
    const uint64_t mine = wi < words ? fr[wi] : 0ull;
    bool has_mid = false;
    for (uint64_t m = mine; m; m &= m - 1) {
      const uint32_t u = uint32_t(wi * 64 + __ffsll((unsigned long long)m) - 1);
      const uint32_t e0 = a.out_ptr[u], e1 = a.out_ptr[u + 1];
      if (e1 - e0 <= kSbfsLaneEdges)
        dsssp_relax(a, nx, a.dist[u], e0, 1, e1 - e0);
      else if (e1 - e0 <= a.hub)
        has_mid = true;
    }
    uint64_t wm = __ballot(has_mid);
    while (wm) {  // wave-uniform
      const int jj = __ffsll((unsigned long long)wm) - 1;
      wm &= wm - 1;
      const bool set = (shfl_t(mine, jj) >> lane) & 1ull;
      const uint32_t u = uint32_t((b0 + jj) * 64 + lane);
      const uint32_t e0 = set ? a.out_ptr[u] : 0u, e1 = set ? a.out_ptr[u + 1] : 0u;
      const bool is_mid = e1 - e0 > kSbfsLaneEdges && e1 - e0 <= a.hub;
      const float du = is_mid ? a.dist[u] : 0.0f;
      uint64_t mid = __ballot(is_mid);
      while (mid) {  // wave-uniform
        const int l = __ffsll((unsigned long long)mid) - 1;
        mid &= mid - 1;
        dsssp_relax_wave(a, nx, __shfl(du, l, kWave), __shfl(e0, l, kWave), __shfl(e1, l, kWave), lane);
      }
    }
  }
  for (uint64_t r = 0; r * nw < a.n_hub_items; r += kWave) {  // hub items, as in k_sbfs_push
    const uint64_t i = (r + lane) * nw + gw;
    uint64_t it = 0;
    bool live = false;
    if (i < a.n_hub_items) {
      it = a.hub_items[i];
      live = bm_get(fr, uint32_t(it >> 32));
    }
    uint64_t lm = __ballot(live);
    while (lm) {  // wave-uniform
      const int l = __ffsll((unsigned long long)lm) - 1;
      lm &= lm - 1;
      const uint64_t t = shfl_t(it, l);
      const uint32_t u = uint32_t(t >> 32);
      const uint32_t e0 = a.out_ptr[u] + uint32_t(t) * a.hub_chunk;
      dsssp_relax_wave(a, nx, a.dist[u], e0, min(e0 + a.hub_chunk, a.out_ptr[u + 1]), lane);
    }
  }
}

// Dense step: k_sssp_pull's reduction on this step's bitmaps, over the
// in-edges whose source lies in a word of the active bitmap that is not
// empty (its summary staged in LDS, when it fits).  A source outside the
// active set has not changed since its out-edges were last relaxed, so
// dist[dst] <= dist[src] + w already holds for it and skipping it changes
// no distance; the reference relaxes every in-edge (sssp.cpp:65 has its
// active test commented out).
__global__ __launch_bounds__(kNT) void k_dsssp_pull(DssspArgs a, uint32_t j) {
  const uint32_t step = *a.st.base + j;
  const SCountView c = count_of(a, step);
  if (!c.nv || dsssp_sparse(a.policy, c.ne, a.E, a.dense_div)) return;  // block-uniform
  extern __shared__ uint64_t s_sum[];  // a.nsum words when they fit kSumMax (launch with that many), else none
  const bool use_sum = a.nsum <= kSumMax;
  if (use_sum) {
    const uint64_t* sum = a.sum[step & 1];
    for (uint32_t i = threadIdx.x; i < a.nsum; i += blockDim.x) s_sum[i] = sum[i];
    __syncthreads();
  }
  uint64_t* nx = a.st.bm[(step + 1) & 1];
  float* dist = a.dist;
  row_reduce_wave<float>(
      first_word(), a.E, a.csc_src, a.rowflag, a.wordrank, 1e9f, OpMin{},
      [&](uint64_t, uint32_t s, float we) {
        return !use_sum || ((s_sum[s >> 12] >> ((s >> 6) & 63)) & 1ull) ? __fadd_rn(dist[s], we) : 1e9f;
      },
      [&](uint32_t row, float m, bool) {
        const uint32_t v = a.row_dst[row];
        if (m < 1e9f && m < dist[v]) {
          const int old = atomicMin(reinterpret_cast<int*>(&dist[v]), __float_as_int(m));
          if (m < __int_as_float(old)) bm_set(nx, v);
        }
      },
      a.csc_w);
}

// Step end: counters of the next active set, this step's bitmap cleared
// (the next step's next); 64 words per wave step, the non-empty ones
// kEndWords at a time.
__global__ __launch_bounds__(kNT) void k_dsssp_end(DssspArgs a, uint32_t j) {
  const uint32_t step = *a.st.base + j;
  const bool live = count_of(a, step).nv != 0;
  if (blockIdx.x == 0) {
    unsigned long long* z = reinterpret_cast<unsigned long long*>(&a.st.cnt[(step + 2) % 3]);
    for (uint32_t i = threadIdx.x; i < sizeof(SCount) / 8; i += blockDim.x) z[i] = 0;
    if (threadIdx.x == 0 && live) *a.st.steps = step + 1;
  }
  if (!live) return;  // block-uniform
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t gw = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  const uint64_t nw = uint64_t(gridDim.x) * blockDim.x / kWave;
  const uint64_t words = (uint64_t(a.V) >> 6) + 1;
  uint64_t* fr = a.st.bm[step & 1];
  const uint64_t* nx = a.st.bm[(step + 1) & 1];
  uint64_t* nsum = a.sum[(step + 1) & 1];
  SCount* nc = &a.st.cnt[(step + 1) % 3];
  unsigned long long nv = 0, ne = 0;
  for (uint64_t b0 = gw * kWave; b0 < words; b0 += nw * kWave) {
    const uint64_t wi = b0 + lane;
    uint64_t mine = 0;
    if (wi < words) {
      mine = nx[wi];
      fr[wi] = 0;
    }
    nv += __popcll(mine);
    uint64_t nz = __ballot(mine != 0);
    // the next step's summary word; the bitmap's last word may lie past
    // the summary (V a multiple of 64: that word holds no vertex and is empty)
    if (lane == 0 && (b0 >> 6) < a.nsum) nsum[b0 >> 6] = nz;
    while (nz) {  // wave-uniform
      int idx[kEndWords];
#pragma unroll
      for (int k = 0; k < kEndWords; ++k) {
        idx[k] = nz ? __ffsll((unsigned long long)nz) - 1 : -1;
        nz &= nz - 1;
      }
      uint32_t d[kEndWords];
#pragma unroll
      for (int k = 0; k < kEndWords; ++k) {
        const uint64_t wk = idx[k] >= 0 ? lane_value(mine, idx[k]) : 0ull;
        d[k] = (wk >> lane) & 1ull ? a.deg8[uint32_t((b0 + uint64_t(idx[k])) * 64 + lane)] : 0u;
      }
#pragma unroll
      for (int k = 0; k < kEndWords; ++k)
        if (d[k] == 255u) d[k] = a.deg[uint32_t((b0 + uint64_t(idx[k])) * 64 + lane)];
#pragma unroll
      for (int k = 0; k < kEndWords; ++k) ne += d[k];
    }
  }
  __shared__ unsigned long long part[2][kNT / kWave];
  nv = wave_reduce(nv, OpSum{});
  ne = wave_reduce(ne, OpSum{});
  if (lane == 0) {
    part[0][threadIdx.x / kWave] = nv;
    part[1][threadIdx.x / kWave] = ne;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long x = 0, y = 0;
    for (int k = 0; k < kNT / kWave; ++k) {
      x += part[0][k];
      y += part[1][k];
    }
    if (x) nc->nv = 1;
    if (y) atomicAdd(&nc->ne[blockIdx.x % kNeSlots][0], y);
  }
}

// Start state in one launch: distances 1e9 but the root's 0, the root the
// only active vertex, the counter ring and scalars.
__global__ void k_dsssp_init(DssspArgs a, uint32_t root) {
  const uint64_t words = (uint64_t(a.V) >> 6) + 1;
  const uint64_t n = a.V > 2 * words ? a.V : 2 * words;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
    if (i < a.V) a.dist[i] = i == root ? 0.0f : 1e9f;
    if (i < 2 * words) a.st.bm[0][i] = i == (root >> 6) ? 1ull << (root & 63) : 0ull;  // bm[0], bm[1] contiguous
    if (i < a.nsum) a.sum[0][i] = i == (root >> 12) ? 1ull << ((root >> 6) & 63) : 0ull;
  }
  if (blockIdx.x != 0) return;
  unsigned long long* z = reinterpret_cast<unsigned long long*>(a.st.cnt);
  for (uint32_t i = threadIdx.x; i < 3 * sizeof(SCount) / 8; i += blockDim.x) z[i] = 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    a.st.cnt[0].nv = 1;
    a.st.cnt[0].ne[0][0] = a.out_ptr[root + 1] - a.out_ptr[root];
    *a.st.steps = 0;
    *a.st.base = 0;
  }
}

// ---------------------------------------------------------------- CC -----
// Pull: min label over in-edges (starting from dst itself, cc.cpp:60-68).
__global__ __launch_bounds__(kNT) void k_cc_pull(uint64_t E, const uint32_t* __restrict__ src,
                                                  const uint64_t* __restrict__ rowflag,
                                                  const uint32_t* __restrict__ wordrank,
                                                  const uint32_t* __restrict__ row_dst, uint32_t* __restrict__ label,
                                                  uint64_t* __restrict__ active_out) {
  row_reduce_wave<uint32_t>(
      first_word(), E, src, rowflag, wordrank, 0xffffffffu, OpMin{}, [&](uint64_t, uint32_t s) { return label[s]; },
      [&](uint32_t row, uint32_t m, bool) {
        const uint32_t v = row_dst[row];
        if (m < v && m < label[v]) {
          atomicMin(&label[v], m);
          bm_set(active_out, v);
        }
      });
}

// --------------------------------------------------------------- SSSP ----
// Pull: min over in-edges of dist[src] + w (sssp.cpp:60-73).
__global__ __launch_bounds__(kNT) void k_sssp_pull(uint64_t E, const uint32_t* __restrict__ src,
                                                    const float* __restrict__ w, const uint64_t* __restrict__ rowflag,
                                                    const uint32_t* __restrict__ wordrank,
                                                    const uint32_t* __restrict__ row_dst, float* __restrict__ dist,
                                                    uint64_t* __restrict__ active_out) {
  row_reduce_wave<float>(
      first_word(), E, src, rowflag, wordrank, 1e9f, OpMin{},
      [&](uint64_t, uint32_t s, float we) { return __fadd_rn(dist[s], we); },
      [&](uint32_t row, float m, bool) {
        const uint32_t v = row_dst[row];
        if (m < 1e9f && m < dist[v]) {
          const int old = atomicMin(reinterpret_cast<int*>(&dist[v]), __float_as_int(m));
          if (m < __int_as_float(old)) bm_set(active_out, v);
        }
      },
      w);
}

// ------------------------------------------- messages between partitions --
// exe_data of a dense step (core/graph.hpp:1736): the messages that the
// dense signals of a partition's rows send to the other partitions' masters.
// Per pull row, the value its program's dense signal tests: PageRank emits
// always (pagerank.cpp:64-70), BFS when the destination is unvisited and a
// source active (bfs.cpp:59-67), CC when a source's label is below the
// destination id (cc.cpp:62-72), SSSP when a relaxation is below 1e9
// (sssp.cpp:62-74).
template <int kProg>
__global__ __launch_bounds__(kNT) void k_msg_rows(uint64_t E, const uint32_t* __restrict__ src,
                                                   const float* __restrict__ w, const uint64_t* __restrict__ rowflag,
                                                   const uint32_t* __restrict__ wordrank,
                                                   const uint64_t* __restrict__ active, const void* __restrict__ values,
                                                   uint32_t* __restrict__ acc) {
  auto put_min = [&](uint32_t row, uint32_t v, bool shared) {
    if (shared)
      atomicMin(&acc[row], v);
    else
      acc[row] = v;
  };
  if constexpr (kProg == EG_PROGRAM_SSSP) {
    const float* dist = static_cast<const float*>(values);
    row_reduce_wave<uint32_t>(
        first_word(), E, src, rowflag, wordrank, 0xFFFFFFFFu, OpMin{},
        [&](uint64_t, uint32_t s, float we) { return __float_as_uint(__fadd_rn(dist[s], we)); }, put_min, w);
  } else if constexpr (kProg == EG_PROGRAM_CC) {
    const uint32_t* label = static_cast<const uint32_t*>(values);
    row_reduce_wave<uint32_t>(first_word(), E, src, rowflag, wordrank, 0xFFFFFFFFu, OpMin{},
                              [&](uint64_t, uint32_t s) { return label[s]; }, put_min);
  } else {
    row_reduce_wave<uint32_t>(
        first_word(), E, src, rowflag, wordrank, 0u, OpOr{},
        [&](uint64_t, uint32_t s) {
          return kProg == EG_PROGRAM_BFS ? uint32_t((active[s >> 6] >> (s & 63)) & 1ull) : 1u;
        },
        [&](uint32_t row, uint32_t v, bool shared) {
          if (shared)
            atomicOr(&acc[row], v);
          else
            acc[row] = v;
        });
  }
}

// Rows whose destination another partition owns and whose signal emits.
template <int kProg>
__global__ __launch_bounds__(kNT) void k_msg_count(uint64_t R, const uint32_t* __restrict__ acc,
                                                    const uint32_t* __restrict__ row_dst,
                                                    const uint32_t* __restrict__ row_start,
                                                    const uint32_t* __restrict__ src, const uint16_t* __restrict__ qmap,
                                                    int S, const uint64_t* __restrict__ visited,
                                                    unsigned long long* __restrict__ out) {
  __shared__ unsigned long long wt[kNT / kWave];
  unsigned long long n = 0;
  for (uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < R; r += uint64_t(gridDim.x) * blockDim.x) {
    const uint32_t d = row_dst[r];
    const uint32_t s0 = src[row_start[r]];
    if (qmap[d >> kPageShift] / S == qmap[s0 >> kPageShift] / S) continue;
    const uint32_t a = acc[r];
    bool emits = true;
    if constexpr (kProg == EG_PROGRAM_BFS) emits = a != 0 && !((visited[d >> 6] >> (d & 63)) & 1ull);
    if constexpr (kProg == EG_PROGRAM_CC) emits = a < d;
    if constexpr (kProg == EG_PROGRAM_SSSP) emits = __uint_as_float(a) < 1e9f;
    n += emits ? 1 : 0;
  }
  n = wave_reduce(n, OpSum{});
  if ((threadIdx.x & (kWave - 1)) == 0) wt[threadIdx.x / kWave] = n;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int i = 0; i < kNT / kWave; ++i) t += wt[i];
    atomicAdd(out, t);
  }
}

template <int kProg>
int count_messages(eg_graph* g, const uint64_t* active, const uint64_t* visited, const void* values, uint64_t* h_count,
                   hipStream_t s) {
  const uint64_t R = g->R;
  EG_TRY(g->e_msg.resize((R + 1) * 4 + 64));
  uint32_t* acc = g->e_msg.as<uint32_t>();
  unsigned long long* cnt = reinterpret_cast<unsigned long long*>(g->e_msg.as<char>() + ((R * 4 + 63) & ~63ull));
  EG_TRY(fill<unsigned long long>(cnt, 1, 0ull, s));
  if (R) {
    const uint32_t id = kProg == EG_PROGRAM_SSSP || kProg == EG_PROGRAM_CC ? 0xFFFFFFFFu : 0u;
    EG_TRY(fill<uint32_t>(acc, R, id, s));
    hipLaunchKernelGGL((k_msg_rows<kProg>), dim3(row_grid(g->E)), dim3(kNT), 0, s, g->E,
                       (const uint32_t*)g->e_csc_src.as<uint32_t>(), (const float*)g->e_csc_w.as<float>(),
                       (const uint64_t*)g->e_rowflag.as<uint64_t>(), (const uint32_t*)g->e_wordrank.as<uint32_t>(),
                       active, values, acc);
    EG_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_msg_count<kProg>), dim3(grid_for(R, kNT, 4096)), dim3(kNT), 0, s, R, (const uint32_t*)acc,
                       (const uint32_t*)g->e_row_dst.as<uint32_t>(), (const uint32_t*)g->e_row_start.as<uint32_t>(),
                       (const uint32_t*)g->e_csc_src.as<uint32_t>(), (const uint16_t*)g->qmap.as<uint16_t>(), g->S,
                       visited, cnt);
    EG_HIP(hipGetLastError());
  }
  EG_HIP(hipMemcpyAsync(h_count, cnt, 8, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  return EG_OK;
}

}  // namespace
}  // namespace eg

// ================================================================ host ===
namespace eg {
namespace {

struct Scratch {
  uint32_t* flist;                 // frontier list, V entries
  uint32_t* foff;                  // prefix of the frontier degrees, V+1 entries
  uint32_t* wcnt;                  // per-word counts, words+1
  unsigned long long* counters;    // [0] new vertices, [1] their out-edges
  double* delta;                   // PageRank residuals, iterations+1
};

uint64_t words_of(uint32_t V) { return (uint64_t(V) >> 6) + 1; }

uint64_t scratch_bytes(uint32_t V, int iters) {
  return uint64_t(V) * 4 + 256 + (uint64_t(V) + 1) * 4 + 256 + (words_of(V) + 1) * 4 + 256 + 64 +
         uint64_t(iters + 2) * 8 + 256;
}

int get_scratch(eg_graph* g, int iters, Scratch* sc) {
  const uint64_t need = scratch_bytes(g->V, iters);
  if (g->e_scratch.bytes < need) EG_TRY(g->e_scratch.resize(need));
  char* p = g->e_scratch.as<char>();
  auto align = [](uint64_t x) { return (x + 255) & ~uint64_t(255); };
  uint64_t o = 0;
  sc->flist = reinterpret_cast<uint32_t*>(p + o);
  o = align(o + uint64_t(g->V) * 4);
  sc->foff = reinterpret_cast<uint32_t*>(p + o);
  o = align(o + (uint64_t(g->V) + 1) * 4);
  sc->wcnt = reinterpret_cast<uint32_t*>(p + o);
  o = align(o + (words_of(g->V) + 1) * 4);
  sc->counters = reinterpret_cast<unsigned long long*>(p + o);
  o = align(o + 64);
  sc->delta = reinterpret_cast<double*>(p + o);
  return EG_OK;
}

// Ascending list of the set bits of bm and the prefix of their out-degrees;
// returns the count and the total degree on the host.
// Frontier list and degree prefix of the set bits of bm.  known_n >= 0 is the
// caller's exact count of set bits (from the previous level end): then
// nothing is read back and *h_total is left unchanged (the caller knows it
// too); otherwise both counts are read back.
int make_frontier(eg_graph* g, const uint64_t* bm, Scratch& sc, uint32_t* h_n, uint32_t* h_total, hipStream_t s,
                  int64_t known_n = -1) {
  const uint64_t words = words_of(g->V);
  hipLaunchKernelGGL(k_bm_count, dim3(grid_for(words, kNT, 65536)), dim3(kNT), 0, s, bm, words, sc.wcnt);
  EG_HIP(hipGetLastError());
  EG_TRY(fill<uint32_t>(sc.wcnt + words, 1, 0u, s));
  EG_TRY(exclusive_scan_u32(sc.wcnt, sc.wcnt, words + 1, g->scan_tmp.as<uint32_t>(), s));
  hipLaunchKernelGGL(k_bm_list, dim3(grid_for(words, kNT, 65536)), dim3(kNT), 0, s, bm, words,
                     (const uint32_t*)sc.wcnt, sc.flist);
  EG_HIP(hipGetLastError());
  if (known_n >= 0) {
    *h_n = uint32_t(known_n);
  } else {
    EG_HIP(hipMemcpyAsync(h_n, sc.wcnt + words, 4, hipMemcpyDeviceToHost, s));
    EG_HIP(hipStreamSynchronize(s));
  }
  const uint32_t fn = *h_n;
  hipLaunchKernelGGL(k_front_deg, dim3(grid_for(uint64_t(fn) + 1, kNT, 65536)), dim3(kNT), 0, s,
                     (const uint32_t*)sc.flist, fn, (const uint32_t*)g->e_out_ptr.as<uint32_t>(), sc.foff);
  EG_HIP(hipGetLastError());
  EG_TRY(exclusive_scan_u32(sc.foff, sc.foff, uint64_t(fn) + 1, g->scan_tmp.as<uint32_t>(), s));
  if (known_n < 0) {
    EG_HIP(hipMemcpyAsync(h_total, sc.foff + fn, 4, hipMemcpyDeviceToHost, s));
    EG_HIP(hipStreamSynchronize(s));
  }
  return EG_OK;
}

template <class Op>
int expand(eg_graph* g, Scratch& sc, uint32_t fn, uint32_t total, Op op, hipStream_t s) {
  if (!fn || !total) return EG_OK;
  hipLaunchKernelGGL((k_expand<Op>), dim3(unsigned(ceil_div(total, kExpSlots))), dim3(kNT), 0, s,
                     (const uint32_t*)sc.flist, (const uint32_t*)sc.foff, fn,
                     (const uint32_t*)g->e_out_ptr.as<uint32_t>(), op);
  EG_HIP(hipGetLastError());
  return EG_OK;
}

// visited |= out (if given), depth of out = level (if given); returns the
// number of vertices in out and the sum of their out-degrees.
int level_end(eg_graph* g, const uint64_t* out, uint64_t* visited, uint32_t* depth, uint32_t level, Scratch& sc,
              uint64_t* h_nv, uint64_t* h_ne, hipStream_t s) {
  const uint64_t words = words_of(g->V);
  EG_TRY(fill<unsigned long long>(sc.counters, 2, 0ull, s));
  hipLaunchKernelGGL(k_level_end, dim3(grid_for(words, kNT, 1024)), dim3(kNT), 0, s, out, visited, words, depth, level,
                     (const uint32_t*)g->out_deg.as<uint32_t>(), sc.counters);
  EG_HIP(hipGetLastError());
  unsigned long long h[2];
  EG_HIP(hipMemcpyAsync(h, sc.counters, 16, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  if (h_nv) *h_nv = h[0];
  if (h_ne) *h_ne = h[1];
  return EG_OK;
}


// core/graph.hpp:1483-1498: rooted programs are sparse under the reference
// policy; undirected ones (and every program under auto) switch on the
// active out-edge count (:1491, R4).
bool choose_sparse(const eg_graph* g, int policy, uint64_t active_edges, bool rooted) {
  if (policy == EG_POLICY_PUSH) return true;
  if (policy == EG_POLICY_PULL) return false;
  if (policy == EG_POLICY_REFERENCE && rooted) return true;
  return active_edges < g->E / 20;
}

// The pull's edge arrays: the built CSC (plain) or the degree-ordered view.
struct PullView {
  uint64_t E, R;
  const uint32_t *src, *wordrank, *row_dst;
  const uint64_t* rowflag;
  int all_atomic;     // a destination has a row per sub-partition
  uint32_t hot;       // ordered view: FP32C values of sources [0, hot) staged in LDS (0: no staging)
  uint32_t hot_grid;  // ... by a persistent grid of this many blocks
};

PullView plain_view(const eg_graph* g) {
  return PullView{g->E, g->R, g->e_csc_src.as<uint32_t>(), g->e_wordrank.as<uint32_t>(), g->e_row_dst.as<uint32_t>(),
                  g->e_rowflag.as<uint64_t>(), g->Q > 1, 0, 0};
}

PullView ordered_view(const eg_graph* g) {  // one row per destination, relabelled
  const bool sorted = g->pv_form == kViewSorted;
  const DevBuf& wr = sorted ? g->pv_wordrank : g->e_wordrank;
  const DevBuf& rf = sorted ? g->pv_rowflag : g->e_rowflag;
  return PullView{g->E, g->pv_R, g->pv_src.as<uint32_t>(), wr.as<uint32_t>(), g->pv_row_dst.as<uint32_t>(),
                  rf.as<uint64_t>(), 0, g->pv_hot, g->pv_hot_grid};
}

int pr_pull(const PullView& pv, uint32_t V, int precision, const void* x, double* next, hipStream_t s) {
  EG_TRY(fill<double>(next, V, 0.0, s));
  if (!pv.E) return EG_OK;
  prof_begin(s);
  if (precision == EG_PR_FP64) {
    hipLaunchKernelGGL(k_pr_pull_f64, dim3(row_grid(pv.E)), dim3(kNT), 0, s, pv.E, pv.src, pv.rowflag, pv.wordrank,
                       pv.row_dst, static_cast<const double*>(x), next, pv.all_atomic);
  } else if (pv.hot) {
    const uint64_t nranges = ceil_div(ceil_div(pv.E, 64), kRowWordsPerWave);
    hipLaunchKernelGGL(k_pr_pull_f32c_hot, dim3(pv.hot_grid), dim3(kHotNT), size_t(pv.hot) * 4, s, pv.E, pv.src,
                       pv.rowflag, pv.wordrank, pv.row_dst, static_cast<const float*>(x), next, pv.all_atomic,
                       pv.hot, nranges);
  } else {
    hipLaunchKernelGGL(k_pr_pull_f32c, dim3(row_grid(pv.E)), dim3(kNT), 0, s, pv.E, pv.src, pv.rowflag, pv.wordrank,
                       pv.row_dst, static_cast<const float*>(x), next, pv.all_atomic);
  }
  // algorithmic bytes: source id + gathered value per edge, row start bits,
  // destination id + partial per row
  const uint64_t xb = precision == EG_PR_FP64 ? 8 : 4;
  prof_end(PK_PR_PULL, pv.E * (4 + xb) + pv.E / 8 + pv.R * (4 + 8), s);
  EG_HIP(hipGetLastError());
  return EG_OK;
}

}  // namespace

int engine_ensure(eg_graph* g, hipStream_t s) { return ensure_engine(g, s); }

}  // namespace eg

using namespace eg;

extern "C" {

int eg_engine_prepare(eg_graph* g, double* h_ms, void* stream) {
  EG_TRY(ensure_engine(g, static_cast<hipStream_t>(stream)));
  if (h_ms) *h_ms = g->eng_prepare_ms;
  return EG_OK;
}

// ---- device helpers ----
int eg_device_fill(void* d_ptr, uint64_t n, const void* h_value, int value_bytes, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!d_ptr || !h_value) return fail(EG_ERR_ARG, "eg_device_fill: NULL argument");
  switch (value_bytes) {
    case 1: return fill<uint8_t>(static_cast<uint8_t*>(d_ptr), n, *static_cast<const uint8_t*>(h_value), s);
    case 2: return fill<uint16_t>(static_cast<uint16_t*>(d_ptr), n, *static_cast<const uint16_t*>(h_value), s);
    case 4: return fill<uint32_t>(static_cast<uint32_t*>(d_ptr), n, *static_cast<const uint32_t*>(h_value), s);
    case 8: return fill<uint64_t>(static_cast<uint64_t*>(d_ptr), n, *static_cast<const uint64_t*>(h_value), s);
    default: return fail(EG_ERR_ARG, "eg_device_fill: value_bytes must be 1, 2, 4 or 8");
  }
}

int eg_device_iota(uint32_t* d_ptr, uint32_t n, void* stream) {
  if (!d_ptr && n) return fail(EG_ERR_ARG, "eg_device_iota: NULL array");
  hipLaunchKernelGGL(k_iota, dim3(grid_for(n, kNT, 65536)), dim3(kNT), 0, static_cast<hipStream_t>(stream), d_ptr, n);
  EG_HIP(hipGetLastError());
  return EG_OK;
}

// ---- bitmaps (core/bitmap.hpp:20-68) ----
int eg_bitmap_or(uint64_t* d_dst, const uint64_t* d_src, uint32_t V, void* stream) {
  const uint64_t words = words_of(V);
  hipLaunchKernelGGL(k_bm_or, dim3(grid_for(words, kNT, 65536)), dim3(kNT), 0, static_cast<hipStream_t>(stream), d_dst,
                     d_src, words);
  EG_HIP(hipGetLastError());
  return EG_OK;
}

int eg_bitmap_count(const uint64_t* d_bm, uint32_t V, uint64_t* h_count, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!h_count) return fail(EG_ERR_ARG, "eg_bitmap_count: h_count is NULL");
  DevBuf c;
  EG_TRY(c.resize(16));
  EG_TRY(fill<unsigned long long>(c.as<unsigned long long>(), 2, 0ull, s));
  hipLaunchKernelGGL(k_level_end, dim3(grid_for(words_of(V), kNT, 8192)), dim3(kNT), 0, s, d_bm, nullptr, words_of(V),
                     nullptr, 0u, nullptr, c.as<unsigned long long>());
  EG_HIP(hipGetLastError());
  unsigned long long h[2];
  EG_HIP(hipMemcpyAsync(h, c.ptr, 16, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  *h_count = h[0];
  return EG_OK;
}

int eg_bitmap_degree_sum(eg_graph* g, const uint64_t* d_bm, uint64_t* h_sum, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  EG_TRY(ensure_engine(g, s));
  if (!h_sum) return fail(EG_ERR_ARG, "eg_bitmap_degree_sum: h_sum is NULL");
  Scratch sc;
  EG_TRY(get_scratch(g, 0, &sc));
  return level_end(g, d_bm, nullptr, nullptr, 0, sc, nullptr, h_sum, s);
}

int eg_bitmap_clear(uint64_t* d_bm, uint32_t V, void* stream) {
  return fill<uint64_t>(d_bm, words_of(V), 0ull, static_cast<hipStream_t>(stream));
}

int eg_bitmap_fill(uint64_t* d_bm, uint32_t V, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint64_t full = uint64_t(V) >> 6;
  EG_TRY(fill<uint64_t>(d_bm, full, ~0ull, s));
  const uint64_t tail = (V & 63) ? ((1ull << (V & 63)) - 1) : 0ull;  // bits [full*64, V)
  EG_TRY(fill<uint64_t>(d_bm + full, 1, tail, s));
  return EG_OK;
}

int eg_bitmap_set_bit(uint64_t* d_bm, uint32_t v, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  uint64_t w = 0;
  EG_HIP(hipMemcpyAsync(&w, d_bm + (v >> 6), 8, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  w |= 1ull << (v & 63);
  EG_HIP(hipMemcpyAsync(d_bm + (v >> 6), &w, 8, hipMemcpyHostToDevice, s));
  EG_HIP(hipStreamSynchronize(s));
  return EG_OK;
}

// ---- PageRank (toolkits/pagerank.cpp) ----
int eg_pr_init(eg_graph* g, double* d_x, float* d_x32, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  EG_TRY(ensure_engine(g, s));
  hipLaunchKernelGGL(k_pr_init, dim3(grid_for(g->V, kNT, 65536)), dim3(kNT), 0, s,
                     (const uint32_t*)g->out_deg.as<uint32_t>(), g->V, d_x, d_x32);
  EG_HIP(hipGetLastError());
  return EG_OK;
}

int eg_pr_pull(eg_graph* g, int precision, const void* d_x, double* d_next, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  EG_TRY(ensure_engine(g, s));
  if (precision != EG_PR_FP64 && precision != EG_PR_FP32C) return fail(EG_ERR_ARG, "unknown precision %d", precision);
  return pr_pull(plain_view(g), g->V, precision, d_x, d_next, s);
}

int eg_pr_apply(eg_graph* g, const double* d_x, double* d_next, float* d_next32, int last, double* d_delta,
                void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  EG_TRY(ensure_engine(g, s));
  if (!d_delta) return fail(EG_ERR_ARG, "d_delta is NULL");
  hipLaunchKernelGGL(k_pr_apply, dim3(grid_for(g->V, kNT, 4096)), dim3(kNT), 0, s,
                     (const uint32_t*)g->out_deg.as<uint32_t>(), g->V, d_x, d_next, d_next32, last, d_delta);
  EG_HIP(hipGetLastError());
  return EG_OK;
}

int eg_pagerank(eg_graph* g, int iterations, int precision, double* d_out, double* h_deltas, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  EG_TRY(ensure_engine(g, s));
  if (iterations < 1 || !d_out) return fail(EG_ERR_ARG, "need iterations >= 1 and an output array");
  const int prec = precision & 0xF;
  const int layout = precision & ~0xF;
  if (prec != EG_PR_FP64 && prec != EG_PR_FP32C) return fail(EG_ERR_ARG, "unknown precision %d", prec);
  if (layout != 0 && layout != EG_PR_ORDERED && layout != EG_PR_PLAIN && layout != EG_PR_SORTED)
    return fail(EG_ERR_ARG, "precision flags: at most one of EG_PR_ORDERED, EG_PR_PLAIN, EG_PR_SORTED");
  const bool ordered = layout == EG_PR_ORDERED || layout == EG_PR_SORTED || (layout == 0 && prec == EG_PR_FP32C);
  if (ordered) EG_TRY(ensure_view(g, view_form_of(g, layout), s));
  const PullView pv = ordered ? ordered_view(g) : plain_view(g);
  const uint32_t* deg = ordered ? g->pv_deg.as<uint32_t>() : g->out_deg.as<uint32_t>();
  Scratch sc;
  EG_TRY(get_scratch(g, iterations, &sc));
  const uint64_t V = g->V;
  const uint64_t vb = (V * 8 + 255) & ~255ull;
  EG_TRY(g->e_pr.resize(2 * vb + V * 4 + 256));
  // ordered: both FP64 vectors live in the view's ids, d_out gets the
  // original order at the end
  double* a = ordered ? g->e_pr.as<double>() : d_out;  // current x (FP64)
  double* b = reinterpret_cast<double*>(g->e_pr.as<char>() + (ordered ? vb : 0));  // next
  float* x32 = prec == EG_PR_FP32C ? reinterpret_cast<float*>(g->e_pr.as<char>() + 2 * vb) : nullptr;
  EG_TRY(fill<double>(sc.delta, uint64_t(iterations) + 1, 0.0, s));
  hipLaunchKernelGGL(k_pr_init, dim3(grid_for(V, kNT, 65536)), dim3(kNT), 0, s, deg, g->V, a, x32);
  EG_HIP(hipGetLastError());
  for (int it = 0; it < iterations; ++it) {
    EG_TRY(pr_pull(pv, g->V, prec, prec == EG_PR_FP64 ? static_cast<const void*>(a) : x32, b, s));
    hipLaunchKernelGGL(k_pr_apply, dim3(grid_for(V, kNT, 4096)), dim3(kNT), 0, s, deg, g->V, (const double*)a, b, x32,
                       it == iterations - 1 ? 1 : 0, sc.delta + it + 1);
    EG_HIP(hipGetLastError());
    std::swap(a, b);
  }
  if (ordered) {
    hipLaunchKernelGGL(k_gather_by, dim3(grid_for(V, kNT, 65536)), dim3(kNT), 0, s,
                       (const uint32_t*)g->pv_perm.as<uint32_t>(), g->V, (const double*)a, d_out);
    EG_HIP(hipGetLastError());
  } else if (a != d_out) {
    EG_HIP(hipMemcpyAsync(d_out, a, V * 8, hipMemcpyDeviceToDevice, s));
  }
  if (h_deltas) {
    EG_HIP(hipMemcpyAsync(h_deltas, sc.delta, sizeof(double) * (iterations + 1), hipMemcpyDeviceToHost, s));
    EG_HIP(hipStreamSynchronize(s));
    h_deltas[0] = 1.0;  // sum of 1 over all vertices / V (pagerank.cpp:36-47)
    for (int it = 1; it <= iterations; ++it) h_deltas[it] /= double(g->V);
  }
  EG_HIP(hipStreamSynchronize(s));
  return EG_OK;
}

int eg_prep_data(const eg_graph* g, uint64_t* h_bytes) {
  if (!g || !h_bytes) return fail(EG_ERR_ARG, "eg_prep_data: NULL argument");
  if (!g->built) return fail(EG_ERR_STATE, "graph has not been built");
  *h_bytes = g->prep_data;
  return EG_OK;
}

int eg_mode_sparse(const eg_graph* g, int policy, int program, uint64_t active_edges, int* h_sparse) {
  if (!g || !h_sparse) return fail(EG_ERR_ARG, "eg_mode_sparse: NULL argument");
  if (policy < EG_POLICY_REFERENCE || policy > EG_POLICY_PULL) return fail(EG_ERR_ARG, "unknown policy %d", policy);
  if (program < EG_PROGRAM_PAGERANK || program > EG_PROGRAM_SSSP) return fail(EG_ERR_ARG, "unknown program %d", program);
  *h_sparse = program == EG_PROGRAM_PAGERANK ? 0 : choose_sparse(g, policy, active_edges, program != EG_PROGRAM_CC);
  return EG_OK;
}

int eg_dense_messages(eg_graph* g, int program, const uint64_t* d_active, const uint64_t* d_visited,
                      const void* d_values, uint64_t* h_count, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  EG_TRY(ensure_engine(g, s));
  if (!h_count) return fail(EG_ERR_ARG, "eg_dense_messages: h_count is NULL");
  switch (program) {
    case EG_PROGRAM_PAGERANK:
      if (g->pr_cross_rows == ~0ull) EG_TRY(count_messages<EG_PROGRAM_PAGERANK>(g, nullptr, nullptr, nullptr,
                                                                               &g->pr_cross_rows, s));
      *h_count = g->pr_cross_rows;
      return EG_OK;
    case EG_PROGRAM_BFS:
      if (!d_active || !d_visited) return fail(EG_ERR_ARG, "BFS messages need the active and visited bitmaps");
      return count_messages<EG_PROGRAM_BFS>(g, d_active, d_visited, nullptr, h_count, s);
    case EG_PROGRAM_CC:
      if (!d_values) return fail(EG_ERR_ARG, "CC messages need the labels");
      return count_messages<EG_PROGRAM_CC>(g, nullptr, nullptr, d_values, h_count, s);
    case EG_PROGRAM_SSSP:
      if (g->rec != 12) return fail(EG_ERR_STATE, "SSSP needs a graph of EdgeType<float> (12-byte) records");
      if (!d_values) return fail(EG_ERR_ARG, "SSSP messages need the distances");
      return count_messages<EG_PROGRAM_SSSP>(g, nullptr, nullptr, d_values, h_count, s);
    default:
      return fail(EG_ERR_ARG, "unknown program %d", program);
  }
}

int eg_pagerank_view_prepare(eg_graph* g, double* h_ms, void* stream) {
  return eg_pagerank_view_prepare_layout(g, EG_PR_ORDERED, h_ms, stream);
}

int eg_pagerank_view_prepare_layout(eg_graph* g, int layout, double* h_ms, void* stream) {
  if (!g) return fail(EG_ERR_ARG, "graph is NULL");
  if (layout != EG_PR_ORDERED && layout != EG_PR_SORTED)
    return fail(EG_ERR_ARG, "view layout must be EG_PR_ORDERED or EG_PR_SORTED");
  EG_TRY(ensure_view(g, view_form_of(g, layout), static_cast<hipStream_t>(stream)));
  if (h_ms) *h_ms = g->pv_prepare_ms;
  return EG_OK;
}

int eg_pagerank_view_release(eg_graph* g) {
  if (!g) return fail(EG_ERR_ARG, "graph is NULL");
  EG_HIP(hipDeviceSynchronize());  // a run still queued may read the view
  g->pv_ready = false;
  g->pv_form = kViewNone;
  g->pv_src.release();
  g->pv_row_dst.release();
  g->pv_rowflag.release();
  g->pv_wordrank.release();
  g->pv_perm.release();
  g->pv_inv.release();
  g->pv_deg.release();
  return EG_OK;
}

int eg_traversal_view_prepare(eg_graph* g, double* h_ms, void* stream) {
  if (!g) return fail(EG_ERR_ARG, "graph is NULL");
  EG_TRY(ensure_tview(g, static_cast<hipStream_t>(stream)));
  if (h_ms) *h_ms = g->tv_prepare_ms;
  return EG_OK;
}

}  // extern "C"

namespace eg {
// ---- BFS (toolkits/bfs.cpp) ----
// One level.  frontier_n >= 0: the exact frontier size, and frontier_edges
// its exact out-degree sum (both from the previous level end, as in eg_bfs's
// loop), so the push step reads nothing back.
int bfs_level(eg_graph* g, int policy, uint32_t level, const uint64_t* d_frontier, uint64_t* d_next,
              uint64_t* d_visited, uint32_t* d_parent, uint32_t* d_depth, uint64_t frontier_edges, int64_t frontier_n,
              uint64_t* h_new_vertices, uint64_t* h_new_edges, int* h_was_sparse, hipStream_t s) {
  EG_TRY(ensure_engine(g, s));
  if (policy < EG_POLICY_REFERENCE || policy > EG_POLICY_PULL) return fail(EG_ERR_ARG, "unknown policy %d", policy);
  Scratch sc;
  EG_TRY(get_scratch(g, 0, &sc));
  const bool sparse = choose_sparse(g, policy, frontier_edges, true);
  EG_TRY(fill<uint64_t>(d_next, words_of(g->V), 0ull, s));
  prof_begin(s);
  if (sparse) {
    uint32_t fn = 0, total = uint32_t(frontier_edges);
    EG_TRY(make_frontier(g, d_frontier, sc, &fn, &total, s, frontier_n));
    EG_TRY(expand(g, sc, fn, total,
                  BfsPushOp{g->e_out_dst.as<uint32_t>(), (const uint64_t*)d_visited, d_parent, d_next}, s));
    prof_end(PK_BFS, uint64_t(total) * 4 + uint64_t(fn) * 8, s);
  } else if (g->E) {
    EG_TRY(ensure_tview(g, s));
    const uint32_t* in_ptr = g->tv_ptr.as<uint32_t>();
    const uint32_t* in_src = g->tv_src.as<uint32_t>();
    hipLaunchKernelGGL(k_bfs_pull_rows, dim3(grid_for(g->V, kNT, 65536)), dim3(kNT), 0, s, g->V, in_ptr, in_src,
                       d_frontier, (const uint64_t*)d_visited, d_parent, d_next);
    EG_HIP(hipGetLastError());
    // first chunks of the large rows, then the others, which an earlier
    // chunk's hit (always smaller: sources ascend) lets skip
    if (g->tv_nbig)
      hipLaunchKernelGGL(k_bfs_pull_items<true>, dim3(grid_for(g->tv_nbig, kNT, 65536)), dim3(kNT), 0, s,
                         (const uint64_t*)g->tv_big.as<uint64_t>(), g->tv_nbig, in_ptr, in_src, d_frontier,
                         (const uint64_t*)d_visited, d_parent, d_next);
    if (g->tv_nbig_rest)
      hipLaunchKernelGGL(k_bfs_pull_items<false>, dim3(grid_for(g->tv_nbig_rest, kNT, 65536)), dim3(kNT), 0, s,
                         (const uint64_t*)g->tv_big_rest.as<uint64_t>(), g->tv_nbig_rest, in_ptr, in_src, d_frontier,
                         (const uint64_t*)d_visited, d_parent, d_next);
    prof_end(PK_BFS, uint64_t(g->V) * 12, s);
  }
  EG_HIP(hipGetLastError());
  if (h_was_sparse) *h_was_sparse = sparse ? 1 : 0;
  return level_end(g, d_next, d_visited, d_depth, level + 1, sc, h_new_vertices, h_new_edges, s);
}
}  // namespace eg

extern "C" {

int eg_bfs_level(eg_graph* g, int policy, uint32_t level, const uint64_t* d_frontier, uint64_t* d_next,
                 uint64_t* d_visited, uint32_t* d_parent, uint32_t* d_depth, uint64_t frontier_edges,
                 uint64_t* h_new_vertices, uint64_t* h_new_edges, int* h_was_sparse, void* stream) {
  return bfs_level(g, policy, level, d_frontier, d_next, d_visited, d_parent, d_depth, frontier_edges, -1,
                   h_new_vertices, h_new_edges, h_was_sparse, static_cast<hipStream_t>(stream));
}

int eg_bfs(eg_graph* g, uint32_t root, int policy, uint32_t* d_parent, uint32_t* d_depth, uint32_t* h_levels,
           uint64_t* h_edges, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  EG_TRY(ensure_engine(g, s));
  if (root >= g->V) return fail(EG_ERR_ARG, "root %u >= vertices_num", root);
  if (!d_parent) return fail(EG_ERR_ARG, "d_parent is NULL");
  if (policy < EG_POLICY_REFERENCE || policy > EG_POLICY_PULL) return fail(EG_ERR_ARG, "unknown policy %d", policy);
  const uint32_t V = g->V;
  const bool pulls = policy == EG_POLICY_AUTO || policy == EG_POLICY_PULL;
  if (pulls) EG_TRY(ensure_tview(g, s));
  // state: frontier / next bitmaps, visited, the counter ring and scalars,
  // and depths when the caller keeps none
  const uint64_t words = words_of(V);
  const uint64_t bytes = 3 * words * 8 + 64 + 3 * sizeof(SCount) + 64 + (d_depth ? 0 : uint64_t(V) * 4);
  EG_TRY(g->e_sbfs.resize(bytes));
  SbfsArgs a{};  // zeroed padding: the bytes key the captured step graph
  SbfsState& st = a.st;
  st.bm[0] = g->e_sbfs.as<uint64_t>();
  st.bm[1] = st.bm[0] + words;
  st.vis = st.bm[1] + words;
  st.cnt = reinterpret_cast<SCount*>((reinterpret_cast<uintptr_t>(st.vis + words) + 63) & ~uintptr_t(63));
  st.total = reinterpret_cast<unsigned long long*>(st.cnt + 3);
  st.steps = reinterpret_cast<uint32_t*>(st.total + 1);
  st.base = st.steps + 1;
  uint32_t* depth = d_depth ? d_depth : reinterpret_cast<uint32_t*>(st.total + 8);
  a.V = V;
  a.E = g->E;
  a.policy = policy;
  a.out_ptr = g->e_out_ptr.as<uint32_t>();
  a.out_dst = g->e_out_dst.as<uint32_t>();
  a.deg = g->out_deg.as<uint32_t>();
  a.deg8 = g->e_deg8.as<uint8_t>();
  EG_TRY(ensure_hub_items(g, s));
  a.hub_items = g->e_hub_items.as<uint64_t>();
  a.n_hub_items = g->e_nhub_items;
  a.dense_div = g->tune.sbfs_div;
  a.hub = g->tune.hub;
  a.hub_chunk = g->tune.hub_chunk;
  if (pulls) {
    a.in_ptr = g->tv_ptr.as<uint32_t>();
    a.in_src = g->tv_src.as<uint32_t>();
    a.big = g->tv_big.as<uint64_t>();
    a.big_rest = g->tv_big_rest.as<uint64_t>();
    a.n_big = g->tv_nbig;
    a.n_big_rest = g->tv_nbig_rest;
  }
  a.parent = d_parent;
  a.depth = depth;
  hipLaunchKernelGGL(k_sbfs_init, dim3(grid_for(std::max<uint64_t>(V, 3 * words), kNT, 65536)), dim3(kNT), 0, s, a,
                     root);
  EG_HIP(hipGetLastError());
  const StepLaunch<SbfsArgs> ls[5] = {{k_sbfs_push, 2048},
                                      {k_sbfs_pull_rows, grid_for(V, kNT, 16384)},
                                      {k_sbfs_pull_items<0>, grid_for(a.n_big, kNT, 65536)},
                                      {k_sbfs_pull_items<1>, grid_for(a.n_big_rest, kNT, 65536)},
                                      {k_sbfs_end, grid_for(ceil_div(words, uint64_t(kEndWords)), kNT / kWave, 2048)}};
  const unsigned long long* const nv[3] = {&st.cnt[0].nv, &st.cnt[1].nv, &st.cnt[2].nv};
  uint32_t steps = 0;
  prof_begin(s);
  unsigned long long total = 0;
  EG_TRY(run_step_graph(g->multi_graph, a, ls, st.base, nv, st.steps, V + 1, &steps, s, st.total, &total));
  prof_end(PK_BFS, 0, s);
  if (h_levels) *h_levels = steps;
  if (h_edges) *h_edges = total;
  return EG_OK;
}

// ---- CC (toolkits/cc.cpp) ----
int eg_cc_iteration(eg_graph* g, int policy, const uint64_t* d_active, uint64_t* d_next, uint32_t* d_label,
                    uint64_t active_edges, uint64_t* h_new_active, uint64_t* h_new_edges, int* h_was_sparse,
                    void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  EG_TRY(ensure_engine(g, s));
  if (policy < EG_POLICY_REFERENCE || policy > EG_POLICY_PULL) return fail(EG_ERR_ARG, "unknown policy %d", policy);
  Scratch sc;
  EG_TRY(get_scratch(g, 0, &sc));
  EG_TRY(fill<uint64_t>(d_next, words_of(g->V), 0ull, s));
  const bool sparse = choose_sparse(g, policy, active_edges, false);
  if (h_was_sparse) *h_was_sparse = sparse ? 1 : 0;
  if (sparse) {
    uint32_t fn = 0, total = 0;
    EG_TRY(make_frontier(g, d_active, sc, &fn, &total, s));
    EG_TRY(expand(g, sc, fn, total, CcPushOp{g->e_out_dst.as<uint32_t>(), d_label, d_next}, s));
  } else if (g->E) {
    hipLaunchKernelGGL(k_cc_pull, dim3(row_grid(g->E)), dim3(kNT), 0, s, g->E,
                       (const uint32_t*)g->e_csc_src.as<uint32_t>(), (const uint64_t*)g->e_rowflag.as<uint64_t>(),
                       (const uint32_t*)g->e_wordrank.as<uint32_t>(), (const uint32_t*)g->e_row_dst.as<uint32_t>(),
                       d_label, d_next);
  }
  EG_HIP(hipGetLastError());
  return level_end(g, d_next, nullptr, nullptr, 0, sc, h_new_active, h_new_edges, s);
}

int eg_cc(eg_graph* g, int policy, uint32_t* d_label, uint32_t* h_iterations, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  EG_TRY(ensure_engine(g, s));
  if (!d_label) return fail(EG_ERR_ARG, "d_label is NULL");
  uint64_t* a = g->e_bm[0].as<uint64_t>();
  uint64_t* b = g->e_bm[1].as<uint64_t>();
  hipLaunchKernelGGL(k_iota, dim3(grid_for(g->V, kNT, 65536)), dim3(kNT), 0, s, d_label, g->V);
  EG_HIP(hipGetLastError());
  EG_TRY(eg_bitmap_fill(a, g->V, stream));
  uint64_t nv = g->V, ne = g->E;
  uint32_t it = 0;
  while (nv > 0) {
    EG_TRY(eg_cc_iteration(g, policy, a, b, d_label, ne, &nv, &ne, nullptr, stream));
    std::swap(a, b);
    ++it;
  }
  if (h_iterations) *h_iterations = it;
  return EG_OK;
}

// ---- SSSP (toolkits/sssp.cpp) ----
int eg_sssp_iteration(eg_graph* g, int policy, const uint64_t* d_active, uint64_t* d_next, float* d_dist,
                      uint64_t active_edges, uint64_t* h_new_active, uint64_t* h_new_edges, int* h_was_sparse,
                      void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  EG_TRY(ensure_engine(g, s));
  if (g->rec != 12) return fail(EG_ERR_STATE, "SSSP needs a graph of EdgeType<float> (12-byte) records");
  if (policy < EG_POLICY_REFERENCE || policy > EG_POLICY_PULL) return fail(EG_ERR_ARG, "unknown policy %d", policy);
  Scratch sc;
  EG_TRY(get_scratch(g, 0, &sc));
  EG_TRY(fill<uint64_t>(d_next, words_of(g->V), 0ull, s));
  const bool sparse = choose_sparse(g, policy, active_edges, true);
  if (h_was_sparse) *h_was_sparse = sparse ? 1 : 0;
  if (sparse) {
    uint32_t fn = 0, total = 0;
    EG_TRY(make_frontier(g, d_active, sc, &fn, &total, s));
    EG_TRY(expand(g, sc, fn, total,
                  SsspPushOp{g->e_out_dst.as<uint32_t>(), g->e_out_w.as<float>(), d_dist, d_next}, s));
  } else if (g->E) {
    hipLaunchKernelGGL(k_sssp_pull, dim3(row_grid(g->E)), dim3(kNT), 0, s, g->E,
                       (const uint32_t*)g->e_csc_src.as<uint32_t>(), (const float*)g->e_csc_w.as<float>(),
                       (const uint64_t*)g->e_rowflag.as<uint64_t>(), (const uint32_t*)g->e_wordrank.as<uint32_t>(),
                       (const uint32_t*)g->e_row_dst.as<uint32_t>(), d_dist, d_next);
  }
  EG_HIP(hipGetLastError());
  return level_end(g, d_next, nullptr, nullptr, 0, sc, h_new_active, h_new_edges, s);
}

int eg_sssp(eg_graph* g, uint32_t root, int policy, float* d_dist, uint32_t* h_iterations, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  EG_TRY(ensure_engine(g, s));
  if (g->rec != 12) return fail(EG_ERR_STATE, "SSSP needs a graph of EdgeType<float> (12-byte) records");
  if (root >= g->V || !d_dist) return fail(EG_ERR_ARG, "bad root or NULL d_dist");
  if (policy < EG_POLICY_REFERENCE || policy > EG_POLICY_PULL) return fail(EG_ERR_ARG, "unknown policy %d", policy);
  const uint32_t V = g->V;
  const uint64_t words = words_of(V);
  const uint64_t nsum = summary_words(V);
  EG_TRY(g->e_dsssp.resize(2 * words * 8 + 64 + 3 * sizeof(SCount) + 64 + 2 * nsum * 8));
  DssspArgs a{};  // zeroed padding: the bytes key the captured step graph
  SbfsState& st = a.st;
  st.bm[0] = g->e_dsssp.as<uint64_t>();
  st.bm[1] = st.bm[0] + words;
  st.cnt = reinterpret_cast<SCount*>((reinterpret_cast<uintptr_t>(st.bm[1] + words) + 63) & ~uintptr_t(63));
  st.steps = reinterpret_cast<uint32_t*>(st.cnt + 3);
  st.base = st.steps + 1;
  a.sum[0] = reinterpret_cast<uint64_t*>(st.cnt + 3) + 8;  // after steps and base, 64 bytes on
  a.sum[1] = a.sum[0] + nsum;
  a.nsum = uint32_t(nsum);
  a.V = V;
  a.E = g->E;
  a.policy = policy;
  a.out_ptr = g->e_out_ptr.as<uint32_t>();
  a.out_dst = g->e_out_dst.as<uint32_t>();
  a.out_w = g->e_out_w.as<float>();
  a.deg = g->out_deg.as<uint32_t>();
  a.deg8 = g->e_deg8.as<uint8_t>();
  EG_TRY(ensure_hub_items(g, s));
  a.hub_items = g->e_hub_items.as<uint64_t>();
  a.n_hub_items = g->e_nhub_items;
  a.dense_div = g->tune.dsssp_div;
  a.hub = g->tune.hub;
  a.hub_chunk = g->tune.hub_chunk;
  a.csc_src = g->e_csc_src.as<uint32_t>();
  a.csc_w = g->e_csc_w.as<float>();
  a.rowflag = g->e_rowflag.as<uint64_t>();
  a.wordrank = g->e_wordrank.as<uint32_t>();
  a.row_dst = g->e_row_dst.as<uint32_t>();
  a.dist = d_dist;
  hipLaunchKernelGGL(k_dsssp_init, dim3(grid_for(std::max<uint64_t>(V, 2 * words), kNT, 65536)), dim3(kNT), 0, s, a,
                     root);
  EG_HIP(hipGetLastError());
  const StepLaunch<DssspArgs> ls[3] = {{k_dsssp_push, 2048},
                                       {k_dsssp_pull, row_grid(g->E), nsum <= kSumMax ? unsigned(nsum * 8) : 0u},
                                       {k_dsssp_end, grid_for(ceil_div(words, uint64_t(kEndWords)), kNT / kWave, 2048)}};
  const unsigned long long* const nv[3] = {&st.cnt[0].nv, &st.cnt[1].nv, &st.cnt[2].nv};
  uint32_t steps = 0;
  EG_TRY(run_step_graph(g->sssp_graph, a, ls, st.base, nv, st.steps, V + 1, &steps, s));
  if (h_iterations) *h_iterations = steps;
  return EG_OK;
}

}  // extern "C"
