// CPU restatement of EndGraph's preprocessing and Gemini's edge-processing
// semantics.  TEST INFRASTRUCTURE ONLY: it is the parity checker for the HIP
// library and the timed host-CPU baseline ("port").  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the product
// path (endgraph_amd/) never does.
//
// Every function cites the reference file:line it follows.  Deviations from
// the reference at HEAD (which does not compile) are the repairs R1..R12:
//   R2  page weight takes an explicit multiplier m instead of the undeclared
//       `dense` / late-set `max_iterations` (core/graph.hpp:733-738).
//   R4  mode switch uses `degree` / `edges_num` (core/graph.hpp:1487,1491).
//   R5  both CSR and CSC are built; `degree` is the out-degree (keyed on src)
//       and the in-degree is kept too (core/graph.hpp:697-706, 986-1007).
//   R6  the dense thread schedule (tuned_chunks_dense) is not reproduced; it
//       does not affect results.
//   R8  a socket with zero edges gets an index holding only the sentinel
//       (core/graph.hpp:1154, 1331 read adj_list[0] unguarded).
//   R9  page boundaries are clamped to the vertex range
//       (core/graph.hpp:749, 778).
//   R10 64-bit counts everywhere (core/graph.hpp:560,562 use int).
//   R12 BFS parent = minimum-id frontier in-neighbour (the reference's CAS
//       race, toolkits/bfs.cpp:54, leaves it unspecified).
//
// Parity pins (tests/test_oracle_*.py): degree vs numpy bincount; CPBT
// bottleneck vs brute-force DP and greedy maximality; CSR/CSC order vs a
// numpy lexsort on the closed-form key tuple; BFS depth vs a textbook queue
// BFS and parents vs the min-id rule computed from depths; CC vs scipy
// connected components; SSSP vs a float32 Dijkstra; PageRank vs a dense FP64
// power iteration of the toolkit formula.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include <omp.h>

namespace {

constexpr uint32_t kPageSize = 4096;  // core/constants.hpp:21 PAGESIZE
constexpr int kMaxRadix = 8;          // core/graph.hpp:424 MAX_RADIX

struct Rec8 {  // core/type.hpp:33-40 EdgeType<Empty>, 8 bytes packed
  uint32_t src, dst;
};
struct Rec12 {  // core/type.hpp:26-31 EdgeType<float>, 12 bytes packed
  uint32_t src, dst;
  float w;
};
static_assert(sizeof(Rec8) == 8, "EdgeType<Empty> is 8 bytes");
static_assert(sizeof(Rec12) == 12, "EdgeType<float> is 12 bytes");

#pragma pack(push, 1)
struct CIdx {  // core/type.hpp:54-59 CompressedAdjIndexType, 12 bytes packed
  uint64_t index;
  uint32_t vertex;
};
#pragma pack(pop)
static_assert(sizeof(CIdx) == 12, "CompressedAdjIndexType is 12 bytes");

// core/graph.hpp:1087-1092
int counting_bits(uint64_t n) {
  int i = 0;
  for (; n > 0; n >>= 1) i++;
  return i;
}

// Stable LSD radix sort, restating load_balance_radix_sort and
// serial_counting_sort (core/graph.hpp:437-584): per round every thread
// counting-sorts its contiguous chunk into a private buffer, then buckets are
// concatenated in thread order.  Returns the last round's bucket offsets
// (buckets+1 entries).  64-bit counts (R10).
template <class E, class F>
std::vector<uint64_t> lb_radix_sort(E* din, uint64_t length, int bits, F getKey,
                                    int threads) {
  int rounds = 1 + (bits - 1) / kMaxRadix;
  int rbits = 1 + (bits - 1) / rounds;
  int bit_offset = 0;
  uint64_t t_length = (length + threads - 1) / threads;
  std::vector<std::vector<E>> t_out(threads);
  std::vector<std::vector<uint8_t>> t_tmp(threads);
  std::vector<std::vector<uint64_t>> t_counts(threads), t_offsets(threads);
  for (int t = 0; t < threads; ++t) {
    t_out[t].resize(t_length);
    t_tmp[t].resize(t_length);
    t_counts[t].resize(size_t(1) << rbits);
    t_offsets[t].resize(size_t(1) << rbits);
  }
  std::vector<uint64_t> g_offsets;
  for (int round = 0; round < rounds; ++round, bit_offset += rbits) {
    if (bit_offset + rbits > bits) rbits = bits - bit_offset;  // :515-516
    const int buckets = 1 << rbits;
    const uint32_t mask = uint32_t(buckets) - 1;
    const int shift = bit_offset;
#pragma omp parallel for num_threads(threads) schedule(static, 1)
    for (int t = 0; t < threads; ++t) {
      uint64_t begin = std::min<uint64_t>(length, uint64_t(t) * t_length);
      uint64_t end = std::min<uint64_t>(length, uint64_t(t + 1) * t_length);
      uint64_t n = end - begin;
      E* in = din + begin;
      uint64_t* counts = t_counts[t].data();
      uint64_t* offs = t_offsets[t].data();
      uint8_t* tmp = t_tmp[t].data();
      E* out = t_out[t].data();
      // serial_counting_sort (:437-456)
      for (int b = 0; b < buckets; ++b) counts[b] = 0;
      for (uint64_t j = 0; j < n; ++j) {
        uint32_t k = tmp[j] = uint8_t(mask & (uint32_t(getKey(in[j])) >> shift));
        counts[k]++;
      }
      uint64_t s = 0;
      for (int b = 0; b < buckets; ++b) {
        offs[b] = s;
        s += counts[b];
      }
      for (uint64_t j = 0; j < n; ++j) out[offs[tmp[j]]++] = in[j];
      // :531-534 rewind offsets to bucket starts
      for (int b = buckets - 1; b > 0; --b) offs[b] = offs[b - 1];
      offs[0] = 0;
    }
    std::vector<uint64_t> g_counts(buckets, 0);
    for (int b = 0; b < buckets; ++b)
      for (int t = 0; t < threads; ++t) g_counts[b] += t_counts[t][b];
    g_offsets.assign(buckets + 1, 0);
    for (int b = 1; b <= buckets; ++b) g_offsets[b] = g_offsets[b - 1] + g_counts[b - 1];
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
    for (int b = 0; b < buckets; ++b) {  // :558-569
      uint64_t bj = g_offsets[b];
      for (int t = 0; t < threads; ++t) {
        const E* src = t_out[t].data() + t_offsets[t][b];
        uint64_t c = t_counts[t][b];
        std::memcpy(din + bj, src, c * sizeof(E));
        bj += c;
      }
    }
  }
  return g_offsets;
}

// judge / optimal_offset: core/graph.hpp:586-626.
bool judge(uint64_t s, const uint64_t* prefix, uint32_t n, uint32_t k) {
  uint64_t sum = 0;
  for (uint32_t i = 0; i < k; ++i)
    sum = *(std::upper_bound(prefix, prefix + n + 1, sum + s) - 1);
  return sum == prefix[n];
}

std::vector<uint32_t> optimal_offset(const uint64_t* a, uint32_t n, uint32_t k) {
  std::vector<uint64_t> prefix(n + 1);
  prefix[0] = 0;
  for (uint32_t i = 0; i < n; ++i) prefix[i + 1] = prefix[i] + a[i];
  uint64_t lb = prefix[0], ub = prefix[n];
  while (lb + 1 < ub) {
    uint64_t mid = lb + ((ub - lb) >> 1);
    if (judge(mid, prefix.data(), n, k))
      ub = mid;
    else
      lb = mid;
  }
  std::vector<uint32_t> offset(k + 1);
  offset[0] = 0;
  offset[k] = n;
  uint64_t sum = 0;
  for (uint32_t i = 0; i + 1 < k; ++i) {
    const uint64_t* ptr = std::upper_bound(prefix.data(), prefix.data() + n + 1, sum + ub) - 1;
    offset[i + 1] = uint32_t(ptr - prefix.data());
    sum = *ptr;
  }
  return offset;
}

struct Graph {
  uint32_t V = 0;
  int P = 1, S = 1;
  uint64_t eta = 0;
  uint32_t m = 1;
  int rec = 8;
  int symmetrize = 0;
  int threads = 1;
  uint64_t E = 0;  // edges after symmetrisation
  double build_seconds = 0;
  std::vector<uint32_t> out_deg, in_deg;
  std::vector<uint64_t> page_amount;
  std::vector<uint32_t> page_offset, partition_offset;
  std::vector<std::vector<uint32_t>> local_partition_offset;  // [P][S+1]
  // per (p,s), index q = p*S + s
  std::vector<std::vector<uint8_t>> csc_list, csr_list;
  std::vector<std::vector<CIdx>> csc_cindex;
  std::vector<std::vector<uint64_t>> csr_index;   // V+1, ~0 = undefined
  std::vector<std::vector<uint64_t>> csr_bitmap;  // (V>>6)+1 words
};

template <class E>
uint32_t sock_of(const Graph& g, int p, uint32_t v) {
  // get_local_partition_id (core/graph.hpp:415-422): linear scan, default 0
  const auto& lo = g.local_partition_offset[p];
  for (int s = 0; s < g.S; ++s)
    if (v >= lo[s] && v < lo[s + 1]) return uint32_t(s);
  return 0;
}

// One direction of the build.  pull=true builds the CSC (prepare_pull_mode,
// core/graph.hpp:1098-1242) from a shuffle keyed on src; pull=false builds
// the CSR (prepare_push_mode, core/graph.hpp:1248-1355) keyed on dst.
template <class E>
void build_direction(Graph& g, const std::vector<std::vector<E>>& locals, bool pull) {
  const int P = g.P, S = g.S, T = g.threads;
  const int vbits = counting_bits(g.V);
  auto owner_key = [pull](const E& e) { return pull ? e.src : e.dst; };
  auto sort_key = [pull](const E& e) { return pull ? e.dst : e.src; };
  // shuffle_edge_list (core/graph.hpp:798-973)
  std::vector<std::vector<E>> part_arr(P);
  if (P == 1) {
    part_arr[0] = locals[0];  // :809-813 identity
  } else {
    std::vector<std::vector<E>> sorted(P);
    for (int r = 0; r < P; ++r) {
      sorted[r] = locals[r];
      lb_radix_sort(sorted[r].data(), sorted[r].size(), vbits,
                    [&](const E& e) { return owner_key(e); }, T);  // :815
    }
    // bucket counts per (r, p) and concatenation by source rank (:819-960)
    std::vector<std::vector<uint64_t>> bstart(P, std::vector<uint64_t>(P + 1, 0));
    for (int r = 0; r < P; ++r) {
      const auto& arr = sorted[r];
      for (int p = 0; p <= P; ++p) {
        uint32_t bound = g.partition_offset[p];
        uint64_t pos = std::lower_bound(arr.begin(), arr.end(), bound,
                                        [&](const E& e, uint32_t b) { return owner_key(e) < b; }) -
                       arr.begin();
        bstart[r][p] = (p == P) ? arr.size() : pos;
      }
    }
    for (int p = 0; p < P; ++p) {
      uint64_t total = 0;
      for (int r = 0; r < P; ++r) total += bstart[r][p + 1] - bstart[r][p];
      part_arr[p].resize(total);
      uint64_t at = 0;
      for (int r = 0; r < P; ++r) {
        uint64_t c = bstart[r][p + 1] - bstart[r][p];
        std::memcpy(part_arr[p].data() + at, sorted[r].data() + bstart[r][p], c * sizeof(E));
        at += c;
      }
    }
  }
  for (int p = 0; p < P; ++p) {
    auto& arr = part_arr[p];
    lb_radix_sort(arr.data(), arr.size(), vbits, [&](const E& e) { return sort_key(e); }, T);
    std::vector<uint64_t> goff =
        lb_radix_sort(arr.data(), arr.size(), counting_bits(uint64_t(S)),
                      [&](const E& e) { return sock_of<E>(g, p, owner_key(e)); }, T);
    for (int s = 0; s < S; ++s) {
      const int q = p * S + s;
      uint64_t b = goff[s], n = goff[s + 1] - goff[s];
      std::vector<uint8_t> bytes(n * sizeof(E));
      std::memcpy(bytes.data(), arr.data() + b, n * sizeof(E));
      const E* adj = reinterpret_cast<const E*>(bytes.data());
      if (pull) {
        // compressed incoming index (:1147-1188); R8 guard for n == 0
        std::vector<CIdx> idx;
        if (n > 0) {
          uint32_t pre = adj[0].dst;
          idx.push_back({0, pre});
          for (uint64_t i = 1; i < n; ++i)
            if (adj[i].dst != pre) {
              pre = adj[i].dst;
              idx.push_back({i, pre});
            }
        }
        idx.push_back({n, g.V});
        g.csc_cindex[q] = std::move(idx);
        g.csc_list[q] = std::move(bytes);
      } else {
        // dense outgoing index + bitmap (:1312-1353); R8 guard for n == 0
        std::vector<uint64_t> index(size_t(g.V) + 1, ~uint64_t(0));
        std::vector<uint64_t> bm((g.V >> 6) + 1, 0);
        if (n > 0) {
          uint32_t pre = adj[0].src;
          bm[pre >> 6] |= 1ull << (pre & 63);
          index[pre] = 0;
          for (uint64_t i = 1; i < n; ++i) {
            uint32_t src = adj[i].src;
            if (pre != src) {
              bm[src >> 6] |= 1ull << (src & 63);
              index[src] = i;
              index[size_t(pre) + 1] = i;
              pre = src;
            }
          }
          index[size_t(pre) + 1] = n;
        }
        g.csr_index[q] = std::move(index);
        g.csr_bitmap[q] = std::move(bm);
        g.csr_list[q] = std::move(bytes);
      }
    }
  }
}

template <class E>
int build_impl(Graph& g, const E* edges, uint64_t E0) {
  const int P = g.P, S = g.S;
  // preload_graph_file (core/graph.hpp:1011-1084): rank slices, per-rank
  // symmetrisation.  Excluded from prep_time in the reference (:1365-1371).
  std::vector<std::vector<E>> locals(P);
  for (int r = 0; r < P; ++r) {
    uint64_t n = E0 / P + (r == P - 1 ? E0 % P : 0);
    uint64_t off = (E0 / P) * uint64_t(r);
    auto& L = locals[r];
    L.resize(g.symmetrize ? 2 * n : n);
    std::memcpy(L.data(), edges + off, n * sizeof(E));
    if (g.symmetrize) {
      for (uint64_t i = 0; i < n; ++i) {
        L[i + n] = L[i];
        L[i + n].src = L[i].dst;
        L[i + n].dst = L[i].src;
      }
    }
  }
  auto t0 = std::chrono::steady_clock::now();
  g.E = 0;
  for (auto& L : locals) g.E += L.size();
  for (auto& L : locals)
    for (const E& e : L)
      if (e.src >= g.V || e.dst >= g.V) return -2;
  // assign_vertices (core/graph.hpp:662-792): degree histogram (R5: out-degree,
  // plus in-degree)
  g.out_deg.assign(g.V, 0);
  g.in_deg.assign(g.V, 0);
  for (int r = 0; r < P; ++r) {
    const auto& L = locals[r];
    const int64_t n = int64_t(L.size());
#pragma omp parallel for num_threads(g.threads)
    for (int64_t i = 0; i < n; ++i) {
      __atomic_fetch_add(&g.out_deg[L[i].src], 1u, __ATOMIC_RELAXED);
      __atomic_fetch_add(&g.in_deg[L[i].dst], 1u, __ATOMIC_RELAXED);
    }
  }
  // page weights (:718-741), R2: explicit m
  const uint32_t pages = (g.V + kPageSize - 1) / kPageSize;
  g.page_amount.assign(pages, 0);
  const uint32_t f1 = uint32_t(1 + (P - 1) / 16);
#pragma omp parallel for num_threads(g.threads)
  for (int64_t pg = 0; pg < int64_t(pages); ++pg) {
    uint64_t acc = 0;
    uint32_t vb = uint32_t(pg) * kPageSize, ve = std::min<uint32_t>(g.V, vb + kPageSize);
    for (uint32_t v = vb; v < ve; ++v) {
      uint32_t d = g.out_deg[v];
      // (1 + (P-1)/16) * degree : 32-bit unsigned; eta * (P-1 + m*degree):
      // the bracket is 32-bit unsigned, the product and sum 64-bit.
      uint32_t a = f1 * d;
      uint32_t b = uint32_t(P - 1) + g.m * d;
      acc += uint64_t(a) + g.eta * uint64_t(b);
    }
    g.page_amount[pg] = acc;
  }
  g.page_offset = optimal_offset(g.page_amount.data(), pages, uint32_t(P));  // :743
  g.partition_offset.assign(P + 1, 0);
  g.partition_offset[P] = g.V;
  for (int i = 0; i + 1 < P; ++i)  // :747-749, R9
    g.partition_offset[i + 1] = uint32_t(std::min<uint64_t>(uint64_t(g.page_offset[i + 1]) * kPageSize, g.V));
  g.local_partition_offset.assign(P, std::vector<uint32_t>(S + 1, 0));
  for (int p = 0; p < P; ++p) {  // :774-780, R9
    uint32_t pb = g.page_offset[p], pe = g.page_offset[p + 1];
    std::vector<uint32_t> lpo = optimal_offset(g.page_amount.data() + pb, pe - pb, uint32_t(S));
    auto& lo = g.local_partition_offset[p];
    lo[0] = g.partition_offset[p];
    lo[S] = g.partition_offset[p + 1];
    for (int i = 0; i + 1 < S; ++i)
      lo[i + 1] = uint32_t(std::min<uint64_t>(uint64_t(lpo[i + 1]) * kPageSize + g.partition_offset[p],
                                              g.partition_offset[p + 1]));
  }
  const int Q = P * S;
  g.csc_list.assign(Q, {});
  g.csr_list.assign(Q, {});
  g.csc_cindex.assign(Q, {});
  g.csr_index.assign(Q, {});
  g.csr_bitmap.assign(Q, {});
  build_direction<E>(g, locals, true);   // CSC (pull)
  build_direction<E>(g, locals, false);  // CSR (push)
  g.build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

// ---------------------------------------------------------------- engine --

template <class E>
const E* csc_edges(const Graph& g, int q) { return reinterpret_cast<const E*>(g.csc_list[q].data()); }
template <class E>
const E* csr_edges(const Graph& g, int q) { return reinterpret_cast<const E*>(g.csr_list[q].data()); }

inline bool bit(const std::vector<uint64_t>& bm, uint32_t v) { return (bm[v >> 6] >> (v & 63)) & 1; }
inline void set_bit(std::vector<uint64_t>& bm, uint32_t v) {
  __atomic_fetch_or(&bm[v >> 6], 1ull << (v & 63), __ATOMIC_RELAXED);
}

// write_min (core/atomic.hpp:35-43): CAS loop, true iff this call lowered *p.
inline bool write_min_u32(uint32_t* p, uint32_t v) {
  uint32_t cur = __atomic_load_n(p, __ATOMIC_RELAXED);
  while (v < cur)
    if (__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) return true;
  return false;
}
// write_min on a non-negative float: its bits order as the values do.
inline bool write_min_f32(float* p, float v) {
  uint32_t vb;
  std::memcpy(&vb, &v, 4);
  return write_min_u32(reinterpret_cast<uint32_t*>(p), vb);
}

// Set bits of a bitmap as an ascending vertex list.
std::vector<uint32_t> bitmap_list(const std::vector<uint64_t>& bm, uint32_t V) {
  std::vector<uint32_t> out;
  for (size_t w = 0; w < bm.size(); ++w)
    for (uint64_t b = bm[w]; b; b &= b - 1) {
      const uint32_t v = uint32_t(w * 64 + __builtin_ctzll(b));
      if (v < V) out.push_back(v);
    }
  return out;
}

// PageRank, toolkits/pagerank.cpp:26-103, always dense (use_iteration).
// Partials of partition i for a destination owned by partition m are
// combined in the receive order of core/graph.hpp:1756-1801:
// i = m-1, m-2, ..., m (mod P), sockets ascending.
int pagerank_impl(const Graph& g, int iterations, double* out, double* deltas) {
  const uint32_t V = g.V;
  const int P = g.P, S = g.S, T = g.threads;
  const double d = 0.85;
  std::vector<double> curr(V), next(V);
  for (uint32_t v = 0; v < V; ++v) {
    curr[v] = 1.0;
    if (g.out_deg[v] > 0) curr[v] /= g.out_deg[v];
  }
  double delta = double(V) / V;
  if (deltas) deltas[0] = delta;
  // partials[q] : one FP64 per compressed destination of sub-partition q
  std::vector<std::vector<double>> partial(P * S);
  for (int it = 0; it < iterations; ++it) {
    std::fill(next.begin(), next.end(), 0.0);
    for (int q = 0; q < P * S; ++q) {
      const auto& ci = g.csc_cindex[q];
      const Rec8* adj = csc_edges<Rec8>(g, q);
      const int64_t nd = int64_t(ci.size()) - 1;
      partial[q].assign(nd, 0.0);
      double* pq = partial[q].data();
#pragma omp parallel for num_threads(T) schedule(dynamic, 64)
      for (int64_t k = 0; k < nd; ++k) {  // dense_signal, pagerank.cpp:64-71
        double sum = 0;
        for (uint64_t e = ci[k].index; e < ci[k + 1].index; ++e) sum += curr[adj[e].src];
        pq[k] = sum;
      }
    }
    for (int m = 0; m < P; ++m) {
      for (int step = 0; step < P; ++step) {
        const int i = (m - step - 1 + P) % P;
        for (int s = 0; s < S; ++s) {
          const int q = i * S + s;
          const auto& ci = g.csc_cindex[q];
          const int64_t nd = int64_t(ci.size()) - 1;
          // only destinations owned by m: they form one contiguous run
          auto lo = std::lower_bound(ci.begin(), ci.begin() + nd, g.partition_offset[m],
                                     [](const CIdx& c, uint32_t v) { return c.vertex < v; }) - ci.begin();
          auto hi = std::lower_bound(ci.begin(), ci.begin() + nd, g.partition_offset[m + 1],
                                     [](const CIdx& c, uint32_t v) { return c.vertex < v; }) - ci.begin();
          const double* pq = partial[q].data();
#pragma omp parallel for num_threads(T)
          for (int64_t k = lo; k < hi; ++k) next[ci[k].vertex] += pq[k];  // dense_slot
        }
      }
    }
    // process_vertices (pagerank.cpp:79-98); sequential reduction order
    delta = 0;
    const bool last = (it == iterations - 1);
    for (uint32_t v = 0; v < V; ++v) {
      double y = (1 - d) + d * next[v];
      if (last) {
        next[v] = y;
      } else if (g.out_deg[v] > 0) {
        next[v] = y / g.out_deg[v];
        delta += std::fabs(next[v] - curr[v]) * g.out_deg[v];
      } else {
        next[v] = y;
        delta += std::fabs(next[v] - curr[v]);
      }
    }
    delta = last ? 0.0 : delta / V;
    if (deltas) deltas[it + 1] = delta;
    std::swap(curr, next);
  }
  std::memcpy(out, curr.data(), sizeof(double) * V);
  return 0;
}

// Sparse (push) step over all partitions for vertex u: calls f(dst, edge) for
// every out-edge, visiting (p, s) in order (core/graph.hpp:1580-1655).
template <class E, class F>
void for_out_edges(const Graph& g, uint32_t u, F&& f) {
  for (int q = 0; q < g.P * g.S; ++q) {
    if (!bit(g.csr_bitmap[q], u)) continue;
    const E* adj = csr_edges<E>(g, q);
    for (uint64_t e = g.csr_index[q][u]; e < g.csr_index[q][size_t(u) + 1]; ++e) f(adj[e]);
  }
}

// BFS, toolkits/bfs.cpp:22-112 with R12 (min-id parent).  policy 0 =
// reference (always push, core/graph.hpp:1497), 1 = direction-optimising
// (push iff the frontier's out-degree sum < E/20).  Level-synchronous and
// OpenMP-parallel like the reference: the push's cas on parent (bfs.cpp:54)
// becomes write_min among the vertices not yet visited, so parents and
// depths do not depend on the policy, the thread count or the schedule.
int bfs_impl(const Graph& g, uint32_t root, int policy, uint32_t* parent, uint32_t* depth,
             uint32_t* levels_out) {
  const uint32_t V = g.V;
  const int T = g.threads;
  if (root >= V) return -2;
  std::vector<uint64_t> visited((V >> 6) + 1, 0), active_in((V >> 6) + 1, 0), active_out((V >> 6) + 1, 0);
#pragma omp parallel for num_threads(T)
  for (int64_t v = 0; v < int64_t(V); ++v) {
    parent[v] = V;
    depth[v] = UINT32_MAX;
  }
  parent[root] = root;
  depth[root] = 0;
  set_bit(visited, root);
  set_bit(active_in, root);
  uint64_t active_vertices = 1;
  uint32_t level = 0;
  while (active_vertices > 0) {
    std::fill(active_out.begin(), active_out.end(), 0);
    const std::vector<uint32_t> front = bitmap_list(active_in, V);
    bool push = true;
    if (policy == 1) {
      uint64_t active_edges = 0;
      for (uint32_t v : front) active_edges += g.out_deg[v];
      push = active_edges < g.E / 20;
    }
    if (push) {
      const int64_t nf = int64_t(front.size());
#pragma omp parallel for num_threads(T) schedule(dynamic, 64)
      for (int64_t k = 0; k < nf; ++k) {
        const uint32_t u = front[k];
        for_out_edges<Rec8>(g, u, [&](const Rec8& e) {
          const uint32_t dst = e.dst;
          if (!bit(visited, dst)) {
            write_min_u32(&parent[dst], u);
            set_bit(active_out, dst);
          }
        });
      }
    } else {
      for (int q = 0; q < g.P * g.S; ++q) {
        const auto& ci = g.csc_cindex[q];
        const Rec8* adj = csc_edges<Rec8>(g, q);
        const int64_t nd = int64_t(ci.size()) - 1;
#pragma omp parallel for num_threads(T) schedule(dynamic, 256)
        for (int64_t k = 0; k < nd; ++k) {
          uint32_t dst = ci[k].vertex;
          if (bit(visited, dst)) continue;
          uint32_t best = V;
          for (uint64_t e = ci[k].index; e < ci[k + 1].index; ++e)
            if (bit(active_in, adj[e].src)) best = std::min(best, adj[e].src);
          if (best < V) {
            // several sub-partitions may propose: keep the minimum
            write_min_u32(&parent[dst], best);
            set_bit(active_out, dst);
          }
        }
      }
    }
    ++level;
    const std::vector<uint32_t> next = bitmap_list(active_out, V);
    for (uint32_t v : next) {
      set_bit(visited, v);
      depth[v] = level;
    }
    active_vertices = next.size();
    std::swap(active_in, active_out);
  }
  if (levels_out) *levels_out = level;
  return 0;
}

// CC, toolkits/cc.cpp:22-112 on the symmetrised graph; mode switch of
// core/graph.hpp:1484-1492 with R4.  Both modes run on OpenMP threads with
// the reference's write_min (cc.cpp:54, :75); labels converge to the
// smallest id of each component whatever the interleaving.
int cc_impl(const Graph& g, uint32_t* label, uint32_t* iters_out) {
  const uint32_t V = g.V;
  const int T = g.threads;
  std::vector<uint64_t> active_in((V >> 6) + 1, 0), active_out((V >> 6) + 1, 0);
#pragma omp parallel for num_threads(T)
  for (int64_t v = 0; v < int64_t(V); ++v) label[v] = uint32_t(v);
  for (uint32_t v = 0; v < V; ++v) set_bit(active_in, v);
  uint64_t active_vertices = V;
  uint32_t it = 0;
  while (active_vertices > 0) {
    std::fill(active_out.begin(), active_out.end(), 0);
    const std::vector<uint32_t> front = bitmap_list(active_in, V);
    uint64_t active_edges = 0;
    for (uint32_t v : front) active_edges += g.out_deg[v];
    const bool sparse = active_edges < g.E / 20;
    if (sparse) {
      const int64_t nf = int64_t(front.size());
#pragma omp parallel for num_threads(T) schedule(dynamic, 64)
      for (int64_t k = 0; k < nf; ++k) {
        const uint32_t u = front[k];
        const uint32_t msg = __atomic_load_n(&label[u], __ATOMIC_RELAXED);
        for_out_edges<Rec8>(g, u, [&](const Rec8& e) {
          if (msg < __atomic_load_n(&label[e.dst], __ATOMIC_RELAXED)) {
            write_min_u32(&label[e.dst], msg);
            set_bit(active_out, e.dst);
          }
        });
      }
    } else {
      for (int q = 0; q < g.P * g.S; ++q) {
        const auto& ci = g.csc_cindex[q];
        const Rec8* adj = csc_edges<Rec8>(g, q);
        const int64_t nd = int64_t(ci.size()) - 1;
#pragma omp parallel for num_threads(T) schedule(dynamic, 256)
        for (int64_t k = 0; k < nd; ++k) {
          const uint32_t dst = ci[k].vertex;
          uint32_t msg = dst;
          for (uint64_t e = ci[k].index; e < ci[k + 1].index; ++e)
            msg = std::min(msg, __atomic_load_n(&label[adj[e].src], __ATOMIC_RELAXED));
          if (msg < dst && msg < __atomic_load_n(&label[dst], __ATOMIC_RELAXED)) {
            write_min_u32(&label[dst], msg);
            set_bit(active_out, dst);
          }
        }
      }
    }
    active_vertices = 0;
    for (uint64_t w : active_out) active_vertices += __builtin_popcountll(w);
    std::swap(active_in, active_out);
    ++it;
  }
  if (iters_out) *iters_out = it;
  return 0;
}

// SSSP, toolkits/sssp.cpp:24-108: float distances, INF = 1e9, push relax
// with write_min (sssp.cpp:53) on OpenMP threads.  One fp32 add per
// relaxation, rounded to nearest: the add is monotone, so the fixpoint is the
// minimum over paths of the left-fold sums whatever the order of updates.
int sssp_impl(const Graph& g, uint32_t root, float* dist, uint32_t* iters_out) {
  const uint32_t V = g.V;
  const int T = g.threads;
  if (root >= V) return -2;
  std::vector<uint64_t> active_in((V >> 6) + 1, 0), active_out((V >> 6) + 1, 0);
#pragma omp parallel for num_threads(T)
  for (int64_t v = 0; v < int64_t(V); ++v) dist[v] = 1e9f;
  dist[root] = 0.0f;
  set_bit(active_in, root);
  uint64_t active_vertices = 1;
  uint32_t it = 0;
  while (active_vertices > 0) {
    std::fill(active_out.begin(), active_out.end(), 0);
    const std::vector<uint32_t> front = bitmap_list(active_in, V);
    const int64_t nf = int64_t(front.size());
#pragma omp parallel for num_threads(T) schedule(dynamic, 64)
    for (int64_t k = 0; k < nf; ++k) {
      const uint32_t u = front[k];
      uint32_t mb = __atomic_load_n(reinterpret_cast<const uint32_t*>(&dist[u]), __ATOMIC_RELAXED);
      float msg;
      std::memcpy(&msg, &mb, 4);
      for_out_edges<Rec12>(g, u, [&](const Rec12& e) {
        volatile float relax = msg + e.w;  // one fp32 add, round to nearest
        if (write_min_f32(&dist[e.dst], relax)) set_bit(active_out, e.dst);
      });
    }
    active_vertices = 0;
    for (uint64_t w : active_out) active_vertices += __builtin_popcountll(w);
    std::swap(active_in, active_out);
    ++it;
  }
  if (iters_out) *iters_out = it;
  return 0;
}

// Batched CSR at P = S = 1: prepare_push_mode (core/graph.hpp:1248-1355) of
// every graph on its own, i.e. a stable counting sort of its records by src
// (:1262) and the row pointer of its dense index (:1312-1353, every entry
// defined).  Graphs run in parallel, one thread each.
template <class E>
void csr_batch_impl(const E* edges, int n, const uint32_t* nv, const uint64_t* ne, E* out, uint64_t* rowptr,
                    int threads) {
  std::vector<uint64_t> eoff(size_t(n) + 1, 0), voff(size_t(n) + 1, 0);
  for (int i = 0; i < n; ++i) {
    eoff[i + 1] = eoff[i] + ne[i];
    voff[i + 1] = voff[i] + uint64_t(nv[i]) + 1;
  }
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
  for (int i = 0; i < n; ++i) {
    const E* in = edges + eoff[i];
    E* o = out + eoff[i];
    uint64_t* rp = rowptr + voff[i];
    const uint32_t V = nv[i];
    std::fill(rp, rp + V + 1, 0);
    for (uint64_t j = 0; j < ne[i]; ++j) rp[in[j].src + 1]++;
    for (uint32_t v = 0; v < V; ++v) rp[v + 1] += rp[v];
    std::vector<uint64_t> at(rp, rp + V);
    for (uint64_t j = 0; j < ne[i]; ++j) o[at[in[j].src]++] = in[j];
  }
}

}  // namespace

extern "C" {

int orc_optimal_offset(const uint64_t* a, uint32_t n, uint32_t k, uint32_t* out) {
  if (k == 0) return -1;
  std::vector<uint32_t> o = optimal_offset(a, n, k);
  std::memcpy(out, o.data(), sizeof(uint32_t) * (k + 1));
  return 0;
}

int orc_build(const void* edges, uint64_t E0, int rec_bytes, uint32_t V, int P, int S, uint64_t eta,
              uint32_t m, int symmetrize, int threads, void** out) {
  if (!out || V == 0 || P < 1 || S < 1 || (rec_bytes != 8 && rec_bytes != 12)) return -1;
  auto g = std::make_unique<Graph>();
  g->V = V;
  g->P = P;
  g->S = S;
  g->eta = eta;
  g->m = m;
  g->rec = rec_bytes;
  g->symmetrize = symmetrize;
  g->threads = threads > 0 ? threads : omp_get_max_threads();
  int rc = rec_bytes == 8 ? build_impl<Rec8>(*g, static_cast<const Rec8*>(edges), E0)
                          : build_impl<Rec12>(*g, static_cast<const Rec12*>(edges), E0);
  if (rc != 0) return rc;
  *out = g.release();
  return 0;
}

void orc_free(void* h) { delete static_cast<Graph*>(h); }
uint64_t orc_num_edges(void* h) { return static_cast<Graph*>(h)->E; }
double orc_build_seconds(void* h) { return static_cast<Graph*>(h)->build_seconds; }
const uint32_t* orc_out_degree(void* h) { return static_cast<Graph*>(h)->out_deg.data(); }
const uint32_t* orc_in_degree(void* h) { return static_cast<Graph*>(h)->in_deg.data(); }
const uint32_t* orc_partition_offset(void* h) { return static_cast<Graph*>(h)->partition_offset.data(); }
const uint32_t* orc_local_partition_offset(void* h, int p) {
  return static_cast<Graph*>(h)->local_partition_offset[p].data();
}
const uint64_t* orc_page_amount(void* h, uint32_t* pages) {
  auto* g = static_cast<Graph*>(h);
  *pages = uint32_t(g->page_amount.size());
  return g->page_amount.data();
}
uint64_t orc_csc(void* h, int p, int s, const void** list, const void** cindex, uint64_t* n_cindex) {
  auto* g = static_cast<Graph*>(h);
  int q = p * g->S + s;
  *list = g->csc_list[q].data();
  *cindex = g->csc_cindex[q].data();
  *n_cindex = g->csc_cindex[q].size();
  return g->csc_list[q].size() / g->rec;
}
uint64_t orc_csr(void* h, int p, int s, const void** list, const uint64_t** index, const uint64_t** bitmap) {
  auto* g = static_cast<Graph*>(h);
  int q = p * g->S + s;
  *list = g->csr_list[q].data();
  *index = g->csr_index[q].data();
  *bitmap = g->csr_bitmap[q].data();
  return g->csr_list[q].size() / g->rec;
}
int orc_pagerank(void* h, int iterations, double* out, double* deltas) {
  auto* g = static_cast<Graph*>(h);
  if (g->rec != 8 || iterations < 1) return -1;
  return pagerank_impl(*g, iterations, out, deltas);
}
int orc_bfs(void* h, uint32_t root, int policy, uint32_t* parent, uint32_t* depth, uint32_t* levels) {
  auto* g = static_cast<Graph*>(h);
  if (g->rec != 8) return -1;
  return bfs_impl(*g, root, policy, parent, depth, levels);
}
int orc_cc(void* h, uint32_t* label, uint32_t* iters) {
  auto* g = static_cast<Graph*>(h);
  if (g->rec != 8) return -1;
  return cc_impl(*g, label, iters);
}
int orc_sssp(void* h, uint32_t root, float* dist, uint32_t* iters) {
  auto* g = static_cast<Graph*>(h);
  if (g->rec != 12) return -1;
  return sssp_impl(*g, root, dist, iters);
}
// n graphs back to back in edges (graph i: local ids < nv[i], ne[i]
// records); out gets every graph's CSR list at its input offset, rowptr the
// row pointers (nv[i] + 1 entries each) back to back.  Ids are not checked
// beyond the caller's guarantee.
int orc_csr_batch(const void* edges, int rec_bytes, int n, const uint32_t* nv, const uint64_t* ne, void* out,
                  uint64_t* rowptr, int threads) {
  if (n < 0 || (rec_bytes != 8 && rec_bytes != 12)) return -1;
  const int T = threads > 0 ? threads : omp_get_max_threads();
  if (rec_bytes == 8)
    csr_batch_impl(static_cast<const Rec8*>(edges), n, nv, ne, static_cast<Rec8*>(out), rowptr, T);
  else
    csr_batch_impl(static_cast<const Rec12*>(edges), n, nv, ne, static_cast<Rec12*>(out), rowptr, T);
  return 0;
}

}  // extern "C"
