// Graph<EdgeData>: the reference's C++ surface (core/graph.hpp:100-1819) over
// the C ABI of libendgraph.so.  Toolkits written against the reference keep
// their structure: construct the graph, load_directed /
// load_undirected_from_directed, allocate vertex arrays and subsets, and call
// process_vertices / process_edges.  Differences, all forced by the device:
//   - vertex arrays and VertexSubsets live in HBM; gather_vertex_array copies
//     to a host vector instead of into the same pointer;
//   - process_vertices / process_edges take a *program* object naming one of
//     the signal/slot pairs the library implements (see namespace programs)
//     instead of std::function lambdas, which cannot run on the GPU;
//   - errors throw endgraph::Error carrying the eg_status and message
//     (the reference asserts or exits).
// The partition and socket counts that the reference reads from MPI and
// libnuma are constructor arguments (all partitions share one GPU).
#pragma once

#include <sys/stat.h>
#include <sys/time.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "endgraph.h"

namespace endgraph {

typedef uint32_t UINT;     // core/type.hpp:23
typedef uint64_t ULLONG;   // core/type.hpp:24
struct Empty {};           // core/type.hpp:20

template <typename EdgeData>
struct EdgeType {          // core/type.hpp:26-31
  UINT src;
  UINT dst;
  EdgeData edge_data;
} __attribute__((packed));

template <>
struct EdgeType<Empty> {   // core/type.hpp:33-40
  UINT src;
  UINT dst;
} __attribute__((packed));

class Error : public std::runtime_error {
 public:
  int status;
  Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

inline void check(int status) {
  if (status != EG_OK) throw Error(status, eg_last_error());
}

inline double get_time() {  // core/time.hpp:22-26
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec + tv.tv_usec / 1e6;
}

// BitmapType / VertexSubset (core/bitmap.hpp:23-68) in HBM.
class BitmapType {
 public:
  size_t size;
  uint64_t* data;  // device words, (size >> 6) + 1 of them
  explicit BitmapType(size_t n) : size(n), data(nullptr) {
    void* p = nullptr;
    check(eg_device_alloc(&p, ((n >> 6) + 1) * 8));
    data = static_cast<uint64_t*>(p);
    clear();
  }
  ~BitmapType() { eg_device_free(data); }
  BitmapType(const BitmapType&) = delete;
  BitmapType& operator=(const BitmapType&) = delete;
  void clear() { check(eg_bitmap_clear(data, UINT(size), nullptr)); }
  void fill() { check(eg_bitmap_fill(data, UINT(size), nullptr)); }
  void set_bit(size_t i) { check(eg_bitmap_set_bit(data, UINT(i), nullptr)); }
  uint64_t count() const {
    uint64_t c = 0;
    check(eg_bitmap_count(data, UINT(size), &c, nullptr));
    return c;
  }
};
typedef BitmapType VertexSubset;

// The signal/slot pairs of toolkits/*.cpp that the library runs on the GPU.
namespace programs {
// pagerank.cpp:36-47: curr = 1/deg; returns 1 per vertex
struct PageRankInit {
  double* curr;
  float* curr32 = nullptr;
};
// pagerank.cpp:53-76: next[dst] += sum of curr[src]; fp32c reads curr32
struct PageRankPull {
  const double* curr;
  double* next;
  const float* curr32 = nullptr;
};
// pagerank.cpp:79-98: next = 0.15 + 0.85 next (/deg unless last); returns |next-curr|*deg
struct PageRankApply {
  const double* curr;
  double* next;
  bool last;
  float* next32 = nullptr;
};
// bfs.cpp:41-78: one level; parents by the minimum frontier id
struct BfsLevel {
  UINT* parent;
  VertexSubset* active_out;
  VertexSubset* visited;
  UINT* depth = nullptr;
  UINT level = 0;
};
// bfs.cpp:79-86: visited |= active_out; returns 1 per vertex
struct BfsVisit {
  VertexSubset* visited;
};
// cc.cpp:35-41: label = vertex id; returns 1 per vertex
struct CcInit {
  UINT* label;
};
// cc.cpp:45-80: min-label propagation
struct CcStep {
  UINT* label;
  VertexSubset* active_out;
};
// sssp.cpp:38-83: Bellman-Ford relax with write_min
struct SsspStep {
  float* distance;
  VertexSubset* active_out;
};
}  // namespace programs

template <typename EdgeData = Empty>
class Graph {
 public:
  // mode flags of the toolkits (core/graph.hpp:104-106 declares has_*; the
  // engine reads use_*: repair R3 keeps the use_* names)
  bool use_undirected = false;
  bool use_iteration = false;
  bool use_root = false;
  int max_iterations = 1;
  size_t eta;

  int partitions_num;
  int partition_idx = 0;  // one process drives every partition
  int sockets_num;
  UINT vertices_num = 0;
  ULLONG edges_num = 0;
  double prep_time = 0;
  ULLONG prep_data = 0;
  ULLONG exe_data = 0;
  std::vector<UINT> degree;  // out-degree, host copy (core/graph.hpp:144)
  const UINT* degree_dev = nullptr;

  // policy for process_edges' sparse/dense switch (EG_POLICY_*)
  int policy = EG_POLICY_REFERENCE;

  explicit Graph(int eta_, int partitions = 1, int sockets = 1)
      : eta(size_t(eta_)), partitions_num(partitions), sockets_num(sockets) {}
  ~Graph() {
    if (g_) eg_graph_destroy(g_);
    if (delta_) eg_device_free(delta_);
  }
  Graph(const Graph&) = delete;
  Graph& operator=(const Graph&) = delete;

  eg_graph* handle() const { return g_; }

  // load_directed (core/graph.hpp:1358-1373): the file read and the upload are
  // outside prep_time, as the reference's preload_graph_file is.
  void load_directed(const std::string& path, UINT vertices) { load(path, vertices, false); }
  // load_undirected_from_directed (core/graph.hpp:629-646)
  void load_undirected_from_directed(const std::string& path, UINT vertices) { load(path, vertices, true); }

  // Build from a device edge list already in HBM.
  void build(const void* d_edges, ULLONG n, UINT vertices, bool symmetric) {
    vertices_num = vertices;
    if (g_) eg_graph_destroy(g_);
    g_ = nullptr;
    check(eg_graph_create(vertices, partitions_num, sockets_num, eta, UINT(max_iterations), int(sizeof(EdgeType<EdgeData>)),
                          &g_));
    double t = -get_time();
    check(eg_graph_build(g_, d_edges, n, symmetric ? 1 : 0, nullptr));
    t += get_time();
    prep_time = t;
    after_build(vertices);
  }

  template <typename T>
  T* alloc_vertex_array() {  // core/graph.hpp:298-305
    void* p = nullptr;
    check(eg_device_alloc(&p, uint64_t(vertices_num) * sizeof(T)));
    return static_cast<T*>(p);
  }
  template <typename T>
  void dealloc_vertex_array(T* array) {  // core/graph.hpp:308-311
    check(eg_device_free(array));
  }
  template <typename T>
  void fill_vertex_array(T* array, T value) {  // core/graph.hpp:289-294
    check(eg_device_fill(array, vertices_num, &value, int(sizeof(T)), nullptr));
  }
  VertexSubset* alloc_vertex_subset() { return new VertexSubset(vertices_num); }  // core/graph.hpp:400-402
  // gather_vertex_array (core/graph.hpp:380-397): the whole array to the host
  template <typename T>
  std::vector<T> gather_vertex_array(const T* array, int /*root*/) {
    std::vector<T> h(vertices_num);
    check(eg_memcpy_d2h(h.data(), array, uint64_t(vertices_num) * sizeof(T), nullptr));
    return h;
  }
  // dump_vertex_array / restore_vertex_array (core/graph.hpp:324-375): the
  // reference's raw V * sizeof(T) file, all partitions at once
  template <typename T>
  void dump_vertex_array(const T* array, const std::string& path) {
    check(eg_dump_vertex_array(array, sizeof(T), vertices_num, path.c_str(), nullptr));
  }
  template <typename T>
  void restore_vertex_array(T* array, const std::string& path) {
    check(eg_restore_vertex_array(array, sizeof(T), vertices_num, path.c_str(), nullptr));
  }

  // ---- process_vertices (core/graph.hpp:1377-1441) ----
  template <typename R>
  R process_vertices(const programs::PageRankInit& p, VertexSubset* active) {
    ensure_engine();
    check(eg_pr_init(g_, p.curr, p.curr32, nullptr));
    return R(active->count());
  }
  template <typename R>
  R process_vertices(const programs::PageRankApply& p, VertexSubset* /*active: all*/) {
    ensure_engine();
    double* d_delta = scratch_double();
    double zero = 0;
    check(eg_memcpy_h2d(d_delta, &zero, 8, nullptr));
    check(eg_pr_apply(g_, p.curr, p.next, p.next32, p.last ? 1 : 0, d_delta, nullptr));
    double delta = 0;
    check(eg_memcpy_d2h(&delta, d_delta, 8, nullptr));
    return R(delta);
  }
  template <typename R>
  R process_vertices(const programs::BfsVisit& p, VertexSubset* active) {
    check(eg_bitmap_or(p.visited->data, active->data, vertices_num, nullptr));
    return R(active->count());
  }
  template <typename R>
  R process_vertices(const programs::CcInit& p, VertexSubset* active) {
    check(eg_device_iota(p.label, vertices_num, nullptr));
    return R(active->count());
  }

  // ---- process_edges (core/graph.hpp:1473-1817) ----
  template <typename R, typename M>
  R process_edges(const programs::PageRankPull& p, VertexSubset* /*active: all, dense*/,
                  VertexSubset* /*dense_selective*/ = nullptr) {
    ensure_engine();
    if (partitions_num > 1) exe_data += 12 * dense_messages(EG_PROGRAM_PAGERANK, nullptr, nullptr, nullptr);
    if (p.curr32)
      check(eg_pr_pull(g_, EG_PR_FP32C, p.curr32, p.next, nullptr));
    else
      check(eg_pr_pull(g_, EG_PR_FP64, p.curr, p.next, nullptr));
    return R(0);
  }
  template <typename R, typename M>
  R process_edges(const programs::BfsLevel& p, VertexSubset* active, VertexSubset* /*dense_selective*/ = nullptr) {
    ensure_engine();
    uint64_t nv = 0, ne = 0;
    const uint64_t ae = active_edges(active);
    // bfs.cpp passes visited as dense_selective
    count_step(EG_PROGRAM_BFS, ae, active, p.visited, nullptr, 8, true);
    check(eg_bfs_level(g_, policy, p.level, active->data, p.active_out->data, p.visited->data, p.parent, p.depth, ae,
                       &nv, &ne, nullptr, nullptr));
    return R(nv);
  }
  template <typename R, typename M>
  R process_edges(const programs::CcStep& p, VertexSubset* active, VertexSubset* /*dense_selective*/ = nullptr) {
    ensure_engine();
    uint64_t nv = 0, ne = 0;
    const uint64_t ae = active_edges(active);
    count_step(EG_PROGRAM_CC, ae, active, nullptr, p.label, 8, false);
    check(eg_cc_iteration(g_, policy, active->data, p.active_out->data, p.label, ae, &nv, &ne, nullptr, nullptr));
    return R(nv);
  }
  template <typename R, typename M>
  R process_edges(const programs::SsspStep& p, VertexSubset* active, VertexSubset* /*dense_selective*/ = nullptr) {
    ensure_engine();
    uint64_t nv = 0, ne = 0;
    const uint64_t ae = active_edges(active);
    count_step(EG_PROGRAM_SSSP, ae, active, nullptr, p.distance, 8, false);
    check(eg_sssp_iteration(g_, policy, active->data, p.active_out->data, p.distance, ae, &nv, &ne, nullptr,
                            nullptr));
    return R(nv);
  }

 private:
  eg_graph* g_ = nullptr;
  double* delta_ = nullptr;

  // The file is read rank slice by rank slice and each slice's first sorts
  // run while the next is read (eg_graph_build_file); prep_time covers the
  // whole call, the read included, since the two overlap.
  void load(const std::string& path, UINT vertices, bool symmetric) {
    const int rec = int(sizeof(EdgeType<EdgeData>));
    uint64_t n = 0;
    check(eg_edge_file_records(path.c_str(), rec, &n));
    void* d = nullptr;
    check(eg_device_alloc(&d, n ? n * uint64_t(rec) : 1));
    try {
      vertices_num = vertices;
      if (g_) eg_graph_destroy(g_);
      g_ = nullptr;
      check(eg_graph_create(vertices, partitions_num, sockets_num, eta, UINT(max_iterations), rec, &g_));
      double t = -get_time();
      check(eg_graph_build_file(g_, path.c_str(), d, n, symmetric ? 1 : 0, &n, nullptr));
      t += get_time();
      prep_time = t;
      after_build(vertices);
    } catch (...) {
      eg_device_free(d);
      throw;
    }
    eg_device_free(d);
  }

  void after_build(UINT vertices) {
    eg_graph_info info;
    check(eg_graph_get_info(g_, &info));
    edges_num = info.edges_num;
    const UINT* in_deg = nullptr;
    check(eg_graph_degree(g_, &degree_dev, &in_deg));
    degree.resize(vertices);
    check(eg_memcpy_d2h(degree.data(), degree_dev, uint64_t(vertices) * 4, nullptr));
    uint64_t bytes = 0;
    check(eg_prep_data(g_, &bytes));
    prep_data = bytes;
  }

  uint64_t dense_messages(int program, VertexSubset* active, VertexSubset* visited, const void* values) {
    uint64_t n = 0;
    check(eg_dense_messages(g_, program, active ? active->data : nullptr, visited ? visited->data : nullptr, values,
                            &n, nullptr));
    return n;
  }

  // exe_data of one step before it runs (core/graph.hpp:1560, :1671, :1736):
  // sparse, each partition's messages (one per active vertex: every toolkit's
  // sparse signal emits) counted once per other partition, as a count, as
  // the reference adds it; dense, the dense_selective bitmap words sent to
  // the other partitions, then the messages' bytes.  One partition sends
  // nothing.
  void count_step(int program, uint64_t ae, VertexSubset* active, VertexSubset* dense_selective, const void* values,
                  uint64_t msg_bytes, bool bitmap) {
    if (partitions_num <= 1) return;
    int sparse = 0;
    check(eg_mode_sparse(g_, policy, program, ae, &sparse));
    if (sparse) {
      exe_data += uint64_t(partitions_num - 1) * active->count();
      return;
    }
    if (bitmap) {
      std::vector<uint32_t> off(partitions_num + 1);
      check(eg_graph_partition_offset(g_, off.data()));
      for (int p = 0; p < partitions_num; ++p)
        exe_data += uint64_t(partitions_num - 1) * ((off[p + 1] - off[p]) / 64) * sizeof(unsigned long);
    }
    exe_data += msg_bytes * dense_messages(program, active, dense_selective, values);
  }

  void ensure_engine() { check(eg_engine_prepare(g_, nullptr, nullptr)); }

  double* scratch_double() {
    if (!delta_) {
      void* p = nullptr;
      check(eg_device_alloc(&p, 8));
      delta_ = static_cast<double*>(p);
    }
    return delta_;
  }

  // sum of out-degrees over the active set (the mode switch's input)
  uint64_t active_edges(VertexSubset* active) {
    uint64_t e = 0;
    check(eg_bitmap_degree_sum(g_, active->data, &e, nullptr));
    return e;
  }
};

// The toolkits' result export: when EG_DUMP names a file, the first call
// writes the array there (dump_vertex_array, core/graph.hpp:324-352) and
// clears the variable, so the repeated timing runs leave it alone.
template <typename EdgeData, typename T>
void dump_result(Graph<EdgeData>* graph, const T* array) {
  const char* path = std::getenv("EG_DUMP");
  if (!path || !*path) return;
  graph->dump_vertex_array(array, std::string(path));
  unsetenv("EG_DUMP");
}

}  // namespace endgraph
