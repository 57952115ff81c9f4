// Widest (maximum-bottleneck) path from a root: width[v] = the largest, over
// the paths root -> v, of the smallest edge weight on the path (width[root]
// = 1e9, unreached 0; weights must be positive).  Not one of the reference's
// toolkits: a new program written only against the Graph<EdgeData> facade
// and the generic device engine (device_engine.hpp), with the reference's
// four signal/slot functors and no kernel of its own in the library.  Its
// updates are min and max of the weights only, so the fixpoint is exact and
// independent of the relaxation order.  With EG_DUMP=<path> the first run
// writes the widths there with dump_vertex_array.
#include <cstdlib>
#include <string>

#include "graph.hpp"
#include "device_engine.hpp"

using namespace endgraph;

typedef float Weight;

// write_max on a non-negative float (its bits order as the values do);
// true iff this call raised *p
__device__ __forceinline__ bool write_max(float* p, float v) {
  return __uint_as_float(atomicMax(reinterpret_cast<unsigned int*>(p), __float_as_uint(v))) < v;
}
__device__ __forceinline__ void set_bit(uint64_t* bm, UINT v) {
  atomicOr(reinterpret_cast<unsigned long long*>(&bm[v >> 6]), 1ull << (v & 63));
}

struct SparseSignal {  // emit(src, width[src])
  const float* width;
  __device__ bool operator()(UINT v, float& msg) const {
    msg = width[v];
    return true;
  }
};
struct SparseSlot {  // relax every out-edge with write_max
  float* width;
  uint64_t* active_out;
  __device__ UINT operator()(UINT src, float msg, AdjacentListType<Weight> out_edges) const {
    UINT activated = 0;
    for (const EdgeType<Weight>* e = out_edges.begin; e != out_edges.end; ++e) {
      const float w = fminf(msg, e->edge_data);
      if (w > width[e->dst] && write_max(&width[e->dst], w)) {
        set_bit(active_out, e->dst);
        activated += 1;
      }
    }
    return activated;
  }
};
struct DenseSignal {  // the widest in-edge of dst
  const float* width;
  __device__ bool operator()(UINT dst, AdjacentListType<Weight> in_edges, float& msg) const {
    float best = 0.0f;
    for (const EdgeType<Weight>* e = in_edges.begin; e != in_edges.end; ++e)
      best = fmaxf(best, fminf(width[e->src], e->edge_data));
    msg = best;
    return best > width[dst];
  }
};
struct DenseSlot {
  float* width;
  uint64_t* active_out;
  __device__ UINT operator()(UINT dst, float msg) const {
    if (msg > width[dst] && write_max(&width[dst], msg)) {
      set_bit(active_out, dst);
      return 1;
    }
    return 0;
  }
};
struct Reached {  // process_vertices: 1 per vertex with a path from the root
  const float* width;
  __device__ UINT operator()(UINT v) const { return width[v] > 0.0f ? 1u : 0u; }
};

void compute(Graph<Weight>* graph, DeviceEngine<Weight>& engine, UINT root) {
  double exec_time = -get_time();
  float* width = graph->alloc_vertex_array<float>();
  VertexSubset* active_in = graph->alloc_vertex_subset();
  VertexSubset* active_out = graph->alloc_vertex_subset();
  VertexSubset* all = graph->alloc_vertex_subset();
  all->fill();
  graph->fill_vertex_array(width, 0.0f);
  const float inf = 1e9f;
  check(eg_memcpy_h2d(width + root, &inf, sizeof(float), nullptr));
  active_in->clear();
  active_in->set_bit(root);
  UINT active_vertices = 1;
  for (int i_i = 0; active_vertices > 0; i_i++) {
    printf("active(%d)>=%u\n", i_i, active_vertices);
    active_out->clear();
    active_vertices = engine.process_edges<UINT, float>(
        SparseSignal{width}, SparseSlot{width, active_out->data}, DenseSignal{width}, DenseSlot{width, active_out->data},
        active_in);
    std::swap(active_in, active_out);
  }
  exec_time += get_time();
  const UINT reached = engine.process_vertices<UINT>(Reached{width}, all);
  dump_result(graph, width);
  printf("reached = %u\n", reached);
  printf("pre time: %lf s, exec time: %lf s\n", graph->prep_time, exec_time);
  graph->dealloc_vertex_array(width);
  delete active_in;
  delete active_out;
  delete all;
}

int main(int argc, char** argv) {
  if (argc < 5) {
    printf("widest [file] [vertices_num] [root] [eta] [partitions=1] [sockets=1] [reference|auto|push|pull]\n");
    return 1;
  }
  const int P = argc > 5 ? std::atoi(argv[5]) : 1;
  const int S = argc > 6 ? std::atoi(argv[6]) : 1;
  try {
    Graph<Weight>* graph = new Graph<Weight>(std::atoi(argv[4]), P, S);
    graph->use_undirected = false;
    graph->use_root = true;
    if (argc > 7) {
      const std::string pol = argv[7];
      graph->policy = pol == "auto" ? EG_POLICY_AUTO : pol == "push" ? EG_POLICY_PUSH
                      : pol == "pull" ? EG_POLICY_PULL : EG_POLICY_REFERENCE;
    }
    graph->load_directed(argv[1], std::atoi(argv[2]));
    DeviceEngine<Weight> engine(*graph);
    const UINT root = UINT(std::atoi(argv[3]));
    compute(graph, engine, root);
    compute(graph, engine, root);
    delete graph;
  } catch (const Error& e) {
    fprintf(stderr, "widest: %s (status %d)\n", e.what(), e.status);
    return 2;
  }
  return 0;
}
