// SSSP drop-in: toolkits/sssp.cpp of the reference on the MI355X library
// (float weights, INF = 1e9, write_min relaxation).  With EG_DUMP=<path> in
// the environment the first run also writes its distances there with
// dump_vertex_array (the reference's raw V * 4-byte format).
#include <cstdlib>

#include "graph.hpp"

using namespace endgraph;

typedef float Weight;

void compute(Graph<Weight>* graph, UINT root) {
  double exec_time = 0;
  exec_time -= get_time();
  graph->exe_data = 0;

  Weight* distance = graph->alloc_vertex_array<Weight>();
  VertexSubset* active_in = graph->alloc_vertex_subset();
  VertexSubset* active_out = graph->alloc_vertex_subset();
  active_in->clear();
  active_in->set_bit(root);
  graph->fill_vertex_array(distance, (Weight)1e9);
  const Weight zero = 0;
  check(eg_memcpy_h2d(distance + root, &zero, sizeof(Weight), nullptr));
  UINT active_vertices = 1;

  for (int i_i = 0; active_vertices > 0; i_i++) {
    printf("active(%d)>=%u\n", i_i, active_vertices);
    active_out->clear();
    active_vertices = graph->process_edges<UINT, Weight>(programs::SsspStep{distance, active_out}, active_in);
    std::swap(active_in, active_out);
  }

  exec_time += get_time();

  dump_result(graph, distance);
  std::vector<Weight> dist = graph->gather_vertex_array(distance, 0);
  UINT max_v_i = root;
  for (UINT v_i = 0; v_i < graph->vertices_num; v_i++)
    if (dist[v_i] < 1e9 && dist[v_i] > dist[max_v_i]) max_v_i = v_i;
  printf("distance[%u]=%f\n", max_v_i, dist[max_v_i]);
  printf("algo data: %lu MB, pre data: %lu MB\npre time: %lfs, exec time: %lf s\n", graph->exe_data >> 20,
         graph->prep_data >> 20, graph->prep_time, exec_time);

  graph->dealloc_vertex_array(distance);
  delete active_in;
  delete active_out;
}

int main(int argc, char** argv) {
  if (argc < 5) {  // R7
    printf("sssp [file] [vertices_num] [root] [eta] [partitions=1] [sockets=1]\n");
    return 1;
  }
  const int P = argc > 5 ? std::atoi(argv[5]) : 1;
  const int S = argc > 6 ? std::atoi(argv[6]) : 1;
  try {
    Graph<Weight>* graph = new Graph<Weight>(std::atoi(argv[4]), P, S);
    graph->use_undirected = false;
    graph->use_root = true;
    graph->load_directed(argv[1], std::atoi(argv[2]));
    UINT root = UINT(std::atoi(argv[3]));
    compute(graph, root);
    for (int run = 0; run < 3; run++) compute(graph, root);
    delete graph;
  } catch (const Error& e) {
    fprintf(stderr, "sssp: %s (status %d)\n", e.what(), e.status);
    return 2;
  }
  return 0;
}
