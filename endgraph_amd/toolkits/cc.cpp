// CC drop-in: toolkits/cc.cpp of the reference on the MI355X library
// (symmetrised graph, minimum-label propagation, undirected mode switch).
// With EG_DUMP=<path> in the environment the first run also writes its labels
// there with dump_vertex_array (the reference's raw V * 4-byte format).
#include <cstdlib>

#include "graph.hpp"

using namespace endgraph;

void compute(Graph<Empty>* graph) {
  double exec_time = 0;
  exec_time -= get_time();
  graph->exe_data = 0;

  UINT* label = graph->alloc_vertex_array<UINT>();
  VertexSubset* active_in = graph->alloc_vertex_subset();
  active_in->fill();
  VertexSubset* active_out = graph->alloc_vertex_subset();

  UINT active_vertices = graph->process_vertices<UINT>(programs::CcInit{label}, active_in);

  for (int i_i = 0; active_vertices > 0; i_i++) {
    printf("active(%d)>=%u\n", i_i, active_vertices);
    active_out->clear();
    active_vertices = graph->process_edges<UINT, UINT>(programs::CcStep{label, active_out}, active_in);
    std::swap(active_in, active_out);
  }

  exec_time += get_time();

  dump_result(graph, label);
  std::vector<UINT> lab = graph->gather_vertex_array(label, 0);
  std::vector<UINT> count(graph->vertices_num, 0);
  for (UINT v_i = 0; v_i < graph->vertices_num; v_i++) count[lab[v_i]] += 1;
  UINT components = 0;
  for (UINT v_i = 0; v_i < graph->vertices_num; v_i++)
    if (count[v_i] > 0) components += 1;
  printf("components = %u\n", components);
  printf("algo data: %lu MB, pre data: %lu MB\npre time: %lf s, exec time: %lf s \n", graph->exe_data >> 20,
         graph->prep_data >> 20, graph->prep_time, exec_time);

  graph->dealloc_vertex_array(label);
  delete active_in;
  delete active_out;
}

int main(int argc, char** argv) {
  if (argc < 5) {  // R7: the reference checks argc < 3 but reads argv[4]
    printf("cc [file] [vertices_num] [null] [eta] [partitions=1] [sockets=1]\n");
    return 1;
  }
  const int P = argc > 5 ? std::atoi(argv[5]) : 1;
  const int S = argc > 6 ? std::atoi(argv[6]) : 1;
  try {
    Graph<Empty>* graph = new Graph<Empty>(std::atoi(argv[4]), P, S);
    graph->use_undirected = true;
    graph->load_undirected_from_directed(argv[1], std::atoi(argv[2]));
    compute(graph);
    for (int run = 0; run < 3; run++) compute(graph);
    delete graph;
  } catch (const Error& e) {
    fprintf(stderr, "cc: %s (status %d)\n", e.what(), e.status);
    return 2;
  }
  return 0;
}
