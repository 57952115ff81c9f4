// BFS drop-in: toolkits/bfs.cpp of the reference on the MI355X library.
// Parents follow the minimum-id rule (the reference's CAS race leaves the
// parent open).  An optional last argument "auto" enables the push/pull
// switch for this rooted program (the reference always pushes).  With
// EG_DUMP=<path> in the environment the first run also writes its parents
// there with dump_vertex_array (the reference's raw V * 4-byte format).
#include <cstdlib>

#include "graph.hpp"

using namespace endgraph;

void compute(Graph<Empty>* graph, UINT root) {
  double exec_time = 0;
  exec_time -= get_time();
  graph->exe_data = 0;

  UINT* parent = graph->alloc_vertex_array<UINT>();
  VertexSubset* visited = graph->alloc_vertex_subset();
  VertexSubset* active_in = graph->alloc_vertex_subset();
  VertexSubset* active_out = graph->alloc_vertex_subset();

  visited->clear();
  visited->set_bit(root);
  active_in->clear();
  active_in->set_bit(root);
  graph->fill_vertex_array(parent, graph->vertices_num);
  check(eg_memcpy_h2d(parent + root, &root, sizeof(UINT), nullptr));

  UINT active_vertices = 1;
  for (int i_i = 0; active_vertices > 0; i_i++) {
    printf("active(%d)>=%u\n", i_i, active_vertices);
    active_out->clear();
    active_vertices = graph->process_edges<UINT, UINT>(
        programs::BfsLevel{parent, active_out, visited, nullptr, UINT(i_i)}, active_in, visited);
    active_vertices = graph->process_vertices<UINT>(programs::BfsVisit{visited}, active_out);
    std::swap(active_in, active_out);
  }

  exec_time += get_time();

  dump_result(graph, parent);
  std::vector<UINT> par = graph->gather_vertex_array(parent, 0);
  UINT found_vertices = 0;
  for (UINT v_i = 0; v_i < graph->vertices_num; v_i++)
    if (par[v_i] < graph->vertices_num) found_vertices += 1;
  printf("found_vertices = %u\n", found_vertices);
  printf("algo data: %lu MB, pre data: %lu MB\npre time: %lf s, exec time: %lf\n", graph->exe_data >> 20,
         graph->prep_data >> 20, graph->prep_time, exec_time);

  graph->dealloc_vertex_array(parent);
  delete active_in;
  delete active_out;
  delete visited;
}

int main(int argc, char** argv) {
  if (argc < 5) {  // R7
    printf("bfs [file] [vertices_num] [root] [eta] [partitions=1] [sockets=1] [reference|auto]\n");
    return 1;
  }
  const int P = argc > 5 ? std::atoi(argv[5]) : 1;
  const int S = argc > 6 ? std::atoi(argv[6]) : 1;
  try {
    Graph<Empty>* graph = new Graph<Empty>(std::atoi(argv[4]), P, S);
    UINT root = UINT(std::atoi(argv[3]));
    graph->use_undirected = false;
    graph->use_root = true;
    if (argc > 7 && std::string(argv[7]) == "auto") graph->policy = EG_POLICY_AUTO;
    graph->load_directed(argv[1], std::atoi(argv[2]));
    compute(graph, root);
    for (int run = 0; run < 3; run++) compute(graph, root);
    delete graph;
  } catch (const Error& e) {
    fprintf(stderr, "bfs: %s (status %d)\n", e.what(), e.status);
    return 2;
  }
  return 0;
}
