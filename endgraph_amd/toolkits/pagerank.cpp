// PageRank drop-in: toolkits/pagerank.cpp of the reference on the MI355X
// library.  Same compute() structure and printouts; the lambdas become the
// PageRank programs of the facade.  An optional last argument "fp32c" runs
// the FP32 gathers with compensated sums (residual and update in FP64).  With
// EG_DUMP=<path> in the environment the first run also writes its ranks
// there with dump_vertex_array (the reference's raw V * 8-byte format).
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "graph.hpp"

using namespace endgraph;

void compute(Graph<Empty>* graph, int iterations, bool fp32c) {
  double exec_time = 0;
  exec_time -= get_time();
  graph->exe_data = 0;

  double* curr = graph->alloc_vertex_array<double>();
  double* next = graph->alloc_vertex_array<double>();
  float* curr32 = fp32c ? graph->alloc_vertex_array<float>() : nullptr;
  VertexSubset* active = graph->alloc_vertex_subset();
  active->fill();

  double delta = graph->process_vertices<double>(programs::PageRankInit{curr, curr32}, active);
  delta /= graph->vertices_num;

  for (int i_i = 0; i_i < iterations; i_i++) {
    printf("delta(%d)=%lf\n", i_i, delta);
    graph->fill_vertex_array(next, (double)0);
    graph->process_edges<int, double>(programs::PageRankPull{curr, next, curr32}, active);
    delta = graph->process_vertices<double>(programs::PageRankApply{curr, next, i_i == iterations - 1, curr32},
                                            active);
    delta /= graph->vertices_num;
    std::swap(curr, next);
  }

  exec_time += get_time();

  dump_result(graph, curr);
  std::vector<double> pr = graph->gather_vertex_array(curr, 0);
  double pr_sum = 0;
  UINT max_v_i = 0;
  for (UINT v_i = 0; v_i < graph->vertices_num; v_i++) {
    pr_sum += pr[v_i];
    if (pr[v_i] > pr[max_v_i]) max_v_i = v_i;
  }
  printf("pr_sum=%lf\n", pr_sum);
  printf("pr[%u]=%lf\n", max_v_i, pr[max_v_i]);
  printf("algo data: %lu MB, pre data: %lu MB\npre time: %lf s exec time: %lf s\n", graph->exe_data >> 20,
         graph->prep_data >> 20, graph->prep_time, exec_time);

  graph->dealloc_vertex_array(curr);
  graph->dealloc_vertex_array(next);
  if (curr32) graph->dealloc_vertex_array(curr32);
  delete active;
}

int main(int argc, char** argv) {
  if (argc < 5) {  // R7: eta is required
    printf("pagerank [file] [vertices_num] [iterations] [eta] [partitions=1] [sockets=1] [fp64|fp32c]\n");
    return 1;
  }
  const int iterations = std::atoi(argv[3]);
  const int P = argc > 5 ? std::atoi(argv[5]) : 1;
  const int S = argc > 6 ? std::atoi(argv[6]) : 1;
  const bool fp32c = argc > 7 && std::string(argv[7]) == "fp32c";
  try {
    Graph<Empty>* graph = new Graph<Empty>(std::atoi(argv[4]), P, S);
    // R2: the mode flags and max_iterations are set before load, where the
    // partition weights read them
    graph->use_undirected = false;
    graph->use_iteration = true;
    graph->max_iterations = iterations;
    graph->load_directed(argv[1], std::atoi(argv[2]));
    compute(graph, iterations, fp32c);
    for (int run = 0; run < 3; run++) compute(graph, iterations, fp32c);
    delete graph;
  } catch (const Error& e) {
    fprintf(stderr, "pagerank: %s (status %d)\n", e.what(), e.status);
    return 2;
  }
  return 0;
}
