// pr_hip.h -- drop-in for include/pr_cuda.h (main at :113-160): the entry gets
// the HOST graph and uploads its own copy, as PR_GPU does.
// usage: prog input_graph
#pragma once
#include <cmath>
#include "igx_harness.hpp"

int main(int argc, char* argv[])
{
  using namespace igx_harness;
  printf("PageRank HIP (%s)\n\n", __FILE__);
  if (argc != 2) {printf("USAGE: %s input_graph\n\n", argv[0]); exit(-1);}
  igx_graph g = read_graph(argv[1]);
  printf("graph: %s\nnodes: %d\nedges: %d\n", argv[1], g.nodes, g.edges);
  printf("avg degree: %.2f\n", 1.0 * g.edges / g.nodes);
  gpu_info(0);
  select_style(IGX_PR, 0, IGX_STYLE);
  printf("style: %s\n", style_name(IGX_PR, 0));
  const float kDamp = 0.85f, EPSILON = 0.0001f;  // pr_cuda.h:42-43
  std::vector<int> degree(g.nodes);
  for (int i = 0; i < g.nodes; i++) degree[i] = g.nindex[i + 1] - g.nindex[i];
  std::vector<float> scores(g.nodes, 1.0f / (float)g.nodes);
  double seconds = 0;
  int iterations = 0;
  check(igx_pr(&g, scores.data(), degree.data(), &seconds, &iterations), "igx_pr");
  printf("GPU iterations = %d.\n", iterations);
  printf("GPU runtime: %.6fs\n\n", seconds);
  printf("Throughput: %.6f gigaedges/s\n", 0.000000001 * g.edges / seconds);
  // pr_cuda.h:139-158: one more push iteration, report the L1 change
  const float base_score = (1.0f - kDamp) / (float)g.nodes;
  std::vector<float> incoming(g.nodes, 0.0f);
  for (int src = 0; src < g.nodes; src++) {
    const float contrib = scores[src] / degree[src];
    for (int i = g.nindex[src]; i < g.nindex[src + 1]; i++) incoming[g.nlist[i]] += contrib;
  }
  double error = 0;
  for (int i = 0; i < g.nodes; i++) error += fabs(base_score + kDamp * incoming[i] - scores[i]);
  if (error < EPSILON) printf("All good.\n");
  else printf("Total Error: %f\n", error);
  return finish(nullptr, &g);
}
