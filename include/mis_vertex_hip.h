// mis_vertex_hip.h -- drop-in for include/mis_vertex_cuda.h (main at :110-171):
// usage: prog input_file_name verify
#pragma once
#include "igx_harness.hpp"

int main(int argc, char* argv[])
{
  using namespace igx_harness;
  printf("mis vertex-based HIP (%s)\n", __FILE__);
  if (argc != 3) {fprintf(stderr, "USAGE: %s input_file_name verify\n", argv[0]); exit(-1);}
  igx_graph g = read_graph(argv[1]);
  printf("input: %s\nnodes: %d\nedges: %d\n", argv[1], g.nodes, g.edges);
  const int runveri = atoi(argv[2]);
  if ((runveri != 0) && (runveri != 1)) printf("has to be 0 (turn off) or 1 (turn on) verification");
  gpu_info(0);
  select_style(IGX_MIS, 0, IGX_STYLE);
  printf("style: %s\n", style_name(IGX_MIS, 0));
  std::vector<unsigned int> priority(g.nodes);
  std::vector<int> status(g.nodes);
  igx_graph d_g = upload(g);
  const int runs = 3;
  double runtimes[runs];
  for (int i = 0; i < runs; i++) {
    check(igx_mis_vertex(&d_g, priority.data(), status.data(), &runtimes[i]), "igx_mis_vertex");
    int cnt = 0;
    for (int v = 0; v < g.nodes; v++) cnt += (status[v] == 1);
    printf("iterations: %d,  elements in set: %d (%.1f%%)\n", last_iterations(), cnt, 100.0 * cnt / g.nodes);
    if (runveri) verify_mis(g, status.data());
    else printf("turn off verification\n\n");
  }
  report(kPlain, runtimes, runs, g.edges);
  return finish(&d_g, &g);
}
