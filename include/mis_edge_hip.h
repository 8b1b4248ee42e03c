// mis_edge_hip.h -- drop-in for include/mis_edge_cuda.h (main at :110-181);
// as in the reference, sp is built on the host and handed over as a device array.
// usage: prog input_file_name verify
#pragma once
#include "igx_harness.hpp"

int main(int argc, char* argv[])
{
  using namespace igx_harness;
  printf("mis edge-based HIP (%s)\n", __FILE__);
  if (argc != 3) {fprintf(stderr, "USAGE: %s input_file_name verify\n", argv[0]); exit(-1);}
  igx_graph g = read_graph(argv[1]);
  printf("input: %s\nnodes: %d\nedges: %d\n", argv[1], g.nodes, g.edges);
  const int runveri = atoi(argv[2]);
  if ((runveri != 0) && (runveri != 1)) printf("has to be 0 (turn off) or 1 (turn on) verification");
  gpu_info(0);
  select_style(IGX_MIS, 1, IGX_STYLE);
  printf("style: %s\n", style_name(IGX_MIS, 1));
  std::vector<int> sp = starting_points(g);
  std::vector<unsigned int> priority(g.nodes);
  std::vector<int> status(g.nodes);
  igx_graph d_g = upload(g);
  void* d_sp = nullptr;
  check(igx_alloc(&d_sp, sizeof(int) * sp.size()), "igx_alloc");
  check(igx_memcpy_h2d(d_sp, sp.data(), sizeof(int) * sp.size()), "igx_memcpy_h2d");
  const int runs = 3;
  double runtimes[runs];
  for (int i = 0; i < runs; i++) {
    check(igx_mis_edge(&d_g, (const int*)d_sp, priority.data(), status.data(), &runtimes[i]), "igx_mis_edge");
    int cnt = 0;
    for (int v = 0; v < g.nodes; v++) cnt += (status[v] == 1);
    printf("iterations: %d,  elements in set: %d (%.1f%%)\n", last_iterations(), cnt, 100.0 * cnt / g.nodes);
    if (runveri) verify_mis(g, status.data());
    else printf("turn off verification\n\n");
  }
  report(kMisEdge, runtimes, runs, g.edges);
  igx_free(d_sp);
  return finish(&d_g, &g);
}
