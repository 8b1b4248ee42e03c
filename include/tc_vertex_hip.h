// tc_vertex_hip.h -- drop-in for include/tc_vertex_cuda.h (main() of that header);
// the count is uint64 (the reference's is int).
// usage: prog input_graph verify
#pragma once
#include "igx_harness.hpp"

int main(int argc, char* argv[])
{
  using namespace igx_harness;
  printf("Triangle counting vertex-centric HIP (%s)\n", __FILE__);
  if (argc != 3) {printf("USAGE: %s input_graph verify\n\n", argv[0]); exit(-1);}
  igx_graph g = read_graph(argv[1]);
  printf("graph: %s\nnodes: %d\nedges: %d\n", argv[1], g.nodes, g.edges);
  printf("avg degree: %.2f\n", 1.0 * g.edges / g.nodes);
  tc_checks(g);
  gpu_info(0);
  select_style(IGX_TC, 0, IGX_STYLE);
  printf("style: %s\n", style_name(IGX_TC, 0));
  igx_graph d_g = upload(g);

  unsigned long long count = 0;
  const int runs = 3;
  double runtimes[runs];
  for (int i = 0; i < runs; i++) check(igx_tc_vertex(&count, g.nodes, d_g.nindex, d_g.nlist, &runtimes[i]), "igx_tc_vertex");
  report(kGpuSpaced, runtimes, runs, g.edges);
  const int verify = atoi(argv[2]);
  if ((verify != 0) && (verify != 1)) printf("has to be 0 (turn off) or 1 (turn on) verification");
  if (verify) {
    const unsigned long long h_count = serial_triangles(g);
    if (h_count != count) printf("ERROR: host %llu device %llu", h_count, count);
    else printf("the pattern occurs %llu times\n\n", count);
  }

  return finish(&d_g, &g);
}
