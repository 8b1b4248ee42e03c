// cc_edge_hip.h -- drop-in for include/cc_edge_cuda.h (main() of that header):
// usage: prog input_file_name verify
#pragma once
#include "igx_harness.hpp"

int main(int argc, char* argv[])
{
  using namespace igx_harness;
  printf("cc edge-based HIP (%s)\n", __FILE__);
  if (argc != 3) {fprintf(stderr, "USAGE: %s input_file_name verify\n", argv[0]); exit(-1);}
  igx_graph g = read_graph(argv[1]);
  printf("input: %s\nnodes: %d\nedges: %d\n", argv[1], g.nodes, g.edges);
  const int runveri = atoi(argv[2]);
  if ((runveri != 0) && (runveri != 1)) printf("has to be 0 (turn off) or 1 (turn on) verification");
  gpu_info(0);
  select_style(IGX_CC, 1, IGX_STYLE);
  printf("style: %s\n", style_name(IGX_CC, 1));
  std::vector<int> sp = starting_points(g);
  std::vector<int> label(g.nodes);
  igx_graph d_g = upload(g);
  const int runs = 3;
  double runtimes[runs];
  for (int i = 0; i < runs; i++) {
    check(igx_cc_edge(&d_g, label.data(), sp.data(), &runtimes[i]), "igx_cc_edge");
    print_iterations();
  }
  report(kCc, runtimes, runs, g.edges);
  std::set<int> s1(label.begin(), label.end());
  printf("number of connected components: %d\n", (int)s1.size());
  if (runveri) verify_components(g, label.data());
  else printf("verification turned off\n\n");
  return finish(&d_g, &g);
}
