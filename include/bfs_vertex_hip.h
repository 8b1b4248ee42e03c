// bfs_vertex_hip.h -- drop-in for include/bfs_vertex_cuda.h (main at :120-191):
// usage: prog input_file_name source_node_number verify
#pragma once
#include "igx_harness.hpp"

int main(int argc, char* argv[])
{
  using namespace igx_harness;
  printf("bfs vertex-based HIP (%s)\n", __FILE__);
  if (argc != 4) {fprintf(stderr, "USAGE: %s input_file_name source_node_number verify\n", argv[0]); exit(-1);}
  igx_graph g = read_graph(argv[1]);
  printf("input: %s\nnodes: %d\nedges: %d\n", argv[1], g.nodes, g.edges);
  const int source = atoi(argv[2]);
  if ((source < 0) || (source >= g.nodes)) {fprintf(stderr, "ERROR: source_node_number must be between 0 and %d\n", g.nodes); exit(-1);}
  printf("source: %d\n", source);
  const int runveri = atoi(argv[3]);
  if ((runveri != 0) && (runveri != 1)) printf("has to be 0 (turn off) or 1 (turn on) verification");
  gpu_info(0);
  select_style(IGX_BFS, 0, IGX_STYLE);
  printf("style: %s\n", style_name(IGX_BFS, 0));
  std::vector<int> distance(g.nodes);
  igx_graph d_g = upload(g);
  timed_runs_with_cutoff([&] {
    double t = 0;
    check(igx_bfs_vertex(source, &d_g, distance.data(), &t), "igx_bfs_vertex");
    print_iterations();
    return t;
  }, g.edges);
  int maxnode = 0;
  for (int v = 1; v < g.nodes; v++) if (distance[maxnode] < distance[v]) maxnode = v;
  printf("vertex %d has maximum distance %d\n", maxnode, distance[maxnode]);
  if (runveri) verify_distances(source, g, false, distance.data());
  else printf("turn off verification\n\n");
  return finish(&d_g, &g);
}
