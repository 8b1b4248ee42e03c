// igx_harness.hpp -- shared host code of the drop-in harness headers
// (include/<algo>_<kind>_hip.h).  Each of those headers plays the role of the
// reference's include/<algo>_<kind>_cuda.h: it defines main() with the same
// command line, prints the same lines ("runtime: ...", "Throughput: ...
// gigaedges/s", "verification passed", ...), runs the entry point three
// times, reports the median, and verifies the result with the harness's serial
// algorithm.  A program is one translation unit that selects a style and
// includes one harness header, e.g.
//
//   #define IGX_STYLE "do_hybrid"
//   #include "bfs_vertex_hip.h"
//
// and is built with g++ -Iinclude prog.cpp -Lindigo2suite_amd -lindigo_hip.
#pragma once

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <queue>
#include <set>
#include <utility>
#include <vector>

#include "indigo_hip.h"

namespace igx_harness {

// readECLgraph (ECLgraph.h:52-80) behaviour: print the counts, exit(-1) on error
inline igx_graph read_graph(const char* path)
{
  setvbuf(stdout, nullptr, _IOLBF, 0);
  igx_graph g = {};
  const int rc = igx_egr_read(path, &g);
  if (rc != IGX_OK) {
    fprintf(stderr, "ERROR: could not read %s (%s)\n\n", path, igx_status_string(rc));
    exit(-1);
  }
  printf("input graph: %d nodes and %d edges\n", g.nodes, g.edges);
  return g;
}

inline void check(int rc, const char* what)
{
  if (rc != IGX_OK) {
    fprintf(stderr, "ERROR: %s failed: %s %s\n", what, igx_status_string(rc), igx_last_error());
    igx_shutdown();
    exit(-1);
  }
}

// end of main(): release the library's device and pinned host resources
// while the HIP runtime is fully alive, before static destruction starts
inline int finish(igx_graph* d_g, igx_graph* g)
{
  if (d_g) igx_graph_free_device(d_g);
  if (g) igx_host_graph_free(g);
  check(igx_shutdown(), "igx_shutdown");
  return 0;
}

// GPUinfo (bfs_vertex_cuda.h:63-73): select the device and print it
inline int gpu_info(int d)
{
  setvbuf(stdout, nullptr, _IOLBF, 0);  // keep lines visible if the program dies

  check(igx_init(d), "igx_init");
  char name[256];
  int cus = 0, mt = 0;
  check(igx_device_info(name, sizeof(name), &cus, &mt), "igx_device_info");
  printf("GPU: %s with %d CUs and %d mTpCU\n", name, cus, mt);
  return cus * mt;
}

// select the style named by IGX_STYLE (if defined) for the entry of this kind
inline void select_style(int algo, int edge_kind, const char* name)
{
  if (name == nullptr || name[0] == 0) return;
  for (int s = 0; s < igx_num_styles(algo); s++) {
    if (strcmp(igx_style_name(algo, s), name) == 0) {
      check(igx_set_entry_style(algo, edge_kind, s), "igx_set_entry_style");
      return;
    }
  }
  fprintf(stderr, "ERROR: unknown style %s\n", name);
  exit(-1);
}

inline const char* style_name(int algo, int edge_kind)
{
  return igx_style_name(algo, igx_get_entry_style(algo, edge_kind));
}

// median (bfs_vertex_cuda.h:111-118)
inline double median(double* a, int n)
{
  std::sort(a, a + n);
  return (n % 2 == 0) ? (a[(n - 1) / 2] + a[n / 2]) / 2.0 : a[n / 2];
}

inline igx_graph upload(const igx_graph& g)
{
  igx_graph d = {};
  check(igx_graph_upload(&g, &d), "igx_graph_upload");
  return d;
}

// harness fallback weights (sssp_vertex_cuda.h:127-137)
inline void ensure_weights(igx_graph& g)
{
  if (g.eweight == nullptr) {
    printf("Generating weights.\n");
    check(igx_ref_weights(&g), "igx_ref_weights");
  }
}

inline std::vector<int> starting_points(const igx_graph& g)
{
  std::vector<int> sp((size_t)std::max(g.edges, 1));
  check(igx_build_sp_host(&g, sp.data()), "igx_build_sp_host");
  return sp;
}

// CPUserialDijkstra (bfs_vertex_cuda.h:75-99 with unit weights,
// sssp_edge_cuda.h:74-99 with eweight)
inline void serial_dijkstra(int src, const igx_graph& g, bool weighted, std::vector<int>& dist)
{
  dist.assign(g.nodes, INT_MAX);
  dist[src] = 0;
  typedef std::pair<int, int> P;
  std::priority_queue<P, std::vector<P>, std::greater<P>> pq;
  pq.push(P(0, src));
  while (!pq.empty()) {
    const P top = pq.top();
    pq.pop();
    const int v = top.second;
    if (top.first > dist[v]) continue;
    for (int i = g.nindex[v]; i < g.nindex[v + 1]; i++) {
      const int n = g.nlist[i];
      const int d = dist[v] + (weighted ? g.eweight[i] : 1);
      if (d < dist[n]) {
        dist[n] = d;
        pq.push(P(d, n));
      }
    }
  }
}

// The reference's result lines differ by harness; each harness header names
// its form:
//   kPlain      "runtime: %.6fs" / "Throughput: %.6f gigaedges/s"  (bfs_*_cuda.h:162-163, mis_vertex_cuda.h:160-161)
//   kGpuSpaced  "GPU runtime: %.6f s" / "GPU Throughput: ..."      (sssp_*_cuda.h:166-167, tc_*_cuda.h:185-186)
//   kCc         "runtime: %.6fs" / "GPU Throughput: ..."           (cc_*_cuda.h:131-132)
//   kMisEdge    "runtime: %.6fs" / "Throughput: %.6f gigabytes/s"  (mis_edge_cuda.h:170-171)
enum ReportForm { kPlain, kGpuSpaced, kCc, kMisEdge };

inline void report_time(ReportForm form, double t, int edges)
{
  const double gteps = 0.000000001 * edges / t;
  switch (form) {
    case kPlain: printf("runtime: %.6fs\nThroughput: %.6f gigaedges/s\n", t, gteps); break;
    case kGpuSpaced: printf("GPU runtime: %.6f s\nGPU Throughput: %.6f gigaedges/s\n", t, gteps); break;
    case kCc: printf("runtime: %.6fs\nGPU Throughput: %.6f gigaedges/s\n", t, gteps); break;
    case kMisEdge: printf("runtime: %.6fs\nThroughput: %.6f gigabytes/s\n", t, gteps); break;
  }
}

inline void report(ReportForm form, double* runtimes, int runs, int edges)
{
  report_time(form, median(runtimes, runs), edges);
}

// iteration count of the entry call just made, printed as the reference's
// variants print it inside their GPU* function (e.g. [C-BVT]:111)
inline int last_iterations()
{
  igx_run_info info = {};
  check(igx_last_run_info(&info), "igx_last_run_info");
  return info.iterations;
}

inline void print_iterations() { printf("iterations: %d\n", last_iterations()); }

// BFS harness timing loop (bfs_vertex_cuda.h:147-164): three runs and their
// median, except that a run of 30 s or more is reported alone and ends the loop
template <typename Run>
inline void timed_runs_with_cutoff(Run run, int edges)
{
  const int runs = 3;
  double runtimes[runs];
  for (int i = 0; i < runs; i++) {
    runtimes[i] = run();
    if (runtimes[i] >= 30.0) {
      report_time(kPlain, runtimes[i], edges);
      return;
    }
  }
  report(kPlain, runtimes, runs, edges);
}

// verify BFS / SSSP distances exactly (bfs_vertex_cuda.h:174-184)
inline void verify_distances(int src, const igx_graph& g, bool weighted, const int* dist)
{
  std::vector<int> ref;
  serial_dijkstra(src, g, weighted, ref);
  for (int v = 0; v < g.nodes; v++) {
    if (dist[v] != ref[v]) {
      fprintf(stderr, "ERROR: verification failed for node %d: %d instead of %d\n", v, dist[v], ref[v]);
      exit(-1);
    }
  }
  printf("verification passed\n\n");
}

// CC verification (cc_vertex_cuda.h:142-167), with an explicit stack
inline void verify_components(const igx_graph& g, const int* label_in)
{
  const int n = g.nodes;
  for (int v = 0; v < n; v++) {
    for (int i = g.nindex[v]; i < g.nindex[v + 1]; i++) {
      if (label_in[g.nlist[i]] != label_in[v]) {fprintf(stderr, "ERROR: found adjacent nodes in different components\n\n"); exit(-1);}
    }
  }
  for (int v = 0; v < n; v++) {
    if (label_in[v] >= n) {fprintf(stderr, "ERROR: found sentinel number\n\n"); exit(-1);}
  }
  std::set<int> s1(label_in, label_in + n);
  std::vector<int> label(label_in, label_in + n);
  std::vector<int> stack;
  int count = 0;
  for (int v = 0; v < n; v++) {
    if (label[v] < n) {
      count++;
      const int id = label[v];
      stack.push_back(v);
      while (!stack.empty()) {
        const int x = stack.back();
        stack.pop_back();
        if (label[x] < n) {
          if (label[x] != id) {fprintf(stderr, "ERROR: found incorrect ID value\n\n"); exit(-1);}
          label[x] = n;
          for (int i = g.nindex[x]; i < g.nindex[x + 1]; i++) stack.push_back(g.nlist[i]);
        }
      }
    }
  }
  if ((int)s1.size() != count) {fprintf(stderr, "ERROR: component IDs are not unique\n\n"); exit(-1);}
  printf("verification passed\n\n");
}

// MIS verification (mis_vertex_cuda.h:80-99)
inline void verify_mis(const igx_graph& g, const int* status)
{
  for (int v = 0; v < g.nodes; v++) {
    if ((status[v] != 1) && (status[v] != 2)) {fprintf(stderr, "ERROR: found undecided node\n"); exit(-1);}
    bool in_nbr = false;
    for (int i = g.nindex[v]; i < g.nindex[v + 1]; i++) {
      if (g.nlist[i] != v && status[g.nlist[i]] == 1) in_nbr = true;
    }
    if (status[v] == 1 && in_nbr) {fprintf(stderr, "ERROR: found adjacent nodes in MIS\n"); exit(-1);}
    if (status[v] == 2 && !in_nbr) {fprintf(stderr, "ERROR: set is not maximal\n"); exit(-1);}
  }
}

// h_triCounting (tc_vertex_cuda.h:98-117) with a uint64 count
inline unsigned long long serial_triangles(const igx_graph& g)
{
  unsigned long long count = 0;
  for (int v = 0; v < g.nodes; v++) {
    const int end1 = g.nindex[v + 1];
    int start1 = end1;
    while (g.nindex[v] < start1 && v < g.nlist[start1 - 1]) start1--;
    for (int j = start1; j < end1; j++) {
      const int u = g.nlist[j];
      const int end2 = g.nindex[u + 1];
      int start2 = end2;
      while (g.nindex[u] < start2 && u < g.nlist[start2 - 1]) start2--;
      int p1 = j + 1, p2 = start2;
      while (p1 < end1 && p2 < end2) {
        if (g.nlist[p1] < g.nlist[p2]) p1++;
        else if (g.nlist[p2] < g.nlist[p1]) p2++;
        else {count++; p1++; p2++;}
      }
    }
  }
  return count;
}

// tc_vertex_cuda.h:150-165: max degree info line and the sortedness check
inline void tc_checks(const igx_graph& g)
{
  int mdeg = 0;
  for (int v = 0; v < g.nodes; v++) mdeg = std::max(mdeg, g.nindex[v + 1] - g.nindex[v]);
  printf("max degree: %d\n\n", mdeg);
  if (igx_check_sorted(&g) != IGX_OK) {
    printf("ERROR: adjacency list not sorted or contains self edge\n");
    exit(-1);
  }
}

}  // namespace igx_harness

#ifndef IGX_STYLE
#define IGX_STYLE ""
#endif
