// ORACLE -- test infrastructure, not product code.
//
// Dump driver for the reference's own OpenMP and C++-threads variants.  oracle/Makefile
// compiles this file once per variant with -DREF_FILE="<variant path under
// the reference tree>" and -DREF_KIND=<n>.  The variant's main() is renamed
// away; this main() reads an .egr file, calls the variant's entry point once
// and writes the raw result array to a file, so tests can compare the
// restatement in oracle.cpp with the reference code itself.  Nothing is copied
// from the reference: the variant is included from where it lies.
//
// usage: ref_<name> input.egr output.bin [source [threads]]
#define main indigo_ref_main
#include REF_FILE
#undef main

#include <cstdio>
#include <cstdlib>

static void dump(const char* path, const void* data, size_t bytes)
{
  FILE* f = fopen(path, "wb");
  if (f == NULL) {fprintf(stderr, "cannot open %s\n", path); exit(2);}
  if (bytes > 0 && fwrite(data, 1, bytes, f) != bytes) {fprintf(stderr, "short write\n"); exit(2);}
  fclose(f);
}

static int ref_threads(int argc, char* argv[]) { return (argc > 4) ? atoi(argv[4]) : 4; }

static int* make_sp(const ECLgraph& g)
{
  int* sp = new int [g.edges > 0 ? g.edges : 1];
  for (int i = 0; i < g.nodes; i++) {
    for (int j = g.nindex[i]; j < g.nindex[i + 1]; j++) sp[j] = i;
  }
  return sp;
}

int main(int argc, char* argv[])
{
  if (argc < 3) {fprintf(stderr, "usage: %s input.egr output.bin [source]\n", argv[0]); return 2;}
  ECLgraph g = readECLgraph(argv[1]);
  const int src = (argc > 3) ? atoi(argv[3]) : 0;
#if REF_KIND == 1   /* BFS vertex: CPUbfs_vertex(src, g, dist) */
  data_type* out = new data_type [g.nodes];
  CPUbfs_vertex(src, g, out);
  dump(argv[2], out, sizeof(data_type) * g.nodes);
#elif REF_KIND == 2 /* SSSP edge: CPUsssp_edge(src, g, dist, sp) */
  if (g.eweight == NULL) {fprintf(stderr, "weights required\n"); return 2;}
  int* sp = make_sp(g);
  data_type* out = new data_type [g.nodes];
  CPUsssp_edge(src, g, out, sp);
  dump(argv[2], out, sizeof(data_type) * g.nodes);
#elif REF_KIND == 3 /* SSSP vertex: CPUsssp_vertex(src, g, dist) */
  if (g.eweight == NULL) {fprintf(stderr, "weights required\n"); return 2;}
  data_type* out = new data_type [g.nodes];
  CPUsssp_vertex(src, g, out);
  dump(argv[2], out, sizeof(data_type) * g.nodes);
#elif REF_KIND == 4 /* CC vertex: CPUcc_vertex(g, label) */
  data_type* out = new data_type [g.nodes];
  CPUcc_vertex(g, out);
  dump(argv[2], out, sizeof(data_type) * g.nodes);
#elif REF_KIND == 5 /* MIS vertex: OMPmis_vertex(g, priority, status) */
  data_type* prio = new data_type [g.nodes];
  unsigned char* out = new unsigned char [g.nodes];
  OMPmis_vertex(g, prio, out);
  dump(argv[2], out, g.nodes);
#elif REF_KIND == 6 /* MIS edge: OMPmis_edge(g, sp, priority, status) */
  int* sp = make_sp(g);
  data_type* prio = new data_type [g.nodes];
  unsigned char* out = new unsigned char [g.nodes];
  OMPmis_edge(g, sp, prio, out);
  dump(argv[2], out, g.nodes);
#elif REF_KIND == 7 /* PR: PR_CPU(g, scores, degree), scores start at 1/n */
  int* degree = new int [g.nodes];
  for (int i = 0; i < g.nodes; i++) degree[i] = g.nindex[i + 1] - g.nindex[i];
  score_type* out = new score_type [g.nodes];
  for (int i = 0; i < g.nodes; i++) out[i] = 1.0f / (score_type)g.nodes;
  PR_CPU(g, out, degree);
  dump(argv[2], out, sizeof(score_type) * g.nodes);
#elif REF_KIND == 8 /* TC vertex: CPUtc_vertex(count, nodes, nindex, nlist) */
  basic_t count = 0;
  CPUtc_vertex(count, g.nodes, g.nindex, g.nlist);
  long long c = count;
  dump(argv[2], &c, sizeof(c));
#elif REF_KIND == 9 /* TC edge: CPUtc_edge(count, edges, nindex, nlist, sp) */
  int* sp = make_sp(g);
  basic_t count = 0;
  CPUtc_edge(count, g.edges, g.nindex, g.nlist, sp);
  long long c = count;
  dump(argv[2], &c, sizeof(c));
/* C++-threads variants (codes/cpp): the entry takes a thread count, argv[4] */
#elif REF_KIND == 10 /* BFS: CPPbfs_vertex(src, g, dist, threads) */
  data_type* out = new data_type [g.nodes];
  CPPbfs_vertex(src, g, out, ref_threads(argc, argv));
  dump(argv[2], out, sizeof(data_type) * g.nodes);
#elif REF_KIND == 11 /* SSSP: CPUsssp_vertex(src, g, dist, threads) */
  if (g.eweight == NULL) {fprintf(stderr, "weights required\n"); return 2;}
  data_type* out = new data_type [g.nodes];
  CPUsssp_vertex(src, g, out, ref_threads(argc, argv));
  dump(argv[2], out, sizeof(data_type) * g.nodes);
#elif REF_KIND == 12 /* CC: CPUcc_vertex(g, label, threads) */
  data_type* out = new data_type [g.nodes];
  CPUcc_vertex(g, out, ref_threads(argc, argv));
  dump(argv[2], out, sizeof(data_type) * g.nodes);
#elif REF_KIND == 13 /* MIS: CPPmis_vertex(g, priority, status, threads) */
  data_type* prio = new data_type [g.nodes];
  unsigned char* out = new unsigned char [g.nodes];
  CPPmis_vertex(g, prio, out, ref_threads(argc, argv));
  dump(argv[2], out, g.nodes);
#elif REF_KIND == 14 /* PR: PR_CPU(g, scores, degree, threads), scores start at 1/n */
  int* degree = new int [g.nodes];
  for (int i = 0; i < g.nodes; i++) degree[i] = g.nindex[i + 1] - g.nindex[i];
  score_type* out = new score_type [g.nodes];
  for (int i = 0; i < g.nodes; i++) out[i] = 1.0f / (score_type)g.nodes;
  PR_CPU(g, out, degree, ref_threads(argc, argv));
  dump(argv[2], out, sizeof(score_type) * g.nodes);
#elif REF_KIND == 15 /* TC: CPUtc_vertex(count, nodes, nindex, nlist, threads) */
  basic_t count = 0;
  CPUtc_vertex(count, g.nodes, g.nindex, g.nlist, ref_threads(argc, argv));
  long long c = count;
  dump(argv[2], &c, sizeof(c));
#elif REF_KIND == 16 /* BFS edge: CPUbfs_edge(src, g, dist, sp) */
  int* sp = make_sp(g);
  data_type* out = new data_type [g.nodes];
  CPUbfs_edge(src, g, out, sp);
  dump(argv[2], out, sizeof(data_type) * g.nodes);
#elif REF_KIND == 17 /* CC edge: CPUcc_edge(g, label, sp) */
  int* sp = make_sp(g);
  data_type* out = new data_type [g.nodes];
  CPUcc_edge(g, out, sp);
  dump(argv[2], out, sizeof(data_type) * g.nodes);
#else
#error "unknown REF_KIND"
#endif
  return 0;
}
