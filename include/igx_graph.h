/*
 * igx_graph.h -- host-side graph substrate for the Indigo2 kernels on MI355X.
 *
 * The CSR type keeps the byte layout of the reference's `struct ECLgraph`
 * (include/ECLgraph.h:44-50): int nodes, int edges, int* nindex (nodes+1),
 * int* nlist (edges), int* eweight (edges or NULL).  `edges` counts directed
 * entries: every undirected edge is stored in both adjacency lists.
 *
 * Functions return an igx_status (0 = success) instead of calling exit(),
 * which is what the reference loader does on every failure
 * (ECLgraph.h:57-75).  Arrays handed out by this library are released with
 * igx_host_graph_free().
 *
 * Library: libindigo_graph.so (host C++/OpenMP, no GPU code).
 */
#ifndef IGX_GRAPH_H
#define IGX_GRAPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same field order, sizes and offsets as ECLgraph (ECLgraph.h:44-50). */
typedef struct igx_graph {
  int nodes;
  int edges;
  int* nindex;
  int* nlist;
  int* eweight;
} igx_graph;

typedef enum igx_status {
  IGX_OK = 0,
  IGX_ERR_ARG = 1,        /* bad argument (null pointer, out-of-range source, ...) */
  IGX_ERR_IO = 2,         /* file could not be opened / short read / short write */
  IGX_ERR_FORMAT = 3,     /* node or edge count too low, or inconsistent CSR */
  IGX_ERR_NOMEM = 4,      /* host or device allocation failed */
  IGX_ERR_HIP = 5,        /* a HIP runtime call failed */
  IGX_ERR_UNSORTED = 6,   /* adjacency not strictly increasing (or self loop) where required */
  IGX_ERR_NO_WEIGHTS = 7, /* algorithm needs eweight but the graph has none */
  IGX_ERR_STYLE = 8,      /* unknown style id for this algorithm */
  IGX_ERR_NODEV = 9,      /* no usable GPU */
  IGX_ERR_LIMIT = 10,     /* size beyond a documented limit */
  IGX_ERR_TIMEOUT = 11    /* run stopped at the caller's time limit (igx_set_time_limit) */
} igx_status;

/* ---- .egr file format (ECLgraph.h:52-96) ----------------------------------
 * int32 nodes, int32 edges, int32 nindex[nodes+1], int32 nlist[edges],
 * optional int32 eweight[edges].  A missing weight section gives eweight=NULL
 * (ECLgraph.h:70-76). */
int igx_egr_read(const char* path, igx_graph* out);            /* ECLgraph.h:52-80 */
int igx_egr_write(const igx_graph* g, const char* path);        /* ECLgraph.h:82-96 */
void igx_host_graph_free(igx_graph* g);                         /* ECLgraph.h:98-106 */

/* Structural checks used by the harnesses.
 * igx_check_csr: nindex monotone, nindex[0]==0, nindex[nodes]==edges, ids in range.
 * igx_check_sorted: every adjacency strictly increasing, i.e. sorted without
 *   duplicates or self loops (tc_vertex_cuda.h:157-165 rejects otherwise). */
int igx_check_csr(const igx_graph* g);
int igx_check_sorted(const igx_graph* g);

/* ---- seeded synthetic generators -----------------------------------------
 * All generators are counter-based (splitmix64 keyed by seed and item index),
 * so the result does not depend on the OpenMP thread count.  Each result is
 * symmetrized, deduplicated, free of self loops, with sorted adjacency, and
 * has eweight == NULL (see igx_gen_weights). */

/* Erdos-Renyi G(n, p) with p = avg_degree / (n - 1). */
int igx_gen_er(int n, double avg_degree, uint64_t seed, igx_graph* out);
/* R-MAT with 2^scale vertices and edge_factor * 2^scale samples, quadrant
 * probabilities a, b, c (d = 1-a-b-c), no noise and no relabelling.  If
 * n_limit > 0 the vertex count is n_limit (<= 2^scale) and samples with an
 * endpoint >= n_limit are dropped; the sample count is then
 * round(n_limit * edge_factor). */
int igx_gen_rmat(int scale, double edge_factor, double a, double b, double c,
                 int n_limit, uint64_t seed, igx_graph* out);
/* side x side 4-neighbour lattice plus, for each vertex with probability
 * long_frac, one uniformly random long-range partner. */
int igx_gen_grid(int side, double long_frac, uint64_t seed, igx_graph* out);
/* Chung-Lu power law: expected degree of vertex i proportional to
 * (i+1)^(-1/(gamma-1)), scaled to avg_degree. */
int igx_gen_chunglu(int n, double avg_degree, double gamma, uint64_t seed, igx_graph* out);
/* Build a graph from an undirected edge list (src[i], dst[i]); self loops
 * and duplicates are dropped and both directions are stored. */
int igx_graph_from_edges(int n, int64_t count, const int* src, const int* dst, igx_graph* out);

/* Symmetric int32 weights uniform in [lo, hi]: w(u,v) = w(v,u) is a hash of
 * (min(u,v), max(u,v), seed).  Allocates g->eweight. */
int igx_gen_weights(igx_graph* g, int lo, int hi, uint64_t seed);
/* The reference harnesses' fallback weights when a file has none:
 * w = |1 + (u * v) % nodes| with 32-bit int arithmetic (sssp_edge_cuda.h:127-137). */
int igx_ref_weights(igx_graph* g);

/* Starting-point array: sp[e] = source vertex of edge e (sssp_edge_cuda.h:149-155). */
int igx_build_sp_host(const igx_graph* g, int* sp);

/* Vertex with the highest degree, lowest id on ties. */
int igx_max_degree_vertex(const igx_graph* g);

/* ---- graph statistics for the style selector ------------------------------
 * Writes IGX_NUM_FEATURES doubles (order documented in graph_host.cpp). */
#define IGX_NUM_FEATURES 16
int igx_graph_features(const igx_graph* g, uint64_t seed, double* features);
/* The raw 64-bit draws of the seeded counter-based stream `item` that the
 * feature sampling uses (wedges: item 0xC1, three draws per wedge; edge
 * locality: item 0x10C, one per edge), so that the device version of the
 * features (igx_graph_features_device in indigo_hip.h) samples the same
 * wedges and edges. */
int igx_feature_draws(uint64_t seed, uint64_t item, int count, uint64_t* out);

#ifdef __cplusplus
}
#endif

#endif /* IGX_GRAPH_H */
