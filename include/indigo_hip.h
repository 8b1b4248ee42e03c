/*
 * indigo_hip.h -- C ABI of libindigo_hip.so, the MI355X (gfx950) implementation
 * of the six Indigo2 graph kernels: BFS, SSSP, CC, MIS, PageRank and triangle
 * counting.
 *
 * The reference binds each algorithm at compile time: a variant file defines
 * the static entry point that its harness in include/<algo>_<kind>_cuda.h
 * forward-declares and calls from main().  The entry points below replace
 * those functions one for one (same argument meaning, same ownership: device
 * CSR in, host result out, elapsed seconds of the timed region out).  They
 * return an igx_status instead of printing and continuing or calling exit().
 *
 * A general entry, igx_run(), exposes every registered style of an algorithm
 * with device-resident input and output; it is what the autotuner, the style
 * selector and the benchmarks use.
 *
 * Threading: one device per process, calls from one host thread.  All work is
 * issued on the library's own stream (default priority).
 */
#ifndef INDIGO_HIP_H
#define INDIGO_HIP_H

#include <stddef.h>
#include <stdint.h>
#include "igx_graph.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------- runtime ---- */
/* Select the device and create the stream, events and control block.
 * Replaces GPUinfo()'s cudaSetDevice (bfs_vertex_cuda.h:63-73).  Idempotent. */
int igx_init(int device);
/* Device name, compute units and max threads per CU (GPUinfo's printout). */
int igx_device_info(char* name, int name_len, int* cus, int* max_threads_per_cu);
/* Wait for all work issued by the library. */
int igx_sync(void);
/* Release cached device workspaces (kept between runs so that allocation
 * stays outside the timed region). */
int igx_release_workspace(void);
/* No reference counterpart.  Release everything the library holds on the
 * device and the host (workspaces, the batched mode's buffers and captured
 * graph, events, the control block and its pinned mirror, the stream), so a
 * program can end with the HIP runtime holding nothing of the library's;
 * the drop-in harnesses call it before main returns.  A later call
 * initialises the library again. */
int igx_shutdown(void);
/* Text of the most recent HIP error recorded by the library. */
const char* igx_last_error(void);
const char* igx_status_string(int status);

/* -------------------------------------------------------- device memory ---- */
int igx_alloc(void** d_ptr, size_t bytes);
int igx_free(void* d_ptr);
int igx_memcpy_h2d(void* d_dst, const void* h_src, size_t bytes);
int igx_memcpy_d2h(void* h_dst, const void* d_src, size_t bytes);
int igx_memset(void* d_ptr, int value, size_t bytes);
/* device-to-device copy on the library stream (waits for completion) */
int igx_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes);

/* Device copy of a host CSR: same struct, device pointers (eweight only if
 * the host graph has weights).  Replaces the harness upload
 * (bfs_vertex_cuda.h:140-144, sssp_edge_cuda.h:160-165). */
int igx_graph_upload(const igx_graph* host, igx_graph* dev);
int igx_graph_free_device(igx_graph* dev);
/* Device starting-point array sp[e] = source of edge e, built on the device
 * (the reference builds it on the host: bfs_edge_cuda.h:140-146).  The caller
 * frees *d_sp with igx_free. */
int igx_build_sp(const igx_graph* dev, int** d_sp);

/* --------------------------------------------------------------- styles ---- */
enum igx_algo { IGX_BFS = 0, IGX_SSSP = 1, IGX_CC = 2, IGX_MIS = 3, IGX_PR = 4, IGX_TC = 5, IGX_NUM_ALGOS = 6 };

int igx_num_styles(int algo);
const char* igx_style_name(int algo, int style);
/* 1 if the style consumes sp[] (edge-based). */
int igx_style_edge_based(int algo, int style);
/* Style used by the reference-signature entries below; settable per algorithm
 * (vertex_kind selects the *_vertex or *_edge entry). */
int igx_get_entry_style(int algo, int edge_kind);
int igx_set_entry_style(int algo, int edge_kind, int style);

typedef struct igx_run_info {
  double seconds;                  /* timed region: first kernel to convergence */
  int iterations;                  /* rounds / levels / PR iterations */
  int launches;                    /* kernels enqueued for the rounds */
  unsigned long long count;        /* TC: triangles */
  double residual;                 /* PR: L1 change of the last iteration (FP64) */
  unsigned long long edges_processed;    /* styles that count their work: edges touched over all rounds */
  unsigned long long vertices_processed; /* and rows / queue entries touched */
  unsigned long long edges_relaxed;      /* of the edges touched, those whose weight and
                                            neighbour state were read (SSSP) */
} igx_run_info;

/* General entry.  `dev` is a device CSR (igx_graph_upload); d_out is device
 * memory:
 *   BFS, SSSP: int32 dist[nodes] (INT_MAX = unreachable), src = source
 *   CC:        int32 label[nodes] (minimum vertex id of the component)
 *   MIS:       int32 status[nodes] (1 included, 2 excluded)
 *   PR:        float scores[nodes]
 *   TC:        unsigned long long count (one element)
 * d_aux: the device sp[] for edge-based styles (NULL: built and cached by the
 * library). */
int igx_run(int algo, int style, const igx_graph* dev, int src, void* d_out, const int* d_aux_sp,
            igx_run_info* info);

/* Run info (iterations, launches, seconds) of the last reference-signature
 * entry call below: what the reference's variant prints inside its GPU*
 * function ("iterations: %d", e.g. [C-BVT]:111; MIS "iterations: %d,
 * elements in set: ...", [C-MVT]:122), which the entries' signatures do not
 * return. */
int igx_last_run_info(igx_run_info* info);

/* Per-kernel timing (for roofline measurements).  After
 * igx_profile_kernel("uf_hook") every launch of that kernel is bracketed by
 * events on the library stream; igx_profile_read() waits for the stream and
 * returns the summed duration and the number of timed launches, then resets.
 * Tags: "uf_hook", "af_link_rest", "bfs_do_level", "sssp_do_round", "sssp_bal_round", "pr_bal_main",
 * "tc_oriented_count", "mis_vdata_pull".  NULL or "" disables. */
int igx_profile_kernel(const char* tag);
int igx_profile_read(double* total_ms, int* launches);

/* PageRank convergence parameters (pr_cuda.h:42-45 fixes eps = 1e-4 and
 * max_iter = 100; those are the defaults). */
int igx_set_pr_params(double eps, int max_iter);

/* No reference counterpart.  Wall-clock budget of one run in seconds (0, the
 * default, = none): a round-based style checks it between its batches of
 * rounds and stops with IGX_ERR_TIMEOUT once past it, leaving partial output.
 * For autotuners that time every style and need not wait for the slow ones;
 * single-launch styles (triangle counting) do not check it. */
int igx_set_time_limit(double seconds);

/* ------------------------------------- reference-signature entry points ---- */
/* Each takes the same arguments as the reference function named in its
 * comment, returns the elapsed seconds through `seconds`, and uses the style
 * set with igx_set_entry_style (default: the fastest measured style). */

/* static double GPUbfs_vertex(const int src, const ECLgraph& g, basic_t* const dist)
 * -- include/bfs_vertex_cuda.h:61.  g: device CSR; dist: host int[nodes]. */
int igx_bfs_vertex(int src, const igx_graph* g, int* dist, double* seconds);
/* static double GPUbfs_edge(int src, const ECLgraph& g, basic_t* dist, const int* sp)
 * -- include/bfs_edge_cuda.h:61.  sp: host array, uploaded by the entry. */
int igx_bfs_edge(int src, const igx_graph* g, int* dist, const int* sp, double* seconds);
/* static double GPUsssp_vertex(int src, const ECLgraph& g, basic_t* dist)
 * -- include/sssp_vertex_cuda.h:60.  g must carry device eweight. */
int igx_sssp_vertex(int src, const igx_graph* g, int* dist, double* seconds);
/* static double GPUsssp_edge(int src, const ECLgraph& g, basic_t* dist, const int* sp)
 * -- include/sssp_edge_cuda.h:60.  sp: host array. */
int igx_sssp_edge(int src, const igx_graph* g, int* dist, const int* sp, double* seconds);
/* static double GPUcc_vertex(const ECLgraph& g, basic_t* label)
 * -- include/cc_vertex_cuda.h:57. */
int igx_cc_vertex(const igx_graph* g, int* label, double* seconds);
/* static double GPUcc_edge(const ECLgraph& g, basic_t* label, const int* sp)
 * -- include/cc_edge_cuda.h:57.  sp: host array. */
int igx_cc_edge(const igx_graph* g, int* label, const int* sp, double* seconds);
/* static double GPUmis_vertex(const ECLgraph& g, data_type* priority, int* status)
 * -- include/mis_vertex_cuda.h:54.  priority: host scratch, filled with the
 * unsigned hash priorities; status: host int[nodes] (1 in set, 2 not). */
int igx_mis_vertex(const igx_graph* g, unsigned int* priority, int* status, double* seconds);
/* static double GPUmis_edge(const ECLgraph& g, const int* sp, data_type* priority, int* status)
 * -- include/mis_edge_cuda.h:54.  sp: DEVICE array (as in the reference). */
int igx_mis_edge(const igx_graph* g, const int* d_sp, unsigned int* priority, int* status, double* seconds);
/* void PR_GPU(const ECLgraph g, score_type* scores, int* degree)
 * -- include/pr_cuda.h:48.  g: HOST CSR (the entry uploads its own copy and
 * frees it); scores: host float[nodes] in/out; degree: host int[nodes].
 * Also reports the iteration count. */
int igx_pr(const igx_graph* host_g, float* scores, const int* degree, double* seconds, int* iterations);
/* static double GPUtc_vertex(basic_t& count, int nodes, const int* nindex, const int* nlist)
 * -- include/tc_vertex_cuda.h:42.  Device arrays; count widened to uint64. */
int igx_tc_vertex(unsigned long long* count, int nodes, const int* d_nindex, const int* d_nlist, double* seconds);
/* static double GPUtc_edge(basic_t& count, int edges, const int* nindex, const int* nlist, const int* sp)
 * -- include/tc_edge_cuda.h:42.  Device arrays; count widened to uint64.
 * `nodes` is not part of the reference signature; it is recovered on the
 * device from sp (nodes = sp[edges-1] + 1 bounds the rows that own edges). */
int igx_tc_edge(unsigned long long* count, int edges, const int* d_nindex, const int* d_nlist, const int* d_sp,
                double* seconds);

/* ---------------------------------------------------------- batched mode ---- */
/* No reference counterpart (the reference runs one program per graph).  Many
 * graphs packed into one device arena; BFS (source 0), CC and MIS run for all
 * of them in one persistent launch per algorithm, one workgroup per graph,
 * largest graphs first.  Outputs are int32[total_nodes] arrays; graph i's
 * values start at its node offset. */
typedef struct igx_batch {
  int graphs;
  long long total_nodes;
  long long total_edges;
  const long long* node_off;  /* [graphs + 1] */
  const long long* edge_off;  /* [graphs + 1] */
  const int* nindex;          /* graph i's rows start at node_off[i] + i (graph-local offsets) */
  const int* nlist;           /* graph i's list starts at edge_off[i] (graph-local ids) */
  const int* order;           /* processing order (largest first) */
} igx_batch;

int igx_batch_upload(int graphs, const igx_graph* host_graphs, igx_batch* dev);
int igx_batch_free(igx_batch* dev);
/* device igx_graph view of graph i inside the arena (for per-graph runs) */
int igx_batch_graph_view(const igx_batch* b, int i, igx_graph* view);
/* algo_mask: 1 BFS, 2 CC, 4 MIS.  block: 256, 512 or 1024 threads per
 * workgroup.  use_graph != 0 captures the launches once as a hipGraph and
 * replays it (re-captured when the batch, mask, block or outputs change). */
int igx_batch_run(const igx_batch* b, int algo_mask, int* d_bfs, int* d_cc, int* d_mis, int block, int use_graph,
                  igx_run_info* info);
/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same as igx_batch_run with use_graph = 0, but the counters and
 * per-vertex scratch come from a fixed device array of the library (about
 * 16 MiB per GPU, held until the process ends): batches of at most 2^20
 * vertices in all (IGX_ERR_LIMIT beyond).  Allocates nothing; igx_init must
 * have been called (IGX_ERR_ARG otherwise).  Calls on one GPU share that
 * array, so calls that overlap on one GPU race. */
int igx_batch_run_noalloc(const igx_batch* b, int algo_mask, int* d_bfs, int* d_cc, int* d_mis, int block,
                          igx_run_info* info);

/* ------------------------------------------------------ selector features --- */
/* No reference counterpart.  The style selector's IGX_NUM_FEATURES graph
 * features (igx_graph_features in igx_graph.h, same order and values up to
 * rounding of two sums) computed on the device from a device-resident CSR;
 * *seconds (may be NULL) is the device time of the whole computation. */
int igx_graph_features_device(const igx_graph* dev, uint64_t seed, double* features, double* seconds);

/* ------------------------------------------------------ selector network ---- */
/* No reference counterpart.  MLP in -> 64 -> 64 -> out (ReLU, FP32) trained
 * and evaluated with hand-written HIP kernels.  Parameters are laid out as
 * W1[in][64], b1[64], W2[64][64], b2[64], W3[64][out], b3[out]. */
int igx_mlp_create(int in, int out, uint64_t seed, void** handle);
int igx_mlp_destroy(void* handle);
int igx_mlp_param_count(void* handle);
int igx_mlp_get_params(void* handle, float* host, size_t count);
int igx_mlp_set_params(void* handle, const float* host, size_t count);
/* host x[batch][in] -> host y[batch][out] */
int igx_mlp_forward(void* handle, const float* host_x, int batch, float* host_y);
/* gradient of sum(mask * (y - t)^2) / norm with respect to the parameters */
int igx_mlp_gradients(void* handle, const float* x, const float* t, const float* mask, int batch, float norm,
                      float* host_grads, float* loss);
/* Adam over `epochs` passes of seeded-shuffled minibatches; *last_loss is the
 * last epoch's mean masked loss */
int igx_mlp_train(void* handle, const float* x, const float* t, const float* mask, int n, int epochs, int batch,
                  float lr, uint64_t seed, float* last_loss);

#ifdef __cplusplus
}
#endif

#endif /* INDIGO_HIP_H */
