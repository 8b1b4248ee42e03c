// ORACLE -- test infrastructure, not product code.
//
// C++-threads restatement of the reference's codes/cpp path: the second CPU
// baseline column beside the OpenMP restatement in oracle.cpp, and a cross
// check of the MIS tie rule of the C++-threads family.  Every function spawns
// `threads` std::threads per round over a block partition of the vertices
// (thread t owns [t*n/T, (t+1)*n/T)), joins them, and repeats until a round
// changes nothing, as the reference's variants do; the shared state is
// accessed with relaxed atomics (GCC __atomic builtins on the plain arrays).
// Runtimes cover the rounds only (the reference's gettimeofday brackets).
//
// Reference files (relative to the reference tree):
//   [P-BFS]  codes/cpp/bfs-cpp/BFS_CPP_V_Topo_Push_IntType_Atomic_ReadModifyWrite_NoBoundsBug_NoLivelockBug_
//            NoOverflowBug_NonDeterm_NoPrecedenceBug_NoFieldBug_Block_NoNbrBoundsBug.cpp
//   [P-SSSP] codes/cpp/sssp-cpp/SSSP_CPP_V_Topo_Pull_IntType_Atomic_NoBoundsBug_NoPrecedenceBug_NonDeterm_
//            NoFieldBug_Block_NoNborBoundsBug_NoRangeBug.cpp
//   [P-CC]   codes/cpp/cc-cpp/CC_CPP_V_Topo_Push_IntType_Atomic_ReadModifyWrite_NoPrecedenceBug_NoLivelockBug_
//            NoBoundsBug_NonDeterm_NoFieldBug_Block_NoNbrBoundsBug.cpp
//   [P-MIS]  codes/cpp/mis-cpp/MIS_CPP_V_Topo_Push_IntType_Atomic_Block_NonDeterm_NoPrecedenceBug_NoFieldBug_
//            NoBoundsBug_NoNborBoundsBug_NoLivelockBug_NoGuardBug_NoShadowBug.cpp
//   [P-PR]   codes/cpp/pr-cpp/PR_CPP_Pull_FloatType_Reduction_Block_NoBoundsBug_NoPrecedenceBug_NoFieldBug_
//            Determin.cpp
//   [P-TC]   codes/cpp/tc-cpp/TC_CPP_V_IntType_Reduction_Block_NoRaceBug_NoBoundsBug_NoPrecedenceBug_
//            NoNbrBoundsBug.cpp
#include <chrono>
#include <climits>
#include <cmath>
#include <thread>
#include <vector>

#include "igx_graph.h"

namespace {

const int kInf = INT_MAX;

inline int ld(const int* p) { return __atomic_load_n(p, __ATOMIC_RELAXED); }
inline void st(int* p, int v) { __atomic_store_n(p, v, __ATOMIC_RELAXED); }
inline unsigned char ldb(const unsigned char* p) { return __atomic_load_n(p, __ATOMIC_RELAXED); }
inline void stb(unsigned char* p, unsigned char v) { __atomic_store_n(p, v, __ATOMIC_RELAXED); }

// atomic minimum; returns the value before (the reference's atomicMin,
// include/bfs_vertex_cpp.h, a compare-exchange loop)
inline int amin(int* p, int v)
{
  int old = ld(p);
  while (old > v && !__atomic_compare_exchange_n(p, &old, v, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
  }
  return old;
}

inline int part_begin(int t, int T, int n) { return (int)((long long)t * n / T); }

// Run body(t, lo, hi) on T threads over the block partition of [0, n).
template <typename F>
void parallel_blocks(int n, int T, F body)
{
  std::vector<std::thread> th;
  th.reserve(T);
  for (int t = 0; t < T; t++) th.emplace_back(body, t, part_begin(t, T, n), part_begin(t + 1, T, n));
  for (auto& x : th) x.join();
}

double seconds_since(std::chrono::steady_clock::time_point t0)
{
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// hash of mis_vertex_cpp.h (same function as mis_vertex_omp.h:127-132)
inline unsigned mis_hash(unsigned val)
{
  val = ((val >> 16) ^ val) * 0x45d9f3bu;
  val = ((val >> 16) ^ val) * 0x45d9f3bu;
  return (val >> 16) ^ val;
}

}  // namespace

extern "C" {

// [P-BFS]:52-101 -- topology-driven push, one pass over every reached vertex
// per round, atomicMin(dist[dst], dist[v] + 1).
double orc_cpp_bfs_topo_push(int src, const igx_graph* g, int* dist, int threads, int* iters)
{
  const int n = g->nodes;
  for (int v = 0; v < n; v++) dist[v] = (v == src) ? 0 : kInf;
  const auto t0 = std::chrono::steady_clock::now();
  int iter = 0;
  bool again;
  do {
    int flag = 0;
    parallel_blocks(n, threads, [&](int, int lo, int hi) {
      for (int v = lo; v < hi; v++) {
        const int s = ld(&dist[v]);
        if (s == kInf) continue;
        bool upd = false;
        for (int i = g->nindex[v]; i < g->nindex[v + 1]; i++) {
          if (amin(&dist[g->nlist[i]], s + 1) > s + 1) upd = true;
        }
        if (upd) st(&flag, 1);
      }
    });
    iter++;
    again = flag != 0;
  } while (again);
  if (iters) *iters = iter;
  return seconds_since(t0);
}

// [P-SSSP]:52-103 -- topology-driven pull: each vertex takes the minimum of
// dist[u] + w over its neighbours, starting from its own distance, and stores
// the running minimum after every neighbour (as the reference does).
double orc_cpp_sssp_topo_pull(int src, const igx_graph* g, int* dist, int threads, int* iters)
{
  const int n = g->nodes;
  for (int v = 0; v < n; v++) dist[v] = (v == src) ? 0 : kInf;
  const auto t0 = std::chrono::steady_clock::now();
  int iter = 0;
  bool again;
  do {
    iter++;
    int flag = 0;
    parallel_blocks(n, threads, [&](int, int lo, int hi) {
      for (int v = lo; v < hi; v++) {
        int d = ld(&dist[v]);
        for (int i = g->nindex[v]; i < g->nindex[v + 1]; i++) {
          const int s = ld(&dist[g->nlist[i]]);
          if (s != kInf && s + g->eweight[i] < d) {
            d = s + g->eweight[i];
            st(&flag, 1);
          }
          st(&dist[v], d);
        }
      }
    });
    again = flag != 0;
  } while (again);
  if (iters) *iters = iter;
  return seconds_since(t0);
}

// [P-CC]:52-104 -- topology-driven push of the own label to every neighbour
// with atomicMin; labels start as the vertex ids.
double orc_cpp_cc_topo_push(const igx_graph* g, int* label, int threads, int* iters)
{
  const int n = g->nodes;
  for (int v = 0; v < n; v++) label[v] = v;
  const auto t0 = std::chrono::steady_clock::now();
  int iter = 0;
  bool again;
  do {
    iter++;
    int flag = 0;
    parallel_blocks(n, threads, [&](int, int lo, int hi) {
      for (int v = lo; v < hi; v++) {
        const int lv = ld(&label[v]);
        bool upd = false;
        for (int i = g->nindex[v]; i < g->nindex[v + 1]; i++) {
          if (amin(&label[g->nlist[i]], lv) > lv) upd = true;
        }
        if (upd) st(&flag, 1);
      }
    });
    again = flag != 0;
  } while (again);
  if (iters) *iters = iter;
  return seconds_since(t0);
}

// [P-MIS]:44-117 -- priorities hash(v + 712313887) compared unsigned; an
// undecided vertex waits while some non-excluded neighbour outranks it
// (higher priority, or equal priority and a higher id: the tie rule at :63),
// otherwise it joins and excludes all its neighbours.  Status: 0 undecided,
// 1 included, 2 excluded.
double orc_cpp_mis_topo_push(const igx_graph* g, unsigned* priority, unsigned char* status, int threads,
                             int* iters)
{
  const int n = g->nodes;
  for (int v = 0; v < n; v++) {
    priority[v] = mis_hash((unsigned)v + 712313887u);
    status[v] = 0;
  }
  const auto t0 = std::chrono::steady_clock::now();
  int iter = 0;
  bool again;
  do {
    iter++;
    int flag = 0;
    parallel_blocks(n, threads, [&](int, int lo, int hi) {
      for (int v = lo; v < hi; v++) {
        if (ldb(&status[v]) != 0) continue;
        bool blocked = false;
        for (int i = g->nindex[v]; i < g->nindex[v + 1]; i++) {
          const int u = g->nlist[i];
          if (ldb(&status[u]) != 2 && (priority[u] > priority[v] || (priority[u] == priority[v] && v < u))) {
            blocked = true;
            break;
          }
        }
        if (blocked) {
          st(&flag, 1);
        } else {
          stb(&status[v], 1);
          for (int i = g->nindex[v]; i < g->nindex[v + 1]; i++) stb(&status[g->nlist[i]], 2);
        }
      }
    });
    again = flag != 0;
  } while (again);
  if (iters) *iters = iter;
  return seconds_since(t0);
}

// [P-PR]:42-93 -- FP32 pull from the contributions of the previous iteration
// (scores / degree, computed serially), the L1 change summed in double per
// thread and then over the threads; stops when it is below eps or after
// max_iter iterations.  *iters follows the reference's report (one more than
// the index of the converging iteration).  scores must hold the start values.
double orc_cpp_pr_pull(const igx_graph* g, float* scores, double eps, int max_iter, int threads, int* iters)
{
  const int n = g->nodes;
  const float kDamp = 0.85f;
  const float base = (1.0f - kDamp) / (float)n;
  std::vector<float> contrib(n);
  std::vector<double> part(threads);
  const auto t0 = std::chrono::steady_clock::now();
  int iter;
  for (iter = 0; iter < max_iter; iter++) {
    for (int v = 0; v < n; v++) contrib[v] = scores[v] / (float)(g->nindex[v + 1] - g->nindex[v]);
    parallel_blocks(n, threads, [&](int t, int lo, int hi) {
      double e = 0;
      for (int v = lo; v < hi; v++) {
        float in = 0;
        for (int i = g->nindex[v]; i < g->nindex[v + 1]; i++) in += contrib[g->nlist[i]];
        const float old = scores[v];
        scores[v] = base + kDamp * in;
        e += std::fabs(scores[v] - old);
      }
      part[t] = e;
    });
    double err = 0;
    for (int t = 0; t < threads; t++) err += part[t];
    if (err < eps) break;
  }
  const double secs = seconds_since(t0);
  if (iter < max_iter) iter++;
  if (iters) *iters = iter;
  return secs;
}

// [P-TC]:41-84 -- per-thread sums of the sorted-list intersections of every
// v < u edge (tc_vertex_cpp.h common()), added after the joins; uint64 here.
double orc_cpp_tc_vertex(const igx_graph* g, unsigned long long* count, int threads)
{
  const int n = g->nodes;
  const int* ni = g->nindex;
  const int* nl = g->nlist;
  std::vector<unsigned long long> part(threads);
  const auto t0 = std::chrono::steady_clock::now();
  parallel_blocks(n, threads, [&](int t, int lo, int hi) {
    unsigned long long c = 0;
    for (int v = lo; v < hi; v++) {
      const int b1 = ni[v], e1 = ni[v + 1];
      int s1 = e1;
      while (b1 < s1 && v < nl[s1 - 1]) s1--;
      for (int j = s1; j < e1; j++) {
        const int u = nl[j];
        const int b2 = ni[u], e2 = ni[u + 1];
        int s2 = e2;
        while (b2 < s2 && u < nl[s2 - 1]) s2--;
        int p = j + 1, q = s2;
        while (p < e1 && q < e2) {
          if (nl[p] < nl[q]) p++;
          else if (nl[q] < nl[p]) q++;
          else {
            c++;
            p++;
            q++;
          }
        }
      }
    }
    part[t] = c;
  });
  unsigned long long total = 0;
  for (int t = 0; t < threads; t++) total += part[t];
  *count = total;
  return seconds_since(t0);
}

}  // extern "C"
