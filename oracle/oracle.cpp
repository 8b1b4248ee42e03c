// ORACLE -- test infrastructure, not product code.
//
// A C++/OpenMP restatement of the reference's CPU path for the six Indigo2
// kernels (codes/omp/* plus the serial verifiers in include/*_omp.h).  Only
// tests/, the smoke check and bench.py's CPU-baseline leg load this library,
// and only as the checker or as the timed CPU baseline; the product path
// (libindigo_hip.so) never calls it.
//
// Every function cites the reference lines it restates.  "[O-BFS]" etc. name
// the reference files:
//   [O-BFS]     codes/omp/bfs-omp/BFS_OMP_V_Topo_Push_IntType_NoRangeBug_ReadModifyWrite_Default_NoBoundsBug_Atomic_NonDeterm.cpp
//   [O-BFSpull] codes/omp/bfs-omp/BFS_OMP_V_Topo_Pull_IntType_NoRangeBug_Default_NoBoundsBug_Atomic_NonDeterm.cpp
//   [O-SSSP]    codes/omp/sssp-omp/SSSP_OMP_E_Data_Push_NonDeterm_IntType_ReadModifyWrite_Atomic_NonDup_NoFieldBug_Default.cpp
//   [O-SSSPpull] codes/omp/sssp-omp/SSSP_OMP_V_Topo_Pull_IntType_NonDeterm_Default_NoBoundsBug_Atomic_NoRangeBug.cpp
//   [O-CC]      codes/omp/cc-omp/CC_OMP_V_Topo_Push_IntType_ReadModifyWrite_NonDeterm_Default_NoBoundsBug_Atomic.cpp
//   [O-CCpull]  codes/omp/cc-omp/CC_OMP_V_Topo_Pull_IntType_NonDeterm_Default_NoBoundsBug_Atomic.cpp
//   [O-MIS]     codes/omp/mis-omp/MIS_OMP_V_Topo_Push_IntType_Atomic_NonDeterm_Default_NoBoundsBug.cpp
//   [O-MISe]    codes/omp/mis-omp/MIS_OMP_E_Topo_IntType_Atomic_NonDeterm_Default_NoBoundsBug.cpp
//   [O-PR]      codes/omp/pr-omp/PR_OMP_Pull_FloatType_Determin_DefaultSched_ClauseRed.cpp
//   [O-TC]      codes/omp/tc-omp/TC_OMP_V_IntType_ClauseRed_Default.cpp
//   [O-TCe]     codes/omp/tc-omp/TC_OMP_E_IntType_ClauseRed_Default.cpp
//
// Deliberate deviations from the reference, each also exercised by a test:
//  * [O-SSSP] zeroes all max(m, n) edge stamps; the reference zeroes only the
//    first n and reads stale heap memory for the rest.
//  * The triangle count is accumulated in uint64 (the reference uses int).
//  * The CC verifier walks components with an explicit stack instead of the
//    recursion of cc_vertex_omp.h, so large components cannot overflow the
//    call stack; it checks the same three properties.
//  * Time is measured with omp_get_wtime() instead of gettimeofday().
#include "igx_graph.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <queue>
#include <vector>
#include <omp.h>

namespace {

template <typename T>
inline T atomic_read(T* addr)
{
  T ret;
  #pragma omp atomic read
  ret = *addr;
  return ret;
}

template <typename T>
inline void atomic_write(T* addr, T val)
{
  #pragma omp atomic write
  *addr = val;
}

// bfs_vertex_omp.h:63-89 -- min/max through one global critical section.
template <typename T>
inline T critical_min(T* addr, T val)
{
  T oldv;
  #pragma omp critical
  {
    oldv = *addr;
    if (oldv > val) *addr = val;
  }
  return oldv;
}

template <typename T>
inline T critical_max(T* addr, T val)
{
  T oldv;
  #pragma omp critical
  {
    oldv = *addr;
    if (oldv < val) *addr = val;
  }
  return oldv;
}

// bfs_vertex_omp.h:91-101
template <typename T>
inline T fetch_and_add(T* addr)
{
  T old;
  #pragma omp atomic capture
  {
    old = *addr;
    (*addr)++;
  }
  return old;
}

// mis_vertex_omp.h:127-132 (integer hash used for MIS priorities)
inline unsigned int mis_hash(unsigned int val)
{
  val = ((val >> 16) ^ val) * 0x45d9f3b;
  val = ((val >> 16) ^ val) * 0x45d9f3b;
  return (val >> 16) ^ val;
}

const unsigned char undecided = 0;  // mis_vertex_omp.h:48-50
const unsigned char included = 1;
const unsigned char excluded = 2;

// tc_vertex_omp.h:44-63 -- merge count of two sorted ranges
inline int common(int beg1, int end1, int beg2, int end2, const int* nlist)
{
  int cnt = 0;
  int pos1 = beg1, pos2 = beg2;
  while ((pos1 < end1) && (pos2 < end2)) {
    while ((pos1 < end1) && (nlist[pos1] < nlist[pos2])) pos1++;
    if (pos1 < end1) {
      while ((pos2 < end2) && (nlist[pos2] < nlist[pos1])) pos2++;
      if ((pos2 < end2) && (nlist[pos1] == nlist[pos2])) {
        pos1++;
        pos2++;
        cnt++;
      } else {
        pos1++;
      }
    }
  }
  return cnt;
}

// tc_edge_omp.h:104-115 -- binary search over the inclusive range [beg, end]
inline bool find_inclusive(int target, int beg, int end, const int* nlist)
{
  int left = beg, right = end;
  while (left <= right) {
    const int middle = (left + right) / 2;
    if (nlist[middle] == target) return true;
    if (nlist[middle] < target) left = middle + 1;
    else right = middle - 1;
  }
  return false;
}

}  // namespace

extern "C" {

// ================================================================== BFS ====
// bfs_vertex_omp.h:105-130 -- serial Dijkstra with unit weights (verifier).
void orc_bfs_serial(int src, const igx_graph* g, int* dist)
{
  for (int i = 0; i < g->nodes; i++) dist[i] = INT_MAX;
  dist[src] = 0;
  std::priority_queue<std::pair<int, int>> pq;  // max-heap on -distance
  pq.push(std::make_pair(0, src));
  while (!pq.empty()) {
    const int v = pq.top().second;
    pq.pop();
    const int dv = dist[v];
    for (int i = g->nindex[v]; i < g->nindex[v + 1]; i++) {
      const int n = g->nlist[i];
      const int d = dv + 1;
      if (d < dist[n]) {
        dist[n] = d;
        pq.push(std::make_pair(-d, n));
      }
    }
  }
}

// [O-BFS]:41-95 -- vertex-based, topology-driven push with critical_min.
double orc_bfs_omp_topo_push(int src, const igx_graph* g, int* dist, int* iters)
{
  const int n = g->nodes;
  for (int v = 0; v < n; v++) dist[v] = (v == src) ? 0 : INT_MAX;  // :41-48
  const double t0 = omp_get_wtime();
  int goagain;
  int iter = 0;
  do {
    goagain = 0;
    #pragma omp parallel for
    for (int v = 0; v < n; v++) {  // :50-72
      const int beg = g->nindex[v], end = g->nindex[v + 1];
      const int s = atomic_read(&dist[v]);
      if (s != INT_MAX) {
        bool updated = false;
        for (int i = beg; i < end; i++) {
          const int dst = g->nlist[i];
          const int new_dist = s + 1;
          if (critical_min(&dist[dst], new_dist) > new_dist) updated = true;
        }
        if (updated) atomic_write(&goagain, 1);
      }
    }
    iter++;
  } while (goagain);
  const double t = omp_get_wtime() - t0;
  if (iters) *iters = iter;
  return t;
}

// [O-BFSpull]:41-93 -- vertex-based, topology-driven pull (no critical).
double orc_bfs_omp_topo_pull(int src, const igx_graph* g, int* dist, int* iters)
{
  const int n = g->nodes;
  for (int v = 0; v < n; v++) dist[v] = (v == src) ? 0 : INT_MAX;
  const double t0 = omp_get_wtime();
  int goagain;
  int iter = 0;
  do {
    goagain = 0;
    #pragma omp parallel for
    for (int v = 0; v < n; v++) {  // :50-70
      const int beg = g->nindex[v], end = g->nindex[v + 1];
      int d = atomic_read(&dist[v]);
      for (int i = beg; i < end; i++) {
        const int s = atomic_read(&dist[g->nlist[i]]);
        if (s != INT_MAX) {
          const int new_dist = s + 1;
          if (new_dist < d) {
            d = new_dist;
            atomic_write(&goagain, 1);
          }
        }
        atomic_write(&dist[v], d);
      }
    }
    iter++;
  } while (goagain);
  const double t = omp_get_wtime() - t0;
  if (iters) *iters = iter;
  return t;
}

// ================================================================= SSSP ====
// sssp_edge_omp.h (CPUserialDijkstra with a min-heap and eweight) -- verifier.
void orc_sssp_serial(int src, const igx_graph* g, int* dist)
{
  for (int i = 0; i < g->nodes; i++) dist[i] = INT_MAX;
  dist[src] = 0;
  typedef std::pair<int, int> pair;
  std::priority_queue<pair, std::vector<pair>, std::greater<pair>> pq;
  pq.push(std::make_pair(0, src));
  while (!pq.empty()) {
    const int v = pq.top().second;
    pq.pop();
    const int dv = dist[v];
    for (int i = g->nindex[v]; i < g->nindex[v + 1]; i++) {
      const int n = g->nlist[i];
      const int d = dv + g->eweight[i];
      if (d < dist[n]) {
        dist[n] = d;
        pq.push(std::make_pair(d, n));
      }
    }
  }
}

// [O-SSSP]:41-113 -- edge-based, data-driven push with NonDup stamps.
// Deviation: every one of the max(m, n) stamps starts at zero.
double orc_sssp_omp_edge_data(int src, const igx_graph* g, const int* sp, int* dist, int* iters)
{
  const int n = g->nodes;
  const int size = std::max(g->edges, n);
  std::vector<int> time(size, 0), wl1(size), wl2(size);
  int wl1size = 0, wl2size = 0;
  for (int v = 0; v < n; v++) dist[v] = (v == src) ? 0 : INT_MAX;  // :41-56
  for (int i = g->nindex[src]; i < g->nindex[src + 1]; i++) wl1[wl1size++] = i;
  const double t0 = omp_get_wtime();
  int iter = 0;
  int* w1 = wl1.data();
  int* w2 = wl2.data();
  int* tm = time.data();
  do {
    iter++;
    wl2size = 0;
    #pragma omp parallel for
    for (int idx = 0; idx < wl1size; idx++) {  // :57-76
      const int e = w1[idx];
      const int s_v = sp[e];
      const int dst = g->nlist[e];
      const int s = atomic_read(&dist[s_v]);
      if (s != INT_MAX) {
        const int new_dist = s + g->eweight[e];
        if (critical_min(&dist[dst], new_dist) > new_dist) {
          for (int j = g->nindex[dst]; j < g->nindex[dst + 1]; j++) {
            if (critical_max(&tm[j], iter) != iter) {
              w2[fetch_and_add(&wl2size)] = j;
            }
          }
        }
      }
    }
    std::swap(w1, w2);
    wl1size = wl2size;
  } while (wl1size > 0);
  const double t = omp_get_wtime() - t0;
  if (iters) *iters = iter;
  return t;
}

// [O-SSSPpull] codes/omp/sssp-omp/SSSP_OMP_V_Topo_Pull_IntType_NonDeterm_Default_NoBoundsBug_Atomic_NoRangeBug.cpp
// :41-90 -- vertex-based, topology-driven pull (the fastest OpenMP SSSP form).
double orc_sssp_omp_topo_pull(int src, const igx_graph* g, int* dist, int* iters)
{
  const int n = g->nodes;
  for (int v = 0; v < n; v++) dist[v] = (v == src) ? 0 : INT_MAX;
  const double t0 = omp_get_wtime();
  int goagain;
  int iter = 0;
  do {
    iter++;
    goagain = 0;
    #pragma omp parallel for
    for (int v = 0; v < n; v++) {  // :51-71
      const int beg = g->nindex[v], end = g->nindex[v + 1];
      int d = atomic_read(&dist[v]);
      for (int i = beg; i < end; i++) {
        const int s = atomic_read(&dist[g->nlist[i]]);
        if (s != INT_MAX) {
          const int new_dist = s + g->eweight[i];
          if (new_dist < d) {
            d = new_dist;
            atomic_write(&goagain, 1);
          }
        }
        atomic_write(&dist[v], d);
      }
    }
  } while (goagain);
  const double t = omp_get_wtime() - t0;
  if (iters) *iters = iter;
  return t;
}

// ================================================================== CC =====
// [O-CC]:41-93 -- vertex-based, topology-driven push of min labels.
double orc_cc_omp_topo_push(const igx_graph* g, int* label, int* iters)
{
  const int n = g->nodes;
  for (int v = 0; v < n; v++) label[v] = v;
  const double t0 = omp_get_wtime();
  int goagain;
  int iter = 0;
  do {
    iter++;
    goagain = 0;
    #pragma omp parallel for
    for (int v = 0; v < n; v++) {  // :49-69
      const int beg = g->nindex[v], end = g->nindex[v + 1];
      const int new_label = atomic_read(&label[v]);
      bool updated = false;
      for (int i = beg; i < end; i++) {
        if (critical_min(&label[g->nlist[i]], new_label) > new_label) updated = true;
      }
      if (updated) atomic_write(&goagain, 1);
    }
  } while (goagain);
  const double t = omp_get_wtime() - t0;
  if (iters) *iters = iter;
  return t;
}

// [O-CCpull]:41-93 -- vertex-based, topology-driven pull of min labels.
double orc_cc_omp_topo_pull(const igx_graph* g, int* label, int* iters)
{
  const int n = g->nodes;
  for (int v = 0; v < n; v++) label[v] = v;
  const double t0 = omp_get_wtime();
  int goagain;
  int iter = 0;
  do {
    iter++;
    goagain = 0;
    #pragma omp parallel for
    for (int v = 0; v < n; v++) {
      const int beg = g->nindex[v], end = g->nindex[v + 1];
      int d = atomic_read(&label[v]);
      for (int i = beg; i < end; i++) {
        const int nl = atomic_read(&label[g->nlist[i]]);
        if (nl < d) {
          d = nl;
          atomic_write(&goagain, 1);
        }
        atomic_write(&label[v], d);
      }
    }
  } while (goagain);
  const double t = omp_get_wtime() - t0;
  if (iters) *iters = iter;
  return t;
}

// cc_vertex_cuda.h:142-167 (same checks in cc_vertex_omp.h): adjacent labels
// equal, no sentinel label >= n, each label names exactly one component.
// Returns 0 when all pass, else 1 (adjacent differ), 2 (sentinel),
// 3 (incorrect id inside a component), 4 (count mismatch).
int orc_cc_verify(const igx_graph* g, const int* label_in)
{
  const int n = g->nodes;
  for (int v = 0; v < n; v++) {
    for (int i = g->nindex[v]; i < g->nindex[v + 1]; i++) {
      if (label_in[g->nlist[i]] != label_in[v]) return 1;
    }
  }
  for (int v = 0; v < n; v++) if (label_in[v] >= n || label_in[v] < 0) return 2;
  std::vector<int> label(label_in, label_in + n);
  std::vector<char> seen(n, 0);
  int count = 0, distinct = 0;
  for (int v = 0; v < n; v++) {
    if (!seen[label_in[v]]) {
      seen[label_in[v]] = 1;
      distinct++;
    }
  }
  std::vector<int> stack;
  for (int v = 0; v < n; v++) {
    if (label[v] < n) {
      count++;
      const int id = label[v];
      stack.push_back(v);
      while (!stack.empty()) {  // cc_vertex_cuda.h:71-80 without recursion
        const int x = stack.back();
        stack.pop_back();
        if (label[x] < n) {
          if (label[x] != id) return 3;
          label[x] = n;
          for (int i = g->nindex[x]; i < g->nindex[x + 1]; i++) stack.push_back(g->nlist[i]);
        }
      }
    }
  }
  return (count == distinct) ? 0 : 4;
}

// ================================================================= MIS =====
unsigned int orc_mis_hash(unsigned int v) { return mis_hash(v); }

// [O-MIS]:41-105 -- vertex-based, topology-driven push; unsigned priorities.
double orc_mis_omp_topo_push(const igx_graph* g, unsigned int* priority, unsigned char* status, int* iters)
{
  const int n = g->nodes;
  for (int v = 0; v < n; v++) {  // :41-49
    priority[v] = mis_hash(v + 712313887);
    status[v] = undecided;
  }
  const double t0 = omp_get_wtime();
  bool goagain;
  int iter = 0;
  do {
    iter++;
    goagain = false;
    #pragma omp parallel for
    for (int v = 0; v < n; v++) {  // :51-74
      if (atomic_read(&status[v]) == undecided) {
        int i = g->nindex[v];
        while ((i < g->nindex[v + 1]) &&
               ((atomic_read(&status[g->nlist[i]]) == excluded) || (priority[v] > priority[g->nlist[i]]) ||
                ((priority[v] == priority[g->nlist[i]]) && (v > g->nlist[i])))) {
          i++;
        }
        if (i < g->nindex[v + 1]) {
          atomic_write(&goagain, true);
        } else {
          atomic_write(&status[v], included);
          for (int j = g->nindex[v]; j < g->nindex[v + 1]; j++) atomic_write(&status[g->nlist[j]], excluded);
        }
      }
    }
  } while (goagain);
  const double t = omp_get_wtime() - t0;
  if (iters) *iters = iter;
  return t;
}

// [O-MISe]:41-138 -- edge-based, topology-driven (edge pass + vertex pass).
double orc_mis_omp_edge(const igx_graph* g, const int* sp, unsigned char* status, int* iters)
{
  const int n = g->nodes;
  std::vector<unsigned int> priority(n);
  std::vector<char> lost(n, 0);
  for (int v = 0; v < n; v++) {
    priority[v] = mis_hash(v + 712313887);
    status[v] = undecided;
  }
  char* lp = lost.data();
  const unsigned int* pr = priority.data();
  const double t0 = omp_get_wtime();
  bool goagain;
  int iter = 0;
  do {
    iter++;
    goagain = false;
    #pragma omp parallel for
    for (int e = 0; e < g->edges; e++) {  // :52-77
      const int src = sp[e];
      const int dst = g->nlist[e];
      if (atomic_read(&status[src]) == included) {
        atomic_write(&status[dst], excluded);
      } else if (atomic_read(&status[dst]) == included) {
        atomic_write(&status[src], excluded);
      } else if (atomic_read(&status[src]) != excluded && atomic_read(&status[dst]) != excluded) {
        if (pr[src] < pr[dst]) atomic_write(&lp[src], (char)1);
        else atomic_write(&lp[dst], (char)1);
      }
    }
    #pragma omp parallel for
    for (int v = 0; v < n; v++) {  // :79-94
      if (lp[v] == 0) {
        if (status[v] == undecided) status[v] = included;
      } else {
        atomic_write(&goagain, true);
        lp[v] = 0;
      }
    }
  } while (goagain);
  #pragma omp parallel for
  for (int i = 0; i < n; i++) {
    if (status[i] == undecided) status[i] = included;
  }
  const double t = omp_get_wtime() - t0;
  if (iters) *iters = iter;
  return t;
}

// Sequential greedy MIS in descending (priority, id) order -- the fixed point
// every MIS style converges to; used to pin the OMP restatement.
void orc_mis_serial_greedy(const igx_graph* g, unsigned char* status)
{
  const int n = g->nodes;
  std::vector<int> order(n);
  for (int v = 0; v < n; v++) order[v] = v;
  std::sort(order.begin(), order.end(), [](int a, int b) {
    const unsigned int pa = mis_hash(a + 712313887), pb = mis_hash(b + 712313887);
    return (pa != pb) ? (pa > pb) : (a > b);
  });
  for (int v = 0; v < n; v++) status[v] = undecided;
  for (int k = 0; k < n; k++) {
    const int v = order[k];
    if (status[v] != undecided) continue;
    status[v] = included;
    for (int i = g->nindex[v]; i < g->nindex[v + 1]; i++) {
      if (status[g->nlist[i]] == undecided) status[g->nlist[i]] = excluded;
    }
  }
}

// mis_vertex_omp.h:54-73 -- 0 ok, 1 undecided, 2 adjacent in set, 3 not maximal.
int orc_mis_verify(const igx_graph* g, const unsigned char* status)
{
  for (int v = 0; v < g->nodes; v++) {
    if ((status[v] != included) && (status[v] != excluded)) return 1;
    if (status[v] == included) {
      for (int i = g->nindex[v]; i < g->nindex[v + 1]; i++) {
        if (status[g->nlist[i]] == included) return 2;
      }
    } else {
      bool flag = true;
      for (int i = g->nindex[v]; i < g->nindex[v + 1]; i++) {
        if (status[g->nlist[i]] == included) {flag = false; break;}
      }
      if (flag) return 3;
    }
  }
  return 0;
}

// ================================================================== PR =====
// [O-PR]:41-71 -- FP32 pull with Jacobi contributions; error in double;
// eps and max_iter are parameters (pr_omp.h:43-46 fixes 1e-4 and 100).
// Returns seconds; *iters gets the reference's reported iteration count.
double orc_pr_omp_pull(const igx_graph* g, float* scores, double eps, int max_iter, int* iters)
{
  const int n = g->nodes;
  const float kDamp = 0.85f;
  const float base_score = (1.0f - kDamp) / (float)n;
  std::vector<float> contrib(n);
  std::vector<int> degree(n);
  for (int i = 0; i < n; i++) degree[i] = g->nindex[i + 1] - g->nindex[i];
  const float init_score = 1.0f / (float)n;  // pr_omp.h:79-81
  for (int i = 0; i < n; i++) scores[i] = init_score;
  const double t0 = omp_get_wtime();
  int iter;
  for (iter = 0; iter < max_iter; iter++) {
    double error = 0;
    for (int i = 0; i < n; i++) contrib[i] = scores[i] / degree[i];
    #pragma omp parallel for reduction(+: error)
    for (int i = 0; i < n; i++) {
      float incoming_total = 0;
      for (int j = g->nindex[i]; j < g->nindex[i + 1]; j++) incoming_total += contrib[g->nlist[j]];
      const float old_score = scores[i];
      scores[i] = base_score + kDamp * incoming_total;
      error += std::fabs(scores[i] - old_score);
    }
    if (error < eps) break;
  }
  const double t = omp_get_wtime() - t0;
  if (iter < max_iter) iter++;  // :65
  if (iters) *iters = iter;
  return t;
}

// FP64 oracle of the same iteration: s0 = 1/n,
// s_v <- (1-d)/n + d * sum_{u in N(v)} s_u / deg(u), no dangling
// redistribution.  Stops after the first iteration whose L1 change is below
// eps, or after max_iter iterations; with fixed_iters > 0 it runs exactly that
// many.  Returns the number of iterations performed.
int orc_pr_fp64(const igx_graph* g, double* scores, double eps, int max_iter, int fixed_iters, double* last_residual)
{
  const int n = g->nodes;
  const double kDamp = 0.85;
  const double base = (1.0 - kDamp) / n;
  std::vector<double> contrib(n);
  for (int i = 0; i < n; i++) scores[i] = 1.0 / n;
  const int limit = (fixed_iters > 0) ? fixed_iters : max_iter;
  int iter = 0;
  double err = 0;
  while (iter < limit) {
    for (int i = 0; i < n; i++) {
      const int d = g->nindex[i + 1] - g->nindex[i];
      contrib[i] = (d > 0) ? scores[i] / d : 0.0;
    }
    err = 0;
    #pragma omp parallel for reduction(+: err)
    for (int i = 0; i < n; i++) {
      double acc = 0;
      for (int j = g->nindex[i]; j < g->nindex[i + 1]; j++) acc += contrib[g->nlist[j]];
      const double nv = base + kDamp * acc;
      err += std::fabs(nv - scores[i]);
      scores[i] = nv;
    }
    iter++;
    if (fixed_iters <= 0 && err < eps) break;
  }
  if (last_residual) *last_residual = err;
  return iter;
}

// pr_omp.h:85-104 -- one extra push iteration; returns the L1 change.
double orc_pr_check(const igx_graph* g, const float* scores)
{
  const int n = g->nodes;
  const float kDamp = 0.85f;
  const float base_score = (1.0f - kDamp) / (float)n;
  std::vector<float> incoming(n, 0.0f);
  for (int src = 0; src < n; src++) {
    const int deg = g->nindex[src + 1] - g->nindex[src];
    const float outgoing = scores[src] / deg;
    for (int i = g->nindex[src]; i < g->nindex[src + 1]; i++) incoming[g->nlist[i]] += outgoing;
  }
  double error = 0;
  for (int i = 0; i < n; i++) error += std::fabs(base_score + kDamp * incoming[i] - scores[i]);
  return error;
}

// ================================================================== TC =====
// tc_vertex_omp.h:86-105 -- serial count (verifier), here in uint64.
unsigned long long orc_tc_serial(const igx_graph* g)
{
  unsigned long long count = 0;
  const int* nindex = g->nindex;
  const int* nlist = g->nlist;
  for (int v = 0; v < g->nodes; v++) {
    const int beg1 = nindex[v], end1 = nindex[v + 1];
    int start1 = end1;
    while ((beg1 < start1) && (v < nlist[start1 - 1])) start1--;
    for (int j = start1; j < end1; j++) {
      const int u = nlist[j];
      const int beg2 = nindex[u], end2 = nindex[u + 1];
      int start2 = end2;
      while ((beg2 < start2) && (u < nlist[start2 - 1])) start2--;
      count += (unsigned long long)common(j + 1, end1, start2, end2, nlist);
    }
  }
  return count;
}

// [O-TC]:41-74 -- vertex-based merge count with an OpenMP reduction.
double orc_tc_omp_vertex(const igx_graph* g, unsigned long long* count_out)
{
  const int* nindex = g->nindex;
  const int* nlist = g->nlist;
  unsigned long long g_count = 0;
  const double t0 = omp_get_wtime();
  #pragma omp parallel for reduction(+: g_count)
  for (int v = 0; v < g->nodes; v++) {
    unsigned long long count = 0;
    const int beg1 = nindex[v], end1 = nindex[v + 1];
    int start1 = end1;
    while ((beg1 < start1) && (v < nlist[start1 - 1])) start1--;
    for (int j = start1; j < end1; j++) {
      const int u = nlist[j];
      const int beg2 = nindex[u], end2 = nindex[u + 1];
      int start2 = end2;
      while ((beg2 < start2) && (u < nlist[start2 - 1])) start2--;
      count += (unsigned long long)common(j + 1, end1, start2, end2, nlist);
    }
    g_count += count;
  }
  const double t = omp_get_wtime() - t0;
  *count_out = g_count;
  return t;
}

// Checker for large graphs (no reference counterpart; the reference's counts
// above take minutes on R-MAT 20): forward counting with degree ordering
// (Schank & Wagner 2005).  Each edge is oriented from the lower to the higher
// (degree, id) rank, and every oriented edge (v, u) adds |N+(v) ∩ N+(u)|; each
// triangle is counted once, at its lowest-ranked vertex and middle vertex.
unsigned long long orc_tc_forward(const igx_graph* g)
{
  const int n = g->nodes;
  const int* nindex = g->nindex;
  const int* nlist = g->nlist;
  auto before = [&](int a, int b) {
    const int da = nindex[a + 1] - nindex[a], db = nindex[b + 1] - nindex[b];
    return (da < db) || (da == db && a < b);
  };
  std::vector<int> oidx(n + 1, 0);
  #pragma omp parallel for schedule(dynamic, 1024)
  for (int v = 0; v < n; v++) {
    int c = 0;
    for (int i = nindex[v]; i < nindex[v + 1]; i++) c += (nlist[i] != v && before(v, nlist[i])) ? 1 : 0;
    oidx[v + 1] = c;
  }
  for (int v = 0; v < n; v++) oidx[v + 1] += oidx[v];
  std::vector<int> out(oidx[n] > 0 ? oidx[n] : 1);
  #pragma omp parallel for schedule(dynamic, 1024)
  for (int v = 0; v < n; v++) {
    int k = oidx[v];
    for (int i = nindex[v]; i < nindex[v + 1]; i++) {
      if (nlist[i] != v && before(v, nlist[i])) out[k++] = nlist[i];  // stays sorted by id
    }
  }
  unsigned long long total = 0;
  #pragma omp parallel for schedule(dynamic, 256) reduction(+: total)
  for (int v = 0; v < n; v++) {
    for (int j = oidx[v]; j < oidx[v + 1]; j++) {
      const int u = out[j];
      int a = oidx[v], ae = oidx[v + 1], b = oidx[u], be = oidx[u + 1];
      while (a < ae && b < be) {
        if (out[a] < out[b]) a++;
        else if (out[a] > out[b]) b++;
        else {
          total++;
          a++;
          b++;
        }
      }
    }
  }
  return total;
}

// [O-TCe] -- edge-based: for each edge src > dst, every u < dst in N(dst) is
// looked up in N(src) by binary search (tc_edge_omp.h find; exclusive upper
// bound here, see the header comment of the reference's inclusive form).
double orc_tc_omp_edge(const igx_graph* g, const int* sp, unsigned long long* count_out)
{
  const int* nindex = g->nindex;
  const int* nlist = g->nlist;
  unsigned long long g_count = 0;
  const double t0 = omp_get_wtime();
  #pragma omp parallel for reduction(+: g_count)
  for (int e = 0; e < g->edges; e++) {
    const int src = sp[e];
    const int dst = nlist[e];
    if (src > dst) {
      unsigned long long count = 0;
      for (int i = nindex[dst]; i < nindex[dst + 1] && nlist[i] < dst; i++) {
        if (find_inclusive(nlist[i], nindex[src], nindex[src + 1] - 1, nlist)) count++;
      }
      g_count += count;
    }
  }
  const double t = omp_get_wtime() - t0;
  *count_out = g_count;
  return t;
}

int orc_num_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
