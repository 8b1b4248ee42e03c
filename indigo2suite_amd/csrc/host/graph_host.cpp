// Host graph substrate: .egr I/O, seeded generators, CSR construction and
// graph statistics.  Built with g++ -O3 -fopenmp into libindigo_graph.so.
//
// The .egr layout and the failure cases follow include/ECLgraph.h:52-106 of the
// reference; errors are returned as igx_status codes instead of exit(-1).
#include "igx_graph.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>
#include <omp.h>

namespace {

// ---------------------------------------------------------------- RNG -----
// splitmix64 finalizer; a counter-based stream is mix(key + k * golden).
inline uint64_t mix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Stream {
  uint64_t key;
  uint64_t ctr;
  Stream(uint64_t seed, uint64_t item) : key(mix64(mix64(seed) ^ (item * 0xD1B54A32D192ED03ull))), ctr(0) {}
  uint64_t next() { return mix64(key + (ctr++) * 0x9E3779B97F4A7C15ull); }
  // uniform double in [0, 1) with 53 random bits
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

int* alloc_ints(size_t count)
{
  return (int*)malloc(std::max<size_t>(count, 1) * sizeof(int));
}

// ------------------------------------------------------- CSR building -----
// Build a symmetric, deduplicated, self-loop-free CSR with sorted adjacency
// from an undirected pair list.  Deterministic for a given pair multiset.
int build_csr(int n, int64_t count, const int* src, const int* dst, igx_graph* out)
{
  if (n < 1 || count < 0 || out == nullptr) return IGX_ERR_ARG;
  std::vector<std::atomic<int64_t>> deg(n);
  #pragma omp parallel for schedule(static)
  for (int v = 0; v < n; v++) deg[v].store(0, std::memory_order_relaxed);
  #pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < count; i++) {
    const int u = src[i], v = dst[i];
    if (u != v) {
      deg[u].fetch_add(1, std::memory_order_relaxed);
      deg[v].fetch_add(1, std::memory_order_relaxed);
    }
  }
  std::vector<int64_t> off(n + 1);
  off[0] = 0;
  for (int v = 0; v < n; v++) off[v + 1] = off[v] + deg[v].load(std::memory_order_relaxed);
  const int64_t total = off[n];
  std::vector<int> tmp(std::max<int64_t>(total, 1));
  #pragma omp parallel for schedule(static)
  for (int v = 0; v < n; v++) deg[v].store(off[v], std::memory_order_relaxed);
  #pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < count; i++) {
    const int u = src[i], v = dst[i];
    if (u != v) {
      tmp[deg[u].fetch_add(1, std::memory_order_relaxed)] = v;
      tmp[deg[v].fetch_add(1, std::memory_order_relaxed)] = u;
    }
  }
  // sort + unique each list, record the unique length
  std::vector<int64_t> ulen(n + 1, 0);
  #pragma omp parallel for schedule(dynamic, 256)
  for (int v = 0; v < n; v++) {
    int* b = tmp.data() + off[v];
    int* e = tmp.data() + off[v + 1];
    std::sort(b, e);
    ulen[v + 1] = std::unique(b, e) - b;
  }
  for (int v = 0; v < n; v++) ulen[v + 1] += ulen[v];
  const int64_t m = ulen[n];
  if (m > (int64_t)std::numeric_limits<int>::max()) return IGX_ERR_LIMIT;
  int* nindex = alloc_ints((size_t)n + 1);
  int* nlist = alloc_ints((size_t)m);
  if (!nindex || !nlist) {free(nindex); free(nlist); return IGX_ERR_NOMEM;}
  #pragma omp parallel for schedule(dynamic, 256)
  for (int v = 0; v < n; v++) {
    nindex[v] = (int)ulen[v];
    std::copy(tmp.data() + off[v], tmp.data() + off[v] + (ulen[v + 1] - ulen[v]), nlist + ulen[v]);
  }
  nindex[n] = (int)m;
  out->nodes = n;
  out->edges = (int)m;
  out->nindex = nindex;
  out->nlist = nlist;
  out->eweight = nullptr;
  return IGX_OK;
}

inline int weight_of(int u, int v, int lo, int hi, uint64_t seed)
{
  const uint64_t a = (uint64_t)std::min(u, v), b = (uint64_t)std::max(u, v);
  const uint64_t h = mix64(mix64(seed ^ 0x5851F42D4C957F2Dull) ^ ((a << 32) | b));
  const uint64_t span = (uint64_t)(hi - lo + 1);
  return lo + (int)(h % span);
}

}  // namespace

extern "C" {

// ------------------------------------------------------------ .egr I/O -----
int igx_egr_read(const char* path, igx_graph* out)
{
  if (path == nullptr || out == nullptr) return IGX_ERR_ARG;
  out->nindex = out->nlist = out->eweight = nullptr;
  FILE* f = fopen(path, "rb");
  if (f == nullptr) return IGX_ERR_IO;
  int hdr[2];
  if (fread(hdr, sizeof(int), 2, f) != 2) {fclose(f); return IGX_ERR_IO;}
  const int n = hdr[0], m = hdr[1];
  if (n < 1 || m < 0) {fclose(f); return IGX_ERR_FORMAT;}
  int* nindex = alloc_ints((size_t)n + 1);
  int* nlist = alloc_ints((size_t)m);
  int* ew = alloc_ints((size_t)m);
  if (!nindex || !nlist || !ew) {free(nindex); free(nlist); free(ew); fclose(f); return IGX_ERR_NOMEM;}
  if (fread(nindex, sizeof(int), (size_t)n + 1, f) != (size_t)n + 1 ||
      fread(nlist, sizeof(int), (size_t)m, f) != (size_t)m) {
    free(nindex); free(nlist); free(ew); fclose(f); return IGX_ERR_IO;
  }
  const size_t cnt = (m > 0) ? fread(ew, sizeof(int), (size_t)m, f) : 0;
  if (cnt == 0) {
    free(ew);
    ew = nullptr;
  } else if (cnt != (size_t)m) {
    free(nindex); free(nlist); free(ew); fclose(f); return IGX_ERR_IO;
  }
  fclose(f);
  out->nodes = n;
  out->edges = m;
  out->nindex = nindex;
  out->nlist = nlist;
  out->eweight = ew;
  return IGX_OK;
}

int igx_egr_write(const igx_graph* g, const char* path)
{
  if (g == nullptr || path == nullptr) return IGX_ERR_ARG;
  if (g->nodes < 1 || g->edges < 0) return IGX_ERR_FORMAT;
  FILE* f = fopen(path, "wb");
  if (f == nullptr) return IGX_ERR_IO;
  const int hdr[2] = {g->nodes, g->edges};
  bool ok = fwrite(hdr, sizeof(int), 2, f) == 2;
  ok = ok && fwrite(g->nindex, sizeof(int), (size_t)g->nodes + 1, f) == (size_t)g->nodes + 1;
  ok = ok && fwrite(g->nlist, sizeof(int), (size_t)g->edges, f) == (size_t)g->edges;
  if (ok && g->eweight != nullptr) ok = fwrite(g->eweight, sizeof(int), (size_t)g->edges, f) == (size_t)g->edges;
  fclose(f);
  return ok ? IGX_OK : IGX_ERR_IO;
}

void igx_host_graph_free(igx_graph* g)
{
  if (g == nullptr) return;
  free(g->nindex);
  free(g->nlist);
  free(g->eweight);
  g->nindex = g->nlist = g->eweight = nullptr;
}

int igx_check_csr(const igx_graph* g)
{
  if (g == nullptr || g->nindex == nullptr || (g->edges > 0 && g->nlist == nullptr)) return IGX_ERR_ARG;
  if (g->nodes < 1 || g->edges < 0) return IGX_ERR_FORMAT;
  if (g->nindex[0] != 0 || g->nindex[g->nodes] != g->edges) return IGX_ERR_FORMAT;
  int bad = 0;
  #pragma omp parallel for reduction(|: bad)
  for (int v = 0; v < g->nodes; v++) {
    if (g->nindex[v] > g->nindex[v + 1]) bad |= 1;
  }
  if (bad) return IGX_ERR_FORMAT;
  #pragma omp parallel for reduction(|: bad)
  for (int e = 0; e < g->edges; e++) {
    if (g->nlist[e] < 0 || g->nlist[e] >= g->nodes) bad |= 1;
  }
  return bad ? IGX_ERR_FORMAT : IGX_OK;
}

int igx_check_sorted(const igx_graph* g)
{
  if (g == nullptr) return IGX_ERR_ARG;
  int bad = 0;
  #pragma omp parallel for reduction(|: bad) schedule(dynamic, 1024)
  for (int v = 0; v < g->nodes; v++) {
    for (int i = g->nindex[v] + 1; i < g->nindex[v + 1]; i++) {
      if (g->nlist[i - 1] >= g->nlist[i]) bad |= 1;
    }
  }
  return bad ? IGX_ERR_UNSORTED : IGX_OK;
}

// --------------------------------------------------------- generators -----
int igx_graph_from_edges(int n, int64_t count, const int* src, const int* dst, igx_graph* out)
{
  if (n < 1 || count < 0 || (count > 0 && (src == nullptr || dst == nullptr))) return IGX_ERR_ARG;
  for (int64_t i = 0; i < count; i++) {
    if (src[i] < 0 || src[i] >= n || dst[i] < 0 || dst[i] >= n) return IGX_ERR_ARG;
  }
  return build_csr(n, count, src, dst, out);
}

int igx_gen_er(int n, double avg_degree, uint64_t seed, igx_graph* out)
{
  if (n < 1 || avg_degree < 0 || out == nullptr) return IGX_ERR_ARG;
  if (n == 1 || avg_degree == 0) return build_csr(n, 0, nullptr, nullptr, out);
  const double p = std::min(1.0, avg_degree / (double)(n - 1));
  // Row u samples partners v > u by geometric skipping (Batagelj-Brandes),
  // each row from its own counter stream.
  std::vector<std::vector<int>> rows(n);
  const double lq = std::log1p(-p);
  #pragma omp parallel for schedule(dynamic, 64)
  for (int u = 0; u < n - 1; u++) {
    Stream rng(seed, (uint64_t)u);
    std::vector<int>& r = rows[u];
    int64_t v = u;
    while (true) {
      if (p >= 1.0) {
        v += 1;
      } else {
        const double x = 1.0 - rng.uniform();  // (0, 1]
        v += 1 + (int64_t)std::floor(std::log(x) / lq);
      }
      if (v >= n) break;
      r.push_back((int)v);
    }
  }
  std::vector<int64_t> off(n + 1, 0);
  for (int u = 0; u < n; u++) off[u + 1] = off[u] + (int64_t)rows[u].size();
  std::vector<int> s(std::max<int64_t>(off[n], 1)), d(std::max<int64_t>(off[n], 1));
  #pragma omp parallel for schedule(dynamic, 64)
  for (int u = 0; u < n; u++) {
    for (size_t k = 0; k < rows[u].size(); k++) {
      s[off[u] + k] = u;
      d[off[u] + k] = rows[u][k];
    }
  }
  return build_csr(n, off[n], s.data(), d.data(), out);
}

int igx_gen_rmat(int scale, double edge_factor, double a, double b, double c,
                 int n_limit, uint64_t seed, igx_graph* out)
{
  if (scale < 1 || scale > 30 || edge_factor < 0 || out == nullptr) return IGX_ERR_ARG;
  if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return IGX_ERR_ARG;
  const int64_t full = (int64_t)1 << scale;
  if (n_limit > full) return IGX_ERR_ARG;
  const int n = (n_limit > 0) ? n_limit : (int)full;
  const int64_t samples = (int64_t)std::llround((double)n * edge_factor);
  std::vector<int> s(std::max<int64_t>(samples, 1)), d(std::max<int64_t>(samples, 1));
  // quadrant thresholds on a 32-bit uniform
  const double t1 = a, t2 = a + b, t3 = a + b + c;
  const uint64_t q1 = (uint64_t)(t1 * 4294967296.0), q2 = (uint64_t)(t2 * 4294967296.0), q3 = (uint64_t)(t3 * 4294967296.0);
  std::vector<unsigned char> keep(std::max<int64_t>(samples, 1), 1);
  #pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < samples; i++) {
    Stream rng(seed, (uint64_t)i);
    uint64_t u = 0, v = 0;
    uint64_t bits = 0;
    for (int l = 0; l < scale; l++) {
      if ((l & 1) == 0) bits = rng.next();
      const uint64_t r = (l & 1) ? (bits >> 32) : (bits & 0xFFFFFFFFull);
      const uint64_t bu = (r >= q2) ? 1 : 0;                       // quadrants c, d
      const uint64_t bv = (r >= q1 && r < q2) || (r >= q3) ? 1 : 0;  // quadrants b, d
      u = (u << 1) | bu;
      v = (v << 1) | bv;
    }
    if ((int64_t)u >= n || (int64_t)v >= n) keep[i] = 0;
    s[i] = (int)(u % (uint64_t)n);
    d[i] = (int)(v % (uint64_t)n);
  }
  if (n_limit > 0) {
    int64_t w = 0;
    for (int64_t i = 0; i < samples; i++) {
      if (keep[i]) {s[w] = s[i]; d[w] = d[i]; w++;}
    }
    return build_csr(n, w, s.data(), d.data(), out);
  }
  return build_csr(n, samples, s.data(), d.data(), out);
}

int igx_gen_grid(int side, double long_frac, uint64_t seed, igx_graph* out)
{
  if (side < 1 || side > 46340 || long_frac < 0 || long_frac > 1 || out == nullptr) return IGX_ERR_ARG;
  const int n = side * side;
  const int64_t lattice = 2ll * side * (side - 1);
  std::vector<int> s, d;
  s.reserve(lattice + (int64_t)(n * long_frac * 1.1) + 16);
  d.reserve(s.capacity());
  for (int r = 0; r < side; r++) {
    for (int col = 0; col < side; col++) {
      const int v = r * side + col;
      if (col + 1 < side) {s.push_back(v); d.push_back(v + 1);}
      if (r + 1 < side) {s.push_back(v); d.push_back(v + side);}
    }
  }
  const bool all = long_frac >= 1.0;
  const uint64_t thresh = all ? 0 : (uint64_t)(long_frac * 18446744073709551616.0);
  for (int v = 0; v < n; v++) {
    Stream rng(seed, (uint64_t)v);
    if (all || rng.next() < thresh) {
      const int w = (int)(rng.next() % (uint64_t)n);
      s.push_back(v);
      d.push_back(w);
    }
  }
  return build_csr(n, (int64_t)s.size(), s.data(), d.data(), out);
}

int igx_gen_chunglu(int n, double avg_degree, double gamma, uint64_t seed, igx_graph* out)
{
  if (n < 1 || avg_degree < 0 || gamma <= 1.0 || out == nullptr) return IGX_ERR_ARG;
  std::vector<double> cdf(n);
  const double ex = -1.0 / (gamma - 1.0);
  double acc = 0.0;
  for (int i = 0; i < n; i++) {
    acc += std::pow((double)(i + 1), ex);
    cdf[i] = acc;
  }
  for (int i = 0; i < n; i++) cdf[i] /= acc;
  const int64_t samples = (int64_t)std::llround((double)n * avg_degree / 2.0);
  std::vector<int> s(std::max<int64_t>(samples, 1)), d(std::max<int64_t>(samples, 1));
  #pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < samples; i++) {
    Stream rng(seed, (uint64_t)i);
    const double x = rng.uniform(), y = rng.uniform();
    s[i] = (int)(std::upper_bound(cdf.begin(), cdf.end(), x) - cdf.begin());
    d[i] = (int)(std::upper_bound(cdf.begin(), cdf.end(), y) - cdf.begin());
    if (s[i] >= n) s[i] = n - 1;
    if (d[i] >= n) d[i] = n - 1;
  }
  return build_csr(n, samples, s.data(), d.data(), out);
}

int igx_gen_weights(igx_graph* g, int lo, int hi, uint64_t seed)
{
  if (g == nullptr || lo > hi || lo < 0) return IGX_ERR_ARG;
  free(g->eweight);
  g->eweight = alloc_ints((size_t)g->edges);
  if (g->eweight == nullptr) return IGX_ERR_NOMEM;
  #pragma omp parallel for schedule(dynamic, 1024)
  for (int v = 0; v < g->nodes; v++) {
    for (int i = g->nindex[v]; i < g->nindex[v + 1]; i++) {
      g->eweight[i] = weight_of(v, g->nlist[i], lo, hi, seed);
    }
  }
  return IGX_OK;
}

int igx_ref_weights(igx_graph* g)
{
  if (g == nullptr) return IGX_ERR_ARG;
  free(g->eweight);
  g->eweight = alloc_ints((size_t)g->edges);
  if (g->eweight == nullptr) return IGX_ERR_NOMEM;
  for (int i = 0; i < g->nodes; i++) {
    for (int j = g->nindex[i]; j < g->nindex[i + 1]; j++) {
      const int nei = g->nlist[j];
      // 32-bit wrap-around product as in the reference (computed unsigned to
      // keep the wrap well defined), then the remainder with C semantics.
      const int prod = (int)((uint32_t)i * (uint32_t)nei);
      int w = 1 + (prod % g->nodes);
      if (w < 0) w = -w;
      g->eweight[j] = w;
    }
  }
  return IGX_OK;
}

int igx_build_sp_host(const igx_graph* g, int* sp)
{
  if (g == nullptr || sp == nullptr) return IGX_ERR_ARG;
  #pragma omp parallel for schedule(dynamic, 1024)
  for (int v = 0; v < g->nodes; v++) {
    for (int i = g->nindex[v]; i < g->nindex[v + 1]; i++) sp[i] = v;
  }
  return IGX_OK;
}

int igx_max_degree_vertex(const igx_graph* g)
{
  if (g == nullptr || g->nodes < 1) return -1;
  int best = 0, bdeg = -1;
  for (int v = 0; v < g->nodes; v++) {
    const int d = g->nindex[v + 1] - g->nindex[v];
    if (d > bdeg) {bdeg = d; best = v;}
  }
  return best;
}

// ------------------------------------------------------------ features -----
// Order of the IGX_NUM_FEATURES values:
//  0 log2(n)                      1 log2(m + 1)
//  2 mean degree                  3 degree standard deviation
//  4 log2(max degree + 1)         5 degree Gini coefficient
//  6 log2(max / mean degree)      7 fraction of isolated vertices
//  8 two-sweep BFS diameter estimate (from the max-degree vertex), log2(1 + d)
//  9 clustering estimate from 1024 sampled wedges
// 10 fraction of vertices reached from the max-degree vertex
// 11 fraction of edge endpoints at the top 1% highest-degree vertices
// 12 mean |u - v| / n over sampled edges (id locality)
// 13 fraction of vertices with degree >= 64
// 14 fraction of edges owned by vertices with degree > 1024
// 15 fraction of vertices with degree <= 4
// Breadth-first levels from src; returns the smallest id at the largest level
// (the device version, igx_graph_features_device, picks the same vertex).
static int bfs_far(const igx_graph* g, int src, std::vector<int>& lvl, int* reached)
{
  const int n = g->nodes;
  lvl.assign(n, -1);
  std::vector<int> q;
  q.reserve(1024);
  q.push_back(src);
  lvl[src] = 0;
  size_t head = 0;
  int far = src;
  while (head < q.size()) {
    const int v = q[head++];
    if (lvl[v] > lvl[far] || (lvl[v] == lvl[far] && v < far)) far = v;  // smallest id at the largest level
    for (int i = g->nindex[v]; i < g->nindex[v + 1]; i++) {
      const int u = g->nlist[i];
      if (lvl[u] < 0) {lvl[u] = lvl[v] + 1; q.push_back(u);}
    }
  }
  if (reached) *reached = (int)q.size();
  return far;
}

int igx_feature_draws(uint64_t seed, uint64_t item, int count, uint64_t* out)
{
  if (count < 0 || (count > 0 && out == nullptr)) return IGX_ERR_ARG;
  Stream rng(seed, item);
  for (int k = 0; k < count; k++) out[k] = rng.next();
  return IGX_OK;
}

int igx_graph_features(const igx_graph* g, uint64_t seed, double* f)
{
  if (g == nullptr || f == nullptr || g->nodes < 1) return IGX_ERR_ARG;
  const int n = g->nodes;
  const double m = g->edges;
  std::vector<int> deg(n);
  double sum = 0, sq = 0;
  int maxd = 0, iso = 0, ge64 = 0, le4 = 0;
  double big_edges = 0;
  for (int v = 0; v < n; v++) {
    const int d = g->nindex[v + 1] - g->nindex[v];
    deg[v] = d;
    sum += d;
    sq += (double)d * d;
    maxd = std::max(maxd, d);
    iso += (d == 0);
    ge64 += (d >= 64);
    le4 += (d <= 4);
    if (d > 1024) big_edges += d;
  }
  const double mean = sum / n;
  const double var = std::max(0.0, sq / n - mean * mean);
  std::vector<int> sorted(deg);
  std::sort(sorted.begin(), sorted.end());
  double gini_num = 0;
  for (int i = 0; i < n; i++) gini_num += (2.0 * (i + 1) - n - 1) * sorted[i];
  const double gini = (sum > 0) ? gini_num / (n * sum) : 0.0;
  const int top = std::max(1, n / 100);
  double top_sum = 0;
  for (int i = n - top; i < n; i++) top_sum += sorted[i];
  const int src = igx_max_degree_vertex(g);
  std::vector<int> lvl;
  int reached = 0;
  const int far = bfs_far(g, src, lvl, &reached);
  bfs_far(g, far, lvl, nullptr);
  int diam = 0;
  for (int v = 0; v < n; v++) diam = std::max(diam, lvl[v]);
  // wedge sampling: pick a vertex with degree >= 2 proportional to C(d,2)
  // approximately (uniform vertex among those with d >= 2), two distinct
  // neighbours, test closure by binary search (adjacency is sorted).
  std::vector<int> cand;
  for (int v = 0; v < n; v++) if (deg[v] >= 2) cand.push_back(v);
  double closed = 0;
  int wedges = 0;
  if (!cand.empty()) {
    Stream rng(seed, 0xC1u);
    for (int k = 0; k < 1024; k++) {
      const int v = cand[rng.next() % cand.size()];
      const int d = deg[v];
      const int i1 = (int)(rng.next() % (uint64_t)d);
      int i2 = (int)(rng.next() % (uint64_t)(d - 1));
      if (i2 >= i1) i2++;
      const int a = g->nlist[g->nindex[v] + i1], b = g->nlist[g->nindex[v] + i2];
      const int* bb = g->nlist + g->nindex[a];
      const int* be = g->nlist + g->nindex[a + 1];
      closed += std::binary_search(bb, be, b) ? 1.0 : 0.0;
      wedges++;
    }
  }
  double loc = 0;
  int locn = 0;
  if (g->edges > 0) {
    Stream rng(seed, 0x10Cu);
    for (int k = 0; k < 4096; k++) {
      const int e = (int)(rng.next() % (uint64_t)g->edges);
      const int u = g->nlist[e];
      const int v = (int)(std::upper_bound(g->nindex, g->nindex + n + 1, e) - g->nindex) - 1;
      loc += std::fabs((double)u - v) / n;
      locn++;
    }
  }
  f[0] = std::log2((double)n);
  f[1] = std::log2(m + 1.0);
  f[2] = mean;
  f[3] = std::sqrt(var);
  f[4] = std::log2(maxd + 1.0);
  f[5] = gini;
  f[6] = (mean > 0) ? std::log2(std::max(1.0, maxd / mean)) : 0.0;
  f[7] = (double)iso / n;
  f[8] = std::log2(1.0 + diam);
  f[9] = wedges ? closed / wedges : 0.0;
  f[10] = (double)reached / n;
  f[11] = (sum > 0) ? top_sum / sum : 0.0;
  f[12] = locn ? loc / locn : 0.0;
  f[13] = (double)ge64 / n;
  f[14] = (m > 0) ? big_edges / m : 0.0;
  f[15] = (double)le4 / n;
  return IGX_OK;
}

}  // extern "C"
