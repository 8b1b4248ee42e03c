// The style selector's 16 graph features computed on the device from a
// device-resident CSR, so that picking a style for a graph already in HBM
// does not copy the graph back to the host.  Same values as the host
// function igx_graph_features (graph_host.cpp): integer statistics are
// exact, the Gini coefficient and the top-1% share agree to rounding (their
// sums run in another order), the diameter estimate and the reached fraction
// come from the library's own BFS (two sweeps from the max-degree vertex, the
// second from the smallest id on the last level, as on the host), and the
// wedge and edge samples use the host stream's draws, so they hit the same
// wedges and edges.
//
// Passes (n vertices, m edges):
//   degree pass       reads nindex: sums, counts, maximum, max-degree vertex
//   histogram         degree counts (per-workgroup LDS bins below kLdsBins)
//   scan + bin terms  positions of each degree in sorted order -> Gini and
//                     top-1% terms per bin, without sorting the degrees
//   two BFS sweeps    the default BFS style; a reduction for reached / far
//   samples           1024 wedges (candidate = k-th vertex of degree >= 2,
//                     found by binary search in a scan of the flags) and
//                     4096 edges (row by binary search in nindex)
#include "igx_internal.hpp"
#include "igx_scan.hpp"
#include "igx_graph.h"

#include <cmath>
#include <vector>

namespace igx {
namespace {

constexpr int kFeatBlocks = 512;
constexpr int kLdsBins = 2048;
constexpr int kWedges = 1024;    // graph_host.cpp igx_graph_features: wedge samples
constexpr int kLocEdges = 4096;  // edge samples of the locality feature
constexpr int kDegParts = 8;     // partial values per workgroup of the degree pass

__device__ __forceinline__ long long block_sum_ll(long long v, long long* lds)
{
  v = wave_sum_ll(v);
  if (lane_id() == 0) lds[threadIdx.x / kWave] = v;
  __syncthreads();
  long long t = 0;
  if (threadIdx.x == 0)
    for (int k = 0; k < kWavesPerBlock; k++) t += lds[k];
  __syncthreads();
  return t;
}

__device__ __forceinline__ long long block_max_ll(long long v, long long* lds)
{
  v = wave_max_ll(v);
  if (lane_id() == 0) lds[threadIdx.x / kWave] = v;
  __syncthreads();
  long long t = v;
  if (threadIdx.x == 0)
    for (int k = 0; k < kWavesPerBlock; k++) t = max(t, lds[k]);
  __syncthreads();
  return t;
}

// part[b * kDegParts + k]: degree sum, sum of squares, isolated, >= 64,
// <= 4, edges of rows above 1024, rows of degree >= 2, and the max key
// (degree << 32 | ~id: the highest degree, lowest id on ties)
__global__ void __launch_bounds__(kBlock) feat_degree(const int* __restrict__ nindex, int n,
                                                       long long* __restrict__ part)
{
  __shared__ long long lds[kWavesPerBlock];
  long long sum = 0, sq = 0, iso = 0, ge64 = 0, le4 = 0, big = 0, cand = 0, key = -1;
  const int stride = gridDim.x * kBlock;
  for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
    const long long d = nindex[v + 1] - nindex[v];
    sum += d;
    sq += d * d;
    iso += d == 0;
    ge64 += d >= 64;
    le4 += d <= 4;
    big += (d > 1024) ? d : 0;
    cand += d >= 2;
    key = max(key, (d << 32) | (long long)(0xFFFFFFFFu - (unsigned)v));
  }
  const long long vals[7] = {sum, sq, iso, ge64, le4, big, cand};
  for (int k = 0; k < 7; k++) {
    const long long t = block_sum_ll(vals[k], lds);
    if (threadIdx.x == 0) part[blockIdx.x * kDegParts + k] = t;
  }
  const long long t = block_max_ll(key, lds);
  if (threadIdx.x == 0) part[blockIdx.x * kDegParts + 7] = t;
}

// hist[d] = number of vertices of degree d (hist has maxd + 1 entries); flag
// and its scan input: cflag[v] = degree >= 2
__global__ void __launch_bounds__(kBlock) feat_hist(const int* __restrict__ nindex, int n, int* __restrict__ hist,
                                                     int bins, int* __restrict__ cflag)
{
  __shared__ int s_bin[kLdsBins];
  for (int i = threadIdx.x; i < kLdsBins; i += kBlock) s_bin[i] = 0;
  __syncthreads();
  const int stride = gridDim.x * kBlock;
  for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
    const int d = nindex[v + 1] - nindex[v];
    cflag[v] = d >= 2;
    if (d < kLdsBins) atomicAdd(&s_bin[d], 1);
    else atomicAdd(&hist[d], 1);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kLdsBins && i < bins; i += kBlock) {
    if (s_bin[i] != 0) atomicAdd(&hist[i], s_bin[i]);
  }
}

__global__ void feat_zero(int* __restrict__ p, long long count)
{
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) p[i] = 0;
}

// Per degree bin d (count c, first position p in ascending order): its Gini
// numerator term, the sum over positions i in [p, p + c) of
// (2 (i + 1) - n - 1) d = d c (2 p + c - n), and its share of the `top`
// largest degrees (positions [n - top, n)).
__global__ void __launch_bounds__(kBlock) feat_bins(const int* __restrict__ hist, const int* __restrict__ pos, int bins,
                                                     int n, int top, double* __restrict__ gini_part,
                                                     long long* __restrict__ top_part)
{
  __shared__ long long lds[kWavesPerBlock];
  __shared__ double ldsd[kWavesPerBlock];
  double g = 0.0;
  long long t = 0;
  const int stride = gridDim.x * kBlock;
  for (int d = blockIdx.x * kBlock + threadIdx.x; d < bins; d += stride) {
    const long long c = hist[d];
    if (c == 0) continue;
    const long long p = pos[d];
    g += (double)d * (double)c * (double)(2 * p + c - n);
    const long long lo = max(p, (long long)(n - top)), hi = min(p + c, (long long)n);
    if (hi > lo) t += (long long)d * (hi - lo);
  }
  g = wave_sum_d(g);
  if (lane_id() == 0) ldsd[threadIdx.x / kWave] = g;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < kWavesPerBlock; k++) s += ldsd[k];
    gini_part[blockIdx.x] = s;
  }
  const long long tt = block_sum_ll(t, lds);
  if (threadIdx.x == 0) top_part[blockIdx.x] = tt;
}

// part[2b]: vertices reached; part[2b+1]: max key (level << 32 | ~id)
__global__ void __launch_bounds__(kBlock) feat_levels(const int* __restrict__ lvl, int n, long long* __restrict__ part)
{
  __shared__ long long lds[kWavesPerBlock];
  long long reached = 0, key = -1;
  const int stride = gridDim.x * kBlock;
  for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
    const int l = lvl[v];
    if (l != kInf) {
      reached++;
      key = max(key, ((long long)l << 32) | (long long)(0xFFFFFFFFu - (unsigned)v));
    }
  }
  const long long r = block_sum_ll(reached, lds);
  const long long k = block_max_ll(key, lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r;
    part[2 * blockIdx.x + 1] = k;
  }
}

// first index i in [0, count) with a[i] > key (a non-decreasing), count if none
__device__ __forceinline__ long long upper_index(const int* __restrict__ a, long long count, long long key)
{
  long long lo = 0, hi = count;
  while (lo < hi) {
    const long long mid = (lo + hi) >> 1;
    if ((long long)a[mid] <= key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// One wedge per thread: candidate j = draw % ncand is the vertex v whose
// inclusive flag count first exceeds j (incl = cpos + cflag), then two
// distinct neighbours a, b of v, closed when b is in N(a) (sorted lists).
__global__ void __launch_bounds__(kBlock) feat_wedges(const int* __restrict__ nindex, const int* __restrict__ nlist,
                                                       int n, const int* __restrict__ cpos,
                                                       const int* __restrict__ cflag, long long ncand,
                                                       const unsigned long long* __restrict__ draws,
                                                       int* __restrict__ closed)
{
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= kWedges) return;
  const long long j = (long long)(draws[3 * k] % (unsigned long long)ncand);
  // incl[v] = cpos[v] + cflag[v] is non-decreasing; find the first v with incl[v] > j
  long long lo = 0, hi = n;
  while (lo < hi) {
    const long long mid = (lo + hi) >> 1;
    if ((long long)cpos[mid] + cflag[mid] <= j) lo = mid + 1;
    else hi = mid;
  }
  const int v = (int)lo;
  const int beg = nindex[v];
  const int d = nindex[v + 1] - beg;
  const int i1 = (int)(draws[3 * k + 1] % (unsigned long long)d);
  int i2 = (int)(draws[3 * k + 2] % (unsigned long long)(d - 1));
  if (i2 >= i1) i2++;
  const int a = nlist[beg + i1], b = nlist[beg + i2];
  int l = nindex[a], h = nindex[a + 1];
  int hit = 0;
  while (l < h) {
    const int mid = (l + h) >> 1;
    const int x = nlist[mid];
    if (x == b) {
      hit = 1;
      break;
    }
    if (x < b) l = mid + 1;
    else h = mid;
  }
  closed[k] = hit;
}

// One sampled edge per thread: |u - v| / n with v the row of edge e.
__global__ void __launch_bounds__(kBlock) feat_locality(const int* __restrict__ nindex, const int* __restrict__ nlist,
                                                         int n, long long m,
                                                         const unsigned long long* __restrict__ draws,
                                                         double* __restrict__ terms)
{
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= kLocEdges) return;
  const long long e = (long long)(draws[k] % (unsigned long long)m);
  const int u = nlist[e];
  const int v = (int)upper_index(nindex, (long long)n + 1, e) - 1;
  terms[k] = fabs((double)u - (double)v) / n;
}

struct FeatBuffers {
  int* lvl;          // n: BFS levels
  int* cflag;        // n + 1: degree >= 2 flags
  int* cpos;         // n + 1: exclusive scan of cflag
  int* hist;         // bins
  int* hpos;         // bins: exclusive scan of hist
  int* tiles;        // scan tile sums
  int* total;        // 1
  long long* part;   // per-workgroup partials
  double* gpart;     // kFeatBlocks Gini partials
  unsigned long long* draws;  // 3 * kWedges + kLocEdges
  int* closed;       // kWedges
  double* terms;     // kLocEdges
};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

int feat_buffers(int n, int bins, FeatBuffers* fb)
{
  const int tiles = std::max(scan_tiles_for(n + 1), scan_tiles_for(bins)) + 1;
  const size_t sizes[] = {sizeof(int) * (size_t)n,
                          sizeof(int) * ((size_t)n + 1),
                          sizeof(int) * ((size_t)n + 1),
                          sizeof(int) * (size_t)bins,
                          sizeof(int) * (size_t)bins,
                          sizeof(int) * (size_t)tiles,
                          sizeof(int),
                          sizeof(long long) * kFeatBlocks * kDegParts,
                          sizeof(double) * kFeatBlocks,
                          sizeof(unsigned long long) * (3 * kWedges + kLocEdges),
                          sizeof(int) * kWedges,
                          sizeof(double) * kLocEdges};
  size_t total = 0;
  for (size_t s : sizes) total += align256(s);
  void* base = nullptr;
  IGX_TRY(ws_get(WS_FEAT, total, &base));
  char* p = (char*)base;
  void* ptrs[12];
  for (int i = 0; i < 12; i++) {
    ptrs[i] = p;
    p += align256(sizes[i]);
  }
  fb->lvl = (int*)ptrs[0];
  fb->cflag = (int*)ptrs[1];
  fb->cpos = (int*)ptrs[2];
  fb->hist = (int*)ptrs[3];
  fb->hpos = (int*)ptrs[4];
  fb->tiles = (int*)ptrs[5];
  fb->total = (int*)ptrs[6];
  fb->part = (long long*)ptrs[7];
  fb->gpart = (double*)ptrs[8];
  fb->draws = (unsigned long long*)ptrs[9];
  fb->closed = (int*)ptrs[10];
  fb->terms = (double*)ptrs[11];
  return IGX_OK;
}

template <typename T>
int fetch(std::vector<T>& h, const T* d, size_t count)
{
  h.resize(count);
  IGX_HIP(hipMemcpyAsync(h.data(), d, sizeof(T) * count, hipMemcpyDeviceToHost, rt().stream));
  IGX_HIP(hipStreamSynchronize(rt().stream));
  return IGX_OK;
}

// BFS levels from src with the library's default BFS entry style; returns
// the reached count and the smallest id on the largest level
int bfs_sweep(const igx_graph* g, int src, FeatBuffers& fb, long long* reached, int* far, int* depth)
{
  igx_run_info info = {};
  IGX_TRY(igx_run(IGX_BFS, igx_get_entry_style(IGX_BFS, 0), g, src, fb.lvl, nullptr, &info));
  hipLaunchKernelGGL(feat_levels, dim3(kFeatBlocks), dim3(kBlock), 0, rt().stream, fb.lvl, g->nodes, fb.part);
  IGX_HIP(hipGetLastError());
  std::vector<long long> h;
  IGX_TRY(fetch(h, fb.part, 2 * (size_t)kFeatBlocks));
  long long r = 0, key = -1;
  for (int b = 0; b < kFeatBlocks; b++) {
    r += h[2 * b];
    key = std::max(key, h[2 * b + 1]);
  }
  *reached = r;
  *far = (int)(0xFFFFFFFFu - (unsigned)(key & 0xFFFFFFFFll));
  *depth = (int)(key >> 32);
  return IGX_OK;
}

}  // namespace
}  // namespace igx

using namespace igx;

extern "C" int igx_graph_features_device(const igx_graph* g, uint64_t seed, double* f, double* seconds)
{
  IGX_TRY(ensure_init());
  if (g == nullptr || f == nullptr || g->nodes < 1 || g->edges < 0 || g->nindex == nullptr) return IGX_ERR_ARG;
  if (g->edges > 0 && g->nlist == nullptr) return IGX_ERR_ARG;
  Runtime& R = rt();
  const int n = g->nodes;
  const long long m = g->edges;
  struct Events {  // destroyed on every return path
    hipEvent_t a = nullptr, b = nullptr;
    ~Events()
    {
      if (a) hipEventDestroy(a);
      if (b) hipEventDestroy(b);
    }
  } ev;
  IGX_HIP(hipEventCreate(&ev.a));
  IGX_HIP(hipEventCreate(&ev.b));
  hipEvent_t t0 = ev.a, t1 = ev.b;
  IGX_HIP(hipEventRecord(t0, R.stream));

  // degree pass: the histogram size needs the maximum degree first
  FeatBuffers fb;
  IGX_TRY(feat_buffers(n, 1, &fb));
  hipLaunchKernelGGL(feat_degree, dim3(kFeatBlocks), dim3(kBlock), 0, R.stream, g->nindex, n, fb.part);
  IGX_HIP(hipGetLastError());
  std::vector<long long> h;
  IGX_TRY(fetch(h, fb.part, (size_t)kFeatBlocks * kDegParts));
  long long acc[7] = {};
  long long key = -1;
  for (int b = 0; b < kFeatBlocks; b++) {
    for (int k = 0; k < 7; k++) acc[k] += h[(size_t)b * kDegParts + k];
    key = std::max(key, h[(size_t)b * kDegParts + 7]);
  }
  const double sum = (double)acc[0], sq = (double)acc[1];
  const long long iso = acc[2], ge64 = acc[3], le4 = acc[4], ncand = acc[6];
  const double big_edges = (double)acc[5];
  const int maxd = (int)(key >> 32);
  const int src = (int)(0xFFFFFFFFu - (unsigned)(key & 0xFFFFFFFFll));
  const int bins = maxd + 1;

  // degree histogram, positions in sorted order, Gini and top-1% terms
  IGX_TRY(feat_buffers(n, bins, &fb));
  hipLaunchKernelGGL(feat_zero, dim3(std::min(blocks_for(bins, kBlock), 1024)), dim3(kBlock), 0, R.stream, fb.hist,
                     (long long)bins);
  hipLaunchKernelGGL(feat_hist, dim3(kFeatBlocks), dim3(kBlock), 0, R.stream, g->nindex, n, fb.hist, bins, fb.cflag);
  exclusive_scan(fb.hist, fb.hpos, bins, fb.tiles, fb.total, R.stream);
  const int top = std::max(1, n / 100);
  hipLaunchKernelGGL(feat_bins, dim3(kFeatBlocks), dim3(kBlock), 0, R.stream, fb.hist, fb.hpos, bins, n, top, fb.gpart,
                     fb.part);
  IGX_HIP(hipGetLastError());
  std::vector<double> hg;
  IGX_TRY(fetch(hg, fb.gpart, (size_t)kFeatBlocks));
  IGX_TRY(fetch(h, fb.part, (size_t)kFeatBlocks));
  double gini_num = 0.0;
  long long top_sum = 0;
  for (int b = 0; b < kFeatBlocks; b++) {
    gini_num += hg[b];
    top_sum += h[b];
  }

  // two BFS sweeps (the graph upload owns nindex/nlist; the BFS needs no weights)
  long long reached = 0, unused = 0;
  int far = 0, depth0 = 0, far2 = 0, diam = 0;
  IGX_TRY(bfs_sweep(g, src, fb, &reached, &far, &depth0));
  IGX_TRY(bfs_sweep(g, far, fb, &unused, &far2, &diam));

  // sampled wedges and edges, drawn from the host feature stream
  std::vector<unsigned long long> draws(3 * kWedges + kLocEdges);
  IGX_TRY(igx_feature_draws(seed, 0xC1u, 3 * kWedges, (uint64_t*)draws.data()));
  IGX_TRY(igx_feature_draws(seed, 0x10Cu, kLocEdges, (uint64_t*)draws.data() + 3 * kWedges));
  IGX_HIP(hipMemcpyAsync(fb.draws, draws.data(), sizeof(unsigned long long) * draws.size(), hipMemcpyHostToDevice,
                         R.stream));
  double closed = 0.0;
  int wedges = 0;
  if (ncand > 0) {
    exclusive_scan(fb.cflag, fb.cpos, n, fb.tiles, fb.total, R.stream);
    hipLaunchKernelGGL(feat_wedges, dim3(blocks_for(kWedges, kBlock)), dim3(kBlock), 0, R.stream, g->nindex, g->nlist,
                       n, fb.cpos, fb.cflag, ncand, fb.draws, fb.closed);
    IGX_HIP(hipGetLastError());
    std::vector<int> hc;
    IGX_TRY(fetch(hc, fb.closed, (size_t)kWedges));
    for (int k = 0; k < kWedges; k++) closed += hc[k];
    wedges = kWedges;
  }
  double loc = 0.0;
  int locn = 0;
  if (m > 0) {
    hipLaunchKernelGGL(feat_locality, dim3(blocks_for(kLocEdges, kBlock)), dim3(kBlock), 0, R.stream, g->nindex,
                       g->nlist, n, m, fb.draws + 3 * kWedges, fb.terms);
    IGX_HIP(hipGetLastError());
    std::vector<double> ht;
    IGX_TRY(fetch(ht, fb.terms, (size_t)kLocEdges));
    for (int k = 0; k < kLocEdges; k++) loc += ht[k];  // in sample order, as on the host
    locn = kLocEdges;
  }
  IGX_HIP(hipEventRecord(t1, R.stream));
  IGX_HIP(hipEventSynchronize(t1));
  float ms = 0.0f;
  IGX_HIP(hipEventElapsedTime(&ms, t0, t1));
  if (seconds) *seconds = ms * 1e-3;

  // the same formulas as igx_graph_features
  const double md = (double)m;
  const double mean = sum / n;
  const double var = std::max(0.0, sq / n - mean * mean);
  const double gini = (sum > 0) ? gini_num / (n * sum) : 0.0;
  f[0] = std::log2((double)n);
  f[1] = std::log2(md + 1.0);
  f[2] = mean;
  f[3] = std::sqrt(var);
  f[4] = std::log2(maxd + 1.0);
  f[5] = gini;
  f[6] = (mean > 0) ? std::log2(std::max(1.0, maxd / mean)) : 0.0;
  f[7] = (double)iso / n;
  f[8] = std::log2(1.0 + diam);
  f[9] = wedges ? closed / wedges : 0.0;
  f[10] = (double)reached / n;
  f[11] = (sum > 0) ? (double)top_sum / sum : 0.0;
  f[12] = locn ? loc / locn : 0.0;
  f[13] = (double)ge64 / n;
  f[14] = (md > 0) ? big_edges / md : 0.0;
  f[15] = (double)le4 / n;
  return IGX_OK;
}
