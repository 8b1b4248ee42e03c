// Runtime of libindigo_hip.so: device selection, stream and events, cached
// workspaces, device CSR upload, the style registry dispatch and the
// reference-signature entry points.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "igx_internal.hpp"
#include "igx_frontier.hpp"
#include "igx_scan.hpp"

namespace igx {

namespace {
Runtime g_rt;
char g_last_error[256] = "";
void* g_ws[WS_COUNT] = {};
size_t g_ws_bytes[WS_COUNT] = {};
PrParams g_pr;
igx_run_info g_last_info = {};  // run info of the last reference-signature entry call
int g_entry_style[IGX_NUM_ALGOS][2];
bool g_entry_style_set = false;

// per-kernel event pairs (ktimer_begin / ktimer_end)
char g_prof_tag[64] = "";
std::vector<std::pair<hipEvent_t, hipEvent_t>> g_prof_events;
size_t g_prof_used = 0;
bool g_prof_open = false;

__global__ void ctl_reset_kernel(Ctl* ctl, int n)
{
  if (threadIdx.x == 0) {
    Ctl c = {};
    c.n = n;
    *ctl = c;
  }
  if ((int)threadIdx.x < kGroups) {
    GroupSlot* gs = ctl_groups(ctl) + threadIdx.x;
    gs->qcount[0] = 0;
    gs->qcount[1] = 0;
    gs->tiles = 0;
  }
}

__global__ void fin_flag_kernel(Ctl* ctl)
{
  if (threadIdx.x == 0 && !round_done(ctl)) finalize_flag_round(ctl);
}

__global__ void fin_worklist_kernel(Ctl* ctl)
{
  if (threadIdx.x == 0 && !round_done(ctl)) finalize_worklist_round(ctl);
}

__global__ void build_sp_kernel(const int* __restrict__ nindex, int nodes, int* __restrict__ sp)
{
  // one wave per vertex; lanes stride over the vertex's edge range
  const int wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
  const int lane = lane_id();
  const int waves = gridDim.x * kWavesPerBlock;
  for (int v = wave; v < nodes; v += waves) {
    const int beg = nindex[v], end = nindex[v + 1];
    for (int e = beg + lane; e < end; e += kWave) sp[e] = v;
  }
}
}  // namespace

void set_last_hip_error(hipError_t e, const char* what)
{
  snprintf(g_last_error, sizeof(g_last_error), "%s: %s", what, hipGetErrorString(e));
}

void snprintf_last_error(const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

Runtime& rt() { return g_rt; }
PrParams& pr_params() { return g_pr; }

int ensure_init()
{
  if (g_rt.ready) return IGX_OK;
  return igx_init(0);
}

int ws_get(int slot, size_t bytes, void** ptr)
{
  if (slot < 0 || slot >= WS_COUNT) return IGX_ERR_ARG;
  if (bytes == 0) bytes = 256;
  if (g_ws_bytes[slot] < bytes) {
    if (g_ws[slot] != nullptr) {
      IGX_HIP(hipStreamSynchronize(g_rt.stream));
      IGX_HIP(hipFree(g_ws[slot]));
      g_ws[slot] = nullptr;
      g_ws_bytes[slot] = 0;
    }
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return IGX_ERR_NOMEM;
    }
    g_ws[slot] = p;
    g_ws_bytes[slot] = bytes;
  }
  *ptr = g_ws[slot];
  return IGX_OK;
}

void launch_fin_flag() { hipLaunchKernelGGL(fin_flag_kernel, dim3(1), dim3(64), 0, g_rt.stream, g_rt.d_ctl); }
void launch_fin_worklist() { hipLaunchKernelGGL(fin_worklist_kernel, dim3(1), dim3(64), 0, g_rt.stream, g_rt.d_ctl); }

int ctl_reset(int n)
{
  hipLaunchKernelGGL(ctl_reset_kernel, dim3(1), dim3(64), 0, g_rt.stream, g_rt.d_ctl, n);
  IGX_HIP(hipGetLastError());
  return IGX_OK;
}

int build_sp_async(const igx_graph* g, int* sp)
{
  hipLaunchKernelGGL(build_sp_kernel, dim3(blocks_for((long long)g->nodes * kWave, kBlock)), dim3(kBlock), 0,
                     g_rt.stream, g->nindex, g->nodes, sp);
  IGX_HIP(hipGetLastError());
  return IGX_OK;
}

int device_sp(const igx_graph* g, const int** d_sp)
{
  // Rebuilt on every call: a cache keyed by device pointers goes stale when
  // a caller frees a graph and uploads another one at the same address.
  void* p = nullptr;
  IGX_TRY(ws_get(WS_SP, sizeof(int) * (size_t)(g->edges > 0 ? g->edges : 1), &p));
  hipLaunchKernelGGL(build_sp_kernel, dim3(blocks_for((long long)g->nodes * kWave, kBlock)), dim3(kBlock), 0,
                     g_rt.stream, g->nindex, g->nodes, (int*)p);
  IGX_HIP(hipGetLastError());
  *d_sp = (const int*)p;
  return IGX_OK;
}

// ---------------------------------------------------- long-row chunk list
// No atomics, balanced however the long rows are placed, in two launches.
// big_rows_count sums the chunks of every range of kRangeRows consecutive
// rows (one workgroup per range).  The work of writing the list is cut into
// units: a range with c chunks is ceil(c / kUnitChunks) units (ranges without
// long rows are none), and big_rows_fill hands the units to its workgroups
// round robin.  For a unit, a workgroup stages the range's row starts and
// chunk prefixes in LDS with one coalesced pass and writes up to kUnitChunks
// entries with 256 lanes; R-MAT's first range alone holds thousands of hub
// chunks, which this way spread over many workgroups.  Each workgroup scans
// the range counts itself (up to kLdsRanges ranges, in LDS); larger graphs
// get the two prefixes from device-wide scans instead.
constexpr int kRangeRows = 2048;
constexpr int kRowsPerLane = kRangeRows / kBlock;
constexpr int kLdsRanges = 4096;
constexpr int kUnitChunks = 1024;

// a row [b, e) longer than kChunk is cut into chunks of kChunk edges
__device__ __forceinline__ int row_chunks(int b, int e)
{
  return (e - b > kChunk) ? chunks_of(e - b) : 0;
}

__device__ __forceinline__ int long_row_chunks(const int* __restrict__ nindex, int n, int v)
{
  if (v >= n) return 0;
  return row_chunks(nindex[v], nindex[v + 1]);
}

__device__ __forceinline__ int range_units(int c) { return (c + kUnitChunks - 1) / kUnitChunks; }

// range_count[r] = chunks of rows [r * kRangeRows, (r + 1) * kRangeRows) and,
// when range_units is given, range_units[r] = its units; entry nranges of
// both is 0, so that an exclusive scan's last entry is the total
__global__ void __launch_bounds__(kBlock) big_rows_count(const int* __restrict__ nindex, int n, int nranges,
                                                          int* __restrict__ range_count, int* __restrict__ range_unit)
{
  __shared__ int lds[kWavesPerBlock];
  for (int r = blockIdx.x; r < nranges; r += gridDim.x) {  // uniform per block
    const long long base = (long long)r * kRangeRows;
    int c = 0;
    #pragma unroll
    for (int i = 0; i < kRowsPerLane; i++) {
      const long long v = base + i * kBlock + threadIdx.x;
      c += (v < n) ? long_row_chunks(nindex, n, (int)v) : 0;
    }
    int tot = 0;
    block_excl_scan(c, lds, &tot);
    if (threadIdx.x == 0) {
      range_count[r] = tot;
      if (range_unit) range_unit[r] = range_units(tot);
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    range_count[nranges] = 0;
    if (range_unit) range_unit[nranges] = 0;
  }
}

// kLdsOff: range_a holds the nranges range counts (nranges <= kLdsRanges) and
// the prefixes are built in dynamic LDS (2 * (nranges + 1) ints); otherwise
// range_a / range_b hold the exclusive prefixes of the chunk and unit counts,
// nranges + 1 entries each, the last one the total.  Writes the total chunk
// count to *d_count.
template <bool kLdsOff>
__global__ void __launch_bounds__(kBlock) big_rows_fill(const int* __restrict__ nindex, int n, int nranges,
                                                         const int* __restrict__ range_a,
                                                         const int* __restrict__ range_b, Entry* __restrict__ bigq,
                                                         int* __restrict__ d_count)
{
  extern __shared__ int s_dyn[];
  __shared__ int s_beg[kRangeRows + 1];
  __shared__ int s_pre[kRangeRows];
  __shared__ int s_red[kWavesPerBlock];
  const int tid = threadIdx.x;
  const int* off = range_a;
  const int* uoff = range_b;
  if (kLdsOff) {
    // coalesced copy of the counts, then each wave scans a contiguous quarter
    // 64 entries at a time and adds the quarters before it
    int* so = s_dyn;
    int* su = s_dyn + nranges + 1;
    for (int r = tid; r < nranges; r += kBlock) so[r] = range_a[r];
    __syncthreads();
    const int w = tid / kWave, lane = lane_id();
    const int quarter = (nranges + kWavesPerBlock - 1) / kWavesPerBlock;
    const int qb = w * quarter, qe = (qb + quarter < nranges) ? qb + quarter : nranges;
    int cc = 0, cu = 0;
    for (int b = qb; b < qe; b += kWave) {  // uniform per wave
      const int r = b + lane;
      const int c = (r < qe) ? so[r] : 0, u = range_units(c);
      const int ic = wave_incl_scan(c), iu = wave_incl_scan(u);
      if (r < qe) {
        so[r] = cc + ic - c;
        su[r] = cu + iu - u;
      }
      cc += wave_bcast(ic, kWave - 1);
      cu += wave_bcast(iu, kWave - 1);
    }
    __shared__ int s_wc[kWavesPerBlock], s_wu[kWavesPerBlock];
    if (lane == 0) {
      s_wc[w] = cc;
      s_wu[w] = cu;
    }
    __syncthreads();
    int oc = 0, ou = 0, tc = 0, tu = 0;
    for (int k = 0; k < kWavesPerBlock; k++) {
      if (k < w) {
        oc += s_wc[k];
        ou += s_wu[k];
      }
      tc += s_wc[k];
      tu += s_wu[k];
    }
    for (int r = qb + lane; r < qe; r += kWave) {
      so[r] += oc;
      su[r] += ou;
    }
    if (tid == 0) {
      so[nranges] = tc;
      su[nranges] = tu;
    }
    __syncthreads();
    off = so;
    uoff = su;
  }
  const int total = off[nranges];
  const int units = uoff[nranges];
  if (blockIdx.x == 0 && tid == 0) *d_count = total;
  for (int j = blockIdx.x; j < units; j += gridDim.x) {  // uniform per block
    int lo = 0, hi = nranges;  // the range of unit j: last r with uoff[r] <= j
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (uoff[mid] <= j) lo = mid;
      else hi = mid;
    }
    const int r = lo;
    const int rb = off[r], c = off[r + 1] - rb;
    const int k = j - uoff[r];
    const long long base = (long long)r * kRangeRows;
    #pragma unroll
    for (int i = 0; i < kRowsPerLane; i++) {
      const long long v = base + i * kBlock + tid;
      s_beg[i * kBlock + tid] = nindex[(v < n) ? v : n];
    }
    if (tid == 0) s_beg[kRangeRows] = nindex[(base + kRangeRows < n) ? base + kRangeRows : n];
    __syncthreads();
    const int row0 = tid * kRowsPerLane;
    int cr[kRowsPerLane];
    int local = 0;
    #pragma unroll
    for (int i = 0; i < kRowsPerLane; i++) {
      cr[i] = (base + row0 + i < n) ? row_chunks(s_beg[row0 + i], s_beg[row0 + i + 1]) : 0;
      local += cr[i];
    }
    int tot = 0;
    int run = block_excl_scan(local, s_red, &tot);
    #pragma unroll
    for (int i = 0; i < kRowsPerLane; i++) {
      s_pre[row0 + i] = run;
      run += cr[i];
    }
    __syncthreads();
    const int qe = (c < (k + 1) * kUnitChunks) ? c : (k + 1) * kUnitChunks;
    for (int q = k * kUnitChunks + tid; q < qe; q += kBlock) {
      int l = 0, h = kRangeRows;  // last row with s_pre <= q (rows without chunks share the next prefix)
      while (h - l > 1) {
        const int mid = (l + h) >> 1;
        if (s_pre[mid] <= q) l = mid;
        else h = mid;
      }
      const int b = s_beg[l] + (q - s_pre[l]) * kChunk, e = s_beg[l + 1];
      bigq[rb + q] = Entry{(int)(base + l), b, (b + kChunk < e) ? b + kChunk : e, 0};
    }
    __syncthreads();
  }
}

int big_rows(const igx_graph* g, void* bigq, int* d_count)
{
  const int nranges = (int)(((long long)g->nodes + kRangeRows - 1) / kRangeRows);
  const int cnt = nranges + 1;
  const int tsum = scan_tiles_for(cnt);
  void* pb = nullptr;
  IGX_TRY(ws_get(WS_BLOCKS, sizeof(int) * (4 * (size_t)cnt + tsum + 2), &pb));
  int* range_count = (int*)pb;
  int* range_unit = range_count + cnt;
  int* count_off = range_unit + cnt;
  int* unit_off = count_off + cnt;
  int* tile_sum = unit_off + cnt;
  int* totals = tile_sum + tsum;
  const int grid = g_rt.persist_blocks;
  const int cgrid = (nranges < grid) ? ((nranges > 0) ? nranges : 1) : grid;
  const bool lds = nranges <= kLdsRanges;
  hipLaunchKernelGGL(big_rows_count, dim3(cgrid), dim3(kBlock), 0, g_rt.stream, g->nindex, g->nodes, nranges,
                     range_count, lds ? (int*)nullptr : range_unit);
  if (lds) {
    hipLaunchKernelGGL(big_rows_fill<true>, dim3(grid), dim3(kBlock), sizeof(int) * 2 * (size_t)cnt, g_rt.stream,
                       g->nindex, g->nodes, nranges, (const int*)range_count, (const int*)nullptr, (Entry*)bigq,
                       d_count);
  } else {
    exclusive_scan(range_count, count_off, cnt, tile_sum, totals, g_rt.stream);
    exclusive_scan(range_unit, unit_off, cnt, tile_sum, totals + 1, g_rt.stream);
    hipLaunchKernelGGL(big_rows_fill<false>, dim3(grid), dim3(kBlock), 0, g_rt.stream, g->nindex, g->nodes, nranges,
                       (const int*)count_off, (const int*)unit_off, (Entry*)bigq, d_count);
  }
  IGX_HIP(hipGetLastError());
  return IGX_OK;
}

void ktimer_begin(const char* tag)
{
  if (g_prof_tag[0] == 0 || strcmp(tag, g_prof_tag) != 0 || g_prof_open) return;
  if (g_prof_used == g_prof_events.size()) {
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
    g_prof_events.push_back({a, b});
  }
  if (hipEventRecord(g_prof_events[g_prof_used].first, g_rt.stream) == hipSuccess) g_prof_open = true;
}

void ktimer_end(const char* tag)
{
  if (!g_prof_open || strcmp(tag, g_prof_tag) != 0) return;
  hipEventRecord(g_prof_events[g_prof_used].second, g_rt.stream);
  g_prof_used++;
  g_prof_open = false;
}

int timer_start()
{
  IGX_HIP(hipEventRecord(g_rt.ev_start, g_rt.stream));
  g_rt.t_start = std::chrono::steady_clock::now();
  return IGX_OK;
}

bool past_time_limit()
{
  if (g_rt.time_limit <= 0.0) return false;
  const std::chrono::duration<double> dt = std::chrono::steady_clock::now() - g_rt.t_start;
  return dt.count() > g_rt.time_limit;
}

int timer_stop(double* seconds)
{
  IGX_HIP(hipEventRecord(g_rt.ev_stop, g_rt.stream));
  IGX_HIP(hipEventSynchronize(g_rt.ev_stop));
  float ms = 0.0f;
  IGX_HIP(hipEventElapsedTime(&ms, g_rt.ev_start, g_rt.ev_stop));
  if (seconds) *seconds = ms * 1e-3;
  return IGX_OK;
}

const StyleEntry* styles_of(int algo, int* count)
{
  switch (algo) {
    case IGX_BFS: return bfs_styles(count);
    case IGX_SSSP: return sssp_styles(count);
    case IGX_CC: return cc_styles(count);
    case IGX_MIS: return mis_styles(count);
    case IGX_PR: return pr_styles(count);
    case IGX_TC: return tc_styles(count);
    default: *count = 0; return nullptr;
  }
}

static int find_style(int algo, const char* name)
{
  int cnt = 0;
  const StyleEntry* s = styles_of(algo, &cnt);
  for (int i = 0; i < cnt; i++) {
    if (strcmp(s[i].name, name) == 0) return i;
  }
  return 0;
}

static void init_entry_styles()
{
  if (g_entry_style_set) return;
  // defaults: fastest measured vertex-kind and edge-kind styles
  g_entry_style[IGX_BFS][0] = find_style(IGX_BFS, "do_hybrid");
  g_entry_style[IGX_BFS][1] = find_style(IGX_BFS, "e_topo_push_thread");
  g_entry_style[IGX_SSSP][0] = find_style(IGX_SSSP, "v_do_balanced");
  g_entry_style[IGX_SSSP][1] = find_style(IGX_SSSP, "e_data_push_thread");
  g_entry_style[IGX_CC][0] = find_style(IGX_CC, "uf_afforest");
  g_entry_style[IGX_CC][1] = find_style(IGX_CC, "e_topo_push_thread");
  g_entry_style[IGX_MIS][0] = find_style(IGX_MIS, "v_resume_pull");
  g_entry_style[IGX_MIS][1] = find_style(IGX_MIS, "e_topo_thread");
  g_entry_style[IGX_PR][0] = find_style(IGX_PR, "pull_balanced_kahan");
  g_entry_style[IGX_PR][1] = g_entry_style[IGX_PR][0];
  g_entry_style[IGX_TC][0] = find_style(IGX_TC, "oriented_lds");
  g_entry_style[IGX_TC][1] = find_style(IGX_TC, "e_bsearch_thread");
  g_entry_style_set = true;
}

}  // namespace igx

using namespace igx;

extern "C" {

int igx_init(int device)
{
  if (g_rt.ready) return (device == g_rt.device) ? IGX_OK : IGX_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return IGX_ERR_NODEV;
  }
  if (device < 0 || device >= ndev) return IGX_ERR_ARG;
  IGX_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  IGX_HIP(hipGetDeviceProperties(&prop, device));
  g_rt.device = device;
  g_rt.cus = prop.multiProcessorCount;
  g_rt.max_threads_per_cu = prop.maxThreadsPerMultiProcessor;
  g_rt.persist_blocks = g_rt.cus * (g_rt.max_threads_per_cu / kBlock);
  if (g_rt.persist_blocks < 1) g_rt.persist_blocks = 1;
  if (g_rt.persist_blocks > kMaxParts) g_rt.persist_blocks = kMaxParts;  // one partial slot per block
  IGX_HIP(hipStreamCreateWithFlags(&g_rt.stream, hipStreamNonBlocking));
  IGX_HIP(hipEventCreate(&g_rt.ev_start));
  IGX_HIP(hipEventCreate(&g_rt.ev_stop));
  IGX_HIP(hipMalloc((void**)&g_rt.d_ctl, kCtlAllocBytes));
  IGX_HIP(hipMemset(g_rt.d_ctl, 0, kCtlAllocBytes));
  IGX_HIP(hipHostMalloc((void**)&g_rt.h_ctl, sizeof(Ctl), hipHostMallocDefault));
  memset(g_rt.h_ctl, 0, sizeof(Ctl));
  IGX_TRY(batch_noalloc_setup(device));
  g_rt.ready = true;
  init_entry_styles();
  return IGX_OK;
}

int igx_device_info(char* name, int name_len, int* cus, int* max_threads_per_cu)
{
  IGX_TRY(ensure_init());
  if (name != nullptr && name_len > 0) {
    hipDeviceProp_t prop;
    IGX_HIP(hipGetDeviceProperties(&prop, g_rt.device));
    snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
  }
  if (cus) *cus = g_rt.cus;
  if (max_threads_per_cu) *max_threads_per_cu = g_rt.max_threads_per_cu;
  return IGX_OK;
}

int igx_sync(void)
{
  IGX_TRY(ensure_init());
  IGX_HIP(hipStreamSynchronize(g_rt.stream));
  return IGX_OK;
}

int igx_release_workspace(void)
{
  if (!g_rt.ready) return IGX_OK;
  IGX_HIP(hipStreamSynchronize(g_rt.stream));
  for (int i = 0; i < WS_COUNT; i++) {
    if (g_ws[i] != nullptr) IGX_HIP(hipFree(g_ws[i]));
    g_ws[i] = nullptr;
    g_ws_bytes[i] = 0;
  }
  return batch_release();
}

int igx_shutdown(void)
{
  if (!g_rt.ready) return IGX_OK;
  int rc = igx_release_workspace();
  for (auto& e : g_prof_events) {
    hipEventDestroy(e.first);
    hipEventDestroy(e.second);
  }
  g_prof_events.clear();
  g_prof_used = 0;
  g_prof_open = false;
  g_prof_tag[0] = 0;
  if (hipEventDestroy(g_rt.ev_start) != hipSuccess) rc = IGX_ERR_HIP;
  if (hipEventDestroy(g_rt.ev_stop) != hipSuccess) rc = IGX_ERR_HIP;
  if (hipHostFree(g_rt.h_ctl) != hipSuccess) rc = IGX_ERR_HIP;
  if (hipFree(g_rt.d_ctl) != hipSuccess) rc = IGX_ERR_HIP;
  if (hipStreamDestroy(g_rt.stream) != hipSuccess) rc = IGX_ERR_HIP;
  (void)hipGetLastError();
  g_rt = Runtime{};
  return rc;
}

const char* igx_last_error(void) { return g_last_error; }

int igx_last_run_info(igx_run_info* info)
{
  if (info == nullptr) return IGX_ERR_ARG;
  *info = g_last_info;
  return IGX_OK;
}

const char* igx_status_string(int status)
{
  switch (status) {
    case IGX_OK: return "ok";
    case IGX_ERR_ARG: return "invalid argument";
    case IGX_ERR_IO: return "file i/o error";
    case IGX_ERR_FORMAT: return "invalid graph format";
    case IGX_ERR_NOMEM: return "out of memory";
    case IGX_ERR_HIP: return "HIP runtime error";
    case IGX_ERR_UNSORTED: return "adjacency list not sorted or contains self edge";
    case IGX_ERR_NO_WEIGHTS: return "graph has no edge weights";
    case IGX_ERR_STYLE: return "unknown style";
    case IGX_ERR_NODEV: return "no GPU device";
    case IGX_ERR_LIMIT: return "size or round limit exceeded";
    case IGX_ERR_TIMEOUT: return "run stopped at the time limit";
    default: return "unknown status";
  }
}

int igx_alloc(void** d_ptr, size_t bytes)
{
  IGX_TRY(ensure_init());
  if (d_ptr == nullptr) return IGX_ERR_ARG;
  if (hipMalloc(d_ptr, bytes > 0 ? bytes : 4) != hipSuccess) {
    (void)hipGetLastError();
    return IGX_ERR_NOMEM;
  }
  return IGX_OK;
}

int igx_free(void* d_ptr)
{
  IGX_TRY(ensure_init());
  if (d_ptr == nullptr) return IGX_OK;
  IGX_HIP(hipStreamSynchronize(g_rt.stream));
  IGX_HIP(hipFree(d_ptr));
  return IGX_OK;
}

int igx_memcpy_h2d(void* d_dst, const void* h_src, size_t bytes)
{
  IGX_TRY(ensure_init());
  if (bytes == 0) return IGX_OK;
  IGX_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, g_rt.stream));
  IGX_HIP(hipStreamSynchronize(g_rt.stream));
  return IGX_OK;
}

int igx_memcpy_d2h(void* h_dst, const void* d_src, size_t bytes)
{
  IGX_TRY(ensure_init());
  if (bytes == 0) return IGX_OK;
  IGX_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, g_rt.stream));
  IGX_HIP(hipStreamSynchronize(g_rt.stream));
  return IGX_OK;
}

int igx_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes)
{
  IGX_TRY(ensure_init());
  if (bytes == 0) return IGX_OK;
  IGX_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, g_rt.stream));
  IGX_HIP(hipStreamSynchronize(g_rt.stream));
  return IGX_OK;
}

int igx_memset(void* d_ptr, int value, size_t bytes)
{
  IGX_TRY(ensure_init());
  if (bytes == 0) return IGX_OK;
  IGX_HIP(hipMemsetAsync(d_ptr, value, bytes, g_rt.stream));
  IGX_HIP(hipStreamSynchronize(g_rt.stream));
  return IGX_OK;
}

int igx_graph_upload(const igx_graph* host, igx_graph* dev)
{
  IGX_TRY(ensure_init());
  if (host == nullptr || dev == nullptr || host->nindex == nullptr) return IGX_ERR_ARG;
  if (host->nodes < 1 || host->edges < 0) return IGX_ERR_FORMAT;
  if (host->edges > 0 && host->nlist == nullptr) return IGX_ERR_ARG;
  int rc = igx_check_csr(host);
  if (rc != IGX_OK) return rc;
  igx_graph d = *host;
  d.nindex = d.nlist = d.eweight = nullptr;
  const size_t nb = sizeof(int) * ((size_t)host->nodes + 1);
  const size_t eb = sizeof(int) * (size_t)(host->edges > 0 ? host->edges : 1);
  if (igx_alloc((void**)&d.nindex, nb) != IGX_OK || igx_alloc((void**)&d.nlist, eb) != IGX_OK ||
      (host->eweight != nullptr && igx_alloc((void**)&d.eweight, eb) != IGX_OK)) {
    igx_graph_free_device(&d);
    return IGX_ERR_NOMEM;
  }
  IGX_HIP(hipMemcpyAsync(d.nindex, host->nindex, nb, hipMemcpyHostToDevice, g_rt.stream));
  if (host->edges > 0) {
    IGX_HIP(hipMemcpyAsync(d.nlist, host->nlist, sizeof(int) * (size_t)host->edges, hipMemcpyHostToDevice,
                           g_rt.stream));
    if (host->eweight != nullptr) {
      IGX_HIP(hipMemcpyAsync(d.eweight, host->eweight, sizeof(int) * (size_t)host->edges, hipMemcpyHostToDevice,
                             g_rt.stream));
    }
  }
  IGX_HIP(hipStreamSynchronize(g_rt.stream));
  *dev = d;
  return IGX_OK;
}

int igx_graph_free_device(igx_graph* dev)
{
  if (dev == nullptr) return IGX_ERR_ARG;
  igx_free(dev->nindex);
  igx_free(dev->nlist);
  igx_free(dev->eweight);
  dev->nindex = dev->nlist = dev->eweight = nullptr;
  return IGX_OK;
}

int igx_build_sp(const igx_graph* dev, int** d_sp)
{
  IGX_TRY(ensure_init());
  if (dev == nullptr || d_sp == nullptr) return IGX_ERR_ARG;
  int* p = nullptr;
  IGX_TRY(igx_alloc((void**)&p, sizeof(int) * (size_t)(dev->edges > 0 ? dev->edges : 1)));
  hipLaunchKernelGGL(build_sp_kernel, dim3(blocks_for((long long)dev->nodes * kWave, kBlock)), dim3(kBlock), 0,
                     g_rt.stream, dev->nindex, dev->nodes, p);
  IGX_HIP(hipGetLastError());
  IGX_HIP(hipStreamSynchronize(g_rt.stream));
  *d_sp = p;
  return IGX_OK;
}

int igx_num_styles(int algo)
{
  int cnt = 0;
  styles_of(algo, &cnt);
  return cnt;
}

const char* igx_style_name(int algo, int style)
{
  int cnt = 0;
  const StyleEntry* s = styles_of(algo, &cnt);
  if (s == nullptr || style < 0 || style >= cnt) return nullptr;
  return s[style].name;
}

int igx_style_edge_based(int algo, int style)
{
  int cnt = 0;
  const StyleEntry* s = styles_of(algo, &cnt);
  if (s == nullptr || style < 0 || style >= cnt) return -1;
  return s[style].edge_based;
}

int igx_get_entry_style(int algo, int edge_kind)
{
  if (algo < 0 || algo >= IGX_NUM_ALGOS) return -1;
  init_entry_styles();
  return g_entry_style[algo][edge_kind ? 1 : 0];
}

int igx_set_entry_style(int algo, int edge_kind, int style)
{
  if (algo < 0 || algo >= IGX_NUM_ALGOS) return IGX_ERR_ARG;
  if (style < 0 || style >= igx_num_styles(algo)) return IGX_ERR_STYLE;
  init_entry_styles();
  g_entry_style[algo][edge_kind ? 1 : 0] = style;
  return IGX_OK;
}

int igx_profile_kernel(const char* tag)
{
  IGX_TRY(ensure_init());
  IGX_HIP(hipStreamSynchronize(g_rt.stream));
  g_prof_used = 0;
  g_prof_open = false;
  if (tag == nullptr) {
    g_prof_tag[0] = 0;
  } else {
    snprintf(g_prof_tag, sizeof(g_prof_tag), "%s", tag);
  }
  return IGX_OK;
}

int igx_profile_read(double* total_ms, int* launches)
{
  IGX_TRY(ensure_init());
  IGX_HIP(hipStreamSynchronize(g_rt.stream));
  double sum = 0.0;
  for (size_t i = 0; i < g_prof_used; i++) {
    float ms = 0.0f;
    IGX_HIP(hipEventElapsedTime(&ms, g_prof_events[i].first, g_prof_events[i].second));
    sum += ms;
  }
  if (total_ms) *total_ms = sum;
  if (launches) *launches = (int)g_prof_used;
  g_prof_used = 0;
  return IGX_OK;
}

int igx_set_time_limit(double seconds)
{
  if (!(seconds >= 0.0)) return IGX_ERR_ARG;
  g_rt.time_limit = seconds;
  return IGX_OK;
}

int igx_set_pr_params(double eps, int max_iter)
{
  if (!(eps >= 0.0) || max_iter < 1) return IGX_ERR_ARG;
  g_pr.eps = eps;
  g_pr.max_iter = max_iter;
  return IGX_OK;
}

int igx_run(int algo, int style, const igx_graph* dev, int src, void* d_out, const int* d_aux_sp, igx_run_info* info)
{
  IGX_TRY(ensure_init());
  int cnt = 0;
  const StyleEntry* s = styles_of(algo, &cnt);
  if (s == nullptr) return IGX_ERR_ARG;
  if (style < 0 || style >= cnt) return IGX_ERR_STYLE;
  if (dev == nullptr || d_out == nullptr || dev->nindex == nullptr || dev->nodes < 1 || dev->edges < 0) {
    return IGX_ERR_ARG;
  }
  if (dev->edges > 0 && dev->nlist == nullptr) return IGX_ERR_ARG;
  if ((algo == IGX_BFS || algo == IGX_SSSP) && (src < 0 || src >= dev->nodes)) return IGX_ERR_ARG;
  if (algo == IGX_SSSP && dev->eweight == nullptr && dev->edges > 0) return IGX_ERR_NO_WEIGHTS;
  igx_run_info local = {};
  const int* sp = d_aux_sp;
  if (s[style].edge_based && sp == nullptr) IGX_TRY(device_sp(dev, &sp));
  int rc = s[style].fn(dev, src, d_out, sp, &local);
  if (info) *info = local;
  return rc;
}

// ---------------------------------------------- reference-signature entries
static int run_to_host(int algo, int edge_kind, const igx_graph* g, int src, void* h_out, size_t out_bytes,
                       const int* d_sp, igx_run_info* info)
{
  void* d_out = nullptr;
  IGX_TRY(ws_get(WS_SCRATCH, out_bytes, &d_out));
  const int style = igx_get_entry_style(algo, edge_kind);
  g_last_info = igx_run_info{};
  IGX_TRY(igx_run(algo, style, g, src, d_out, d_sp, info));
  if (info) g_last_info = *info;
  if (h_out != nullptr && out_bytes > 0) IGX_TRY(igx_memcpy_d2h(h_out, d_out, out_bytes));
  return IGX_OK;
}

static int upload_sp(const igx_graph* g, const int* sp_host, int** d_sp)
{
  *d_sp = nullptr;
  if (sp_host == nullptr) return IGX_OK;
  IGX_TRY(igx_alloc((void**)d_sp, sizeof(int) * (size_t)(g->edges > 0 ? g->edges : 1)));
  return igx_memcpy_h2d(*d_sp, sp_host, sizeof(int) * (size_t)g->edges);
}

int igx_bfs_vertex(int src, const igx_graph* g, int* dist, double* seconds)
{
  IGX_TRY(ensure_init());
  if (g == nullptr || dist == nullptr) return IGX_ERR_ARG;
  igx_run_info info = {};
  IGX_TRY(run_to_host(IGX_BFS, 0, g, src, dist, sizeof(int) * (size_t)g->nodes, nullptr, &info));
  if (seconds) *seconds = info.seconds;
  return IGX_OK;
}

int igx_bfs_edge(int src, const igx_graph* g, int* dist, const int* sp, double* seconds)
{
  IGX_TRY(ensure_init());
  if (g == nullptr || dist == nullptr) return IGX_ERR_ARG;
  int* d_sp = nullptr;
  IGX_TRY(upload_sp(g, sp, &d_sp));
  igx_run_info info = {};
  const int rc = run_to_host(IGX_BFS, 1, g, src, dist, sizeof(int) * (size_t)g->nodes, d_sp, &info);
  igx_free(d_sp);
  if (seconds) *seconds = info.seconds;
  return rc;
}

int igx_sssp_vertex(int src, const igx_graph* g, int* dist, double* seconds)
{
  IGX_TRY(ensure_init());
  if (g == nullptr || dist == nullptr) return IGX_ERR_ARG;
  igx_run_info info = {};
  IGX_TRY(run_to_host(IGX_SSSP, 0, g, src, dist, sizeof(int) * (size_t)g->nodes, nullptr, &info));
  if (seconds) *seconds = info.seconds;
  return IGX_OK;
}

int igx_sssp_edge(int src, const igx_graph* g, int* dist, const int* sp, double* seconds)
{
  IGX_TRY(ensure_init());
  if (g == nullptr || dist == nullptr) return IGX_ERR_ARG;
  int* d_sp = nullptr;
  IGX_TRY(upload_sp(g, sp, &d_sp));
  igx_run_info info = {};
  const int rc = run_to_host(IGX_SSSP, 1, g, src, dist, sizeof(int) * (size_t)g->nodes, d_sp, &info);
  igx_free(d_sp);
  if (seconds) *seconds = info.seconds;
  return rc;
}

int igx_cc_vertex(const igx_graph* g, int* label, double* seconds)
{
  IGX_TRY(ensure_init());
  if (g == nullptr || label == nullptr) return IGX_ERR_ARG;
  igx_run_info info = {};
  IGX_TRY(run_to_host(IGX_CC, 0, g, 0, label, sizeof(int) * (size_t)g->nodes, nullptr, &info));
  if (seconds) *seconds = info.seconds;
  return IGX_OK;
}

int igx_cc_edge(const igx_graph* g, int* label, const int* sp, double* seconds)
{
  IGX_TRY(ensure_init());
  if (g == nullptr || label == nullptr) return IGX_ERR_ARG;
  int* d_sp = nullptr;
  IGX_TRY(upload_sp(g, sp, &d_sp));
  igx_run_info info = {};
  const int rc = run_to_host(IGX_CC, 1, g, 0, label, sizeof(int) * (size_t)g->nodes, d_sp, &info);
  igx_free(d_sp);
  if (seconds) *seconds = info.seconds;
  return rc;
}

static void fill_priorities(unsigned int* priority, int n)
{
  if (priority == nullptr) return;
  for (int v = 0; v < n; v++) {
    unsigned int val = (unsigned int)v + 712313887u;
    val = ((val >> 16) ^ val) * 0x45d9f3bu;
    val = ((val >> 16) ^ val) * 0x45d9f3bu;
    priority[v] = (val >> 16) ^ val;
  }
}

int igx_mis_vertex(const igx_graph* g, unsigned int* priority, int* status, double* seconds)
{
  IGX_TRY(ensure_init());
  if (g == nullptr || status == nullptr) return IGX_ERR_ARG;
  fill_priorities(priority, g->nodes);
  igx_run_info info = {};
  IGX_TRY(run_to_host(IGX_MIS, 0, g, 0, status, sizeof(int) * (size_t)g->nodes, nullptr, &info));
  if (seconds) *seconds = info.seconds;
  return IGX_OK;
}

int igx_mis_edge(const igx_graph* g, const int* d_sp, unsigned int* priority, int* status, double* seconds)
{
  IGX_TRY(ensure_init());
  if (g == nullptr || status == nullptr) return IGX_ERR_ARG;
  fill_priorities(priority, g->nodes);
  igx_run_info info = {};
  IGX_TRY(run_to_host(IGX_MIS, 1, g, 0, status, sizeof(int) * (size_t)g->nodes, d_sp, &info));
  if (seconds) *seconds = info.seconds;
  return IGX_OK;
}

int igx_pr(const igx_graph* host_g, float* scores, const int* degree, double* seconds, int* iterations)
{
  IGX_TRY(ensure_init());
  if (host_g == nullptr || scores == nullptr) return IGX_ERR_ARG;
  (void)degree;  // degrees are recomputed on the device from nindex
  igx_graph d = {};
  IGX_TRY(igx_graph_upload(host_g, &d));
  igx_run_info info = {};
  const int rc = run_to_host(IGX_PR, 0, &d, 0, scores, sizeof(float) * (size_t)host_g->nodes, nullptr, &info);
  igx_graph_free_device(&d);
  if (seconds) *seconds = info.seconds;
  if (iterations) *iterations = info.iterations;
  return rc;
}

int igx_tc_vertex(unsigned long long* count, int nodes, const int* d_nindex, const int* d_nlist, double* seconds)
{
  IGX_TRY(ensure_init());
  if (count == nullptr || d_nindex == nullptr || nodes < 1) return IGX_ERR_ARG;
  int edges = 0;
  IGX_TRY(igx_memcpy_d2h(&edges, d_nindex + nodes, sizeof(int)));
  igx_graph g = {nodes, edges, const_cast<int*>(d_nindex), const_cast<int*>(d_nlist), nullptr};
  igx_run_info info = {};
  IGX_TRY(run_to_host(IGX_TC, 0, &g, 0, count, sizeof(unsigned long long), nullptr, &info));
  if (seconds) *seconds = info.seconds;
  return IGX_OK;
}

int igx_tc_edge(unsigned long long* count, int edges, const int* d_nindex, const int* d_nlist, const int* d_sp,
                double* seconds)
{
  IGX_TRY(ensure_init());
  if (count == nullptr || d_nindex == nullptr || d_sp == nullptr || edges < 0) return IGX_ERR_ARG;
  // the rows that own edges end at sp[edges-1]; nodes only bounds row ids here
  int last = 0;
  if (edges > 0) IGX_TRY(igx_memcpy_d2h(&last, d_sp + edges - 1, sizeof(int)));
  igx_graph g = {last + 1, edges, const_cast<int*>(d_nindex), const_cast<int*>(d_nlist), nullptr};
  igx_run_info info = {};
  IGX_TRY(run_to_host(IGX_TC, 1, &g, 0, count, sizeof(unsigned long long), d_sp, &info));
  if (seconds) *seconds = info.seconds;
  return IGX_OK;
}

}  // extern "C"
