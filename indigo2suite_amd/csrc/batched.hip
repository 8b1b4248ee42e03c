// Batched mode: BFS, CC and MIS over hundreds of uneven graphs packed into one
// arena, in one persistent launch, optionally replayed as one captured
// hipGraph.  (No reference counterpart: the reference runs one program per
// graph.)
//
// Arena layout (device memory):
//   node_off[G+1], edge_off[G+1]  int64 prefix offsets of the graphs
//   nindex  int32[sum(n_i + 1)]   graph-local row offsets; graph i's rows
//                                 start at node_off[i] + i
//   nlist   int32[sum(m_i)]       graph-local neighbour ids; graph i's list
//                                 starts at edge_off[i]
//   outputs int32[sum(n_i)]       graph i's values start at node_off[i]
//
// Execution: a persistent grid; every workgroup repeatedly takes the next
// (graph, algorithm) item from an atomic counter (graphs ordered largest
// first, so the long ones start early) and runs the whole algorithm on that
// graph with workgroup barriers between rounds.  No workgroup ever waits for
// another, so every wave reaches the loop exit once the counter passes the
// item count.
//
// Per-graph algorithms (results identical to the single-graph styles):
//   BFS from vertex 0: level-synchronous top-down; each level expands the
//     frontier queue with a workgroup-wide merge path over the degrees of a
//     tile of kTile queue entries (tile = workgroup size), CAS-claimed
//     discoveries pushed through an LDS counter.
//   CC: union-find (larger root hooked under smaller with a CAS), rows walked
//     with the same workgroup merge path, then full compression.
//   MIS: pull rounds over a compacted worklist of undecided vertices with the
//     unsigned hash priorities.
#include "igx_internal.hpp"

namespace igx {
namespace {

struct Arena {
  const long long* node_off;
  const long long* edge_off;
  const int* nindex;
  const int* nlist;
  int graphs;
};

__device__ __forceinline__ void graph_view(const Arena& A, int g, int& n, const int*& nidx, const int*& nl,
                                           long long& voff)
{
  voff = A.node_off[g];
  n = (int)(A.node_off[g + 1] - voff);
  nidx = A.nindex + voff + g;
  nl = A.nlist + A.edge_off[g];
}

// exclusive scan over the workgroup; returns this thread's prefix, total via *tot
template <int B>
__device__ __forceinline__ int wg_excl_scan(int v, int* lds_w, int* tot)
{
  constexpr int W = B / kWave;
  const int w = threadIdx.x / kWave;
  const int incl = wave_incl_scan(v);
  if (lane_id() == kWave - 1) lds_w[w] = incl;
  __syncthreads();
  int off = 0, t = 0;
  #pragma unroll
  for (int k = 0; k < W; k++) {
    const int x = lds_w[k];
    if (k < w) off += x;
    t += x;
  }
  __syncthreads();
  *tot = t;
  return off + incl - v;
}

// Walk all edges of `items` rows given by row(k) for k in [0, count), tile by
// tile: the workgroup scans the tile's degrees and every thread takes one
// edge per step.  visit(row, e) for each edge.  All threads must call.
template <int B, typename RowOf, typename F>
__device__ __forceinline__ void wg_expand(int count, const int* __restrict__ nidx, RowOf row_of, int* s_excl,
                                          int* s_row, int* s_beg, int* s_w, F visit)
{
  for (int base = 0; base < count; base += B) {
    const int k = base + (int)threadIdx.x;
    int v = 0, beg = 0, len = 0;
    if (k < count) {
      v = row_of(k);
      beg = nidx[v];
      len = nidx[v + 1] - beg;
    }
    int total = 0;
    const int ex = wg_excl_scan<B>(len, s_w, &total);
    s_excl[threadIdx.x] = ex;
    s_row[threadIdx.x] = v;
    s_beg[threadIdx.x] = beg;
    __syncthreads();
    for (int tb = 0; tb < total; tb += B) {
      const int t = tb + (int)threadIdx.x;
      if (t < total) {
        int lo = 0, hi = B;
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (s_excl[mid] <= t) lo = mid;
          else hi = mid;
        }
        visit(s_row[lo], s_beg[lo] + (t - s_excl[lo]));
      }
    }
    __syncthreads();
  }
}

__device__ __forceinline__ int next_graph(int* counter, int* s_g)
{
  if (threadIdx.x == 0) *s_g = atomicAdd(counter, 1);
  __syncthreads();
  const int k = *s_g;
  __syncthreads();
  return k;
}

// Workgroup-shared scratch of the per-graph routines (one set per workgroup;
// the routines run one after another in a workgroup, never nested).
template <int B>
struct WgLds {
  int excl[B], row[B], beg[B];
  int w[B / kWave];
  int next;
};

// ------------------------------------------------------------------ BFS ----
// BFS of one graph from vertex 0 by the whole workgroup (see the file head).
template <int B>
__device__ void bfs_graph(const Arena& A, int g, int* __restrict__ dist_all, int* __restrict__ queue_all, WgLds<B>& L)
{
  int n;
  const int* nidx;
  const int* nl;
  long long voff;
  graph_view(A, g, n, nidx, nl, voff);
  int* dist = dist_all + voff;
  int* q0 = queue_all + 2 * voff;
  int* q1 = q0 + n;
  for (int v = threadIdx.x; v < n; v += B) dist[v] = (v == 0) ? 0 : kInf;
  if (threadIdx.x == 0) q0[0] = 0;
  __syncthreads();
  int qsize = 1;
  int level = 0;
  while (qsize > 0) {
    if (threadIdx.x == 0) L.next = 0;
    __syncthreads();
    const int* qin = (level & 1) ? q1 : q0;
    int* qout = (level & 1) ? q0 : q1;
    const int nd = level + 1;
    wg_expand<B>(qsize, nidx, [&](int i) { return qin[i]; }, L.excl, L.row, L.beg, L.w, [&](int, int e) {
      const int u = nl[e];
      bool disc = false;
      if (dist[u] == kInf) {
        int expected = kInf;
        disc = __hip_atomic_compare_exchange_strong(&dist[u], &expected, nd, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      const uint64_t mask = ballot(disc);
      if (mask != 0) {
        const int leader = __ffsll((unsigned long long)mask) - 1;
        int base = 0;
        if (lane_id() == leader) base = atomicAdd(&L.next, __popcll(mask));
        base = __shfl(base, leader, kWave);
        if (disc) qout[base + mbcnt(mask)] = u;
      }
    });
    __syncthreads();
    qsize = L.next;
    level++;
    __syncthreads();
  }
}

// ------------------------------------------------------------------- CC ----
__device__ __forceinline__ int wg_find(int* __restrict__ parent, int x)
{
  while (true) {
    const int p = parent[x];
    if (p == x) return x;
    const int gp = parent[p];
    if (gp == p) return p;
    parent[x] = gp;
    x = gp;
  }
}

__device__ __forceinline__ void wg_unite(int* __restrict__ parent, int a, int b)
{
  int ra = wg_find(parent, a), rb = wg_find(parent, b);
  while (ra != rb) {
    const int hi = max(ra, rb), lo = min(ra, rb);
    int expected = hi;
    if (__hip_atomic_compare_exchange_strong(&parent[hi], &expected, lo, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_WORKGROUP)) {
      return;
    }
    ra = wg_find(parent, expected);
    rb = wg_find(parent, lo);
  }
}

template <int B>
__device__ void cc_graph(const Arena& A, int g, int* __restrict__ label_all, WgLds<B>& L)
{
  int n;
  const int* nidx;
  const int* nl;
  long long voff;
  graph_view(A, g, n, nidx, nl, voff);
  int* parent = label_all + voff;
  for (int v = threadIdx.x; v < n; v += B) {
    const int u0 = (nidx[v + 1] > nidx[v]) ? nl[nidx[v]] : v;
    parent[v] = (u0 < v) ? u0 : v;
  }
  __syncthreads();
  wg_expand<B>(n, nidx, [&](int i) { return i; }, L.excl, L.row, L.beg, L.w, [&](int v, int e) {
    const int u = nl[e];
    if (u < v) wg_unite(parent, v, u);
  });
  __syncthreads();
  // read-only walks; each thread writes only its own vertex (see uf_compress)
  for (int v = threadIdx.x; v < n; v += B) {
    int x = parent[v];
    int p = parent[x];
    while (p != x) {
      x = p;
      p = parent[x];
    }
    parent[v] = x;
  }
  __syncthreads();
}

// ------------------------------------------------------------------ MIS ----
__device__ __forceinline__ bool mis_outranks(uint32_t pu, int u, uint32_t pv, int v)
{
  return (pu > pv) || (pu == pv && u > v);
}

template <int B>
__device__ void mis_graph(const Arena& A, int g, int* __restrict__ status_all, int* __restrict__ wl_all, WgLds<B>& L)
{
  int n;
  const int* nidx;
  const int* nl;
  long long voff;
  graph_view(A, g, n, nidx, nl, voff);
  int* status = status_all + voff;
  int* w0 = wl_all + 2 * voff;
  int* w1 = w0 + n;
  for (int v = threadIdx.x; v < n; v += B) {
    status[v] = 0;
    w0[v] = v;
  }
  __syncthreads();
  int count = n;
  int round = 0;
  while (count > 0) {
    if (threadIdx.x == 0) L.next = 0;
    __syncthreads();
    const int* win = (round & 1) ? w1 : w0;
    int* wout = (round & 1) ? w0 : w1;
    for (int base = 0; base < count; base += B) {
      const int idx = base + (int)threadIdx.x;
      bool keep = false;
      int v = 0;
      if (idx < count) {
        v = win[idx];
        const uint32_t pv = mis_priority(v);
        int decision = 1;  // included unless blocked
        for (int i = nidx[v]; i < nidx[v + 1]; i++) {
          const int u = nl[i];
          if (u == v) continue;
          const int su = __hip_atomic_load(&status[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          if (su == 1) {
            decision = 2;
            break;
          }
          if (su != 2 && mis_outranks(mis_priority(u), u, pv, v)) decision = 0;
        }
        if (decision == 0) keep = true;
        else __hip_atomic_store(&status[v], decision, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      const uint64_t mask = ballot(keep);
      if (mask != 0) {
        const int leader = __ffsll((unsigned long long)mask) - 1;
        int b0 = 0;
        if (lane_id() == leader) b0 = atomicAdd(&L.next, __popcll(mask));
        b0 = __shfl(b0, leader, kWave);
        if (keep) wout[b0 + mbcnt(mask)] = v;
      }
    }
    __syncthreads();
    count = L.next;
    round++;
    __syncthreads();
  }
}

// One persistent launch for every (graph, algorithm) pair of the batch: a
// workgroup takes the next item from one counter, items ordered by graph
// (largest first) and then BFS, CC, MIS, so the three algorithms of the
// largest graphs run side by side on different workgroups from the start
// instead of one algorithm's largest graph being every launch's tail.  The
// per-vertex scratch of BFS (two queues) and of MIS (two worklists) are
// disjoint halves of `scratch`, so a graph's BFS and MIS may run at once.
template <int B>
__global__ void __launch_bounds__(B) batched_all(Arena A, const int* __restrict__ order, int* counter, int mask,
                                                  int* __restrict__ d_bfs, int* __restrict__ d_cc,
                                                  int* __restrict__ d_mis, int* __restrict__ scratch,
                                                  long long total_nodes)
{
  __shared__ WgLds<B> L;
  __shared__ int s_k;
  const int per = __popc((unsigned)mask);
  const long long items = (long long)A.graphs * per;
  while (true) {
    if (threadIdx.x == 0) s_k = atomicAdd(counter, 1);
    __syncthreads();
    const int k = s_k;
    __syncthreads();
    if (k >= items) break;
    const int g = order[k / per];
    int which = k % per;  // the which-th algorithm of the mask
    int algo = 0;
    for (int a = 0; a < 3; a++) {
      if ((mask >> a) & 1) {
        if (which == 0) {
          algo = a;
          break;
        }
        which--;
      }
    }
    if (algo == 0) bfs_graph<B>(A, g, d_bfs, scratch, L);
    else if (algo == 1) cc_graph<B>(A, g, d_cc, L);
    else mis_graph<B>(A, g, d_mis, scratch + 2 * total_nodes, L);
  }
}

__global__ void batched_reset(int* counters, int count)
{
  if ((int)threadIdx.x < count) counters[threadIdx.x] = 0;
}

// ------------------------------------------------------------ host side ----
// The batched mode owns its buffers (the three queue counters and the
// per-vertex scratch) instead of borrowing the library's shared workspaces: a
// captured graph keeps the addresses it was recorded with, and a shared slot
// may be freed and regrown by any other entry point between two replays.
// The buffers only ever change together with dropping the captured graph.
struct BatchState {
  igx_batch key = {};
  int block = 0;
  int mask = 0;
  int* outs[3] = {nullptr, nullptr, nullptr};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int* counters = nullptr;     // kCounterInts ints
  int* scratch = nullptr;      // scratch_ints ints
  size_t scratch_ints = 0;
};
BatchState g_state;
constexpr int kCounterInts = 64;

// scratch ints of a batch: BFS levels queue pair (2 per vertex), MIS
// worklist pair (2 per vertex), and a small tail
inline size_t scratch_ints_for(const igx_batch* b) { return 4 * (size_t)b->total_nodes + 64; }

void drop_graph()
{
  if (g_state.exec) hipGraphExecDestroy(g_state.exec);
  if (g_state.graph) hipGraphDestroy(g_state.graph);
  g_state.exec = nullptr;
  g_state.graph = nullptr;
}

template <int B>
void launch_all(const igx_batch* b, const Arena& A, int mask, int* const* outs, int* counters, int* scratch, int grid,
                hipStream_t s)
{
  hipLaunchKernelGGL(batched_reset, dim3(1), dim3(64), 0, s, counters, 1);
  hipLaunchKernelGGL((batched_all<B>), dim3(grid), dim3(B), 0, s, A, b->order, counters, mask, outs[0], outs[1],
                     outs[2], scratch, (long long)b->total_nodes);
}

// Make the batch buffers hold at least `ints` scratch ints.  Growing them
// drops the captured graph first (it refers to the old addresses).
int batch_buffers(size_t ints)
{
  if (g_state.counters == nullptr) {
    if (hipMalloc((void**)&g_state.counters, sizeof(int) * kCounterInts) != hipSuccess) {
      (void)hipGetLastError();
      g_state.counters = nullptr;
      return IGX_ERR_NOMEM;
    }
  }
  if (g_state.scratch_ints >= ints && g_state.scratch != nullptr) return IGX_OK;
  drop_graph();
  g_state.key = igx_batch{};
  if (g_state.scratch != nullptr) {
    IGX_HIP(hipStreamSynchronize(rt().stream));
    IGX_HIP(hipFree(g_state.scratch));
    g_state.scratch = nullptr;
    g_state.scratch_ints = 0;
  }
  if (hipMalloc((void**)&g_state.scratch, sizeof(int) * ints) != hipSuccess) {
    (void)hipGetLastError();
    g_state.scratch = nullptr;
    return IGX_ERR_NOMEM;
  }
  g_state.scratch_ints = ints;
  return IGX_OK;
}

// Enqueue the batch's kernels on stream s with the given counter and scratch
// buffers (kCounterInts and scratch_ints_for(b) ints).
int enqueue_with(const igx_batch* b, int mask, int* const* outs, int block, int* counters, int* scratch,
                 hipStream_t s)
{
  Runtime& R = rt();
  Arena A{b->node_off, b->edge_off, b->nindex, b->nlist, b->graphs};
  const int per_cu = R.max_threads_per_cu / block;
  int grid = R.cus * (per_cu > 0 ? per_cu : 1);
  const int items = b->graphs * __builtin_popcount((unsigned)mask);  // (graph, algorithm) pairs
  if (grid > items) grid = items;
  if (grid < 1) grid = 1;
  switch (block) {
    case 256: launch_all<256>(b, A, mask, outs, counters, scratch, grid, s); break;
    case 512: launch_all<512>(b, A, mask, outs, counters, scratch, grid, s); break;
    case 1024: launch_all<1024>(b, A, mask, outs, counters, scratch, grid, s); break;
    default: return IGX_ERR_ARG;
  }
  IGX_HIP(hipGetLastError());
  return IGX_OK;
}

int enqueue(const igx_batch* b, int mask, int* const* outs, int block, hipStream_t s)
{
  IGX_TRY(batch_buffers(scratch_ints_for(b)));
  return enqueue_with(b, mask, outs, block, g_state.counters, g_state.scratch, s);
}

bool same_batch(const igx_batch& a, const igx_batch& b)
{
  return a.graphs == b.graphs && a.total_nodes == b.total_nodes && a.total_edges == b.total_edges &&
         a.node_off == b.node_off && a.edge_off == b.edge_off && a.nindex == b.nindex && a.nlist == b.nlist &&
         a.order == b.order;
}

}  // namespace

// ---------------------------------------------- allocation-free variant ----
// igx_batch_run_noalloc takes its counters and per-vertex scratch from this
// file-scope device array instead of allocating them: batches of at most
// kNoAllocMaxNodes vertices in all.  Every GPU the library runs on holds its
// own copy (about 16 MiB) until the process ends.  Host code never uses the
// array's name (a host-side stand-in, not a device address): igx_init looks
// up its device address once per device with hipGetSymbolAddress and the
// entry point passes that address to the kernels.
constexpr long long kNoAllocMaxNodes = 1LL << 20;
constexpr size_t kNoAllocInts = kCounterInts + 4 * (size_t)kNoAllocMaxNodes + 64;
static __device__ __align__(256) int scratch[kNoAllocInts];

constexpr int kMaxDevices = 64;
int* g_scratch_ptr[kMaxDevices] = {};

int batch_release()
{
  drop_graph();
  g_state.key = igx_batch{};
  if (g_state.counters != nullptr || g_state.scratch != nullptr) IGX_HIP(hipStreamSynchronize(rt().stream));
  if (g_state.counters != nullptr) IGX_HIP(hipFree(g_state.counters));
  if (g_state.scratch != nullptr) IGX_HIP(hipFree(g_state.scratch));
  g_state.counters = nullptr;
  g_state.scratch = nullptr;
  g_state.scratch_ints = 0;
  return IGX_OK;
}

int batch_noalloc_setup(int device)
{
  if (device < 0 || device >= kMaxDevices) return IGX_ERR_ARG;
  if (g_scratch_ptr[device] != nullptr) return IGX_OK;
  void* ptr = nullptr;
  IGX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  g_scratch_ptr[device] = (int*)ptr;
  return IGX_OK;
}

}  // namespace igx

using namespace igx;

extern "C" {

int igx_batch_upload(int graphs, const igx_graph* host, igx_batch* out)
{
  IGX_TRY(ensure_init());
  if (graphs < 1 || host == nullptr || out == nullptr) return IGX_ERR_ARG;
  std::vector<long long> noff(graphs + 1, 0), eoff(graphs + 1, 0);
  for (int i = 0; i < graphs; i++) {
    if (host[i].nodes < 1 || host[i].edges < 0 || host[i].nindex == nullptr) return IGX_ERR_ARG;
    const int rc = igx_check_csr(&host[i]);
    if (rc != IGX_OK) return rc;
    noff[i + 1] = noff[i] + host[i].nodes;
    eoff[i + 1] = eoff[i] + host[i].edges;
  }
  // largest graphs first so the longest workgroups start early
  std::vector<int> order(graphs);
  for (int i = 0; i < graphs; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return (long long)host[a].edges + host[a].nodes > (long long)host[b].edges + host[b].nodes;
  });
  igx_batch b = {};
  b.graphs = graphs;
  b.total_nodes = noff[graphs];
  b.total_edges = eoff[graphs];
  void *p1 = nullptr, *p2 = nullptr, *p3 = nullptr, *p4 = nullptr, *p5 = nullptr;
  if (igx_alloc(&p1, sizeof(long long) * (graphs + 1)) != IGX_OK ||
      igx_alloc(&p2, sizeof(long long) * (graphs + 1)) != IGX_OK ||
      igx_alloc(&p3, sizeof(int) * (size_t)(b.total_nodes + graphs)) != IGX_OK ||
      igx_alloc(&p4, sizeof(int) * (size_t)(b.total_edges > 0 ? b.total_edges : 1)) != IGX_OK ||
      igx_alloc(&p5, sizeof(int) * (size_t)graphs) != IGX_OK) {
    igx_free(p1); igx_free(p2); igx_free(p3); igx_free(p4); igx_free(p5);
    return IGX_ERR_NOMEM;
  }
  IGX_TRY(igx_memcpy_h2d(p1, noff.data(), sizeof(long long) * (graphs + 1)));
  IGX_TRY(igx_memcpy_h2d(p2, eoff.data(), sizeof(long long) * (graphs + 1)));
  IGX_TRY(igx_memcpy_h2d(p5, order.data(), sizeof(int) * graphs));
  for (int i = 0; i < graphs; i++) {
    IGX_TRY(igx_memcpy_h2d((int*)p3 + noff[i] + i, host[i].nindex, sizeof(int) * ((size_t)host[i].nodes + 1)));
    if (host[i].edges > 0) {
      IGX_TRY(igx_memcpy_h2d((int*)p4 + eoff[i], host[i].nlist, sizeof(int) * (size_t)host[i].edges));
    }
  }
  b.node_off = (const long long*)p1;
  b.edge_off = (const long long*)p2;
  b.nindex = (const int*)p3;
  b.nlist = (const int*)p4;
  b.order = (const int*)p5;
  *out = b;
  return IGX_OK;
}

int igx_batch_free(igx_batch* b)
{
  if (b == nullptr) return IGX_ERR_ARG;
  if (same_batch(g_state.key, *b)) {
    drop_graph();
    g_state.key = igx_batch{};
  }
  igx_free((void*)b->node_off);
  igx_free((void*)b->edge_off);
  igx_free((void*)b->nindex);
  igx_free((void*)b->nlist);
  igx_free((void*)b->order);
  *b = igx_batch{};
  return IGX_OK;
}

int igx_batch_graph_view(const igx_batch* b, int i, igx_graph* view)
{
  if (b == nullptr || view == nullptr || i < 0 || i >= b->graphs) return IGX_ERR_ARG;
  long long no[2], eo[2];
  IGX_TRY(igx_memcpy_d2h(no, b->node_off + i, sizeof(no)));
  IGX_TRY(igx_memcpy_d2h(eo, b->edge_off + i, sizeof(eo)));
  view->nodes = (int)(no[1] - no[0]);
  view->edges = (int)(eo[1] - eo[0]);
  view->nindex = const_cast<int*>(b->nindex) + no[0] + i;
  view->nlist = const_cast<int*>(b->nlist) + eo[0];
  view->eweight = nullptr;
  return IGX_OK;
}

int igx_batch_run_noalloc(const igx_batch* b, int algo_mask, int* d_bfs, int* d_cc, int* d_mis, int block,
                          igx_run_info* info)
{
  // no ensure_init(): initialising would allocate
  Runtime& R = rt();
  if (!R.ready) {
    snprintf_last_error("igx_batch_run_noalloc: call igx_init first (the entry allocates nothing, not even at "
                        "initialisation)");
    return IGX_ERR_ARG;
  }
  if (b == nullptr || b->graphs < 1 || (algo_mask & ~7) != 0 || algo_mask == 0) return IGX_ERR_ARG;
  if (((algo_mask & 1) && !d_bfs) || ((algo_mask & 2) && !d_cc) || ((algo_mask & 4) && !d_mis)) return IGX_ERR_ARG;
  if (block != 256 && block != 512 && block != 1024) return IGX_ERR_ARG;
  if (b->total_nodes > kNoAllocMaxNodes) {
    snprintf_last_error("igx_batch_run_noalloc: the batch has %lld vertices; the allocation-free entry holds at "
                        "most %lld (use igx_batch_run)", b->total_nodes, kNoAllocMaxNodes);
    return IGX_ERR_LIMIT;
  }
  int* ptr = g_scratch_ptr[R.device];
  if (ptr == nullptr) {
    snprintf_last_error("igx_batch_run_noalloc: no scratch address for device %d", R.device);
    return IGX_ERR_ARG;
  }
  int* outs[3] = {d_bfs, d_cc, d_mis};
  igx_run_info local = {};
  IGX_TRY(timer_start());
  IGX_TRY(enqueue_with(b, algo_mask, outs, block, ptr, ptr + kCounterInts, R.stream));
  IGX_TRY(timer_stop(&local.seconds));
  local.launches = 2;  // counter reset and the persistent launch
  local.iterations = b->graphs;
  if (info) *info = local;
  return IGX_OK;
}

int igx_batch_run(const igx_batch* b, int algo_mask, int* d_bfs, int* d_cc, int* d_mis, int block, int use_graph,
                  igx_run_info* info)
{
  IGX_TRY(ensure_init());
  Runtime& R = rt();
  if (b == nullptr || b->graphs < 1 || (algo_mask & ~7) != 0 || algo_mask == 0) return IGX_ERR_ARG;
  if (((algo_mask & 1) && !d_bfs) || ((algo_mask & 2) && !d_cc) || ((algo_mask & 4) && !d_mis)) return IGX_ERR_ARG;
  if (block != 256 && block != 512 && block != 1024) return IGX_ERR_ARG;
  int* outs[3] = {d_bfs, d_cc, d_mis};
  igx_run_info local = {};
  if (!use_graph) {
    IGX_TRY(timer_start());
    IGX_TRY(enqueue(b, algo_mask, outs, block, R.stream));
    IGX_TRY(timer_stop(&local.seconds));
    local.launches = 2;  // counter reset and the persistent launch
  } else {
    const bool same = same_batch(g_state.key, *b) && g_state.block == block &&
                      g_state.mask == algo_mask && g_state.outs[0] == d_bfs && g_state.outs[1] == d_cc &&
                      g_state.outs[2] == d_mis && g_state.exec != nullptr;
    if (!same) {
      drop_graph();
      // buffers must exist before capture (no allocation inside it)
      IGX_TRY(batch_buffers(scratch_ints_for(b)));
      IGX_HIP(hipStreamSynchronize(R.stream));
      IGX_HIP(hipStreamBeginCapture(R.stream, hipStreamCaptureModeThreadLocal));
      const int rc = enqueue(b, algo_mask, outs, block, R.stream);
      hipGraph_t graph = nullptr;
      const hipError_t e = hipStreamEndCapture(R.stream, &graph);
      if (rc != IGX_OK) {
        if (graph) hipGraphDestroy(graph);
        return rc;
      }
      IGX_HIP(e);
      g_state.graph = graph;
      IGX_HIP(hipGraphInstantiate(&g_state.exec, graph, nullptr, nullptr, 0));
      g_state.key = *b;
      g_state.block = block;
      g_state.mask = algo_mask;
      for (int k = 0; k < 3; k++) g_state.outs[k] = outs[k];
    }
    IGX_TRY(timer_start());
    IGX_HIP(hipGraphLaunch(g_state.exec, R.stream));
    IGX_TRY(timer_stop(&local.seconds));
    local.launches = 1;
  }
  local.iterations = b->graphs;
  if (info) *info = local;
  return IGX_OK;
}

}  // extern "C"
