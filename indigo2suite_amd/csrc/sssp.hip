// Single-source shortest paths on gfx950: int32 distances with the graph's
// int32 edge weights (eweight), INT_MAX for unreachable vertices.  Every style
// relaxes to the same fixed point (the shortest-path distances).
//
// Reference family: codes/cuda/sssp-cuda (168 variants).
//   v_topo_push_*   [C-SVT] SSSP_CUDA_V_Topo_Push_..._Thread:56-120
//   v_topo_pull_*   V_Topo_Pull siblings
//   e_topo_push_*   E_Topo_Push siblings (one edge per item)
//   v_data_push_*   V_Data_Push NonDup ([C-SVDw]:62-87 for the warp form)
//   e_data_push_thread     [C-SED]:45-140 edge worklist, NonDup stamps on edges
//   e_data_push_rw_thread  [C-SEDrw]:63-87 ReadWrite: racy store, the edge
//                   itself is re-queued to repair a lost update
//   v_data_push_balanced   (no reference counterpart) vertex worklist of edge
//                   chunks expanded by wave-level merge path
//   v_topo_{push,pull}_{thread,wave}_determ  V_Topo_{Push,Pull}_Determ
//                   siblings: Jacobi rounds over two distance buffers
//   v_do_balanced   (no reference counterpart) direction-optimizing push/pull
//                   rounds on bitmap frontiers with weight-filtered pulls and a
//                   partial push round before the first pull over a hub
//                   frontier
#include "igx_internal.hpp"
#include "igx_frontier.hpp"

namespace igx {
namespace {

enum Gran { kThread = 0, kWaveG = 1, kBlockG = 2 };

__global__ void __launch_bounds__(kBlock) sssp_init(int src, int* __restrict__ dist, int n)
{
  const int v = blockIdx.x * kBlock + threadIdx.x;
  if (v < n) dist[v] = (v == src) ? 0 : kInf;
}

__global__ void __launch_bounds__(kBlock) sssp_fill(int* __restrict__ p, int value, long long count)
{
  const long long stride = (long long)gridDim.x * kBlock;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < count; i += stride) p[i] = value;
}

__device__ __forceinline__ void sssp_raise_flag(Ctl* ctl, bool updated)
{
  const uint64_t any = ballot(updated);
  if (any != 0 && lane_id() == (__ffsll((unsigned long long)any) - 1)) store_relaxed(&ctl->flag, 1);
}

template <int G, int AK>
__global__ void __launch_bounds__(kBlock) sssp_vtopo_push(const int* __restrict__ nindex,
                                                           const int* __restrict__ nlist,
                                                           const int* __restrict__ ew, int n, int* __restrict__ dist,
                                                           Ctl* ctl)
{
  if (round_done(ctl)) return;
  bool updated = false;
  if constexpr (G == kThread) {
    const int stride = gridDim.x * kBlock;
    for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
      const int s = dist[v];
      if (s != kInf) {
        for (int i = nindex[v]; i < nindex[v + 1]; i++) {
          const int nd = s + ew[i];
          const int dst = nlist[i];
          if (dist[dst] > nd && atomic_min_k<AK>(&dist[dst], nd) > nd) updated = true;
        }
      }
    }
  } else if constexpr (G == kWaveG) {
    const int waves = gridDim.x * kWavesPerBlock;
    const int lane = lane_id();
    for (int v = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; v < n; v += waves) {
      const int s = dist[v];
      if (s != kInf) {
        for (int i = nindex[v] + lane; i < nindex[v + 1]; i += kWave) {
          const int nd = s + ew[i];
          const int dst = nlist[i];
          if (dist[dst] > nd && atomic_min_k<AK>(&dist[dst], nd) > nd) updated = true;
        }
      }
    }
  } else {
    for (int v = blockIdx.x; v < n; v += gridDim.x) {
      const int s = dist[v];
      if (s != kInf) {
        for (int i = nindex[v] + (int)threadIdx.x; i < nindex[v + 1]; i += kBlock) {
          const int nd = s + ew[i];
          const int dst = nlist[i];
          if (dist[dst] > nd && atomic_min_k<AK>(&dist[dst], nd) > nd) updated = true;
        }
      }
    }
  }
  sssp_raise_flag(ctl, updated);
}

__global__ void __launch_bounds__(kBlock) sssp_vtopo_pull(const int* __restrict__ nindex,
                                                           const int* __restrict__ nlist,
                                                           const int* __restrict__ ew, int n, int* __restrict__ dist,
                                                           Ctl* ctl)
{
  if (round_done(ctl)) return;
  bool updated = false;
  const int stride = gridDim.x * kBlock;
  for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
    const int d0 = dist[v];
    int d = d0;
    for (int i = nindex[v]; i < nindex[v + 1]; i++) {
      const int s = load_relaxed(&dist[nlist[i]]);
      if (s != kInf && s + ew[i] < d) d = s + ew[i];
    }
    if (d < d0) {
      store_relaxed(&dist[v], d);
      updated = true;
    }
  }
  sssp_raise_flag(ctl, updated);
}

__global__ void __launch_bounds__(kBlock) sssp_etopo_push(const int* __restrict__ nlist,
                                                           const int* __restrict__ ew, const int* __restrict__ sp,
                                                           int m, int* __restrict__ dist, Ctl* ctl)
{
  if (round_done(ctl)) return;
  bool updated = false;
  const int stride = gridDim.x * kBlock;
  for (int e = blockIdx.x * kBlock + threadIdx.x; e < m; e += stride) {
    const int s = dist[sp[e]];
    if (s != kInf) {
      const int nd = s + ew[e];
      const int dst = nlist[e];
      if (dist[dst] > nd && atomic_min_relaxed(&dist[dst], nd) > nd) updated = true;
    }
  }
  sssp_raise_flag(ctl, updated);
}

// ------------------------------------------------ deterministic (Jacobi) ----
// SSSP_CUDA_V_Topo_Push_Determ_IntType_ReadModifyWrite_NonPersist_Atomic_
// Thread_NoBoundsBug2_NoBoundsBug1.cu sssp():57-80 and the V_Topo_Pull_Determ
// sibling (..._ReadWrite_NonPersist_Atomic_Thread_... sssp():57-79), Warp
// forms alike.  A round reads only `cur` and writes only `nxt`, and the two
// trade places every round, so the distances after round r are the r-hop
// Bellman-Ford values whatever the schedule, and the round count is fixed by
// the graph and the source.
//
// Push: nxt enters a round holding the values of two rounds ago, never below
// cur's (every entry only decreases), so a reached vertex first lowers its own
// nxt entry to its cur value and then relaxes its edges into nxt.  A
// relaxation counts as progress only if it goes below the destination's cur
// value; otherwise it merely caught nxt up.  The last round therefore changes
// no value of cur, and leaves nxt equal to cur: both buffers hold the result.
// Pull: every vertex writes min(cur[v], min over edges of cur[u] + w) to nxt,
// so again both buffers are equal after the last round.
template <int G>
__global__ void __launch_bounds__(kBlock) sssp_determ_push(const int* __restrict__ nindex,
                                                            const int* __restrict__ nlist,
                                                            const int* __restrict__ ew, int n,
                                                            const int* __restrict__ cur, int* __restrict__ nxt,
                                                            Ctl* ctl)
{
  if (round_done(ctl)) return;
  bool updated = false;
  auto relax = [&](int i, int s) {
    const int nd = s + ew[i];
    const int dst = nlist[i];
    if (load_relaxed(&nxt[dst]) > nd && atomic_min_relaxed(&nxt[dst], nd) > nd && nd < cur[dst]) updated = true;
  };
  if constexpr (G == kThread) {
    const int stride = gridDim.x * kBlock;
    for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
      const int s = cur[v];
      if (s != kInf) {
        if (load_relaxed(&nxt[v]) > s) atomic_min_relaxed(&nxt[v], s);
        for (int i = nindex[v]; i < nindex[v + 1]; i++) relax(i, s);
      }
    }
  } else {
    const int waves = gridDim.x * kWavesPerBlock;
    const int lane = lane_id();
    for (int v = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; v < n; v += waves) {
      const int s = cur[v];
      if (s != kInf) {
        if (lane == 0 && load_relaxed(&nxt[v]) > s) atomic_min_relaxed(&nxt[v], s);
        for (int i = nindex[v] + lane; i < nindex[v + 1]; i += kWave) relax(i, s);
      }
    }
  }
  sssp_raise_flag(ctl, updated);
}

template <int G>
__global__ void __launch_bounds__(kBlock) sssp_determ_pull(const int* __restrict__ nindex,
                                                            const int* __restrict__ nlist,
                                                            const int* __restrict__ ew, int n,
                                                            const int* __restrict__ cur, int* __restrict__ nxt,
                                                            Ctl* ctl)
{
  if (round_done(ctl)) return;
  bool updated = false;
  if constexpr (G == kThread) {
    const int stride = gridDim.x * kBlock;
    for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
      const int d0 = cur[v];
      int d = d0;
      for (int i = nindex[v]; i < nindex[v + 1]; i++) {
        const int s = cur[nlist[i]];
        if (s != kInf && s + ew[i] < d) d = s + ew[i];
      }
      nxt[v] = d;
      updated |= d < d0;
    }
  } else {
    const int waves = gridDim.x * kWavesPerBlock;
    const int lane = lane_id();
    for (int v = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; v < n; v += waves) {
      const int d0 = cur[v];
      long long d = d0;
      for (int i = nindex[v] + lane; i < nindex[v + 1]; i += kWave) {
        const int s = cur[nlist[i]];
        if (s != kInf && s + ew[i] < d) d = s + ew[i];
      }
      d = wave_min_ll(d);
      if (lane == 0) {
        nxt[v] = (int)d;
        updated |= d < d0;
      }
    }
  }
  sssp_raise_flag(ctl, updated);
}

// ------------------------------------------------------- vertex, data ----
__global__ void sssp_vdata_seed(int src, int* __restrict__ wl, Ctl* ctl)
{
  if (threadIdx.x == 0) {
    wl[0] = src;
    ctl->wl_count[0] = 1;
    ctl->wl_count[1] = 0;
  }
}

template <int G>
__global__ void __launch_bounds__(kBlock) sssp_vdata_push(const int* __restrict__ nindex,
                                                           const int* __restrict__ nlist,
                                                           const int* __restrict__ ew, int* __restrict__ dist,
                                                           int* __restrict__ wl0, int* __restrict__ wl1,
                                                           int* __restrict__ stamp, Ctl* ctl)
{
  if (round_done(ctl)) return;
  const int iter = ctl->iter;
  const int p = iter & 1;
  const int count = ctl->wl_count[p];
  const int* wl_in = p ? wl1 : wl0;
  int* wl_out = p ? wl0 : wl1;
  int* out_count = &ctl->wl_count[p ^ 1];
  const int tag = iter + 1;
  if constexpr (G == kThread) {
    const int stride = gridDim.x * kBlock;
    for (int idx = blockIdx.x * kBlock + threadIdx.x; idx < count; idx += stride) {
      const int v = wl_in[idx];
      const int s = dist[v];
      for (int i = nindex[v]; i < nindex[v + 1]; i++) {
        const int nd = s + ew[i];
        const int dst = nlist[i];
        bool push = false;
        if (dist[dst] > nd && atomic_min_relaxed(&dist[dst], nd) > nd) {
          push = atomic_max_relaxed(&stamp[dst], tag) != tag;
        }
        const int slot = wave_push_slot(out_count, push);
        if (push) wl_out[slot] = dst;
      }
    }
  } else {
    const int waves = gridDim.x * kWavesPerBlock;
    const int lane = lane_id();
    for (int idx = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; idx < count; idx += waves) {
      const int v = wl_in[idx];
      const int s = dist[v];
      const int beg = nindex[v], end = nindex[v + 1];
      for (int base = beg; base < end; base += kWave) {
        const int i = base + lane;
        if (i < end) {
          const int nd = s + ew[i];
          const int dst = nlist[i];
          bool push = false;
          if (dist[dst] > nd && atomic_min_relaxed(&dist[dst], nd) > nd) {
            push = atomic_max_relaxed(&stamp[dst], tag) != tag;
          }
          const int slot = wave_push_slot(out_count, push);
          if (push) wl_out[slot] = dst;
        }
      }
    }
  }
}

// --------------------------------------------------------- edge, data ----
__global__ void __launch_bounds__(kBlock) sssp_edata_seed(int src, const int* __restrict__ nindex,
                                                           int* __restrict__ wl, Ctl* ctl)
{
  const int beg = nindex[src], end = nindex[src + 1];
  for (int i = beg + (int)threadIdx.x; i < end; i += kBlock) wl[i - beg] = i;
  if (threadIdx.x == 0) {
    ctl->wl_count[0] = end - beg;
    ctl->wl_count[1] = 0;
  }
}

// One thread per worklist edge.  An edge of dst is queued for the next round
// when dst improves; every edge of dst is queued together, so the NonDup test
// is kept per source vertex (vstamp[dst] == tag: dst's edges are queued this
// round), which admits exactly the edges the per-edge stamps admit.  The
// queueing itself is wave-cooperative (wave_push_edge_ranges): a hub's edge
// ids are written 64 per step instead of one by one by the improving thread.
// RW: the racy load/store may lose an update to a concurrent store, so the
// edge itself is queued again (per-edge stamp estamp).
template <bool RW>
__global__ void __launch_bounds__(kBlock) sssp_edata_push(const int* __restrict__ nindex,
                                                           const int* __restrict__ nlist,
                                                           const int* __restrict__ ew, const int* __restrict__ sp,
                                                           int* __restrict__ dist, int* __restrict__ wl0,
                                                           int* __restrict__ wl1, int* __restrict__ vstamp,
                                                           int* __restrict__ estamp, Ctl* ctl)
{
  __shared__ int s_excl[kWavesPerBlock][kWave];
  __shared__ int s_beg[kWavesPerBlock][kWave];
  if (round_done(ctl)) return;
  const int iter = ctl->iter;
  const int p = iter & 1;
  const int count = ctl->wl_count[p];
  const int* wl_in = p ? wl1 : wl0;
  int* wl_out = p ? wl0 : wl1;
  int* out_count = &ctl->wl_count[p ^ 1];
  const int tag = iter + 1;
  const int w = threadIdx.x / kWave;
  const int stride = gridDim.x * kBlock;
  for (int base = blockIdx.x * kBlock; base < count; base += stride) {  // uniform per wave
    const int idx = base + (int)threadIdx.x;
    int rbeg = 0, rlen = 0;
    if (idx < count) {
      const int e = wl_in[idx];
      const int s = load_relaxed(&dist[sp[e]]);
      if (s != kInf) {
        const int nd = s + ew[e];
        const int dst = nlist[e];
        bool improved;
        if (RW) {
          improved = load_relaxed(&dist[dst]) > nd;
          if (improved) {
            store_relaxed(&dist[dst], nd);
            if (atomic_max_relaxed(&estamp[e], tag) != tag) {
              const int slot = atomicAdd(out_count, 1);
              wl_out[slot] = e;
            }
          }
        } else {
          improved = dist[dst] > nd && atomic_min_relaxed(&dist[dst], nd) > nd;
        }
        if (improved && atomic_max_relaxed(&vstamp[dst], tag) != tag) {
          rbeg = nindex[dst];
          rlen = nindex[dst + 1] - rbeg;
        }
      }
    }
    wave_push_edge_ranges(wl_out, out_count, rbeg, rlen, s_excl[w], s_beg[w]);
  }
}

// --------------------------------------------- balanced vertex worklist ----
__global__ void __launch_bounds__(kBlock) sssp_bal_seed(int src, const int* __restrict__ nindex,
                                                         Entry* __restrict__ q0, Ctl* ctl)
{
  const int beg = nindex[src], end = nindex[src + 1];
  const int c = chunks_of(end - beg);
  for (int k = threadIdx.x; k < c; k += kBlock) {
    const int b = beg + k * kChunk;
    q0[k] = Entry{src, b, (b + kChunk < end) ? b + kChunk : end, 0};
  }
  if (threadIdx.x == 0) {
    ctl->wl_count[0] = c;
    ctl->wl_count[1] = 0;
  }
}

__global__ void __launch_bounds__(kBlock) sssp_bal_round(const int* __restrict__ nindex,
                                                          const int* __restrict__ nlist,
                                                          const int* __restrict__ ew, int* __restrict__ dist,
                                                          Entry* __restrict__ q0, Entry* __restrict__ q1,
                                                          int* __restrict__ stamp, Ctl* ctl)
{
  __shared__ int s_excl[kWavesPerBlock][kWave];
  __shared__ Entry s_ent[kWavesPerBlock][kWave];
  if (round_done(ctl)) return;
  const int iter = ctl->iter;
  const int p = iter & 1;
  const Entry* q_in = p ? q1 : q0;
  Entry* q_out = p ? q0 : q1;
  int* out_count = &ctl->wl_count[p ^ 1];
  const int qsize = ctl->wl_count[p];
  const int tag = iter + 1;
  const int w = threadIdx.x / kWave;
  const int gw = blockIdx.x * kWavesPerBlock + w;
  const int nwaves = gridDim.x * kWavesPerBlock;
  wave_expand(q_in, qsize, gw, nwaves, s_excl[w], s_ent[w], [&](const Entry& en, int e) {
    const int s = dist[en.v];
    const int nd = s + ew[e];
    const int u = nlist[e];
    bool push = false;
    int ub = 0, ue = 0;
    if (dist[u] > nd && atomic_min_relaxed(&dist[u], nd) > nd) {
      push = atomic_max_relaxed(&stamp[u], tag) != tag;
      if (push) {
        ub = nindex[u];
        ue = nindex[u + 1];
      }
    }
    push_vertex_chunks(q_out, out_count, push && ue > ub, u, ub, ue);
  });
}

// ------------------------------------------ direction-optimizing rounds ----
// The frontier of a round is the set of vertices whose distance dropped in
// the round before; it is always held by a bitmap (three rotate:
// bm[iter % 3] current, bm[(iter + 1) % 3] next, bm[(iter + 2) % 3] cleared
// for the round after), and a push round whose frontier has at most
// kQueueEdgesS edges also lists the next frontier's chunks in a segmented
// queue.  The first setter of a vertex's bit (atomicOr's old value) is the
// one that counts and queues it, so no stamps are needed.
//
// Push round: relax the frontier's edges, read from the queue when the round
// before wrote one, else from the bitmap (64-vertex tiles of rows with at most
// kChunk edges, one wave per tile, plus the precomputed chunks of the longer
// rows that are in the frontier).
// Pull round: every row takes the minimum of dist[u] + w over the neighbours u
// in the frontier (the only ones that can lower it; the others were taken
// into account when they last changed) -- runs of rows of at most kChunk
// edges streamed by one wave per 64-row tile, the longer rows by their chunks
// with a global atomicMin.  Rows already final are skipped: with weights >= 0
// every value a round writes is at least the smallest distance of the
// frontier it starts from (`settled`, the minimum the previous round lowered
// anything to), so a row with dist <= settled cannot improve, now or later.
// The same bound filters single edges before their neighbour's distance is
// gathered: a candidate through an edge of weight w is at least settled + w
// when the neighbour is in the frontier or was lowered in this round, and a
// neighbour that is neither offers nothing new, so an edge with
// settled + w >= dist[v] (dist[v] as read at the row's start; it only
// decreases) cannot lower its row.  With uniform weights and distances that
// soon fall well below the largest weight, this skips most gathers (a
// synchronous model of R-MAT 20 from the hub: 3.6 m gathers in its pull
// rounds without the filter, 0.8 m with it); ids and weights are streamed
// together, so the filter adds no dependent load.
// The next round pulls when the next frontier's edges exceed m / kAlphaS.  A
// round that lowers nothing ends the run.
//
// Device work counters (algorithmic bytes of the round kernel): edges scanned
// (4 B id + 4 B weight each), edges relaxed, i.e. neighbour distances
// gathered (4 B each), rows visited (8 B offsets + 4 B distance).
#ifndef IGX_SSSP_ALPHA
#define IGX_SSSP_ALPHA 8
#endif
constexpr int kAlphaS = IGX_SSSP_ALPHA;
#ifndef IGX_SSSP_SEED_DIV
#define IGX_SSSP_SEED_DIV 32
#endif
constexpr int kSeedDiv = IGX_SSSP_SEED_DIV;  // 0: no partial push round
#ifndef IGX_SSSP_HUB_RATIO
#define IGX_SSSP_HUB_RATIO 4.0
#endif
constexpr double kHubRatio = IGX_SSSP_HUB_RATIO;
constexpr long long kQueueEdgesS = 1 << 16;
constexpr long long kEdgesPerBlock = 2048;  // frontier edges per workgroup of a push round from a queue
// A pull round whose frontier holds more than this percentage of the edges
// reads every neighbour's distance instead of testing its frontier bit first:
// a neighbour outside the frontier only offers a candidate its row has
// already seen (the row took it when that neighbour last changed), so the
// result is the same, and the test is a dependent random access that mostly
// passes.
#ifndef IGX_SSSP_ALLFRONT_PCT
#define IGX_SSSP_ALLFRONT_PCT 25
#endif
constexpr long long kAllFrontPct = IGX_SSSP_ALLFRONT_PCT;
constexpr int kSlice = 512;  // edges per streamed slice of a pull run

enum { kStatFrontierEdges = 0, kStatScanned = 1, kStatRelaxed = 2, kStatRows = 3 };

__global__ void __launch_bounds__(kBlock) sssp_do_init(int src, const int* __restrict__ nindex, Entry* __restrict__ q0,
                                                        unsigned* __restrict__ bm, int words, Ctl* ctl)
{
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < 2 * words; i += stride) {
    bm[i] = (i == (src >> 5)) ? (1u << (src & 31)) : 0u;
  }
  if (blockIdx.x != 0) return;
  const int beg = nindex[src], end = nindex[src + 1];
  const int c = chunks_of(end - beg);
  for (int k = threadIdx.x; k < c; k += kBlock) {
    const int b = beg + k * kChunk;
    q0[k] = Entry{src, b, (b + kChunk < end) ? b + kChunk : end, 0};
  }
  if (threadIdx.x == 0) {
    ctl_groups(ctl)[0].qcount[0] = c;
    ctl->mode = 0;
    ctl->in_bits = 0;
    ctl->edges_next = end - beg;  // edges of the current frontier
  }
}

__device__ __forceinline__ bool bit_of(const unsigned* __restrict__ bits, int v)
{
  return (bits[v >> 5] >> (v & 31)) & 1u;
}

__global__ void __launch_bounds__(kBlock) sssp_do_round(const int* __restrict__ nindex,
                                                         const int* __restrict__ nlist, const int* __restrict__ ew,
                                                         int n, long long m, int* __restrict__ dist, SegQ q0,
                                                         SegQ q1, const Entry* __restrict__ bigq,
                                                         const int* __restrict__ bigq_count,
                                                         unsigned* __restrict__ bm, int words, int grab, Ctl* ctl)
{
  __shared__ int s_excl[kWavesPerBlock][kWave];
  __shared__ int s_acc[kWavesPerBlock][kWave];
  __shared__ int s_rows[kWavesPerBlock][kWave];
  __shared__ unsigned long long s_head[kWavesPerBlock][kSlice / kWave];
  __shared__ Entry s_ent[kWavesPerBlock][kWave];
  __shared__ long long s_red[kWavesPerBlock * kStats];
  __shared__ double s_redd[2][kWavesPerBlock];
  if (round_done(ctl)) return;
  const int iter = ctl->iter;
  const int mode = ctl->mode;
  const int in_b = ctl->in_bits;
  // a partial push round (push_cut != 0) always writes its frontier as a bitmap
  const bool out_q = (mode == 0) && (ctl->edges_next <= kQueueEdgesS) && ctl->push_cut == 0;
  // (a round that lowered nothing leaves LLONG_MAX; no distance exceeds INT_MAX)
  const long long settled = min(ctl->aux_ll[1], (long long)INT_MAX);
  const bool all_front = (mode == 1) && (ctl->edges_next * 100 > m * kAllFrontPct || ctl->force_all);
  const int p = iter & 1;
  const SegQ q_in = p ? q1 : q0;
  const SegQ q_out = p ? q0 : q1;
  Entry* q_mine = segq_mine(q_out);
  int* out_count = segq_counter(ctl, p ^ 1, my_group());
  const unsigned* cur = bm + (size_t)(iter % 3) * words;
  unsigned* next = bm + (size_t)((iter + 1) % 3) * words;
  unsigned* stale = bm + (size_t)((iter + 2) % 3) * words;
  const int w = threadIdx.x / kWave;
  const int lane = lane_id();
  const int gw = blockIdx.x * kWavesPerBlock + w;
  // a push round from a small queue runs on the first `active` workgroups
  // only (set by the round before); the others leave at once
  const int active = (ctl->active > 0) ? ctl->active : (int)gridDim.x;
  if ((int)blockIdx.x >= active) return;
  const int nwaves = active * kWavesPerBlock;
  {
    const int stride = active * kBlock;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < words; i += stride) stale[i] = 0u;
  }
  long long front_e = 0, scanned = 0, relaxed = 0, rows = 0;
  long long low = LLONG_MAX;  // smallest distance this round lowered something to
  double lsum = 0.0, lcnt = 0.0;  // sum and count of the values this round lowered distances to
  // first marker of u in the next frontier: counts its edges, returns its row
  auto mark = [&](int u, int& ub, int& ue) -> bool {
    const unsigned bit = 1u << (u & 31);
    if ((atomicOr(&next[u >> 5], bit) & bit) != 0) return false;
    ub = nindex[u];
    ue = nindex[u + 1];
    front_e += ue - ub;
    return true;
  };
  const int cut = (mode == 0) ? ctl->push_cut : 0;
  // A frontier vertex at or above the cut of a partial push round is not
  // pushed: it stays in the frontier.  The pull after a partial push round
  // reads every neighbour (force_all), so neither the pending vertices nor
  // the ones the round lowers are marked in a bitmap; only the smallest
  // distance of the next frontier (pending ones included) is kept.
  auto keep_pending = [&](int v) -> bool {
    if (cut == 0) return false;
    const int dv = dist[v];
    if (dv < cut) return false;
    low = min(low, (long long)dv);
    front_e += 1;
    return true;
  };
  if (mode == 0) {
    auto relax = [&](const Entry& en, int e) {
      scanned += 1;
      relaxed += 1;
      const int nd = dist[en.v] + ew[e];
      const int u = nlist[e];
      bool first = false;
      int ub = 0, ue = 0;
      if (cut != 0) {
        // partial push round: the atomic's old value is not needed, so it
        // does not return one and the wave does not wait on it -- most
        // targets start at infinity and the hubs among them take many
        // candidates at once.  No bitmap mark (see keep_pending); a lowered
        // vertex counts 1 in the round's frontier statistic, which only has
        // to be nonzero.  The test before the atomic reads the distance from
        // L2 (an agent-scope load): a copy in this CU's L1 would not see
        // other CUs' lowered values, and a hub target would take an atomic
        // for every candidate.
        if (load_relaxed(&dist[u]) > nd) {
          atomic_min_relaxed(&dist[u], nd);
          low = min(low, (long long)nd);
          front_e += 1;
        }
        return;
      }
      if (dist[u] > nd && atomic_min_relaxed(&dist[u], nd) > nd) {
        low = min(low, (long long)nd);
        lsum += nd;
        lcnt += 1.0;
        first = mark(u, ub, ue);
      }
      if (out_q) push_vertex_chunks(q_mine, out_count, first && ue > ub, u, ub, ue);
    };
    if (!in_b) {
      const SegPrefix pre = segq_prefix(ctl, p);
      if (blockIdx.x == 0 && threadIdx.x == 0) rows = pre.pre[kGroups];
      wave_expand_with(pre.pre[kGroups], gw, nwaves, s_excl[w], s_ent[w],
                       [&](int i) {
                         Entry en = segq_at(q_in, pre, i);
                         if (en.end > en.beg && keep_pending(en.v)) en.end = en.beg;
                         return en;
                       },
                       relax);
    } else {
      for_each_bitmap_tile(cur, n, gw, nwaves, [&](int v0, bool in) {
        const int v = v0 + lane;
        Entry ent = Entry{0, 0, 0, 0};
        if (in) {
          const int b = nindex[v], e = nindex[v + 1];
          rows += 1;
          if (!keep_pending(v) && e - b <= kChunk) ent = Entry{v, b, e, 0};
        }
        const int len = ent.end - ent.beg;
        const int incl = wave_incl_scan(len);
        const int total = wave_bcast(incl, kWave - 1);
        if (total == 0) return;  // uniform
        __builtin_amdgcn_wave_barrier();
        s_excl[w][lane] = incl - len;
        s_ent[w][lane] = ent;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        for (int tb = 0; tb < total; tb += kWave) {
          const int t = tb + lane;
          if (t < total) {
            int lo = 0, hi = kWave;
            while (hi - lo > 1) {
              const int mid = (lo + hi) >> 1;
              if (s_excl[w][mid] <= t) lo = mid;
              else hi = mid;
            }
            const Entry& en = s_ent[w][lo];
            relax(en, en.beg + (t - s_excl[w][lo]));
          }
        }
        __builtin_amdgcn_wave_barrier();
      });
      // chunks dealt out round-robin: the live ones are the runs of chunks
      // of the frontier's long rows, which would otherwise fall to a few waves
      wave_expand_strided(*bigq_count, gw, nwaves, s_excl[w], s_ent[w],
                          [&](int i) {
                            Entry en = bigq[i];
                            // rows of pending vertices were kept by the tile walk above
                            if (!bit_of(cur, en.v) || (cut != 0 && dist[en.v] >= cut)) en.end = en.beg;
                            return en;
                          },
                          relax);
    }
  } else {
    // rows longer than kChunk (most of the edges of a power-law graph): one
    // wave per chunk, two chunks per step, their 2 x kChunk edges loaded
    // together, a wave minimum per chunk, and one atomicMin on each row's
    // distance
    const int nbig = *bigq_count;
    constexpr int kPer = kChunk / kWave;
    constexpr int kPair = 2;
    for (int c0 = gw; c0 < nbig; c0 += kPair * nwaves) {  // uniform per wave
      Entry en[kPair];
      #pragma unroll
      for (int h = 0; h < kPair; h++) {
        const int ci = c0 + h * nwaves;
        en[h] = (ci < nbig) ? bigq[ci] : Entry{0, 0, 0, 0};
      }
      int dv[kPair];
      #pragma unroll
      for (int h = 0; h < kPair; h++) {  // chunks of final rows are skipped
        dv[h] = (en[h].end > en[h].beg) ? dist[en[h].v] : 0;
        if (dv[h] <= settled) en[h].end = en[h].beg;
      }
      // ids and weights of the chunk's edges, loaded together
      int u[kPair][kPer], wt[kPair][kPer];
      #pragma unroll
      for (int h = 0; h < kPair; h++) {
        #pragma unroll
        for (int k = 0; k < kPer; k++) {
          const int e = en[h].beg + k * kWave + lane;
          u[h][k] = (e < en[h].end) ? nlist[e] : -1;
          wt[h][k] = (e < en[h].end) ? ew[e] : 0;
        }
      }
      // weight filter (see kernel comment): only an edge with
      // settled + w < dist[v] can lower its row
      bool hit[kPair][kPer];
      #pragma unroll
      for (int h = 0; h < kPair; h++) {
        #pragma unroll
        for (int k = 0; k < kPer; k++) hit[h][k] = u[h][k] >= 0 && settled + wt[h][k] < (long long)dv[h];
      }
      if (!all_front) {  // uniform
        // the frontier words of the edges that pass are loaded together
        unsigned bw[kPair][kPer];
        #pragma unroll
        for (int h = 0; h < kPair; h++) {
          #pragma unroll
          for (int k = 0; k < kPer; k++) bw[h][k] = hit[h][k] ? cur[u[h][k] >> 5] : 0u;
        }
        #pragma unroll
        for (int h = 0; h < kPair; h++) {
          #pragma unroll
          for (int k = 0; k < kPer; k++) hit[h][k] = hit[h][k] && ((bw[h][k] >> (u[h][k] & 31)) & 1u);
        }
      }
      // distances of the hits, all issued before the first use
      int du[kPair][kPer];
      #pragma unroll
      for (int h = 0; h < kPair; h++) {
        #pragma unroll
        for (int k = 0; k < kPer; k++) du[h][k] = hit[h][k] ? dist[u[h][k]] : kInf;
      }
      int cand[kPair];
      #pragma unroll
      for (int h = 0; h < kPair; h++) {
        cand[h] = kInf;
        #pragma unroll
        for (int k = 0; k < kPer; k++) {
          relaxed += hit[h][k] ? 1 : 0;
          if (du[h][k] != kInf) cand[h] = min(cand[h], du[h][k] + wt[h][k]);
        }
      }
      #pragma unroll
      for (int h = 0; h < kPair; h++) {
        #pragma unroll
        for (int d = kWave / 2; d > 0; d >>= 1) cand[h] = min(cand[h], __shfl_xor(cand[h], d, kWave));
      }
      if (lane == 0) {
        #pragma unroll
        for (int h = 0; h < kPair; h++) {
          scanned += en[h].end - en[h].beg;
          const int v = en[h].v;
          if (en[h].end > en[h].beg && cand[h] < dist[v] && atomic_min_relaxed(&dist[v], cand[h]) > cand[h]) {
            int ub, ue;
            low = min(low, (long long)cand[h]);
            lsum += cand[h];
            lcnt += 1.0;
            mark(v, ub, ue);
          }
        }
      }
    }
    // 64-row tiles.  Consecutive rows of at most kChunk edges form runs whose
    // edges are contiguous in nlist; a run is streamed in slices of kSlice
    // edges: the wave loads kSlice / 64 edges per lane with independent loads
    // (ids, frontier bits, distances, weights) and folds each candidate into
    // its row's minimum in LDS with an LDS atomicMin.  A position's row is
    // found without a search: the run's non-empty rows start at distinct
    // edges, which the owners mark in a head bitmap of the slice (one bit per
    // position, kSlice / 64 words); the number of heads up to a position,
    // plus the heads of the run's earlier slices, indexes the run's list of
    // non-empty rows in lane order.  Every lane does the same work per slice,
    // whatever the row lengths.
    // The tiles are claimed dynamically, `grab` at a time: the blocks of group g
    // take the tiles g, g + kGroups, g + 2 kGroups, ... in turn (tile weights
    // vary by more than 4x on R-MAT and drift with the vertex id; the long
    // rows' chunks above by little), so waves finish together.
    int* racc = s_acc[w];
    int* rdv = s_excl[w];  // each lane's row distance (the push paths' scan space, unused here)
    int* rlist = s_rows[w];
    unsigned long long* heads = s_head[w];
    const int ntiles = (n + kWave - 1) / kWave;
    const int grp = my_group();
    int* t_next = &ctl_groups(ctl)[grp].tiles;
    for (;;) {  // uniform per wave
      int k0 = 0;
      if (lane == 0) k0 = atomicAdd(t_next, grab);
      k0 = wave_bcast(k0, 0);
      if (grp + (long long)k0 * kGroups >= ntiles) break;
      for (int kc = k0; kc < k0 + grab; kc++) {
        const long long tl = grp + (long long)kc * kGroups;
        if (tl >= ntiles) break;  // uniform
        const int v0 = (int)tl * kWave;
        const int v = v0 + lane;
        int rb = 0, re = 0, dv = 0;
        if (v < n) {
          rb = nindex[v];
          re = nindex[v + 1];
          dv = dist[v];
          rows += 1;
        }
        // rows at most `settled` are final (see above) and skip their edges
        const bool small = (v < n) && (re - rb <= kChunk) && dv > settled;
        racc[lane] = kInf;
        rdv[lane] = dv;
        uint64_t todo = ballot(small);
        while (todo != 0) {  // uniform: one run of consecutive small rows
          const int first = __ffsll((unsigned long long)todo) - 1;
          const uint64_t from = todo >> first;
          const int cnt = (~from == 0) ? (kWave - first) : (__ffsll((unsigned long long)~from) - 1);
          const int last = first + cnt - 1;
          todo &= (cnt + first >= kWave) ? ((1ull << first) - 1) : ~(((1ull << cnt) - 1) << first);
          const int r0 = __shfl(rb, first, kWave), r1 = __shfl(re, last, kWave);
          const bool own = lane >= first && lane <= last && re > rb;
          const uint64_t om = ballot(own);
          __builtin_amdgcn_wave_barrier();
          if (own) rlist[mbcnt(om)] = lane;
          if (lane == 0) scanned += r1 - r0;
          int before = 0;  // heads of the run's earlier slices
          constexpr int kPerLane = kSlice / kWave;
          for (int c0 = r0; c0 < r1; c0 += kSlice) {
            const int c1 = (r1 - c0 < kSlice) ? r1 : c0 + kSlice;
            __builtin_amdgcn_wave_barrier();
            if (lane < kPerLane) heads[lane] = 0ull;
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            if (own && rb >= c0 && rb < c1) {
              const int q = rb - c0;
              atomicOr(&heads[q >> 6], 1ull << (q & 63));
            }
            // ids and weights of the slice, loaded together
            int u[kPerLane], wt[kPerLane];
            #pragma unroll
            for (int k = 0; k < kPerLane; k++) {
              const int e = c0 + k * kWave + lane;
              u[k] = (e < c1) ? nlist[e] : -1;
              wt[k] = (e < c1) ? ew[e] : 0;
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            // each position's row (lane of the tile), then the weight filter
            // against that row's distance
            const uint64_t upto = ~0ull >> (kWave - 1 - lane);  // positions 0..lane of a word
            int row[kPerLane];
            bool hit[kPerLane];
            int pk = before;
            #pragma unroll
            for (int k = 0; k < kPerLane; k++) {
              const uint64_t hw = heads[k];
              row[k] = (u[k] >= 0) ? rlist[pk + __popcll(hw & upto) - 1] : 0;
              hit[k] = u[k] >= 0 && settled + wt[k] < (long long)rdv[row[k]];
              pk += __popcll(hw);
            }
            before = pk;
            if (!all_front) {  // uniform
              unsigned bw[kPerLane];  // issued together, see the long rows above
              #pragma unroll
              for (int k = 0; k < kPerLane; k++) bw[k] = hit[k] ? cur[u[k] >> 5] : 0u;
              #pragma unroll
              for (int k = 0; k < kPerLane; k++) hit[k] = hit[k] && ((bw[k] >> (u[k] & 31)) & 1u);
            }
            int du[kPerLane];
            #pragma unroll
            for (int k = 0; k < kPerLane; k++) du[k] = hit[k] ? dist[u[k]] : kInf;
            #pragma unroll
            for (int k = 0; k < kPerLane; k++) {
              relaxed += hit[k] ? 1 : 0;
              if (du[k] != kInf) atomicMin(&racc[row[k]], du[k] + wt[k]);
            }
          }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        const int acc = racc[lane];
        const int len = re - rb;
        const bool improved = small && len > 0 && acc < dv;
        if (improved) {
          dist[v] = acc;
          front_e += len;
          low = min(low, (long long)acc);
          lsum += acc;
          lcnt += 1.0;
        }
        // this wave owns the tile's bits except those of rows longer than
        // kChunk, which other waves' chunk walks may set concurrently
        const uint64_t mask = ballot(improved);
        if (mask != 0) {
          if (lane == 0 && (unsigned)mask != 0u) atomicOr(&next[v0 >> 5], (unsigned)mask);
          if (lane == 1 && (unsigned)(mask >> 32) != 0u) atomicOr(&next[(v0 >> 5) + 1], (unsigned)(mask >> 32));
        }
        __builtin_amdgcn_wave_barrier();
      }
    }
  }
  round_partials(ctl, front_e, scanned, relaxed, rows, s_red, low);
  if (kSeedDiv > 0 && mode == 0 && !ctl->seeded) {  // uniform: only push rounds before the partial push need the mean
    round_partial_d(ctl, 0, lsum, s_redd[0]);
    round_partial_d(ctl, 1, lcnt, s_redd[1]);
  }
}

// Round decision (run by round_finalize after each sssp_do_round).
struct FinSsspDo {
  long long m;
  int nblocks;  // the round kernel's grid
  __device__ void operator()(Ctl* ctl, const RoundTotals& t) const
  {
    const long long ef = t.s[kStatFrontierEdges];
    const bool out_q = (ctl->mode == 0) && (ctl->edges_next <= kQueueEdgesS) && ctl->push_cut == 0;
    const int p = ctl->iter & 1;
    ctl->work_edges += t.s[kStatScanned];
    ctl->aux_ll[0] += t.s[kStatRelaxed];
    ctl->work_vertices += t.s[kStatRows];
    ctl->aux_ll[1] = t.s[kStatMin];  // the next frontier's smallest distance
    GroupSlot* gs = ctl_groups(ctl);
    for (int g = 0; g < kGroups; g++) {
      gs[g].qcount[p] = 0;
      gs[g].tiles = 0;
    }
    ctl->iter += 1;
    ctl->in_bits = out_q ? 0 : 1;
    ctl->edges_next = ef;
    // ef == 0: nothing was lowered, or only vertices without edges, from
    // which nothing can propagate
    const bool was_partial = ctl->push_cut != 0;
    ctl->push_cut = 0;
    ctl->force_all = 0;
    if (was_partial) {
      // the pull that follows a partial push reads every neighbour: the
      // round's frontier (pending and lowered vertices) is not in a bitmap
      ctl->mode = 1;
      ctl->force_all = 1;
    } else if (ef == 0) {
      ctl->done = 1;
    } else if (ef > m / kAlphaS) {
      // The first frontier large enough for a pull, reached by pushes: most
      // rows are still unreached, so a pull would gather a distance for
      // nearly every edge (no edge can be filtered against an infinite
      // row).  Instead, one push round from the closest part of the
      // frontier (distance below lowest + (mean - lowest) / kSeedDiv) gives
      // most rows a finite distance first; the rest of the frontier stays
      // in it for the pull that follows.
      const double cnt = t.d[1];
      const double lowest = (double)t.s[kStatMin];
      const double mean = (cnt > 0.0) ? t.d[0] / cnt : 0.0;
      // Only where the frontier is made of hubs (its edges per lowered
      // vertex exceed kHubRatio times the graph's average degree: a power-law
      // frontier next to the source), since there the pull would scan most
      // edges for few reached rows; on a frontier of ordinary vertices the
      // extra round costs more than it saves.
      const bool hubs = ef > kHubRatio * cnt * ((double)m / ctl->n);
      if (kSeedDiv > 0 && ctl->mode == 0 && !ctl->seeded && hubs && mean - lowest >= 2.0 * kSeedDiv) {
        ctl->push_cut = (int)(lowest + (mean - lowest) / kSeedDiv) + 1;
        ctl->seeded = 1;
        ctl->mode = 0;
      } else {
        ctl->mode = 1;
      }
    } else {
      ctl->mode = 0;
    }
    // a push from a queue (the frontier was small) needs few workgroups:
    // one per kEdgesPerBlock edges of the frontier, at least 16
    int act = 0;
    if (!ctl->done && ctl->mode == 0 && ctl->in_bits == 0) {
      const long long want = (ef + kEdgesPerBlock - 1) / kEdgesPerBlock;
      act = (int)min(max(want, 16LL), (long long)nblocks);
      if (act >= nblocks) act = 0;
    }
    ctl->active = act;
  }
};

// ------------------------------------------------------------ host side ----
template <typename K>
int run_sssp_rounds(const igx_graph* g, igx_run_info* info, long long max_rounds, K launch, bool worklist)
{
  int rounds = 0, launches = 0;
  const int rc = run_rounds(
      [&](long long) {
        launch();
        if (worklist) launch_fin_worklist();
        else launch_fin_flag();
      },
      max_rounds, &rounds, &launches);
  IGX_TRY(timer_stop(&info->seconds));
  info->iterations = rounds;
  info->launches = launches;
  (void)g;
  return rc;
}

// Bellman-Ford style rounds: at most n rounds change anything on a graph
// with non-negative weights; data-driven rounds can revisit, so allow more.
long long sssp_round_cap(const igx_graph* g) { return 4LL * g->nodes + 64; }

template <int G, int AK>
int style_vtopo_push(const igx_graph* g, int src, void* out, const void*, igx_run_info* info)
{
  Runtime& R = rt();
  int* dist = (int*)out;
  IGX_TRY(timer_start());
  IGX_TRY(ctl_reset(g->nodes));
  hipLaunchKernelGGL(sssp_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, src, dist, g->nodes);
  return run_sssp_rounds(g, info, sssp_round_cap(g), [&] {
    hipLaunchKernelGGL((sssp_vtopo_push<G, AK>), dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nindex,
                       g->nlist, g->eweight, g->nodes, dist, R.d_ctl);
  }, false);
}

int style_vtopo_pull(const igx_graph* g, int src, void* out, const void*, igx_run_info* info)
{
  Runtime& R = rt();
  int* dist = (int*)out;
  IGX_TRY(timer_start());
  IGX_TRY(ctl_reset(g->nodes));
  hipLaunchKernelGGL(sssp_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, src, dist, g->nodes);
  return run_sssp_rounds(g, info, sssp_round_cap(g), [&] {
    hipLaunchKernelGGL(sssp_vtopo_pull, dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nindex, g->nlist,
                       g->eweight, g->nodes, dist, R.d_ctl);
  }, false);
}

// Both buffers start equal and end equal (see sssp_determ_push), so the
// result is in `out` whichever buffer the last round wrote.
template <bool PULL, int G>
int style_determ(const igx_graph* g, int src, void* out, const void*, igx_run_info* info)
{
  Runtime& R = rt();
  int* d0 = (int*)out;
  void* p1 = nullptr;
  IGX_TRY(ws_get(WS_A, sizeof(int) * (size_t)g->nodes, &p1));
  int* d1 = (int*)p1;
  IGX_TRY(timer_start());
  IGX_TRY(ctl_reset(g->nodes));
  hipLaunchKernelGGL(sssp_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, src, d0, g->nodes);
  hipLaunchKernelGGL(sssp_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, src, d1, g->nodes);
  int rounds = 0, launches = 0;
  const int rc = run_rounds(
      [&](long long r) {
        const int* cur = (r & 1) ? d1 : d0;
        int* nxt = (r & 1) ? d0 : d1;
        if constexpr (PULL) {
          hipLaunchKernelGGL((sssp_determ_pull<G>), dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nindex,
                             g->nlist, g->eweight, g->nodes, cur, nxt, R.d_ctl);
        } else {
          hipLaunchKernelGGL((sssp_determ_push<G>), dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nindex,
                             g->nlist, g->eweight, g->nodes, cur, nxt, R.d_ctl);
        }
        launch_fin_flag();
      },
      sssp_round_cap(g), &rounds, &launches);
  IGX_TRY(timer_stop(&info->seconds));
  info->iterations = rounds;
  info->launches = launches;
  return rc;
}

int style_etopo_push(const igx_graph* g, int src, void* out, const void* aux, igx_run_info* info)
{
  Runtime& R = rt();
  int* dist = (int*)out;
  const int* sp = (const int*)aux;
  IGX_TRY(timer_start());
  IGX_TRY(ctl_reset(g->nodes));
  hipLaunchKernelGGL(sssp_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, src, dist, g->nodes);
  return run_sssp_rounds(g, info, sssp_round_cap(g), [&] {
    hipLaunchKernelGGL(sssp_etopo_push, dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nlist, g->eweight, sp,
                       g->edges, dist, R.d_ctl);
  }, false);
}

template <int G>
int style_vdata_push(const igx_graph* g, int src, void* out, const void*, igx_run_info* info)
{
  Runtime& R = rt();
  int* dist = (int*)out;
  const size_t items = (size_t)std::max(g->edges, g->nodes);
  void *p0 = nullptr, *p1 = nullptr, *p2 = nullptr;
  IGX_TRY(ws_get(WS_A, sizeof(int) * items, &p0));
  IGX_TRY(ws_get(WS_B, sizeof(int) * items, &p1));
  IGX_TRY(ws_get(WS_C, sizeof(int) * (size_t)g->nodes, &p2));
  IGX_TRY(timer_start());
  IGX_TRY(ctl_reset(g->nodes));
  hipLaunchKernelGGL(sssp_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, src, dist, g->nodes);
  hipLaunchKernelGGL(sssp_fill, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, (int*)p2, 0,
                     (long long)g->nodes);
  hipLaunchKernelGGL(sssp_vdata_seed, dim3(1), dim3(64), 0, R.stream, src, (int*)p0, R.d_ctl);
  return run_sssp_rounds(g, info, sssp_round_cap(g), [&] {
    hipLaunchKernelGGL((sssp_vdata_push<G>), dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nindex, g->nlist,
                       g->eweight, dist, (int*)p0, (int*)p1, (int*)p2, R.d_ctl);
  }, true);
}

template <bool RW>
int style_edata_push(const igx_graph* g, int src, void* out, const void* aux, igx_run_info* info)
{
  Runtime& R = rt();
  int* dist = (int*)out;
  const int* sp = (const int*)aux;
  // an edge is queued at most once per round (its source vertex's stamp), and
  // in RW at most once more as a re-queued edge (its own stamp)
  const size_t items = (size_t)std::max(g->edges, g->nodes) * (RW ? 2 : 1);
  const size_t estamps = RW ? (size_t)std::max(g->edges, 1) : 1;
  void *p0 = nullptr, *p1 = nullptr, *p2 = nullptr, *p3 = nullptr;
  IGX_TRY(ws_get(WS_A, sizeof(int) * items, &p0));
  IGX_TRY(ws_get(WS_B, sizeof(int) * items, &p1));
  IGX_TRY(ws_get(WS_C, sizeof(int) * (size_t)g->nodes, &p2));
  IGX_TRY(ws_get(WS_F, sizeof(int) * estamps, &p3));
  IGX_TRY(timer_start());
  IGX_TRY(ctl_reset(g->nodes));
  hipLaunchKernelGGL(sssp_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, src, dist, g->nodes);
  hipLaunchKernelGGL(sssp_fill, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, (int*)p2, 0,
                     (long long)g->nodes);
  hipLaunchKernelGGL(sssp_fill, dim3(blocks_for((long long)estamps, kBlock)), dim3(kBlock), 0, R.stream, (int*)p3, 0,
                     (long long)estamps);
  hipLaunchKernelGGL(sssp_edata_seed, dim3(1), dim3(kBlock), 0, R.stream, src, g->nindex, (int*)p0, R.d_ctl);
  return run_sssp_rounds(g, info, sssp_round_cap(g), [&] {
    hipLaunchKernelGGL((sssp_edata_push<RW>), dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nindex, g->nlist,
                       g->eweight, sp, dist, (int*)p0, (int*)p1, (int*)p2, (int*)p3, R.d_ctl);
  }, true);
}

int style_balanced(const igx_graph* g, int src, void* out, const void*, igx_run_info* info)
{
  Runtime& R = rt();
  int* dist = (int*)out;
  const size_t cap = (size_t)g->nodes + (size_t)g->edges / kChunk + 64;
  void *p0 = nullptr, *p1 = nullptr, *p2 = nullptr;
  IGX_TRY(ws_get(WS_D, sizeof(Entry) * cap, &p0));
  IGX_TRY(ws_get(WS_E, sizeof(Entry) * cap, &p1));
  IGX_TRY(ws_get(WS_C, sizeof(int) * (size_t)g->nodes, &p2));
  IGX_TRY(timer_start());
  IGX_TRY(ctl_reset(g->nodes));
  hipLaunchKernelGGL(sssp_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, src, dist, g->nodes);
  hipLaunchKernelGGL(sssp_fill, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, (int*)p2, 0,
                     (long long)g->nodes);
  hipLaunchKernelGGL(sssp_bal_seed, dim3(1), dim3(kBlock), 0, R.stream, src, g->nindex, (Entry*)p0, R.d_ctl);
  return run_sssp_rounds(g, info, sssp_round_cap(g), [&] {
    ktimer_begin("sssp_bal_round");
    hipLaunchKernelGGL(sssp_bal_round, dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nindex, g->nlist,
                       g->eweight, dist, (Entry*)p0, (Entry*)p1, (int*)p2, R.d_ctl);
    ktimer_end("sssp_bal_round");
  }, true);
}

int style_do(const igx_graph* g, int src, void* out, const void*, igx_run_info* info)
{
  Runtime& R = rt();
  int* dist = (int*)out;
  const long long cap = (long long)g->nodes + g->edges / kChunk + 64;  // chunks of all rows
  const int words = ((g->nodes + kWave - 1) / kWave) * 2;
  const int grab = tile_grab(g);
  void *p0 = nullptr, *p1 = nullptr, *pm = nullptr, *pb = nullptr, *pc = nullptr;
  IGX_TRY(ws_get(WS_D, sizeof(Entry) * (size_t)cap * kGroups, &p0));
  IGX_TRY(ws_get(WS_E, sizeof(Entry) * (size_t)cap * kGroups, &p1));
  IGX_TRY(ws_get(WS_F, sizeof(unsigned) * 3 * (size_t)words, &pm));
  // chunks of the rows longer than kChunk: at most 2 * edges / kChunk
  IGX_TRY(ws_get(WS_G, sizeof(Entry) * (2 * ((size_t)g->edges / kChunk) + 64), &pb));
  IGX_TRY(ws_get(WS_B, sizeof(int) * 64, &pc));
  const SegQ q0{(Entry*)p0, cap}, q1{(Entry*)p1, cap};
  IGX_TRY(timer_start());
  IGX_TRY(ctl_reset(g->nodes));
  hipLaunchKernelGGL(sssp_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, src, dist, g->nodes);
  IGX_TRY(big_rows(g, pb, (int*)pc));
  hipLaunchKernelGGL(sssp_do_init, dim3(blocks_for(2 * (long long)words, kBlock)), dim3(kBlock), 0, R.stream, src,
                     g->nindex, (Entry*)p0, (unsigned*)pm, words, R.d_ctl);
  int rounds = 0, launches = 0;
  const int rc = run_rounds(
      [&](long long) {
        ktimer_begin("sssp_do_round");
        hipLaunchKernelGGL(sssp_do_round, dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nindex, g->nlist,
                           g->eweight, g->nodes, (long long)g->edges, dist, q0, q1, (const Entry*)pb,
                           (const int*)pc, (unsigned*)pm, words, grab, R.d_ctl);
        ktimer_end("sssp_do_round");
        launch_finalize(R.persist_blocks, 0xF | (1 << kStatMin) | (3 << kStats),
                        FinSsspDo{(long long)g->edges, R.persist_blocks});
      },
      sssp_round_cap(g), &rounds, &launches, 8);
  IGX_TRY(timer_stop(&info->seconds));
  info->iterations = rounds;
  info->launches = launches;
  info->edges_processed = (unsigned long long)R.h_ctl->work_edges;
  info->edges_relaxed = (unsigned long long)R.h_ctl->aux_ll[0];
  info->vertices_processed = (unsigned long long)R.h_ctl->work_vertices;
  return rc;
}

const StyleEntry kStyles[] = {
    {"v_topo_push_thread", style_vtopo_push<kThread, kAtomicRelaxed>, 0},
    {"v_topo_push_wave", style_vtopo_push<kWaveG, kAtomicRelaxed>, 0},
    {"v_topo_push_block", style_vtopo_push<kBlockG, kAtomicRelaxed>, 0},
    {"v_topo_push_thread_cas", style_vtopo_push<kThread, kAtomicCas>, 0},
    {"v_topo_pull_thread", style_vtopo_pull, 0},
    {"e_topo_push_thread", style_etopo_push, 1},
    {"v_data_push_thread", style_vdata_push<kThread>, 0},
    {"v_data_push_wave", style_vdata_push<kWaveG>, 0},
    {"e_data_push_thread", style_edata_push<false>, 1},
    {"e_data_push_rw_thread", style_edata_push<true>, 1},
    {"v_data_push_balanced", style_balanced, 0},
    {"v_do_balanced", style_do, 0},
    {"v_topo_push_thread_determ", style_determ<false, kThread>, 0},
    {"v_topo_push_wave_determ", style_determ<false, kWaveG>, 0},
    {"v_topo_pull_thread_determ", style_determ<true, kThread>, 0},
    {"v_topo_pull_wave_determ", style_determ<true, kWaveG>, 0},
};

}  // namespace

const StyleEntry* sssp_styles(int* count)
{
  *count = (int)(sizeof(kStyles) / sizeof(kStyles[0]));
  return kStyles;
}

}  // namespace igx
