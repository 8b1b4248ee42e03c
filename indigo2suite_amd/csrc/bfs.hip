// BFS on gfx950: unweighted levels from `src`, INT_MAX for unreachable.
//
// Reference family: codes/cuda/bfs-cuda (180 variants).  The styles below
// cover its axes on wave64 hardware:
//   v_topo_push_*   [C-BVT] BFS_CUDA_V_Topo_Push_..._Thread:46-119 (warp/block
//                   siblings map to one wave64 / one 256-lane workgroup)
//   v_topo_pull_*   BFS_CUDA_V_Topo_Pull_..._Thread:56-78
//   *_determ        level-synchronous rule `if (s == iter)` (Determ siblings :65)
//   *_cas / *_seqcst  CAS-loop and seq_cst/system-scope atomicMin (the
//                   reference's Atomic vs CudaAtomic axis)
//   v_data_push_*   [C-BVD]:46-139 worklist with NonDup stamps
//   e_topo_push_*   [C-BET]:56-115 one edge per item
//   e_data_push_*   [C-BED]:46-140 edge worklist, stamps on edge ids
//   do_hybrid       direction-optimizing BFS (no reference counterpart):
//                   top-down from a chunked queue, bottom-up over unvisited
//                   vertices, switched per level on the device
// All styles produce identical level arrays (BFS levels are unique).
#include "igx_internal.hpp"
#include "igx_frontier.hpp"

namespace igx {
namespace {

enum Gran { kThread = 0, kWaveG = 1, kBlockG = 2 };

__global__ void __launch_bounds__(kBlock) bfs_init(int src, int* __restrict__ dist, int n)
{
  const int v = blockIdx.x * kBlock + threadIdx.x;
  if (v < n) dist[v] = (v == src) ? 0 : kInf;
}

__global__ void __launch_bounds__(kBlock) fill_int(int* __restrict__ p, int value, long long count)
{
  const long long stride = (long long)gridDim.x * kBlock;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < count; i += stride) p[i] = value;
}

__device__ __forceinline__ void raise_flag(Ctl* ctl, bool updated)
{
  const uint64_t any = ballot(updated);
  if (any != 0 && lane_id() == (__ffsll((unsigned long long)any) - 1)) store_relaxed(&ctl->flag, 1);
}

// --------------------------------------------------- vertex, topology ----
template <int G, int AK, bool DETERM, bool PERSIST>
__global__ void __launch_bounds__(kBlock) bfs_vtopo_push(const int* __restrict__ nindex,
                                                          const int* __restrict__ nlist, int n,
                                                          int* __restrict__ dist, Ctl* ctl)
{
  if (round_done(ctl)) return;
  const int iter = ctl->iter;
  bool updated = false;
  if constexpr (G == kThread) {
    const int stride = gridDim.x * kBlock;
    for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
      const int s = dist[v];
      if (s != kInf && (!DETERM || s == iter)) {
        const int nd = s + 1;
        for (int i = nindex[v]; i < nindex[v + 1]; i++) {
          if (atomic_min_k<AK>(&dist[nlist[i]], nd) > nd) updated = true;
        }
      }
    }
  } else if constexpr (G == kWaveG) {
    const int waves = gridDim.x * kWavesPerBlock;
    const int lane = lane_id();
    for (int v = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; v < n; v += waves) {
      const int s = dist[v];
      if (s != kInf && (!DETERM || s == iter)) {
        const int nd = s + 1;
        for (int i = nindex[v] + lane; i < nindex[v + 1]; i += kWave) {
          if (atomic_min_k<AK>(&dist[nlist[i]], nd) > nd) updated = true;
        }
      }
    }
  } else {
    for (int v = blockIdx.x; v < n; v += gridDim.x) {
      const int s = dist[v];
      if (s != kInf && (!DETERM || s == iter)) {
        const int nd = s + 1;
        for (int i = nindex[v] + (int)threadIdx.x; i < nindex[v + 1]; i += kBlock) {
          if (atomic_min_k<AK>(&dist[nlist[i]], nd) > nd) updated = true;
        }
      }
    }
  }
  raise_flag(ctl, updated);
}

template <int G>
__global__ void __launch_bounds__(kBlock) bfs_vtopo_pull(const int* __restrict__ nindex,
                                                          const int* __restrict__ nlist, int n,
                                                          int* __restrict__ dist, Ctl* ctl)
{
  if (round_done(ctl)) return;
  bool updated = false;
  if constexpr (G == kThread) {
    const int stride = gridDim.x * kBlock;
    for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
      const int d0 = dist[v];
      int d = d0;
      for (int i = nindex[v]; i < nindex[v + 1]; i++) {
        const int s = load_relaxed(&dist[nlist[i]]);
        if (s != kInf && s + 1 < d) d = s + 1;
      }
      if (d < d0) {
        store_relaxed(&dist[v], d);
        updated = true;
      }
    }
  } else {
    const int waves = gridDim.x * kWavesPerBlock;
    const int lane = lane_id();
    for (int v = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; v < n; v += waves) {
      const int d0 = dist[v];
      int d = d0;
      for (int i = nindex[v] + lane; i < nindex[v + 1]; i += kWave) {
        const int s = load_relaxed(&dist[nlist[i]]);
        if (s != kInf && s + 1 < d) d = s + 1;
      }
      #pragma unroll
      for (int k = kWave / 2; k > 0; k >>= 1) d = min(d, __shfl_xor(d, k, kWave));
      if (d < d0) {
        if (lane == 0) store_relaxed(&dist[v], d);
        updated = true;
      }
    }
  }
  raise_flag(ctl, updated);
}

// ------------------------------------------------------- edge, topology ----
template <int AK>
__global__ void __launch_bounds__(kBlock) bfs_etopo_push(const int* __restrict__ nlist, const int* __restrict__ sp,
                                                          int m, int* __restrict__ dist, Ctl* ctl)
{
  if (round_done(ctl)) return;
  bool updated = false;
  const int stride = gridDim.x * kBlock;
  for (int e = blockIdx.x * kBlock + threadIdx.x; e < m; e += stride) {
    const int s = dist[sp[e]];
    if (s != kInf) {
      const int nd = s + 1;
      const int dst = nlist[e];
      if (dist[dst] > nd && atomic_min_k<AK>(&dist[dst], nd) > nd) updated = true;
    }
  }
  raise_flag(ctl, updated);
}

// ------------------------------------------------------- vertex, data ----
__global__ void __launch_bounds__(kBlock) bfs_vdata_init(int src, int* __restrict__ wl, Ctl* ctl)
{
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    wl[0] = src;
    ctl->wl_count[0] = 1;
    ctl->wl_count[1] = 0;
  }
}

template <int G>
__global__ void __launch_bounds__(kBlock) bfs_vdata_push(const int* __restrict__ nindex,
                                                          const int* __restrict__ nlist, int* __restrict__ dist,
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
      if (s != kInf) {
        const int nd = s + 1;
        for (int i = nindex[v]; i < nindex[v + 1]; i++) {
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
  } else {
    const int waves = gridDim.x * kWavesPerBlock;
    const int lane = lane_id();
    for (int idx = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; idx < count; idx += waves) {
      const int v = wl_in[idx];
      const int s = dist[v];
      if (s != kInf) {
        const int nd = s + 1;
        const int beg = nindex[v], end = nindex[v + 1];
        for (int base = beg; base < end; base += kWave) {
          const int i = base + lane;
          bool push = false;
          if (i < end) {
            const int dst = nlist[i];
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
}

// --------------------------------------------------------- edge, data ----
__global__ void __launch_bounds__(kBlock) bfs_edata_init(int src, const int* __restrict__ nindex,
                                                          int* __restrict__ wl, Ctl* ctl)
{
  if (blockIdx.x != 0) return;
  const int beg = nindex[src], end = nindex[src + 1];
  for (int i = beg + (int)threadIdx.x; i < end; i += kBlock) wl[i - beg] = i;
  if (threadIdx.x == 0) {
    ctl->wl_count[0] = end - beg;
    ctl->wl_count[1] = 0;
  }
}

// One thread per worklist edge.  When dst improves, all of dst's edges are
// queued for the next round at once, so the NonDup test is kept per source
// vertex (vstamp[dst] == tag), admitting exactly the edges that per-edge
// stamps admit; the ids are written wave-cooperatively.
__global__ void __launch_bounds__(kBlock) bfs_edata_push(const int* __restrict__ nindex,
                                                          const int* __restrict__ nlist, const int* __restrict__ sp,
                                                          int* __restrict__ dist, int* __restrict__ wl0,
                                                          int* __restrict__ wl1, int* __restrict__ vstamp, Ctl* ctl)
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
      const int s = dist[sp[e]];
      if (s != kInf) {
        const int nd = s + 1;
        const int dst = nlist[e];
        if (dist[dst] > nd && atomic_min_relaxed(&dist[dst], nd) > nd &&
            atomic_max_relaxed(&vstamp[dst], tag) != tag) {
          rbeg = nindex[dst];
          rlen = nindex[dst + 1] - rbeg;
        }
      }
    }
    wave_push_edge_ranges(wl_out, out_count, rbeg, rlen, s_excl[w], s_beg[w]);
  }
}

// ------------------------------------------ direction-optimizing hybrid ----
// Per level one kernel: top-down expansion of the frontier, or bottom-up over
// every vertex.  The last block picks the next direction with Beamer's rule
// (top-down -> bottom-up when the next frontier's edges exceed the unvisited
// edges / kAlpha; back when the frontier has < n / kBeta vertices).
//
// Frontier representation.  Three bitmaps rotate: bm[level % 3] holds the
// current frontier, bm[(level + 1) % 3] receives the next one and
// bm[(level + 2) % 3] (the previous frontier) is cleared for the level after.
// The bitmaps are always complete.  A top-down level whose frontier has at
// most kQueueEdges edges also appends the discovered vertices' chunks to a
// segmented queue (few appends, so the queue is cheaper to read than the
// bitmap); larger top-down levels and all bottom-up levels produce the bitmap
// only -- no queue atomics at all.  A top-down level reads the queue when the
// previous level wrote one (ctl->in_bits == 0), else the bitmap: 64-vertex
// tiles whose rows of at most kChunk edges one wave expands, and the
// precomputed chunks of the longer rows, filtered by the bitmap.
constexpr int kAlpha = 14;
constexpr int kBeta = 24;
constexpr long long kQueueEdges = 1 << 16;

__global__ void __launch_bounds__(kBlock) bfs_do_init(int src, const int* __restrict__ nindex, int m,
                                                       Entry* __restrict__ q0, unsigned* __restrict__ bm, int words,
                                                       Ctl* ctl)
{
  // bitmaps 0 and 1 start empty (2 is cleared during level 0); src is in 0;
  // src's chunks go to queue segment 0
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
    ctl->level = 0;
    ctl->in_bits = 0;
    ctl->edges_next = end - beg;  // edges of the current frontier
    ctl->edges_unvisited = (long long)m - (end - beg);
  }
}

__device__ __forceinline__ bool in_bits(const unsigned* __restrict__ bits, int w)
{
  return (bits[w >> 5] >> (w & 31)) & 1u;
}

__global__ void __launch_bounds__(kBlock) bfs_do_level(const int* __restrict__ nindex,
                                                        const int* __restrict__ nlist, int n,
                                                        int* __restrict__ dist, SegQ q0, SegQ q1,
                                                        const Entry* __restrict__ bigq,
                                                        const int* __restrict__ bigq_count,
                                                        unsigned* __restrict__ bm, int words, Ctl* ctl)
{
  __shared__ int s_excl[kWavesPerBlock][kWave];
  __shared__ Entry s_ent[kWavesPerBlock][kWave];
  __shared__ long long s_red[kWavesPerBlock * kStats];
  if (round_done(ctl)) return;
  const int level = ctl->level;
  const int mode = ctl->mode;
  const int in_b = ctl->in_bits;
  const bool out_b = (mode == 1) || (ctl->edges_next > kQueueEdges);
  const int p = level & 1;
  const SegQ q_in = p ? q1 : q0;
  const SegQ q_out = p ? q0 : q1;
  Entry* q_mine = segq_mine(q_out);
  int* out_count = segq_counter(ctl, p ^ 1, my_group());
  const unsigned* cur = bm + (size_t)(level % 3) * words;
  unsigned* next = bm + (size_t)((level + 1) % 3) * words;
  unsigned* stale = bm + (size_t)((level + 2) % 3) * words;
  const int nd = level + 1;
  const int w = threadIdx.x / kWave;
  const int lane = lane_id();
  const int gw = blockIdx.x * kWavesPerBlock + w;
  const int nwaves = gridDim.x * kWavesPerBlock;
  {
    const int stride = gridDim.x * kBlock;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < words; i += stride) stale[i] = 0u;
  }
  long long found_v = 0, found_e = 0;
  if (mode == 0) {
    auto discover = [&](const Entry&, int e) {
      const int u = nlist[e];
      bool disc = false;
      int ub = 0, ue = 0;
      if (dist[u] == kInf) {
        int expected = kInf;
        disc = __hip_atomic_compare_exchange_strong(&dist[u], &expected, nd, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
        if (disc) {
          atomicOr(&next[u >> 5], 1u << (u & 31));
          ub = nindex[u];
          ue = nindex[u + 1];
          found_v += 1;
          found_e += ue - ub;
        }
      }
      if (!out_b) push_vertex_chunks(q_mine, out_count, disc && ue > ub, u, ub, ue);
    };
    if (!in_b) {
      const SegPrefix pre = segq_prefix(ctl, p);
      wave_expand_with(pre.pre[kGroups], gw, nwaves, s_excl[w], s_ent[w],
                       [&](int i) { return segq_at(q_in, pre, i); }, discover);
    } else {
      // rows of at most kChunk edges, the non-empty 64-vertex tiles of the bitmap
      for_each_bitmap_tile(cur, n, gw, nwaves, [&](int v0, bool in) {
        const int u = v0 + lane;
        Entry ent = Entry{0, 0, 0, 0};
        if (in) {
          const int b = nindex[u], e = nindex[u + 1];
          if (e - b <= kChunk) ent = Entry{u, b, e, 0};
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
            discover(en, en.beg + (t - s_excl[w][lo]));
          }
        }
        __builtin_amdgcn_wave_barrier();
      });
      // longer rows: their precomputed chunks, kept when the row is in the frontier
      wave_expand_with(*bigq_count, gw, nwaves, s_excl[w], s_ent[w],
                       [&](int i) {
                         Entry en = bigq[i];
                         if (!in_bits(cur, en.v)) en.end = en.beg;
                         return en;
                       },
                       discover);
    }
  } else {
    // kSub 64-vertex sub-tiles per step, their loads issued together (the
    // chain dist -> row -> first neighbours -> frontier bits is latency-bound)
    constexpr int kFirst = 4;
    constexpr int kSub = 4;
    for (int base = gw * kWave * kSub; base < n; base += nwaves * kWave * kSub) {
      bool unvisited[kSub], found[kSub];
      int ub[kSub], ue[kSub];
      #pragma unroll
      for (int s = 0; s < kSub; s++) {
        const int u = base + s * kWave + lane;
        unvisited[s] = (u < n) && dist[u] == kInf;
        found[s] = false;
        ub[s] = ue[s] = 0;
      }
      #pragma unroll
      for (int s = 0; s < kSub; s++) {
        const int u = base + s * kWave + lane;
        if (unvisited[s]) {
          ub[s] = nindex[u];
          ue[s] = nindex[u + 1];
        }
      }
      int nb[kSub][kFirst];
      #pragma unroll
      for (int s = 0; s < kSub; s++) {
        #pragma unroll
        for (int k = 0; k < kFirst; k++) nb[s][k] = (unvisited[s] && ub[s] + k < ue[s]) ? nlist[ub[s] + k] : -1;
      }
      #pragma unroll
      for (int s = 0; s < kSub; s++) {
        #pragma unroll
        for (int k = 0; k < kFirst; k++) found[s] = found[s] || (nb[s][k] >= 0 && in_bits(cur, nb[s][k]));
      }
      #pragma unroll
      for (int s = 0; s < kSub; s++) {
        if (base + s * kWave >= n) break;  // uniform
        uint64_t pending = ballot(unvisited[s] && !found[s] && ue[s] - ub[s] > kFirst);
        while (pending != 0) {
          const int L = __ffsll((unsigned long long)pending) - 1;
          pending &= pending - 1;
          const int s0 = __shfl(ub[s], L, kWave) + kFirst;
          const int e0 = __shfl(ue[s], L, kWave);
          bool hit = false;
          for (int i0 = s0; i0 < e0; i0 += kWave) {
            const int i = i0 + lane;
            const bool h = (i < e0) && in_bits(cur, nlist[i]);
            if (ballot(h) != 0) {
              hit = true;
              break;
            }
          }
          if (lane == L) found[s] = hit;
        }
        const int u = base + s * kWave + lane;
        if (found[s]) {
          dist[u] = nd;
          found_v += 1;
          found_e += ue[s] - ub[s];
        }
        const uint64_t mask = ballot(found[s]);
        const int wd = (base + s * kWave) >> 5;
        if (lane == 0) next[wd] = (unsigned)mask;
        if (lane == 1) next[wd + 1] = (unsigned)(mask >> 32);
      }
    }
  }
  round_partials(ctl, found_v, found_e, 0, 0, s_red);
}

// Level decision (run by round_finalize after each bfs_do_level): totals s[0]
// = vertices found, s[1] = their edges.
struct FinBfsDo {
  __device__ void operator()(Ctl* ctl, const RoundTotals& t) const
  {
    const long long nf = t.s[0];
    const long long mf = t.s[1];
    const int p = ctl->level & 1;
    const bool out_b = (ctl->mode == 1) || (ctl->edges_next > kQueueEdges);
    ctl->edges_unvisited -= mf;
    if (nf == 0) {
      ctl->done = 1;
    } else if (ctl->mode == 0) {
      if (mf > ctl->edges_unvisited / kAlpha) ctl->mode = 1;
    } else {
      if (nf < (long long)ctl->n / kBeta) ctl->mode = 0;
    }
    ctl->in_bits = out_b ? 1 : 0;
    ctl->edges_next = mf;
    GroupSlot* gs = ctl_groups(ctl);
    for (int g = 0; g < kGroups; g++) gs[g].qcount[p] = 0;
    ctl->level += 1;
    ctl->iter += 1;
  }
};

// ------------------------------------------------------------ host side ----
struct Bufs {
  int* wl0;
  int* wl1;
  int* stamp;
};

template <typename K>
int run_topo(const igx_graph* g, int src, int* d_dist, igx_run_info* info, int grid, long long max_rounds,
             K kernel_launch)
{
  Runtime& R = rt();
  IGX_TRY(timer_start());
  IGX_TRY(ctl_reset(g->nodes));
  hipLaunchKernelGGL(bfs_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, src, d_dist, g->nodes);
  int rounds = 0, launches = 0;
  const int rc = run_rounds(
      [&](long long) {
        kernel_launch(grid);
        launch_fin_flag();
      },
      max_rounds, &rounds, &launches);
  IGX_TRY(timer_stop(&info->seconds));
  info->iterations = rounds;
  info->launches = launches;
  return rc;
}

template <int G, int AK, bool DETERM, bool PERSIST>
int style_vtopo_push(const igx_graph* g, int src, void* out, const void*, igx_run_info* info)
{
  Runtime& R = rt();
  int* dist = (int*)out;
  int grid = R.persist_blocks;
  if (!PERSIST) {
    grid = (G == kThread) ? blocks_for(g->nodes, kBlock)
                          : (G == kWaveG ? blocks_for(g->nodes, kWavesPerBlock) : g->nodes);
  }
  return run_topo(g, src, dist, info, grid, (long long)g->nodes + 2,
                  [&](int gr) {
                    hipLaunchKernelGGL((bfs_vtopo_push<G, AK, DETERM, PERSIST>), dim3(gr), dim3(kBlock), 0, R.stream,
                                       g->nindex, g->nlist, g->nodes, dist, R.d_ctl);
                  });
}

template <int G>
int style_vtopo_pull(const igx_graph* g, int src, void* out, const void*, igx_run_info* info)
{
  Runtime& R = rt();
  int* dist = (int*)out;
  return run_topo(g, src, dist, info, R.persist_blocks, (long long)g->nodes + 2,
                  [&](int gr) {
                    hipLaunchKernelGGL((bfs_vtopo_pull<G>), dim3(gr), dim3(kBlock), 0, R.stream, g->nindex, g->nlist,
                                       g->nodes, dist, R.d_ctl);
                  });
}

template <int AK>
int style_etopo_push(const igx_graph* g, int src, void* out, const void* aux, igx_run_info* info)
{
  Runtime& R = rt();
  int* dist = (int*)out;
  const int* sp = (const int*)aux;
  return run_topo(g, src, dist, info, R.persist_blocks, (long long)g->nodes + 2,
                  [&](int gr) {
                    hipLaunchKernelGGL((bfs_etopo_push<AK>), dim3(gr), dim3(kBlock), 0, R.stream, g->nlist, sp,
                                       g->edges, dist, R.d_ctl);
                  });
}

int get_worklists(size_t items, size_t stamp_items, Bufs* b)
{
  void* p = nullptr;
  IGX_TRY(ws_get(WS_A, sizeof(int) * items, &p));
  b->wl0 = (int*)p;
  IGX_TRY(ws_get(WS_B, sizeof(int) * items, &p));
  b->wl1 = (int*)p;
  IGX_TRY(ws_get(WS_C, sizeof(int) * stamp_items, &p));
  b->stamp = (int*)p;
  return IGX_OK;
}

template <int G>
int style_vdata_push(const igx_graph* g, int src, void* out, const void*, igx_run_info* info)
{
  Runtime& R = rt();
  int* dist = (int*)out;
  const size_t items = (size_t)std::max(g->edges, g->nodes);
  Bufs b;
  IGX_TRY(get_worklists(items, (size_t)g->nodes, &b));
  IGX_TRY(timer_start());
  IGX_TRY(ctl_reset(g->nodes));
  hipLaunchKernelGGL(bfs_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, src, dist, g->nodes);
  hipLaunchKernelGGL(fill_int, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, b.stamp, 0,
                     (long long)g->nodes);
  hipLaunchKernelGGL(bfs_vdata_init, dim3(1), dim3(64), 0, R.stream, src, b.wl0, R.d_ctl);
  int rounds = 0, launches = 0;
  const int rc = run_rounds(
      [&](long long) {
        hipLaunchKernelGGL((bfs_vdata_push<G>), dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nindex,
                           g->nlist, dist, b.wl0, b.wl1, b.stamp, R.d_ctl);
        launch_fin_worklist();
      },
      (long long)g->nodes + 2, &rounds, &launches);
  IGX_TRY(timer_stop(&info->seconds));
  info->iterations = rounds;
  info->launches = launches;
  return rc;
}

int style_edata_push(const igx_graph* g, int src, void* out, const void* aux, igx_run_info* info)
{
  Runtime& R = rt();
  int* dist = (int*)out;
  const int* sp = (const int*)aux;
  const size_t items = (size_t)std::max(g->edges, g->nodes);
  Bufs b;
  IGX_TRY(get_worklists(items, (size_t)g->nodes, &b));
  IGX_TRY(timer_start());
  IGX_TRY(ctl_reset(g->nodes));
  hipLaunchKernelGGL(bfs_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, src, dist, g->nodes);
  hipLaunchKernelGGL(fill_int, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, b.stamp, 0,
                     (long long)g->nodes);
  hipLaunchKernelGGL(bfs_edata_init, dim3(1), dim3(kBlock), 0, R.stream, src, g->nindex, b.wl0, R.d_ctl);
  int rounds = 0, launches = 0;
  const int rc = run_rounds(
      [&](long long) {
        hipLaunchKernelGGL(bfs_edata_push, dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nindex, g->nlist,
                           sp, dist, b.wl0, b.wl1, b.stamp, R.d_ctl);
        launch_fin_worklist();
      },
      (long long)g->nodes + 2, &rounds, &launches);
  IGX_TRY(timer_stop(&info->seconds));
  info->iterations = rounds;
  info->launches = launches;
  return rc;
}

int style_do_hybrid(const igx_graph* g, int src, void* out, const void*, igx_run_info* info)
{
  Runtime& R = rt();
  int* dist = (int*)out;
  const long long cap = (long long)g->nodes + g->edges / kChunk + 64;  // chunks of all rows
  const int words = ((g->nodes + kWave - 1) / kWave) * 2;
  void *p0 = nullptr, *p1 = nullptr, *pb = nullptr, *pq = nullptr, *pc = nullptr;
  IGX_TRY(ws_get(WS_D, sizeof(Entry) * (size_t)cap * kGroups, &p0));
  IGX_TRY(ws_get(WS_E, sizeof(Entry) * (size_t)cap * kGroups, &p1));
  IGX_TRY(ws_get(WS_F, sizeof(unsigned) * 3 * (size_t)words, &pb));
  // chunks of the rows longer than kChunk: at most 2 * edges / kChunk
  IGX_TRY(ws_get(WS_G, sizeof(Entry) * (2 * (size_t)(g->edges / kChunk) + 64), &pq));
  IGX_TRY(ws_get(WS_H, sizeof(int) * 64, &pc));
  const SegQ q0{(Entry*)p0, cap}, q1{(Entry*)p1, cap};
  IGX_TRY(timer_start());
  IGX_TRY(ctl_reset(g->nodes));
  IGX_TRY(big_rows(g, pq, (int*)pc));
  hipLaunchKernelGGL(bfs_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, src, dist, g->nodes);
  hipLaunchKernelGGL(bfs_do_init, dim3(blocks_for(2 * (long long)words, kBlock)), dim3(kBlock), 0, R.stream, src,
                     g->nindex, g->edges, (Entry*)p0, (unsigned*)pb, words, R.d_ctl);
  int rounds = 0, launches = 0;
  const int rc = run_rounds(
      [&](long long) {
        ktimer_begin("bfs_do_level");
        hipLaunchKernelGGL(bfs_do_level, dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nindex, g->nlist,
                           g->nodes, dist, q0, q1, (const Entry*)pq, (const int*)pc, (unsigned*)pb, words, R.d_ctl);
        ktimer_end("bfs_do_level");
        launch_finalize(R.persist_blocks, 0x3, FinBfsDo{});
      },
      (long long)g->nodes + 2, &rounds, &launches, 8);
  IGX_TRY(timer_stop(&info->seconds));
  info->iterations = rounds;
  info->launches = launches;
  return rc;
}

const StyleEntry kStyles[] = {
    {"v_topo_push_thread", style_vtopo_push<kThread, kAtomicRelaxed, false, true>, 0},
    {"v_topo_push_wave", style_vtopo_push<kWaveG, kAtomicRelaxed, false, true>, 0},
    {"v_topo_push_block", style_vtopo_push<kBlockG, kAtomicRelaxed, false, true>, 0},
    {"v_topo_push_thread_nonpersist", style_vtopo_push<kThread, kAtomicRelaxed, false, false>, 0},
    {"v_topo_push_thread_determ", style_vtopo_push<kThread, kAtomicRelaxed, true, true>, 0},
    {"v_topo_push_thread_cas", style_vtopo_push<kThread, kAtomicCas, false, true>, 0},
    {"v_topo_push_thread_seqcst", style_vtopo_push<kThread, kAtomicSeqCst, false, true>, 0},
    {"v_topo_pull_thread", style_vtopo_pull<kThread>, 0},
    {"v_topo_pull_wave", style_vtopo_pull<kWaveG>, 0},
    {"v_data_push_thread", style_vdata_push<kThread>, 0},
    {"v_data_push_wave", style_vdata_push<kWaveG>, 0},
    {"e_topo_push_thread", style_etopo_push<kAtomicRelaxed>, 1},
    {"e_data_push_thread", style_edata_push, 1},
    {"do_hybrid", style_do_hybrid, 0},
};

}  // namespace

const StyleEntry* bfs_styles(int* count)
{
  *count = (int)(sizeof(kStyles) / sizeof(kStyles[0]));
  return kStyles;
}

}  // namespace igx
