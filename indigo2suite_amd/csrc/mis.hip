// Maximal independent set on gfx950.  Priorities are the reference's hash of
// (v + 712313887) compared as UNSIGNED with ties broken toward the larger id
// (codes/omp MIS init and mis_vertex_omp.h:127-132; the CUDA variants compare
// them as signed int and produce a different set).  With fixed priorities
// every style converges to the same set: the greedy MIS in descending
// (priority, id) order.  Output: int32 status, 1 = in the set, 2 = not.
// Self loops are ignored (the reference never terminates on them).
//
// Reference family: codes/cuda/mis-cuda (112 variants), codes/omp/mis-omp.
//   v_topo_push_*  [C-MVT] MIS_CUDA_V_Topo_Push_..._Thread:45-128 / [O-MIS]
//   v_topo_pull_*  V_Topo_Pull siblings (exclude self when a neighbour is in)
//   v_data_pull_*  V_Data_Pull siblings: worklist of undecided vertices
//   e_topo_thread  [C-MET]:57-127 / [O-MISe]:52-94 edge pass + vertex pass
//   e_data_thread  [C-MED]:45-213 worklist of edges, edge pass + vertex pass
//   e_data_thread_{copy,merge}determ  MIS_OMP_E_Data_..._{Copy,Merge}Determ:
//                  the same passes over a second status buffer, published
//                  after each pass
//   v_resume_pull  (no reference counterpart) pull over the outranking
//                  neighbours only, resuming where the last round stopped
#include "igx_internal.hpp"

namespace igx {
namespace {

constexpr int kUndecided = 0;
constexpr int kIncluded = 1;
constexpr int kExcluded = 2;

// u outranks v under the unsigned-priority order with id tie-break
__device__ __forceinline__ bool outranks(uint32_t pu, int u, uint32_t pv, int v)
{
  return (pu > pv) || (pu == pv && u > v);
}

__global__ void __launch_bounds__(kBlock) mis_init(int* __restrict__ status, int n)
{
  const int v = blockIdx.x * kBlock + threadIdx.x;
  if (v < n) status[v] = kUndecided;
}

__device__ __forceinline__ void mis_raise_flag(Ctl* ctl, bool pending)
{
  const uint64_t any = ballot(pending);
  if (any != 0 && lane_id() == (__ffsll((unsigned long long)any) - 1)) store_relaxed(&ctl->flag, 1);
}

// [C-MVT] mis():56-80 with unsigned priorities
template <bool WAVE>
__global__ void __launch_bounds__(kBlock) mis_vtopo_push(const int* __restrict__ nindex,
                                                          const int* __restrict__ nlist, int n,
                                                          int* __restrict__ status, Ctl* ctl)
{
  if (round_done(ctl)) return;
  bool pending = false;
  if constexpr (!WAVE) {
    const int stride = gridDim.x * kBlock;
    for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
      if (load_relaxed(&status[v]) != kUndecided) continue;
      const uint32_t pv = mis_priority(v);
      const int beg = nindex[v], end = nindex[v + 1];
      int i = beg;
      while (i < end) {
        const int u = nlist[i];
        if (u != v && load_relaxed(&status[u]) != kExcluded && outranks(mis_priority(u), u, pv, v)) break;
        i++;
      }
      if (i < end) {
        pending = true;
      } else {
        store_relaxed(&status[v], kIncluded);
        for (int j = beg; j < end; j++) {
          const int u = nlist[j];
          if (u != v) store_relaxed(&status[u], kExcluded);
        }
      }
    }
  } else {
    const int waves = gridDim.x * kWavesPerBlock;
    const int lane = lane_id();
    for (int v = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; v < n; v += waves) {
      if (load_relaxed(&status[v]) != kUndecided) continue;
      const uint32_t pv = mis_priority(v);
      const int beg = nindex[v], end = nindex[v + 1];
      bool blocked = false;
      for (int i = beg + lane; i < end; i += kWave) {
        const int u = nlist[i];
        if (u != v && load_relaxed(&status[u]) != kExcluded && outranks(mis_priority(u), u, pv, v)) blocked = true;
      }
      if (ballot(blocked) != 0) {
        if (lane == 0) pending = true;
      } else {
        if (lane == 0) store_relaxed(&status[v], kIncluded);
        for (int j = beg + lane; j < end; j += kWave) {
          const int u = nlist[j];
          if (u != v) store_relaxed(&status[u], kExcluded);
        }
      }
    }
  }
  mis_raise_flag(ctl, pending);
}

// pull: a vertex excludes itself when a neighbour is in the set and includes
// itself when every outranking neighbour is excluded
__device__ __forceinline__ int mis_pull_decide(const int* __restrict__ nindex, const int* __restrict__ nlist,
                                               const int* __restrict__ status, int v)
{
  const uint32_t pv = mis_priority(v);
  bool best = true;
  for (int i = nindex[v]; i < nindex[v + 1]; i++) {
    const int u = nlist[i];
    if (u == v) continue;
    const int su = load_relaxed(&status[u]);
    if (su == kIncluded) return kExcluded;
    if (su != kExcluded && outranks(mis_priority(u), u, pv, v)) best = false;
  }
  return best ? kIncluded : kUndecided;
}

__global__ void __launch_bounds__(kBlock) mis_vtopo_pull(const int* __restrict__ nindex,
                                                          const int* __restrict__ nlist, int n,
                                                          int* __restrict__ status, Ctl* ctl)
{
  if (round_done(ctl)) return;
  bool pending = false;
  const int stride = gridDim.x * kBlock;
  for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
    if (load_relaxed(&status[v]) != kUndecided) continue;
    const int d = mis_pull_decide(nindex, nlist, status, v);
    if (d == kUndecided) pending = true;
    else store_relaxed(&status[v], d);
  }
  mis_raise_flag(ctl, pending);
}

__global__ void __launch_bounds__(kBlock) mis_seed(int* __restrict__ wl, int n, Ctl* ctl)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) wl[i] = i;
  if (i == 0) {
    ctl->wl_count[0] = n;
    ctl->wl_count[1] = 0;
  }
}

__global__ void __launch_bounds__(kBlock) mis_vdata_pull(const int* __restrict__ nindex,
                                                          const int* __restrict__ nlist, int* __restrict__ status,
                                                          int* __restrict__ wl0, int* __restrict__ wl1, Ctl* ctl)
{
  if (round_done(ctl)) return;
  const int p = ctl->iter & 1;
  const int count = ctl->wl_count[p];
  const int* wl_in = p ? wl1 : wl0;
  int* wl_out = p ? wl0 : wl1;
  int* out_count = &ctl->wl_count[p ^ 1];
  const int stride = gridDim.x * kBlock;
  for (int base = blockIdx.x * kBlock; base < count; base += stride) {
    const int idx = base + threadIdx.x;
    bool keep = false;
    int v = 0;
    if (idx < count) {
      v = wl_in[idx];
      if (load_relaxed(&status[v]) == kUndecided) {
        const int d = mis_pull_decide(nindex, nlist, status, v);
        if (d == kUndecided) keep = true;
        else store_relaxed(&status[v], d);
      }
    }
    const int slot = wave_push_slot(out_count, keep);
    if (keep) wl_out[slot] = v;
  }
}

// [O-MISe]:52-77 edge pass: exclude the neighbours of included vertices and
// mark the lower-ranked endpoint of an undecided pair as lost
__global__ void __launch_bounds__(kBlock) mis_edge_pass(const int* __restrict__ nlist, const int* __restrict__ sp,
                                                         int m, int* __restrict__ status, int* __restrict__ lost,
                                                         Ctl* ctl)
{
  if (round_done(ctl)) return;
  const int stride = gridDim.x * kBlock;
  for (int e = blockIdx.x * kBlock + threadIdx.x; e < m; e += stride) {
    const int src = sp[e];
    const int dst = nlist[e];
    if (src == dst) continue;
    const int ss = load_relaxed(&status[src]);
    const int sd = load_relaxed(&status[dst]);
    if (ss == kIncluded) {
      store_relaxed(&status[dst], kExcluded);
    } else if (sd == kIncluded) {
      store_relaxed(&status[src], kExcluded);
    } else if (ss != kExcluded && sd != kExcluded) {
      const bool dst_wins = outranks(mis_priority(dst), dst, mis_priority(src), src);
      store_relaxed(&lost[dst_wins ? src : dst], 1);
    }
  }
}

// [O-MISe]:79-94 vertex pass: include every undecided vertex that did not lose
__global__ void __launch_bounds__(kBlock) mis_vertex_pass(int n, int* __restrict__ status, int* __restrict__ lost,
                                                           Ctl* ctl)
{
  if (round_done(ctl)) return;
  bool pending = false;
  const int stride = gridDim.x * kBlock;
  for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
    if (lost[v] == 0) {
      if (status[v] == kUndecided) status[v] = kIncluded;
    } else {
      pending = true;
      lost[v] = 0;
    }
  }
  mis_raise_flag(ctl, pending);
}

// ------------------------------------------------------ edge, data-driven ----
// [C-MED] init():45-62, mis():64-91, mis_vertex_pass():93-126,
// mis_last_pass():128-137: a worklist of edges, each
// undirected edge once (src < dst), shrinks as both endpoints get decided.
// Per round an edge pass excludes the other endpoint of an included vertex and
// marks the lower-ranked endpoint of an undecided pair as lost; a vertex pass
// over the same edges includes every undecided endpoint that did not lose and
// keeps the edges that still have an undecided endpoint.  A vertex left
// undecided at the end has no edge to anyone else and joins.
//
// `lost` holds the number of the round in which a vertex lost (rounds count
// from 1), so it never needs clearing; the reference clears a 0/1 flag array
// every round.  Statuses use one array, as in the topology-driven edge style:
// an edge pass sees every inclusion of earlier rounds (the kernel boundary
// publishes them) and, within the pass, a stale "undecided" only adds a lost
// mark, which delays a vertex by a round and never decides one wrongly.  The
// worklist never holds an edge twice, so the reference's per-edge time stamps
// that suppress duplicate pushes are not needed.
__global__ void __launch_bounds__(kBlock) mis_edata_seed(const int* __restrict__ nlist, const int* __restrict__ sp,
                                                          int m, int* __restrict__ wl, Ctl* ctl)
{
  const int stride = gridDim.x * kBlock;
  for (int base = blockIdx.x * kBlock; base < m; base += stride) {  // uniform per wave
    const int e = base + threadIdx.x;
    const bool keep = e < m && sp[e] < nlist[e];
    const int slot = wave_push_slot(&ctl->wl_count[0], keep);
    if (keep) wl[slot] = e;
  }
}

__global__ void __launch_bounds__(kBlock) mis_edata_edge(const int* __restrict__ nlist, const int* __restrict__ sp,
                                                          int* __restrict__ status, int* __restrict__ lost,
                                                          const int* __restrict__ wl0, const int* __restrict__ wl1,
                                                          Ctl* ctl)
{
  if (round_done(ctl)) return;
  const int p = ctl->iter & 1;
  const int count = ctl->wl_count[p];
  const int* wl_in = p ? wl1 : wl0;
  const int stamp = ctl->iter + 1;
  const int stride = gridDim.x * kBlock;
  for (int idx = blockIdx.x * kBlock + threadIdx.x; idx < count; idx += stride) {
    const int e = wl_in[idx];
    const int src = sp[e];
    const int dst = nlist[e];
    const int ss = status[src];
    const int sd = status[dst];
    if (ss == kIncluded) {
      status[dst] = kExcluded;
    } else if (sd == kIncluded) {
      status[src] = kExcluded;
    } else if (ss == kUndecided && sd == kUndecided) {
      const bool dst_wins = outranks(mis_priority(dst), dst, mis_priority(src), src);
      lost[dst_wins ? src : dst] = stamp;
    }
  }
}

__global__ void __launch_bounds__(kBlock) mis_edata_vertex(const int* __restrict__ nlist, const int* __restrict__ sp,
                                                            int* __restrict__ status, const int* __restrict__ lost,
                                                            int* __restrict__ wl0, int* __restrict__ wl1, Ctl* ctl)
{
  if (round_done(ctl)) return;
  const int p = ctl->iter & 1;
  const int count = ctl->wl_count[p];
  const int* wl_in = p ? wl1 : wl0;
  int* wl_out = p ? wl0 : wl1;
  int* out_count = &ctl->wl_count[p ^ 1];
  const int stamp = ctl->iter + 1;
  const int stride = gridDim.x * kBlock;
  for (int base = blockIdx.x * kBlock; base < count; base += stride) {  // uniform per wave
    const int idx = base + threadIdx.x;
    bool keep = false;
    int e = 0;
    if (idx < count) {
      e = wl_in[idx];
      const int src = sp[e];
      const int dst = nlist[e];
      const int ss = status[src];
      const int sd = status[dst];
      if (ss == kUndecided && lost[src] != stamp) status[src] = kIncluded;
      if (sd == kUndecided && lost[dst] != stamp) status[dst] = kIncluded;
      keep = ss == kUndecided || sd == kUndecided;
    }
    const int slot = wave_push_slot(out_count, keep);
    if (keep) wl_out[slot] = e;
  }
}

__global__ void __launch_bounds__(kBlock) mis_edata_last(int* __restrict__ status, int n)
{
  const int v = blockIdx.x * kBlock + threadIdx.x;
  if (v < n && status[v] == kUndecided) status[v] = kIncluded;
}

// Determ forms of the edge worklist style: MIS_OMP_E_Data_IntType_Atomic_
// CopyDeterm_Default_NoFieldBug.cpp OMPmis_edge() and its MergeDeterm sibling
// (include/mis_edge_omp.h:151-163 updateFromWorklist).  Each pass reads
// `status` only and writes its decisions into `status_n`; a publish kernel
// after the pass copies them to `status`: every vertex (Copy), or the
// endpoints of the round's worklist edges, the only vertices a pass can
// change (Merge).  No read in a pass sees a write of the same pass, so every
// round's decisions are fixed by the round before.  status_n equals status
// at the start of every pass.
__global__ void __launch_bounds__(kBlock) mis_determ_edge(const int* __restrict__ nlist, const int* __restrict__ sp,
                                                           const int* __restrict__ status,
                                                           int* __restrict__ status_n, int* __restrict__ lost,
                                                           const int* __restrict__ wl0, const int* __restrict__ wl1,
                                                           Ctl* ctl)
{
  if (round_done(ctl)) return;
  const int p = ctl->iter & 1;
  const int count = ctl->wl_count[p];
  const int* wl_in = p ? wl1 : wl0;
  const int stamp = ctl->iter + 1;
  const int stride = gridDim.x * kBlock;
  for (int idx = blockIdx.x * kBlock + threadIdx.x; idx < count; idx += stride) {
    const int e = wl_in[idx];
    const int src = sp[e];
    const int dst = nlist[e];
    const int ss = status[src];
    const int sd = status[dst];
    if (ss == kIncluded) {
      status_n[dst] = kExcluded;
    } else if (sd == kIncluded) {
      status_n[src] = kExcluded;
    } else if (ss == kUndecided && sd == kUndecided) {
      const bool dst_wins = outranks(mis_priority(dst), dst, mis_priority(src), src);
      lost[dst_wins ? src : dst] = stamp;
    }
  }
}

__global__ void __launch_bounds__(kBlock) mis_determ_vertex(const int* __restrict__ nlist, const int* __restrict__ sp,
                                                             const int* __restrict__ status,
                                                             int* __restrict__ status_n, const int* __restrict__ lost,
                                                             int* __restrict__ wl0, int* __restrict__ wl1, Ctl* ctl)
{
  if (round_done(ctl)) return;
  const int p = ctl->iter & 1;
  const int count = ctl->wl_count[p];
  const int* wl_in = p ? wl1 : wl0;
  int* wl_out = p ? wl0 : wl1;
  int* out_count = &ctl->wl_count[p ^ 1];
  const int stamp = ctl->iter + 1;
  const int stride = gridDim.x * kBlock;
  for (int base = blockIdx.x * kBlock; base < count; base += stride) {  // uniform per wave
    const int idx = base + threadIdx.x;
    bool keep = false;
    int e = 0;
    if (idx < count) {
      e = wl_in[idx];
      const int src = sp[e];
      const int dst = nlist[e];
      const int ss = status[src];
      const int sd = status[dst];
      if (ss == kUndecided && lost[src] != stamp) status_n[src] = kIncluded;
      if (sd == kUndecided && lost[dst] != stamp) status_n[dst] = kIncluded;
      keep = ss == kUndecided || sd == kUndecided;
    }
    const int slot = wave_push_slot(out_count, keep);
    if (keep) wl_out[slot] = e;
  }
}

// status <- status_n after a pass: all n vertices (Copy) or the endpoints of
// the round's input worklist edges (Merge)
template <bool MERGE>
__global__ void __launch_bounds__(kBlock) mis_determ_publish(const int* __restrict__ nlist,
                                                              const int* __restrict__ sp,
                                                              const int* __restrict__ status_n,
                                                              int* __restrict__ status, int n,
                                                              const int* __restrict__ wl0,
                                                              const int* __restrict__ wl1, Ctl* ctl)
{
  if (round_done(ctl)) return;
  const int stride = gridDim.x * kBlock;
  if constexpr (MERGE) {
    const int p = ctl->iter & 1;
    const int count = ctl->wl_count[p];
    const int* wl_in = p ? wl1 : wl0;
    for (int idx = blockIdx.x * kBlock + threadIdx.x; idx < count; idx += stride) {
      const int e = wl_in[idx];
      const int src = sp[e];
      const int dst = nlist[e];
      status[src] = status_n[src];
      status[dst] = status_n[dst];
    }
  } else {
    for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) status[v] = status_n[v];
  }
}

// ------------------------------------------------------- resumable pull ----
// Only the neighbours that outrank v matter for v joining the set: v joins
// once all of them are excluded.  A vertex that joins excludes all its
// neighbours itself (a push: plain stores, every one of them is either
// lower-ranked, and so cannot join while v is undecided or in, or already
// excluded), so an undecided vertex only has to wait: it walks its
// outranking neighbours in list order and stops at the first one that is not
// excluded (in the set -> v is excluded; undecided -> v waits).  Where it
// stopped is kept in pos[v] (every outranking neighbour before it is excluded,
// which is final), so later rounds resume there instead of rescanning.
// Statuses are read and written with plain loads and stores, not agent-scope
// atomics (on gfx950 those bypass the XCD's L2 and made the first round 8x
// slower): a status only ever moves from undecided to in or out, so a stale
// read can only say "undecided", and then v waits a round; every decision is
// taken on a value that was true.  The kernel boundary publishes the round.
// A lane walks at most kMisLane neighbours on its own per round and pushes
// rows of at most kMisLane edges itself; longer walks and pushes are finished
// by the whole wave, 64 neighbours per step.  The fixed point is the same
// greedy set; only the round in which a vertex is decided can differ from the
// full-scan styles.
constexpr int kMisLane = 32;

__global__ void __launch_bounds__(kBlock) mis_pos_init(const int* __restrict__ nindex, int* __restrict__ pos, int n)
{
  const int v = blockIdx.x * kBlock + threadIdx.x;
  if (v < n) pos[v] = nindex[v];
}

__global__ void __launch_bounds__(kBlock) mis_resume_pull(const int* __restrict__ nindex,
                                                           const int* __restrict__ nlist, int n,
                                                           int* __restrict__ status, int* __restrict__ pos, Ctl* ctl)
{
  if (round_done(ctl)) return;
  const int lane = lane_id();
  const int gw = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  const int nwaves = gridDim.x * kWavesPerBlock;
  bool pending = false;
  for (int base = gw * kWave; base < n; base += nwaves * kWave) {  // uniform per wave
    const int v = base + lane;
    const bool undec = (v < n) && status[v] == kUndecided;
    const uint32_t pv = mis_priority((uint32_t)v);
    int i = 0, end = 0, result = kUndecided;
    bool more = false;
    if (undec) {
      i = pos[v];
      end = nindex[v + 1];
      int steps = 0;
      for (; i < end && steps < kMisLane; i++, steps++) {
        const int u = nlist[i];
        if (u == v || !outranks(mis_priority((uint32_t)u), u, pv, v)) continue;
        const int su = status[u];
        if (su == kIncluded) {
          result = kExcluded;
          break;
        }
        if (su == kUndecided) break;  // waits for u; resumes here
      }
      if (result == kUndecided && i >= end) result = kIncluded;
      more = result == kUndecided && i < end && steps >= kMisLane;
    }
    uint64_t todo = ballot(more);
    while (todo != 0) {
      const int L = __ffsll((unsigned long long)todo) - 1;
      todo &= todo - 1;
      const int vL = base + L;
      const uint32_t pvL = mis_priority((uint32_t)vL);
      const int s0 = __shfl(i, L, kWave), e0 = __shfl(end, L, kWave);
      int res = kIncluded, stop = e0;
      for (int i0 = s0; i0 < e0; i0 += kWave) {
        const int idx = i0 + lane;
        int su = kExcluded;
        if (idx < e0) {
          const int u = nlist[idx];
          if (u != vL && outranks(mis_priority((uint32_t)u), u, pvL, vL)) su = status[u];
        }
        const uint64_t hit = ballot(su != kExcluded);
        if (hit != 0) {
          const int first = __ffsll((unsigned long long)hit) - 1;
          const int sf = __shfl(su, first, kWave);
          res = (sf == kIncluded) ? kExcluded : kUndecided;
          stop = i0 + first;
          break;
        }
      }
      if (lane == L) {
        result = res;
        i = stop;
      }
    }
    int rb = 0;
    if (undec) {
      if (result != kUndecided) {
        status[v] = result;
        rb = nindex[v];
      } else {
        pos[v] = i;
        pending = true;
      }
    }
    // a vertex that joined excludes its neighbours
    const bool joined = undec && result == kIncluded;
    const bool short_row = joined && end - rb <= kMisLane;
    if (short_row) {
      for (int j = rb; j < end; j++) {
        const int u = nlist[j];
        if (u != v) status[u] = kExcluded;
      }
    }
    uint64_t big = ballot(joined && !short_row);
    while (big != 0) {
      const int L = __ffsll((unsigned long long)big) - 1;
      big &= big - 1;
      const int vL = base + L;
      const int s0 = __shfl(rb, L, kWave), e0 = __shfl(end, L, kWave);
      for (int j = s0 + lane; j < e0; j += kWave) {
        const int u = nlist[j];
        if (u != vL) status[u] = kExcluded;
      }
    }
  }
  mis_raise_flag(ctl, pending);
}

// ------------------------------------------------------------ host side ----
// worklist: the round ends by the data-driven rule, else by the flag rule
template <typename K>
int mis_rounds(const igx_graph* g, igx_run_info* info, K launch, bool worklist = false)
{
  int rounds = 0, launches = 0;
  const int rc = run_rounds(
      [&](long long) {
        launch();
        if (worklist) launch_fin_worklist();
        else launch_fin_flag();
      },
      (long long)g->nodes + 2, &rounds, &launches);
  IGX_TRY(timer_stop(&info->seconds));
  info->iterations = rounds;
  info->launches = launches;
  return rc;
}

template <bool WAVE>
int style_vtopo_push(const igx_graph* g, int, void* out, const void*, igx_run_info* info)
{
  Runtime& R = rt();
  int* status = (int*)out;
  IGX_TRY(timer_start());
  IGX_TRY(ctl_reset(g->nodes));
  hipLaunchKernelGGL(mis_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, status, g->nodes);
  return mis_rounds(g, info, [&] {
    hipLaunchKernelGGL((mis_vtopo_push<WAVE>), dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nindex,
                       g->nlist, g->nodes, status, R.d_ctl);
  });
}

int style_vtopo_pull(const igx_graph* g, int, void* out, const void*, igx_run_info* info)
{
  Runtime& R = rt();
  int* status = (int*)out;
  IGX_TRY(timer_start());
  IGX_TRY(ctl_reset(g->nodes));
  hipLaunchKernelGGL(mis_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, status, g->nodes);
  return mis_rounds(g, info, [&] {
    hipLaunchKernelGGL(mis_vtopo_pull, dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nindex, g->nlist,
                       g->nodes, status, R.d_ctl);
  });
}

int style_vdata_pull(const igx_graph* g, int, void* out, const void*, igx_run_info* info)
{
  Runtime& R = rt();
  int* status = (int*)out;
  void *p0 = nullptr, *p1 = nullptr;
  IGX_TRY(ws_get(WS_A, sizeof(int) * (size_t)g->nodes, &p0));
  IGX_TRY(ws_get(WS_B, sizeof(int) * (size_t)g->nodes, &p1));
  IGX_TRY(timer_start());
  IGX_TRY(ctl_reset(g->nodes));
  hipLaunchKernelGGL(mis_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, status, g->nodes);
  hipLaunchKernelGGL(mis_seed, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, (int*)p0, g->nodes,
                     R.d_ctl);
  return mis_rounds(g, info, [&] {
    ktimer_begin("mis_vdata_pull");
    hipLaunchKernelGGL(mis_vdata_pull, dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nindex, g->nlist,
                       status, (int*)p0, (int*)p1, R.d_ctl);
    ktimer_end("mis_vdata_pull");
  }, true);
}

int style_resume_pull(const igx_graph* g, int, void* out, const void*, igx_run_info* info)
{
  Runtime& R = rt();
  int* status = (int*)out;
  void* pp = nullptr;
  IGX_TRY(ws_get(WS_A, sizeof(int) * (size_t)g->nodes, &pp));
  IGX_TRY(timer_start());
  IGX_TRY(ctl_reset(g->nodes));
  hipLaunchKernelGGL(mis_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, status, g->nodes);
  hipLaunchKernelGGL(mis_pos_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, g->nindex,
                     (int*)pp, g->nodes);
  return mis_rounds(g, info, [&] {
    ktimer_begin("mis_resume_pull");
    hipLaunchKernelGGL(mis_resume_pull, dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nindex, g->nlist,
                       g->nodes, status, (int*)pp, R.d_ctl);
    ktimer_end("mis_resume_pull");
  });
}

int style_etopo(const igx_graph* g, int, void* out, const void* aux, igx_run_info* info)
{
  Runtime& R = rt();
  int* status = (int*)out;
  const int* sp = (const int*)aux;
  void* pl = nullptr;
  IGX_TRY(ws_get(WS_A, sizeof(int) * (size_t)g->nodes, &pl));
  int* lost = (int*)pl;
  IGX_TRY(timer_start());
  IGX_TRY(ctl_reset(g->nodes));
  hipLaunchKernelGGL(mis_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, status, g->nodes);
  hipLaunchKernelGGL(mis_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, lost, g->nodes);
  return mis_rounds(g, info, [&] {
    hipLaunchKernelGGL(mis_edge_pass, dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nlist, sp, g->edges,
                       status, lost, R.d_ctl);
    hipLaunchKernelGGL(mis_vertex_pass, dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nodes, status, lost,
                       R.d_ctl);
  });
}

int style_edata(const igx_graph* g, int, void* out, const void* aux, igx_run_info* info)
{
  Runtime& R = rt();
  int* status = (int*)out;
  const int* sp = (const int*)aux;
  const size_t wl_items = (size_t)(g->edges > 0 ? g->edges : 1);
  void *pl = nullptr, *p0 = nullptr, *p1 = nullptr;
  IGX_TRY(ws_get(WS_A, sizeof(int) * (size_t)g->nodes, &pl));
  IGX_TRY(ws_get(WS_B, sizeof(int) * wl_items, &p0));
  IGX_TRY(ws_get(WS_C, sizeof(int) * wl_items, &p1));
  int* lost = (int*)pl;
  IGX_TRY(timer_start());
  IGX_TRY(ctl_reset(g->nodes));
  hipLaunchKernelGGL(mis_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, status, g->nodes);
  hipLaunchKernelGGL(mis_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, lost, g->nodes);
  hipLaunchKernelGGL(mis_edata_seed, dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nlist, sp, g->edges,
                     (int*)p0, R.d_ctl);
  int rounds = 0, launches = 0;
  const int rc = run_rounds(
      [&](long long) {
        hipLaunchKernelGGL(mis_edata_edge, dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nlist, sp, status,
                           lost, (const int*)p0, (const int*)p1, R.d_ctl);
        hipLaunchKernelGGL(mis_edata_vertex, dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nlist, sp,
                           status, (const int*)lost, (int*)p0, (int*)p1, R.d_ctl);
        launch_fin_worklist();
      },
      (long long)g->nodes + 2, &rounds, &launches);
  hipLaunchKernelGGL(mis_edata_last, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, status,
                     g->nodes);
  IGX_TRY(timer_stop(&info->seconds));
  info->iterations = rounds;
  info->launches = launches;
  return rc;
}

template <bool MERGE>
int style_edata_determ(const igx_graph* g, int, void* out, const void* aux, igx_run_info* info)
{
  Runtime& R = rt();
  int* status = (int*)out;
  const int* sp = (const int*)aux;
  const size_t wl_items = (size_t)(g->edges > 0 ? g->edges : 1);
  void *pl = nullptr, *p0 = nullptr, *p1 = nullptr, *pn = nullptr;
  IGX_TRY(ws_get(WS_A, sizeof(int) * (size_t)g->nodes, &pl));
  IGX_TRY(ws_get(WS_B, sizeof(int) * wl_items, &p0));
  IGX_TRY(ws_get(WS_C, sizeof(int) * wl_items, &p1));
  IGX_TRY(ws_get(WS_D, sizeof(int) * (size_t)g->nodes, &pn));
  int* lost = (int*)pl;
  int* status_n = (int*)pn;
  IGX_TRY(timer_start());
  IGX_TRY(ctl_reset(g->nodes));
  const dim3 vb(blocks_for(g->nodes, kBlock));
  hipLaunchKernelGGL(mis_init, vb, dim3(kBlock), 0, R.stream, status, g->nodes);
  hipLaunchKernelGGL(mis_init, vb, dim3(kBlock), 0, R.stream, status_n, g->nodes);
  hipLaunchKernelGGL(mis_init, vb, dim3(kBlock), 0, R.stream, lost, g->nodes);
  hipLaunchKernelGGL(mis_edata_seed, dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nlist, sp, g->edges,
                     (int*)p0, R.d_ctl);
  int rounds = 0, launches = 0;
  const int rc = run_rounds(
      [&](long long) {
        const dim3 grid(R.persist_blocks), blk(kBlock);
        hipLaunchKernelGGL(mis_determ_edge, grid, blk, 0, R.stream, g->nlist, sp, (const int*)status, status_n, lost,
                           (const int*)p0, (const int*)p1, R.d_ctl);
        hipLaunchKernelGGL((mis_determ_publish<MERGE>), grid, blk, 0, R.stream, g->nlist, sp,
                           (const int*)status_n, status, g->nodes, (const int*)p0, (const int*)p1, R.d_ctl);
        hipLaunchKernelGGL(mis_determ_vertex, grid, blk, 0, R.stream, g->nlist, sp, (const int*)status, status_n,
                           (const int*)lost, (int*)p0, (int*)p1, R.d_ctl);
        hipLaunchKernelGGL((mis_determ_publish<MERGE>), grid, blk, 0, R.stream, g->nlist, sp,
                           (const int*)status_n, status, g->nodes, (const int*)p0, (const int*)p1, R.d_ctl);
        launch_fin_worklist();
      },
      (long long)g->nodes + 2, &rounds, &launches);
  hipLaunchKernelGGL(mis_edata_last, vb, dim3(kBlock), 0, R.stream, status, g->nodes);
  IGX_TRY(timer_stop(&info->seconds));
  info->iterations = rounds;
  info->launches = launches;
  return rc;
}

const StyleEntry kStyles[] = {
    {"v_topo_push_thread", style_vtopo_push<false>, 0},
    {"v_topo_push_wave", style_vtopo_push<true>, 0},
    {"v_topo_pull_thread", style_vtopo_pull, 0},
    {"v_data_pull_thread", style_vdata_pull, 0},
    {"e_topo_thread", style_etopo, 1},
    {"v_resume_pull", style_resume_pull, 0},
    {"e_data_thread", style_edata, 1},
    {"e_data_thread_copydeterm", style_edata_determ<false>, 1},
    {"e_data_thread_mergedeterm", style_edata_determ<true>, 1},
};

}  // namespace

const StyleEntry* mis_styles(int* count)
{
  *count = (int)(sizeof(kStyles) / sizeof(kStyles[0]));
  return kStyles;
}

}  // namespace igx
