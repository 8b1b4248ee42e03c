// Connected components on gfx950: label[v] = minimum vertex id of v's
// component (the fixed point of the reference's min-label propagation,
// `label[v] = v` then repeated atomicMin along edges).  Inputs are undirected
// (both directions stored), as all Indigo2 inputs are.
//
// Reference family: codes/cuda/cc-cuda (168 variants).
//   v_topo_push_*   [C-CVT] CC_CUDA_V_Topo_Push_..._Thread:47-116
//   v_topo_pull_*   CC_CUDA_V_Topo_Pull siblings (read neighbours, write self)
//   e_topo_push_*   [C-CET]:57-69 one edge per item
//   v_data_push_*   [C-CVD]:46-136 worklist seeded with every vertex
//   e_data_push_*   [C-CED]:48-80 edge worklist seeded with every edge
//   v_topo_{push,pull}_{thread,wave}_determ  V_Topo_{Push,Pull}_Determ
//                   siblings: Jacobi rounds over label / label_n buffers
//   uf_hook_compress  union-find (no reference counterpart): parents start at
//                   a smaller neighbour, every edge (v, u<v) hooks the larger
//                   root under the smaller with a CAS, and a final pass points
//                   every vertex at its root.  Roots only ever link to smaller
//                   ids, so each root is its component's minimum id.
#include "igx_internal.hpp"
#include "igx_frontier.hpp"

namespace igx {
namespace {

enum Gran { kThread = 0, kWaveG = 1, kBlockG = 2 };

__global__ void __launch_bounds__(kBlock) cc_init(int* __restrict__ label, int n)
{
  const int v = blockIdx.x * kBlock + threadIdx.x;
  if (v < n) label[v] = v;
}

__global__ void __launch_bounds__(kBlock) cc_fill(int* __restrict__ p, int value, long long count)
{
  const long long stride = (long long)gridDim.x * kBlock;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < count; i += stride) p[i] = value;
}


__device__ __forceinline__ void cc_raise_flag(Ctl* ctl, bool updated)
{
  const uint64_t any = ballot(updated);
  if (any != 0 && lane_id() == (__ffsll((unsigned long long)any) - 1)) store_relaxed(&ctl->flag, 1);
}

template <int G, int AK, bool PERSIST>
__global__ void __launch_bounds__(kBlock) cc_vtopo_push(const int* __restrict__ nindex,
                                                         const int* __restrict__ nlist, int n,
                                                         int* __restrict__ label, Ctl* ctl)
{
  if (round_done(ctl)) return;
  bool updated = false;
  if constexpr (G == kThread) {
    const int stride = gridDim.x * kBlock;
    for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
      const int lv = label[v];
      for (int i = nindex[v]; i < nindex[v + 1]; i++) {
        const int dst = nlist[i];
        if (label[dst] > lv && atomic_min_k<AK>(&label[dst], lv) > lv) updated = true;
      }
    }
  } else if constexpr (G == kWaveG) {
    const int waves = gridDim.x * kWavesPerBlock;
    const int lane = lane_id();
    for (int v = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; v < n; v += waves) {
      const int lv = label[v];
      for (int i = nindex[v] + lane; i < nindex[v + 1]; i += kWave) {
        const int dst = nlist[i];
        if (label[dst] > lv && atomic_min_k<AK>(&label[dst], lv) > lv) updated = true;
      }
    }
  } else {
    for (int v = blockIdx.x; v < n; v += gridDim.x) {
      const int lv = label[v];
      for (int i = nindex[v] + (int)threadIdx.x; i < nindex[v + 1]; i += kBlock) {
        const int dst = nlist[i];
        if (label[dst] > lv && atomic_min_k<AK>(&label[dst], lv) > lv) updated = true;
      }
    }
  }
  cc_raise_flag(ctl, updated);
}

__global__ void __launch_bounds__(kBlock) cc_vtopo_pull(const int* __restrict__ nindex,
                                                         const int* __restrict__ nlist, int n,
                                                         int* __restrict__ label, Ctl* ctl)
{
  if (round_done(ctl)) return;
  bool updated = false;
  const int stride = gridDim.x * kBlock;
  for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
    const int l0 = label[v];
    int l = l0;
    for (int i = nindex[v]; i < nindex[v + 1]; i++) l = min(l, load_relaxed(&label[nlist[i]]));
    if (l < l0) {
      store_relaxed(&label[v], l);
      updated = true;
    }
  }
  cc_raise_flag(ctl, updated);
}

// ------------------------------------------------ deterministic (Jacobi) ----
// CC_CUDA_V_Topo_Push_Determ_IntType_ReadModifyWrite_NonPersist_Atomic_Thread_
// NoBoundsBug2_NoBoundsBug1.cu cc():57-77 and the V_Topo_Pull_Determ sibling
// (..._ReadWrite_NonPersist_Atomic_Thread_... cc():58-80), Warp forms alike: a
// round reads only `cur` and writes only `nxt`, the two trading places every
// round, so the labels after round r are the minima over r-hop
// neighbourhoods whatever the schedule.  Push: nxt enters a round with the
// labels of two rounds ago (never below cur's), every vertex lowers its own
// nxt entry to its cur label and then pushes that label to its neighbours; a
// push counts as progress only below the neighbour's cur label, so the last
// round changes no label and leaves both buffers equal.  Pull: every vertex
// writes the minimum of its own and its neighbours' cur labels to nxt.
template <int G>
__global__ void __launch_bounds__(kBlock) cc_determ_push(const int* __restrict__ nindex,
                                                          const int* __restrict__ nlist, int n,
                                                          const int* __restrict__ cur, int* __restrict__ nxt,
                                                          Ctl* ctl)
{
  if (round_done(ctl)) return;
  bool updated = false;
  auto push = [&](int i, int l) {
    const int dst = nlist[i];
    if (load_relaxed(&nxt[dst]) > l && atomic_min_relaxed(&nxt[dst], l) > l && l < cur[dst]) updated = true;
  };
  if constexpr (G == kThread) {
    const int stride = gridDim.x * kBlock;
    for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
      const int l = cur[v];
      if (load_relaxed(&nxt[v]) > l) atomic_min_relaxed(&nxt[v], l);
      for (int i = nindex[v]; i < nindex[v + 1]; i++) push(i, l);
    }
  } else {
    const int waves = gridDim.x * kWavesPerBlock;
    const int lane = lane_id();
    for (int v = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; v < n; v += waves) {
      const int l = cur[v];
      if (lane == 0 && load_relaxed(&nxt[v]) > l) atomic_min_relaxed(&nxt[v], l);
      for (int i = nindex[v] + lane; i < nindex[v + 1]; i += kWave) push(i, l);
    }
  }
  cc_raise_flag(ctl, updated);
}

template <int G>
__global__ void __launch_bounds__(kBlock) cc_determ_pull(const int* __restrict__ nindex,
                                                          const int* __restrict__ nlist, int n,
                                                          const int* __restrict__ cur, int* __restrict__ nxt,
                                                          Ctl* ctl)
{
  if (round_done(ctl)) return;
  bool updated = false;
  if constexpr (G == kThread) {
    const int stride = gridDim.x * kBlock;
    for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
      const int l0 = cur[v];
      int l = l0;
      for (int i = nindex[v]; i < nindex[v + 1]; i++) l = min(l, cur[nlist[i]]);
      nxt[v] = l;
      updated |= l < l0;
    }
  } else {
    const int waves = gridDim.x * kWavesPerBlock;
    const int lane = lane_id();
    for (int v = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; v < n; v += waves) {
      const int l0 = cur[v];
      long long l = l0;
      for (int i = nindex[v] + lane; i < nindex[v + 1]; i += kWave) l = min(l, (long long)cur[nlist[i]]);
      l = wave_min_ll(l);
      if (lane == 0) {
        nxt[v] = (int)l;
        updated |= l < l0;
      }
    }
  }
  cc_raise_flag(ctl, updated);
}

__global__ void __launch_bounds__(kBlock) cc_etopo_push(const int* __restrict__ nlist, const int* __restrict__ sp,
                                                         int m, int* __restrict__ label, Ctl* ctl)
{
  if (round_done(ctl)) return;
  bool updated = false;
  const int stride = gridDim.x * kBlock;
  for (int e = blockIdx.x * kBlock + threadIdx.x; e < m; e += stride) {
    const int lv = label[sp[e]];
    const int dst = nlist[e];
    if (label[dst] > lv && atomic_min_relaxed(&label[dst], lv) > lv) updated = true;
  }
  cc_raise_flag(ctl, updated);
}

// worklist seeded with all items [0, count)
__global__ void __launch_bounds__(kBlock) cc_seed(int* __restrict__ wl, int count, Ctl* ctl)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < count) wl[i] = i;
  if (i == 0) {
    ctl->wl_count[0] = count;
    ctl->wl_count[1] = 0;
  }
}

__global__ void __launch_bounds__(kBlock) cc_vdata_push(const int* __restrict__ nindex,
                                                         const int* __restrict__ nlist, int* __restrict__ label,
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
  const int stride = gridDim.x * kBlock;
  for (int idx = blockIdx.x * kBlock + threadIdx.x; idx < count; idx += stride) {
    const int v = wl_in[idx];
    const int lv = label[v];
    for (int i = nindex[v]; i < nindex[v + 1]; i++) {
      const int dst = nlist[i];
      bool push = false;
      if (label[dst] > lv && atomic_min_relaxed(&label[dst], lv) > lv) {
        push = atomic_max_relaxed(&stamp[dst], tag) != tag;
      }
      const int slot = wave_push_slot(out_count, push);
      if (push) wl_out[slot] = dst;
    }
  }
}

// One thread per worklist edge; when dst's label drops, all of dst's edges
// are queued at once (NonDup per source vertex, see bfs_edata_push) and
// written wave-cooperatively.
__global__ void __launch_bounds__(kBlock) cc_edata_push(const int* __restrict__ nindex,
                                                         const int* __restrict__ nlist, const int* __restrict__ sp,
                                                         int* __restrict__ label, int* __restrict__ wl0,
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
      const int lv = label[sp[e]];
      const int dst = nlist[e];
      if (label[dst] > lv && atomic_min_relaxed(&label[dst], lv) > lv &&
          atomic_max_relaxed(&vstamp[dst], tag) != tag) {
        rbeg = nindex[dst];
        rlen = nindex[dst + 1] - rbeg;
      }
    }
    wave_push_edge_ranges(wl_out, out_count, rbeg, rlen, s_excl[w], s_beg[w]);
  }
}

// ------------------------------------------------------------- union-find --
// Parent pointers satisfy parent[x] <= x; roots point at themselves.  Path
// halving only rewrites non-roots, and roots change only through the CAS in
// uf_unite, so concurrent finds and unions keep the forest valid.
__device__ __forceinline__ int uf_find(int* __restrict__ parent, int x)
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

__device__ __forceinline__ void uf_unite(int* __restrict__ parent, int a, int b)
{
  int ra = uf_find(parent, a);
  int rb = uf_find(parent, b);
  while (ra != rb) {
    const int hi = max(ra, rb), lo = min(ra, rb);
    int expected = hi;
    if (__hip_atomic_compare_exchange_strong(&parent[hi], &expected, lo, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT)) {
      return;
    }
    // hi was linked meanwhile: `expected` holds its current parent
    ra = uf_find(parent, expected);
    rb = uf_find(parent, lo);
  }
}

// parent[v] = the first neighbour if it is smaller (else v).
__global__ void __launch_bounds__(kBlock) uf_init(const int* __restrict__ nindex, const int* __restrict__ nlist,
                                                   int n, int* __restrict__ parent)
{
  const int stride = gridDim.x * kBlock;
  for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
    const int beg = nindex[v], end = nindex[v + 1];
    // any smaller neighbour is a valid first parent; with sorted adjacency
    // the first neighbour is the smallest one
    const int u0 = (end > beg) ? nlist[beg] : v;
    parent[v] = (u0 < v) ? u0 : v;
  }
}

// Hook every edge (v, u) with u < v.  Rows of at most kChunk edges are walked
// 64 rows per wave over their contiguous edge range (coalesced nlist reads);
// larger rows come from the chunk queue so hub edges spread over all waves.
__global__ void __launch_bounds__(kBlock) uf_hook(const int* __restrict__ nindex, const int* __restrict__ nlist,
                                                   int n, int* __restrict__ parent, const Entry* __restrict__ bigq,
                                                   const int* __restrict__ bigq_count)
{
  __shared__ int s_excl[kWavesPerBlock][kWave];
  __shared__ int s_beg[kWavesPerBlock][kWave];
  __shared__ Entry s_ent[kWavesPerBlock][kWave];
  const int w = threadIdx.x / kWave;
  const int lane = lane_id();
  const int gw = blockIdx.x * kWavesPerBlock + w;
  const int nwaves = gridDim.x * kWavesPerBlock;
  __shared__ int s_root[kWavesPerBlock][kWave];
  int* excl = s_excl[w];
  int* rbeg = s_beg[w];
  int* rroot = s_root[w];
  for (int v0 = gw * kWave; v0 < n; v0 += nwaves * kWave) {
    const int v = v0 + lane;
    int beg = 0, len = 0;
    if (v < n) {
      beg = nindex[v];
      len = nindex[v + 1] - beg;
      if (len > kChunk) len = 0;
    }
    const int incl = wave_incl_scan(len);
    const int total = wave_bcast(incl, kWave - 1);
    excl[lane] = incl - len;
    rbeg[lane] = beg;
    // the row's current root, reused as a hint by every edge of the row
    rroot[lane] = (len > 0) ? uf_find(parent, v) : v;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    // kIlp independent edges per lane per step so their loads overlap
    constexpr int kIlp = 4;
    for (int tb = 0; tb < total; tb += kIlp * kWave) {
      int u[kIlp], row[kIlp], hint[kIlp], pu[kIlp], ph[kIlp], ppu[kIlp];
      bool need[kIlp];
      #pragma unroll
      for (int k = 0; k < kIlp; k++) {
        const int t = tb + k * kWave + lane;
        need[k] = false;
        row[k] = 0;
        hint[k] = 0;
        u[k] = 0;
        if (t < total) {
          int lo = 0, hi = kWave;
          while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (excl[mid] <= t) lo = mid;
            else hi = mid;
          }
          row[k] = v0 + lo;
          hint[k] = rroot[lo];
          u[k] = nlist[rbeg[lo] + (t - excl[lo])];
          need[k] = true;
        }
      }
      #pragma unroll
      for (int k = 0; k < kIlp; k++) {
        need[k] = need[k] && u[k] < row[k];
        pu[k] = need[k] ? parent[u[k]] : 0;
        ph[k] = need[k] ? parent[hint[k]] : 0;
      }
      #pragma unroll
      for (int k = 0; k < kIlp; k++) ppu[k] = (need[k] && pu[k] != u[k]) ? parent[pu[k]] : pu[k];
      #pragma unroll
      for (int k = 0; k < kIlp; k++) {
        if (!need[k]) continue;
        // two levels are resolved above; deeper trees fall back to the loop
        int ru = (pu[k] == u[k]) ? u[k] : ((ppu[k] == pu[k]) ? pu[k] : uf_find(parent, ppu[k]));
        int ra = (ph[k] == hint[k]) ? hint[k] : uf_find(parent, ph[k]);
        while (ra != ru) {
          const int hi2 = max(ra, ru), lo2 = min(ra, ru);
          int expected = hi2;
          if (__hip_atomic_compare_exchange_strong(&parent[hi2], &expected, lo2, __ATOMIC_RELAXED,
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            break;
          }
          ra = uf_find(parent, expected);
          ru = uf_find(parent, lo2);
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  const int qsize = *bigq_count;
  wave_expand(bigq, qsize, gw, nwaves, s_excl[w], s_ent[w], [&](const Entry& en, int e) {
    const int u = nlist[e];
    if (u < en.v) uf_unite(parent, en.v, u);
  });
}

// Final labels.  The walk is read-only and each thread writes only its own
// vertex: a concurrent path-halving write could otherwise replace another
// vertex's final root with a non-root ancestor.  Every value read is the
// vertex's old ancestor or its root, so each walk ends at the root.
__global__ void __launch_bounds__(kBlock) uf_compress(int n, int* __restrict__ parent)
{
  const int stride = gridDim.x * kBlock;
  for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
    int x = parent[v];
    int p = parent[x];
    while (p != x) {
      x = p;
      p = parent[x];
    }
    parent[v] = x;
  }
}

// -------------------------------------------------- sampled union-find ----
// Link every vertex to its first kAfRounds neighbours, compress, find the
// root shared by most of kAfSamples sampled vertices (the giant component), then
// hook the remaining edges of only those vertices outside it.  An edge with
// an endpoint outside the sampled component is processed from that endpoint;
// an edge with both endpoints inside joins nothing new.  Skipping is decided
// by reading parent[v] == c, which only ever holds for vertices already
// connected to c.  The root of every tree is still its minimum id.
constexpr int kAfRounds = 2;
constexpr int kAfSamples = 256;

// parent[v] = v's first neighbour when that is smaller: linking a vertex to a
// smaller vertex needs no atomics while every vertex is still a root, and the
// pointers stay monotone (parent[x] <= x).  A first neighbour that is larger
// is linked by af_link below, with the second neighbour.
__global__ void __launch_bounds__(kBlock) af_init(const int* __restrict__ nindex, const int* __restrict__ nlist,
                                                   int n, int* __restrict__ parent)
{
  const int stride = gridDim.x * kBlock;
  for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
    const int beg = nindex[v], end = nindex[v + 1];
    const int u0 = (end > beg) ? nlist[beg] : v;
    parent[v] = (u0 < v) ? u0 : v;
  }
}

__global__ void __launch_bounds__(kBlock) af_link(const int* __restrict__ nindex, const int* __restrict__ nlist,
                                                   int n, int* __restrict__ parent)
{
  const int stride = gridDim.x * kBlock;
  for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
    const int beg = nindex[v], end = nindex[v + 1];
    if (end > beg && nlist[beg] > v) uf_unite(parent, v, nlist[beg]);
    #pragma unroll
    for (int r = 1; r < kAfRounds; r++) {
      if (beg + r < end) uf_unite(parent, v, nlist[beg + r]);
    }
  }
}

// one workgroup: the most frequent root among kAfSamples hashed vertices
// (bitonic sort in LDS, then each run's length by a binary search for its end)
__global__ void __launch_bounds__(kBlock) af_sample(int n, const int* __restrict__ parent, int* __restrict__ out)
{
  __shared__ int s_key[kAfSamples];
  __shared__ int s_len[kBlock];
  __shared__ int s_val[kBlock];
  for (int i = threadIdx.x; i < kAfSamples; i += kBlock) {
    uint32_t h = (uint32_t)i * 2654435761u + 0x9E3779B9u;
    h ^= h >> 16;
    h *= 0x45d9f3bu;
    h ^= h >> 16;
    s_key[i] = parent[(int)(h % (uint32_t)n)];
  }
  __syncthreads();
  for (int k = 2; k <= kAfSamples; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < kAfSamples; i += kBlock) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const int a = s_key[i], b = s_key[ixj];
          const bool up = (i & k) == 0;
          if ((a > b) == up) {
            s_key[i] = b;
            s_key[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  int best_len = 0, best_val = s_key[0];
  for (int i = threadIdx.x; i < kAfSamples; i += kBlock) {
    const int x = s_key[i];
    if (i > 0 && s_key[i - 1] == x) continue;  // not the start of a run
    int lo = i, hi = kAfSamples;                // first index with key > x
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (s_key[mid] <= x) lo = mid + 1;
      else hi = mid;
    }
    if (lo - i > best_len) {
      best_len = lo - i;
      best_val = x;
    }
  }
  s_len[threadIdx.x] = best_len;
  s_val[threadIdx.x] = best_val;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off && s_len[threadIdx.x + off] > s_len[threadIdx.x]) {
      s_len[threadIdx.x] = s_len[threadIdx.x + off];
      s_val[threadIdx.x] = s_val[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s_val[0];
}

__global__ void __launch_bounds__(kBlock) af_link_rest(const int* __restrict__ nindex, const int* __restrict__ nlist,
                                                        int n, int* __restrict__ parent,
                                                        const Entry* __restrict__ bigq,
                                                        const int* __restrict__ bigq_count,
                                                        const int* __restrict__ giant)
{
  __shared__ int s_excl[kWavesPerBlock][kWave];
  __shared__ int s_beg[kWavesPerBlock][kWave];
  __shared__ Entry s_ent[kWavesPerBlock][kWave];
  const int c = *giant;
  const int w = threadIdx.x / kWave;
  const int lane = lane_id();
  const int gw = blockIdx.x * kWavesPerBlock + w;
  const int nwaves = gridDim.x * kWavesPerBlock;
  int* excl = s_excl[w];
  int* rbeg = s_beg[w];
  for (int v0 = gw * kWave; v0 < n; v0 += nwaves * kWave) {
    const int v = v0 + lane;
    int beg = 0, len = 0;
    if (v < n && parent[v] != c) {
      const int b0 = nindex[v], e0 = nindex[v + 1];
      if (e0 - b0 <= kChunk) {
        beg = b0 + kAfRounds;
        len = (e0 > beg) ? e0 - beg : 0;
      }
    }
    const int incl = wave_incl_scan(len);
    const int total = wave_bcast(incl, kWave - 1);
    if (total == 0) continue;  // uniform
    __builtin_amdgcn_wave_barrier();
    excl[lane] = incl - len;
    rbeg[lane] = beg;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    for (int tb = 0; tb < total; tb += kWave) {
      const int t = tb + lane;
      if (t < total) {
        int lo = 0, hi = kWave;
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (excl[mid] <= t) lo = mid;
          else hi = mid;
        }
        uf_unite(parent, v0 + lo, nlist[rbeg[lo] + (t - excl[lo])]);
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  // longer rows by chunks; a chunk whose row is in the sampled component (at
  // the time it is fetched) is skipped whole, the first kAfRounds edges of a
  // row were linked before
  wave_expand_with(*bigq_count, gw, nwaves, s_excl[w], s_ent[w],
                   [&](int i) {
                     Entry en = bigq[i];
                     if (parent[en.v] == c) {
                       en.end = en.beg;
                     } else {
                       const int first = nindex[en.v] + kAfRounds;
                       if (en.beg < first) en.beg = (first < en.end) ? first : en.end;
                     }
                     return en;
                   },
                   [&](const Entry& en, int e) { uf_unite(parent, en.v, nlist[e]); });
}

// ------------------------------------------------------------ host side ----
template <typename K>
int run_cc_topo(const igx_graph* g, int* d_label, igx_run_info* info, K launch)
{
  Runtime& R = rt();
  IGX_TRY(timer_start());
  IGX_TRY(ctl_reset(g->nodes));
  hipLaunchKernelGGL(cc_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, d_label, g->nodes);
  int rounds = 0, launches = 0;
  const int rc = run_rounds(
      [&](long long) {
        launch();
        launch_fin_flag();
      },
      (long long)g->nodes + 2, &rounds, &launches);
  IGX_TRY(timer_stop(&info->seconds));
  info->iterations = rounds;
  info->launches = launches;
  return rc;
}

template <int G, int AK, bool PERSIST>
int style_vtopo_push(const igx_graph* g, int, void* out, const void*, igx_run_info* info)
{
  Runtime& R = rt();
  int* label = (int*)out;
  int grid = R.persist_blocks;
  if (!PERSIST) {
    grid = (G == kThread) ? blocks_for(g->nodes, kBlock)
                          : (G == kWaveG ? blocks_for(g->nodes, kWavesPerBlock) : g->nodes);
  }
  return run_cc_topo(g, label, info, [&] {
    hipLaunchKernelGGL((cc_vtopo_push<G, AK, PERSIST>), dim3(grid), dim3(kBlock), 0, R.stream, g->nindex, g->nlist,
                       g->nodes, label, R.d_ctl);
  });
}

int style_vtopo_pull(const igx_graph* g, int, void* out, const void*, igx_run_info* info)
{
  Runtime& R = rt();
  int* label = (int*)out;
  return run_cc_topo(g, label, info, [&] {
    hipLaunchKernelGGL(cc_vtopo_pull, dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nindex, g->nlist,
                       g->nodes, label, R.d_ctl);
  });
}

int style_etopo_push(const igx_graph* g, int, void* out, const void* aux, igx_run_info* info)
{
  Runtime& R = rt();
  int* label = (int*)out;
  const int* sp = (const int*)aux;
  return run_cc_topo(g, label, info, [&] {
    hipLaunchKernelGGL(cc_etopo_push, dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nlist, sp, g->edges,
                       label, R.d_ctl);
  });
}

template <bool EDGE>
int style_data_push(const igx_graph* g, int, void* out, const void* aux, igx_run_info* info)
{
  Runtime& R = rt();
  int* label = (int*)out;
  const int* sp = (const int*)aux;
  const size_t items = (size_t)std::max(g->edges, g->nodes);
  const int seed_count = EDGE ? g->edges : g->nodes;
  const size_t stamp_items = (size_t)g->nodes;  // per-vertex NonDup stamps in both kinds
  void *p0 = nullptr, *p1 = nullptr, *p2 = nullptr;
  IGX_TRY(ws_get(WS_A, sizeof(int) * items, &p0));
  IGX_TRY(ws_get(WS_B, sizeof(int) * items, &p1));
  IGX_TRY(ws_get(WS_C, sizeof(int) * stamp_items, &p2));
  int* wl0 = (int*)p0;
  int* wl1 = (int*)p1;
  int* stamp = (int*)p2;
  IGX_TRY(timer_start());
  IGX_TRY(ctl_reset(g->nodes));
  hipLaunchKernelGGL(cc_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, label, g->nodes);
  hipLaunchKernelGGL(cc_fill, dim3(blocks_for((long long)stamp_items, kBlock)), dim3(kBlock), 0, R.stream, stamp, 0,
                     (long long)stamp_items);
  hipLaunchKernelGGL(cc_seed, dim3(blocks_for(seed_count, kBlock)), dim3(kBlock), 0, R.stream, wl0, seed_count,
                     R.d_ctl);
  int rounds = 0, launches = 0;
  const int rc = run_rounds(
      [&](long long) {
        if (EDGE) {
          hipLaunchKernelGGL(cc_edata_push, dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nindex, g->nlist,
                             sp, label, wl0, wl1, stamp, R.d_ctl);
        } else {
          hipLaunchKernelGGL(cc_vdata_push, dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nindex, g->nlist,
                             label, wl0, wl1, stamp, R.d_ctl);
        }
        launch_fin_worklist();
      },
      (long long)g->nodes + 2, &rounds, &launches);
  IGX_TRY(timer_stop(&info->seconds));
  info->iterations = rounds;
  info->launches = launches;
  return rc;
}

int style_uf(const igx_graph* g, int, void* out, const void*, igx_run_info* info)
{
  Runtime& R = rt();
  int* parent = (int*)out;
  // rows with more than kChunk edges need ceil(deg / kChunk) < 2 * deg / kChunk slots
  const size_t cap = 2 * ((size_t)g->edges / kChunk) + 64;
  void *pq = nullptr, *pc = nullptr;
  IGX_TRY(ws_get(WS_D, sizeof(Entry) * cap, &pq));
  IGX_TRY(ws_get(WS_E, sizeof(int) * 64, &pc));
  const int grid = R.persist_blocks;
  IGX_TRY(timer_start());
  hipLaunchKernelGGL(cc_fill, dim3(1), dim3(kBlock), 0, R.stream, (int*)pc, 0, 1LL);
  hipLaunchKernelGGL(uf_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, g->nindex, g->nlist,
                     g->nodes, parent);
  IGX_TRY(big_rows(g, pq, (int*)pc));
  ktimer_begin("uf_hook");
  hipLaunchKernelGGL(uf_hook, dim3(grid), dim3(kBlock), 0, R.stream, g->nindex, g->nlist, g->nodes, parent,
                     (const Entry*)pq, (const int*)pc);
  ktimer_end("uf_hook");
  hipLaunchKernelGGL(uf_compress, dim3(grid), dim3(kBlock), 0, R.stream, g->nodes, parent);
  IGX_HIP(hipGetLastError());
  IGX_TRY(timer_stop(&info->seconds));
  info->iterations = 1;
  info->launches = 5;
  return IGX_OK;
}

int style_afforest(const igx_graph* g, int, void* out, const void*, igx_run_info* info)
{
  Runtime& R = rt();
  int* parent = (int*)out;
  // chunks of the rows longer than kChunk: at most 2 * edges / kChunk
  const size_t cap = 2 * ((size_t)g->edges / kChunk) + 64;
  void *pq = nullptr, *pc = nullptr;
  IGX_TRY(ws_get(WS_D, sizeof(Entry) * cap, &pq));
  IGX_TRY(ws_get(WS_E, sizeof(int) * 64, &pc));
  int* counters = (int*)pc;  // [0] big chunk count, [1] sampled giant root
  const int grid = R.persist_blocks;
  IGX_TRY(timer_start());
  hipLaunchKernelGGL(cc_fill, dim3(1), dim3(kBlock), 0, R.stream, counters, 0, 2LL);
  hipLaunchKernelGGL(af_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, g->nindex, g->nlist,
                     g->nodes, parent);
  IGX_TRY(big_rows(g, pq, counters));
  hipLaunchKernelGGL(af_link, dim3(grid), dim3(kBlock), 0, R.stream, g->nindex, g->nlist, g->nodes, parent);
  hipLaunchKernelGGL(uf_compress, dim3(grid), dim3(kBlock), 0, R.stream, g->nodes, parent);
  hipLaunchKernelGGL(af_sample, dim3(1), dim3(kBlock), 0, R.stream, g->nodes, parent, counters + 1);
  ktimer_begin("af_link_rest");
  hipLaunchKernelGGL(af_link_rest, dim3(grid), dim3(kBlock), 0, R.stream, g->nindex, g->nlist, g->nodes, parent,
                     (const Entry*)pq, (const int*)counters, (const int*)(counters + 1));
  ktimer_end("af_link_rest");
  hipLaunchKernelGGL(uf_compress, dim3(grid), dim3(kBlock), 0, R.stream, g->nodes, parent);
  IGX_HIP(hipGetLastError());
  IGX_TRY(timer_stop(&info->seconds));
  info->iterations = 1;
  info->launches = 8;
  return IGX_OK;
}

// Both buffers start equal and end equal (see cc_determ_push), so the result
// is in `out` whichever buffer the last round wrote.
template <bool PULL, int G>
int style_determ(const igx_graph* g, int, void* out, const void*, igx_run_info* info)
{
  Runtime& R = rt();
  int* l0 = (int*)out;
  void* p1 = nullptr;
  IGX_TRY(ws_get(WS_A, sizeof(int) * (size_t)g->nodes, &p1));
  int* l1 = (int*)p1;
  IGX_TRY(timer_start());
  IGX_TRY(ctl_reset(g->nodes));
  hipLaunchKernelGGL(cc_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, l0, g->nodes);
  hipLaunchKernelGGL(cc_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, l1, g->nodes);
  int rounds = 0, launches = 0;
  const int rc = run_rounds(
      [&](long long r) {
        const int* cur = (r & 1) ? l1 : l0;
        int* nxt = (r & 1) ? l0 : l1;
        if constexpr (PULL) {
          hipLaunchKernelGGL((cc_determ_pull<G>), dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nindex,
                             g->nlist, g->nodes, cur, nxt, R.d_ctl);
        } else {
          hipLaunchKernelGGL((cc_determ_push<G>), dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nindex,
                             g->nlist, g->nodes, cur, nxt, R.d_ctl);
        }
        launch_fin_flag();
      },
      (long long)g->nodes + 2, &rounds, &launches);
  IGX_TRY(timer_stop(&info->seconds));
  info->iterations = rounds;
  info->launches = launches;
  return rc;
}

const StyleEntry kStyles[] = {
    {"v_topo_push_thread", style_vtopo_push<kThread, kAtomicRelaxed, true>, 0},
    {"v_topo_push_wave", style_vtopo_push<kWaveG, kAtomicRelaxed, true>, 0},
    {"v_topo_push_block", style_vtopo_push<kBlockG, kAtomicRelaxed, true>, 0},
    {"v_topo_push_thread_nonpersist", style_vtopo_push<kThread, kAtomicRelaxed, false>, 0},
    {"v_topo_push_thread_cas", style_vtopo_push<kThread, kAtomicCas, true>, 0},
    {"v_topo_push_thread_seqcst", style_vtopo_push<kThread, kAtomicSeqCst, true>, 0},
    {"v_topo_pull_thread", style_vtopo_pull, 0},
    {"e_topo_push_thread", style_etopo_push, 1},
    {"v_data_push_thread", style_data_push<false>, 0},
    {"e_data_push_thread", style_data_push<true>, 1},
    {"uf_hook_compress", style_uf, 0},
    {"uf_afforest", style_afforest, 0},
    {"v_topo_push_thread_determ", style_determ<false, kThread>, 0},
    {"v_topo_push_wave_determ", style_determ<false, kWaveG>, 0},
    {"v_topo_pull_thread_determ", style_determ<true, kThread>, 0},
    {"v_topo_pull_wave_determ", style_determ<true, kWaveG>, 0},
};

}  // namespace

const StyleEntry* cc_styles(int* count)
{
  *count = (int)(sizeof(kStyles) / sizeof(kStyles[0]));
  return kStyles;
}

}  // namespace igx
