// Frontier machinery shared by the data-driven kernels (BFS, SSSP).
//
// A frontier is a queue of edge chunks: each entry names a vertex and a range
// of at most kChunk of its edges.  Vertices with more edges are split into
// several entries when they are pushed, so the expansion below can balance
// hubs (R-MAT vertex 0 has ~1.6e5 neighbours at scale 22) across many waves.
//
// Expansion is wave-level merge path: a wave loads 64 entries, scans their
// lengths, and then walks the concatenated edge space 64 edges at a time;
// every lane finds its entry by binary search over the 64 exclusive prefixes
// kept in LDS.  All lanes stay busy regardless of the degree mix.
#pragma once

#include "igx_device.hpp"

namespace igx {

constexpr int kChunk = 256;

struct Entry {
  int v;    // vertex owning the edges
  int beg;  // first edge (index into nlist)
  int end;  // one past the last edge
  int pad;
};

__device__ __forceinline__ int chunks_of(int deg) { return (deg + kChunk - 1) / kChunk; }

// Push the chunks of vertex v (edge range [beg, end)) into q.  Lanes with
// exactly one chunk share one wave-aggregated atomic; larger vertices reserve
// their own slots.  Safe in divergent code.
__device__ __forceinline__ void push_vertex_chunks(Entry* __restrict__ q, int* counter, bool pred, int v, int beg,
                                                   int end)
{
  const int deg = end - beg;
  const int c = pred ? chunks_of(deg) : 0;
  const int slot = wave_push_slot(counter, c == 1);
  if (c == 1) q[slot] = Entry{v, beg, end, 0};
  uint64_t big = ballot(c > 1);
  if (big != 0) {
    // each multi-chunk vertex reserves its run of slots; then all active
    // lanes write that run together (a hub has hundreds of chunks)
    const int base = (c > 1) ? atomicAdd(counter, c) : 0;
    const uint64_t act = ballot(true);
    const int P = __popcll(act);
    const int r = mbcnt(act);
    while (big != 0) {
      const int L = __ffsll((unsigned long long)big) - 1;
      big &= big - 1;
      const int bv = __shfl(v, L, kWave), bb = __shfl(beg, L, kWave), be = __shfl(end, L, kWave);
      const int bbase = __shfl(base, L, kWave), bc = __shfl(c, L, kWave);
      for (int k = r; k < bc; k += P) {
        const int b = bb + k * kChunk;
        q[bbase + k] = Entry{bv, b, (b + kChunk < be) ? b + kChunk : be, 0};
      }
    }
  }
}

// Walk all edges of the queue entries [0, qsize) with the calling wave among
// `nwaves` waves (wave index `gw`).  lds_excl / lds_ent must point at this
// wave's 64-slot LDS arrays.  visit(ent, e) is called once per edge; all lanes
// of the wave call this function together (the loops are wave-uniform).
// Entries per wave group: 64 when the queue is long; fewer (a power of two)
// when it is short, so that a handful of hub chunks still spreads over all
// waves instead of over qsize / 64 of them.
__device__ __forceinline__ int group_entries(int qsize, int nwaves)
{
  if (qsize >= nwaves * kWave) return kWave;
  const int per = (qsize + nwaves - 1) / nwaves;
  int g = 1;
  while (g < per) g <<= 1;
  return g;
}

template <typename F>
__device__ __forceinline__ void wave_expand(const Entry* __restrict__ q, int qsize, int gw, int nwaves,
                                            int* lds_excl, Entry* lds_ent, F visit)
{
  const int lane = lane_id();
  const int G = group_entries(qsize, nwaves);
  for (int gbase = gw * G; gbase < qsize; gbase += nwaves * G) {
    const int idx = gbase + lane;
    Entry ent = Entry{0, 0, 0, 0};
    if (lane < G && idx < qsize) ent = q[idx];
    const int len = ent.end - ent.beg;
    const int incl = wave_incl_scan(len);
    const int total = wave_bcast(incl, kWave - 1);
    lds_excl[lane] = incl - len;
    lds_ent[lane] = ent;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    for (int tb = 0; tb < total; tb += kWave) {
      const int t = tb + lane;
      if (t < total) {
        // last k with excl[k] <= t (zero-length entries share the next prefix)
        int lo = 0, hi = kWave;
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (lds_excl[mid] <= t) lo = mid;
          else hi = mid;
        }
        const Entry& en = lds_ent[lo];
        visit(en, en.beg + (t - lds_excl[lo]));
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
}

}  // namespace igx
