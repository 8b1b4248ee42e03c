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
// exactly one chunk share one wave-aggregated atomic, lanes with several
// share a second one (their runs are laid out in lane order), then all active
// lanes write the multi-chunk runs together (a hub has hundreds of chunks).
// Safe in divergent code.
__device__ __forceinline__ void push_vertex_chunks(Entry* __restrict__ q, int* counter, bool pred, int v, int beg,
                                                   int end)
{
  const int deg = end - beg;
  const int c = pred ? chunks_of(deg) : 0;
  const int slot = wave_push_slot(counter, c == 1);
  if (c == 1) q[slot] = Entry{v, beg, end, 0};
  uint64_t big = ballot(c > 1);
  if (big != 0) {
    // lane-ordered offsets of the multi-chunk runs (serial over those lanes)
    const int lane = lane_id();
    int total = 0, off = 0;
    for (uint64_t b = big; b != 0; b &= b - 1) {
      const int L = __ffsll((unsigned long long)b) - 1;
      if (lane == L) off = total;
      total += __shfl(c, L, kWave);
    }
    const int leader = __ffsll((unsigned long long)big) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, total);
    base = __shfl(base, leader, kWave) + off;
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

// ----------------------------------------------------- segmented queues ----
// A chunk queue split into kGroups segments of `cap` entries each.  Workgroup
// b appends to segment b % kGroups, whose size lives in that group's control
// slot, so pushes from the whole grid spread over kGroups counters; readers
// walk the concatenation of the segments.  Each segment can hold a full
// round's output on its own.
struct SegQ {
  Entry* base;
  long long cap;
};

__device__ __forceinline__ Entry* segq_mine(const SegQ& q) { return q.base + (long long)my_group() * q.cap; }
__device__ __forceinline__ int* segq_counter(Ctl* ctl, int parity, int seg)
{
  return &ctl_groups(ctl)[seg].qcount[parity];
}

// prefix offsets of the kGroups segment sizes (kept in registers by the reader)
struct SegPrefix {
  int pre[kGroups + 1];
};
__device__ __forceinline__ SegPrefix segq_prefix(Ctl* ctl, int parity)
{
  SegPrefix sp;
  sp.pre[0] = 0;
  #pragma unroll
  for (int g = 0; g < kGroups; g++) sp.pre[g + 1] = sp.pre[g] + load_relaxed(segq_counter(ctl, parity, g));
  return sp;
}
__device__ __forceinline__ Entry segq_at(const SegQ& q, const SegPrefix& sp, int idx)
{
  // constant indices only, so that the prefixes stay in registers
  int s = 0, base = 0;
  #pragma unroll
  for (int g = 1; g < kGroups; g++) {
    if (idx >= sp.pre[g]) {
      s = g;
      base = sp.pre[g];
    }
  }
  return q.base[(long long)s * q.cap + (idx - base)];
}

// Visit every 64-vertex tile [v0, v0 + 64) that holds at least one set bit of
// the bitmap, one wave per tile: body(v0, bit) with `bit` this lane's vertex
// v0 + lane.  A wave loads 64 words (2048 vertices) at once and skips empty
// tiles without further loads, so a sparse frontier costs one coalesced pass
// over the bitmap.  All lanes of the wave call this together; bits at or past
// n must be clear.
template <typename F>
__device__ __forceinline__ void for_each_bitmap_tile(const unsigned* __restrict__ bits, int n, int gw, int nwaves,
                                                     F body)
{
  const int lane = lane_id();
  const int nwords = (n + 31) >> 5;
  for (int wb = gw * kWave; wb < nwords; wb += nwaves * kWave) {  // uniform
    const int wi = wb + lane;
    const unsigned word = (wi < nwords) ? bits[wi] : 0u;
    const uint64_t nz = ballot(word != 0u);
    uint64_t pair = (nz | (nz >> 1)) & 0x5555555555555555ull;  // even j: word j or j + 1 non-empty
    while (pair != 0) {
      const int j = __ffsll((unsigned long long)pair) - 1;
      pair &= pair - 1;
      const unsigned w0 = __shfl(word, j, kWave), w1 = __shfl(word, j + 1, kWave);
      const unsigned mine = (lane < 32) ? w0 : w1;
      body((wb + j) * 32, ((mine >> (lane & 31)) & 1u) != 0u);
    }
  }
}

// Append the edge ids [beg, beg + len) of every lane to an edge worklist:
// one atomic per wave reserves the wave's total, then the 64 lanes write the
// concatenated ranges together (each lane finds its range by binary search
// over the exclusive prefixes in LDS), so a hub's ids go out 64 at a time in
// contiguous stores.  All lanes of the wave call this together; lanes with
// nothing to push pass len = 0.  lds_excl / lds_beg: this wave's 64-slot LDS
// rows.
__device__ __forceinline__ void wave_push_edge_ranges(int* __restrict__ out, int* counter, int beg, int len,
                                                      int* lds_excl, int* lds_beg)
{
  const int lane = lane_id();
  const int incl = wave_incl_scan(len);
  const int total = wave_bcast(incl, kWave - 1);
  if (total == 0) return;
  int base = 0;
  if (lane == kWave - 1) base = atomicAdd(counter, total);
  base = wave_bcast(base, kWave - 1);
  __builtin_amdgcn_wave_barrier();
  lds_excl[lane] = incl - len;
  lds_beg[lane] = beg;
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  for (int tb = 0; tb < total; tb += kWave) {
    const int t = tb + lane;
    if (t < total) {
      int lo = 0, hi = kWave;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (lds_excl[mid] <= t) lo = mid;
        else hi = mid;
      }
      out[base + t] = lds_beg[lo] + (t - lds_excl[lo]);
    }
  }
  __builtin_amdgcn_wave_barrier();
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

// Expansion over any indexable source of entries: fetch(idx) returns entry
// idx of [0, qsize); entries with end == beg are skipped at no cost.
template <typename Fetch, typename F>
__device__ __forceinline__ void wave_expand_with(int qsize, int gw, int nwaves, int* lds_excl, Entry* lds_ent,
                                                 Fetch fetch, F visit)
{
  const int lane = lane_id();
  const int G = group_entries(qsize, nwaves);
  for (int gbase = gw * G; gbase < qsize; gbase += nwaves * G) {
    const int idx = gbase + lane;
    Entry ent = Entry{0, 0, 0, 0};
    if (lane < G && idx < qsize) ent = fetch(idx);
    const int len = ent.end - ent.beg;
    const int incl = wave_incl_scan(len);
    const int total = wave_bcast(incl, kWave - 1);
    if (total == 0) continue;  // uniform
    __builtin_amdgcn_wave_barrier();
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

// The same walk with the entries dealt out round-robin: lane l of wave gw
// takes entry gw + (64 s + l) * nwaves in step s, so consecutive entries go
// to consecutive waves.  For a list in which the live entries come in long
// runs (the chunks of a few hub rows among the chunks of every long row),
// the consecutive grouping above hands a whole run to one wave; here a run
// of c chunks spreads over c waves.
template <typename Fetch, typename F>
__device__ __forceinline__ void wave_expand_strided(int qsize, int gw, int nwaves, int* lds_excl, Entry* lds_ent,
                                                    Fetch fetch, F visit)
{
  const int lane = lane_id();
  const long long step = (long long)kWave * nwaves;
  for (long long base = gw; base < qsize; base += step) {
    const long long idx = base + (long long)lane * nwaves;
    Entry ent = Entry{0, 0, 0, 0};
    if (idx < qsize) ent = fetch((int)idx);
    const int len = ent.end - ent.beg;
    const int incl = wave_incl_scan(len);
    const int total = wave_bcast(incl, kWave - 1);
    if (total == 0) continue;  // uniform
    __builtin_amdgcn_wave_barrier();
    lds_excl[lane] = incl - len;
    lds_ent[lane] = ent;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    for (int tb = 0; tb < total; tb += kWave) {
      const int t = tb + lane;
      if (t < total) {
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

template <typename F>
__device__ __forceinline__ void wave_expand(const Entry* __restrict__ q, int qsize, int gw, int nwaves,
                                            int* lds_excl, Entry* lds_ent, F visit)
{
  wave_expand_with(qsize, gw, nwaves, lds_excl, lds_ent, [&](int i) { return q[i]; }, visit);
}

}  // namespace igx
