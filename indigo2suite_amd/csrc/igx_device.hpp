// Device-side building blocks shared by the Indigo2 kernels on gfx950.
//
// * wave64 helpers: lane id, 64-bit ballots, exclusive bit counts (mbcnt)
// * wave-aggregated worklist push: one global atomic per wave, not per lane
// * the per-run control block and the "last block finishes the round" step
//   that lets the host enqueue several rounds without a round trip per round
#pragma once

#include <hip/hip_runtime.h>
#include <climits>
#include <cstdint>

namespace igx {

constexpr int kWave = 64;
constexpr int kBlock = 256;        // 4 waves per workgroup
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kInf = INT_MAX;

// ------------------------------------------------------------ wave64 ------
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ uint64_t ballot(bool p) { return (uint64_t)__ballot(p); }

// number of set bits of `mask` in lanes below this lane
__device__ __forceinline__ int mbcnt(uint64_t mask)
{
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

template <typename T>
__device__ __forceinline__ T wave_bcast(T v, int src_lane) { return __shfl(v, src_lane, kWave); }

// inclusive prefix sum over the wave (Hillis-Steele with shuffles)
__device__ __forceinline__ int wave_incl_scan(int v)
{
  const int lane = lane_id();
  #pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const int t = __shfl_up(v, d, kWave);
    if (lane >= d) v += t;
  }
  return v;
}

__device__ __forceinline__ long long wave_sum_ll(long long v)
{
  #pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
  return v;
}

__device__ __forceinline__ int wave_sum(int v)
{
  #pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
  return v;
}

__device__ __forceinline__ double wave_sum_d(double v)
{
  #pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
  return v;
}

// ----------------------------------------------------------- atomics ------
// Relaxed, agent-scope atomics (the "Atomic" style); seq_cst, system-scope
// variants model the reference's cuda::atomic default ("CudaAtomic").
__device__ __forceinline__ int atomic_min_relaxed(int* p, int v)
{
  return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int atomic_min_seqcst_sys(int* p, int v)
{
  return __hip_atomic_fetch_min(p, v, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_SYSTEM);
}
// CAS-loop minimum (the "CAS" style)
__device__ __forceinline__ int atomic_min_cas(int* p, int v)
{
  int old = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (old > v) {
    int expected = old;
    if (__hip_atomic_compare_exchange_strong(p, &expected, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT)) {
      break;
    }
    old = expected;
  }
  return old;
}
__device__ __forceinline__ int atomic_max_relaxed(int* p, int v)
{
  return __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T load_relaxed(const T* p)
{
  return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void store_relaxed(T* p, T v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// atomic-min flavours selectable per style
enum AtomicKind { kAtomicRelaxed = 0, kAtomicSeqCst = 1, kAtomicCas = 2 };

template <int K>
__device__ __forceinline__ int atomic_min_k(int* p, int v)
{
  if constexpr (K == kAtomicRelaxed) return atomic_min_relaxed(p, v);
  else if constexpr (K == kAtomicSeqCst) return atomic_min_seqcst_sys(p, v);
  else return atomic_min_cas(p, v);
}

// -------------------------------------------- wave-aggregated worklist ------
// Every active lane with `pred` gets a distinct slot; one atomicAdd per wave.
// Returns the slot, or -1 for lanes without pred.
__device__ __forceinline__ int wave_push_slot(int* counter, bool pred)
{
  const uint64_t mask = ballot(pred);
  if (mask == 0) return -1;
  const int leader = __ffsll((unsigned long long)mask) - 1;
  int base = 0;
  if (lane_id() == leader) base = atomicAdd(counter, __popcll(mask));
  base = wave_bcast(base, leader);
  return pred ? base + mbcnt(mask) : -1;
}

// Variable count per lane: lane reserves `cnt` consecutive slots.
__device__ __forceinline__ int wave_reserve(int* counter, int cnt)
{
  const int incl = wave_incl_scan(cnt);
  const int total = wave_bcast(incl, kWave - 1);
  int base = 0;
  if (total > 0) {
    // the last lane always exists in a full wave; use the highest active lane
    const uint64_t act = ballot(true);
    const int last = 63 - __clzll((long long)act);
    if (lane_id() == last) base = atomicAdd(counter, total);
    base = wave_bcast(base, last);
  }
  return base + incl - cnt;
}

// --------------------------------------------------------- control ------
// One control block per run.  `done` is checked by every round kernel at entry
// so the host can enqueue rounds speculatively; the last block to finish a
// round decides whether another round is needed.
struct Ctl {
  int done;              // 1 once converged
  int iter;              // rounds completed
  int flag;              // "goagain" of the current round
  unsigned arrived;      // groups finished in the current round
  int wl_count[2];       // worklist sizes, indexed by round parity
  int mode;              // algorithm-specific (BFS: 0 top-down, 1 bottom-up)
  int level;             // BFS level being expanded
  long long edges_next;  // BFS: edges of the current frontier; SSSP: of the next one
  long long edges_unvisited;
  long long aux_ll[2];
  long long work_edges;     // edges processed by the round kernels of this run
  long long work_vertices;  // rows / queue entries processed by them
  double res;            // PR residual of the current round (FP64)
  double res_last;       // PR residual of the last completed round
  int n;                 // vertex count (for heuristics)
  int in_bits;           // BFS: the current frontier is held by its bitmap only
};

// Same-address atomics from thousands of workgroups serialise (one every
// ~15 ns on MI355X), so a round's arrivals, statistics and queue counters are
// spread over kGroups slots, each on its own 4 KiB page: workgroup b uses slot
// b % kGroups, which is also the XCD that the hardware deals it to.  The slots
// follow the control block in the same allocation.
constexpr int kGroups = 8;
constexpr int kCtlBytes = 4096;
constexpr int kStats = 4;

struct alignas(4096) GroupSlot {
  unsigned arrived;            // workgroups of this group finished in the round
  unsigned pad0[31];
  long long stat[kStats];      // per-round partial sums (reset by the finaliser)
  long long pad1[12];
  int qcount[2];               // queue segment sizes, indexed by round parity
  int pad2[30];
};
static_assert(sizeof(Ctl) <= kCtlBytes, "control block exceeds its page");
constexpr size_t kCtlAllocBytes = kCtlBytes + kGroups * sizeof(GroupSlot);

__device__ __forceinline__ GroupSlot* ctl_groups(Ctl* ctl)
{
  return reinterpret_cast<GroupSlot*>(reinterpret_cast<char*>(ctl) + kCtlBytes);
}
__device__ __forceinline__ int my_group() { return (int)(blockIdx.x % kGroups); }

// Block-level epilogue of one round.  All threads of the block must call it.
// Arrival is two-level: a block counts in its group's slot; the last block of
// a group counts in ctl->arrived; `finalize` runs on exactly one thread of the
// last block overall, after device-scope fences, and must reset per-round
// fields it uses (group stats: stat_take).
template <typename F>
__device__ __forceinline__ void finish_round(Ctl* ctl, F finalize)
{
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned nb = gridDim.x;
    const unsigned grp = blockIdx.x % kGroups;
    const unsigned ngroups = nb < (unsigned)kGroups ? nb : (unsigned)kGroups;
    const unsigned gsize = (nb - 1 - grp) / kGroups + 1;
    GroupSlot* gs = ctl_groups(ctl) + grp;
    __threadfence();
    if (atomicAdd(&gs->arrived, 1u) == gsize - 1) {
      gs->arrived = 0;  // every block of the group has arrived: reusable next round
      __threadfence();
      if (atomicAdd(&ctl->arrived, 1u) == ngroups - 1) {
        __threadfence();
        finalize();
        ctl->arrived = 0;
        __threadfence();
      }
    }
  }
}

// Add a block's value to statistic k of its group (one atomic per block).
__device__ __forceinline__ void block_stat_add(Ctl* ctl, int k, long long v, long long* lds4)
{
  v = wave_sum_ll(v);
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) lds4[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long t = 0;
    for (int i = 0; i < (int)(blockDim.x / kWave); i++) t += lds4[i];  // lds4: one slot per wave
    if (t != 0) atomicAdd((unsigned long long*)&ctl_groups(ctl)[my_group()].stat[k], (unsigned long long)t);
  }
  __syncthreads();
}

// Finaliser side: the round's total of statistic k, clearing the partials.
__device__ __forceinline__ long long stat_take(Ctl* ctl, int k)
{
  GroupSlot* gs = ctl_groups(ctl);
  long long t = 0;
  for (int g = 0; g < kGroups; g++) {
    t += load_relaxed(&gs[g].stat[k]);
    gs[g].stat[k] = 0;
  }
  return t;
}

// the common topology-driven rule: stop after a round without updates
__device__ __forceinline__ void finalize_flag_round(Ctl* ctl)
{
  const int f = load_relaxed(&ctl->flag);
  ctl->iter += 1;
  if (f == 0) ctl->done = 1;
  ctl->flag = 0;
}

// the common data-driven rule: stop after a round that pushed nothing
__device__ __forceinline__ void finalize_worklist_round(Ctl* ctl)
{
  const int p = ctl->iter & 1;
  const int produced = load_relaxed(&ctl->wl_count[p ^ 1]);
  ctl->wl_count[p] = 0;  // becomes the output counter of the next round
  ctl->iter += 1;
  if (produced == 0) ctl->done = 1;
}

__device__ __forceinline__ bool round_done(const Ctl* ctl)
{
  return *(volatile const int*)&ctl->done != 0;
}

// MIS priority: the reference's integer hash of (v + 712313887), compared as
// unsigned (mis_vertex_omp.h:127-132 and codes/omp MIS init).
__device__ __forceinline__ uint32_t mis_priority(uint32_t v)
{
  uint32_t val = v + 712313887u;
  val = ((val >> 16) ^ val) * 0x45d9f3bu;
  val = ((val >> 16) ^ val) * 0x45d9f3bu;
  return (val >> 16) ^ val;
}

}  // namespace igx
