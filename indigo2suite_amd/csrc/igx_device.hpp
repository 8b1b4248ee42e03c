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

__device__ __forceinline__ long long wave_min_ll(long long v)
{
  #pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) v = min(v, (long long)__shfl_xor(v, d, kWave));
  return v;
}

__device__ __forceinline__ long long wave_max_ll(long long v)
{
  #pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) v = max(v, (long long)__shfl_xor(v, d, kWave));
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
  unsigned pad0;
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
  int push_cut;          // SSSP: a push round relaxes only frontier vertices below this distance (0: all)
  int seeded;            // SSSP: the run has had its partial push round
  int force_all;         // SSSP: the next pull reads every neighbour (set after the partial push round)
  int active;            // workgroups that take part in the next round (0: all); round_finalize reduces that many
};

// Round epilogue.  Ending a round inside the round kernel ("the last block to
// arrive finalises") needs a device-scope fence and an atomic per block; on
// MI355X the fences alone cost ~30 us per 2048-block launch (each one writes
// back the XCD's L2), and same-address atomics serialise at ~11 ns each.  So a
// round is two launches instead: the round kernel leaves per-block partial
// statistics with plain stores (no fence: the kernel boundary publishes them),
// and a one-workgroup finalize kernel reduces them in block order
// (deterministic, also for the FP64 residual) and updates the control block.
//
// Queue appends still need counters: those are spread over kGroups slots on
// their own 4 KiB pages (workgroup b appends to slot b % kGroups, which is also
// the XCD the hardware deals it to).  Slots and partials follow the control
// block in the same allocation.
constexpr int kGroups = 8;
constexpr int kCtlBytes = 4096;
constexpr int kStats = 5;        // integer partials per block
constexpr int kStatMin = 4;      // the one reduced by minimum instead of sum
constexpr int kStatsD = 2;       // FP64 partials per block
constexpr int kMaxParts = 8192;  // largest grid of a round kernel

struct alignas(4096) GroupSlot {
  int qcount[2];  // queue segment sizes, indexed by round parity
  int tiles;      // next unclaimed tile of the group's share (pull rounds)
  int pad[1021];
};
static_assert(sizeof(Ctl) <= kCtlBytes, "control block exceeds its page");
constexpr size_t kPartsOffset = kCtlBytes + kGroups * sizeof(GroupSlot);
constexpr size_t kCtlAllocBytes = kPartsOffset + (size_t)(kStats + kStatsD) * kMaxParts * 8;

__device__ __forceinline__ GroupSlot* ctl_groups(Ctl* ctl)
{
  return reinterpret_cast<GroupSlot*>(reinterpret_cast<char*>(ctl) + kCtlBytes);
}
__device__ __forceinline__ long long* ctl_parts(Ctl* ctl)
{
  return reinterpret_cast<long long*>(reinterpret_cast<char*>(ctl) + kPartsOffset);
}
__device__ __forceinline__ double* ctl_parts_d(Ctl* ctl)
{
  return reinterpret_cast<double*>(ctl_parts(ctl) + (size_t)kStats * kMaxParts);
}
__device__ __forceinline__ int my_group() { return (int)(blockIdx.x % kGroups); }

// Round kernel side: every block of a launch that did not exit on `done` calls
// this once, with all threads, and stores its sums of a..d and its minimum of
// e (integer slots 0..3 and kStatMin).  lds: kWavesPerBlock * kStats slots.
__device__ __forceinline__ void round_partials(Ctl* ctl, long long a, long long b, long long c, long long d,
                                               long long* lds, long long e = LLONG_MAX)
{
  a = wave_sum_ll(a);
  b = wave_sum_ll(b);
  c = wave_sum_ll(c);
  d = wave_sum_ll(d);
  e = wave_min_ll(e);
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) {
    lds[w * kStats + 0] = a;
    lds[w * kStats + 1] = b;
    lds[w * kStats + 2] = c;
    lds[w * kStats + 3] = d;
    lds[w * kStats + kStatMin] = e;
  }
  __syncthreads();
  if ((int)threadIdx.x < kStats) {
    const bool is_min = (int)threadIdx.x == kStatMin;
    long long t = is_min ? LLONG_MAX : 0;
    for (int i = 0; i < (int)(blockDim.x / kWave); i++) {
      const long long x = lds[i * kStats + threadIdx.x];
      t = is_min ? min(t, x) : t + x;
    }
    ctl_parts(ctl)[(size_t)threadIdx.x * kMaxParts + blockIdx.x] = t;
  }
}

// Integer partial `slot` of this block, summed over its threads (all threads
// call; lds: kWavesPerBlock); leaves the other slots alone.
__device__ __forceinline__ void round_partial_ll(Ctl* ctl, int slot, long long v, long long* lds)
{
  v = wave_sum_ll(v);
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) lds[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long t = 0;
    for (int i = 0; i < (int)(blockDim.x / kWave); i++) t += lds[i];
    ctl_parts(ctl)[(size_t)slot * kMaxParts + blockIdx.x] = t;
  }
}

// FP64 partial `slot` of this block (all threads call; lds: kWavesPerBlock).
__device__ __forceinline__ void round_partial_d(Ctl* ctl, int slot, double v, double* lds)
{
  v = wave_sum_d(v);
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) lds[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < (int)(blockDim.x / kWave); i++) t += lds[i];
    ctl_parts_d(ctl)[(size_t)slot * kMaxParts + blockIdx.x] = t;
  }
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

struct RoundTotals {
  long long s[kStats];
  double d[kStatsD];
};

constexpr int kFinBlock = 1024;  // 16 waves: two partials per thread for a 2048-block round

// Finalize kernel (one workgroup of kFinBlock threads): reduces the partials
// of the nparts blocks of the round kernel for the statistics in `stat_mask`
// (bit k: integer k, summed, or the minimum for kStatMin; bit kStats + j:
// FP64 j, summed), then runs fin(ctl, totals) on one thread.  All statistics
// are reduced in one pass: every thread loads its strided partials of each
// (independent loads), the waves reduce by shuffles and thread 0 adds the
// waves in order, so FP64 totals are the same on every run.  Does nothing once
// the run is done.
template <typename Fin>
__global__ void __launch_bounds__(kFinBlock) round_finalize(Ctl* ctl, int nparts, int stat_mask, Fin fin)
{
  constexpr int kFinWaves = kFinBlock / kWave;
  __shared__ long long s_l[kFinWaves][kStats];
  __shared__ double s_d[kFinWaves][kStatsD];
  if (round_done(ctl)) return;
  long long acc[kStats];
  double accd[kStatsD];
  #pragma unroll
  for (int k = 0; k < kStats; k++) acc[k] = (k == kStatMin) ? LLONG_MAX : 0;
  #pragma unroll
  for (int k = 0; k < kStatsD; k++) accd[k] = 0.0;
  const long long* pl = ctl_parts(ctl);
  const double* pd = ctl_parts_d(ctl);
  // a round run by fewer workgroups (Ctl::active) left that many partials
  if (ctl->active > 0 && ctl->active < nparts) nparts = ctl->active;
  for (int b0 = threadIdx.x; b0 < nparts; b0 += 2 * kFinBlock) {
    const int b1 = b0 + kFinBlock;  // two partials per statistic, loaded together
    #pragma unroll
    for (int k = 0; k < kStats; k++) {
      if ((stat_mask >> k) & 1) {
        const bool is_min = k == kStatMin;
        const long long x0 = pl[(size_t)k * kMaxParts + b0];
        const long long x1 = (b1 < nparts) ? pl[(size_t)k * kMaxParts + b1] : (is_min ? LLONG_MAX : 0);
        acc[k] = is_min ? min(acc[k], min(x0, x1)) : acc[k] + x0 + x1;
      }
    }
    #pragma unroll
    for (int k = 0; k < kStatsD; k++) {
      if ((stat_mask >> (kStats + k)) & 1) {
        const double x0 = pd[(size_t)k * kMaxParts + b0];
        const double x1 = (b1 < nparts) ? pd[(size_t)k * kMaxParts + b1] : 0.0;
        accd[k] += x0 + x1;
      }
    }
  }
  const int w = threadIdx.x / kWave;
  #pragma unroll
  for (int k = 0; k < kStats; k++) {
    const long long x = (k == kStatMin) ? wave_min_ll(acc[k]) : wave_sum_ll(acc[k]);
    if (lane_id() == 0) s_l[w][k] = x;
  }
  #pragma unroll
  for (int k = 0; k < kStatsD; k++) {
    const double x = wave_sum_d(accd[k]);
    if (lane_id() == 0) s_d[w][k] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    RoundTotals tot;
    for (int k = 0; k < kStats; k++) {
      long long t = (k == kStatMin) ? LLONG_MAX : 0;
      for (int i = 0; i < kFinWaves; i++) t = (k == kStatMin) ? min(t, s_l[i][k]) : t + s_l[i][k];
      tot.s[k] = ((stat_mask >> k) & 1) ? t : 0;
    }
    for (int k = 0; k < kStatsD; k++) {
      double t = 0.0;
      for (int i = 0; i < kFinWaves; i++) t += s_d[i][k];
      tot.d[k] = ((stat_mask >> (kStats + k)) & 1) ? t : 0.0;
    }
    fin(ctl, tot);
  }
}


// finalizers of the two common round rules
struct FinFlag {
  __device__ void operator()(Ctl* ctl, const RoundTotals&) const { finalize_flag_round(ctl); }
};
struct FinWorklist {
  __device__ void operator()(Ctl* ctl, const RoundTotals&) const { finalize_worklist_round(ctl); }
};

// Visiting order for one-wave-per-vertex loops: k -> (k * mult) mod n with
// mult coprime to n is a permutation of [0, n).  R-MAT's hubs sit at ids with
// many zero bits, so a plain round-robin (v = wave + k * waves) hands dozens of
// them to the same wave; the multiplicative order spreads them out.
__device__ __forceinline__ int spread_vertex(int k, int n, unsigned mult)
{
  return (int)(((unsigned long long)(unsigned)k * mult) % (unsigned long long)n);
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
