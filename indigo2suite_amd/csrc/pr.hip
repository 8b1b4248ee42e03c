// PageRank (pull, Jacobi) on gfx950, following pr_cuda.h / [C-PRt]:
//   s0 = 1/n;  s_v <- (1 - d)/n + d * sum_{u in N(v)} s_u / deg(u),  d = 0.85
// with no dangling-mass redistribution, stopping when the L1 change of an
// iteration falls below eps or after max_iter iterations.
//
// Precision: scores and contributions are FP32; every neighbour sum is
// compensated (TwoSum per addition; partial sums are merged as (sum, error)
// pairs), and the per-iteration L1 residual that decides convergence is
// accumulated exactly: each row's |change| is computed in FP64 and rounded to
// a multiple of 2^-50, and those integers are summed in 64 bits, so the
// residual does not depend on which wave handled which row (below 1e-9 from
// the FP64 sum for a million rows).  Results are checked against an FP64 host
// oracle.
//
// Reference family: codes/cuda/pr-cuda (54 variants).
//   pull_thread_kahan   [C-PRt] contrib:41-47 + pull:49-70 (thread per vertex)
//   pull_wave_kahan     PR_CUDA_Pull_..._Warp siblings (one wave64 per vertex)
//   pull_block_kahan    PR_CUDA_Pull_..._Block siblings (one workgroup per vertex)
//   pull_balanced_kahan (no reference counterpart) rows classed by degree:
//                       small rows streamed in runs of consecutive rows and
//                       summed in edge order by their lane, longer rows split
//                       into chunks of at most kChunk edges (one wave each)
//                       whose compensated partial sums are merged in chunk
//                       order
//   push_fixed_{thread,wave,block}  [C-PRpush] Push siblings: contributions
//                       pushed to the neighbours' totals with 64-bit
//                       fixed-point atomic adds (exact, order-independent)
// The contribution of the next iteration (score / degree) is written by the
// same kernel that produces the score, so one pass per iteration remains.
#include "igx_internal.hpp"
#include "igx_frontier.hpp"

namespace igx {
namespace {

constexpr float kDamp = 0.85f;  // pr_cuda.h:43
constexpr int kSmallRow = 32;    // rows streamed in runs by pr_bal_main; longer ones by chunks
constexpr int kSlice = 512;      // edges per streamed slice
constexpr double kResScale = 1125899906842624.0;  // 2^50: fixed-point unit of the residual

struct Pair {
  float s;  // running sum
  float e;  // accumulated rounding error (true sum ~ s + e)
};

__device__ __forceinline__ void two_sum_add(Pair& p, float x)
{
  const float t = p.s + x;
  const float bp = t - p.s;
  const float err = (p.s - (t - bp)) + (x - bp);
  p.s = t;
  p.e += err;
}

__device__ __forceinline__ Pair pair_merge(Pair a, Pair b)
{
  Pair r;
  r.s = a.s + b.s;
  const float bp = r.s - a.s;
  const float err = (a.s - (r.s - bp)) + (b.s - bp);
  r.e = a.e + b.e + err;
  return r;
}

__device__ __forceinline__ Pair wave_merge(Pair p)
{
  #pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    Pair o;
    o.s = __shfl_xor(p.s, d, kWave);
    o.e = __shfl_xor(p.e, d, kWave);
    // merge in a lane-order-independent way: lower half first
    p = ((lane_id() & d) == 0) ? pair_merge(p, o) : pair_merge(o, p);
  }
  return p;
}

__global__ void __launch_bounds__(kBlock) pr_init(const int* __restrict__ nindex, int n, float* __restrict__ score,
                                                   float* __restrict__ contrib)
{
  const int v = blockIdx.x * kBlock + threadIdx.x;
  if (v < n) {
    const float s0 = 1.0f / (float)n;  // pr_cuda.h:133
    score[v] = s0;
    contrib[v] = s0 / (float)(nindex[v + 1] - nindex[v]);
  }
}

// finish one vertex: new score, next contribution, residual term (|change|
// in units of 2^-50, rounded to nearest)
__device__ __forceinline__ long long pr_finish(int v, Pair acc, float base, const int* __restrict__ nindex,
                                               float* __restrict__ score, float* __restrict__ contrib_next)
{
  const float sum = acc.s + acc.e;
  const float value = base + kDamp * sum;
  const float old = score[v];
  score[v] = value;
  contrib_next[v] = value / (float)(nindex[v + 1] - nindex[v]);
  return (long long)(fabs((double)value - (double)old) * kResScale + 0.5);
}


// End of an iteration (run by round_finalize): the L1 residual is the exact
// integer sum of the per-block partials (slot 0: rows scored by the main
// kernel, slot 1: rows merged from chunks), scaled back from 2^-50 units.
struct FinPr {
  double eps;
  int max_iter;
  __device__ void operator()(Ctl* ctl, const RoundTotals& t) const
  {
    const double r = (double)(t.s[0] + t.s[1]) / kResScale;
    ctl->res_last = r;
    ctl->iter += 1;
    if (r < eps || ctl->iter >= max_iter) ctl->done = 1;
  }
};

enum Gran { kThread = 0, kWaveG = 1, kBlockG = 2 };

template <int G>
__global__ void __launch_bounds__(kBlock) pr_pull(const int* __restrict__ nindex, const int* __restrict__ nlist,
                                                   int n, float base, float* __restrict__ score,
                                                   const float* __restrict__ contrib_cur,
                                                   float* __restrict__ contrib_next, Ctl* ctl, double eps,
                                                   int max_iter)
{
  __shared__ long long s_red[kWavesPerBlock];
  __shared__ Pair s_pair[kWavesPerBlock];
  if (round_done(ctl)) return;
  long long err = 0;
  if constexpr (G == kThread) {
    const int stride = gridDim.x * kBlock;
    for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
      Pair acc{0.0f, 0.0f};
      for (int i = nindex[v]; i < nindex[v + 1]; i++) two_sum_add(acc, contrib_cur[nlist[i]]);
      err += pr_finish(v, acc, base, nindex, score, contrib_next);
    }
  } else if constexpr (G == kWaveG) {
    const int waves = gridDim.x * kWavesPerBlock;
    const int lane = lane_id();
    for (int v = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; v < n; v += waves) {
      Pair acc{0.0f, 0.0f};
      for (int i = nindex[v] + lane; i < nindex[v + 1]; i += kWave) two_sum_add(acc, contrib_cur[nlist[i]]);
      acc = wave_merge(acc);
      if (lane == 0) err += pr_finish(v, acc, base, nindex, score, contrib_next);
    }
  } else {
    const int w = threadIdx.x / kWave;
    for (int v = blockIdx.x; v < n; v += gridDim.x) {
      Pair acc{0.0f, 0.0f};
      for (int i = nindex[v] + (int)threadIdx.x; i < nindex[v + 1]; i += kBlock) two_sum_add(acc, contrib_cur[nlist[i]]);
      acc = wave_merge(acc);
      if (lane_id() == 0) s_pair[w] = acc;
      __syncthreads();
      if (threadIdx.x == 0) {
        Pair t = s_pair[0];
        for (int k = 1; k < kWavesPerBlock; k++) t = pair_merge(t, s_pair[k]);
        err += pr_finish(v, t, base, nindex, score, contrib_next);
      }
      __syncthreads();
    }
  }
  round_partial_ll(ctl, 0, err, s_red);
}

// ------------------------------------------------------------- balanced ----
// Row classes are built once per run: every row longer than kSmallRow is cut
// into chunks of at most kChunk edges (its pieces in edge order, in a
// contiguous run of chunk slots reserved with one atomic per wave).  A row of
// one chunk is finished by the wave that sums it (Entry::pad = -1); the others
// leave compensated partial sums that pr_bal_big merges in slot order, so the
// result does not depend on the schedule.
struct BigRow {
  int v;
  int first;  // first chunk slot
  int count;  // number of chunks
};

__global__ void pr_zero(int* __restrict__ p, int count)
{
  for (int i = threadIdx.x; i < count; i += blockDim.x) p[i] = 0;
}

// counts[1]: rows of several chunks (big); counts[2]: chunks
__global__ void __launch_bounds__(kBlock) pr_classify(const int* __restrict__ nindex, int n,
                                                       int* __restrict__ counts, Entry* __restrict__ chunks,
                                                       BigRow* __restrict__ big)
{
  const int stride = gridDim.x * kBlock;
  for (int base = blockIdx.x * kBlock; base < n; base += stride) {
    const int v = base + threadIdx.x;
    int deg = 0, beg = 0, end = 0;
    if (v < n) {
      beg = nindex[v];
      end = nindex[v + 1];
      deg = end - beg;
    }
    const int c = (deg > kSmallRow) ? chunks_of(deg) : 0;
    const int first = wave_reserve(&counts[2], c);
    const int bs = wave_push_slot(&counts[1], c > 1);
    if (c > 1) big[bs] = BigRow{v, first, c};
    for (int k = 0; k < c; k++) {
      const int b = beg + k * kChunk;
      chunks[first + k] = Entry{v, b, (b + kChunk < end) ? b + kChunk : end, (c == 1) ? -1 : 0};
    }
  }
}

// One iteration's sums: first the chunks (one wave each, kChunk / 64 loads per
// lane issued together; chunks are at most kChunk edges, so a fixed
// round-robin share balances), then the small rows in 64-row tiles, also
// round robin.  Small rows:
// consecutive small rows form runs whose edges are contiguous, streamed kSlice
// at a time (independent loads, values parked in LDS); each lane then adds its
// own row's values in edge order.
__global__ void __launch_bounds__(kBlock) pr_bal_main(const int* __restrict__ nindex, const int* __restrict__ nlist,
                                                       int n, float base, float* __restrict__ score,
                                                       const float* __restrict__ contrib_cur,
                                                       float* __restrict__ contrib_next,
                                                       const int* __restrict__ counts,
                                                       const Entry* __restrict__ chunks, Pair* __restrict__ partial,
                                                       Ctl* ctl)
{
  __shared__ long long s_red[kWavesPerBlock];
  __shared__ float s_val[kWavesPerBlock][kSlice];
  if (round_done(ctl)) return;
  long long err = 0;
  const int waves = gridDim.x * kWavesPerBlock;
  const int gw = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  const int lane = lane_id();
  constexpr int kPer = kChunk / kWave;
  const int nch = counts[2];
  for (int k = gw; k < nch; k += waves) {  // uniform per wave
    const Entry en = chunks[k];
    int u[kPer];
    #pragma unroll
    for (int j = 0; j < kPer; j++) {
      const int i = en.beg + j * kWave + lane;
      u[j] = (i < en.end) ? nlist[i] : -1;
    }
    float x[kPer];
    #pragma unroll
    for (int j = 0; j < kPer; j++) x[j] = (u[j] >= 0) ? contrib_cur[u[j]] : 0.0f;
    Pair acc{0.0f, 0.0f};
    #pragma unroll
    for (int j = 0; j < kPer; j++) {
      if (u[j] >= 0) two_sum_add(acc, x[j]);
    }
    acc = wave_merge(acc);
    if (lane == 0) {
      if (en.pad < 0) err += pr_finish(en.v, acc, base, nindex, score, contrib_next);
      else partial[k] = acc;
    }
  }
  float* vbuf = s_val[threadIdx.x / kWave];
  for (int v0 = gw * kWave; v0 < n; v0 += waves * kWave) {  // uniform per wave
    const int v = v0 + lane;
    int rb = 0, re = 0;
    if (v < n) {
      rb = nindex[v];
      re = nindex[v + 1];
    }
    const bool small = (v < n) && (re - rb <= kSmallRow);
    Pair acc{0.0f, 0.0f};
    uint64_t todo = ballot(small);
    while (todo != 0) {  // one run of consecutive small rows
      const int first = __ffsll((unsigned long long)todo) - 1;
      const uint64_t from = todo >> first;
      const int cnt = (~from == 0) ? (kWave - first) : (__ffsll((unsigned long long)~from) - 1);
      const int last = first + cnt - 1;
      todo &= (cnt + first >= kWave) ? ((1ull << first) - 1) : ~(((1ull << cnt) - 1) << first);
      const int r0 = __shfl(rb, first, kWave), r1 = __shfl(re, last, kWave);
      const bool mine = lane >= first && lane <= last;
      for (int c0 = r0; c0 < r1; c0 += kSlice) {
        const int c1 = (r1 - c0 < kSlice) ? r1 : c0 + kSlice;
        int uu[kSlice / kWave];
        #pragma unroll
        for (int k = 0; k < kSlice / kWave; k++) {
          const int e = c0 + k * kWave + lane;
          uu[k] = (e < c1) ? nlist[e] : -1;
        }
        float xx[kSlice / kWave];
        #pragma unroll
        for (int k = 0; k < kSlice / kWave; k++) xx[k] = (uu[k] >= 0) ? contrib_cur[uu[k]] : 0.0f;
        __builtin_amdgcn_wave_barrier();
        #pragma unroll
        for (int k = 0; k < kSlice / kWave; k++) vbuf[k * kWave + lane] = xx[k];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        if (mine) {
          const int lo = (rb > c0) ? rb : c0;
          const int hi = (re < c1) ? re : c1;
          for (int i = lo; i < hi; i++) two_sum_add(acc, vbuf[i - c0]);
        }
        __builtin_amdgcn_wave_barrier();
      }
    }
    if (small) err += pr_finish(v, acc, base, nindex, score, contrib_next);
  }
  round_partial_ll(ctl, 0, err, s_red);
}

__global__ void __launch_bounds__(kBlock) pr_bal_big(const int* __restrict__ nindex, float base,
                                                      float* __restrict__ score, float* __restrict__ contrib_next,
                                                      const BigRow* __restrict__ big, const int* __restrict__ counts,
                                                      const Pair* __restrict__ partial, Ctl* ctl)
{
  __shared__ long long s_red[kWavesPerBlock];
  if (round_done(ctl)) return;
  long long err = 0;
  const int nb = counts[1];
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < nb; i += stride) {
    const BigRow b = big[i];
    Pair acc{0.0f, 0.0f};
    for (int k = 0; k < b.count; k++) acc = pair_merge(acc, partial[b.first + k]);
    err += pr_finish(b.v, acc, base, nindex, score, contrib_next);
  }
  round_partial_ll(ctl, 1, err, s_red);
}

// ----------------------------------------------------------------- push ----
// [C-PRpush] contrib():41-48, push():50-63, compute():65-77: every vertex adds
// its contribution score / degree to each neighbour's incoming total, then
// every vertex turns its total into its new score.  The reference adds floats
// atomically, so its totals depend on the order in which the atomics land.
// Here each contribution is rounded once to a multiple of 2^-60 (a float
// contribution of at least 2^-37 already is one, so typical contributions are
// not rounded at all) and added as a 64-bit integer: integer addition is
// associative, so every total, and with it the iteration count, is the same
// whatever the schedule, and it is the exact sum of the rounded contributions.
// A total is at most the sum of all scores (at most 1 = 2^60 units), far from
// the int64 limit.  The compute pass reads each total once, clears it for the
// next iteration, and writes the score, the next contribution and the
// residual exactly like the pull styles.
constexpr double kFixScale = 1152921504606846976.0;  // 2^60: fixed-point unit of the totals

__device__ __forceinline__ unsigned long long pr_fixed(float c)
{
  return (unsigned long long)((double)c * kFixScale + 0.5);
}

template <int G>
__global__ void __launch_bounds__(kBlock) pr_push(const int* __restrict__ nindex, const int* __restrict__ nlist,
                                                   int n, const float* __restrict__ contrib,
                                                   unsigned long long* __restrict__ incoming, Ctl* ctl)
{
  if (round_done(ctl)) return;
  if constexpr (G == kThread) {
    const int stride = gridDim.x * kBlock;
    for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
      const int beg = nindex[v], end = nindex[v + 1];
      if (beg == end) continue;  // a vertex of degree 0 has an infinite contribution and no edge
      const unsigned long long q = pr_fixed(contrib[v]);
      for (int i = beg; i < end; i++) atomicAdd(&incoming[nlist[i]], q);
    }
  } else if constexpr (G == kWaveG) {
    const int waves = gridDim.x * kWavesPerBlock;
    const int lane = lane_id();
    for (int v = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; v < n; v += waves) {
      const int beg = nindex[v], end = nindex[v + 1];
      if (beg == end) continue;
      const unsigned long long q = pr_fixed(contrib[v]);
      for (int i = beg + lane; i < end; i += kWave) atomicAdd(&incoming[nlist[i]], q);
    }
  } else {
    for (int v = blockIdx.x; v < n; v += gridDim.x) {
      const int beg = nindex[v], end = nindex[v + 1];
      if (beg == end) continue;
      const unsigned long long q = pr_fixed(contrib[v]);
      for (int i = beg + (int)threadIdx.x; i < end; i += kBlock) atomicAdd(&incoming[nlist[i]], q);
    }
  }
}

__global__ void __launch_bounds__(kBlock) pr_push_compute(const int* __restrict__ nindex, int n, float base,
                                                           float* __restrict__ score, float* __restrict__ contrib,
                                                           unsigned long long* __restrict__ incoming, Ctl* ctl)
{
  __shared__ long long s_red[kWavesPerBlock];
  if (round_done(ctl)) return;
  long long err = 0;
  const int stride = gridDim.x * kBlock;
  for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
    const float total = (float)((double)incoming[v] / kFixScale);
    incoming[v] = 0;
    err += pr_finish(v, Pair{total, 0.0f}, base, nindex, score, contrib);
  }
  round_partial_ll(ctl, 0, err, s_red);
}

__global__ void __launch_bounds__(kBlock) pr_zero_ull(unsigned long long* __restrict__ p, int n)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ------------------------------------------------------------ host side ----
int pr_buffers(const igx_graph* g, float** c0, float** c1)
{
  void *p0 = nullptr, *p1 = nullptr;
  IGX_TRY(ws_get(WS_F, sizeof(float) * (size_t)g->nodes, &p0));
  IGX_TRY(ws_get(WS_G, sizeof(float) * (size_t)g->nodes, &p1));
  *c0 = (float*)p0;
  *c1 = (float*)p1;
  return IGX_OK;
}

template <int G>
int style_pull(const igx_graph* g, int, void* out, const void*, igx_run_info* info)
{
  Runtime& R = rt();
  float* score = (float*)out;
  float *c0 = nullptr, *c1 = nullptr;
  IGX_TRY(pr_buffers(g, &c0, &c1));
  const PrParams P = pr_params();
  const float base = (1.0f - kDamp) / (float)g->nodes;  // pr_cuda.h:140
  IGX_TRY(timer_start());
  IGX_TRY(ctl_reset(g->nodes));
  hipLaunchKernelGGL(pr_init, dim3(blocks_for(g->nodes, kBlock)), dim3(kBlock), 0, R.stream, g->nindex, g->nodes,
                     score, c0);
  int rounds = 0, launches = 0;
  const int rc = run_rounds(
      [&](long long r) {
        float* cur = (r & 1) ? c1 : c0;
        float* nxt = (r & 1) ? c0 : c1;
        hipLaunchKernelGGL((pr_pull<G>), dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nindex, g->nlist,
                           g->nodes, base, score, cur, nxt, R.d_ctl, P.eps, P.max_iter);
        launch_finalize(R.persist_blocks, 0x1, FinPr{P.eps, P.max_iter});
      },
      P.max_iter, &rounds, &launches);
  IGX_TRY(timer_stop(&info->seconds));
  info->iterations = rounds;
  info->launches = launches;
  info->residual = R.h_ctl->res_last;
  return rc == IGX_ERR_LIMIT ? IGX_OK : rc;
}

int style_balanced(const igx_graph* g, int, void* out, const void*, igx_run_info* info)
{
  Runtime& R = rt();
  float* score = (float*)out;
  float *c0 = nullptr, *c1 = nullptr;
  IGX_TRY(pr_buffers(g, &c0, &c1));
  const int n = g->nodes;
  // rows above kSmallRow: at most edges / (kSmallRow + 1) of them, each with
  // at most deg / kChunk + 1 chunks
  const size_t nrows = (size_t)g->edges / (kSmallRow + 1) + 64;
  const size_t nchunks = (size_t)g->edges / kChunk + nrows;
  void *pcnt = nullptr, *pch = nullptr, *pbig = nullptr, *ppart = nullptr;
  IGX_TRY(ws_get(WS_B, sizeof(int) * 64, &pcnt));
  IGX_TRY(ws_get(WS_D, sizeof(Entry) * nchunks, &pch));
  IGX_TRY(ws_get(WS_C, sizeof(BigRow) * nrows, &pbig));
  IGX_TRY(ws_get(WS_H, sizeof(Pair) * nchunks, &ppart));
  const PrParams P = pr_params();
  const float base = (1.0f - kDamp) / (float)n;
  IGX_TRY(timer_start());
  IGX_TRY(ctl_reset(n));
  hipLaunchKernelGGL(pr_zero, dim3(1), dim3(64), 0, R.stream, (int*)pcnt, 64);
  hipLaunchKernelGGL(pr_init, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, R.stream, g->nindex, n, score, c0);
  hipLaunchKernelGGL(pr_classify, dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nindex, n, (int*)pcnt,
                     (Entry*)pch, (BigRow*)pbig);
  int rounds = 0, launches = 0;
  const int rc = run_rounds(
      [&](long long r) {
        float* cur = (r & 1) ? c1 : c0;
        float* nxt = (r & 1) ? c0 : c1;
        ktimer_begin("pr_bal_main");
        hipLaunchKernelGGL(pr_bal_main, dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nindex, g->nlist, n,
                           base, score, cur, nxt, (const int*)pcnt, (const Entry*)pch, (Pair*)ppart, R.d_ctl);
        ktimer_end("pr_bal_main");
        hipLaunchKernelGGL(pr_bal_big, dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nindex, base, score,
                           nxt, (const BigRow*)pbig, (const int*)pcnt, (const Pair*)ppart, R.d_ctl);
        launch_finalize(R.persist_blocks, 0x3, FinPr{P.eps, P.max_iter});
      },
      P.max_iter, &rounds, &launches);
  IGX_TRY(timer_stop(&info->seconds));
  info->iterations = rounds;
  info->launches = launches;
  info->residual = R.h_ctl->res_last;
  return rc == IGX_ERR_LIMIT ? IGX_OK : rc;
}

// One contribution buffer suffices: the push pass reads it and the compute
// pass, a separate launch, rewrites it.
template <int G>
int style_push(const igx_graph* g, int, void* out, const void*, igx_run_info* info)
{
  Runtime& R = rt();
  float* score = (float*)out;
  const int n = g->nodes;
  void *pc = nullptr, *pin = nullptr;
  IGX_TRY(ws_get(WS_F, sizeof(float) * (size_t)n, &pc));
  IGX_TRY(ws_get(WS_H, sizeof(unsigned long long) * (size_t)n, &pin));
  float* contrib = (float*)pc;
  unsigned long long* incoming = (unsigned long long*)pin;
  const PrParams P = pr_params();
  const float base = (1.0f - kDamp) / (float)n;
  IGX_TRY(timer_start());
  IGX_TRY(ctl_reset(n));
  hipLaunchKernelGGL(pr_init, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, R.stream, g->nindex, n, score, contrib);
  hipLaunchKernelGGL(pr_zero_ull, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, R.stream, incoming, n);
  int rounds = 0, launches = 0;
  const int rc = run_rounds(
      [&](long long) {
        ktimer_begin("pr_push");
        hipLaunchKernelGGL((pr_push<G>), dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nindex, g->nlist, n,
                           contrib, incoming, R.d_ctl);
        ktimer_end("pr_push");
        hipLaunchKernelGGL(pr_push_compute, dim3(R.persist_blocks), dim3(kBlock), 0, R.stream, g->nindex, n, base,
                           score, contrib, incoming, R.d_ctl);
        launch_finalize(R.persist_blocks, 0x1, FinPr{P.eps, P.max_iter});
      },
      P.max_iter, &rounds, &launches);
  IGX_TRY(timer_stop(&info->seconds));
  info->iterations = rounds;
  info->launches = launches;
  info->residual = R.h_ctl->res_last;
  return rc == IGX_ERR_LIMIT ? IGX_OK : rc;
}

const StyleEntry kStyles[] = {
    {"pull_thread_kahan", style_pull<kThread>, 0},
    {"pull_wave_kahan", style_pull<kWaveG>, 0},
    {"pull_block_kahan", style_pull<kBlockG>, 0},
    {"pull_balanced_kahan", style_balanced, 0},
    {"push_fixed_thread", style_push<kThread>, 0},
    {"push_fixed_wave", style_push<kWaveG>, 0},
    {"push_fixed_block", style_push<kBlockG>, 0},
};

}  // namespace

const StyleEntry* pr_styles(int* count)
{
  *count = (int)(sizeof(kStyles) / sizeof(kStyles[0]));
  return kStyles;
}

}  // namespace igx
