// Host-side internals of libindigo_hip.so: runtime state, error handling,
// cached device workspaces, the round loop and the style registry types.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <chrono>
#include <vector>

#include "igx_graph.h"
#include "indigo_hip.h"
#include "igx_device.hpp"

#define IGX_HIP(call)                                  \
  do {                                                 \
    hipError_t igx_e_ = (call);                        \
    if (igx_e_ != hipSuccess) {                        \
      igx::set_last_hip_error(igx_e_, #call);          \
      return IGX_ERR_HIP;                              \
    }                                                  \
  } while (0)

#define IGX_TRY(call)                \
  do {                               \
    int igx_s_ = (call);             \
    if (igx_s_ != IGX_OK) return igx_s_; \
  } while (0)

namespace igx {

void set_last_hip_error(hipError_t e, const char* what);

struct Runtime {
  bool ready = false;
  int device = 0;
  int cus = 0;                 // compute units
  int max_threads_per_cu = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  Ctl* d_ctl = nullptr;        // device control block
  Ctl* h_ctl = nullptr;        // pinned host mirror
  int persist_blocks = 0;      // grid of persistent kernels
  double time_limit = 0.0;     // igx_set_time_limit: seconds per run, 0 = none
  std::chrono::steady_clock::time_point t_start;  // host clock at timer_start
};

Runtime& rt();
int ensure_init();

// Cached device buffers, one per slot, grown on demand and kept until
// igx_release_workspace() -- allocation stays outside the timed region.
enum WsSlot {
  WS_A = 0, WS_B, WS_C, WS_D, WS_E, WS_F, WS_G, WS_H, WS_SP, WS_SCRATCH, WS_BLOCKS, WS_FEAT, WS_COUNT
};
int ws_get(int slot, size_t bytes, void** ptr);

// Enqueue round kernels in growing batches until the control block reports
// done or max_rounds were enqueued.  `launch(r)` enqueues round r.  Batches:
// first_batch, then 4, 8, 16, ... 64.  A round enqueued after convergence
// returns at once on the device but still costs its host-side submission
// (about 10 us, more than its device time), so the batch after the first
// stays small: a run that needs a few rounds more than the first batch wastes
// at most three submissions instead of a doubled batch.
// After each batch the run also stops, with IGX_ERR_TIMEOUT, once the host
// clock since timer_start() passes the time limit set by igx_set_time_limit.
// Returns IGX_OK, IGX_ERR_LIMIT when max_rounds did not converge, or an error.
bool past_time_limit();
template <typename F>
int run_rounds(F launch, long long max_rounds, int* rounds_done, int* launches, int first_batch = 2)
{
  Runtime& R = rt();
  long long launched = 0;
  int batch = first_batch;
  while (true) {
    const long long k = (launched + batch > max_rounds) ? (max_rounds - launched) : batch;
    for (long long i = 0; i < k; i++) {
      launch(launched + i);
    }
    IGX_HIP(hipGetLastError());
    launched += k;
    IGX_HIP(hipMemcpyAsync(R.h_ctl, R.d_ctl, sizeof(Ctl), hipMemcpyDeviceToHost, R.stream));
    IGX_HIP(hipStreamSynchronize(R.stream));
    if (R.h_ctl->done) break;
    if (launched >= max_rounds || past_time_limit()) {
      if (rounds_done) *rounds_done = R.h_ctl->iter;
      if (launches) *launches = (int)launched;
      return (launched >= max_rounds) ? IGX_ERR_LIMIT : IGX_ERR_TIMEOUT;
    }
    batch = (launched == k) ? 4 : ((batch < 64) ? batch * 2 : 64);
  }
  if (rounds_done) *rounds_done = R.h_ctl->iter;
  if (launches) *launches = (int)launched;
  return IGX_OK;
}

// Clear the control block on the stream (kernel, not a memset node).
int ctl_reset(int n);

// End of a round, enqueued after the round kernel: the topology-driven rule
// (stop after a round without updates), the data-driven rule (stop after a
// round that appended nothing), or a reducing finalizer over the round
// kernel's per-block partials (round_finalize in igx_device.hpp).
void launch_fin_flag();
void launch_fin_worklist();
template <typename Fin>
inline void launch_finalize(int nparts, int stat_mask, Fin fin)
{
  hipLaunchKernelGGL((round_finalize<Fin>), dim3(1), dim3(kFinBlock), 0, rt().stream, rt().d_ctl, nparts, stat_mask,
                     fin);
}

// multiplier for spread_vertex: odd, coprime to n
inline unsigned spread_mult(int n)
{
  auto gcd = [](unsigned long long a, unsigned long long b) {
    while (b != 0) {
      const unsigned long long t = a % b;
      a = b;
      b = t;
    }
    return a;
  };
  unsigned m = 2654435761u % (unsigned)(n > 1 ? n : 2);
  if (m == 0) m = 1;
  while (gcd(m, (unsigned)n) != 1) m++;
  return m;
}

inline int blocks_for(long long items, int per_block)
{
  long long b = (items + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > 0x7fffffffLL) b = 0x7fffffffLL;
  return (int)b;
}

// ---------------------------------------------------------------- styles --
typedef int (*RunFn)(const igx_graph* g, int src, void* d_out, const void* d_aux, igx_run_info* info);

struct StyleEntry {
  const char* name;
  RunFn fn;
  int edge_based;  // needs or benefits from sp[]
};

const StyleEntry* styles_of(int algo, int* count);

// per-algorithm registries (defined in each algorithm's translation unit)
const StyleEntry* bfs_styles(int* count);
const StyleEntry* sssp_styles(int* count);
const StyleEntry* cc_styles(int* count);
const StyleEntry* mis_styles(int* count);
const StyleEntry* pr_styles(int* count);
const StyleEntry* tc_styles(int* count);

// PR parameters (eps on the L1 residual, iteration cap)
struct PrParams {
  double eps = 1e-4;   // pr_cuda.h:42 (EPSILON)
  int max_iter = 100;  // pr_cuda.h:45 (MAX_ITER)
};
PrParams& pr_params();

// device sp[] of the graph, built on the device into the WS_SP workspace
int device_sp(const igx_graph* g, const int** d_sp);
// enqueue the construction of sp[] (edge -> row) into `sp` (edges ints)
int build_sp_async(const igx_graph* g, int* sp);

// Free the batched mode's own buffers and its captured graph.
int batch_release();

// Look up the device address of the allocation-free batched entry's scratch
// array on `device` (the current device); called once per device by igx_init.
int batch_noalloc_setup(int device);

// printf-style message for igx_last_error()
void snprintf_last_error(const char* fmt, ...);

// Enqueue the construction of the chunk list of every row longer than the
// frontier chunk (igx_frontier.hpp): bigq receives the chunks, *d_count their
// number.  Two kernels, no atomics: per-block counts, then each block writes
// at the sum of the counts before it.  bigq needs 2 * edges / chunk + 64
// entries.
int big_rows(const igx_graph* g, void* bigq, int* d_count);

// Tiles of 64 rows a wave claims at once in the kernels that hand out tiles
// dynamically: about 4096 edges of work per claim (two tiles on R-MAT 22,
// sixteen on a 2-D grid), so the shared claim counters stay cold.
inline int tile_grab(const igx_graph* g)
{
  const long long tiles = ((long long)g->nodes + 63) / 64;
  const long long grab = 4096LL * tiles / (g->edges > 0 ? (long long)g->edges : 1LL);
  return (int)(grab < 1 ? 1 : (grab > 32 ? 32 : grab));
}

// events bracketing the timed region
int timer_start();
int timer_stop(double* seconds);

// Per-kernel timing for the roofline: when igx_profile_kernel() selected
// `tag`, launches wrapped in ktimer_begin/ktimer_end with that tag are
// bracketed by events on the library stream and their durations summed.
void ktimer_begin(const char* tag);
void ktimer_end(const char* tag);

}  // namespace igx
