// Device-wide exclusive prefix sum of int32 counts in three launches:
// per-tile scans (1024 items per 256-lane workgroup), one workgroup scanning
// the tile totals, and a pass adding each tile's offset.  Used for building
// derived CSR arrays; not on any per-round path.
#pragma once

#include "igx_device.hpp"

namespace igx {

constexpr int kScanTile = 4 * kBlock;

__device__ __forceinline__ int block_excl_scan(int v, int* lds, int* total)
{
  const int w = threadIdx.x / kWave;
  const int incl = wave_incl_scan(v);
  if (lane_id() == kWave - 1) lds[w] = incl;
  __syncthreads();
  int off = 0, tot = 0;
  for (int k = 0; k < kWavesPerBlock; k++) {
    if (k < w) off += lds[k];
    tot += lds[k];
  }
  __syncthreads();
  *total = tot;
  return off + incl - v;
}

// out[i] = sum of in[0..i) within the tile; tile_sum[b] = tile total
static __global__ void __launch_bounds__(kBlock) scan_tiles(const int* __restrict__ in, int* __restrict__ out, int count,
                                                      int* __restrict__ tile_sum)
{
  __shared__ int lds[kWavesPerBlock];
  const long long base = (long long)blockIdx.x * kScanTile;
  int vals[4];
  int local = 0;
  #pragma unroll
  for (int k = 0; k < 4; k++) {
    const long long i = base + threadIdx.x * 4 + k;
    vals[k] = (i < count) ? in[i] : 0;
    local += vals[k];
  }
  int total = 0;
  int run = block_excl_scan(local, lds, &total);
  #pragma unroll
  for (int k = 0; k < 4; k++) {
    const long long i = base + threadIdx.x * 4 + k;
    if (i < count) out[i] = run;
    run += vals[k];
  }
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

// one workgroup: exclusive scan of the tile totals in place; total[0] = sum
static __global__ void __launch_bounds__(kBlock) scan_tile_sums(int* __restrict__ tile_sum, int tiles, int* __restrict__ total)
{
  __shared__ int lds[kWavesPerBlock];
  int carry = 0;
  for (int base = 0; base < tiles; base += kBlock) {
    const int i = base + threadIdx.x;
    const int v = (i < tiles) ? tile_sum[i] : 0;
    int chunk = 0;
    const int ex = block_excl_scan(v, lds, &chunk);
    if (i < tiles) tile_sum[i] = carry + ex;
    carry += chunk;
  }
  if (threadIdx.x == 0) *total = carry;
}

static __global__ void __launch_bounds__(kBlock) scan_add(int* __restrict__ out, int count, const int* __restrict__ tile_sum)
{
  const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (i < count) out[i] += tile_sum[i / kScanTile];
}

inline int scan_tiles_for(int count) { return (count + kScanTile - 1) / kScanTile > 0 ? (count + kScanTile - 1) / kScanTile : 1; }

// enqueue the three launches; tile_sum needs scan_tiles_for(count) ints
inline void exclusive_scan(const int* in, int* out, int count, int* tile_sum, int* total, hipStream_t s)
{
  const int tiles = scan_tiles_for(count);
  hipLaunchKernelGGL(scan_tiles, dim3(tiles), dim3(kBlock), 0, s, in, out, count, tile_sum);
  hipLaunchKernelGGL(scan_tile_sums, dim3(1), dim3(kBlock), 0, s, tile_sum, tiles, total);
  hipLaunchKernelGGL(scan_add, dim3((count + kBlock - 1) / kBlock > 0 ? (count + kBlock - 1) / kBlock : 1), dim3(kBlock),
                     0, s, out, count, tile_sum);
}

}  // namespace igx
