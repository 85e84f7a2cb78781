// Device-wide exclusive prefix sum over an int/uint8 array whose live length
// sits in device memory (clamped to the host-side capacity `cap`).  Grids are sized from a host-side capacity so the
// launch sequence does not depend on data (it can be graph-captured); blocks
// past the live length return at once.
//
// Three kernels: per-block totals, one-block scan of the totals, per-block
// rescan with the block offset (one kernel for arrays of at most eight tiles).  out[i] = sum(in[0..i)), out[n] = total,
// and *total (if given) = total.
#pragma once
#include <hip/hip_runtime.h>

namespace rbh {

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

__device__ __forceinline__ int block_exclusive_scan(int v, int* sh, int& total) {
  // sh: SCAN_BLOCK / 64 ints; wave64 inclusive scan via shuffles.
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh[wid] = x;
  __syncthreads();
  int base = 0, tot = 0;
  for (int w = 0; w < SCAN_BLOCK / 64; ++w) {
    int s = sh[w];
    if (w < wid) base += s;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return base + x - v;
}

template <typename T>
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_reduce(const T* __restrict__ in, const int* __restrict__ n_dev,
                                                           int cap, int* __restrict__ part) {
  __shared__ int sh[SCAN_BLOCK / 64];
  const int n = min(max(*n_dev, 0), cap);
  const long base = (long)blockIdx.x * SCAN_TILE;
  if (base >= n) return;
  int s = 0;
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    long idx = base + (long)i * SCAN_BLOCK + threadIdx.x;
    if (idx < n) s += (int)in[idx];
  }
  int tot;
  block_exclusive_scan(s, sh, tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_top(int* __restrict__ part, const int* __restrict__ n_dev, int cap,
                                                        int* __restrict__ total_out) {
  __shared__ int sh[SCAN_BLOCK / 64];
  const int n = min(max(*n_dev, 0), cap);
  const int nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  int carry = 0;
  for (int base = 0; base < nb; base += SCAN_BLOCK) {
    int i = base + threadIdx.x;
    int v = i < nb ? part[i] : 0;
    int tot;
    int ex = block_exclusive_scan(v, sh, tot);
    if (i < nb) part[i] = carry + ex;
    carry += tot;
  }
  // the block whose first element is index n (n a multiple of the tile)
  // writes out[n] from part[nb]
  if (threadIdx.x == 0) {
    part[nb] = carry;
    if (total_out) *total_out = carry;
  }
}

template <typename T>
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_down(const T* __restrict__ in, const int* __restrict__ n_dev,
                                                         int cap, const int* __restrict__ part, int* __restrict__ out) {
  __shared__ int sh[SCAN_BLOCK / 64];
  const int n = min(max(*n_dev, 0), cap);
  const long base = (long)blockIdx.x * SCAN_TILE;
  if (base > n) return;
  // each thread owns SCAN_ITEMS consecutive elements
  int vals[SCAN_ITEMS];
  int s = 0;
  const long t0 = base + (long)threadIdx.x * SCAN_ITEMS;
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    long idx = t0 + i;
    vals[i] = idx < n ? (int)in[idx] : 0;
    s += vals[i];
  }
  int tot;
  int ex = block_exclusive_scan(s, sh, tot);
  int run = part[blockIdx.x] + ex;
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    long idx = t0 + i;
    if (idx <= n) out[idx] = run;   // out[n] = total
    run += vals[i];
  }
}

// Short arrays (bodies, units, scenes): one workgroup scans the tiles in
// turn with a running carry, in one launch instead of three.  Same outputs as
// the three kernels above.
constexpr int SCAN_SINGLE_TILES = 8;

template <typename T>
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_single(const T* __restrict__ in, const int* __restrict__ n_dev,
                                                           int cap, int* __restrict__ out, int* __restrict__ total_out) {
  __shared__ int sh[SCAN_BLOCK / 64];
  const int n = min(max(*n_dev, 0), cap);
  int carry = 0;
  // base <= n: the tile that holds out[n] runs even when n is a multiple of the tile
  for (long base = 0; base <= n; base += SCAN_TILE) {
    int vals[SCAN_ITEMS];
    int s = 0;
    const long t0 = base + (long)threadIdx.x * SCAN_ITEMS;
    for (int i = 0; i < SCAN_ITEMS; ++i) {
      const long idx = t0 + i;
      vals[i] = idx < n ? (int)in[idx] : 0;
      s += vals[i];
    }
    int tot;
    int run = carry + block_exclusive_scan(s, sh, tot);
    for (int i = 0; i < SCAN_ITEMS; ++i) {
      const long idx = t0 + i;
      if (idx <= n) out[idx] = run;   // out[n] = total
      run += vals[i];
    }
    carry += tot;
  }
  if (threadIdx.x == 0 && total_out) *total_out = carry;
}

}  // namespace rbh
