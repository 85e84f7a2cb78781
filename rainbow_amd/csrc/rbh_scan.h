// Device-wide exclusive prefix sum over an int/uint8 array whose live length
// sits in device memory (clamped to the host-side capacity `cap`).  Grids are sized from a host-side capacity so the
// launch sequence does not depend on data (it can be graph-captured); blocks
// past the live length return at once.
//
// One kernel: a single workgroup for arrays of at most eight tiles, a
// single-pass decoupled look-back for longer ones.  out[i] = sum(in[0..i)),
// out[n] = total, and *total (if given) = total.
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

// Tiles of the look-back scan: 32 ints or 64 bytes per thread (8 KB / 16 KB
// of input per workgroup), so that a long array has few tiles to look back
// over.
template <typename T> struct LbTile {
  static constexpr int ITEMS = sizeof(T) == 1 ? 64 : 32;
  static constexpr int TILE = SCAN_BLOCK * ITEMS;
};

// Loads of one thread's N consecutive elements at t0; a full tile (every
// element below n) is read with 16-byte vector loads.
template <int N>
__device__ __forceinline__ void scan_load(const int* __restrict__ in, long t0, int n, bool full, int (&v)[N]) {
  if (full) {
    const int4* p = reinterpret_cast<const int4*>(in + t0);
#pragma unroll
    for (int q = 0; q < N / 4; ++q) {
      const int4 x = p[q];
      v[4 * q] = x.x; v[4 * q + 1] = x.y; v[4 * q + 2] = x.z; v[4 * q + 3] = x.w;
    }
  } else {
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = t0 + i < n ? in[t0 + i] : 0;
  }
}

template <int N>
__device__ __forceinline__ void scan_load(const unsigned char* __restrict__ in, long t0, int n, bool full,
                                          int (&v)[N]) {
  if (full) {
    const uint4* p = reinterpret_cast<const uint4*>(in + t0);
#pragma unroll
    for (int q = 0; q < N / 16; ++q) {
      const uint4 x = p[q];
      const unsigned w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
      for (int i = 0; i < 16; ++i) v[16 * q + i] = (int)((w[i >> 2] >> (8 * (i & 3))) & 0xffu);
    }
  } else {
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = t0 + i < n ? (int)in[t0 + i] : 0;
  }
}

// Single-pass scan of arrays of more than SCAN_SINGLE_TILES tiles, by
// decoupled look-back.  A workgroup takes a ticket as it starts, and tiles go
// by ticket, so a tile only ever waits for tiles whose workgroups are already
// running.  Each tile publishes its total, then sums the values its
// predecessors published (tile totals, back to the nearest tile that has
// published its inclusive prefix) and publishes its own inclusive prefix.
// A status word holds flag (bits 62-63) and value (bits 0-31).
//
// state: word 0 the ticket counter, word 1 + t the status of tile t (the
// grid has one workgroup per tile of the capacity, LbTile<T>); all
// zero before the launch (the step's first kernel clears them).  The wait for
// a predecessor is bounded: past SCAN_SPIN_LIMIT polls the kernel sets
// SCAN_ERR_STALL in err[0] and finishes (the host reports the error).
constexpr unsigned long long SCAN_FLAG_AGG = 1ull << 62, SCAN_FLAG_INC = 2ull << 62;
constexpr int SCAN_SPIN_LIMIT = 1 << 20;
constexpr int SCAN_ERR_STALL = 128;

__device__ __forceinline__ int scan_wave_sum(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <typename T>
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_lookback(const T* __restrict__ in, const int* __restrict__ n_dev,
                                                             int cap, int* __restrict__ out,
                                                             int* __restrict__ total_out,
                                                             unsigned long long* __restrict__ state,
                                                             int* __restrict__ err) {
  constexpr int ITEMS = LbTile<T>::ITEMS, TILE = LbTile<T>::TILE;
  __shared__ int sh[SCAN_BLOCK / 64];
  __shared__ int s_tile, s_excl;
  const int n = min(max(*n_dev, 0), cap);
  // the grid covers the capacity; workgroups past the tile that holds out[n]
  // leave before taking a ticket, so the tickets are exactly the live tiles
  if ((long)blockIdx.x * TILE > n) return;
  if (threadIdx.x == 0) s_tile = (int)atomicAdd(reinterpret_cast<unsigned int*>(state), 1u);
  __syncthreads();
  const int tile = s_tile;
  const long base = (long)tile * TILE;
  const bool full = base + TILE <= n;
  const long t0 = base + (long)threadIdx.x * ITEMS;
  int vals[ITEMS];
  scan_load<ITEMS>(in, t0, n, full, vals);
  int s = 0;
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) s += vals[i];
  int tot;
  const int ex = block_exclusive_scan(s, sh, tot);
  if (threadIdx.x < 64) {
    const int lane = (int)threadIdx.x;
    unsigned long long* status = state + 1;
    int excl = 0;
    if (tile > 0) {
      if (lane == 0)
        __hip_atomic_store(status + tile, SCAN_FLAG_AGG | (unsigned)tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      // lane k reads tile j - k; before tile 0 the inclusive prefix is 0
      int j = tile - 1, polls = 0;
      while (true) {
        const int t = j - lane;
        const unsigned long long w =
            t >= 0 ? __hip_atomic_load(status + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : SCAN_FLAG_INC;
        if (!__all((w >> 62) != 0ull)) {
          if (++polls < SCAN_SPIN_LIMIT) {
            __builtin_amdgcn_s_sleep(1);
            continue;
          }
          if (lane == 0) atomicOr(err, SCAN_ERR_STALL);
          break;
        }
        const unsigned long long inc = __ballot((w >> 62) == 2ull);
        const int last = inc ? __ffsll((unsigned long long)inc) - 1 : 63;   // nearest inclusive prefix
        excl += scan_wave_sum(lane <= last ? (int)(unsigned)w : 0);
        if (inc) break;
        j -= 64;
      }
    }
    if (lane == 0) {
      __hip_atomic_store(status + tile, SCAN_FLAG_INC | (unsigned)(excl + tot), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      s_excl = excl;
    }
  }
  __syncthreads();
  int run = s_excl + ex;
  if (full) {
    int4* o = reinterpret_cast<int4*>(out + t0);
#pragma unroll
    for (int q = 0; q < ITEMS / 4; ++q) {
      int4 x;
      x.x = run; run += vals[4 * q];
      x.y = run; run += vals[4 * q + 1];
      x.z = run; run += vals[4 * q + 2];
      x.w = run; run += vals[4 * q + 3];
      o[q] = x;
    }
  } else {
    for (int i = 0; i < ITEMS; ++i) {
      if (t0 + i <= n) out[t0 + i] = run;   // out[n] = total
      run += vals[i];
    }
    if (threadIdx.x == 0 && total_out) *total_out = s_excl + tot;   // this tile holds index n
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
