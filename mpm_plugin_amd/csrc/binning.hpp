// Particle binning on gfx950: stable LSD radix sort of (cell key, slot) pairs, cell offsets and the list
// of occupied bricks.
//
// A particle's cell is its stencil base node floor(x / dx - 0.5) (the node its quadratic stencil starts
// at, kernels.py:26-27). Cells are numbered brick-major: key = brick * 64 + (cell within the 4x4x4-node
// brick), so one brick's particles are contiguous after the sort and, inside the brick, grouped by cell.
// The sort is stable: particles of one cell keep their previous relative order, and the permutation equals
// numpy.argsort(keys, kind="stable") bit for bit.
//
// Radix sort: 8-bit digits, least significant first. Each pass runs
//   k_radix_hist     per-tile digit histogram, stored digit-major (hist[d * tiles + t])
//   k_scan_tiles     exclusive scan over the digit-major histogram, in two parts (tile-local scans and the
//                    scanned tile totals) that give the output offset of (digit d, tile t) when added
//   k_radix_scatter  per-tile stable ranking (wave64 ballots + per-wave digit counts in LDS)
// A tile is 256 threads x kItems items, walked in item order so the ranking preserves input order.
#pragma once

#include "kernels.hpp"

namespace mpm {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

// word i of a two-part scan (k_scan_tiles)
__device__ __forceinline__ uint32_t scanned(const uint32_t* __restrict__ out, const uint32_t* __restrict__ tile_sums,
                                            uint32_t i) {
  return out[i] + tile_sums[i / (uint32_t)kScanTile];
}

// keys[p] = cell key of particle slot p; vals[p] = p
__global__ void __launch_bounds__(256) k_cell_keys(Particles P, GridGeom g, int n, uint32_t* __restrict__ keys,
                                                   uint32_t* __restrict__ vals) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float x[3] = {P.x[p], P.x[P.cap + p], P.x[2 * P.cap + p]};
  int base[3];
  float fx[3];
  stencil_base(g, x, base, fx);
  keys[p] = cell_key(g, base);
  vals[p] = (uint32_t)p;
}

__global__ void __launch_bounds__(kSortThreads) k_radix_hist(const uint32_t* __restrict__ keys, int n, int shift,
                                                             int tiles, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int t = blockIdx.x;
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const int i = t * kSortTile + j * kSortThreads + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[threadIdx.x * tiles + t] = h[threadIdx.x];
}

__global__ void __launch_bounds__(kSortThreads) k_radix_scatter(const uint32_t* __restrict__ keys_in,
                                                                const uint32_t* __restrict__ vals_in, int n,
                                                                int shift, int tiles,
                                                                const uint32_t* __restrict__ offsets,
                                                                const uint32_t* __restrict__ tile_sums,
                                                                uint32_t* __restrict__ keys_out,
                                                                uint32_t* __restrict__ vals_out) {
  constexpr int kWaves = kSortThreads / 64;
  __shared__ uint32_t running[256];
  __shared__ uint32_t wave_count[kWaves][256];
  const int t = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  running[threadIdx.x] = scanned(offsets, tile_sums, threadIdx.x * tiles + t);
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int j = 0; j < kSortItems; ++j) {
#pragma unroll
    for (int w = 0; w < kWaves; ++w) wave_count[w][threadIdx.x] = 0;
    __syncthreads();
    const int i = t * kSortTile + j * kSortThreads + threadIdx.x;
    const bool valid = i < n;
    const uint32_t key = valid ? keys_in[i] : 0u;
    const uint32_t val = valid ? vals_in[i] : 0u;
    const uint32_t d = (key >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint64_t m = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? m : ~m;
    }
    const uint32_t rank_in_wave = (uint32_t)__popcll(peers & lt_mask);
    if (valid && rank_in_wave == 0) wave_count[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = running[d] + rank_in_wave;
      for (int w = 0; w < wave; ++w) pos += wave_count[w][d];
      keys_out[pos] = key;
      vals_out[pos] = val;
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) add += wave_count[w][threadIdx.x];
    running[threadIdx.x] += add;
    __syncthreads();
  }
}

// ---- device-wide exclusive scan of uint32 (reduce, scan of block sums, apply) -----------------
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_warp, uint32_t& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) s_warp[wave] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (int w = 0; w < waves; ++w) {
      const uint32_t s = s_warp[w];
      s_warp[w] = acc;
      acc += s;
    }
    s_warp[waves] = acc;
  }
  __syncthreads();
  total = s_warp[waves];
  const uint32_t r = x - v + s_warp[wave];
  __syncthreads();
  return r;
}

// Called by every thread of every workgroup of a kernel at its end (no workgroup may have returned
// early): true, in all threads of one workgroup, for the workgroup that finishes last. `done` counts
// tickets; it starts at 0 and the last workgroup resets it for the kernel's next launch. Nothing waits: a
// workgroup that is not last just ends. `s_flag` is a shared word used for nothing else.
// What the last workgroup may read of the other workgroups' work: counters they updated with returning
// agent-scope atomics (each wave uses the returned value, so the update is performed before the wave
// reaches the barrier below, hence before its workgroup's ticket), read back with load_agent. There is
// no device-scope fence: on gfx950 one writes back the XCD's L2, and a fence per wave of a persistent
// grid cost hundreds of microseconds per substep. Plain stores of this kernel reach the next kernel
// through the kernel boundary as usual.
__device__ __forceinline__ bool last_workgroup(uint32_t* done, uint32_t* s_flag) {
  __syncthreads();
  if (threadIdx.x == 0) {
    const bool last = atomicAdd(done, 1u) == gridDim.x - 1;
    if (last) atomicExch(done, 0u);
    *s_flag = last ? 1u : 0u;
  }
  __syncthreads();
  return *s_flag != 0u;
}

// a counter updated by other workgroups' agent-scope atomics, read at the coherence point those use
__device__ __forceinline__ uint32_t load_agent(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive scan of n words in two parts: each kScanTile-word tile of `in` scanned on its own into `out`,
// and the tiles' totals (up to kScanTile tiles), which the workgroup that finishes last scans in place, so
// that the scan of word i is scanned(out, tile_sums, i). `ticket` (a last_workgroup counter) starts at 0
// and is left at 0. Each total is published with a returning agent-scope exchange, which the last
// workgroup reads back with load_agent.
__global__ void __launch_bounds__(kScanThreads) k_scan_tiles(const uint32_t* __restrict__ in, int n,
                                                             uint32_t* __restrict__ out,
                                                             uint32_t* __restrict__ tile_sums,
                                                             uint32_t* __restrict__ ticket) {
  __shared__ uint32_t s_warp[kScanThreads / 64 + 1];
  __shared__ uint32_t s_last;
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0u;
    sum += v[k];
  }
  uint32_t total;
  uint32_t acc = block_exclusive_scan(sum, s_warp, total);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = acc;
    acc += v[k];
  }
  if (threadIdx.x == 0) {
    // the old value is used (an empty asm takes it), so the exchange returns and the wave waits for it
    // before the ticket's barrier
    const uint32_t prev = atomicExch(&tile_sums[blockIdx.x], total);
    asm volatile("" ::"v"(prev));
  }
  if (!last_workgroup(ticket, &s_last)) return;
  const int tiles = (int)gridDim.x;
  const int tb = threadIdx.x * kScanItems;
  uint32_t w[kScanItems];
  uint32_t tsum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    w[k] = (tb + k < tiles) ? load_agent(&tile_sums[tb + k]) : 0u;
    tsum += w[k];
  }
  uint32_t all;
  uint32_t tacc = block_exclusive_scan(tsum, s_warp, all);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (tb + k < tiles) tile_sums[tb + k] = tacc;
    tacc += w[k];
  }
}

__global__ void __launch_bounds__(kScanThreads) k_scan_add(uint32_t* __restrict__ out, int n,
                                                           const uint32_t* __restrict__ tile_sums) {
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  const uint32_t add = tile_sums[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) out[base + k] += add;
}

// Occupied-brick list in decreasing order of work. A P2G wave takes one brick and its lanes walk their
// cells' particles in lockstep, so a brick costs in proportion to its fullest cell; the waves of the
// costliest bricks start first and the cheap ones fill the machine behind them (longest-processing-time
// order). Cost class q = 15 - min(fullest cell, 15): class 0 first.
constexpr int kCostClasses = 16;

// first sorted position with key >= k (binary search of [lo, hi))
__device__ __forceinline__ uint32_t lower_bound_key(const uint32_t* __restrict__ keys, uint32_t lo, uint32_t hi,
                                                    uint32_t k) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (keys[mid] < k) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// offsets[64 b] for every brick boundary b in [0, n_bricks] (the last one is n)
__global__ void __launch_bounds__(256) k_brick_bounds(const uint32_t* __restrict__ sorted_keys, int n, int n_bricks,
                                                      uint32_t* __restrict__ offsets) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > n_bricks) return;
  offsets[64 * (size_t)b] = lower_bound_key(sorted_keys, 0u, (uint32_t)n, 64u * (uint32_t)b);
}

// Occupied bricks are ordered by a counting sort on the cost class, stable in brick index, with no atomics:
//   k_brick_cost   one wave per brick (lane = cell): the brick's cell offsets and class into slot[b]
//   k_brick_rank   workgroup g (256 bricks): each brick's rank within its class in the workgroup, by
//                  ballots (slot[b] = class | rank << 4), and the class counts hist[q][g]
//   k_brick_order  every workgroup scans hist class-major (it is small) and places its own bricks
constexpr int kBrickGroup = 256;    // bricks per k_brick_rank / k_brick_order workgroup
constexpr int kBrickGroups = 1024;  // most workgroups (2^18 bricks): hist fits k_brick_order's LDS

// After k_brick_bounds. Each cell's offset is searched within its brick's own range (nothing to search in
// an empty brick); the brick's cost is its fullest cell.
__global__ void __launch_bounds__(256) k_brick_cost(const uint32_t* __restrict__ sorted_keys, int n_bricks,
                                                    uint32_t* __restrict__ offsets, int32_t* __restrict__ slot) {
  const int b = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (b >= n_bricks) return;
  uint32_t* o = offsets + 64 * (size_t)b;
  const uint32_t lo = o[0], hi = o[64];
  const uint32_t mine = lane == 0 ? lo : lower_bound_key(sorted_keys, lo, hi, 64u * (uint32_t)b + (uint32_t)lane);
  if (lane > 0) o[lane] = mine;
  const uint32_t next = (uint32_t)__shfl_down((int)mine, 1);
  uint32_t most = (lane == 63 ? hi : next) - mine;
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) most = max(most, (uint32_t)__shfl_xor((int)most, d));
  if (lane == 0)
    slot[b] = most == 0 ? -1 : kCostClasses - 1 - (int)min(most, (uint32_t)(kCostClasses - 1));
}

__global__ void __launch_bounds__(kBrickGroup) k_brick_rank(int n_bricks, int32_t* __restrict__ slot,
                                                           uint32_t* __restrict__ hist) {
  constexpr int kWaves = kBrickGroup / 64;
  __shared__ uint32_t s_cnt[kCostClasses][kWaves];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int b = blockIdx.x * kBrickGroup + threadIdx.x;
  const int32_t cls = b < n_bricks ? slot[b] : -1;
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  uint32_t in_wave = 0;
#pragma unroll
  for (int q = 0; q < kCostClasses; ++q) {
    const uint64_t m = __ballot(cls == q);
    if (cls == q) in_wave = (uint32_t)__popcll(m & lt);
    if (lane == 0) s_cnt[q][wave] = (uint32_t)__popcll(m);
  }
  __syncthreads();
  if (cls >= 0) {
    uint32_t before = 0;
    for (int w = 0; w < wave; ++w) before += s_cnt[cls][w];
    slot[b] = cls | (int32_t)((before + in_wave) << 4);
  }
  if (threadIdx.x < kCostClasses) {
    uint32_t t = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) t += s_cnt[threadIdx.x][w];
    hist[threadIdx.x * gridDim.x + blockIdx.x] = t;
  }
}

// Writes active[] (bricks by class, then brick index), slot[] (brick -> position) and the count
__global__ void __launch_bounds__(kBrickGroup) k_brick_order(int n_bricks, const uint32_t* __restrict__ hist,
                                                            uint32_t* __restrict__ active,
                                                            uint32_t* __restrict__ n_active,
                                                            int32_t* __restrict__ slot) {
  __shared__ uint32_t s_off[kCostClasses * kBrickGroups];
  __shared__ uint32_t s_warp[kBrickGroup / 64 + 1];
  const int groups = gridDim.x, entries = kCostClasses * groups;
  const int per_thread = (entries + kBrickGroup - 1) / kBrickGroup;
  uint32_t sum = 0;
  for (int k = 0; k < per_thread; ++k) {
    const int e = threadIdx.x * per_thread + k;
    sum += e < entries ? hist[e] : 0u;
  }
  uint32_t total;
  uint32_t acc = block_exclusive_scan(sum, s_warp, total);
  for (int k = 0; k < per_thread; ++k) {
    const int e = threadIdx.x * per_thread + k;
    if (e < entries) {
      s_off[e] = acc;
      acc += hist[e];
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *n_active = total;
  __syncthreads();
  const int b = blockIdx.x * kBrickGroup + threadIdx.x;
  if (b >= n_bricks) return;
  const int32_t packed = slot[b];
  if (packed < 0) return;
  const uint32_t pos = s_off[(packed & (kCostClasses - 1)) * groups + blockIdx.x] + ((uint32_t)packed >> 4);
  active[pos] = (uint32_t)b;
  slot[b] = (int32_t)pos;
}

// The 64-byte records of the 256 slots [r0, r0 + 256) handled by a 256-thread workgroup, dst[i] =
// src[perm[i]] for i < n: four lanes per record, one float4 each, so that each load and store instruction
// covers 16 whole records (with one float4 of 64 records per instruction, 64 records 64 bytes apart, the
// gather took 79 instead of 53 us on the 1M mix).
__device__ __forceinline__ void reorder_records(const float4* __restrict__ src, float4* __restrict__ dst,
                                                const uint32_t* __restrict__ perm, int r0, int n) {
  const int e = (int)threadIdx.x & 3;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int i = r0 + q * 64 + ((int)threadIdx.x >> 2);
    if (i < n) dst[4 * (int64_t)i + e] = src[4 * (int64_t)perm[i] + e];
  }
}

// gather every per-slot particle array into sorted order: dst[i] = src[perm[i]]
__global__ void __launch_bounds__(256) k_reorder(Particles src, Particles dst, const uint32_t* __restrict__ perm,
                                                 int n) {
  reorder_records(src.rec, dst.rec, perm, blockIdx.x * blockDim.x, n);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t s = perm[i];
  const int64_t cs = src.cap, cd = dst.cap;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    dst.x[a * cd + i] = src.x[a * cs + s];
    dst.v[a * cd + i] = src.v[a * cs + s];
  }
#pragma unroll
  for (int e = 0; e < 9; ++e) dst.F[e * cd + i] = src.F[e * cs + s];
  dst.Jp[i] = src.Jp[s];
  dst.mat[i] = src.mat[s];
  dst.pid[i] = src.pid[s];
}

}  // namespace mpm
