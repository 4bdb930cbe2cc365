// Batched substep: many independent scenes advance through one launch sequence.
//
// Every scene keeps its own particle sets, grid, bins and colliders (those of its context). A table of
// SceneDev records in device memory describes them, and a table of SceneStep records holds what changes
// from substep to substep (substep index, dt, the live particle set, sort bookkeeping, the speed maximum).
// The kernels read both tables, so one captured graph per substep kind serves every substep of every
// frame, whatever each scene's CFL step count: a scene that has run its `steps` substeps of the frame
// takes no further part until the next frame.
//
// Work is flattened across scenes on the device. At the start of a substep k_batch_plan computes, over
// the scenes that run (or sort) in it, prefix sums of their work units: 256-slot particle tiles, radix
// tiles, 256-cell blocks, groups of 64 bricks. Every kernel then runs a fixed, persistent grid that
// strides over the concatenated units of the running scenes only and finds each unit's scene in the
// prefix (searching forward from the previous one). A substep in which only a few scenes still run so
// costs only their work, and no launch size depends on the scenes' state, which is what lets the
// sequence be captured once. P2G strides the same way over the concatenation of the running scenes'
// occupied-brick lists (prefix computed after the sort), so large and small scenes share the machine at
// brick granularity.
//
// One substep (semantics of solver.py:176-211 per scene):
//   k_batch_plan            per scene: runs?, frame fraction s, sort now?, work prefixes, collider frames
//   [k_batch_cell_keys, radix passes, k_batch_cell_bounds, k_batch_reorder]   only in the sorting graph
//   k_batch_drift           scenes that did not sort: drifted particles to their current cell's lane;
//                           its last workgroup then scans the running scenes' occupied-brick counts
//   k_batch_p2g             over those bricks
//   k_batch_grid_scan       dirty bricks to per-scene work lists; its last workgroup scans their counts
//   k_batch_grid
//   k_batch_g2p             its last workgroup scans the counts of particles deferred to the FP64 pass
//   [k_batch_fixup]         when a scene's precision has an FP64 pass
//   k_batch_advance         bookkeeping; a scene that sorted switches to its reordered set here
// A prefix needed between two kernels is computed by whichever workgroup of the first kernel finishes
// last (a ticket counter tells it), not by a kernel of its own: one launch and one drain fewer each.
// Between the sort and k_batch_advance a scene's live particle set is live_set(): the reordered one if
// it sorted in this substep.
#pragma once

#include "binning.hpp"
#include "substep.hpp"

namespace mpm {

constexpr int kTile = 256;        // particle slots per unit of a per-particle kernel
constexpr int kGroupBricks = 64;  // bricks whose dirty flags one wave of the grid update examines
constexpr int kPlanThreads = 1024;
// per scene, the speed maximum of a substep's G2P is gathered in this many words (workgroup u takes word
// u mod kSpeedSlots): thousands of workgroups of one large scene on one word would serialise in L2
constexpr int kSpeedSlots = 64;

struct SceneDev {
  Particles set[2];  // set[cur] is live, the other is the reorder target
  MPM_GLOBAL const MatDev* mats;
  GridGeom g;
  StepConsts sc;
  ColliderSet cs;
  MPM_GLOBAL float4* accum;
  MPM_GLOBAL float4* vel;
  MPM_GLOBAL uint32_t* dirty;
  MPM_GLOBAL uint8_t* live;
  MPM_GLOBAL uint32_t *keys0, *keys1, *keys2, *vals1, *vals2;
  MPM_GLOBAL uint32_t* offsets;   // per cell
  MPM_GLOBAL uint32_t* active;    // occupied bricks
  MPM_GLOBAL uint32_t* n_active;  // their count
  MPM_GLOBAL uint32_t* dirty_list;  // bricks P2G wrote this substep (grid update work list)
  MPM_GLOBAL uint32_t* fix_list;    // particles whose material update is redone in FP64 this substep
  MPM_GLOBAL int32_t* slot;         // per brick: position in `active`, or -1
  MPM_GLOBAL float4* stage;         // per occupied brick: its P2G window (8^3 nodes)
  MPM_GLOBAL uint32_t* in_cnt;      // drift routing (DriftLists): per cell of an occupied brick
  MPM_GLOBAL uint32_t* in_lst;
  MPM_GLOBAL uint8_t* fwd;          // per particle
  int32_t n;           // particles
  int32_t n_bricks, n_cells, passes;
  int32_t sort_interval;
};

struct SceneStep {
  float dt, inv_frame_dt;
  int32_t k, steps;
  int32_t run;          // this substep advances the scene
  int32_t sort_now;     // this substep sorts the scene's particles first
  int32_t since_sort;   // substeps since the last sort
  int32_t sorted;
  int32_t cur;          // live particle set
  uint32_t speed_bits;  // max |v| of the scene's last G2P, as float bits (NaN recorded as +inf)
  uint32_t n_dirty;     // entries of dirty_list this substep
  uint32_t n_fix;       // entries of fix_list this substep
};

// the particle set the kernels after the sort work on: a scene that sorted in this substep has its
// particles in the reordered set (k_batch_advance makes that the current one)
__device__ __forceinline__ int live_set(const SceneStep& T) { return T.cur ^ T.sort_now; }

// work-unit prefixes, each n_scenes + 1 entries, in one array: pre + q * (n_scenes + 1)
enum PrefixKind {
  kPreTiles = 0,       // particle tiles of running scenes (G2P; the tile is the G2P workgroup size)
  kPreGroups = 1,      // brick groups of running scenes (grid update)
  kPreSortTiles = 2,   // particle tiles of sorting scenes (keys, reorder)
  kPreRadixTiles = 3,  // radix tiles of sorting scenes (histogram segment = 256 * this prefix)
  kPreCellBlocks = 4,  // 256-cell blocks of sorting scenes (cell offsets)
  kPreBricks = 5,      // occupied bricks of running scenes (P2G)
  kPreDirty = 6,       // bricks written by P2G in running scenes (grid update)
  kPreFix = 7,         // particles deferred to the FP64 pass in running scenes
  kPrefixKinds = 8
};

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint32_t uniform_u(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }

// the scene that owns unit t of a prefix (pre[i] <= t holds for the current guess i; units are visited
// in increasing order, so the search only moves forward)
__device__ __forceinline__ int scene_of(const uint32_t* __restrict__ pre, int n_scenes, uint32_t t, int i) {
  if (pre[i + 1] > t) return i;
  int lo = i + 1, hi = n_scenes - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ uint32_t ceil_div_u(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// Whole workgroup: pre[0..n_scenes] = exclusive prefix of count(i) over the scenes (pre[n_scenes] is the
// total). Any workgroup size that is a multiple of 64.
template <typename Count>
__device__ void workgroup_prefix(uint32_t* __restrict__ pre, int n_scenes, uint32_t* s_warp, Count count) {
  uint32_t carry = 0;
  for (int base = 0; base < n_scenes; base += blockDim.x) {
    const int i = base + threadIdx.x;
    const uint32_t c = i < n_scenes ? count(i) : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive_scan(c, s_warp, total);
    if (i < n_scenes) pre[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) pre[n_scenes] = carry;
}

// ticket counters of the kernels whose last workgroup computes a prefix
enum TicketKind { kTicketDrift = 0, kTicketGridScan = 1, kTicketG2p = 2, kTicketKinds = 3 };

// radix pass `pass` reads keys0/vals2 (pass 0), keys1/vals1 (odd) or keys2/vals2 (even) and writes the
// other pair; the sort's result is the output of pass passes-1
__device__ __forceinline__ const uint32_t* sorted_keys(const SceneDev& S) {
  return ((S.passes - 1) % 2 == 0) ? S.keys1 : S.keys2;
}
__device__ __forceinline__ const uint32_t* sorted_vals(const SceneDev& S) {
  return ((S.passes - 1) % 2 == 0) ? S.vals1 : S.vals2;
}

// Single workgroup. Per scene: does it run this substep (k < steps), its frame fraction, does it sort
// (never when allow_sort == 0, the graph without the binning kernels); the work prefixes of the
// running and sorting scenes; then the collider frames of every running scene at its fraction s.
__global__ void __launch_bounds__(kPlanThreads) k_batch_plan(const SceneDev* __restrict__ sd,
                                                             SceneStep* __restrict__ st,
                                                             ColliderFrame* __restrict__ frames, int n_scenes,
                                                             int allow_sort, int g2p_tile, uint32_t* __restrict__ pre,
                                                             uint32_t* __restrict__ speed_slots) {
  __shared__ uint32_t s_warp[kPlanThreads / 64 + 1];
  const int stride = n_scenes + 1;
  uint32_t carry[5] = {0, 0, 0, 0, 0};
  for (int base = 0; base < n_scenes; base += kPlanThreads) {
    const int i = base + threadIdx.x;
    uint32_t count[5] = {0, 0, 0, 0, 0};
    if (i < n_scenes) {
      const SceneDev& S = sd[i];
      SceneStep& T = st[i];
      const bool run = T.k < T.steps;
      const bool sort_now = allow_sort && run && (T.sorted == 0 || T.since_sort >= S.sort_interval);
      T.run = run ? 1 : 0;
      T.sort_now = sort_now ? 1 : 0;
      if (run) {
        T.speed_bits = 0u;
        T.n_dirty = 0u;
        T.n_fix = 0u;
        count[kPreTiles] = ceil_div_u((uint32_t)S.n, (uint32_t)g2p_tile);
        count[kPreGroups] = ceil_div_u((uint32_t)S.n_bricks, kGroupBricks);
      }
      if (sort_now) {
        *S.n_active = 0u;
        count[kPreSortTiles] = ceil_div_u((uint32_t)S.n, kTile);
        count[kPreRadixTiles] = ceil_div_u((uint32_t)S.n, kSortTile);
        count[kPreCellBlocks] = ceil_div_u((uint32_t)S.n_cells + 1u, 256u);
      }
    }
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      uint32_t total;
      const uint32_t ex = block_exclusive_scan(count[q], s_warp, total);
      if (i < n_scenes) pre[q * stride + i] = carry[q] + ex;
      carry[q] += total;
    }
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 5; ++q) pre[q * stride + n_scenes] = carry[q];
  }
  __syncthreads();  // every scene's `run` is set
  for (int t = threadIdx.x; t < n_scenes * kSpeedSlots; t += kPlanThreads)
    if (st[t / kSpeedSlots].run) speed_slots[t] = 0u;
  for (int t = threadIdx.x; t < n_scenes * kMaxColliders; t += kPlanThreads) {
    const int i = t / kMaxColliders, j = t % kMaxColliders;
    const int32_t k = st[i].k, steps = st[i].steps;
    if (k < steps && j < sd[i].cs.count) {
      // substep k of `steps` ends at frame fraction (k + 1) / steps (solver.py:217-218)
      const float s = (float)((double)(k + 1) / (double)steps);
      collider_frame(sd[i].cs.cols[j], s, st[i].inv_frame_dt, frames[t]);
    }
  }
}

// ---- binning of the scenes that sort this substep (binning.hpp, per scene) ---------------------
__global__ void __launch_bounds__(kTile) k_batch_cell_keys(const SceneDev* __restrict__ sd,
                                                           const SceneStep* __restrict__ st,
                                                           const uint32_t* __restrict__ pre, int n_scenes) {
  const uint32_t* P0 = pre + kPreSortTiles * (n_scenes + 1);
  const uint32_t total = P0[n_scenes];
  int i = 0;
  for (uint32_t u = blockIdx.x; u < total; u += gridDim.x) {
    i = uniform(scene_of(P0, n_scenes, u, i));
    const SceneDev& S = sd[i];
    const int p = (int)(u - P0[i]) * kTile + threadIdx.x;
    if (p >= S.n) continue;
    const Particles P = S.set[st[i].cur];
    const float x[3] = {P.x[p], P.x[P.cap + p], P.x[2 * P.cap + p]};
    int base[3];
    float fx[3];
    stencil_base(S.g, x, base, fx);
    S.keys0[p] = cell_key(S.g, base);
    S.vals2[p] = (uint32_t)p;
  }
}

__device__ __forceinline__ void radix_io(const SceneDev& S, int pass, const uint32_t*& kin, const uint32_t*& vin,
                                         uint32_t*& kout, uint32_t*& vout) {
  kin = pass == 0 ? S.keys0 : (pass % 2 ? S.keys1 : S.keys2);
  vin = pass == 0 ? S.vals2 : (pass % 2 ? S.vals1 : S.vals2);
  kout = pass % 2 == 0 ? S.keys1 : S.keys2;
  vout = pass % 2 == 0 ? S.vals1 : S.vals2;
}

// digit histogram of one radix tile; a sorting scene's histogram is the segment that starts at
// 256 * (its radix-tile prefix), digit-major
__global__ void __launch_bounds__(kSortThreads) k_batch_radix_hist(const SceneDev* __restrict__ sd,
                                                                   const uint32_t* __restrict__ pre, int n_scenes,
                                                                   int pass, uint32_t* __restrict__ hist) {
  const uint32_t* P0 = pre + kPreRadixTiles * (n_scenes + 1);
  const uint32_t total = P0[n_scenes];
  __shared__ uint32_t h[256];
  int i = 0;
  for (uint32_t u = blockIdx.x; u < total; u += gridDim.x) {
    i = uniform(scene_of(P0, n_scenes, u, i));
    const SceneDev& S = sd[i];
    if (pass >= S.passes) continue;
    const int tiles = (S.n + kSortTile - 1) / kSortTile;
    const int t = (int)(u - P0[i]);
    const uint32_t *kin, *vin;
    uint32_t *kout, *vout;
    radix_io(S, pass, kin, vin, kout, vout);
    h[threadIdx.x] = 0;
    __syncthreads();
    const int shift = 8 * pass;
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
      const int p = t * kSortTile + j * kSortThreads + threadIdx.x;
      if (p < S.n) atomicAdd(&h[(kin[p] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[256u * P0[i] + threadIdx.x * tiles + t] = h[threadIdx.x];
    __syncthreads();
  }
}

// stable scatter of one radix tile. (offsets, tile_sums) is the two-part exclusive scan of the whole
// histogram buffer (k_scan_tiles), so the scene's own exclusive scan is scanned(e) - scanned(first entry of
// its segment) (modulo 2^32, exact).
__global__ void __launch_bounds__(kSortThreads) k_batch_radix_scatter(const SceneDev* __restrict__ sd,
                                                                      const uint32_t* __restrict__ pre, int n_scenes,
                                                                      int pass, const uint32_t* __restrict__ offsets,
                                                                      const uint32_t* __restrict__ tile_sums) {
  constexpr int kWaves = kSortThreads / 64;
  const uint32_t* P0 = pre + kPreRadixTiles * (n_scenes + 1);
  const uint32_t total = P0[n_scenes];
  __shared__ uint32_t running[256];
  __shared__ uint32_t wave_count[kWaves][256];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int shift = 8 * pass;
  int i = 0;
  for (uint32_t u = blockIdx.x; u < total; u += gridDim.x) {
    i = uniform(scene_of(P0, n_scenes, u, i));
    const SceneDev& S = sd[i];
    if (pass >= S.passes) continue;
    const int tiles = (S.n + kSortTile - 1) / kSortTile;
    const int t = (int)(u - P0[i]);
    const uint32_t *kin, *vin;
    uint32_t *kout, *vout;
    radix_io(S, pass, kin, vin, kout, vout);
    const uint32_t seg = 256u * P0[i];
    running[threadIdx.x] = scanned(offsets, tile_sums, seg + threadIdx.x * tiles + t) - scanned(offsets, tile_sums, seg);
    for (int j = 0; j < kSortItems; ++j) {
#pragma unroll
      for (int w = 0; w < kWaves; ++w) wave_count[w][threadIdx.x] = 0;
      __syncthreads();
      const int p = t * kSortTile + j * kSortThreads + threadIdx.x;
      const bool valid = p < S.n;
      const uint32_t key = valid ? kin[p] : 0u;
      const uint32_t val = valid ? vin[p] : 0u;
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
        kout[pos] = key;
        vout[pos] = val;
      }
      __syncthreads();
      uint32_t add = 0;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) add += wave_count[w][threadIdx.x];
      running[threadIdx.x] += add;
      __syncthreads();
    }
  }
}

__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t* __restrict__ keys, int n, uint32_t c) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (keys[mid] < c) lo = mid + 1;
    else hi = mid;
  }
  return (uint32_t)lo;
}

// offsets[c] for c in [0, n_cells] (first sorted position with key >= c), one wave per brick (a 256-cell
// block = four bricks): two lanes find the brick's first and past-the-end positions over all keys, the
// other cells are searched within that range only (nothing to search in an empty brick), and lane 0
// appends the brick to the occupied list when it holds a particle
__global__ void __launch_bounds__(256) k_batch_cell_bounds(const SceneDev* __restrict__ sd,
                                                           const uint32_t* __restrict__ pre, int n_scenes) {
  const uint32_t* P0 = pre + kPreCellBlocks * (n_scenes + 1);
  const uint32_t total = P0[n_scenes];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int i = 0;
  for (uint32_t u = blockIdx.x; u < total; u += gridDim.x) {
    i = uniform(scene_of(P0, n_scenes, u, i));
    const SceneDev& S = sd[i];
    const int b = (int)(u - P0[i]) * 4 + wave;
    const int n_bricks = S.n_cells >> 6;
    if (b > n_bricks) continue;
    const uint32_t* keys = sorted_keys(S);
    if (b == n_bricks) {  // the terminal offset
      if (lane == 0) S.offsets[S.n_cells] = lower_bound_u32(keys, S.n, (uint32_t)S.n_cells);
      continue;
    }
    const uint32_t c0 = 64u * (uint32_t)b;
    const uint32_t edge = lane < 2 ? lower_bound_u32(keys, S.n, c0 + 64u * (uint32_t)lane) : 0u;
    const uint32_t lo = (uint32_t)__shfl((int)edge, 0), hi = (uint32_t)__shfl((int)edge, 1);
    uint32_t mine = lo;
    if (lane > 0) {
      uint32_t l = lo, h = hi;
      while (l < h) {
        const uint32_t mid = (l + h) >> 1;
        if (keys[mid] < c0 + (uint32_t)lane) l = mid + 1;
        else h = mid;
      }
      mine = l;
    }
    S.offsets[c0 + lane] = mine;
    if (lane == 0) {
      int32_t sl = -1;
      if (hi > lo) {
        sl = (int32_t)atomicAdd(S.n_active, 1u);
        S.active[sl] = (uint32_t)b;
      }
      S.slot[b] = sl;
    }
  }
}

// gather into the other particle set in sorted order (k_reorder)
__global__ void __launch_bounds__(kTile) k_batch_reorder(const SceneDev* __restrict__ sd,
                                                         const SceneStep* __restrict__ st,
                                                         const uint32_t* __restrict__ pre, int n_scenes) {
  const uint32_t* P0 = pre + kPreSortTiles * (n_scenes + 1);
  const uint32_t total = P0[n_scenes];
  int i = 0;
  for (uint32_t u = blockIdx.x; u < total; u += gridDim.x) {
    i = uniform(scene_of(P0, n_scenes, u, i));
    const SceneDev& S = sd[i];
    const int r0 = (int)(u - P0[i]) * kTile;
    const int cur = st[i].cur;
    const Particles src = S.set[cur], dst = S.set[1 - cur];
    static_assert(kTile == 256, "reorder_records moves the 256 records of a 256-thread workgroup");
    reorder_records(src.rec, dst.rec, sorted_vals(S), r0, S.n);
    const int p = r0 + threadIdx.x;
    if (p >= S.n) continue;
    const int64_t s = sorted_vals(S)[p];
    const int64_t cs = src.cap, cd = dst.cap;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      dst.x[a * cd + p] = src.x[a * cs + s];
      dst.v[a * cd + p] = src.v[a * cs + s];
    }
#pragma unroll
    for (int e = 0; e < 9; ++e) dst.F[e * cd + p] = src.F[e * cs + s];
    dst.Jp[p] = src.Jp[s];
    dst.mat[p] = src.mat[s];
    dst.pid[p] = src.pid[s];
  }
}

// drift routing (k_drift) of the running scenes that did not sort this substep, over their G2P tiles;
// the last workgroup then writes the prefix of the running scenes' occupied bricks (final now: the
// sorting graph's binning kernels have run) for P2G
__global__ void __launch_bounds__(256) k_batch_drift(const SceneDev* __restrict__ sd,
                                                     const SceneStep* __restrict__ st,
                                                     uint32_t* __restrict__ pre, int n_scenes, int tile,
                                                     uint32_t* __restrict__ tickets) {
  __shared__ uint32_t s_warp[256 / 64 + 1];
  __shared__ uint32_t s_last;
  const uint32_t* P0 = pre + kPreTiles * (n_scenes + 1);
  const uint32_t total = P0[n_scenes];
  int i = 0;
  for (uint32_t u = blockIdx.x; u < total; u += gridDim.x) {
    i = uniform(scene_of(P0, n_scenes, u, i));
    if (st[i].sort_now) continue;
    const SceneDev& S = sd[i];
    const Particles P = S.set[st[i].cur];
    const int end = min(S.n, (int)(u - P0[i] + 1) * tile);
    for (int p = (int)(u - P0[i]) * tile + threadIdx.x; p < end; p += blockDim.x)
      drift_route_pair(P, S.g, p, -1, sorted_keys(S), S.slot, S.in_cnt, S.in_lst, S.fwd);
  }
  if (last_workgroup(tickets + kTicketDrift, &s_last))
    workgroup_prefix(pre + kPreBricks * (n_scenes + 1), n_scenes, s_warp,
                     [&](int j) { return st[j].run ? *sd[j].n_active : 0u; });
}

// P2G over every running scene's occupied bricks, one wave per brick (p2g_brick)
// (one wave per workgroup, so a brick starts as soon as any wave slot frees)
template <int kWaves>
__global__ void __launch_bounds__(64 * kWaves) k_batch_p2g(const SceneDev* __restrict__ sd,
                                                           const SceneStep* __restrict__ st,
                                                           const uint32_t* __restrict__ pre, int n_scenes) {
  __shared__ float s_win[kWaves][4][kWinNodes];
  const uint32_t* P0 = pre + kPreBricks * (n_scenes + 1);
  const int wave = uniform(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const uint32_t total = P0[n_scenes];
  int i = 0;
  for (uint32_t w = blockIdx.x * kWaves + wave; w < total; w += gridDim.x * kWaves) {
    i = uniform(scene_of(P0, n_scenes, w, i));
    const SceneDev& S = sd[i];
    const int cur = live_set(st[i]);
    const float k = -st[i].dt * S.sc.p_vol * 4.0f * S.g.inv_dx * S.g.inv_dx;
    // a scene that sorted this substep has no drifted particles (and k_batch_drift skipped it)
    const DriftLists dl = st[i].sort_now ? DriftLists{} : DriftLists{S.in_cnt, S.in_lst, S.fwd};
    p2g_brick<1>(S.set[cur], S.g, k, S.offsets, S.active[w - P0[i]], &s_win[wave], S.accum, S.dirty, lane,
              FixedAccum{nullptr, 0.0, 0.0}, S.stage, w - P0[i], dl);
  }
}

// The same with kSplit waves per brick (one workgroup per brick at a time; wave h takes every kSplit-th
// particle of each cell): shorter chains per brick for the launches that carry only a few scenes
template <int kSplit>
__global__ void __launch_bounds__(64 * kSplit) k_batch_p2g_split(const SceneDev* __restrict__ sd,
                                                                 const SceneStep* __restrict__ st,
                                                                 const uint32_t* __restrict__ pre, int n_scenes) {
  __shared__ float s_win[kSplit][4][kWinNodes];
  const uint32_t* P0 = pre + kPreBricks * (n_scenes + 1);
  const int half = uniform(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const uint32_t total = P0[n_scenes];
  int i = 0;
  for (uint32_t w = blockIdx.x; w < total; w += gridDim.x) {
    i = uniform(scene_of(P0, n_scenes, w, i));
    const SceneDev& S = sd[i];
    const int cur = live_set(st[i]);
    const float k = -st[i].dt * S.sc.p_vol * 4.0f * S.g.inv_dx * S.g.inv_dx;
    const DriftLists dl = st[i].sort_now ? DriftLists{} : DriftLists{S.in_cnt, S.in_lst, S.fwd};
    p2g_brick<kSplit>(S.set[cur], S.g, k, S.offsets, S.active[w - P0[i]], s_win, S.accum, S.dirty, lane,
                      FixedAccum{nullptr, 0.0, 0.0}, S.stage, w - P0[i], dl, half);
  }
}

// Grid update, first pass: one wave per group of 64 bricks of a running scene; each lane reads one
// brick's dirty flag, records whether the brick holds a fresh velocity, clears the flag, and the dirty
// bricks are appended to the scene's work list (one atomic per wave). The last workgroup then writes
// the prefix of the running scenes' dirty-brick counts for the second pass.
__global__ void __launch_bounds__(256) k_batch_grid_scan(const SceneDev* __restrict__ sd, SceneStep* __restrict__ st,
                                                         uint32_t* __restrict__ pre, int n_scenes,
                                                         uint32_t* __restrict__ tickets) {
  __shared__ uint32_t s_warp[256 / 64 + 1];
  __shared__ uint32_t s_last;
  const uint32_t* P0 = pre + kPreGroups * (n_scenes + 1);
  const uint32_t total = P0[n_scenes];
  const int wave = uniform(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int i = 0;
  for (uint32_t gw = blockIdx.x * (blockDim.x >> 6) + wave; gw < total; gw += gridDim.x * (blockDim.x >> 6)) {
    i = uniform(scene_of(P0, n_scenes, gw, i));
    const SceneDev& S = sd[i];
    const int brick = (int)(gw - P0[i]) * kGroupBricks + lane;
    bool d = false;
    if (brick < S.n_bricks) {
      d = S.dirty[brick] != 0u;
      S.live[brick] = d ? 1 : 0;
      if (d) S.dirty[brick] = 0u;
    }
    const uint64_t mask = __ballot(d);
    if (mask == 0) continue;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(&st[i].n_dirty, (uint32_t)__popcll(mask));
    base = __shfl(base, 0, 64);
    if (d) S.dirty_list[base + (uint32_t)__popcll(mask & lt_mask)] = (uint32_t)brick;
  }
  if (last_workgroup(tickets + kTicketGridScan, &s_last))
    workgroup_prefix(pre + kPreDirty * (n_scenes + 1), n_scenes, s_warp,
                     [&](int j) { return st[j].run ? load_agent(&st[j].n_dirty) : 0u; });
}

// Grid update, second pass: one wave per dirty brick of every running scene (grid_brick)
__global__ void __launch_bounds__(256) k_batch_grid(const SceneDev* __restrict__ sd, const SceneStep* __restrict__ st,
                                                    const uint32_t* __restrict__ pre, int n_scenes,
                                                    const ColliderFrame* __restrict__ frames) {
  const uint32_t* P0 = pre + kPreDirty * (n_scenes + 1);
  const uint32_t total = P0[n_scenes];
  const int wave = uniform(threadIdx.x >> 6), lane = threadIdx.x & 63;
  int i = 0;
  for (uint32_t w = blockIdx.x * (blockDim.x >> 6) + wave; w < total; w += gridDim.x * (blockDim.x >> 6)) {
    i = uniform(scene_of(P0, n_scenes, w, i));
    const SceneDev& S = sd[i];
    const int b = (int)S.dirty_list[w - P0[i]];
    grid_brick(S.g, S.sc, st[i].dt, b, S.accum, S.vel, S.cs, frames + i * kMaxColliders, lane,
               FixedAccum{nullptr, 0.0, 0.0}, S.stage, S.slot);
  }
}

// G2P of every running scene (g2p_particle), one workgroup per tile of blockDim.x particles (64, 128 or
// 256: a launch-shape choice); each tile's speed maximum goes to the scene's record. With an FP64 pass
// the last workgroup then writes the prefix of the running scenes' deferred-particle counts.
__global__ void __launch_bounds__(256) k_batch_g2p(const SceneDev* __restrict__ sd, SceneStep* __restrict__ st,
                                                   uint32_t* __restrict__ pre, int n_scenes,
                                                   const ColliderFrame* __restrict__ frames,
                                                   uint32_t* __restrict__ speed_slots, int fp64_pass,
                                                   uint32_t* __restrict__ tickets) {
  const uint32_t* P0 = pre + kPreTiles * (n_scenes + 1);
  const uint32_t total = P0[n_scenes];
  const int tile = (int)blockDim.x, waves = tile >> 6;
  __shared__ float s_max[4];
  __shared__ uint32_t s_warp[256 / 64 + 1];
  __shared__ uint32_t s_last;
  __shared__ float s_c[9][256];  // each thread's C across its SVD (g2p_particle)
  int i = 0;
  for (uint32_t u = blockIdx.x; u < total; u += gridDim.x) {
    i = uniform(scene_of(P0, n_scenes, u, i));
    const SceneDev& S = sd[i];
    const int p = (int)(u - P0[i]) * tile + threadIdx.x;
    const int cur = live_set(st[i]);
    Particles P = S.set[cur];
    float speed = 0.0f;
    bool defer = false;
    if (p < S.n)
      defer = g2p_particle(P, S.mats, S.g, S.sc, st[i].dt, S.vel, p, S.cs, frames + i * kMaxColliders, speed,
                           &s_c[0][threadIdx.x], 256);
    if (S.sc.precision != kPrecisionFp32) wave_append(FixList{S.fix_list, &st[i].n_fix}, defer, (uint32_t)p);
    if (!(speed == speed)) speed = __int_as_float(0x7f800000);
    const float wm = wave_max(speed);
    uint32_t* slot_max = speed_slots + (size_t)i * kSpeedSlots + (u % kSpeedSlots);
    if (waves == 1) {
      if (threadIdx.x == 0) atomicMax(slot_max, __float_as_uint(wm));
      continue;
    }
    if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = wm;
    __syncthreads();
    if (threadIdx.x == 0) {
      float b = s_max[0];
      for (int w = 1; w < waves; ++w) b = fmaxf(b, s_max[w]);
      // non-negative floats order like their bit patterns
      atomicMax(slot_max, __float_as_uint(b));
    }
    __syncthreads();
  }
  if (fp64_pass && last_workgroup(tickets + kTicketG2p, &s_last))
    workgroup_prefix(pre + kPreFix * (n_scenes + 1), n_scenes, s_warp,
                     [&](int j) { return st[j].run ? load_agent(&st[j].n_fix) : 0u; });
}

// the FP64 pass over every running scene's deferred particles (k_fp64_fixup)
__global__ void __launch_bounds__(256) k_batch_fixup(const SceneDev* __restrict__ sd, const SceneStep* __restrict__ st,
                                                     const uint32_t* __restrict__ pre, int n_scenes) {
  const uint32_t* P0 = pre + kPreFix * (n_scenes + 1);
  const uint32_t total = P0[n_scenes];
  int i = 0;
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < total; u += gridDim.x * blockDim.x) {
    i = scene_of(P0, n_scenes, u, i);
    const SceneDev& S = sd[i];
    Particles P = S.set[live_set(st[i])];
    const int64_t p = S.fix_list[u - P0[i]];
    material_update_fp64(S.mats[P.mat[p]], P, p);
  }
}

// end of a substep: bookkeeping (a scene that sorted now lives in its reordered set), and the speed
// maximum of the scene's G2P from its slots
__global__ void __launch_bounds__(256) k_batch_advance(SceneStep* __restrict__ st, int n_scenes,
                                                       const uint32_t* __restrict__ speed_slots) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_scenes || !st[i].run) return;
  if (st[i].sort_now) {
    st[i].cur = 1 - st[i].cur;
    st[i].sorted = 1;
    st[i].since_sort = 0;
  }
  st[i].since_sort += 1;
  st[i].k += 1;
  uint32_t m = 0u;
  for (int j = 0; j < kSpeedSlots; ++j) m = max(m, speed_slots[(size_t)i * kSpeedSlots + j]);
  st[i].speed_bits = m;
}

}  // namespace mpm
