// Hand-written gfx950 kernels for the sparse-grid MLS-MPM substep.
//
// Substep pipeline (all counts stay on the device; every kernel is a
// grid-stride loop over a device-side count, so the launch sequence never
// depends on host readback):
//   1 k_bin_insert        particle -> 4^3-cell block key, lock-free hash insert
//                         by the head lane of each run of equal keys, counting-
//                         sort rank; each workgroup lists the blocks it claimed
//   2 k_bin_blocks        particle-block ids from prefix sums of those lists;
//                         insert the 7 upper neighbours of every particle block
//                         (grid blocks a block's P2G tile touches)
//   3 k_bin_offsets       counting-sort offsets, grid slots of neighbour blocks
//   4 k_bin_scatter       per particle block: grid slots of its 8 tile blocks;
//                         particle -> position in block-sorted order (every few
//                         substeps the particle state itself moves: reorder)
//   5 k_p2g               per particle block: stress, 27-node B-spline scatter
//                         into a 6^3 LDS tile accumulated in 64-bit fixed point
//                         (integer LDS atomics); each of the tile's 8 sub-boxes
//                         is stored in the area of the grid block it covers
//   6 k_grid_update       per node: sum of the covering tile parts in a fixed
//                         order, v = p/m + g dt, walls, obstacles
//   7 k_g2p               per particle block: 6^3 velocity tile in LDS, gather,
//                         v/C/x/F update, material projection, validity flags;
//                         clears the block hash for the next substep
// The reference computes the same substep with an open-addressed hash of
// all 27N stencil nodes (gpu_topology.wgsl:18-103) and per-particle CAS-loop
// float atomics (gpu.wgsl:711-1019); here the hash holds blocks, not nodes,
// and P2G/G2P work on LDS tiles.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "mpm_device.h"
#include "mpm_kernels.h"

namespace tmpm {

// ---------------------------------------------------------------------------
// Small helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ int readlane_i(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
// A value the caller knows to be uniform over the wavefront, in a scalar
// register (loads indexed by it can then be scalar loads).
__device__ __forceinline__ int4 uniform4(int4 v) {
  return make_int4(__builtin_amdgcn_readfirstlane(v.x), __builtin_amdgcn_readfirstlane(v.y),
                   __builtin_amdgcn_readfirstlane(v.z), __builtin_amdgcn_readfirstlane(v.w));
}
__device__ __forceinline__ uint32_t readlane_u(uint32_t v, int lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}

__device__ __forceinline__ void flag_error(uint32_t* counters, bool cond, uint32_t bit) {
  uint64_t m = __ballot(cond);
  if (m != 0 && lane_id() == (uint32_t)(__ffsll((unsigned long long)m) - 1)) atomicOr(&counters[CNT_ERROR], bit);
}

// Error of one scene's particles or blocks (`scene` uniform over the wave):
// recorded for the scene and in the session-wide word.
__device__ __forceinline__ void flag_scene_error(const Topo& t, uint32_t scene, bool cond, uint32_t bit) {
  uint64_t m = __ballot(cond);
  if (m != 0 && lane_id() == (uint32_t)(__ffsll((unsigned long long)m) - 1)) {
    atomicOr(&t.counters[CNT_ERROR], bit);
    atomicOr(&t.scene_err[scene], bit);
  }
}

// Cell key of one coordinate: floor(x * inv_h - 0.5) with both operations
// rounded to f32 and no fused multiply-add (gpu.rs:1145, the host packing of
// the reference).  Valid when finite and within [i32::MIN, i32::MAX - 2].
__device__ __forceinline__ bool cell_of(float x, float inv_h, int& cell) {
#pragma clang fp contract(off)
  float gp = x * inv_h;
  float shifted = gp - 0.5f;
  float f = floorf(shifted);
  bool ok = (double)f >= -2147483648.0 && (double)f <= 2147483645.0;
  cell = ok ? (int)f : 0;
  return ok;
}

__device__ __forceinline__ bool cell3(float4 pm, float inv_h, int3& c) {
  bool ok = cell_of(pm.x, inv_h, c.x);
  ok = cell_of(pm.y, inv_h, c.y) && ok;
  ok = cell_of(pm.z, inv_h, c.z) && ok;
  return ok;
}

__device__ __forceinline__ uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}
__device__ __forceinline__ uint32_t hash_key(int4 k) {
  uint32_t h = (uint32_t)k.x * 0x9E3779B1u;
  h ^= (uint32_t)k.y * 0x85EBCA77u;
  h ^= (uint32_t)k.z * 0xC2B2AE3Du;
  h ^= (uint32_t)k.w * 0x27D4EB2Fu;
  return fmix32(h);
}
__device__ __forceinline__ bool key_eq(int4 a, int4 b) { return a.x == b.x && a.y == b.y && a.z == b.z && a.w == b.w; }
// With CPIC the block key has a fifth word, the particles' side mask (the
// reference's transfer key, world.rs:470-490): blocks of one cell range but
// different masks are distinct and never share nodes.
__device__ __forceinline__ uint32_t hash_key_m(int4 k, uint32_t m) {
  uint32_t h = (uint32_t)k.x * 0x9E3779B1u;
  h ^= (uint32_t)k.y * 0x85EBCA77u;
  h ^= (uint32_t)k.z * 0xC2B2AE3Du;
  h ^= (uint32_t)k.w * 0x27D4EB2Fu;
  h ^= m * 0x165667B1u;
  return fmix32(h);
}

// Block key of a particle that binning accepted (cell valid).
__device__ __forceinline__ int4 particle_key(const Parts& p, const DevClass* classes, const DevScene* scenes,
                                             uint32_t i) {
  float4 pm = p.pos_mass[i];
  const DevClass& c = classes[p.cls[i]];
  float inv_h = c.inv_h;
  int3 cell;
  cell3(pm, inv_h, cell);
  return make_int4(cell.x >> 2, cell.y >> 2, cell.z >> 2, (int)c.region);
}

__device__ __forceinline__ int4 owner_key(const Parts& p, const Topo& t, const DevClass* classes,
                                          const DevScene* scenes, uint32_t owner) {
  if (owner <= p.n) return particle_key(p, classes, scenes, owner - 1);
  uint32_t r = owner - p.n - 1;
  uint32_t bucket = r >> 3, d = r & 7;
  // A neighbour insert's source is a particle block listed by the previous
  // kernel, whose key is visible by now: one load instead of rebuilding it.
  int4 k = t.bkey[bucket];
  if (k.x == kNoKey) k = particle_key(p, classes, scenes, t.owner[bucket] - 1);
  k.x += (int)(d >> 2);
  k.y += (int)((d >> 1) & 1);
  k.z += (int)(d & 1);
  return k;
}

// Side mask of the key an owner refers to (CPIC): a particle's own mask, or
// for a neighbour insert the mask of the particle that owns the source bucket.
__device__ __forceinline__ uint32_t owner_mask(const Parts& p, const Topo& t, uint32_t owner) {
  if (owner <= p.n) return t.pmask[owner - 1];
  const uint32_t r = owner - p.n - 1;
  return t.pmask[t.owner[r >> 3] - 1];
}

// Lock-free linear-probing insert.  Owners refer to keys that are immutable
// during the kernel (particle positions, or a particle block's key plus a
// fixed offset), so a losing thread compares against the winner's key
// without waiting for anything to be published.  The winner also stores the
// key in bkey; a loser that already sees it there skips recomputing the key
// from the owner (a chain of dependent loads).  Bounded by the table size.
// With CPIC a loser always rebuilds the winner's key and mask from the owner
// (the claimer's two stores are not published together).
// Whether bucket b, whose compare-and-swap returned owner `prev`, holds `key`.
template <bool CPIC>
__device__ __forceinline__ bool bucket_holds(const Parts& p, const Topo& t, const DevClass* classes,
                                             const DevScene* scenes, uint32_t b, uint32_t prev, int4 key,
                                             uint32_t mask) {
  if (CPIC) return key_eq(owner_key(p, t, classes, scenes, prev), key) && owner_mask(p, t, prev) == mask;
  int4 k = t.bkey[b];
  if (k.x == kNoKey) k = owner_key(p, t, classes, scenes, prev);
  return key_eq(k, key);
}

__device__ __forceinline__ void claim_bucket(const Topo& t, uint32_t b, int4 key, bool cpic, uint32_t mask) {
  t.bkey[b] = key;
  if (cpic) t.bmask[b] = mask;
}

// Probes from bucket b, the probe-th bucket of the key's sequence.  LOOKUP
// (without CPIC) reads bkey before trying to claim: where most keys are
// already in the table and their bkey is visible (the neighbour inserts after
// the particle inserts), a plain load finds them without an atomic.  bkey only
// ever holds its bucket's key or kNoKey, so a different key sends the probe on
// and kNoKey falls through to the claim.
template <bool CPIC, bool LOOKUP = false>
__device__ uint32_t hash_insert_from(const Parts& p, const Topo& t, const DevClass* classes, const DevScene* scenes,
                                     int4 key, uint32_t me, uint32_t b, uint32_t probe, bool& claimed,
                                     uint32_t mask) {
  claimed = false;
  for (; probe <= t.bucket_mask; ++probe) {
    if (LOOKUP && !CPIC) {
      const int4 k = t.bkey[b];
      if (k.x != kNoKey) {
        if (key_eq(k, key)) return b;
        b = (b + 1) & t.bucket_mask;
        continue;
      }
    }
    const uint32_t prev = atomicCAS(&t.owner[b], 0u, me);
    if (prev == 0u) {
      claimed = true;
      claim_bucket(t, b, key, CPIC, mask);
      return b;
    }
    if (bucket_holds<CPIC>(p, t, classes, scenes, b, prev, key, mask)) return b;
    b = (b + 1) & t.bucket_mask;
  }
  return kInvalid;
}

template <bool CPIC>
__device__ __forceinline__ uint32_t key_bucket(int4 key, uint32_t mask, uint32_t bucket_mask) {
  return (CPIC ? hash_key_m(key, mask) : hash_key(key)) & bucket_mask;
}

template <bool CPIC = false, bool LOOKUP = false>
__device__ uint32_t hash_insert(const Parts& p, const Topo& t, const DevClass* classes, const DevScene* scenes,
                                int4 key, uint32_t me, bool& claimed, uint32_t mask = 0) {
  return hash_insert_from<CPIC, LOOKUP>(p, t, classes, scenes, key, me, key_bucket<CPIC>(key, mask, t.bucket_mask),
                                        0, claimed, mask);
}

// Sums of `a[j]` over j < upto and over all j < n, for every thread of a
// 256-thread workgroup (the per-workgroup prefix of the binning lists).
constexpr int kBinThreads = 256;

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int K>
__device__ __forceinline__ void group_prefix(const uint32_t* const (&a)[K], uint32_t n, uint32_t upto,
                                             uint32_t (&before)[K], uint32_t (&total)[K]) {
  __shared__ uint32_t red[2 * K][kBinThreads / 64];
  uint32_t bs[K], ts[K];
#pragma unroll
  for (int k = 0; k < K; ++k) bs[k] = ts[k] = 0;
  // All of a thread's loads are issued before any is added: the session keeps
  // n <= 2048, i.e. one pass of kPer loads per list (a loop that adds as it
  // loads waits on each one in turn, ~3 us per workgroup).
  constexpr int kPer = 2048 / kBinThreads;
  for (uint32_t j0 = 0; j0 < n; j0 += kPer * kBinThreads) {
    uint32_t v[K][kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const uint32_t j = j0 + u * kBinThreads + threadIdx.x;
#pragma unroll
      for (int k = 0; k < K; ++k) v[k][u] = j < n ? a[k][j] : 0u;
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const uint32_t j = j0 + u * kBinThreads + threadIdx.x;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        ts[k] += v[k][u];
        if (j < upto) bs[k] += v[k][u];
      }
    }
  }
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    bs[k] = wave_sum(bs[k]);
    ts[k] = wave_sum(ts[k]);
    if (lane == 0) {
      red[2 * k][wave] = bs[k];
      red[2 * k + 1][wave] = ts[k];
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    before[k] = total[k] = 0;
#pragma unroll
    for (int w = 0; w < kBinThreads / 64; ++w) {
      before[k] += red[2 * k][w];
      total[k] += red[2 * k + 1][w];
    }
  }
  __syncthreads();
}

__device__ __forceinline__ void clear_rev(const Topo& t, uint32_t slot) {
#pragma unroll
  for (int d = 0; d < 8; ++d) t.grid_rev[8 * slot + d] = kInvalid;
}

// ---------------------------------------------------------------------------
// 1. reset
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_topo_clear(Topo t) {
  uint32_t nb = t.bucket_mask + 1;
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
    t.owner[b] = 0;
    t.bucket_count[b] = 0;
    t.bkey[b].x = kNoKey;
  }
}

// ---------------------------------------------------------------------------
// 2. particle -> block insert + counting-sort rank
// ---------------------------------------------------------------------------
// Workgroup g bins particles [g * bin_span, (g + 1) * bin_span).  Lanes of a
// wave hold consecutive particles, which are mostly block-sorted: the first
// lane of each run of equal keys (a "head") inserts the key and reserves the
// run's ranks with one atomic; all heads of a wave probe in parallel.  Blocks
// this workgroup claims are listed in its part of bin_stage; ids are handed
// out from prefix sums of the per-workgroup counts (k_bin_blocks), since
// returning atomics on one shared counter serialize device-wide.
//
// Every 256-particle round waits on a chain of device-scope atomics (claim,
// key check, rank reservation), and at most 2048 such workgroups are resident
// at once, so a 1M-particle binning is two rounds of that chain and 4M eight.
// Each thread therefore bins K particles per pass (sub-rounds k * 256 apart)
// and issues their loads, first-probe compare-and-swaps and rank atomics
// together: one chain per K rounds.  Capped at 64 VGPRs and the matching
// SGPRs so that 8 waves fit per SIMD and all 2048 workgroups are resident
// together: at 106 SGPRs only 7 fit, and the rest ran as a second
// generation that doubled the kernel's span.
constexpr int kInsertK = 4;  // at most; the launch picks 1, 2 or 4 by the workgroup's rounds

template <bool CPIC, int K>
__global__ __launch_bounds__(kBinThreads) __attribute__((amdgpu_waves_per_eu(8))) void k_bin_insert(Parts p, Topo t, const DevClass* classes,
                                                            const DevScene* scenes) {
  __shared__ uint32_t s_claims;
  const uint32_t lane = lane_id();
  if (threadIdx.x == 0) s_claims = 0;
  __syncthreads();
  const uint32_t begin = blockIdx.x * t.bin_span;
  uint32_t* stage = t.bin_stage + (size_t)blockIdx.x * t.bin_span;
  for (uint32_t r = 0; r < t.bin_span; r += K * kBinThreads) {
    uint32_t idx[K];
    bool in[K];
    float4 pm[K];
    uint32_t ci[K], mask[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      idx[k] = begin + r + k * kBinThreads + threadIdx.x;
      in[k] = r + k * kBinThreads < t.bin_span && idx[k] < p.n;
      pm[k] = in[k] ? p.pos_mass[idx[k]] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      ci[k] = in[k] ? p.cls[idx[k]] : 0u;
      mask[k] = CPIC && in[k] ? t.pmask[idx[k]] : 0u;
    }
    int4 key[K];
    bool active[K], head[K];
    uint64_t heads[K];
    uint32_t len[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      active[k] = false;
      key[k] = make_int4(0, 0, 0, 0);
      if (in[k] && pm[k].w > 0.0f) {
        const DevClass& c = classes[ci[k]];
        int3 cell;
        // A batch scene that failed in an earlier substep is not binned any
        // more: its particles stop like retired ones (synchronize restores
        // its last synchronized world), so a diverging scene cannot exhaust
        // the hash or grid capacity that every scene of the batch shares.
        if (t.isolate && t.scene_err[c.region >> 16] != 0u) {
          // not binned: pslot stays kInvalid below
        } else if (cell3(pm[k], c.inv_h, cell)) {
          active[k] = true;
          key[k] = make_int4(cell.x >> 2, cell.y >> 2, cell.z >> 2, (int)c.region);
        } else {
          atomicOr(&t.counters[CNT_ERROR], ERR_RESIDENT);
          atomicOr(&t.scene_err[c.region >> 16], ERR_RESIDENT);
        }
      }
      const int4 prev = make_int4(__shfl_up(key[k].x, 1, 64), __shfl_up(key[k].y, 1, 64),
                                  __shfl_up(key[k].z, 1, 64), __shfl_up(key[k].w, 1, 64));
      const bool prev_active = __shfl_up((int)active[k], 1, 64) != 0;
      const bool same_mask = !CPIC || __shfl_up(mask[k], 1, 64) == mask[k];
      head[k] = active[k] && (lane == 0 || !prev_active || !key_eq(key[k], prev) || !same_mask);
      heads[k] = __ballot(head[k]);
      const uint64_t cont = __ballot(active[k] && !head[k]);
      // run length: this lane plus the continuing lanes directly above it
      const uint64_t above = lane == 63 ? 0ull : (cont >> (lane + 1));
      len[k] = 1u + (uint32_t)__builtin_ctzll(~above);
    }
    // first probe of every head, all in flight together
    uint32_t bucket[K], prev[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      bucket[k] = kInvalid;
      prev[k] = 0u;
      if (head[k]) {
        bucket[k] = key_bucket<CPIC>(key[k], mask[k], t.bucket_mask);
        prev[k] = atomicCAS(&t.owner[bucket[k]], 0u, idx[k] + 1);
      }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (!head[k]) continue;
      bool claimed = prev[k] == 0u;
      if (claimed) {
        claim_bucket(t, bucket[k], key[k], CPIC, mask[k]);
      } else if (!bucket_holds<CPIC>(p, t, classes, scenes, bucket[k], prev[k], key[k], mask[k])) {
        bucket[k] = hash_insert_from<CPIC>(p, t, classes, scenes, key[k], idx[k] + 1,
                                           (bucket[k] + 1) & t.bucket_mask, 1, claimed, mask[k]);
      }
      if (bucket[k] == kInvalid) atomicOr(&t.counters[CNT_ERROR], ERR_TOPOLOGY);
      else if (claimed) stage[atomicAdd(&s_claims, 1u)] = bucket[k];
    }
    uint32_t first[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      first[k] = 0u;
      if (head[k] && bucket[k] != kInvalid) first[k] = atomicAdd(&t.bucket_count[bucket[k]], len[k]);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      // every active lane belongs to the run of the nearest head at or below it
      const uint64_t upto = heads[k] & ((2ull << lane) - 1ull);
      const int hl = upto ? 63 - __clzll((long long)upto) : (int)lane;
      const uint32_t b = __shfl(bucket[k], hl, 64);
      const uint32_t f = __shfl(first[k], hl, 64);
      if (in[k]) {
        const bool ok = active[k] && b != kInvalid;
        t.pslot[idx[k]] = ok ? b : kInvalid;
        if (ok) t.prank[idx[k]] = f + (lane - (uint32_t)hl);
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) t.bin_claims[blockIdx.x] = s_claims;
}

// ---------------------------------------------------------------------------
// 3. particle-block ids, grid slots of particle blocks, neighbour inserts
// ---------------------------------------------------------------------------
// Workgroup g numbers the blocks listed by insert workgroup g after those of
// workgroups < g.  A particle block's grid slot is its particle-block id.
// It then inserts the 7 upper neighbours of each of its blocks (the grid
// blocks its 6^3 P2G tile touches) and lists the ones it claims.
template <bool CPIC>
__global__ __launch_bounds__(kBinThreads) __attribute__((amdgpu_waves_per_eu(8))) void k_bin_blocks(Parts p, Topo t, const DevClass* classes,
                                                            const DevScene* scenes) {
  __shared__ uint32_t s_claims;
  __shared__ uint32_t s_csum[kBinThreads / 64];
  __shared__ uint32_t s_b[kBinThreads];  // the first kBinThreads listed buckets and their keys
  __shared__ int4 s_key[kBinThreads];
  const uint32_t g = blockIdx.x;
  if (threadIdx.x == 0) s_claims = 0;
  const uint32_t* stage = t.bin_stage + (size_t)g * t.bin_span;
  uint32_t* nstage = t.bin_nstage + (size_t)g * 7 * t.bin_span;
  // The list's length and first entries load together with the prefix sums
  // (the list has room for bin_span >= kBinThreads entries, so the load stays
  // in bounds; entries past the length are stale and unused).
  const uint32_t len = t.bin_claims[g];
  const uint32_t b0 = stage[threadIdx.x];
  const uint32_t* const lists[1] = {t.bin_claims};
  uint32_t before[1], total[1];
  group_prefix<1>(lists, t.bin_groups, g, before, total);
  uint32_t csum = 0;
  bool over = false;
  for (uint32_t j = threadIdx.x; j < len; j += kBinThreads) {
    const uint32_t b = j < kBinThreads ? b0 : stage[j], pb = before[0] + j;
    const int4 key = t.bkey[b];
    if (j < kBinThreads) {
      s_b[j] = b;
      s_key[j] = key;
    }
    if (pb < t.pblock_cap) {
      t.bucket_pblock[b] = pb;
      t.bucket_slot[b] = pb;
      t.pblock_nbr[8 * pb] = b;  // tile block d = 0: the block itself
      t.grid_key[pb] = key;
      clear_rev(t, pb);
      // the block's particle count, where k_bin_offsets reads it in id order
      const uint32_t cnt = t.bucket_count[b];
      t.pblock_len[pb] = cnt;
      csum += cnt;
    } else {
      t.bucket_pblock[b] = kInvalid;
      t.bucket_slot[b] = kInvalid;
      t.bucket_start[b] = kInvalid;
      over = true;
    }
  }
  __syncthreads();
  for (uint32_t q = threadIdx.x; q < 7 * len; q += kBinThreads) {
    const uint32_t j = q / 7u, d = q % 7u + 1u;
    const uint32_t b = j < kBinThreads ? s_b[j] : stage[j], pb = before[0] + j;
    if (pb >= t.pblock_cap) continue;
    int4 k = j < kBinThreads ? s_key[j] : t.bkey[b];
    k.x += (int)(d >> 2);
    k.y += (int)((d >> 1) & 1);
    k.z += (int)(d & 1);
    bool claimed;
    const uint32_t got =
        hash_insert<CPIC, true>(p, t, classes, scenes, k, p.n + 1 + 8 * b + d, claimed, CPIC ? t.bmask[b] : 0u);
    if (got == kInvalid) {
      over = true;
    } else if (claimed) {
      t.bucket_pblock[got] = kInvalid;
      nstage[atomicAdd(&s_claims, 1u)] = got;
    }
    t.pblock_nbr[8 * pb + d] = got;
  }
  flag_error(t.counters, over, ERR_TOPOLOGY);
  csum = wave_sum(csum);
  if ((threadIdx.x & 63) == 0) s_csum[threadIdx.x >> 6] = csum;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < kBinThreads / 64; ++w) c += s_csum[w];
    t.bin_csum[g] = c;
    t.bin_nclaims[g] = s_claims;
  }
}

// ---------------------------------------------------------------------------
// 4. counting-sort offsets and grid slots of neighbour-only blocks
// ---------------------------------------------------------------------------
// One wave per binning group, kOffsetGroups groups per workgroup: the
// workgroup sums the three per-group lists over the groups before its first
// one (the lists are read once per workgroup, not once per group: at 2048
// groups that read was ~48 MB of L2 traffic and half this kernel's time),
// and each wave adds the groups of its workgroup before its own and scans
// its group's blocks with wave shuffles.
constexpr int kOffsetGroups = kBinThreads / 64;

__device__ __forceinline__ uint32_t wave_exclusive_scan(uint32_t v, uint32_t& sum) {
  const uint32_t lane = lane_id();
  uint32_t incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t u = __shfl_up(incl, off, 64);
    if (lane >= (uint32_t)off) incl += u;
  }
  sum = __shfl(incl, 63, 64);
  return incl - v;
}

__global__ __launch_bounds__(kBinThreads) void k_bin_offsets(Parts p, Topo t, const DevClass* classes,
                                                             const DevScene* scenes) {
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t g0 = blockIdx.x * kOffsetGroups, g = g0 + wave;
  const uint32_t* const lists[3] = {t.bin_claims, t.bin_nclaims, t.bin_csum};
  uint32_t before[3], total[3];
  group_prefix<3>(lists, t.bin_groups, g0, before, total);
  for (uint32_t u = g0; u < g && u < t.bin_groups; ++u) {
    before[0] += t.bin_claims[u];
    before[1] += t.bin_nclaims[u];
    before[2] += t.bin_csum[u];
  }
  const uint32_t npb = min(total[0], t.pblock_cap);
  bool over = false;
  if (g < t.bin_groups) {
    // particle offsets of this group's particle blocks, in id order
    const uint32_t len = t.bin_claims[g];
    const uint32_t* stage = t.bin_stage + (size_t)g * t.bin_span;
    uint32_t carry = before[2];
    for (uint32_t c0 = 0; c0 < len; c0 += 64) {
      const uint32_t j = c0 + lane, pb = before[0] + j;
      const bool mine = j < len && pb < t.pblock_cap;
      // count (k_bin_blocks) and bucket load together
      const uint32_t cnt = mine ? t.pblock_len[pb] : 0u, b = mine ? stage[j] : 0u;
      uint32_t sum;
      const uint32_t excl = wave_exclusive_scan(cnt, sum);
      if (mine) {
        t.pblock_start[pb] = carry + excl;
        t.bucket_start[b] = carry + excl;  // the scatter's one look-up per particle
      }
      carry += sum;
    }
    // grid slots of the neighbour-only blocks this group claimed
    const uint32_t nlen = t.bin_nclaims[g];
    const uint32_t* nstage = t.bin_nstage + (size_t)g * 7 * t.bin_span;
    for (uint32_t j = lane; j < nlen; j += 64) {
      const uint32_t b = nstage[j], slot = npb + before[1] + j;
      if (slot < t.grid_cap) {
        t.bucket_slot[b] = slot;
        t.grid_key[slot] = t.bkey[b];
        clear_rev(t, slot);
      } else {
        t.bucket_slot[b] = kInvalid;
        over = true;
      }
    }
  }
  flag_error(t.counters, over, ERR_TOPOLOGY);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    t.counters[CNT_PARTICLE_BLOCKS] = npb;
    t.counters[CNT_GRID_BLOCKS] = min(npb + total[1], t.grid_cap);
    t.counters[CNT_BINNED] = total[2];
  }
}

// ---------------------------------------------------------------------------
// 5. per particle block: grid slots of the 8 tile blocks
// ---------------------------------------------------------------------------
__device__ __forceinline__ void pblock_finalize(const Topo& t) {
  uint32_t npb = t.counters[CNT_PARTICLE_BLOCKS];
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < npb * 8; q += gridDim.x * blockDim.x) {
    uint32_t k = q >> 3, d = q & 7;
    uint32_t nb = t.pblock_nbr[q];
    uint32_t slot = nb == kInvalid ? kInvalid : t.bucket_slot[nb];
    t.pblock_slot8[q] = slot;
    // grid block `slot` = this block + d: exactly one particle block per (slot, d)
    if (slot != kInvalid) t.grid_rev[8 * slot + d] = k;
  }
}

// ---------------------------------------------------------------------------
// 6. scatter particle indices into block-sorted order (optionally the state)
// ---------------------------------------------------------------------------
// Each thread scatters kScatterK particles per pass with their three
// dependent loads (slot and rank, block id, block start) in flight together,
// on at most as many workgroups as are resident at once.
constexpr int kScatterK = 4;
__global__ __launch_bounds__(256) void k_bin_scatter(Parts p, Topo t) {
  pblock_finalize(t);
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x; i0 < p.n; i0 += kScatterK * stride) {
    uint32_t b[kScatterK], r[kScatterK], st[kScatterK];
#pragma unroll
    for (int u = 0; u < kScatterK; ++u) {
      const uint32_t i = i0 + u * stride;
      b[u] = i < p.n ? t.pslot[i] : kInvalid;
      r[u] = i < p.n ? t.prank[i] : 0u;
    }
#pragma unroll
    for (int u = 0; u < kScatterK; ++u) st[u] = b[u] != kInvalid ? t.bucket_start[b[u]] : kInvalid;
    bool bad = false;
#pragma unroll
    for (int u = 0; u < kScatterK; ++u) {
      if (st[u] == kInvalid) continue;  // a block past the particle-block capacity (flagged)
      const uint32_t pos = st[u] + r[u];
      if (pos < p.n) t.sorted[pos] = i0 + u * stride;
      else bad = true;  // offsets and ranks disagree: never write past the list
    }
    flag_error(t.counters, bad, ERR_INTERNAL);
  }
}

// ---------------------------------------------------------------------------
// Diagnostics: live-node masks
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mask_clear(Topo t) {
  uint32_t ng = t.counters[CNT_GRID_BLOCKS];
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < 2 * ng; q += gridDim.x * blockDim.x)
    t.node_mask[q] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) t.counters[CNT_LIVE_NODES] = 0;
}

// ---------------------------------------------------------------------------
// Material and geometry helpers (f32)
// ---------------------------------------------------------------------------
struct M3 {
  float a[9];  // column-major: a[c*3 + r]
  __device__ float& at(int r, int c) { return a[c * 3 + r]; }
  __device__ float at(int r, int c) const { return a[c * 3 + r]; }
};
__device__ __forceinline__ float3 col(const M3& m, int c) { return make_float3(m.a[3 * c], m.a[3 * c + 1], m.a[3 * c + 2]); }
__device__ __forceinline__ float3 operator+(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 operator-(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 operator*(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float3 operator/(float3 a, float s) { return make_float3(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 cross3(float3 a, float3 b) {
  return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float len3(float3 a) { return sqrtf(dot3(a, a)); }
__device__ __forceinline__ float comp(float3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
__device__ __forceinline__ void set_comp(float3& v, int i, float s) {
  if (i == 0) v.x = s; else if (i == 1) v.y = s; else v.z = s;
}

__device__ __forceinline__ float det3(const M3& m) {
  return dot3(col(m, 0), cross3(col(m, 1), col(m, 2)));
}

// gpu.wgsl:451-465: the corotated stress uses the f32 Newton polar iteration
// only where it converges in 10 steps.
__device__ __forceinline__ bool regular_corotated(const M3& F) {
  float3 a = col(F, 0), b = col(F, 1), c = col(F, 2);
  float det = dot3(a, cross3(b, c));
  float nrm = sqrtf(dot3(a, a) + dot3(b, b) + dot3(c, c));
  if (fabsf(det) < 1e-5f || nrm < 1e-5f || nrm > 20.0f) return false;
  const float id = 1.0f / det;
  float3 i0 = cross3(b, c) * id, i1 = cross3(c, a) * id, i2 = cross3(a, b) * id;
  float inrm = sqrtf(dot3(i0, i0) + dot3(i1, i1) + dot3(i2, i2));
  return nrm * inrm <= 100.0f;
}

// Newton iteration R <- (R + R^-T) / 2 (gpu.wgsl:467-482), at most 10 steps.
// The gate above bounds the condition number by 100, where the iteration
// converges quadratically.  A lane stops updating once its last step moved
// no entry by more than 1e-6 (the next step would be below f32 resolution),
// so its result depends on its own F only, not on the other particles of
// its wavefront; the wavefront leaves the loop when every lane has stopped.
// Near F = I that is 3-4 steps instead of 10.
__device__ void polar_newton(const M3& F, M3& R) {
  R = F;
  bool done = false;
  for (int it = 0; it < 10; ++it) {
    if (!done) {
      float3 a = col(R, 0), b = col(R, 1), c = col(R, 2);
      float det = dot3(a, cross3(b, c));
      float inv = __builtin_amdgcn_rcpf(det);
      float3 t0 = cross3(b, c) * inv, t1 = cross3(c, a) * inv, t2 = cross3(a, b) * inv;
      float delta = 0.0f;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const float n0 = 0.5f * (comp(a, q) + comp(t0, q));
        const float n1 = 0.5f * (comp(b, q) + comp(t1, q));
        const float n2 = 0.5f * (comp(c, q) + comp(t2, q));
        delta = fmaxf(delta, fmaxf(fabsf(n0 - R.a[q]), fmaxf(fabsf(n1 - R.a[3 + q]), fabsf(n2 - R.a[6 + q]))));
        R.a[q] = n0;
        R.a[3 + q] = n1;
        R.a[6 + q] = n2;
      }
      done = delta <= 1e-6f;
    }
    if (__all(done)) break;
  }
}

// FP64 fallback for near-degenerate F: one-sided Jacobi SVD, R = U V^T (the
// rotation the reference CPU takes from nalgebra's SVD, material.rs:447-450).
// Taken and returned by value: a reference would make every caller keep F in
// scratch memory for this rarely taken call.
// Singular value decomposition F = U diag(s) V^T in f64 by one-sided Jacobi
// rotations of F's columns (singular values >= 0; U or V may include a
// reflection, as nalgebra's SVD does).  Degenerate columns of U are completed
// to an orthonormal frame.
__device__ __noinline__ void svd_fp64(const M3& Ff, double U[3][3], double s[3], double V[3][3]) {
  double A[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      A[r][c] = (double)Ff.at(r, c);
      V[r][c] = r == c ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 12; ++sweep) {
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        double alpha = 0, beta = 0, gamma = 0;
        for (int i = 0; i < 3; ++i) {
          alpha += A[i][p] * A[i][p];
          beta += A[i][q] * A[i][q];
          gamma += A[i][p] * A[i][q];
        }
        if (fabs(gamma) <= 1e-300) continue;
        double zeta = (beta - alpha) / (2.0 * gamma);
        double tt = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        double cs = 1.0 / sqrt(1.0 + tt * tt), sn = cs * tt;
        for (int i = 0; i < 3; ++i) {
          double ap = A[i][p], aq = A[i][q];
          A[i][p] = cs * ap - sn * aq;
          A[i][q] = sn * ap + cs * aq;
          double vp = V[i][p], vq = V[i][q];
          V[i][p] = cs * vp - sn * vq;
          V[i][q] = sn * vp + cs * vq;
        }
      }
  }
  double smax = 0;
  for (int j = 0; j < 3; ++j) {
    s[j] = sqrt(A[0][j] * A[0][j] + A[1][j] * A[1][j] + A[2][j] * A[2][j]);
    smax = fmax(smax, s[j]);
  }
  int good[3], ng = 0, bad = -1, nbad = 0;
  for (int j = 0; j < 3; ++j) {
    if (s[j] > smax * 1e-12 && s[j] > 0) {
      for (int i = 0; i < 3; ++i) U[i][j] = A[i][j] / s[j];
      good[ng++] = j;
    } else {
      bad = j;
      ++nbad;
    }
  }
  if (nbad == 1) {
    int a = good[0], b = good[1];
    U[0][bad] = U[1][a] * U[2][b] - U[2][a] * U[1][b];
    U[1][bad] = U[2][a] * U[0][b] - U[0][a] * U[2][b];
    U[2][bad] = U[0][a] * U[1][b] - U[1][a] * U[0][b];
  } else if (nbad > 1) {
    // Rank <= 1: no unique rotation; use the nearest rotation to V's frame.
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) U[r][c] = V[r][c];
  }
}

// Polar rotation R = U V^T in f64 (fallback of the corotated stress).  The
// same one-sided Jacobi as svd_fp64, kept in one frame: a call from here
// would raise the register count of every P2G instantiation.
__device__ __noinline__ M3 polar_fp64(M3 Ff) {
  M3 R;
  double A[3][3], V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) A[r][c] = (double)Ff.at(r, c);
  for (int sweep = 0; sweep < 12; ++sweep) {
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        double alpha = 0, beta = 0, gamma = 0;
        for (int i = 0; i < 3; ++i) {
          alpha += A[i][p] * A[i][p];
          beta += A[i][q] * A[i][q];
          gamma += A[i][p] * A[i][q];
        }
        if (fabs(gamma) <= 1e-300) continue;
        double zeta = (beta - alpha) / (2.0 * gamma);
        double tt = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        double cs = 1.0 / sqrt(1.0 + tt * tt), sn = cs * tt;
        for (int i = 0; i < 3; ++i) {
          double ap = A[i][p], aq = A[i][q];
          A[i][p] = cs * ap - sn * aq;
          A[i][q] = sn * ap + cs * aq;
          double vp = V[i][p], vq = V[i][q];
          V[i][p] = cs * vp - sn * vq;
          V[i][q] = sn * vp + cs * vq;
        }
      }
  }
  double s[3], U[3][3], smax = 0;
  for (int j = 0; j < 3; ++j) {
    s[j] = sqrt(A[0][j] * A[0][j] + A[1][j] * A[1][j] + A[2][j] * A[2][j]);
    smax = fmax(smax, s[j]);
  }
  int good[3], ng = 0, bad = -1, nbad = 0;
  for (int j = 0; j < 3; ++j) {
    if (s[j] > smax * 1e-12 && s[j] > 0) {
      for (int i = 0; i < 3; ++i) U[i][j] = A[i][j] / s[j];
      good[ng++] = j;
    } else {
      bad = j;
      ++nbad;
    }
  }
  if (nbad == 1) {
    int a = good[0], b = good[1];
    U[0][bad] = U[1][a] * U[2][b] - U[2][a] * U[1][b];
    U[1][bad] = U[2][a] * U[0][b] - U[0][a] * U[2][b];
    U[2][bad] = U[0][a] * U[1][b] - U[1][a] * U[0][b];
  } else if (nbad > 1) {
    // Rank <= 1: no unique rotation; use the nearest rotation to V's frame.
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) U[r][c] = V[r][c];
  }
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) R.at(r, c) = (float)(U[r][0] * V[c][0] + U[r][1] * V[c][1] + U[r][2] * V[c][2]);
  return R;
}

// Kirchhoff stress (material.rs:234-284; f32 op order of gpu.wgsl:638-699).
// plastic_det is read for snow only; PLASTIC = false instantiations (no snow,
// sand or sand-NH class in the session) leave the plastic materials out.
template <bool PLASTIC>
__device__ void kirchhoff(const DevClass& c, const M3& F, const M3& C, float plastic_det, M3& S) {
  float j = det3(F);
  if (c.kind == DM_NEO_HOOKEAN || (PLASTIC && c.kind == DM_SAND_NH)) {
    float iso = c.lambda * logf(fmaxf(j, 1e-10f)) - c.mu;
    for (int r = 0; r < 3; ++r)
      for (int q = 0; q < 3; ++q) {
        float s = F.at(r, 0) * F.at(q, 0) + F.at(r, 1) * F.at(q, 1) + F.at(r, 2) * F.at(q, 2);
        S.at(r, q) = s * c.mu + (r == q ? iso : 0.0f);
      }
  } else if (c.kind == DM_FLUID) {
    float sj = fmaxf(j, 1e-6f);
    float pressure = sj <= 1.0f ? c.bulk * (powf(sj, -c.gamma) - 1.0f) : -c.bulk * c.tensile * (sj - 1.0f);
    float tr3 = (C.at(0, 0) + C.at(1, 1) + C.at(2, 2)) / 3.0f;
    float diagonal = -2.0f * c.viscosity * sj * tr3 - pressure * sj;
    float k = 2.0f * c.viscosity * sj;
    for (int r = 0; r < 3; ++r)
      for (int q = 0; q < 3; ++q) {
        float sr = (C.at(r, q) + C.at(q, r)) * 0.5f;
        S.at(r, q) = sr * k + (r == q ? diagonal : 0.0f);
      }
  } else {  // corotated: elastic, sand, snow (hardened by the plastic volume)
    M3 R;
    if (regular_corotated(F)) polar_newton(F, R);
    else R = polar_fp64(F);
    const float factor = PLASTIC && c.kind == DM_SNOW
                             ? fminf(fmaxf(expf(c.hardening * (1.0f - plastic_det)), 0.01f), 100.0f)
                             : 1.0f;
    float iso = c.lambda * factor * (j - 1.0f) * j;
    float k = 2.0f * c.mu * factor;
    for (int r = 0; r < 3; ++r)
      for (int q = 0; q < 3; ++q) {
        float s = (F.at(r, 0) - R.at(r, 0)) * F.at(q, 0) + (F.at(r, 1) - R.at(r, 1)) * F.at(q, 1) +
                  (F.at(r, 2) - R.at(r, 2)) * F.at(q, 2);
        S.at(r, q) = s * k + (r == q ? iso : 0.0f);
      }
  }
}

// ---------------------------------------------------------------------------
// Plastic return mappings (material.rs:287-371 and 488-532; gpu.wgsl:489-636)
// ---------------------------------------------------------------------------
// New principal stretches of snow, sand and sand-NH for stretches s and the
// plastic state pl = {plastic_det, hardening, log_volume_gain}; false when F
// is left unchanged (sand with lambda = 0, or inside the yield surface).  T
// is float on the f32 path and double on the f64 fallback.
template <class T>
__device__ bool plastic_stretches(const DevClass& c, const T s[3], T pl[3], T out[3]) {
  if (c.kind == DM_SNOW) {
    for (int a = 0; a < 3; ++a) out[a] = fmin(fmax(s[a], T(1) - (T)c.crit_compression), T(1) + (T)c.crit_stretch);
    const T old_det = s[0] * s[1] * s[2];
    const T new_det = fmax(out[0] * out[1] * out[2], T(1e-12));
    pl[0] = fmin(fmax(pl[0] * old_det / new_det, T(0.1)), T(4));
    return true;
  }
  if (c.lambda == 0.0f) return false;
  const bool nh = c.kind == DM_SAND_NH;
  T strain[3];
  for (int a = 0; a < 3; ++a) strain[a] = log(fmax(s[a], T(1e-12))) + pl[2] / T(3);
  const T tr = strain[0] + strain[1] + strain[2];
  T dev[3];
  for (int a = 0; a < 3; ++a) dev[a] = strain[a] - tr / T(3);
  const T norm = sqrt(dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2]);
  const T cohesion = (T)c.cohesion;
  const T shifted = tr - cohesion;
  const T q = pl[1];
  const T deg = T(0.017453292519943295);
  const T angle = (T)c.friction_angle + (T(9) * deg * q - T(10) * deg) * exp(T(-0.2) * q);
  const T sine = sin(angle);
  const T alpha = sqrt(T(2) / T(3)) * (T(2) * sine) / (T(3) - sine);
  T logs[3], inc;
  // the isotropic test uses the f32-aware threshold 4 * FLT_EPSILON on both paths
  if (shifted > T(0) || norm <= T(4.76837158203125e-7)) {
    for (int a = 0; a < 3; ++a) logs[a] = cohesion / T(3);
    inc = sqrt(strain[0] * strain[0] + strain[1] * strain[1] + strain[2] * strain[2]);
  } else {
    const T lambda = (T)c.lambda, mu = (T)c.mu;
    const T gamma = norm + (T(3) * lambda + T(2) * mu) / (T(2) * mu) * shifted * alpha;
    if (gamma <= T(0)) return false;
    for (int a = 0; a < 3; ++a) logs[a] = strain[a] - dev[a] * (gamma / norm);
    inc = gamma;
  }
  for (int a = 0; a < 3; ++a) out[a] = exp(logs[a]);
  T old_det = s[0] * s[1] * s[2], new_det = out[0] * out[1] * out[2];
  if (!nh) {
    old_det = fmax(old_det, T(1e-12));
    new_det = fmax(new_det, T(1e-12));
  }
  pl[0] *= old_det / new_det;
  if (!nh) pl[0] = fmin(fmax(pl[0], T(0.01)), T(100));
  pl[2] += log(old_det) - log(new_det);
  pl[1] += inc;
  return true;
}

// One cyclic Jacobi rotation zeroing A[p][q] of the symmetric A; V collects
// the rotations (its columns are A's eigenvectors).  gpu.wgsl jacobi_pair.
__device__ __forceinline__ void jacobi_rotate(float A[3][3], float V[3][3], int p, int q) {
  const float off = A[p][q];
  if (fabsf(off) < 1e-8f) return;
  const float tau = (A[q][q] - A[p][p]) / (2.0f * off);
  const float t = (tau >= 0.0f ? 1.0f : -1.0f) / (fabsf(tau) + sqrtf(1.0f + tau * tau));
  const float cs = 1.0f / sqrtf(1.0f + t * t), sn = t * cs;
  const float app = A[p][p], aqq = A[q][q];
  const int r = 3 - p - q;
  const float arp = A[r][p], arq = A[r][q];
  A[r][p] = A[p][r] = cs * arp - sn * arq;
  A[r][q] = A[q][r] = sn * arp + cs * arq;
  A[p][p] = app - t * off;
  A[q][q] = aqq + t * off;
  A[p][q] = A[q][p] = 0.0f;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float vp = V[i][p], vq = V[i][q];
    V[i][p] = cs * vp - sn * vq;
    V[i][q] = sn * vp + cs * vq;
  }
}

// Projects F and the plastic state in place.  f32 path: 8 Jacobi sweeps of
// F^T F give the right singular axes V and stretches s, and
// F V diag(s'/s) V^T replaces the stretches (gpu.wgsl principal_axes,
// replace_stretches).  Squaring F squares its condition number, so F with a
// stretch below 1e-6 or a stretch ratio above 8 takes an f64 SVD instead and
// is rebuilt as U diag(s') V^T, as the reference's CPU path does.
__device__ void project_plastic(const DevClass& c, M3& F, float pl[3]) {
  float A[3][3], V[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      A[r][q] = F.at(0, r) * F.at(0, q) + F.at(1, r) * F.at(1, q) + F.at(2, r) * F.at(2, q);
      V[r][q] = r == q ? 1.0f : 0.0f;
    }
#pragma unroll
  for (int sweep = 0; sweep < 8; ++sweep) {
    jacobi_rotate(A, V, 0, 1);
    jacobi_rotate(A, V, 0, 2);
    jacobi_rotate(A, V, 1, 2);
  }
  float sv[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) sv[a] = sqrtf(fmaxf(A[a][a], 0.0f));
  const float smin = fminf(fminf(sv[0], sv[1]), sv[2]), smax = fmaxf(fmaxf(sv[0], sv[1]), sv[2]);
  if (smin >= 1e-6f && smax <= 8.0f * smin) {
    float out[3];
    if (!plastic_stretches<float>(c, sv, pl, out)) return;
    float scale[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) scale[a] = out[a] / sv[a];
    // M = V diag(scale) V^T; F <- F M
    M3 M, Fo = F;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int q = 0; q < 3; ++q)
        M.at(r, q) = V[r][0] * scale[0] * V[q][0] + V[r][1] * scale[1] * V[q][1] + V[r][2] * scale[2] * V[q][2];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int q = 0; q < 3; ++q)
        F.at(r, q) = Fo.at(r, 0) * M.at(0, q) + Fo.at(r, 1) * M.at(1, q) + Fo.at(r, 2) * M.at(2, q);
    return;
  }
  double U[3][3], sd[3], Vd[3][3], pd[3] = {pl[0], pl[1], pl[2]}, out[3];
  svd_fp64(F, U, sd, Vd);
  if (!plastic_stretches<double>(c, sd, pd, out)) return;
  for (int r = 0; r < 3; ++r)
    for (int q = 0; q < 3; ++q)
      F.at(r, q) = (float)(U[r][0] * out[0] * Vd[q][0] + U[r][1] * out[1] * Vd[q][1] + U[r][2] * out[2] * Vd[q][2]);
  for (int a = 0; a < 3; ++a) pl[a] = (float)pd[a];
}

__device__ __forceinline__ void quadratic_weights(float fx, float w[3]) {
  float t0 = 1.5f - fx, t1 = fx - 1.0f, t2 = fx - 0.5f;
  w[0] = 0.5f * t0 * t0;
  w[1] = 0.75f - t1 * t1;
  w[2] = 0.5f * t2 * t2;
}

// Obstacles (obstacle.rs:373-539), f32.
__device__ __forceinline__ float3 qrot(const float q[4], float3 v) {
  float3 qv = make_float3(q[0], q[1], q[2]);
  float3 t = cross3(qv, v) * 2.0f;
  return v + t * q[3] + cross3(qv, t);
}
__device__ __forceinline__ float3 qrot_inv(const float q[4], float3 v) {
  float qc[4] = {-q[0], -q[1], -q[2], q[3]};
  return qrot(qc, v);
}
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

__device__ void obstacle_surface(const DevObstacle& o, float3 point, float& dist, float3& normal) {
  float3 rel = point - make_float3(o.center[0], o.center[1], o.center[2]);
  const float* d = o.dims;
  switch (o.shape) {
    case 0: {  // sphere
      float n = len3(rel);
      normal = n > 1e-12f ? rel / n : make_float3(0, 0, 1);
      dist = n - d[0];
      return;
    }
    case 1: {  // box
      float3 l = qrot_inv(o.q, rel);
      float3 cl = make_float3(clampf(l.x, -d[0], d[0]), clampf(l.y, -d[1], d[1]), clampf(l.z, -d[2], d[2]));
      float3 out = l - cl;
      float n = len3(out);
      if (n > 1e-12f) {
        dist = n;
        normal = qrot(o.q, out / n);
      } else {
        float3 dd = make_float3(d[0] - fabsf(l.x), d[1] - fabsf(l.y), d[2] - fabsf(l.z));
        int axis = (dd.x <= dd.y && dd.x <= dd.z) ? 0 : (dd.y <= dd.z ? 1 : 2);
        float3 nn = make_float3(0, 0, 0);
        set_comp(nn, axis, comp(l, axis) < 0.0f ? -1.0f : 1.0f);
        dist = -comp(dd, axis);
        normal = qrot(o.q, nn);
      }
      return;
    }
    case 2: {  // capsule
      float3 l = qrot_inv(o.q, rel);
      float cz = clampf(l.z, -d[0], d[0]);
      float3 delta = l - make_float3(0, 0, cz);
      float n = len3(delta);
      float eps = fmaxf(d[0], d[1]) * 1e-6f;
      float3 nn = n > eps ? delta / n : make_float3(1, 0, 0);
      dist = n - d[1];
      normal = qrot(o.q, nn);
      return;
    }
    case 3: {  // cylinder
      float3 l = qrot_inv(o.q, rel);
      float hh = d[0], r = d[1];
      float radial = sqrtf(l.x * l.x + l.y * l.y);
      float3 rn = radial > 1e-12f ? make_float3(l.x / radial, l.y / radial, 0.0f) : make_float3(1, 0, 0);
      float rd = radial - r;
      float cd = fabsf(l.z) - hh;
      float eps = fmaxf(hh, r) * 1e-6f;
      float3 nn;
      if (fabsf(rd) <= eps && fabsf(cd) <= eps) {
        float3 capn = make_float3(0, 0, l.z < 0.0f ? -1.0f : 1.0f);
        float3 s = rn + capn;
        dist = fmaxf(rd, cd);
        nn = s / len3(s);
      } else {
        float ro = fmaxf(rd, 0.0f), co = fmaxf(cd, 0.0f);
        float3 out = rn * ro + make_float3(0, 0, l.z < 0.0f ? -co : co);
        float ol = len3(out);
        if (ol > 1e-12f) {
          dist = ol;
          nn = out / ol;
        } else if (rd >= cd) {
          dist = rd;
          nn = rn;
        } else {
          dist = cd;
          nn = make_float3(0, 0, l.z < 0.0f ? -1.0f : 1.0f);
        }
      }
      normal = qrot(o.q, nn);
      return;
    }
    case 4: {  // cone
      float3 l = qrot_inv(o.q, rel);
      float hh = d[0], r = d[1];
      float radial = sqrtf(l.x * l.x + l.y * l.y);
      float3 rn = radial > 1e-12f ? make_float3(l.x / radial, l.y / radial, 0.0f) : make_float3(1, 0, 0);
      float px = radial, pz = l.z;
      float bx = clampf(radial, 0.0f, r), bz = -hh;
      float sx = r, sz = -hh, dx = -r, dz = 2.0f * hh;
      float tt = clampf(((px - sx) * dx + (pz - sz) * dz) / (dx * dx + dz * dz), 0.0f, 1.0f);
      float cx = sx + dx * tt, cz = sz + dz * tt;
      float eps = fmaxf(hh, r) * 1e-6f;
      float bd2 = (px - bx) * (px - bx) + (pz - bz) * (pz - bz);
      float sd2 = (px - cx) * (px - cx) + (pz - cz) * (pz - cz);
      bool use_base = bd2 <= sd2;
      float clx = use_base ? bx : cx, clz = use_base ? bz : cz;
      float ddx = px - clx, ddz = pz - clz;
      float nr = sqrtf(ddx * ddx + ddz * ddz);
      bool inside = l.z >= -hh - eps && l.z <= hh + eps && radial <= r * (hh - l.z) / (2.0f * hh) + eps;
      float snl = sqrtf(4.0f * hh * hh + r * r);
      float snx = 2.0f * hh / snl, snz = r / snl;
      bool corner = bd2 <= eps * eps && sd2 <= eps * eps;
      float fnx, fnz;
      if (corner) {
        float ax = snx, az = -1.0f + snz;
        float al = sqrtf(ax * ax + az * az);
        fnx = ax / al;
        fnz = az / al;
      } else if (use_base) {
        fnx = 0.0f;
        fnz = -1.0f;
      } else {
        fnx = snx;
        fnz = snz;
      }
      float n2x, n2z;
      if (inside || nr <= eps) {
        n2x = fnx;
        n2z = fnz;
      } else {
        n2x = ddx / nr;
        n2z = ddz / nr;
      }
      normal = qrot(o.q, rn * n2x + make_float3(0, 0, n2z));
      dist = inside ? -nr : nr;
      return;
    }
    default: {  // 5: ground
      float3 l = qrot_inv(o.q, rel);
      float ox = l.x - clampf(l.x, -d[0], d[0]);
      float oy = l.y - clampf(l.y, -d[1], d[1]);
      float horizontal = sqrtf(ox * ox + oy * oy);
      if (horizontal <= 1e-12f) {
        dist = l.z;
        normal = qrot(o.q, make_float3(0, 0, 1));
      } else {
        float3 out = make_float3(ox, oy, l.z);
        float dd = len3(out);
        dist = dd;
        normal = qrot(o.q, out / dd);
      }
      return;
    }
  }
}

// Closest point of triangle abc to p (Voronoi-region walk; gpu.wgsl
// closest_point_triangle, obstacle.rs closest_point_triangle).
__device__ float3 closest_point_triangle(float3 p, float3 a, float3 b, float3 c) {
  const float3 ab = b - a, ac = c - a, ap = p - a;
  const float d1 = dot3(ab, ap), d2 = dot3(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) return a;
  const float3 bp = p - b;
  const float d3 = dot3(ab, bp), d4 = dot3(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) return b;
  const float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) return a + ab * (d1 / (d1 - d3));
  const float3 cp = p - c;
  const float d5 = dot3(ab, cp), d6 = dot3(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) return c;
  const float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) return a + ac * (d2 / (d2 - d6));
  const float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && d4 - d3 >= 0.0f && d5 - d6 >= 0.0f)
    return b + (c - b) * ((d4 - d3) / ((d4 - d3) + (d5 - d6)));
  const float r = 1.0f / (va + vb + vc);
  return a + ab * (vb * r) + ac * (vc * r);
}

__device__ __forceinline__ float3 unit3(float3 v) { return v / len3(v); }

// Thin triangular prism in its local frame (obstacle.rs:633-694): the closest
// of its 8 triangular faces; ties within epsilon average the face normals.
__device__ void prism_surface(const float* tri, float ht, float3 point, float& dist, float3& outward) {
  const float3 a = make_float3(tri[0], tri[1], tri[2]);
  const float3 b = make_float3(tri[3], tri[4], tri[5]);
  const float3 c = make_float3(tri[6], tri[7], tri[8]);
  const float3 normal = unit3(cross3(b - a, c - a));
  const float3 off = normal * ht;
  const float3 pr[6] = {a + off, b + off, c + off, a - off, b - off, c - off};
  constexpr int faces[8][3] = {{0, 1, 2}, {3, 5, 4}, {3, 4, 1}, {3, 1, 0}, {4, 5, 2}, {4, 2, 1}, {5, 3, 0}, {5, 0, 2}};
  const float edge = fmaxf(fmaxf(len3(b - a), len3(c - b)), fmaxf(len3(a - c), ht));
  const float eps = edge * 1e-6f;
  bool inside = true;
  float3 closest = make_float3(0.f, 0.f, 0.f), best_n = normal, tied = make_float3(0.f, 0.f, 0.f);
  float best = 1e30f;
#pragma unroll
  for (int f = 0; f < 8; ++f) {
    const float3 va = pr[faces[f][0]], vb = pr[faces[f][1]], vc = pr[faces[f][2]];
    const float3 fn = unit3(cross3(vb - va, vc - va));
    inside = inside && dot3(point - va, fn) <= eps;
    const float3 cand = closest_point_triangle(point, va, vb, vc);
    const float3 dd = point - cand;
    const float d2 = dot3(dd, dd);
    if (d2 + eps * eps < best) {
      best = d2;
      closest = cand;
      best_n = fn;
      tied = fn;
    } else if (fabsf(d2 - best) <= eps * eps) {
      tied = tied + fn;
    }
  }
  const float d = sqrtf(best);
  const float tl = len3(tied);
  const float3 feature = tl > 1e-12f ? tied / tl : best_n;
  outward = (!inside && d > eps) ? (point - closest) / d : feature;
  dist = inside ? -d : d;
}

// Convex polyhedron in its local frame (obstacle.rs:696-745): inside, the
// largest plane distance; outside, the distance to the Dykstra projection
// onto the intersection of the half-spaces (64 sweeps, f32 tolerance as in
// gpu.wgsl convex_surface).  Each plane's correction is kept as the scalar
// it removed along that plane's normal.
__device__ void convex_surface(const float4* planes, int count, float bound, float3 point, float& dist,
                               float3& normal) {
  float maximum = -1e30f;
  float3 nn = make_float3(0.f, 0.f, 1.f);
  for (int i = 0; i < count; ++i) {
    const float4 pl = planes[i];
    const float sd = pl.x * point.x + pl.y * point.y + pl.z * point.z - pl.w;
    if (sd > maximum) {
      maximum = sd;
      nn = make_float3(pl.x, pl.y, pl.z);
    }
  }
  if (maximum <= 0.0f) {
    dist = maximum;
    normal = nn;
    return;
  }
  float corr[kMaxConvexPlanes];
  for (int i = 0; i < count; ++i) corr[i] = 0.0f;
  float3 proj = point;
  const float tol = fmaxf(bound, 1.0f) * 1e-5f;
  for (int it = 0; it < 64; ++it) {
    const float3 prev = proj;
    for (int i = 0; i < count; ++i) {
      const float4 pl = planes[i];
      const float3 n = make_float3(pl.x, pl.y, pl.z);
      const float3 shifted = proj + n * corr[i];
      const float viol = fmaxf(dot3(n, shifted) - pl.w, 0.0f);
      proj = shifted - n * viol;
      corr[i] = viol;
    }
    if (len3(proj - prev) <= tol) break;
  }
  const float3 offset = point - proj;
  const float sd = len3(offset);
  if (sd > tol) {
    dist = sd;
    normal = offset / sd;
  } else {
    dist = maximum;
    normal = nn;
  }
}

// Surface of any obstacle.  SHAPES = false instantiations (no triangle prism
// or convex obstacle in the session) leave those paths out.
template <bool SHAPES>
__device__ __forceinline__ void surface_of(const DevObstacle& o, const float4* planes, float3 point, float& dist,
                                           float3& normal) {
  if (SHAPES && o.shape >= 6) {
    const float3 l = qrot_inv(o.q, point - make_float3(o.center[0], o.center[1], o.center[2]));
    float3 nn;
    if (o.shape == 6) prism_surface(o.tri, o.dims[0], l, dist, nn);
    else convex_surface(planes + o.plane_begin, o.plane_count, o.bound, l, dist, nn);
    normal = qrot(o.q, nn);
    return;
  }
  obstacle_surface(o, point, dist, normal);
}

__device__ __forceinline__ bool obstacle_may_contact(const DevObstacle& o, float3 point, float margin) {
  if (o.shape == 2) {
    float3 rel = point - make_float3(o.center[0], o.center[1], o.center[2]);
    float b = o.dims[0] + o.dims[1] + margin;
    return dot3(rel, rel) <= b * b;
  }
  if (o.shape >= 6) {  // gpu.wgsl obstacle_may_contact: bounding sphere
    float3 rel = point - make_float3(o.center[0], o.center[1], o.center[2]);
    return len3(rel) <= o.bound + margin;
  }
  return true;
}

__device__ __forceinline__ float3 contact_velocity(const DevObstacle& o, float3 point, float3 v, float3 n) {
  float3 c = make_float3(o.center[0], o.center[1], o.center[2]);
  float3 sv = make_float3(o.lin[0], o.lin[1], o.lin[2]) + cross3(make_float3(o.ang[0], o.ang[1], o.ang[2]), point - c);
  if (o.boundary == 3) return v;
  if (o.boundary == 1) return sv;
  float3 rel = v - sv;
  float ns = dot3(rel, n);
  if (o.boundary == 4) return v - n * ns;
  if (ns >= 0.0f) return v;
  float3 tangent = rel - n * ns;
  if (o.boundary == 2) return sv + tangent;
  float ts = len3(tangent);
  float fs = ts > 1e-12f ? fmaxf(1.0f - o.friction * -ns / ts, 0.0f) : 0.0f;
  return sv + tangent * fs;
}

__device__ __forceinline__ bool finite_abs(float v) { return fabsf(v) < 1e30f; }

// ---------------------------------------------------------------------------
// 7. CPIC side masks, before binning (gpu.wgsl update_cpic_colors, cpic_triangle_side)
// ---------------------------------------------------------------------------
// Signed side of `point` on a CPIC triangle if its projection falls inside
// the triangle (barycentric tolerance 1e-5 in f32, as the reference's GPU pass).
__device__ bool cpic_side(const DevObstacle& o, float3 point, float& dist, bool& positive) {
  const float3 l = qrot_inv(o.q, point - make_float3(o.center[0], o.center[1], o.center[2]));
  const float3 a = make_float3(o.tri[0], o.tri[1], o.tri[2]);
  const float3 eb = make_float3(o.tri[3], o.tri[4], o.tri[5]) - a;
  const float3 ec = make_float3(o.tri[6], o.tri[7], o.tri[8]) - a;
  const float3 n = unit3(cross3(eb, ec));
  const float sd = dot3(l - a, n);
  const float3 pr = l - n * sd - a;
  const float d00 = dot3(eb, eb), d01 = dot3(eb, ec), d11 = dot3(ec, ec);
  const float d20 = dot3(pr, eb), d21 = dot3(pr, ec);
  const float den = d00 * d11 - d01 * d01;
  const float v = (d11 * d20 - d01 * d21) / den;
  const float w = (d00 * d21 - d01 * d20) / den;
  if (v >= -1e-5f && w >= -1e-5f && v + w <= 1.0f + 1e-5f) {
    dist = fabsf(sd);
    positive = sd >= 0.0f;
    return true;
  }
  return false;
}

// Per particle: for every CPIC group, the side of the closest covering
// triangle of that group (the first one on equal distance) sets the group's
// bit (world.rs:470-490).  Groups are compared pairwise, so no per-group
// arrays are needed.
__global__ __launch_bounds__(256) void k_cpic_mask(Parts p, Topo t, const DevClass* classes, const DevScene* scenes,
                                                   const DevObstacle* obstacles) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p.n; i += gridDim.x * blockDim.x) {
    const float4 pm = p.pos_mass[i];
    uint32_t mask = 0;
    if (pm.w > 0.0f) {
      const DevScene& sc = scenes[classes[p.cls[i]].region >> 16];
      const float3 x = make_float3(pm.x, pm.y, pm.z);
      for (int j = 0; j < sc.obstacle_count; ++j) {
        const DevObstacle& oj = obstacles[sc.obstacle_begin + j];
        if (oj.cpic_group < 0) continue;
        float dj;
        bool sj;
        if (!cpic_side(oj, x, dj, sj)) continue;
        bool wins = true;
        for (int k = 0; k < sc.obstacle_count && wins; ++k) {
          const DevObstacle& ok = obstacles[sc.obstacle_begin + k];
          if (k == j || ok.cpic_group != oj.cpic_group) continue;
          float dk;
          bool sk;
          if (cpic_side(ok, x, dk, sk) && (k < j ? dk <= dj : dk < dj)) wins = false;
        }
        if (wins && sj) mask |= 1u << oj.cpic_group;
      }
    }
    t.pmask[i] = mask;
  }
}

// ---------------------------------------------------------------------------
// 8. P2G
// ---------------------------------------------------------------------------
__device__ __forceinline__ int tile_index(int x, int y, int z) { return (x * kTileEdge + y) * kTileEdge + z; }
// Tile placement.  A 6^3 tile covers 8 grid blocks (x, y, z each in [0, 4)
// or [4, 6)); sub-box d of particle block k's tile lies on grid block
// slot8[d] and is stored in that grid block's own 216-node area, at the
// offset it has in a tile (sub-boxes of 64, 32, 32, 16, 32, 16, 16, 8 nodes,
// z fastest inside each).  The area of grid block s thus holds, at sub-box d,
// the part of the tile of the particle block s - d that covers it: the grid
// update reads its 8 covering parts at addresses that depend on s alone, in
// the same round trip as the map of which of them exist (grid_rev), instead
// of looking up the covering tiles first.  Stored node by node in tile
// order, a sub-box row was 2 or 4 nodes of a 128-byte line whose other part
// another grid block reads, often on another XCD with its own L2: the tiles
// were fetched ~1.8 times.
__device__ __forceinline__ int tile_store_index(int x, int y, int z) {
  const int dx = x >> 2, dy = y >> 2, dz = z >> 2;
  const int sx = dx ? 2 : 4, sy = dy ? 2 : 4, sz = dz ? 2 : 4;
  const int offset = dx * 144 + dy * (sx * 24) + dz * (sx * sy * 4);
  return offset + ((x & 3) * sy + (y & 3)) * sz + (z & 3);
}

// Grid-velocity layout inside a grid block: the 8 octants of 2^3 nodes (one
// 128-byte line each) one after the other.  A particle block's G2P reads a
// neighbour grid block's first 2 nodes along each axis it crosses, i.e. whole
// octants, never parts of a line that another reader (often on another XCD)
// fetches too.
__device__ __forceinline__ int grid_store_index(int x, int y, int z) {
  return (((x >> 1) << 2) | ((y >> 1) << 1) | (z >> 1)) * 8 + (((x & 1) << 2) | ((y & 1) << 1) | (z & 1));
}

__device__ __forceinline__ void tile_node(int t, int& d, int& local) {
  int x = t / 36, y = (t / 6) % 6, z = t % 6;
  d = ((x >> 2) << 2) | ((y >> 2) << 1) | (z >> 2);
  local = grid_store_index(x & 3, y & 3, z & 3);
}

// P2G scatter.  One workgroup per particle block: every particle computes
// its stress and scatters its 27 quadratic B-spline contributions into a 6^3
// LDS tile.  Contributions are accumulated as 64-bit fixed-point integers in
// both precision modes: gfx950's float LDS atomics run at ~1/30 the rate of
// its integer ones (and f64 ones slow down more than integer ones when lanes
// hit the same node), and integer sums do not depend on atomic order.
//
// Each scaled contribution s = a * w (|s| < 2^45 units, checked per
// particle) is formed in f64 and rounded to the nearest integer by one fused
// multiply-add with M = 1.5 * 2^52: for |s| < 2^51, fma(a, w, M) is
// M + round(s) exactly, and its bit pattern is bits(M) + round(s) as a
// two's-complement integer.  The low word of bits(M) is zero, so one 32-bit
// subtract on the high word leaves round(s) as an int64 for one ds_add_u64.
// The session scales mass to at most 2^43 units, so the unbiased rounding
// error per contribution is at most 2^-44 of the largest particle mass.  A
// block of at most kP2GMaxBlock particles keeps a tile sum below 2^59, and
// the FIXED64 grid update's sum of 8 tiles below 2^62; larger blocks raise
// Capacity.
//
// COPIES = 2 gives the block two tile copies, and the particles of one cell
// in a wavefront alternate between them (p2g_particle): LDS serializes atomics
// of one instruction that hit the same address.  The copies are summed in a
// fixed order.  FP32 mode converts each node's sum to f32; FIXED64 mode stores the
// integers and the grid update sums them exactly.
constexpr uint32_t kP2GMaxBlock = 1u << 14;
constexpr float kFixedBound = 0x1p45f;

// round(a * w) as a 64-bit integer, |a * w| < 2^51 (see above).
__device__ __forceinline__ unsigned long long fixed64_round(double a, double w) {
  const double r = __builtin_fma(a, w, 0x1.8p52);
  return (unsigned long long)__double_as_longlong(r) - 0x4338000000000000ull;
}

// v = s * 2^-32 -> floor(s) as a 64-bit integer: hi = floor(v),
// lo = floor((v - rint(v)) * 2^32) mod 2^32.  v - rint(v) is exact
// (|v| < 2^13) and both halves round the same way, so hi * 2^32 + lo is
// floor(s) for either sign.
__device__ __forceinline__ unsigned long long fixed64_of(float v) {
  // No contraction: the subtraction must see the same rounded v as floor(v)
  // (fused into the caller's product it would see the unrounded one).
#pragma clang fp contract(off)
  const float d = (v - rintf(v)) * 0x1p32f;  // exact, in [-2^31, 2^31]
  int hi, lo;  // v_cvt_flr_i32_f32: floor to integer in one instruction
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(hi) : "v"(v));
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(lo) : "v"(d));
  return ((unsigned long long)(uint32_t)hi << 32) | (uint32_t)lo;
}

// One particle of a P2G block: its stress and its 27 quadratic B-spline
// contributions (mass and APIC momentum) into one of the block's COPIES LDS
// tile copies (tacc), as rounded 64-bit fixed-point integers
// (fixed64_round).  `bad` records a particle outside the block, `range` a
// contribution beyond the fixed-point bound.  snow_p is the snow hardening
// factor (1 otherwise).
//
// LDS serializes the atomics of one instruction that hit the same address,
// i.e. lanes whose particles share a cell and a tile copy.  With two copies a
// lane takes copy (r & 1), r = the number of lower lanes of its wavefront
// whose particle is in the same cell, so the particles of a cell alternate
// between the copies.  The copies' sums are integers, so the choice does not
// change any result.
template <bool PLASTIC, int COPIES>
__device__ __forceinline__ void p2g_particle(unsigned long long* tacc, const DevClass& cls, float4 pm, float4 vv,
                                             const M3& F, const M3& C, float snow_p, int ox, int oy, int oz, float h,
                                             float inv_h, float stress_scale, float mscale, float pscale, bool& bad,
                                             bool& range) {
  int3 cell;
  cell3(pm, inv_h, cell);
  const int lx = cell.x - ox, ly = cell.y - oy, lz = cell.z - oz;
  if ((unsigned)lx > 3u || (unsigned)ly > 3u || (unsigned)lz > 3u) {
    bad = true;
    return;
  }
  unsigned long long* mine = tacc;
  if (COPIES == 2) {
    const uint32_t cid = (uint32_t)((lx << 4) | (ly << 2) | lz);
    uint64_t same = __ballot(true);  // the lanes with a particle in this block
#pragma unroll
    for (int b = 0; b < 6; ++b) {
      const bool bit = (cid >> b) & 1u;
      const uint64_t m = __ballot(bit);
      same &= bit ? m : ~m;
    }
    const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(same >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)same, 0u));
    mine = tacc + (below & 1u) * (4 * kTileNodes);
  }
  const float fx = pm.x * inv_h - (float)cell.x;
  const float fy = pm.y * inv_h - (float)cell.y;
  const float fz = pm.z * inv_h - (float)cell.z;
  float wx[3], wy[3], wz[3];
  quadratic_weights(fx, wx);
  quadratic_weights(fy, wy);
  quadratic_weights(fz, wz);
  M3 S;
  kirchhoff<PLASTIC>(cls, F, C, snow_p, S);
  const float m = pm.w;
  const float sscale = stress_scale * vv.w;
  // node o receives w_o (m v + A (o - fx) h) with A = m C + stress (-dt V0 inv_d):
  // pre-scaled as q + o.x a0 + o.y a1 + o.z a2 (a = h A columns, q = m v - h A fx)
  const float3 a0 = (col(C, 0) * m + col(S, 0) * sscale) * (h * pscale);
  const float3 a1 = (col(C, 1) * m + col(S, 1) * sscale) * (h * pscale);
  const float3 a2 = (col(C, 2) * m + col(S, 2) * sscale) * (h * pscale);
  const float3 q = make_float3(vv.x, vv.y, vv.z) * (m * pscale) - (a0 * fx + a1 * fy + a2 * fz);
  const float ms = m * mscale;
  // weights are at most 1, so |contribution| <= |q| + 2 (|a0| + |a1| + |a2|)
  const float bound = fmaxf(fmaxf(fabsf(q.x), fabsf(q.y)), fabsf(q.z)) +
                      2.0f * (fmaxf(fmaxf(fabsf(a0.x), fabsf(a0.y)), fabsf(a0.z)) +
                              fmaxf(fmaxf(fabsf(a1.x), fabsf(a1.y)), fabsf(a1.z)) +
                              fmaxf(fmaxf(fabsf(a2.x), fabsf(a2.y)), fabsf(a2.z)));
  range |= !(ms < kFixedBound && bound < kFixedBound);
  // the scatter runs in f64: gfx950 issues f64 adds and FMAs at its f32 rate
  const double msd = ms;
  const double q0[3] = {q.x, q.y, q.z};
  const double d0[3] = {a0.x, a0.y, a0.z}, d1[3] = {a1.x, a1.y, a1.z}, d2[3] = {a2.x, a2.y, a2.z};
#pragma unroll
  for (int x = 0; x < 3; ++x) {
#pragma unroll
    for (int y = 0; y < 3; ++y) {
      const double wxy = (double)wx[x] * (double)wy[y];
#pragma unroll
      for (int z = 0; z < 3; ++z) {
        const double w = wxy * (double)wz[z];
        double mom[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) mom[c] = q0[c] + d0[c] * x + d1[c] * y + d2[c] * z;
        const int n = tile_index(lx + x, ly + y, lz + z);
        atomicAdd(&mine[n], fixed64_round(msd, w));
        atomicAdd(&mine[kTileNodes + n], fixed64_round(mom[0], w));
        atomicAdd(&mine[2 * kTileNodes + n], fixed64_round(mom[1], w));
        atomicAdd(&mine[3 * kTileNodes + n], fixed64_round(mom[2], w));
      }
    }
  }
}

// Writes a block's tile from LDS, each sub-box d into the area of grid block
// slots[d] (the block's 8 tile slots): per node the sum over the tile copies,
// as the integers (FIXED64) or converted to f32 (FP32).  Each thread takes
// node pairs (n, n + 1), n even.
template <int FIXED, int THREADS, int COPIES>
__device__ __forceinline__ void p2g_flush(unsigned long long* tacc, const Topo& t, const uint32_t* slots,
                                          const DevScene& sc) {
  constexpr int kTile = 4 * kTileNodes;
  const int tid = threadIdx.x;
  // sums of nodes q and q + 1 (q even) over the copies
  auto pair_sum = [&](int q) -> longlong2 {
    ulonglong2 v = *reinterpret_cast<const ulonglong2*>(tacc + q);
    if (COPIES == 2) {
      const ulonglong2 w = *reinterpret_cast<const ulonglong2*>(tacc + kTile + q);
      v.x += w.x;
      v.y += w.y;
    }
    return make_longlong2((long long)v.x, (long long)v.y);
  };
  for (int n = 2 * tid; n < kTileNodes; n += 2 * THREADS) {
    const longlong2 m = pair_sum(n), px = pair_sum(kTileNodes + n), py = pair_sum(2 * kTileNodes + n),
                    pz = pair_sum(3 * kTileNodes + n);
    // nodes n and n + 1 differ in z only (z even), so they share a sub-box
    // and are adjacent in it
    const int x = n / 36, y = (n / 6) % 6, z = n % 6;
    const int o = tile_store_index(x, y, z);
    const uint32_t slot = slots[((x >> 2) << 2) | ((y >> 2) << 1) | (z >> 2)];
    if (slot == kInvalid) continue;  // grid capacity exhausted (flagged by binning)
    if (FIXED) {
      longlong2* out = reinterpret_cast<longlong2*>(t.ptile) + ((size_t)slot * kTileNodes + o) * 2;
      out[0] = make_longlong2(m.x, px.x);
      out[1] = make_longlong2(py.x, pz.x);
      out[2] = make_longlong2(m.y, px.y);
      out[3] = make_longlong2(py.y, pz.y);
    } else {
      const float im = exp2f(-(float)sc.mass_shift), ip = exp2f(-(float)sc.mom_shift);
      float4* out = reinterpret_cast<float4*>(t.ptile) + (size_t)slot * kTileNodes + o;
      out[0] = make_float4(__ll2float_rn(m.x) * im, __ll2float_rn(px.x) * ip, __ll2float_rn(py.x) * ip,
                           __ll2float_rn(pz.x) * ip);
      out[1] = make_float4(__ll2float_rn(m.y) * im, __ll2float_rn(px.y) * ip, __ll2float_rn(py.y) * ip,
                           __ll2float_rn(pz.y) * ip);
    }
  }
}

template <int FIXED, int THREADS, int COPIES, bool PLASTIC>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(THREADS == 64 && COPIES == 2 ? 3 : 4))) void k_p2g(Parts p, Topo t,
                                                                                         const DevClass* __restrict__ classes,
                                                                                         const DevScene* __restrict__ scenes) {
  constexpr int kTile = 4 * kTileNodes;  // m, px, py, pz planes of one tile copy
  // 16-byte aligned: cleared and read back two nodes (one ds_*_b128) at a time
  __shared__ __attribute__((aligned(16))) unsigned long long tacc[COPIES * kTile];
  static_assert(kTileNodes % 2 == 0 && kTile % 2 == 0, "tile planes hold whole node pairs");
  __shared__ uint32_t s_slot[8];
  const uint32_t npb = t.counters[CNT_PARTICLE_BLOCKS];
  const int tid = threadIdx.x;
  for (uint32_t k = blockIdx.x; k < npb; k += gridDim.x) {
    // The block's particle range, slot list and the particle indices of the
    // first two passes load together, and the first particles' state loads
    // while the header (key, then scene) resolves: three dependent round trips
    // before any compute instead of five, and none for the second pass's
    // indices.
    const uint32_t start = t.pblock_start[k], len = t.pblock_len[k];
    const uint32_t slot = tid < 8 ? t.pblock_slot8[8 * k + tid] : 0u;
    const uint32_t i_first = (uint32_t)tid < len ? t.sorted[start + tid] : 0u;
    const uint32_t i_second = (uint32_t)(tid + THREADS) < len ? t.sorted[start + tid + THREADS] : 0u;
    if (tid < 8) s_slot[tid] = slot;
    for (int q = tid; q < COPIES * kTile / 2; q += THREADS) reinterpret_cast<ulonglong2*>(tacc)[q] = make_ulonglong2(0, 0);
    __syncthreads();
    // uniform over the workgroup: read once into a scalar register, so the
    // block key and the scene parameters load as scalars
    const uint32_t own = __builtin_amdgcn_readfirstlane(s_slot[0]);
    if (own == kInvalid) {  // grid capacity exhausted (flagged by binning)
      __syncthreads();
      continue;
    }
    const int4 key = uniform4(t.grid_key[own]);
    const DevScene& sc = scenes[(uint32_t)key.w >> 16];
    const float h = sc.h, inv_h = sc.inv_h;
    const float stress_scale = -sc.dt * sc.inv_d;
    const float mscale = exp2f((float)sc.mass_shift), pscale = exp2f((float)sc.mom_shift);
    const int ox = key.x * kBlockCells, oy = key.y * kBlockCells, oz = key.z * kBlockCells;
    bool bad = false, range = len > kP2GMaxBlock;
    for (uint32_t j = tid; j < len; j += THREADS) {
      const uint32_t i = j == (uint32_t)tid ? i_first : j == (uint32_t)(tid + THREADS) ? i_second : t.sorted[start + j];
      const float4 pm = p.pos_mass[i];
      const float4 vv = p.vel_vol[i];
      const DevClass& cls = classes[p.cls[i]];
      M3 F, C;
#pragma unroll
      for (int q = 0; q < 9; ++q) {
        F.a[q] = p.F[q * p.stride + i];
        C.a[q] = p.C[q * p.stride + i];
      }
      p2g_particle<PLASTIC, COPIES>(tacc, cls, pm, vv, F, C, PLASTIC && cls.kind == DM_SNOW ? p.P[i] : 1.0f, ox,
                                    oy, oz, h, inv_h, stress_scale, mscale, pscale, bad, range);
    }
    __syncthreads();
    flag_error(t.counters, bad, ERR_INTERNAL);
    flag_scene_error(t, (uint32_t)key.w >> 16, range, ERR_FIXED_RANGE);
    p2g_flush<FIXED, THREADS, COPIES>(tacc, t, s_slot, sc);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// 9. grid update (world.rs:401-417; gpu.wgsl:772-819)
// ---------------------------------------------------------------------------
// Reactions (ObstacleReaction::record, world.rs:198-204): an obstacle
// receives the linear impulse J = (v_before - v_after) m and the angular
// impulse (point - center) x J of every node (gpu.wgsl:1056-1108) and
// particle (gpu.wgsl:1110-1142) it deflects.  Node side: the wavefront of a
// grid block sums its node lanes in a fixed butterfly order and adds the six
// sums to the obstacle's int64 bins.  Particle side: each impulse is rounded
// to the bins' integer units first (reaction_fixed).  Either way the totals
// do not depend on scheduling; k_react_finalize folds the bins.
__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Called by every lane of a wavefront (uniform control flow).
__device__ __forceinline__ void add_reaction(const Topo& t, const DevObstacle& o, uint32_t oidx, uint32_t bin,
                                             float3 lin, float3 ang) {
  const float c[6] = {wave_sum_f(lin.x), wave_sum_f(lin.y), wave_sum_f(lin.z),
                      wave_sum_f(ang.x), wave_sum_f(ang.y), wave_sum_f(ang.z)};
  if (lane_id() == 0) {
    unsigned long long* dst =
        reinterpret_cast<unsigned long long*>(t.react_bins) + ((size_t)oidx * kReactBins + bin) * 6;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      atomicAdd(&dst[k], fixed64_of(c[k] * o.rscale_lin));
      atomicAdd(&dst[3 + k], fixed64_of(c[3 + k] * o.rscale_ang));
    }
  }
}

// One particle's impulse in the obstacle's int64 units (momentum * 2^rshift),
// rounded to nearest: G2P's particle-side reactions.  Per-particle integers
// sum to the same total in any order, so which lane or wave takes a particle
// (the workgroup size, the storage order) changes no reaction.  The scale
// maps the heaviest particle at 64 cells per substep to 2^43 units; an
// impulse past 2^51 units (a surface over 16,000 cells per substep faster
// than the particle) saturates there instead of wrapping, and `clamped`
// records it (the session counts such impulses).  v * 2^rshift is exact in
// f64 (a power-of-two scale), so clamping it first and adding the rounding
// constant rounds once, like fixed64_round.
__device__ __forceinline__ unsigned long long reaction_unit(float v, double w, bool& clamped) {
  constexpr double kMax = 0x1p51 - 1.0;
  const double u = (double)v * w;
  clamped |= !(fabs(u) <= kMax);
  const double x = fmin(fmax(u, -kMax), kMax);
  return (unsigned long long)__double_as_longlong(x + 0x1.8p52) - 0x4338000000000000ull;
}

// Returns whether a component saturated.
__device__ __forceinline__ bool reaction_fixed(const DevObstacle& o, float3 lin, float3 ang,
                                               unsigned long long r[6]) {
  const double wl = (double)o.rscale_lin * 0x1p32, wa = (double)o.rscale_ang * 0x1p32;
  bool clamped = false;
  r[0] = reaction_unit(lin.x, wl, clamped);
  r[1] = reaction_unit(lin.y, wl, clamped);
  r[2] = reaction_unit(lin.z, wl, clamped);
  r[3] = reaction_unit(ang.x, wa, clamped);
  r[4] = reaction_unit(ang.y, wa, clamped);
  r[5] = reaction_unit(ang.z, wa, clamped);
  return clamped;
}

// One lane's rounded impulse straight into the obstacle's bins (divergent
// control flow allowed): obstacles past the first kMaxReactObstacles of a
// scene, which have no workgroup sums in G2P.
__device__ __forceinline__ void add_reaction_lane(const Topo& t, uint32_t oidx, uint32_t bin,
                                                  const unsigned long long r[6]) {
  unsigned long long* dst = reinterpret_cast<unsigned long long*>(t.react_bins) + ((size_t)oidx * kReactBins + bin) * 6;
#pragma unroll
  for (int c = 0; c < 6; ++c)
    if (r[c]) atomicAdd(&dst[c], r[c]);
}

template <int FIXED, int REACT, bool SHAPES>
__global__ __launch_bounds__(256) void k_grid_update(Topo t, const DevScene* __restrict__ scenes,
                                                     const DevObstacle* __restrict__ obstacles) {
  const uint32_t ng = t.counters[CNT_GRID_BLOCKS];
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t waves = blockDim.x >> 6;
  const int lx = (int)(lane >> 4), ly = (int)((lane >> 2) & 3), lz = (int)(lane & 3);
  // the wavefront's grid block, uniform over it (scalar loads of its key and scene)
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (uint32_t s = blockIdx.x * waves + wave; s < ng; s += gridDim.x * waves) {
    const int4 key = uniform4(t.grid_key[s]);
    const DevScene& sc = scenes[(uint32_t)key.w >> 16];
    // Sum the covering tile parts in the fixed order d = 0..7 (the part of
    // block key - d's tile, node at tile coordinates local + 4 d), all stored
    // in this grid block's area.  The parts load together with grid_rev,
    // which says which of them this substep wrote; a part with no particle
    // block adds nothing (selected, not multiplied, so stale values never
    // enter the sum: + 0 leaves every sum's bits unchanged, and no sum here
    // is -0, since tile values are converted integers).
    float m = 0.0f;
    float3 mom = make_float3(0.f, 0.f, 0.f);
    long long im = 0, ix = 0, iy = 0, iz = 0;
#pragma unroll
    for (int d = 0; d < 8; ++d) {
      const int dx = d >> 2, dy = (d >> 1) & 1, dz = d & 1;
      const bool has = t.grid_rev[8 * s + d] != kInvalid;
      const int tx = lx + 4 * dx, ty = ly + 4 * dy, tz = lz + 4 * dz;
      if (tx >= kTileEdge || ty >= kTileEdge || tz >= kTileEdge) continue;
      const size_t n = (size_t)s * kTileNodes + tile_store_index(tx, ty, tz);
      if (FIXED) {
        const longlong2* src = reinterpret_cast<const longlong2*>(t.ptile) + 2 * n;
        const longlong2 u = src[0], w = src[1];
        im += has ? u.x : 0ll;
        ix += has ? u.y : 0ll;
        iy += has ? w.x : 0ll;
        iz += has ? w.y : 0ll;
      } else {
        const float4 a = reinterpret_cast<const float4*>(t.ptile)[n];
        m += has ? a.x : 0.0f;
        mom.x += has ? a.y : 0.0f;
        mom.y += has ? a.z : 0.0f;
        mom.z += has ? a.w : 0.0f;
      }
    }
    if (FIXED) {
      const float sm = exp2f(-(float)sc.mass_shift), sp = exp2f(-(float)sc.mom_shift);
      m = __ll2float_rn(im) * sm;
      mom = make_float3(__ll2float_rn(ix) * sp, __ll2float_rn(iy) * sp, __ll2float_rn(iz) * sp);
    }
    float3 v = make_float3(0, 0, 0);
    const bool live = m > 0.0f;
    const int cx = key.x * kBlockCells + lx;
    const int cy = key.y * kBlockCells + ly;
    const int cz = key.z * kBlockCells + lz;
    const float3 pos = make_float3((float)cx * sc.h, (float)cy * sc.h, (float)cz * sc.h);
    if (live) {
      v = mom / m + make_float3(sc.gravity[0], sc.gravity[1], sc.gravity[2]) * sc.dt;
      if (sc.has_bounds) {
        for (int a = 0; a < 3; ++a) {
          float pa = comp(pos, a), va = comp(v, a);
          if ((pa <= sc.bmin[a] + sc.h && va < 0.0f) || (pa >= sc.bmax[a] - sc.h && va > 0.0f)) set_comp(v, a, 0.0f);
        }
      }
    }
    for (int oi = 0; oi < sc.obstacle_count; ++oi) {  // uniform over the wavefront (one scene)
      const uint32_t oidx = sc.obstacle_begin + oi;
      const DevObstacle& o = obstacles[oidx];
      bool contact = false;
      float dist = 0.0f;
      float3 nrm = make_float3(0.f, 0.f, 0.f);
      const float3 before = v;
      if (live && obstacle_may_contact(o, pos, sc.h)) {
        surface_of<SHAPES>(o, t.planes, pos, dist, nrm);
        if (dist <= sc.h * (1.0f + 1e-5f)) {
          contact = true;
          v = contact_velocity(o, pos, v, nrm);
        }
      }
      if (REACT && __any(contact)) {
        const float3 J = contact ? (before - v) * m : make_float3(0.f, 0.f, 0.f);
        const float3 point = pos - nrm * dist;
        const float3 L = contact ? cross3(point - make_float3(o.center[0], o.center[1], o.center[2]), J)
                                 : make_float3(0.f, 0.f, 0.f);
        add_reaction(t, o, oidx, s % kReactBins, J, L);
      }
    }
    t.grid_vel[(size_t)s * kBlockNodes + grid_store_index(lx, ly, lz)] = make_float4(v.x, v.y, v.z, 0.0f);
  }
}

// ---------------------------------------------------------------------------
// 10. G2P (world.rs:419-464; gpu.wgsl:821-1019)
// ---------------------------------------------------------------------------
// REACT = 1 also records the particle-side obstacle reactions: each contact's
// impulse, rounded to the obstacle's int64 units, goes into per-workgroup LDS
// sums that are added to the obstacle's bins once per block.  Any workgroup
// size computes the same bits.
// G2P waits on memory most of its time, so waves per SIMD matter: the plain
// variants fit 96 VGPRs without spilling (5 waves; 64 us instead of 68 on
// the 1M scene at 4 waves), while shapes and plastics would spill there.
//
// XCD-aware block order.  Consecutive particle blocks share the cache lines
// of the particle arrays at their boundaries, and G2P both reads and writes
// those lines.  Workgroup w runs on XCD w % 8 (round-robin dispatch) and each
// XCD has its own L2, so with block k on workgroup k every boundary line was
// fetched, and partly written back, by two XCDs.  Here runs of kXcdChunk
// consecutive blocks go to one XCD, the runs rotating over the XCDs, so only
// run boundaries are shared while every XCD still gets an even sample of the
// scene (one contiguous eighth per XCD was 2% slower on uneven workloads: a
// batch of scenes, the dam break).  A single pass when the grid covers npb
// rounded up to whole rotations; the grid-stride order below kXcdMinBlocks
// blocks or on a smaller grid.  Measured on the 1M-particle four-cube scene:
// the substep 1% faster, bit-identical results; P2G, which reads those lines
// but writes only its own tile, keeps block k on workgroup k.
constexpr uint32_t kXcds = 8, kXcdChunk = 8, kXcdMinBlocks = 1024;
__device__ __forceinline__ bool xcd_order(uint32_t npb, uint32_t grid, uint32_t& span) {
  constexpr uint32_t g = kXcds * kXcdChunk;
  span = (npb + g - 1) / g * g;
  return npb >= kXcdMinBlocks && grid >= span;
}
__device__ __forceinline__ uint32_t xcd_span(uint32_t npb, uint32_t grid) {
  uint32_t span;
  return xcd_order(npb, grid, span) ? span : npb;
}
__device__ __forceinline__ uint32_t xcd_block(uint32_t v, uint32_t npb, uint32_t grid) {
  uint32_t span;
  if (!xcd_order(npb, grid, span)) return v;
  const uint32_t x = v % kXcds, i = v / kXcds;
  return ((i / kXcdChunk) * kXcds + x) * kXcdChunk + i % kXcdChunk;
}

// Reorder.  In the MOVE variant (q set), G2P also moves every particle's whole state
// into q, into its block's range [start, start + len): the next substeps read
// particles in nearly sequential order.  Inside the range the particles go
// round-robin by cell, the cell of their position before this substep's
// update (first one particle of each of the block's cells, then a second, ...;
// blocks of more than kMoveOrderMax particles, or of more than kMoveLevels in
// one cell, keep their order), so that a P2G wavefront's lanes rarely share a
// cell, whose atomics LDS would serialize.  Slots past the binned count hold
// no particle afterwards (retired particles were not binned).  P2G has read
// the unmoved state through the sorted list; its sums do not depend on the
// order of the particles, so results do not depend on which substep reorders.
// Otherwise q is empty and p is updated in place.
constexpr uint32_t kMoveOrderMax = 1024;
constexpr uint32_t kMoveLevels = 128;  // particles per cell beyond which a block keeps its order

template <int REACT, bool SHAPES, bool PLASTIC, bool MOVE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(!SHAPES && !PLASTIC ? 5 : 1))) void k_g2p(
    Parts p, Parts q, Topo t, const DevClass* __restrict__ classes, const DevScene* __restrict__ scenes,
    const DevObstacle* __restrict__ obstacles) {
  __shared__ float vt[3 * kTileNodes];
  __shared__ uint32_t s_slot[8];
  // MOVE: particles per cell of the block (64 cells, then particles outside
  // it); per particle its cell and rank in the cell, then its offset in the
  // block's range; per rank level the cells that reach it and the offset
  // where the level starts
  __shared__ uint32_t s_cnt[MOVE ? 65 : 1];
  __shared__ uint32_t s_dest[MOVE ? kMoveOrderMax : 1];
  __shared__ unsigned long long s_level_cells[MOVE ? kMoveLevels : 1];
  __shared__ uint32_t s_level_start[MOVE ? kMoveLevels : 1];
  __shared__ uint32_t s_levels;
  // Particle-side reactions (REACT) on a scene's first kMaxReactObstacles
  // obstacles: each contact adds its rounded impulse (reaction_fixed) into
  // these workgroup sums, which go to the obstacle's bins once per block.
  constexpr int kReactSums = REACT ? kMaxReactObstacles * 6 : 1;
  __shared__ unsigned long long rsum[kReactSums];
  const uint32_t npb = t.counters[CNT_PARTICLE_BLOCKS];
  constexpr bool move = MOVE;  // a separate variant: the plain one keeps its registers
  const Parts o = move ? q : p;  // where updated particles are written
  if (move) {
    const uint32_t binned = min(t.counters[CNT_BINNED], p.n);
    for (uint32_t i = binned + blockIdx.x * blockDim.x + threadIdx.x; i < p.n; i += gridDim.x * blockDim.x) {
      q.pos_mass[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      q.ins_id[i] = kInvalid;
    }
  }
  // The block hash is not read after the scatter of this substep: clear it
  // here for the next substep's binning (saves a launch).
  {
    const uint32_t nb = t.bucket_mask + 1;
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
      t.owner[b] = 0;
      t.bucket_count[b] = 0;
      t.bkey[b].x = kNoKey;
    }
  }
  const uint32_t span = xcd_span(npb, gridDim.x);
  for (uint32_t v = blockIdx.x; v < span; v += gridDim.x) {
    const uint32_t k = xcd_block(v, npb, gridDim.x);
    if (k >= npb) continue;  // uniform over the workgroup
    // As in P2G: the block's particle range, slot list and the particle
    // indices of the first two passes load together, and the block key loads
    // with the velocity tile.
    const uint32_t start = t.pblock_start[k], len = t.pblock_len[k];
    const uint32_t slot0 = threadIdx.x < 8 ? t.pblock_slot8[8 * k + threadIdx.x] : 0u;
    const uint32_t i_first = threadIdx.x < len ? t.sorted[start + threadIdx.x] : 0u;
    const uint32_t i_second = threadIdx.x + blockDim.x < len ? t.sorted[start + threadIdx.x + blockDim.x] : 0u;
    if (threadIdx.x < 8) s_slot[threadIdx.x] = slot0;
    if (REACT)  // the previous block's sums were flushed before the loop's last barrier
      for (int q = threadIdx.x; q < kReactSums; q += blockDim.x) rsum[q] = 0ull;
    __syncthreads();
    // uniform over the workgroup: the block key, the scene and its obstacles
    // then load as scalars
    const uint32_t own = __builtin_amdgcn_readfirstlane(s_slot[0]);
    const int4 key = own != kInvalid ? uniform4(t.grid_key[own]) : make_int4(0, 0, 0, 0);
    for (int n = threadIdx.x; n < kTileNodes; n += blockDim.x) {
      int d, local;
      tile_node(n, d, local);
      uint32_t slot = s_slot[d];
      float4 v = make_float4(0, 0, 0, 0);
      if (own != kInvalid && slot != kInvalid) v = t.grid_vel[(size_t)slot * kBlockNodes + local];
      vt[n] = v.x;
      vt[kTileNodes + n] = v.y;
      vt[2 * kTileNodes + n] = v.z;
    }
    __syncthreads();
    if (own != kInvalid) {
      const DevScene& sc = scenes[(uint32_t)key.w >> 16];
      const float h = sc.h, inv_h = sc.inv_h, dt = sc.dt, inv_d = sc.inv_d;
      const int ox = key.x * kBlockCells, oy = key.y * kBlockCells, oz = key.z * kBlockCells;
      bool failed = false, bad = false;
      const bool spread = move && len <= kMoveOrderMax;  // uniform over the workgroup
      if (spread) {
        for (uint32_t c = threadIdx.x; c < 65; c += blockDim.x) s_cnt[c] = 0;
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < len; j += blockDim.x) {
          const uint32_t i = j == threadIdx.x ? i_first : j == threadIdx.x + blockDim.x ? i_second : t.sorted[start + j];
          int3 cell;
          cell3(p.pos_mass[i], inv_h, cell);
          const int lx = cell.x - ox, ly = cell.y - oy, lz = cell.z - oz;
          const uint32_t lc = ((unsigned)lx > 3u || (unsigned)ly > 3u || (unsigned)lz > 3u)
                                  ? 64u
                                  : (uint32_t)((lx << 4) | (ly << 2) | lz);
          s_dest[j] = (lc << 16) | atomicAdd(&s_cnt[lc], 1u);
        }
        __syncthreads();
        // The order is by (rank r, cell c): level r holds the rank-r particle
        // of every cell with more than r particles, in cell order (cell 64,
        // the particles outside the block, last).  The first wavefront, one
        // lane per cell, lists each level's cells and where the level starts.
        if (threadIdx.x < 64) {
          const uint32_t n = s_cnt[threadIdx.x], outside = s_cnt[64];
          uint32_t level = 0, at = 0;
          for (; level < kMoveLevels; ++level) {
            const uint64_t cells = __ballot(n > level);
            if (cells == 0 && outside <= level) break;  // uniform over the wavefront
            if (threadIdx.x == 0) {
              s_level_cells[level] = cells;
              s_level_start[level] = at;
            }
            at += (uint32_t)__popcll(cells) + (outside > level ? 1u : 0u);
          }
          const uint64_t beyond = __ballot(n > level);  // after a full loop: cells past the last level
          if (threadIdx.x == 0) s_levels = level < kMoveLevels || (beyond == 0 && outside <= level) ? level : kMoveLevels + 1;
        }
        __syncthreads();
        const bool by_level = s_levels <= kMoveLevels;  // else a cell holds too many particles
        for (uint32_t j = threadIdx.x; j < len; j += blockDim.x) {
          const uint32_t lc = s_dest[j] >> 16, r = min(s_dest[j] & 0xffffu, kMoveLevels - 1);
          const unsigned long long lower = lc < 64 ? (1ull << lc) - 1ull : ~0ull;
          // entry j is only read by this thread
          s_dest[j] = by_level ? s_level_start[r] + (uint32_t)__popcll(s_level_cells[r] & lower) : j;
        }
      }
      for (uint32_t j = threadIdx.x; j < len; j += blockDim.x) {
        const uint32_t i = j == threadIdx.x ? i_first : j == threadIdx.x + blockDim.x ? i_second : t.sorted[start + j];
        // the slot the particle is written to
        const uint32_t w = move ? start + (spread ? s_dest[j] : j) : i;
        const float4 pm = p.pos_mass[i];
        const uint32_t ci = p.cls[i];
        const DevClass& cl = classes[ci];
        // a moved particle takes along the state this kernel does not rewrite
        auto carry = [&](bool with_plastic) {
          o.cls[w] = ci;
          o.ins_id[w] = p.ins_id[i];
          if (with_plastic && p.P) {
#pragma unroll
            for (int a = 0; a < 3; ++a) o.P[a * o.stride + w] = p.P[a * p.stride + i];
          }
        };
        if (cl.fixed) {
          const float4 vv = p.vel_vol[i];
          o.vel_vol[w] = make_float4(0.f, 0.f, 0.f, vv.w);
          if (sc.apply_force) p.force[p.ins_id[i]] = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
          for (int a = 0; a < 9; ++a) o.C[a * o.stride + w] = 0.0f;
          if (move) {
            o.pos_mass[w] = pm;
#pragma unroll
            for (int a = 0; a < 9; ++a) o.F[a * o.stride + w] = p.F[a * p.stride + i];
            carry(true);
          }
          continue;
        }
        int3 cell;
        cell3(pm, inv_h, cell);
        const int lx = cell.x - ox, ly = cell.y - oy, lz = cell.z - oz;
        if ((unsigned)lx > 3u || (unsigned)ly > 3u || (unsigned)lz > 3u) {
          bad = true;
          if (move) {  // binning disagrees with this block: keep the particle as it was
            o.pos_mass[w] = pm;
            o.vel_vol[w] = p.vel_vol[i];
#pragma unroll
            for (int a = 0; a < 9; ++a) {
              o.C[a * o.stride + w] = p.C[a * p.stride + i];
              o.F[a * o.stride + w] = p.F[a * p.stride + i];
            }
            carry(true);
          }
          continue;
        }
        const float fx = pm.x * inv_h - (float)cell.x;
        const float fy = pm.y * inv_h - (float)cell.y;
        const float fz = pm.z * inv_h - (float)cell.z;
        float wx[3], wy[3], wz[3];
        quadratic_weights(fx, wx);
        quadratic_weights(fy, wy);
        quadratic_weights(fz, wz);
        // v = sum w v_i; C = sum v_i (w (o - fx) h inv_d)^T, with h inv_d applied once
        float3 v = make_float3(0, 0, 0), c0 = v, c1 = v, c2 = v;
        const float ex[3] = {-fx, 1.0f - fx, 2.0f - fx};
        const float ey[3] = {-fy, 1.0f - fy, 2.0f - fy};
        const float ez[3] = {-fz, 1.0f - fz, 2.0f - fz};
        const int n0 = tile_index(lx, ly, lz);
#pragma unroll 1
        for (int x = 0; x < 3; ++x) {
#pragma unroll
          for (int y = 0; y < 3; ++y) {
            const float wxy = wx[x] * wy[y];
#pragma unroll
            for (int z = 0; z < 3; ++z) {
              const int n = n0 + (x * kTileEdge + y) * kTileEdge + z;
              const float w = wxy * wz[z];
              const float3 nvw = make_float3(vt[n], vt[kTileNodes + n], vt[2 * kTileNodes + n]) * w;
              v = v + nvw;
              c0 = c0 + nvw * ex[x];
              c1 = c1 + nvw * ey[y];
              c2 = c2 + nvw * ez[z];
            }
          }
        }
        const float hd = h * inv_d;
        c0 = c0 * hd;
        c1 = c1 * hd;
        c2 = c2 * hd;
        if (sc.apply_force) {
          const uint32_t ii = p.ins_id[i];
          const float4 f = p.force[ii];
          v = v + make_float3(f.x, f.y, f.z) / pm.w * dt;
          p.force[ii] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        v = v * expf(-cl.damping * dt);
        float3 x = make_float3(pm.x, pm.y, pm.z) + v * dt;
        const float radius = cl.radius;
        if (sc.has_bounds) {
          for (int a = 0; a < 3; ++a) {
            float lower = sc.bmin[a] + radius, upper = sc.bmax[a] - radius;
            float xa = comp(x, a), va = comp(v, a);
            if (xa < lower) {
              set_comp(x, a, lower);
              set_comp(v, a, fmaxf(va, 0.0f));
            } else if (xa > upper) {
              set_comp(x, a, upper);
              set_comp(v, a, fminf(va, 0.0f));
            }
          }
        }
        for (int oi = 0; oi < sc.obstacle_count; ++oi) {
          const DevObstacle& o = obstacles[sc.obstacle_begin + oi];
          if (!obstacle_may_contact(o, x, radius)) continue;
          float dist;
          float3 nrm;
          surface_of<SHAPES>(o, t.planes, x, dist, nrm);
          if (dist < radius) {
            x = x + nrm * (radius - dist);
            const float3 before = v;
            v = contact_velocity(o, x, v, nrm);
            if (REACT) {
              const float3 J = (before - v) * pm.w;
              const float3 L = cross3(x - nrm * radius - make_float3(o.center[0], o.center[1], o.center[2]), J);
              unsigned long long r[6];
              if (reaction_fixed(o, J, L, r)) atomicAdd(&t.counters[CNT_REACT_SATURATED], 1u);
              if (oi < kMaxReactObstacles) {
#pragma unroll
                for (int c = 0; c < 6; ++c)
                  if (r[c]) atomicAdd(&rsum[oi * 6 + c], r[c]);
              } else {
                add_reaction_lane(t, sc.obstacle_begin + oi, k % kReactBins, r);
              }
            }
          }
        }
        // F <- (I + C dt) F
        M3 Fo, Fn;
#pragma unroll
        for (int q = 0; q < 9; ++q) Fo.a[q] = p.F[q * p.stride + i];
        const float3 cc[3] = {c0, c1, c2};
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            float s = Fo.at(r, q);
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) s += comp(cc[kk], r) * dt * Fo.at(kk, q);
            Fn.at(r, q) = s;
          }
        if (cl.kind == DM_FLUID) {
          const float sc3 = cbrtf(fmaxf(det3(Fn), 1e-9f));
#pragma unroll
          for (int q = 0; q < 9; ++q) Fn.a[q] = (q % 4 == 0) ? sc3 : 0.0f;
        }
        bool plastic_written = false;
        if (PLASTIC && dm_plastic(cl.kind)) {
          float pl[3] = {p.P[i], p.P[p.stride + i], p.P[2 * p.stride + i]};
          project_plastic(cl, Fn, pl);
          o.P[w] = pl[0];
          o.P[o.stride + w] = pl[1];
          o.P[2 * o.stride + w] = pl[2];
          plastic_written = true;
        }
        if (move) carry(!plastic_written);
        // Resident validity (gpu.wgsl:946-993).
        int3 nb;
        bool ok = cell3(make_float4(x.x, x.y, x.z, 0.f), inv_h, nb);
        bool fin = finite_abs(x.x) && finite_abs(x.y) && finite_abs(x.z) && finite_abs(v.x) && finite_abs(v.y) &&
                   finite_abs(v.z);
#pragma unroll
        for (int q = 0; q < 9; ++q) fin = fin && finite_abs(Fn.a[q]);
        fin = fin && finite_abs(c0.x) && finite_abs(c0.y) && finite_abs(c0.z) && finite_abs(c1.x) &&
              finite_abs(c1.y) && finite_abs(c1.z) && finite_abs(c2.x) && finite_abs(c2.y) && finite_abs(c2.z);
        bool out = !ok || nb.x < sc.first_base[0] || nb.y < sc.first_base[1] || nb.z < sc.first_base[2] ||
                   nb.x > sc.last_base[0] || nb.y > sc.last_base[1] || nb.z > sc.last_base[2];
        const float density = fmaxf(pm.w / p.vel_vol[i].w, 1e-12f);
        const float wave = sqrtf(cl.wave_numerator / density);
        // sand-NH: wave speed at the current volume, 0.5 h bound (material.rs:388-396)
        const bool cfl = PLASTIC && cl.kind == DM_SAND_NH
                             ? dt * fmaxf(wave * sqrtf(fmaxf(det3(Fn), 1e-6f)), len3(v)) > 0.5f * h * (1.0f + 1e-4f)
                             : dt * (wave + len3(v)) > 0.4f * h * (1.0f + 1e-4f);
        failed |= out || !fin || cfl;
        o.pos_mass[w] = make_float4(x.x, x.y, x.z, pm.w);
        const float vol = p.vel_vol[i].w;
        o.vel_vol[w] = make_float4(v.x, v.y, v.z, vol);
        o.C[0 * o.stride + w] = c0.x; o.C[1 * o.stride + w] = c0.y; o.C[2 * o.stride + w] = c0.z;
        o.C[3 * o.stride + w] = c1.x; o.C[4 * o.stride + w] = c1.y; o.C[5 * o.stride + w] = c1.z;
        o.C[6 * o.stride + w] = c2.x; o.C[7 * o.stride + w] = c2.y; o.C[8 * o.stride + w] = c2.z;
#pragma unroll
        for (int a = 0; a < 9; ++a) o.F[a * o.stride + w] = Fn.a[a];
      }
      flag_scene_error(t, (uint32_t)key.w >> 16, failed, ERR_RESIDENT);
      flag_error(t.counters, bad, ERR_INTERNAL);
      if (REACT) {  // own is uniform over the workgroup: every thread reaches the barrier
        __syncthreads();
        const int sums = min(sc.obstacle_count, kMaxReactObstacles) * 6;
        unsigned long long* bins = reinterpret_cast<unsigned long long*>(t.react_bins);
        for (int q = threadIdx.x; q < sums; q += blockDim.x) {
          const unsigned long long v = rsum[q];
          if (v) atomicAdd(&bins[((size_t)(sc.obstacle_begin + q / 6) * kReactBins + k % kReactBins) * 6 + q % 6], v);
        }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Reactions and rigid coupling
// ---------------------------------------------------------------------------
// Folds each obstacle's bins (exact int64 sums) into this substep's reaction
// and the running total since the last readback, and clears the bins.
__global__ __launch_bounds__(64) void k_react_finalize(Topo t, const DevObstacle* obstacles) {
  for (uint32_t o = blockIdx.x * blockDim.x + threadIdx.x; o < t.obstacle_total; o += gridDim.x * blockDim.x) {
    const DevObstacle& ob = obstacles[o];
    long long* bins = t.react_bins + (size_t)o * kReactBins * 6;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      long long sum = 0;
      for (int b = 0; b < kReactBins; ++b) {
        sum += bins[b * 6 + c];
        bins[b * 6 + c] = 0;
      }
      const double value = ldexp((double)sum, -(c < 3 ? ob.rshift_lin : ob.rshift_ang));
      t.react_sub[o * 6 + c] = (float)value;
      t.react_total[o * 6 + c] += value;
    }
  }
}

__device__ __forceinline__ float4 quat_mul(float4 a, float4 b) {
  const float3 av = make_float3(a.x, a.y, a.z), bv = make_float3(b.x, b.y, b.z);
  const float3 v = bv * a.w + av * b.w + cross3(av, bv);
  return make_float4(v.x, v.y, v.z, a.w * b.w - dot3(av, bv));
}

// Obstacle pose from its rigid body (gpu_rigid_sync.wgsl:43-62): center =
// body position + R local_center, orientation = q_body q_local, surface
// velocity of the body at the center.
__global__ __launch_bounds__(64) void k_rigid_sync(const DevRigid* bodies, const DevMapping* maps,
                                                  DevObstacle* obstacles, uint32_t count) {
  for (uint32_t o = blockIdx.x * blockDim.x + threadIdx.x; o < count; o += gridDim.x * blockDim.x) {
    const DevMapping mp = maps[o];
    if (mp.body < 0) continue;
    const DevRigid st = bodies[mp.body];
    const float3 offset = qrot(st.q, make_float3(mp.local_center[0], mp.local_center[1], mp.local_center[2]));
    DevObstacle& ob = obstacles[o];
    ob.center[0] = st.pos[0] + offset.x;
    ob.center[1] = st.pos[1] + offset.y;
    ob.center[2] = st.pos[2] + offset.z;
    const float4 q = quat_mul(make_float4(st.q[0], st.q[1], st.q[2], st.q[3]),
                              make_float4(mp.local_q[0], mp.local_q[1], mp.local_q[2], mp.local_q[3]));
    ob.q[0] = q.x;
    ob.q[1] = q.y;
    ob.q[2] = q.z;
    ob.q[3] = q.w;
    const float3 w = make_float3(st.ang[0], st.ang[1], st.ang[2]);
    const float3 lv = make_float3(st.lin[0], st.lin[1], st.lin[2]) + cross3(w, offset);
    ob.lin[0] = lv.x;
    ob.lin[1] = lv.y;
    ob.lin[2] = lv.z;
    ob.ang[0] = w.x;
    ob.ang[1] = w.y;
    ob.ang[2] = w.z;
  }
}

// Applies this substep's reactions of a body's obstacles to it
// (gpu_rigid_reaction.wgsl:72-150): J / m, R diag(I^-1) R^T L, then, for an
// MPM-owned step, gravity, the speed cap and the pose update.
__global__ __launch_bounds__(64) void k_rigid_apply(Topo t, const DevScene* scenes, DevRigid* bodies,
                                                   const DevMapping* maps, const DevObstacle* obstacles,
                                                   const uint32_t* body_scene, uint32_t count) {
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < count; b += gridDim.x * blockDim.x) {
    DevRigid st = bodies[b];
    const float inv_mass = st.inv_mass;
    if (inv_mass <= 0.0f) continue;
    const DevScene& sc = scenes[body_scene[b]];
    const float3 pos = make_float3(st.pos[0], st.pos[1], st.pos[2]);
    float3 J = make_float3(0.f, 0.f, 0.f), L = J;
    bool mapped = false;
    for (int oi = 0; oi < sc.obstacle_count; ++oi) {
      const uint32_t o = sc.obstacle_begin + oi;
      if (maps[o].body != (int32_t)b) continue;
      mapped = true;
      const float* r = t.react_sub + o * 6;
      const float3 imp = make_float3(r[0], r[1], r[2]);
      J = J + imp;
      L = L + make_float3(r[3], r[4], r[5]) +
          cross3(make_float3(obstacles[o].center[0], obstacles[o].center[1], obstacles[o].center[2]) - pos, imp);
    }
    const bool finite = fabsf(J.x) < 1e30f && fabsf(J.y) < 1e30f && fabsf(J.z) < 1e30f && fabsf(L.x) < 1e30f &&
                        fabsf(L.y) < 1e30f && fabsf(L.z) < 1e30f;
    if (!finite || !mapped) continue;
    if (dot3(J, J) == 0.0f && dot3(L, L) == 0.0f && !sc.rigid_integrate) continue;
    const float qc[4] = {-st.q[0], -st.q[1], -st.q[2], st.q[3]};
    const float3 torque_body = qrot(qc, L);
    const float3 dw = qrot(st.q, make_float3(torque_body.x * st.inv_inertia[0], torque_body.y * st.inv_inertia[1],
                                             torque_body.z * st.inv_inertia[2]));
    float3 lin = make_float3(st.lin[0], st.lin[1], st.lin[2]) + J * inv_mass;
    const float3 ang = make_float3(st.ang[0], st.ang[1], st.ang[2]) + dw;
    if (sc.rigid_integrate)
      lin = lin + make_float3(sc.rigid_gravity[0], sc.rigid_gravity[1], sc.rigid_gravity[2]) * sc.dt;
    if (sc.rigid_max_speed > 0.0f) {
      const float speed = len3(lin);
      if (speed > sc.rigid_max_speed) lin = lin * (sc.rigid_max_speed / speed);
    }
    st.lin[0] = lin.x;
    st.lin[1] = lin.y;
    st.lin[2] = lin.z;
    st.ang[0] = ang.x;
    st.ang[1] = ang.y;
    st.ang[2] = ang.z;
    if (sc.rigid_integrate) {
      st.pos[0] += lin.x * sc.dt;
      st.pos[1] += lin.y * sc.dt;
      st.pos[2] += lin.z * sc.dt;
      // q <- normalize(exp(w dt) q), exp as in quat_from_rotation
      const float3 rot = ang * sc.dt;
      const float angle = len3(rot);
      float4 dq;
      if (angle < 1e-5f) {
        const float n = sqrtf(0.25f * dot3(rot, rot) + 1.0f);
        dq = make_float4(0.5f * rot.x / n, 0.5f * rot.y / n, 0.5f * rot.z / n, 1.0f / n);
      } else {
        const float sh = sinf(0.5f * angle) / angle;
        dq = make_float4(rot.x * sh, rot.y * sh, rot.z * sh, cosf(0.5f * angle));
      }
      float4 q = quat_mul(dq, make_float4(st.q[0], st.q[1], st.q[2], st.q[3]));
      const float qn = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
      st.q[0] = q.x / qn;
      st.q[1] = q.y / qn;
      st.q[2] = q.z / qn;
      st.q[3] = q.w / qn;
    }
    bodies[b] = st;
  }
}

// ---------------------------------------------------------------------------
// Snapshot export (gpu_snapshot.wgsl:6-11): 32-byte records, insertion order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_snapshot(Parts p, float4* out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p.n; i += gridDim.x * blockDim.x) {
    uint32_t dst = p.ins_id[i];
    if (dst == kInvalid) continue;  // retired slot
    out[2 * dst] = p.pos_mass[i];
    out[2 * dst + 1] = p.vel_vol[i];
  }
}

// Records of disabled particles (kept on the host; mass 0): {pos, 0; vel, vol}
// at their insertion index, stored as 3 float4 {x, y, z, index bits; ...}.
__global__ __launch_bounds__(256) void k_snapshot_disabled(const float4* rec, uint32_t count, float4* out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
    const float4 a = rec[2 * i], b = rec[2 * i + 1];
    const uint32_t dst = __float_as_uint(a.w);
    out[2 * dst] = make_float4(a.x, a.y, a.z, 0.0f);
    out[2 * dst + 1] = b;
  }
}

// ---------------------------------------------------------------------------
// Batches: retiring a scene, compacting insertion indices, per-scene gather
// ---------------------------------------------------------------------------
// Particles of insertion range [first, first + count) leave the simulation:
// mass 0 (binning skips them; the next reorder drops them) and no index.
__global__ __launch_bounds__(256) void k_retire(Parts p, uint32_t first, uint32_t count) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p.n; i += gridDim.x * blockDim.x) {
    const uint32_t g = p.ins_id[i];
    if (g != kInvalid && g - first < count) {
      p.pos_mass[i].w = 0.0f;
      p.ins_id[i] = kInvalid;
    }
  }
}

// Insertion indices above a removed range move down by its length.
__global__ __launch_bounds__(256) void k_renumber(uint32_t* ins_id, uint32_t n, uint32_t first, uint32_t count) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t g = ins_id[i];
    if (g == kInvalid) continue;
    if (g - first < count) ins_id[i] = kInvalid;
    else if (g >= first + count) ins_id[i] = g - count;
  }
}

// One scene's particles in insertion order, kGatherFloats each: x, y, z,
// mass, vx, vy, vz, volume, C (9, column-major), F (9), plastic state (3;
// 1, 0, 0 when the session stores none).
__global__ __launch_bounds__(256) void k_gather_scene(Parts p, uint32_t first, uint32_t count, float* out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p.n; i += gridDim.x * blockDim.x) {
    const uint32_t g = p.ins_id[i];
    if (g == kInvalid || g - first >= count) continue;
    float* r = out + (size_t)(g - first) * kGatherFloats;
    r[26] = p.P ? p.P[i] : 1.0f;
    r[27] = p.P ? p.P[p.stride + i] : 0.0f;
    r[28] = p.P ? p.P[2 * p.stride + i] : 0.0f;
    const float4 a = p.pos_mass[i], b = p.vel_vol[i];
    r[0] = a.x; r[1] = a.y; r[2] = a.z; r[3] = a.w;
    r[4] = b.x; r[5] = b.y; r[6] = b.z; r[7] = b.w;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      r[8 + q] = p.C[q * p.stride + i];
      r[17 + q] = p.F[q * p.stride + i];
    }
  }
}

// ---------------------------------------------------------------------------
// FP64 audit of the particle state after a substep, per scene: mass,
// momentum, kinetic energy, first moment of mass and angular momentum about
// the origin (the reference's momentum balance, world.rs:784-807, checks the
// same sums on the host).  Every product and sum is f64.  A wavefront whose
// particles belong to one scene (nearly all of them: blocks never mix scenes)
// reduces its lanes with shuffles and adds the sums with one lane; otherwise
// each lane adds its own.  The sums are exact to f64 rounding but their order
// follows the scheduler, so the last bits may differ run to run.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__global__ __launch_bounds__(256) void k_audit(Parts p, const DevClass* classes, double* out) {
  const uint32_t lane = lane_id();
  for (uint32_t i0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); i0 < p.n; i0 += gridDim.x * blockDim.x) {
    const uint32_t i = i0 + lane;
    double v[kAuditFields];
#pragma unroll
    for (int f = 0; f < kAuditFields; ++f) v[f] = 0.0;
    uint32_t scene = kInvalid;
    if (i < p.n) {
      const float4 pm = p.pos_mass[i];
      if (pm.w > 0.0f && p.ins_id[i] != kInvalid) {
        const float4 vv = p.vel_vol[i];
        scene = classes[p.cls[i]].region >> 16;
        const double m = pm.w, x = pm.x, y = pm.y, z = pm.z, vx = vv.x, vy = vv.y, vz = vv.z;
        v[0] = m;
        v[1] = m * vx;
        v[2] = m * vy;
        v[3] = m * vz;
        v[4] = 0.5 * m * (vx * vx + vy * vy + vz * vz);
        v[5] = m * x;
        v[6] = m * y;
        v[7] = m * z;
        v[8] = m * (y * vz - z * vy);
        v[9] = m * (z * vx - x * vz);
        v[10] = m * (x * vy - y * vx);
        v[11] = 1.0;
      }
    }
    const uint64_t valid = __ballot(scene != kInvalid);
    if (valid == 0) continue;
    const uint32_t s0 = readlane_u(scene, __ffsll((unsigned long long)valid) - 1);
    if (__ballot(scene != kInvalid && scene != s0) == 0) {
#pragma unroll
      for (int f = 0; f < kAuditFields; ++f) v[f] = wave_sum_d(v[f]);
      if (lane == 0) {
        double* dst = out + (size_t)s0 * kAuditFields;
#pragma unroll
        for (int f = 0; f < kAuditFields; ++f) atomicAdd(&dst[f], v[f]);
      }
    } else if (scene != kInvalid) {
      double* dst = out + (size_t)scene * kAuditFields;
#pragma unroll
      for (int f = 0; f < kAuditFields; ++f) atomicAdd(&dst[f], v[f]);
    }
  }
}

// ---------------------------------------------------------------------------
// Diagnostic: exact live-node set (union of 27-node stencils) as per-block
// 64-bit masks.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_live_mask(Parts p, Topo t, const DevClass* classes, const DevScene* scenes) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p.n; i += gridDim.x * blockDim.x) {
    uint32_t b = t.pslot[i];
    if (b == kInvalid) continue;
    uint32_t k = t.bucket_pblock[b];
    if (k == kInvalid) continue;
    const float4 pm = p.pos_mass[i];
    const DevClass& cl = classes[p.cls[i]];
    const DevScene& sc = scenes[cl.region >> 16];
    int3 cell;
    cell3(pm, sc.inv_h, cell);
    const int4 key = t.grid_key[t.pblock_slot8[8 * k]];
    const int lx = cell.x - key.x * 4, ly = cell.y - key.y * 4, lz = cell.z - key.z * 4;
    for (int x = 0; x < 3; ++x)
      for (int y = 0; y < 3; ++y)
        for (int z = 0; z < 3; ++z) {
          // grid block d of the tile and the node's row-major index in it (the
          // mask layout the node export decodes)
          const int tx = lx + x, ty = ly + y, tz = lz + z;
          const int d = ((tx >> 2) << 2) | ((ty >> 2) << 1) | (tz >> 2);
          const int local = ((tx & 3) << 4) | ((ty & 3) << 2) | (tz & 3);
          uint32_t slot = t.pblock_slot8[8 * k + d];
          if (slot == kInvalid) continue;
          atomicOr(&t.node_mask[2 * slot + (local >> 5)], 1u << (local & 31));
        }
  }
}

__global__ __launch_bounds__(256) void k_live_count(Topo t) {
  uint32_t ng = t.counters[CNT_GRID_BLOCKS];
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < 2 * ng; q += gridDim.x * blockDim.x) {
    uint32_t c = __popc(t.node_mask[q]);
    if (c) atomicAdd(&t.counters[CNT_LIVE_NODES], c);
  }
}

// ---------------------------------------------------------------------------
// Closed-mesh volume sampling (sampling.rs:47-146): lattice centres whose
// solid angle subtended by the mesh exceeds 2 pi (winding number 1, either
// orientation).  One thread per candidate, f64 throughout, triangles in the
// reference's order; flags[c] = 1 for an accepted candidate.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mesh_inside(const double* __restrict__ verts,
                                                     const uint32_t* __restrict__ tris, uint64_t ntri,
                                                     double lx, double ly, double lz, double spacing, uint64_t cy,
                                                     uint64_t cz, double ux, double uy, double uz, uint64_t count,
                                                     uint8_t* flags) {
  for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < count; c += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t x = c / (cy * cz), y = (c / cz) % cy, z = c % cz;
    const double px = lx + ((double)x + 0.5) * spacing;
    const double py = ly + ((double)y + 0.5) * spacing;
    const double pz = lz + ((double)z + 0.5) * spacing;
    if (px >= ux || py >= uy || pz >= uz) {
      flags[c] = 0;
      continue;
    }
    double angle = 0.0;
    for (uint64_t t = 0; t < ntri; ++t) {
      double v[3][3];
      for (int k = 0; k < 3; ++k) {
        const double* q = verts + 3ull * tris[3 * t + k];
        v[k][0] = q[0] - px;
        v[k][1] = q[1] - py;
        v[k][2] = q[2] - pz;
      }
      auto dot = [](const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
      const double la = sqrt(dot(v[0], v[0])), lb = sqrt(dot(v[1], v[1])), lc = sqrt(dot(v[2], v[2]));
      const double bc[3] = {v[1][1] * v[2][2] - v[1][2] * v[2][1], v[1][2] * v[2][0] - v[1][0] * v[2][2],
                            v[1][0] * v[2][1] - v[1][1] * v[2][0]};
      const double num = dot(v[0], bc);
      const double den = la * lb * lc + dot(v[0], v[1]) * lc + dot(v[1], v[2]) * la + dot(v[2], v[0]) * lb;
      angle += 2.0 * atan2(num, den);
    }
    flags[c] = fabs(angle) > 6.283185307179586 ? 1 : 0;
  }
}

// Closed, consistently oriented surface test of the sampler (sampling.rs:66-93)
// without a host edge map: every directed edge (a, b) of a triangle must occur
// exactly once and its reverse (b, a) exactly once among all 3 * ntri directed
// edges.  One thread per directed edge counts both over every triangle
// (O(ntri^2), for meshes of at most tens of thousands of triangles); a failing
// edge stores 1 into *bad (plain stores of the same value).
__global__ __launch_bounds__(256) void k_mesh_edges(const uint32_t* __restrict__ tris, uint64_t ntri,
                                                    uint32_t* bad) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < 3 * ntri;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t t = e / 3;
    const int k = (int)(e % 3);
    const uint32_t a = tris[3 * t + k], b = tris[3 * t + (k + 1) % 3];
    uint32_t same = 0, rev = 0;
    for (uint64_t u = 0; u < ntri; ++u) {
      const uint32_t p = tris[3 * u], q = tris[3 * u + 1], r = tris[3 * u + 2];
      same += (p == a && q == b) + (q == a && r == b) + (r == a && p == b);
      rev += (p == b && q == a) + (q == b && r == a) + (r == b && p == a);
    }
    if (same != 1u || rev != 1u) *bad = 1u;
  }
}


// ---------------------------------------------------------------------------
// The first n particles' whole state from one particle buffer set to another
// (each with its own column stride): every array in one launch instead of one
// copy per array and column.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_copy_state(Parts dst, Parts src, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    dst.pos_mass[i] = src.pos_mass[i];
    dst.vel_vol[i] = src.vel_vol[i];
    dst.cls[i] = src.cls[i];
    dst.ins_id[i] = src.ins_id[i];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      dst.C[(size_t)q * dst.stride + i] = src.C[(size_t)q * src.stride + i];
      dst.F[(size_t)q * dst.stride + i] = src.F[(size_t)q * src.stride + i];
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) dst.P[(size_t)q * dst.stride + i] = src.P[(size_t)q * src.stride + i];
  }
}

// ---------------------------------------------------------------------------
// Bandwidth probe: a plain float4 copy of known size (reads n*16 B, writes
// n*16 B), used to measure achievable HBM bandwidth and to calibrate
// hardware byte counters.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// Launch wrappers
// ---------------------------------------------------------------------------
static inline uint32_t grid_for(uint64_t items, uint32_t per_block, uint32_t max_blocks) {
  uint64_t b = (items + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > max_blocks) b = max_blocks;
  return (uint32_t)b;
}

void launch_topo_clear(const Topo& t, hipStream_t s) {
  hipLaunchKernelGGL(k_topo_clear, dim3(grid_for(t.bucket_mask + 1ull, 256, 4096)), dim3(256), 0, s, t);
}
void launch_bin_insert(const Parts& p, const Topo& t, const DevClass* c, const DevScene* sc, int cpic,
                       hipStream_t s) {
  // sub-rounds per pass: as many as the workgroup's rounds, at most kInsertK
  // (a 1M-particle binning has 2 rounds; unused sub-rounds would still shuffle)
  const uint32_t rounds = t.bin_span / kBinThreads;
#define TMPM_INSERT(C, K) hipLaunchKernelGGL((k_bin_insert<C, K>), dim3(t.bin_groups), dim3(kBinThreads), 0, s, p, t, c, sc)
  if (cpic) {
    if (rounds >= kInsertK) TMPM_INSERT(true, kInsertK); else if (rounds == 2 || rounds == 3) TMPM_INSERT(true, 2); else TMPM_INSERT(true, 1);
  } else {
    if (rounds >= kInsertK) TMPM_INSERT(false, kInsertK); else if (rounds == 2 || rounds == 3) TMPM_INSERT(false, 2); else TMPM_INSERT(false, 1);
  }
#undef TMPM_INSERT
}
void launch_bin_blocks(const Parts& p, const Topo& t, const DevClass* c, const DevScene* sc, int cpic,
                       hipStream_t s) {
  if (cpic) hipLaunchKernelGGL(k_bin_blocks<true>, dim3(t.bin_groups), dim3(kBinThreads), 0, s, p, t, c, sc);
  else hipLaunchKernelGGL(k_bin_blocks<false>, dim3(t.bin_groups), dim3(kBinThreads), 0, s, p, t, c, sc);
}
void launch_cpic_mask(const Parts& p, const Topo& t, const DevClass* c, const DevScene* sc, const DevObstacle* o,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_cpic_mask, dim3(grid_for(p.n, 256, 8192)), dim3(256), 0, s, p, t, c, sc, o);
}
void launch_bin_offsets(const Parts& p, const Topo& t, const DevClass* c, const DevScene* sc, hipStream_t s) {
  hipLaunchKernelGGL(k_bin_offsets, dim3((t.bin_groups + kOffsetGroups - 1) / kOffsetGroups), dim3(kBinThreads), 0, s,
                     p, t, c, sc);
}
void launch_bin_scatter(const Parts& p, const Topo& t, hipStream_t s) {
  hipLaunchKernelGGL(k_bin_scatter, dim3(grid_for(p.n, 256 * kScatterK, 2048)), dim3(256), 0, s, p, t);
}
void launch_p2g(const Parts& p, const Topo& t, const DevClass* c, const DevScene* sc, int fixed64, uint32_t blocks,
                uint32_t threads, int copies, int plastic, hipStream_t s) {
  // sessions with plastic materials run the default shape (64 threads, 2 tile copies)
  if (plastic) {
    if (fixed64) hipLaunchKernelGGL((k_p2g<1, 64, 2, true>), dim3(blocks), dim3(64), 0, s, p, t, c, sc);
    else hipLaunchKernelGGL((k_p2g<0, 64, 2, true>), dim3(blocks), dim3(64), 0, s, p, t, c, sc);
    return;
  }
  // threads per workgroup: 64 or 128 (anything else runs 128); LDS tile copies: 1 or 2
#define TMPM_P2G(F, T, K) hipLaunchKernelGGL((k_p2g<F, T, K, false>), dim3(blocks), dim3(T), 0, s, p, t, c, sc)
#define TMPM_P2G_T(F, K)                  \
  do {                                    \
    if (threads == 64) TMPM_P2G(F, 64, K); \
    else TMPM_P2G(F, 128, K);              \
  } while (0)
  if (fixed64) {
    if (copies == 2) TMPM_P2G_T(1, 2); else TMPM_P2G_T(1, 1);
  } else {
    if (copies == 2) TMPM_P2G_T(0, 2); else TMPM_P2G_T(0, 1);
  }
#undef TMPM_P2G_T
#undef TMPM_P2G
}
template <bool SHAPES>
static void grid_update_variant(const Topo& t, const DevScene* sc, const DevObstacle* o, int fixed64, int react,
                                uint32_t blocks, hipStream_t s) {
  if (fixed64) {
    if (react) hipLaunchKernelGGL((k_grid_update<1, 1, SHAPES>), dim3(blocks), dim3(256), 0, s, t, sc, o);
    else hipLaunchKernelGGL((k_grid_update<1, 0, SHAPES>), dim3(blocks), dim3(256), 0, s, t, sc, o);
  } else {
    if (react) hipLaunchKernelGGL((k_grid_update<0, 1, SHAPES>), dim3(blocks), dim3(256), 0, s, t, sc, o);
    else hipLaunchKernelGGL((k_grid_update<0, 0, SHAPES>), dim3(blocks), dim3(256), 0, s, t, sc, o);
  }
}
void launch_grid_update(const Topo& t, const DevScene* sc, const DevObstacle* o, int fixed64, int react, int shapes,
                        hipStream_t s) {
  uint32_t blocks = grid_for(t.grid_cap, 4, 16384);
  if (shapes) grid_update_variant<true>(t, sc, o, fixed64, react, blocks, s);
  else grid_update_variant<false>(t, sc, o, fixed64, react, blocks, s);
}
template <bool PLASTIC>
static void g2p_variant(const Parts& p, const Parts& q, const Topo& t, const DevClass* c, const DevScene* sc,
                        const DevObstacle* o, uint32_t blocks, uint32_t threads, int react, int shapes,
                        hipStream_t s) {
  const bool mv = q.pos_mass != nullptr;
#define TMPM_G2P(R, SH, M, T) \
  hipLaunchKernelGGL((k_g2p<R, SH, PLASTIC, M>), dim3(blocks), dim3(T), 0, s, p, q, t, c, sc, o)
#define TMPM_G2P_M(R, SH, T)          \
  do {                                \
    if (mv) TMPM_G2P(R, SH, true, T); \
    else TMPM_G2P(R, SH, false, T);   \
  } while (0)
  if (react) {
    if (shapes) TMPM_G2P_M(1, true, threads);
    else TMPM_G2P_M(1, false, threads);
  } else {
    if (shapes) TMPM_G2P_M(0, true, threads);
    else TMPM_G2P_M(0, false, threads);
  }
#undef TMPM_G2P_M
#undef TMPM_G2P
}
void launch_g2p(const Parts& p, const Parts& q, const Topo& t, const DevClass* c, const DevScene* sc,
                const DevObstacle* o, uint32_t blocks, uint32_t threads, int react, int shapes, int plastic,
                hipStream_t s) {
  if (plastic) g2p_variant<true>(p, q, t, c, sc, o, blocks, threads, react, shapes, s);
  else g2p_variant<false>(p, q, t, c, sc, o, blocks, threads, react, shapes, s);
}
void launch_react_finalize(const Topo& t, const DevObstacle* o, hipStream_t s) {
  hipLaunchKernelGGL(k_react_finalize, dim3(grid_for(t.obstacle_total, 64, 1024)), dim3(64), 0, s, t, o);
}
void launch_rigid_sync(const DevRigid* bodies, const DevMapping* maps, DevObstacle* o, uint32_t count, hipStream_t s) {
  hipLaunchKernelGGL(k_rigid_sync, dim3(grid_for(count, 64, 1024)), dim3(64), 0, s, bodies, maps, o, count);
}
void launch_rigid_apply(const Topo& t, const DevScene* sc, DevRigid* bodies, const DevMapping* maps,
                        const DevObstacle* o, const uint32_t* body_scene, uint32_t count, hipStream_t s) {
  hipLaunchKernelGGL(k_rigid_apply, dim3(grid_for(count, 64, 1024)), dim3(64), 0, s, t, sc, bodies, maps, o,
                     body_scene, count);
}
void launch_snapshot(const Parts& p, float4* out, hipStream_t s) {
  hipLaunchKernelGGL(k_snapshot, dim3(grid_for(p.n, 256, 8192)), dim3(256), 0, s, p, out);
}
void launch_snapshot_disabled(const float4* rec, uint32_t count, float4* out, hipStream_t s) {
  if (count) hipLaunchKernelGGL(k_snapshot_disabled, dim3(grid_for(count, 256, 4096)), dim3(256), 0, s, rec, count, out);
}
void launch_retire(const Parts& p, uint32_t first, uint32_t count, hipStream_t s) {
  hipLaunchKernelGGL(k_retire, dim3(grid_for(p.n, 256, 8192)), dim3(256), 0, s, p, first, count);
}
void launch_renumber(uint32_t* ins_id, uint32_t n, uint32_t first, uint32_t count, hipStream_t s) {
  hipLaunchKernelGGL(k_renumber, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, ins_id, n, first, count);
}
void launch_gather_scene(const Parts& p, uint32_t first, uint32_t count, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_gather_scene, dim3(grid_for(p.n, 256, 8192)), dim3(256), 0, s, p, first, count, out);
}
void launch_audit(const Parts& p, const DevClass* c, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_audit, dim3(grid_for(p.n, 256, 4096)), dim3(256), 0, s, p, c, out);
}
void launch_live_nodes(const Parts& p, const Topo& t, const DevClass* c, const DevScene* sc, hipStream_t s) {
  hipLaunchKernelGGL(k_mask_clear, dim3(grid_for(2ull * t.grid_cap, 256, 4096)), dim3(256), 0, s, t);
  hipLaunchKernelGGL(k_live_mask, dim3(grid_for(p.n, 256, 8192)), dim3(256), 0, s, p, t, c, sc);
  hipLaunchKernelGGL(k_live_count, dim3(grid_for(2ull * t.grid_cap, 256, 4096)), dim3(256), 0, s, t);
}

void launch_mesh_inside(const double* verts, const uint32_t* tris, uint64_t ntri, const double lower[3],
                        const double upper[3], double spacing, const uint64_t counts[3], uint8_t* flags,
                        hipStream_t s) {
  const uint64_t count = counts[0] * counts[1] * counts[2];
  hipLaunchKernelGGL(k_mesh_inside, dim3(grid_for(count, 256, 65536)), dim3(256), 0, s, verts, tris, ntri, lower[0],
                     lower[1], lower[2], spacing, counts[1], counts[2], upper[0], upper[1], upper[2], count, flags);
}
void launch_mesh_edges(const uint32_t* tris, uint64_t ntri, uint32_t* bad, hipStream_t s) {
  hipLaunchKernelGGL(k_mesh_edges, dim3(grid_for(3 * ntri, 256, 4096)), dim3(256), 0, s, tris, ntri, bad);
}
void launch_copy_state(const Parts& dst, const Parts& src, uint32_t n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_copy_state, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, dst, src, n);
}
void launch_copy(const float4* src, float4* dst, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_copy, dim3(grid_for(n, 256 * 8, 16384)), dim3(256), 0, s, src, dst, n);
}

}  // namespace tmpm
