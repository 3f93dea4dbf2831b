// cs_p2g.hip — tile-binned P2G: particles sorted into 8^3-cell tiles, one
// workgroup per tile accumulating its node footprint in LDS, one flush to HBM.
//
// Restates p2g (shader.rs:955-1052).  The reference adds every particle's 162
// fixed-point terms (27 nodes x 6 lanes) to the grid with global i32 atomics.
// Here each active particle is binned by the tile of its base node
// floor(g - 0.5): counts staged per workgroup in LDS, then one global atomic
// per workgroup and tile, give its rank in the tile; a one-block scan turns
// counts into offsets, and a scatter
// writes the tile-sorted particle list.  One workgroup per tile then adds the
// terms of its particles into a 10^3-node x 6-lane LDS block (nodes base-tile
// origin + [0, 10) in each axis) and flushes the nonzero entries with one
// global i32 atomic each: neighbouring tiles share two node layers per axis,
// so the flush adds rather than stores.  Integer addition is associative, so
// the lanes equal the reference's bit for bit whatever the order.
#include "cs_common.h"
#include "cs_kernels.h"

#include "cs_batch.h"
#include "cs_device.h"

namespace {

// Tile of a particle's base node; bases one cell outside the grid (the
// clamped domain reaches g = -0.5 .. n + 0.5) go to the edge tile.
__device__ __forceinline__ uint32_t tile_of(const cs_params& u, int bx, int by, int bz) {
  const uint32_t tx = (uint32_t)min(max(bx, 0), (int)u.gx - 1) / TB;
  const uint32_t ty = (uint32_t)min(max(by, 0), (int)u.gy - 1) / TB;
  const uint32_t tz = (uint32_t)min(max(bz, 0), (int)u.gz - 1) / TB;
  return (tz * u.nby + ty) * u.nbx + tx;
}

}  // namespace

// Pass 1: tile and in-tile rank of every particle, R particles per thread
// (pid = first + r * 256 + thread).  Lanes of a wavefront that share a tile
// form a group; each group adds to a slot of a small LDS table keyed by
// tile, then one global atomic per slot reserves the workgroup's run of ranks
// in that tile (an atomic on one address is serialised device-wide, and
// neighbouring particles share tiles).  A group that finds the table full takes its ranks
// from the tile counter directly.  Ranks stay unique and dense per tile.
constexpr uint32_t kBinSlots = 64;
template <int R>
__device__ __forceinline__ void p2g_bin_block(const cs_params& u, const float4* __restrict__ part,
                                              const float4* __restrict__ aff, uint32_t* __restrict__ tile_count,
                                              uint2* __restrict__ key, uint32_t first) {
  __shared__ uint32_t s_tile[kBinSlots], s_cnt[kBinSlots], s_base[kBinSlots];
  const uint32_t t = threadIdx.x;
  if (t < kBinSlots) {
    s_tile[t] = kNoTile;
    s_cnt[t] = 0u;
  }
  __syncthreads();
  const uint32_t lane = t & 63u;
  const uint64_t below = lane ? (~0ull >> (64u - lane)) : 0ull;
  uint32_t tile[R], rank[R];
  int slot[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t pid = first + (uint32_t)r * 256u + t;
    bool active = false;
    tile[r] = kNoTile;
    if (pid < u.num_particles) {
      P2GParticle q;
      active = p2g_setup(u, part, aff, pid, q);
      if (active) tile[r] = tile_of(u, q.bx, q.by, q.bz);
    }
    uint64_t todo = __ballot(active);
    rank[r] = 0u;
    slot[r] = -1;
    while (todo) {
      const uint32_t leader = (uint32_t)__ffsll((unsigned long long)todo) - 1u;
      const uint32_t lt = __shfl(tile[r], (int)leader, 64);
      const uint64_t group = __ballot(active && tile[r] == lt);
      uint32_t base = 0u;
      int sl = -1;
      if (lane == leader) {
        uint32_t h = (lt * 2654435761u) >> 26;
        for (uint32_t probe = 0; probe < kBinSlots; ++probe, h = (h + 1u) & (kBinSlots - 1u)) {
          const uint32_t prev = atomicCAS(&s_tile[h], kNoTile, lt);
          if (prev == kNoTile || prev == lt) {
            sl = (int)h;
            break;
          }
        }
        base = sl >= 0 ? atomicAdd(&s_cnt[sl], (uint32_t)__popcll(group))
                       : atomicAdd(&tile_count[lt], (uint32_t)__popcll(group));
      }
      base = __shfl(base, (int)leader, 64);
      sl = __shfl(sl, (int)leader, 64);
      if (active && tile[r] == lt) {
        rank[r] = base + (uint32_t)__popcll(group & below);
        slot[r] = sl;
      }
      todo &= ~group;
    }
  }
  __syncthreads();
  if (t < kBinSlots && s_tile[t] != kNoTile) s_base[t] = atomicAdd(&tile_count[s_tile[t]], s_cnt[t]);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t pid = first + (uint32_t)r * 256u + t;
    if (pid >= u.num_particles) continue;
    const uint32_t rk = (tile[r] != kNoTile && slot[r] >= 0) ? s_base[slot[r]] + rank[r] : rank[r];
    key[pid] = make_uint2(tile[r], rk);
  }
}
constexpr int kBinPerThread = 4;  // particles per thread of k_p2g_bin: 1024 per workgroup
__global__ void __launch_bounds__(256) k_p2g_bin(const cs_dparams* __restrict__ dp, const float4* __restrict__ part,
                                                 const float4* __restrict__ aff, uint32_t* __restrict__ tile_count,
                                                 uint2* __restrict__ key) {
  CS_PARAMS_FROM(dp);
  p2g_bin_block<kBinPerThread>(u, part, aff, tile_count, key, blockIdx.x * (256u * kBinPerThread));
}

// Pass 2: exclusive scan of the tile counts (one workgroup).  With `chunk`,
// also the exclusive scan of the tiles' chunk counts ceil(count / kChunk),
// chunk[ntiles] holding the total: the work items of the chunked tile pass.
__device__ __forceinline__ void p2g_scan_block(const uint32_t* __restrict__ count, uint32_t ntiles,
                                               uint32_t* __restrict__ start, uint32_t* __restrict__ chunk,
                                               uint32_t* s_sum, uint32_t* s_ch) {
  const uint32_t t = threadIdx.x;
  const uint32_t per = (ntiles + 1023u) / 1024u;
  const uint32_t b0 = min(t * per, ntiles), b1 = min(b0 + per, ntiles);
  uint32_t c = 0, cc = 0;
  for (uint32_t b = b0; b < b1; ++b) {
    c += count[b];
    cc += (count[b] + kChunk - 1u) / kChunk;
  }
  s_sum[t] = c;
  s_ch[t] = cc;
  __syncthreads();
  for (uint32_t d = 1; d < 1024u; d <<= 1) {
    const uint32_t v = t >= d ? s_sum[t - d] : 0u;
    const uint32_t w = t >= d ? s_ch[t - d] : 0u;
    __syncthreads();
    s_sum[t] += v;
    s_ch[t] += w;
    __syncthreads();
  }
  uint32_t run = s_sum[t] - c, crun = s_ch[t] - cc;
  for (uint32_t b = b0; b < b1; ++b) {
    start[b] = run;
    run += count[b];
    if (chunk) {
      chunk[b] = crun;
      crun += (count[b] + kChunk - 1u) / kChunk;
    }
  }
  if (chunk && t == 1023u) chunk[ntiles] = s_ch[t];
}
__global__ void __launch_bounds__(1024) k_p2g_scan(const cs_dparams* __restrict__ dp,
                                                   const uint32_t* __restrict__ count, uint32_t ntiles,
                                                   uint32_t* __restrict__ start, uint32_t* __restrict__ chunk) {
  __shared__ uint32_t s_sum[1024], s_ch[1024];
  if (dp->halt) return;  // the bins stay as a halted substep's G2P will read them
  p2g_scan_block(count, ntiles, start, chunk, s_sum, s_ch);
}

// Pass 3: the tile-sorted particle list.
__device__ __forceinline__ void p2g_scatter_particle(uint32_t n, const uint2* __restrict__ key,
                                                     const uint32_t* __restrict__ start, uint32_t* __restrict__ sorted,
                                                     uint32_t pid) {
  if (pid >= n) return;
  const uint2 k = key[pid];
  if (k.x != kNoTile) sorted[start[k.x] + k.y] = pid;
}
__global__ void __launch_bounds__(256) k_p2g_scatter(const cs_dparams* __restrict__ dp, const uint2* __restrict__ key,
                                                     const uint32_t* __restrict__ start, uint32_t* __restrict__ sorted) {
  CS_PARAMS_FROM(dp);
  p2g_scatter_particle(u.num_particles, key, start, sorted, blockIdx.x * blockDim.x + threadIdx.x);
}

// Pass 4: one workgroup per tile.  The tile's particles are dealt to threads
// in a strided order (thread t takes list positions t*K, t*K + 1, ...), so the
// lanes of one wavefront work on particles far apart in the list and rarely
// hit the same LDS word in one instruction.
constexpr uint32_t kTileThreads = 256;  // the lane dealing below assumes 256
// The particles [first, first + n) of `tile`'s sorted list into the tile's
// 10^3-node LDS block, then the block's nonzero entries to the lanes.
__device__ __forceinline__ void p2g_tile_range(const cs_params& u, const float4* __restrict__ part,
                                               const float4* __restrict__ aff, const uint32_t* __restrict__ sorted,
                                               uint32_t first, uint32_t n, int32_t* __restrict__ lanes,
                                               uint32_t* __restrict__ metrics, uint32_t tile, int32_t* s_acc) {
  constexpr uint32_t NT = kTileThreads;
  for (uint32_t i = threadIdx.x; i < 6u * FP3; i += NT) s_acc[i] = 0;
  const uint32_t tx = tile % u.nbx, ty = (tile / u.nbx) % u.nby, tz = tile / (u.nbx * u.nby);
  const int ox = (int)(tx * TB), oy = (int)(ty * TB), oz = (int)(tz * TB);
  __syncthreads();
  const uint32_t per = (n + NT - 1) / NT;
  // thread t takes list positions rev(t) * per + r: the bit-reversed thread
  // index spreads a wavefront's lanes over the whole range, so they rarely
  // share a cell and hit the same LDS word in one instruction
  const uint32_t lead = __brev(threadIdx.x) >> (32 - 8);
  uint32_t overflow = 0;
  for (uint32_t r = 0; r < per; ++r) {
    const uint32_t pos = lead * per + r;
    if (pos >= n) break;
    const uint32_t pid = sorted[first + pos];
    P2GParticle q;
    p2g_setup(u, part, aff, pid, q);  // binned particles are active
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const int cx = q.bx + i, cy = q.by + j, cz = q.bz + k;
          if (!on_grid(u, cx, cy, cz)) continue;
          int32_t v[6];
          overflow += p2g_node(u, q, i, j, k, v);
          const int lx = cx - ox, ly = cy - oy, lz = cz - oz;
          if ((uint32_t)lx < (uint32_t)FP && (uint32_t)ly < (uint32_t)FP && (uint32_t)lz < (uint32_t)FP) {
            const int li = (lz * FP + ly) * FP + lx;
#pragma unroll
            for (int l = 0; l < 6; ++l) atomicAdd(&s_acc[l * FP3 + li], v[l]);
          } else {
            // a node outside the footprint (a particle far outside the domain): straight to HBM
            const uint32_t ci = cell_index(u, (uint32_t)cx, (uint32_t)cy, (uint32_t)cz);
#pragma unroll
            for (int l = 0; l < 6; ++l) atomicAdd(&lanes[(size_t)l * u.nc + ci], v[l]);
          }
        }
  }
  __syncthreads();
  for (uint32_t li = threadIdx.x; li < (uint32_t)FP3; li += NT) {
    const int gx = ox + (int)(li % FP), gy = oy + (int)((li / FP) % FP), gz = oz + (int)(li / (FP * FP));
    if (!on_grid(u, gx, gy, gz)) continue;
    const uint32_t ci = cell_index(u, (uint32_t)gx, (uint32_t)gy, (uint32_t)gz);
#pragma unroll
    for (int l = 0; l < 6; ++l) {
      const int32_t v = s_acc[l * FP3 + li];
      if (v != 0) atomicAdd(&lanes[(size_t)l * u.nc + ci], v);
    }
  }
  if (overflow) atomicAdd(&metrics[CS_M_OVERFLOW], overflow);
}
// Pass 4 (batched scenes): one workgroup per tile.
__device__ __forceinline__ void p2g_tile_block(const cs_params& u, const float4* __restrict__ part,
                                               const float4* __restrict__ aff,
                                               const uint32_t* __restrict__ tile_count,
                                               const uint32_t* __restrict__ start,
                                               const uint32_t* __restrict__ sorted, int32_t* __restrict__ lanes,
                                               uint32_t* __restrict__ metrics, uint32_t tile, int32_t* s_acc) {
  const uint32_t n = tile_count[tile];
  if (n == 0) return;
  p2g_tile_range(u, part, aff, sorted, start[tile], n, lanes, metrics, tile, s_acc);
}
// Pass 4 (one scene): one workgroup per chunk of at most kChunk particles of a
// tile's list, so a tile holding several thousand particles is spread over
// several workgroups (and CUs) instead of serialising on one; each chunk
// flushes its own block (the lanes are added with i32 atomics, so the result
// does not depend on the split).  Workgroup w finds its tile as the last one
// whose chunk prefix is at most w.
__global__ void __launch_bounds__(kTileThreads) k_p2g_tile_chunks(const cs_dparams* __restrict__ dp,
                                                                  const float4* __restrict__ part,
                                                                  const float4* __restrict__ aff,
                                                                  const uint32_t* __restrict__ tile_count,
                                                                  const uint32_t* __restrict__ start,
                                                                  const uint32_t* __restrict__ chunk,
                                                                  const uint32_t* __restrict__ sorted,
                                                                  int32_t* __restrict__ lanes,
                                                                  uint32_t* __restrict__ metrics) {
  CS_PARAMS_FROM(dp);
  __shared__ int32_t s_acc[6 * FP3];
  const uint32_t w = blockIdx.x;
  if (w >= chunk[u.nbricks]) return;
  uint32_t lo = 0, hi = u.nbricks;  // chunk[lo] <= w < chunk[hi]
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    if (chunk[mid] <= w) lo = mid;
    else hi = mid;
  }
  const uint32_t c = w - chunk[lo];
  const uint32_t n = tile_count[lo];
  const uint32_t off = c * kChunk;
  p2g_tile_range(u, part, aff, sorted, start[lo] + off, min(kChunk, n - off), lanes, metrics, lo, s_acc);
}

// ---- batched passes (one launch over every scene of a batch) -------------------------------
#define CS_BATCH_SCENE(space)                                                                         \
  const uint32_t* st_ = starts + (space) * (n + 1u);                                                  \
  const uint32_t scene_ = cs_batch_scene_of(st_, n, blockIdx.x);                                      \
  const cs_scene_dev& d = sc[scene_];                                                                 \
  const uint32_t blk = blockIdx.x - st_[scene_];                                                      \
  const cs_params u = d.u

// Binning for tiled scenes; the whole P2G for the others (k_p2g_atomic).
__global__ void __launch_bounds__(256) k_batch_p2g_bin(const cs_scene_dev* __restrict__ sc,
                                                       const uint32_t* __restrict__ starts, uint32_t n) {
  CS_BATCH_SCENE(CS_SPACE_PARTICLES);
  const uint32_t pid = blk * 256u + threadIdx.x;
  if (d.p2g_tiled)
    p2g_bin_block<1>(u, reinterpret_cast<const float4*>(d.part), reinterpret_cast<const float4*>(d.aff),
                     d.tile_count, d.p2g_key, blk * 256u);
  else
    p2g_atomic_particle(u, reinterpret_cast<const float4*>(d.part), reinterpret_cast<const float4*>(d.aff), d.lanes,
                        d.metrics, pid);
}

__global__ void __launch_bounds__(1024) k_batch_p2g_scan(const cs_scene_dev* __restrict__ sc,
                                                         const uint32_t* __restrict__ starts, uint32_t n) {
  __shared__ uint32_t s_sum[1024], s_ch[1024];
  CS_BATCH_SCENE(CS_SPACE_SCENES);
  (void)blk;
  if (!d.p2g_tiled) return;
  p2g_scan_block(d.tile_count, u.nbricks, d.tile_start, nullptr, s_sum, s_ch);
}

__global__ void __launch_bounds__(256) k_batch_p2g_scatter(const cs_scene_dev* __restrict__ sc,
                                                           const uint32_t* __restrict__ starts, uint32_t n) {
  CS_BATCH_SCENE(CS_SPACE_PARTICLES);
  if (!d.p2g_tiled) return;
  p2g_scatter_particle(u.num_particles, d.p2g_key, d.tile_start, d.p2g_sorted, blk * 256u + threadIdx.x);
}

__global__ void __launch_bounds__(kTileThreads) k_batch_p2g_tile(const cs_scene_dev* __restrict__ sc,
                                                                 const uint32_t* __restrict__ starts, uint32_t n) {
  __shared__ int32_t s_acc[6 * FP3];
  CS_BATCH_SCENE(CS_SPACE_BRICKS);
  if (!d.p2g_tiled) return;
  p2g_tile_block(u, reinterpret_cast<const float4*>(d.part), reinterpret_cast<const float4*>(d.aff), d.tile_count,
                 d.tile_start, d.p2g_sorted, d.lanes, d.metrics, blk, s_acc);
}

hipError_t cs_batch_launch_p2g(const cs_batch_launch& b) {
  const uint32_t pb = cs_batch_blocks(b, CS_SPACE_PARTICLES);
  if (pb == 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_p2g_bin, dim3(pb), dim3(256), 0, b.s, b.sc, b.starts, b.n);
  hipLaunchKernelGGL(k_batch_p2g_scan, dim3(b.n), dim3(1024), 0, b.s, b.sc, b.starts, b.n);
  hipLaunchKernelGGL(k_batch_p2g_scatter, dim3(pb), dim3(256), 0, b.s, b.sc, b.starts, b.n);
  hipLaunchKernelGGL(k_batch_p2g_tile, dim3(cs_batch_blocks(b, CS_SPACE_BRICKS)), dim3(kTileThreads), 0, b.s, b.sc,
                     b.starts, b.n);
  return hipGetLastError();
}

hipError_t cs_launch_p2g_tiled(hipStream_t s, const cs_pblock& P, const float* part, const float* aff,
                               int32_t* lanes, uint32_t* metrics, uint32_t* tile_count, uint32_t* tile_start,
                               uint32_t* tile_chunk, uint2* key, uint32_t* sorted) {
  const cs_params& u = P.u;
  if (u.num_particles == 0) return hipSuccess;
  const dim3 pb((u.num_particles + 255u) / 256u);
  hipError_t e = cs_launch_fill_i32_gated(s, P.d, reinterpret_cast<int32_t*>(tile_count), u.nbricks, 0);
  if (e != hipSuccess) return e;
  const uint32_t per_block = 256u * kBinPerThread;
  hipLaunchKernelGGL(k_p2g_bin, dim3((u.num_particles + per_block - 1u) / per_block), dim3(256), 0, s, P.d,
                     reinterpret_cast<const float4*>(part), reinterpret_cast<const float4*>(aff), tile_count, key);
  hipLaunchKernelGGL(k_p2g_scan, dim3(1), dim3(1024), 0, s, P.d, tile_count, u.nbricks, tile_start, tile_chunk);
  hipLaunchKernelGGL(k_p2g_scatter, pb, dim3(256), 0, s, P.d, key, tile_start, sorted);
  // at most one partial chunk per tile beyond the full ones
  const uint32_t items = u.nbricks + (u.num_particles + kChunk - 1u) / kChunk;
  hipLaunchKernelGGL(k_p2g_tile_chunks, dim3(items), dim3(kTileThreads), 0, s, P.d,
                     reinterpret_cast<const float4*>(part),
                     reinterpret_cast<const float4*>(aff), tile_count, tile_start, tile_chunk, sorted, lanes, metrics);
  return hipGetLastError();
}
