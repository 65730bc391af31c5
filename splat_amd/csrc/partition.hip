// partition.hip -- stable live/terminated compaction (replaces partition!, src/partition.jl:10-132).
//
// The reference claims output ranges with two atomics per 256-ray block, so block order -- and
// the permutation -- depends on scheduling. Here the permutation is a fixed function of the
// input: per-tile ACTIVE counts (written by the evolve kernel, or by k_count_active), one
// exclusive scan over tiles, then a scatter whose in-tile ranks come from wave ballots
// (v_mbcnt) and a 4-entry LDS prefix. The scatter stages the tile's 136-byte records through LDS
// so that both the read of the tile and the writes of its two output runs are coalesced.
#include "scatter.hpp"

namespace spl {


__device__ __forceinline__ int64_t live_count_p(int64_t R, const int64_t* __restrict__ d_count) {
  if (!d_count) return R;
  int64_t c = *d_count;
  return c < R ? c : R;
}

__global__ __launch_bounds__(TILE) void k_count_active(const splat_adray_t* __restrict__ rays, int64_t R,
                                                       const int64_t* __restrict__ d_count,
                                                       int32_t* __restrict__ tile_counts) {
  const int64_t Rl = live_count_p(R, d_count);
  const int64_t r = (int64_t)blockIdx.x * TILE + threadIdx.x;
  const bool a = r < Rl && rays[r].status == SPLAT_RAY_ACTIVE;
  int c = __syncthreads_count(a);
  if (threadIdx.x == 0) tile_counts[blockIdx.x] = c;
}

// single workgroup, 1024 threads: each thread owns a contiguous chunk of tiles
// exclusive scan of the tile counts of the live tiles (the first ceil(min(*d_count, R) / TILE)
// of n_tiles; all of them without d_count)
__global__ __launch_bounds__(1024) void k_scan_tiles(const int32_t* __restrict__ tile_counts, int64_t n_tiles,
                                                     const int64_t* __restrict__ d_count, int64_t R,
                                                     int32_t* __restrict__ tile_offsets, int64_t* __restrict__ d_total) {
  __shared__ int64_t wave_sums[1024 / 64];
  if (d_count) {
    const int64_t A = *d_count < R ? *d_count : R;
    const int64_t live_tiles = (A + TILE - 1) / TILE;
    n_tiles = live_tiles < n_tiles ? live_tiles : n_tiles;
  }
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int64_t chunk = (n_tiles + 1023) / 1024;
  const int64_t b = t * chunk, e = b + chunk < n_tiles ? b + chunk : n_tiles;
  int64_t s = 0;
  for (int64_t i = b; i < e; ++i) s += tile_counts[i];
  // inclusive scan of the chunk sums: within each wave by shuffles, then over the 16 wave sums
  int64_t incl = s;
  for (int off = 1; off < 64; off <<= 1) {
    const int64_t u = __shfl_up(incl, off);
    if (lane >= off) incl += u;
  }
  if (lane == 63) wave_sums[wave] = incl;
  __syncthreads();
  int64_t before = 0, total = 0;
  for (int w = 0; w < 1024 / 64; ++w) {
    const int64_t ws = wave_sums[w];
    before += w < wave ? ws : 0;
    total += ws;
  }
  int64_t run = before + incl - s;  // exclusive prefix of this chunk
  for (int64_t i = b; i < e; ++i) {
    tile_offsets[i] = (int32_t)run;
    run += tile_counts[i];
  }
  if (t == 0) *d_total = total;
}

__global__ __launch_bounds__(TILE) void k_scatter(const splat_adray_t* __restrict__ in, int64_t R,
                                                  const int64_t* __restrict__ d_count,
                                                  const int32_t* __restrict__ tile_offsets,
                                                  const int64_t* __restrict__ d_total,
                                                  splat_adray_t* __restrict__ out_active,
                                                  splat_adray_t* __restrict__ out_rest,
                                                  const int32_t* __restrict__ carry_in,
                                                  const uint64_t* __restrict__ carry_keys,
                                                  int32_t* __restrict__ carry_rest) {
  const int64_t Rl = live_count_p(R, d_count);
  const int64_t tile_base = (int64_t)blockIdx.x * TILE;
  if (tile_base >= Rl) return;  // uniform over the workgroup
  const int nvalid = (int)(Rl - tile_base < TILE ? Rl - tile_base : TILE);
  const int64_t off_a = tile_offsets[blockIdx.x];
  const int64_t off_r = *d_total + (tile_base - off_a);  // non-ACTIVE rays before this tile
  scatter_tile(in + tile_base, nvalid, off_a, off_r, out_active, out_rest, carry_in ? carry_in + tile_base : nullptr,
               carry_rest, carry_keys ? carry_keys + tile_base : nullptr);
}

void launch_count_active(const splat_adray_t* rays, int64_t R, const int64_t* d_count, int32_t* tile_counts,
                         hipStream_t s) {
  if (R <= 0) return;
  hipLaunchKernelGGL(k_count_active, dim3((unsigned)((R + TILE - 1) / TILE)), dim3(TILE), 0, s, rays, R, d_count,
                     tile_counts);
}
void launch_scan_tiles(const int32_t* tile_counts, int64_t n_tiles, const int64_t* d_count, int64_t R,
                       int32_t* tile_offsets, int64_t* d_total, hipStream_t s) {
  hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, s, tile_counts, n_tiles, d_count, R, tile_offsets, d_total);
}
void launch_scatter(const splat_adray_t* in, int64_t R, const int64_t* d_count, const int32_t* tile_offsets,
                    const int64_t* d_total, splat_adray_t* out_active, splat_adray_t* out_rest, const int32_t* carry_in,
                    const uint64_t* carry_keys, int32_t* carry_rest, hipStream_t s) {
  if (R <= 0) return;
  hipLaunchKernelGGL(k_scatter, dim3((unsigned)((R + TILE - 1) / TILE)), dim3(TILE), 0, s, in, R, d_count, tile_offsets,
                     d_total, out_active, out_rest, carry_in, carry_keys, carry_rest);
}

}  // namespace spl
