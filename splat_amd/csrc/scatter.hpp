// scatter.hpp -- the stable LDS-staged tile scatter shared by partition.hip and engine.hip.
#pragma once

#include "internal.hpp"

namespace spl {

constexpr int WORDS = sizeof(splat_adray_t) / 8;  // 17 eight-byte words per record

// Coalesced copy of n 8-byte words src -> stage (LDS, 16-byte aligned) by the whole TILE-thread
// workgroup: 16-byte loads when src is 16-byte aligned (a tile of 256 records always is).
__device__ __forceinline__ void stage_words(uint2* __restrict__ stage, const uint2* __restrict__ src, int n) {
  if ((reinterpret_cast<uintptr_t>(src) & 15) == 0) {
    const int np = n / 2;
    for (int k = threadIdx.x; k < np; k += TILE) reinterpret_cast<uint4*>(stage)[k] = reinterpret_cast<const uint4*>(src)[k];
    if ((n & 1) && threadIdx.x == 0) stage[n - 1] = src[n - 1];
  } else {
    for (int e = threadIdx.x; e < n; e += TILE) stage[e] = src[e];
  }
}

// Words [w0, w1) of the permuted tile (word e: word e % WORDS of record order[e / WORDS] in
// stage) written to dst in order, by the whole workgroup: 16-byte stores from dst's first
// 16-byte boundary on (a leading and a trailing 8-byte word alone).
__device__ __forceinline__ void write_words(const uint2* __restrict__ stage, const int16_t* __restrict__ order, int w0,
                                            int w1, uint2* __restrict__ dst) {
  const int n = w1 - w0;
  if (n <= 0) return;
  auto word = [&](int r) {
    const int e = w0 + r, slot = e / WORDS;
    return stage[order[slot] * WORDS + (e - slot * WORDS)];
  };
  const int lead = (reinterpret_cast<uintptr_t>(dst) & 15) ? 1 : 0;
  if (lead && threadIdx.x == 0) dst[0] = word(0);
  const int np = (n - lead) / 2;
  uint4* d4 = reinterpret_cast<uint4*>(dst + lead);
  for (int k = threadIdx.x; k < np; k += TILE) {
    const uint2 a = word(lead + 2 * k), b = word(lead + 2 * k + 1);
    d4[k] = make_uint4(a.x, a.y, b.x, b.y);
  }
  const int tail = lead + 2 * np;
  if (tail < n && threadIdx.x == TILE - 1) dst[tail] = word(tail);
}

// Stable scatter of one tile of `nvalid` rays starting at `in`: ACTIVE rays to out_active[off_a..],
// the others to out_rest[off_r..], both in input order; the carry array (hit indices, or the index
// parts of packed keys) moves with the others.
// Called by a whole TILE-thread workgroup.
__device__ __forceinline__ void scatter_tile(const splat_adray_t* __restrict__ in, int nvalid, int64_t off_a,
                                             int64_t off_r, splat_adray_t* __restrict__ out_active,
                                             splat_adray_t* __restrict__ out_rest, const int32_t* __restrict__ carry_in,
                                             int32_t* __restrict__ carry_rest,
                                             const uint64_t* __restrict__ carry_keys = nullptr) {
  __shared__ __align__(16) uint2 stage[TILE * WORDS];  // 34,816 bytes
  __shared__ int16_t order[TILE];
  __shared__ int32_t wave_a[TILE / 64], wave_r[TILE / 64];
  // coalesced read of the tile
  stage_words(stage, reinterpret_cast<const uint2*>(in), nvalid * WORDS);
  __syncthreads();

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const bool valid = tid < nvalid;
  const uint8_t status = valid ? reinterpret_cast<const splat_adray_t*>(stage)[tid].status : 0;
  const bool act = valid && status == SPLAT_RAY_ACTIVE;
  const bool rest = valid && !act;
  const uint64_t ma = __builtin_amdgcn_ballot_w64(act), mr = __builtin_amdgcn_ballot_w64(rest);
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const int ra = __builtin_popcountll(ma & below), rr = __builtin_popcountll(mr & below);
  if (lane == 0) {
    wave_a[wave] = __builtin_popcountll(ma);
    wave_r[wave] = __builtin_popcountll(mr);
  }
  __syncthreads();
  int pa = 0, pr = 0, na = 0;
  for (int w = 0; w < TILE / 64; ++w) {
    if (w < wave) { pa += wave_a[w]; pr += wave_r[w]; }
    na += wave_a[w];
  }
  if (act) order[pa + ra] = (int16_t)tid;
  if (rest) order[na + pr + rr] = (int16_t)tid;
  if (rest && carry_rest) carry_rest[off_r + pr + rr] = carry_keys ? key_index(carry_keys[tid]) : carry_in[tid];
  __syncthreads();

  // coalesced writes of the two output runs
  write_words(stage, order, 0, na * WORDS, reinterpret_cast<uint2*>(out_active + off_a));
  write_words(stage, order, na * WORDS, nvalid * WORDS, reinterpret_cast<uint2*>(out_rest + off_r));
  __syncthreads();  // the LDS stage is reused by the caller's next tile
}

}  // namespace spl
