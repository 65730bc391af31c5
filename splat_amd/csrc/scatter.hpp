// scatter.hpp -- the stable LDS-staged tile scatter shared by partition.hip and engine.hip.
#pragma once

#include "internal.hpp"

namespace spl {

constexpr int WORDS = sizeof(splat_adray_t) / 8;  // 17 eight-byte words per record

// Stable scatter of one tile of `nvalid` rays starting at `in`: ACTIVE rays to out_active[off_a..],
// the others to out_rest[off_r..], both in input order; the carry array (hit indices, or the index
// parts of packed keys) moves with the others.
// Called by a whole TILE-thread workgroup.
__device__ __forceinline__ void scatter_tile(const splat_adray_t* __restrict__ in, int nvalid, int64_t off_a,
                                             int64_t off_r, splat_adray_t* __restrict__ out_active,
                                             splat_adray_t* __restrict__ out_rest, const int32_t* __restrict__ carry_in,
                                             int32_t* __restrict__ carry_rest,
                                             const uint64_t* __restrict__ carry_keys = nullptr) {
  __shared__ uint2 stage[TILE * WORDS];  // 34,816 bytes
  __shared__ int16_t order[TILE];
  __shared__ int32_t wave_a[TILE / 64], wave_r[TILE / 64];
  // coalesced read of the tile
  const uint2* src = reinterpret_cast<const uint2*>(in);
  for (int e = threadIdx.x; e < nvalid * WORDS; e += TILE) stage[e] = src[e];
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
  uint2* dst_a = reinterpret_cast<uint2*>(out_active + off_a);
  uint2* dst_r = reinterpret_cast<uint2*>(out_rest + off_r);
  for (int e = threadIdx.x; e < nvalid * WORDS; e += TILE) {
    const int slot = e / WORDS, w = e - slot * WORDS;
    const uint2 v = stage[order[slot] * WORDS + w];
    if (slot < na) dst_a[e] = v;
    else dst_r[e - na * WORDS] = v;
  }
  __syncthreads();  // the LDS stage is reused by the caller's next tile
}

}  // namespace spl
