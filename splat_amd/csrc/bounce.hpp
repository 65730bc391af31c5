// bounce.hpp -- one bounce of one tile of live rays in a single pass over the ray data, shared by
// the single-frame loop (trace.hip) and the multi-frame engine (engine.hip):
//
//   count pass (bounce_count):  per tile, how many rays stay ACTIVE -- from the hit keys alone
//                               (status_after), so it reads 8 bytes per ray, not the ray;
//   scan:                       tile offsets and the live count A_k (partition.hip);
//   bounce pass (bounce_tile):  coalesced read of the tile into LDS -> evolve every ray in place
//                               (dual-number refract/reflect) -> rays that became DIFFUSE get their
//                               final evolution and splat into the light map at once -> stable
//                               partition by ballot ranks -> coalesced writes of the two runs:
//                               ACTIVE rays to the next live buffer, the others straight to their
//                               final slots with their hit index.
//
// This replaces evolve (read + write every ray), scatter (read + write every ray) and the light
// map's read of every final ray with one read and one write; at the first bounce the tile is
// generated from the pinhole in LDS, so the rays are never written before their first bounce.
// Positions come from status_after in both passes (same function of the same inputs), so they
// always agree with the counts.
#pragma once

#include "hitcore.hpp"
#include "path.hpp"
#include "scatter.hpp"

namespace spl {

// What a bounce pass needs besides the rays and the light map.
struct BounceParams {
  const splat_ftri_t* tris;
  int32_t n_tris, first_diffuse, bounce;
  float force_rand;
  uint64_t seed;
  splat_optics_t optics;
  float* tex;  // non-null: splat the rays that become DIFFUSE into this light map
};

__device__ __forceinline__ int32_t clamp_index(int32_t i, int32_t N) { return (i < 1 || i > N) ? 1 : i; }

// ACTIVE rays after this bounce among the tile's first nvalid rays, counted by one wave (every
// lane returns the count): lane l takes rays l, l + 64, l + 128, l + 192, whose key loads are
// independent, so a wave moves a whole tile's keys per memory round trip. all_active: every ray of
// the tile is known to be ACTIVE (the loop's live rays), so no ray is read unless the sentinel
// test needs it. CAM: the rays are the pinhole's rays base, base + 1, ... (not yet in memory).
template <bool CAM>
__device__ __forceinline__ int bounce_count(const splat_adray_t* __restrict__ in, const uint64_t* __restrict__ keys,
                                            int nvalid, const BounceParams& P, bool all_active,
                                            const splat_pinhole_t& cam, int64_t base) {
  static_assert(TILE == 256, "four rays per lane");
  const int lane = threadIdx.x & 63;
  uint64_t k0 = KEY_MISS, k1 = KEY_MISS, k2 = KEY_MISS, k3 = KEY_MISS;
  if (lane < nvalid) k0 = keys[lane];
  if (lane + 64 < nvalid) k1 = keys[lane + 64];
  if (lane + 128 < nvalid) k2 = keys[lane + 128];
  if (lane + 192 < nvalid) k3 = keys[lane + 192];
  const bool live = P.first_diffuse > 1 && sentinel_live(P.tris);  // uniform
  auto active = [&](int t, uint64_t key) {
    if (t >= nvalid || !(CAM || all_active || in[t].status == SPLAT_RAY_ACTIVE)) return false;
    const int32_t i = clamp_index(key_index(key), P.n_tris);
    bool tri1 = false;
    if (i == 1 && live) {
      if constexpr (CAM) {
        const splat_adray_t g = raygen_ray(cam, base + t);
        tri1 = sentinel_hit(P.tris, ld3(g.pos), ld3(g.dir), g.ignore_tri);
      } else {
        tri1 = sentinel_hit(P.tris, ld3(in[t].pos), ld3(in[t].dir), in[t].ignore_tri);
      }
    }
    return status_after(i, P.first_diffuse, tri1) == SPLAT_RAY_ACTIVE;
  };
  return __builtin_popcountll(__builtin_amdgcn_ballot_w64(active(lane, k0))) +
         __builtin_popcountll(__builtin_amdgcn_ballot_w64(active(lane + 64, k1))) +
         __builtin_popcountll(__builtin_amdgcn_ballot_w64(active(lane + 128, k2))) +
         __builtin_popcountll(__builtin_amdgcn_ballot_w64(active(lane + 192, k3)));
}

// The bounce pass's LDS: one instance per workgroup, whichever bounce_tile variant runs. After the
// tile's records are written out, `stage` holds the privatised light-map bins.
constexpr int MULTI_BINS = 832;  // bins of the multi-sample splat (its own area: records still staged)
struct BounceLDS {
  __align__(16) uint2 stage[TILE * WORDS];  // 34,816 bytes: the tile's records
  // 3,328 bytes (4 engine workgroups per CU still fit): the multi-sample splat's bins, or, for
  // single-sample splats weighted by curves, the curves' segment table (CurveTab)
  float mbins[MULTI_BINS];
  int16_t order[TILE];
  int32_t wave_a[TILE / 64], wave_r[TILE / 64];
  int32_t box[TILE / 64][4];  // per wave: min x, max x, min y, max y of its splat texels
};
constexpr int SPLAT_BINS = TILE * WORDS * 2;  // floats of `stage`: 8,704
static_assert(sizeof(CurveTab) <= MULTI_BINS * sizeof(float), "the curve table shares the multi-sample bins");

// max/min over the wave (all 64 lanes take part)
__device__ __forceinline__ int32_t wave_max_i32(int32_t v) {
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
  return v;
}
__device__ __forceinline__ int32_t wave_min_i32(int32_t v) {
  for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off));
  return v;
}
__device__ __forceinline__ BounceLDS& bounce_lds() {
  __shared__ BounceLDS lds;
  return lds;
}

// The tile's pending single-sample splats (has, texel (sx, sy), values sv per plane). When they
// crowd a small texel window (at least two splats per texel of it on average: caustics), they are
// summed in LDS bins over that window (the `stage` area, free once the records are written out)
// and flushed with one global atomic per touched bin; otherwise -- spread-out splats, where a bin
// would absorb nothing -- they go straight to the map. Called by the whole workgroup after the
// records' copy-out.
__device__ __forceinline__ void splat_tile(BounceLDS& lds, bool has, int32_t sx, int32_t sy, const float (&sv)[3],
                                           const BounceParams& P, const splat_lightmap_params_t& lm) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int np = splat_planes(lm) < 3 ? splat_planes(lm) : 3;
  const int64_t plane = (int64_t)lm.width * lm.height;
  const int32_t x0 = wave_min_i32(has ? sx : INT32_MAX), x1 = wave_max_i32(has ? sx : INT32_MIN);
  const int32_t y0 = wave_min_i32(has ? sy : INT32_MAX), y1 = wave_max_i32(has ? sy : INT32_MIN);
  if (lane == 0) {
    lds.box[wave][0] = x0; lds.box[wave][1] = x1; lds.box[wave][2] = y0; lds.box[wave][3] = y1;
  }
  const int n_splat = __syncthreads_count(has);  // also: every record has left `stage`
  int32_t bx0 = INT32_MAX, bx1 = INT32_MIN, by0 = INT32_MAX, by1 = INT32_MIN;
  for (int w = 0; w < TILE / 64; ++w) {
    bx0 = min(bx0, lds.box[w][0]); bx1 = max(bx1, lds.box[w][1]);
    by0 = min(by0, lds.box[w][2]); by1 = max(by1, lds.box[w][3]);
  }
  if (bx0 > bx1) return;  // no splat in the tile (uniform)
  const int64_t ww = (int64_t)bx1 - bx0 + 1, wh = (int64_t)by1 - by0 + 1;
  if (ww * wh * np > SPLAT_BINS || 2 * ww * wh > n_splat) {  // uniform: straight to the map
    if (has)
      for (int c = 0; c < np; ++c)
        unsafeAtomicAdd(P.tex + sx + (int64_t)lm.width * sy + plane * splat_plane(lm, 0, c), sv[c]);
    return;
  }
  float* bins = reinterpret_cast<float*>(lds.stage);
  const int nb = (int)(ww * wh * np);
  for (int e = threadIdx.x; e < nb; e += TILE) bins[e] = 0.0f;
  __syncthreads();
  if (has)
    for (int c = 0; c < np; ++c)
      atomicAdd(bins + (c * wh + (sy - by0)) * ww + (sx - bx0), sv[c]);
  __syncthreads();
  for (int e = threadIdx.x; e < nb; e += TILE) {
    const float v = bins[e];
    if (v != 0.0f) {
      const int c = e / (int)(ww * wh), rem = e - c * (int)(ww * wh);
      const int32_t y = by0 + rem / (int)ww, x = bx0 + rem % (int)ww;
      unsafeAtomicAdd(P.tex + x + (int64_t)lm.width * y + plane * splat_plane(lm, 0, c), v);
    }
  }
}

// The union over the workgroup of per-wave texel windows [x0, x1] x [y0, y1] (lanes without
// splats pass an empty window) and the workgroup's count of `n`; false when the union is empty.
__device__ __forceinline__ bool tile_window(BounceLDS& lds, int32_t x0, int32_t x1, int32_t y0, int32_t y1, int n,
                                            int32_t& bx0, int32_t& bx1, int32_t& by0, int32_t& by1, int& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  x0 = wave_min_i32(x0); x1 = wave_max_i32(x1); y0 = wave_min_i32(y0); y1 = wave_max_i32(y1);
  int32_t cnt = n;
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
  if (lane == 0) {
    lds.box[wave][0] = x0; lds.box[wave][1] = x1; lds.box[wave][2] = y0; lds.box[wave][3] = y1;
    lds.wave_r[wave] = cnt;
  }
  __syncthreads();
  bx0 = INT32_MAX; bx1 = INT32_MIN; by0 = INT32_MAX; by1 = INT32_MIN; total = 0;
  for (int w = 0; w < TILE / 64; ++w) {
    bx0 = min(bx0, lds.box[w][0]); bx1 = max(bx1, lds.box[w][1]);
    by0 = min(by0, lds.box[w][2]); by1 = max(by1, lds.box[w][3]);
    total += lds.wave_r[w];
  }
  __syncthreads();  // box and wave_r are reused after this
  return bx0 <= bx1;
}

// Every splat sample of the tile's DIFFUSE rays with several samples each (spectral planes,
// sub-pixel offsets): `diff` marks this lane's final-evolved ray r on triangle F. The texture
// coordinates of a ray's samples are affine in its expansion offsets (lambda - lambda0, dx, dy),
// so the samples at the corners of its offset box (at most 8; 2 for a wavelength grid alone)
// bound the texels of all of them. When the tile's samples crowd the union of those windows (at
// least two plane writes per bin), they are summed in the LDS bins `mbins` and flushed with one
// global atomic per touched bin (a sample outside the window, after rounding, goes straight to
// the map); otherwise every sample goes straight to the map. Called by the whole workgroup.
__device__ __forceinline__ void splat_tile_multi(BounceLDS& lds, bool diff, const RayF& r, const FTriF& F,
                                                 const BounceParams& P, const splat_lightmap_params_t& lm) {
  const int nl = lm.n_lambda > 0 ? lm.n_lambda : 1, nd = lm.n_delta;
  const int np = splat_planes(lm);
  const int64_t plane = (int64_t)lm.width * lm.height;
  int32_t x0 = INT32_MAX, x1 = INT32_MIN, y0 = INT32_MAX, y1 = INT32_MIN;
  if (diff) {
    for (int kc = 0; kc < (nl > 1 ? 2 : 1); ++kc)
      for (int yc = 0; yc < (nd > 1 ? 2 : 1); ++yc)
        for (int xc = 0; xc < (nd > 1 ? 2 : 1); ++xc) {
          const int k = kc ? nl - 1 : 0;
          const float lam = lm.n_lambda > 0 ? lm.spectrum[k] : r.lambda;
          int32_t tx, ty;
          float I;
          if (splat_sample(r, F, lm, lam, xc ? nd - 1 : 0, yc ? nd - 1 : 0, &tx, &ty, &I)) {
            x0 = min(x0, tx); x1 = max(x1, tx); y0 = min(y0, ty); y1 = max(y1, ty);
          }
        }
  }
  int32_t bx0, bx1, by0, by1;
  int n_diff;
  if (!tile_window(lds, x0, x1, y0, y1, diff ? 1 : 0, bx0, bx1, by0, by1, n_diff)) {
    // no corner sample landed; any sample that does goes straight to the map (uniform branch)
    bx0 = 1;
    bx1 = 0;
  }
  const int64_t ww = (int64_t)bx1 - bx0 + 1, wh = (int64_t)by1 - by0 + 1;
  const int64_t writes = (int64_t)n_diff * nl * nd * nd * np;
  const int64_t nb = bx0 <= bx1 ? ww * wh * lm.channels : 0;
  const bool binned = nb > 0 && nb <= MULTI_BINS && writes >= 2 * nb;  // uniform
  if (binned) {
    for (int e = threadIdx.x; e < (int)nb; e += TILE) lds.mbins[e] = 0.0f;
    __syncthreads();
  }
  if (diff) {
    for (int k = 0; k < nl; ++k) {
      const float lam = lm.n_lambda > 0 ? lm.spectrum[k] : r.lambda;
      for (int iy = 0; iy < nd; ++iy)
        for (int ix = 0; ix < nd; ++ix) {
          int32_t tx, ty;
          float I;
          if (!splat_sample(r, F, lm, lam, ix, iy, &tx, &ty, &I)) continue;
          const bool in = binned && tx >= bx0 && tx <= bx1 && ty >= by0 && ty <= by1;
          for (int c = 0; c < np; ++c) {
            const int32_t pl = splat_plane(lm, k, c);
            const float v = splat_value(lm, k, c, lam, I);
            if (in) atomicAdd(lds.mbins + (pl * wh + (ty - by0)) * ww + (tx - bx0), v);
            else unsafeAtomicAdd(P.tex + tx + (int64_t)lm.width * ty + plane * pl, v);
          }
        }
    }
  }
  if (binned) {
    __syncthreads();
    for (int e = threadIdx.x; e < (int)nb; e += TILE) {
      const float v = lds.mbins[e];
      if (v != 0.0f) {
        const int pl = e / (int)(ww * wh), rem = e - pl * (int)(ww * wh);
        const int32_t y = by0 + rem / (int)ww, x = bx0 + rem % (int)ww;
        unsafeAtomicAdd(P.tex + x + (int64_t)lm.width * y + plane * pl, v);
      }
    }
  }
}

// The bounce pass of one tile (see the top of this file). off_a / off_r: where the tile's ACTIVE
// and other rays start in out_active / out_rest; hit_rest gets the hit index of the others (either
// may be null: a frame whose retired rays nobody reads -- only its light map -- skips them);
// keys_next (nullable) gets the miss key (Inf, 1) at each ACTIVE ray's new position, ready for
// the next bounce's hitter. CAM: the tile is generated from the pinhole (rays base, base + 1,
// ...) instead of read from `in`.
template <bool CAM>
__device__ __forceinline__ void bounce_tile(const splat_adray_t* __restrict__ in, const uint64_t* __restrict__ keys,
                                            uint64_t* __restrict__ keys_next, int nvalid, const BounceParams& P,
                                            const splat_lightmap_params_t& lm, const splat_pinhole_t& cam,
                                            int64_t base, int64_t off_a, int64_t off_r,
                                            splat_adray_t* __restrict__ out_active, splat_adray_t* __restrict__ out_rest,
                                            int32_t* __restrict__ hit_rest) {
  BounceLDS& lds = bounce_lds();
  uint2* stage = lds.stage;
  int16_t* order = lds.order;
  int32_t* wave_a = lds.wave_a;
  int32_t* wave_r = lds.wave_r;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const bool valid = tid < nvalid;
  splat_adray_t* rec = reinterpret_cast<splat_adray_t*>(stage) + tid;
  // One splat sample per ray (one wavelength, no sub-pixel offsets): the tile's splats are summed
  // in LDS bins over the texel window they cover, then flushed with one global atomic per
  // touched bin; otherwise every sample goes straight to the map.
  const bool one_sample = P.tex && lm.n_delta == 1 && lm.n_lambda <= 1;
  // single-sample curve weights read the curves' segment table (in the multi-sample bins' area)
  const bool curve_tab = one_sample && lm.weight_mode == SPLAT_WEIGHTS_CURVES;
  CurveTab& ctab = *reinterpret_cast<CurveTab*>(lds.mbins);
  if (curve_tab) curve_tab_fill(ctab, lm, tid, TILE);
  if constexpr (CAM) {
    if (valid) *rec = raygen_ray(cam, base + tid);
  } else {
    stage_words(stage, reinterpret_cast<const uint2*>(in), nvalid * WORDS);
  }
  __syncthreads();

  bool has = false;   // this ray has a single-sample splat pending
  bool multi = false;  // this ray is final-evolved in `rec` with several samples pending
  int32_t sx = 0, sy = 0;
  float sv[3] = {0.f, 0.f, 0.f};
  int32_t i = 1;
  uint8_t placed = 0;  // the status that decides where the ray goes
  if (valid) {
    RayF ray = load_rayf(rec);
    if (ray.status == SPLAT_RAY_ACTIVE) {
      i = clamp_index(key_index(keys[tid]), P.n_tris);
      placed = predict_status(i, P.first_diffuse, P.tris, ray.pos, ray.dir, ray.ignore_tri);
      const FTriF F = load_ftri(P.tris + (i - 1));
      const float u = bounce_random(P.force_rand, nullptr, 0, P.seed, P.bounce, ray.dest);
      // a ray that becomes DIFFUSE is final-evolved with evolve's own hit test (same ray, same
      // triangle)
      bool dhit = false;
      D3<float> dtD;
      RayF o = evolve_one(ray, i, F, u, P.first_diffuse, P.optics, &dhit, &dtD);
      if (P.tex && o.status == SPLAT_RAY_DIFFUSE) {
        if (one_sample) {
          final_evolve_hit(o, i, dhit, dtD);
          const float lam = lm.n_lambda > 0 ? lm.spectrum[0] : o.lambda;
          float I;
          has = splat_sample(o, F, lm, lam, 0, 0, &sx, &sy, &I);
          if (has)
#pragma unroll
            for (int c = 0; c < 3; ++c)
              if (c < splat_planes(lm)) sv[c] = curve_tab ? I * curve_eval_tab(ctab, lm, c, lam) : splat_value(lm, 0, c, lam, I);
        } else {
          final_evolve_hit(o, i, dhit, dtD);
          multi = true;
        }
      }
      store_rayf(rec, o);
    } else {
      placed = ray.status;
    }
  }
  const bool act = valid && placed == SPLAT_RAY_ACTIVE;
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
  if (act) {
    order[pa + ra] = (int16_t)tid;
    if (keys_next) keys_next[off_a + pa + ra] = KEY_MISS;
  }
  if (rest) {
    order[na + pr + rr] = (int16_t)tid;
    if (hit_rest) hit_rest[off_r + pr + rr] = i;
  }
  __syncthreads();
  if (P.tex && !one_sample) {  // uniform: the final-evolved records are still staged
    RayF o{};
    FTriF F{};
    if (multi) {
      o = load_rayf(rec);
      F = load_ftri(P.tris + (i - 1));
    }
    splat_tile_multi(lds, multi, o, F, P, lm);
  }
  write_words(stage, order, 0, na * WORDS, reinterpret_cast<uint2*>(out_active + off_a));
  if (out_rest) write_words(stage, order, na * WORDS, nvalid * WORDS, reinterpret_cast<uint2*>(out_rest + off_r));
  if (one_sample) splat_tile(lds, has, sx, sy, sv, P, lm);
  __syncthreads();  // the LDS stage is reused by the caller's next tile
}

}  // namespace spl
