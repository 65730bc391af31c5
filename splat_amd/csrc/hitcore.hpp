// hitcore.hpp -- the per-wave ray/triangle traversal shared by the single-frame hitter kernels
// (hit.hip) and the multi-frame engine (engine.hip).
#pragma once

#include "internal.hpp"

namespace spl {

// ---- ray loading ----------------------------------------------------------------------------
struct RayLanes {
  float ox, oy, oz, dx, dy, dz;
  int64_t ig;
};
__device__ __forceinline__ RayLanes load_ray_lanes(const splat_adray_t* __restrict__ rays, int64_t r, bool valid) {
  RayLanes L{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0};
  if (valid) {
    const splat_adray_t* p = rays + r;
    L.ox = p->pos[0]; L.oy = p->pos[1]; L.oz = p->pos[2];
    L.dx = p->dir[0]; L.dy = p->dir[1]; L.dz = p->dir[2];
    L.ig = p->ignore_tri;
  }
  return L;
}

__device__ __forceinline__ int64_t live_count(int64_t R, const int64_t* __restrict__ d_count) {
  if (!d_count) return R;
  int64_t c = *d_count;
  return c < R ? c : R;
}

// ---- bounding volumes -------------------------------------------------------------------------
// The per-lane slab test against a conservative box (padded member bounds). Reciprocals are
// precomputed per ray; |d| < 1e-30 is clamped so 0 * inf never appears.
__device__ __forceinline__ bool slab_enter(const float4& lo, const float4& hi, const RayLanes& L, float ix, float iy,
                                           float iz, float best) {
  float t1 = (lo.x - L.ox) * ix, t2 = (hi.x - L.ox) * ix;
  float tmin = fmaxf(0.0f, fminf(t1, t2)), tmax = fmaxf(t1, t2);
  t1 = (lo.y - L.oy) * iy;
  t2 = (hi.y - L.oy) * iy;
  tmin = fmaxf(tmin, fminf(t1, t2));
  tmax = fminf(tmax, fmaxf(t1, t2));
  t1 = (lo.z - L.oz) * iz;
  t2 = (hi.z - L.oz) * iz;
  tmin = fmaxf(tmin, fminf(t1, t2));
  tmax = fminf(tmax, fmaxf(t1, t2));
  return (tmin <= tmax) && (tmin <= best);
}

// Prepared scene of the BV hitter as the kernels see it.
struct BVScene {
  const float4* boxes;
  const int2* ranges;
  const TriRec* recs;
  const float4* groups;
  int32_t n_bv, gsize, n_groups;
};

struct HitStats {
  unsigned long long n_box = 0, n_enter = 0, n_alg = 0, n_exec = 0;
};

// One wave, one ray per lane: the lexicographic (t, index) minimum over the scene's BV leaves,
// visited in order, optionally under the coarse level of groups of `gsize` consecutive leaves.
// A box no lane enters (one ballot) is skipped with its members; otherwise all lanes test its
// member records (a triangle held by two leaves ties with itself). Must be called by whole waves.
template <bool STATS, bool GROUPS>
__device__ __forceinline__ int32_t bv_traverse(const RayLanes& L, bool valid, const BVScene& S, HitStats& st) {
  const float4* __restrict__ boxes = S.boxes;
  const int2* __restrict__ ranges = S.ranges;
  const TriRec* __restrict__ recs = S.recs;
  const float4* __restrict__ groups = S.groups;
  const int32_t n_bv = S.n_bv, gsize = S.gsize, n_groups = S.n_groups;
  unsigned long long& n_box = st.n_box;
  unsigned long long& n_enter = st.n_enter;
  unsigned long long& n_alg = st.n_alg;
  unsigned long long& n_exec = st.n_exec;
  auto safe_inv = [](float d) {
    float a = fabsf(d) < 1e-30f ? copysignf(1e-30f, d) : d;
    return 1.0f / a;
  };
  const float ix = safe_inv(L.dx), iy = safe_inv(L.dy), iz = safe_inv(L.dz);
  float best = INFINITY;
  int32_t bi = 1;

  auto scan_leaf = [&](int32_t b) {
    const bool enter = valid && slab_enter(boxes[2 * b], boxes[2 * b + 1], L, ix, iy, iz, best);
    if constexpr (STATS) n_box += valid ? 1ull : 0ull;
    if (__builtin_amdgcn_ballot_w64(enter) == 0) return;
    const int2 rg = ranges[b];
    if constexpr (STATS) {
      n_enter += enter ? 1ull : 0ull;
      n_alg += enter ? (unsigned long long)rg.y : 0ull;
      n_exec += (unsigned long long)rg.y;
    }
#pragma unroll 4
    for (int32_t m = 0; m < rg.y; ++m) {
      const TriRec& T = recs[rg.x + m];
      float t;
      bool h = hit_rec(T.q0, T.q1, T.q2, T.q3, L.ox, L.oy, L.oz, L.dx, L.dy, L.dz, L.ig, t);
      int32_t idx = __float_as_int(T.q1.w);
      if (h && (t < best || (t == best && idx < bi))) {
        best = t;
        bi = idx;
      }
    }
  };

  if constexpr (GROUPS) {
    for (int32_t g = 0; g < n_groups; ++g) {
      const bool enter = valid && slab_enter(groups[2 * g], groups[2 * g + 1], L, ix, iy, iz, best);
      if constexpr (STATS) n_box += valid ? 1ull : 0ull;
      if (__builtin_amdgcn_ballot_w64(enter) == 0) continue;
      const int32_t b1 = (g + 1) * gsize < n_bv ? (g + 1) * gsize : n_bv;
      for (int32_t b = g * gsize; b < b1; ++b) scan_leaf(b);
    }
  } else {
    for (int32_t b = 0; b < n_bv; ++b) scan_leaf(b);
  }
  return bi;
}

}  // namespace spl
