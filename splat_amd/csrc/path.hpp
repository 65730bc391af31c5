// path.hpp -- per-ray device functions of the light-map path, shared by the single-frame kernels
// (trace.hip) and the multi-frame engine (engine.hip).
#pragma once

#include "internal.hpp"

namespace spl {

struct RayF {
  V3<float> pos, dp[3], dir, dd[3];
  bool in_medium;
  int64_t ignore_tri, dest;
  float lambda, x, y;
  uint8_t status;
};

__device__ __forceinline__ RayF load_rayf(const splat_adray_t* p) {
  RayF r;
  r.pos = ld3(p->pos); r.dp[0] = ld3(p->dpdl); r.dp[1] = ld3(p->dpdx); r.dp[2] = ld3(p->dpdy);
  r.dir = ld3(p->dir); r.dd[0] = ld3(p->dddl); r.dd[1] = ld3(p->dddx); r.dd[2] = ld3(p->dddy);
  r.in_medium = p->in_medium != 0; r.ignore_tri = p->ignore_tri; r.dest = p->dest;
  r.lambda = p->lambda; r.x = p->x; r.y = p->y; r.status = p->status;
  return r;
}
// writes every field a bounce can change (the rest -- dest, lambda, x, y -- is left in place)
__device__ __forceinline__ void store_rayf(splat_adray_t* p, const RayF& r) {
  st3(p->pos, r.pos); st3(p->dpdl, r.dp[0]); st3(p->dpdx, r.dp[1]); st3(p->dpdy, r.dp[2]);
  st3(p->dir, r.dir); st3(p->dddl, r.dd[0]); st3(p->dddx, r.dd[1]); st3(p->dddy, r.dd[2]);
  p->in_medium = r.in_medium ? 1 : 0; p->ignore_tri = r.ignore_tri; p->status = r.status;
}

__device__ __forceinline__ int64_t live_count_t(int64_t R, const int64_t* __restrict__ d_count) {
  if (!d_count) return R;
  int64_t c = *d_count;
  return c < R ? c : R;
}

__device__ __forceinline__ V3<D3<float>> pos_dual(const RayF& r) {
  return V3<D3<float>>{{r.pos.x, {r.dp[0].x, r.dp[1].x, r.dp[2].x}},
                       {r.pos.y, {r.dp[0].y, r.dp[1].y, r.dp[2].y}},
                       {r.pos.z, {r.dp[0].z, r.dp[1].z, r.dp[2].z}}};
}
__device__ __forceinline__ V3<D3<float>> dir_dual(const RayF& r) {
  return V3<D3<float>>{{r.dir.x, {r.dd[0].x, r.dd[1].x, r.dd[2].x}},
                       {r.dir.y, {r.dd[0].y, r.dd[1].y, r.dd[2].y}},
                       {r.dir.z, {r.dd[0].z, r.dd[1].z, r.dd[2].z}}};
}
// next_p at the expansion point (src/tracer.jl:109-116): p + d t, as duals
__device__ __forceinline__ V3<D3<float>> next_p_dual(const V3<D3<float>>& P0, const V3<D3<float>>& D0, const D3<float>& t) {
  return V3<D3<float>>{FMA(D0.x, t, P0.x), FMA(D0.y, t, P0.y), FMA(D0.z, t, P0.z)};
}
__device__ __forceinline__ void split(const V3<D3<float>>& a, V3<float>* v, V3<float>* d3) {
  *v = vval(a);
#pragma unroll
  for (int k = 0; k < 3; ++k) d3[k] = V3<float>{a.x.d[k], a.y.d[k], a.z.d[k]};
}

// get_hit((i, T), ::ADRay) (src/tracer.jl:61-81): hit test on the value, t as a dual
__device__ __forceinline__ bool get_hit_ad(const RayF& r, const FTriF& F, int32_t i, D3<float>* tD) {
  TriRec P = tri_prep(F.v[0], F.v[1], F.v[2], F.v[3], i);
  float tv;
  if (!hit_rec(P.q0, P.q1, P.q2, P.q3, r.pos.x, r.pos.y, r.pos.z, r.dir.x, r.dir.y, r.dir.z, r.ignore_tri, tv)) {
    *tD = D3<float>{INFINITY, {INFINITY, INFINITY, INFINITY}};
    return false;
  }
  V3<float> n{P.q0.x, P.q0.y, P.q0.z};
  D3<float> no = dot3(n, pos_dual(r));
  D3<float> nd = dot3(n, dir_dual(r));
  D3<float> num = P.q0.w - no;
  *tD = num / nd;
  return true;
}

// The status evolve_one gives a live ray whose nearest hit (from the hitter) is i, known before
// the evolve math: for i != 1 evolve's own hit test on triangle i repeats the hitter's winning
// test (same record, same ray) and succeeds; for i == 1 it tests the sentinel triangle, which
// only matters when first_diffuse > 1 (tri1_hit: that test, done by the caller only then).
__device__ __forceinline__ uint8_t status_after(int32_t i, int32_t first_diffuse, bool tri1_hit) {
  if (i != 1) return i >= first_diffuse ? SPLAT_RAY_DIFFUSE : SPLAT_RAY_ACTIVE;
  if (1 >= first_diffuse) return SPLAT_RAY_DIFFUSE;
  return tri1_hit ? SPLAT_RAY_ACTIVE : SPLAT_RAY_INFINITY;
}
// evolve's hit test of the sentinel triangle (tris[0]) for a ray
__device__ __forceinline__ bool sentinel_hit(const splat_ftri_t* __restrict__ tris, const V3<float>& pos,
                                             const V3<float>& dir, int64_t ignore_tri) {
  const FTriF F1 = load_ftri(tris);
  const TriRec P = tri_prep(F1.v[0], F1.v[1], F1.v[2], F1.v[3], 1);
  float t;
  return hit_rec(P.q0, P.q1, P.q2, P.q3, pos.x, pos.y, pos.z, dir.x, dir.y, dir.z, ignore_tri, t);
}
// Whether the sentinel's test can succeed at all: with a zero normal it computes t = 0/0 (NaN)
// for every ray, so it never does (the usual, degenerate sentinel) and no ray need be read.
__device__ __forceinline__ bool sentinel_live(const splat_ftri_t* __restrict__ tris) {
  return tris[0].v[0][0] != 0.0f || tris[0].v[0][1] != 0.0f || tris[0].v[0][2] != 0.0f;
}
__device__ __forceinline__ uint8_t predict_status(int32_t i, int32_t first_diffuse, const splat_ftri_t* __restrict__ tris,
                                                  const V3<float>& pos, const V3<float>& dir, int64_t ignore_tri) {
  const bool need = i == 1 && first_diffuse > 1 && sentinel_live(tris);
  return status_after(i, first_diffuse, need && sentinel_hit(tris, pos, dir, ignore_tri));
}

// evolve_ray + handle_optics (src/tracer.jl:181-201, :118-179). A ray that becomes DIFFUSE keeps
// its position and direction, so its hit test (hit, dual distance) is what final_evolution's own
// test on the same triangle gives: returned in *diffuse_hit / *diffuse_tD when asked for.
__device__ __forceinline__ RayF evolve_one(const RayF& r, int32_t i, const FTriF& F, float rndm, int32_t first_diffuse,
                                           const splat_optics_t& opt, bool* diffuse_hit = nullptr,
                                           D3<float>* diffuse_tD = nullptr) {
  D3<float> tD;
  bool hit = get_hit_ad(r, F, i, &tD);
  int32_t i2 = hit ? i : 1;
  RayF out = r;
  if (i2 >= first_diffuse) {
    out.status = SPLAT_RAY_DIFFUSE;
    if (diffuse_hit) {
      *diffuse_hit = hit;
      *diffuse_tD = tD;
    }
    return out;
  }
  if (!hit) { out.status = SPLAT_RAY_INFINITY; return out; }

  D3<float> g = glass_ior(opt, r.lambda), a{1.0f, {0.0f, 0.0f, 0.0f}};
  D3<float> n1 = r.in_medium ? g : a, n2 = r.in_medium ? a : g;

  V3<D3<float>> P0 = pos_dual(r), D0 = dir_dual(r);
  V3<D3<float>> P = next_p_dual(P0, D0, tD);
  V3<D3<float>> Nd = optical_normal(F, P);
  V3<float> Nv = vval(Nd);
  bool refracts = can_refract(r.dir, Nv, n1.v, n2.v) && (rndm > reflectance(r.dir, Nv, n1.v, n2.v));

  V3<D3<float>> Dout;
  if (refracts) {
    V3<D3<float>> Nr = Nd;
    if (opt.compat & SPLAT_COMPAT_REFRACT_NORMAL_LAMBDA) { Nr.x.d[0] = 0.0f; Nr.y.d[0] = 0.0f; Nr.z.d[0] = 0.0f; }
    Dout = refract(D0, Nr, n1, n2);
  } else {
    Dout = reflect(D0, Nd);
  }
  split(P, &out.pos, out.dp);
  split(Dout, &out.dir, out.dd);

  if ((opt.compat & SPLAT_COMPAT_DY_AT_X) && r.x != r.y) {
    // d dir / d y evaluated at y = r.x, as the reference does (src/tracer.jl:138-141, :166-169)
    float delta = r.x - r.y;
    float z = 0.0f;
    V3<D3<float>> Py0{{FMA(r.dp[2].x, delta, r.pos.x), {z, z, r.dp[2].x}},
                      {FMA(r.dp[2].y, delta, r.pos.y), {z, z, r.dp[2].y}},
                      {FMA(r.dp[2].z, delta, r.pos.z), {z, z, r.dp[2].z}}};
    V3<D3<float>> Dy0{{FMA(r.dd[2].x, delta, r.dir.x), {z, z, r.dd[2].x}},
                      {FMA(r.dd[2].y, delta, r.dir.y), {z, z, r.dd[2].y}},
                      {FMA(r.dd[2].z, delta, r.dir.z), {z, z, r.dd[2].z}}};
    D3<float> sy{FMA(tD.d[2], delta, tD.v), {z, z, tD.d[2]}};
    V3<D3<float>> Py = next_p_dual(Py0, Dy0, sy);
    V3<D3<float>> Ny = optical_normal(F, Py);
    V3<D3<float>> Dy = refracts ? refract(Dy0, Ny, n1.v, n2.v) : reflect(Dy0, Ny);
    out.dd[2] = V3<float>{Dy.x.d[2], Dy.y.d[2], Dy.z.d[2]};
  }
  out.in_medium = refracts ? !r.in_medium : r.in_medium;
  out.ignore_tri = i;
  out.status = SPLAT_RAY_ACTIVE;
  return out;
}

__device__ __forceinline__ float bounce_random(float force_rand, const float* __restrict__ rndm, int64_t r,
                                              uint64_t seed, int32_t bounce, int64_t dest) {
  return !__builtin_isnan(force_rand) ? force_rand
         : (rndm ? rndm[r] : u01(seed, RNG_RNDM | (uint32_t)bounce, (uint64_t)dest));
}

// ---- final_evolution + atomic splat (src/ray_imagers.jl:4-37, :78-113) ---------------------
// Piecewise-linear curve c at lam: segment m = the first with lam < x[m + 1] (at most n - 2).
// The breakpoints are the same for every lane, so the segment is found by one uniform pass
// over them that advances with selects (no per-lane loop, no per-lane indexing of the
// parameters); the same segment and the same arithmetic as a scan from the left.
__device__ __forceinline__ float curve_eval(const splat_lightmap_params_t& lm, int c, float lam) {
  const int n = lm.curve_n[c];
  float x0 = lm.curve_x[c][0], x1 = lm.curve_x[c][1];
  float y0 = lm.curve_y[c][0], y1 = lm.curve_y[c][1];
  bool go = true;
#pragma unroll 4
  for (int j = 1; j < n - 1; ++j) {
    {
      // values first, then selects (a select of the parameters' addresses would put them in
      // private memory)
      const float xa = lm.curve_x[c][j], ya = lm.curve_y[c][j], xb = lm.curve_x[c][j + 1], yb = lm.curve_y[c][j + 1];
      go = go && lam >= xa;
      x0 = go ? xa : x0;
      y0 = go ? ya : y0;
      x1 = go ? xb : x1;
      y1 = go ? yb : y1;
    }
  }
  float slope = (y1 - y0) / (x1 - x0);
  return FMA(slope, lam - x1, y1);
}

// The curves' segments as a table a workgroup fills once: for curve c and segment m, its right
// end (x1, y1) and slope (y1 - y0) / (x1 - x0), the same FP32 operations curve_eval performs, so
// curve_eval_tab returns the same value. Finding the segment then takes one compare and one
// select per breakpoint (the abscissae stay wave-uniform operands) instead of carrying four
// values through every breakpoint.
struct CurveTab {
  float x1[3][SPLAT_MAX_CURVE - 1], y1[3][SPLAT_MAX_CURVE - 1], slope[3][SPLAT_MAX_CURVE - 1];
};
// thread tid of nthreads fills its share of T (the caller synchronises before any read)
__device__ __forceinline__ void curve_tab_fill(CurveTab& T, const splat_lightmap_params_t& lm, int tid, int nthreads) {
  for (int e = tid; e < 3 * (SPLAT_MAX_CURVE - 1); e += nthreads) {
    const int c = e / (SPLAT_MAX_CURVE - 1), m = e - c * (SPLAT_MAX_CURVE - 1);
    if (m + 1 < lm.curve_n[c]) {
      const float x0 = lm.curve_x[c][m], x1 = lm.curve_x[c][m + 1], y0 = lm.curve_y[c][m], y1 = lm.curve_y[c][m + 1];
      T.x1[c][m] = x1;
      T.y1[c][m] = y1;
      T.slope[c][m] = (y1 - y0) / (x1 - x0);
    }
  }
}
// curve_eval through the table: segment m = the last with lam >= x[m] (x ascending, so the same
// segment as curve_eval's scan; 0 when lam is NaN or below x[1])
__device__ __forceinline__ float curve_eval_tab(const CurveTab& T, const splat_lightmap_params_t& lm, int c, float lam) {
  const int n = lm.curve_n[c];
  int m = 0;
  // unrolled over the largest curve, so the abscissae load together (a loop over n would wait
  // for each one's scalar load in turn)
#pragma unroll
  for (int j = 1; j < SPLAT_MAX_CURVE - 1; ++j) m = j < n - 1 && lam >= lm.curve_x[c][j] ? j : m;
  return FMA(T.slope[c][m], lam - T.x1[c][m], T.y1[c][m]);
}

// final_evolution (src/ray_imagers.jl:4-37) of one DIFFUSE ray, in place, given its hit test on
// triangle i (hit, dual distance tD)
__device__ __forceinline__ void final_evolve_hit(RayF& r, int32_t i, bool hit, const D3<float>& tD) {
  if (hit) {
    V3<D3<float>> P = next_p_dual(pos_dual(r), dir_dual(r), tD);
    split(P, &r.pos, r.dp);
  } else {
    r.pos = V3<float>{NAN, NAN, NAN};
    r.dp[0] = r.dp[1] = r.dp[2] = r.pos;
  }
  r.ignore_tri = hit ? i : 1;
}
__device__ __forceinline__ void final_evolve(RayF& r, int32_t i, const FTriF& F) {
  D3<float> tD;
  const bool hit = get_hit_ad(r, F, i, &tD);
  final_evolve_hit(r, i, hit, tD);
}

// One splat sample (src/ray_imagers.jl:78-113) of a final-evolved DIFFUSE ray: wavelength k
// (lam), sub-pixel offsets (ix, iy). Returns false when the texture coordinate is not finite
// (no splat); else the texel (x, y) (0-based) and the cosine weight I.
__device__ __forceinline__ bool splat_sample(const RayF& r, const FTriF& F, const splat_lightmap_params_t& lm, float lam,
                                             int ix, int iy, int32_t* tx, int32_t* ty, float* I) {
  const float W = (float)lm.width, H = (float)lm.height;
  float xs = r.x + lm.delta[ix], ys = r.y + lm.delta[iy];
  float dl = lam - r.lambda, dx = xs - r.x, dy = ys - r.y;
  V3<float> p{FMA(r.dp[2].x, dy, FMA(r.dp[1].x, dx, FMA(r.dp[0].x, dl, r.pos.x))),
              FMA(r.dp[2].y, dy, FMA(r.dp[1].y, dx, FMA(r.dp[0].y, dl, r.pos.y))),
              FMA(r.dp[2].z, dy, FMA(r.dp[1].z, dx, FMA(r.dp[0].z, dl, r.pos.z)))};
  V3<float> d{FMA(r.dd[2].x, dy, FMA(r.dd[1].x, dx, FMA(r.dd[0].x, dl, r.dir.x))),
              FMA(r.dd[2].y, dy, FMA(r.dd[1].y, dx, FMA(r.dd[0].y, dl, r.dir.y))),
              FMA(r.dd[2].z, dy, FMA(r.dd[1].z, dx, FMA(r.dd[0].z, dl, r.dir.z)))};
  float bu, bv;
  reverse_uv(F, p, &bu, &bv);
  float w = (1.0f - bu) - bv;
  float U = FMA(F.v[9].x, w, FMA(F.v[8].x, bv, F.v[7].x * bu));
  float V = FMA(F.v[9].y, w, FMA(F.v[8].y, bv, F.v[7].y * bu));
  if (!__builtin_isfinite(U) || !__builtin_isfinite(V)) return false;
  float fi = ceilf(U * W), fj = ceilf(V * H);
  fi = fi < 1.0f ? 1.0f : (fi > W ? W : fi);
  fj = fj < 1.0f ? 1.0f : (fj > H ? H : fj);
  *tx = (int32_t)fi - 1;
  *ty = (int32_t)fj - 1;
  V3<float> nrm = optical_normal(F, p);
  *I = fabsf(dot3(nrm, normalize3(d)));
  return true;
}

// Light-map planes a sample of wavelength k writes, and the value for each: plane k with I
// (identity weights), or each channel c with I times the table or curve weight of lam.
__device__ __forceinline__ int splat_planes(const splat_lightmap_params_t& lm) {
  return lm.weight_mode == SPLAT_WEIGHTS_IDENTITY ? 1 : lm.channels;
}
__device__ __forceinline__ int32_t splat_plane(const splat_lightmap_params_t& lm, int k, int c) {
  return lm.weight_mode == SPLAT_WEIGHTS_IDENTITY ? k : c;
}
__device__ __forceinline__ float splat_value(const splat_lightmap_params_t& lm, int k, int c, float lam, float I) {
  if (lm.weight_mode == SPLAT_WEIGHTS_IDENTITY) return I;
  if (lm.weight_mode == SPLAT_WEIGHTS_TABLE) return I * lm.weights[k * 3 + c];
  return I * curve_eval(lm, c, lam);
}

// final_evolution of one DIFFUSE ray, in place, then all its splats (k, dy, dx) straight into
// the map with global atomics, in the reference's thread order
__device__ __forceinline__ void final_evolve_splat(RayF& r, int32_t i, const FTriF& F, const splat_lightmap_params_t& lm,
                                                   float* __restrict__ tex) {
  final_evolve(r, i, F);
  const int nl = lm.n_lambda > 0 ? lm.n_lambda : 1;
  const int64_t plane = (int64_t)lm.width * lm.height;
  for (int k = 0; k < nl; ++k) {
    const float lam = lm.n_lambda > 0 ? lm.spectrum[k] : r.lambda;
    for (int iy = 0; iy < lm.n_delta; ++iy) {
      for (int ix = 0; ix < lm.n_delta; ++ix) {
        int32_t tx, ty;
        float I;
        if (!splat_sample(r, F, lm, lam, ix, iy, &tx, &ty, &I)) continue;
        const int64_t base = (int64_t)tx + (int64_t)lm.width * ty;
        for (int c = 0; c < splat_planes(lm); ++c)
          unsafeAtomicAdd(tex + base + plane * splat_plane(lm, k, c), splat_value(lm, k, c, lam, I));
      }
    }
  }
}

// final evolution + splat of one DIFFUSE ray loaded from ray_out, written back in place: final
// evolution changes only the position, its three tangents and ignore_tri, so only those 56 of the
// record's 136 bytes are stored (the rest of the record already holds the ray)
__device__ __forceinline__ void final_evolve_and_splat(RayF r, int32_t i, const FTriF& F,
                                                       const splat_lightmap_params_t& lm, float* __restrict__ tex,
                                                       splat_adray_t* ray_out) {
  final_evolve_splat(r, i, F, lm, tex);
  st3(ray_out->pos, r.pos); st3(ray_out->dpdl, r.dp[0]); st3(ray_out->dpdx, r.dp[1]); st3(ray_out->dpdy, r.dp[2]);
  ray_out->ignore_tri = r.ignore_tri;
}

// ---- backward shading: shade_tex4_kernel (src/ray_imagers.jl:381-436) ------------------------
// Bilinear sample of one float plane (W along u, H along v, u fastest) at 1-based (u, v), clamp
// to edge: the reference's CUDA.jl linear-interpolation texture reads, with exact FP32 weights.
__device__ __forceinline__ float tex_bilinear(const float* __restrict__ plane, int64_t W, int64_t H, float u, float v) {
  float fu = floorf(u), fv = floorf(v);
  const float a = u - fu, b = v - fv;
  fu = fu < 0.f ? 0.f : (fu > (float)W ? (float)W : fu);
  fv = fv < 0.f ? 0.f : (fv > (float)H ? (float)H : fv);
  int64_t i0 = (int64_t)fu, j0 = (int64_t)fv;
  int64_t i1 = i0 + 1 > W ? W : i0 + 1, j1 = j0 + 1 > H ? H : j0 + 1;
  i0 = i0 < 1 ? 1 : i0;
  j0 = j0 < 1 ? 1 : j0;
  i1 = i1 < 1 ? 1 : i1;
  j1 = j1 < 1 ? 1 : j1;
  const float t00 = plane[(i0 - 1) + W * (j0 - 1)], t10 = plane[(i1 - 1) + W * (j0 - 1)];
  const float t01 = plane[(i0 - 1) + W * (j1 - 1)], t11 = plane[(i1 - 1) + W * (j1 - 1)];
  const float r0 = FMA(a, t10 - t00, t00), r1 = FMA(a, t11 - t01, t01);
  return FMA(b, r1 - r0, r0);
}

// One (ray, a, b) sample for a final-evolved ray r on triangle F: an INFINITY ray reads the
// skydome (3 planes SW x SH) at its direction's polar angles; a DIFFUSE ray sums the retina-
// weighted (lm.weights[k*3 + c]) light-map values over the spectrum at its texture coordinate;
// channels clamped to [0, 1].
__device__ __forceinline__ void shade_sample(const RayF& r, const FTriF& F, const splat_lightmap_params_t& lm,
                                             const float* __restrict__ tex, const float* __restrict__ sky, int64_t SW,
                                             int64_t SH, int a, int b, float rgb[3]) {
  rgb[0] = rgb[1] = rgb[2] = 0.f;
  const float pi = 3.14159265358979323846f;
  if (r.status == SPLAT_RAY_INFINITY && sky) {
    const float xs = r.x + lm.delta[a], ys = r.y + lm.delta[b];
    const float dx = xs - r.x, dy = ys - r.y;
    const V3<float> d{FMA(r.dd[2].x, dy, FMA(r.dd[1].x, dx, r.dir.x)), FMA(r.dd[2].y, dy, FMA(r.dd[1].y, dx, r.dir.y)),
                      FMA(r.dd[2].z, dy, FMA(r.dd[1].z, dx, r.dir.z))};
    const float th = acosf(d.z), ph = atan2f(d.y, d.x);
    float v = ((ph / pi) + 1.f) * 0.5f, u = th / pi;
    u = u * (float)SW;
    v = v * (float)SH;
    if (__builtin_isfinite(u) && __builtin_isfinite(v))
      for (int c = 0; c < 3; ++c) rgb[c] = tex_bilinear(sky + SW * SH * c, SW, SH, u, v);
  } else if (r.status == SPLAT_RAY_DIFFUSE) {
    const float W = (float)lm.width, H = (float)lm.height;
    for (int k = 0; k < lm.n_lambda; ++k) {
      const float lam = lm.spectrum[k];
      const float xs = r.x + lm.delta[a], ys = r.y + lm.delta[b];
      const float dl = lam - r.lambda, dx = xs - r.x, dy = ys - r.y;
      const V3<float> p{FMA(r.dp[2].x, dy, FMA(r.dp[1].x, dx, FMA(r.dp[0].x, dl, r.pos.x))),
                        FMA(r.dp[2].y, dy, FMA(r.dp[1].y, dx, FMA(r.dp[0].y, dl, r.pos.y))),
                        FMA(r.dp[2].z, dy, FMA(r.dp[1].z, dx, FMA(r.dp[0].z, dl, r.pos.z)))};
      float bu, bv;
      reverse_uv(F, p, &bu, &bv);
      const float w = (1.0f - bu) - bv;
      const float U = FMA(F.v[9].x, w, FMA(F.v[8].x, bv, F.v[7].x * bu));
      const float V = FMA(F.v[9].y, w, FMA(F.v[8].y, bv, F.v[7].y * bu));
      if (__builtin_isnan(U) || __builtin_isnan(V)) continue;
      const float X = tex_bilinear(tex + (int64_t)lm.width * lm.height * k, lm.width, lm.height, U * W, V * H);
      for (int c = 0; c < 3; ++c) rgb[c] = FMA(lm.weights[k * 3 + c], X, rgb[c]);
    }
  }
  for (int c = 0; c < 3; ++c) rgb[c] = rgb[c] > 1.f ? 1.f : (rgb[c] < 0.f ? 0.f : rgb[c]);
}

// camera_ray (src/ray_generators.jl:18-50) for ray q of a pinhole frame, evaluated in Float64 and
// rounded to the Float32 ADRay fields as the reference does
__device__ __forceinline__ splat_adray_t raygen_ray(const splat_pinhole_t& cam, int64_t q) {

  const int64_t H = cam.height, Wd = cam.width, L = cam.samples_per_pixel;
  const int64_t P = q / L, l = q % L;
  int64_t x, y;
  if (cam.order == SPLAT_ORDER_TILED8 && H % 8 == 0 && Wd % 8 == 0) {
    int64_t tile = P / 64, w = P % 64, tx = tile % (H / 8), ty = tile / (H / 8);
    x = tx * 8 + (w % 8) + 1;
    y = ty * 8 + (w / 8) + 1;
  } else {
    x = P % H + 1;
    y = P / H + 1;
  }
  const int64_t dest = ((y - 1) * H + (x - 1)) * L + l + 1;
  float lam = cam.lambda_min;
  const float span = cam.lambda_max - cam.lambda_min;
  if (cam.lambda_mode == SPLAT_LAMBDA_UNIFORM) {
    lam = FMA(u01(cam.seed, RNG_LAMBDA, (uint64_t)dest), span, cam.lambda_min);
  } else if (cam.lambda_mode == SPLAT_LAMBDA_STRATIFIED) {
    float u = u01(cam.seed, RNG_LAMBDA, (uint64_t)dest);
    lam = FMA(((float)l + u) / (float)L, span, cam.lambda_min);
  }
  float jit[3] = {0.f, 0.f, 0.f};
  if (cam.jitter != 0.f) {
    float mx = (float)(H > Wd ? H : Wd);
#pragma unroll
    for (int k = 0; k < 3; ++k) jit[k] = (u01(cam.seed, RNG_JITTER | (uint64_t)k, (uint64_t)dest) * cam.jitter) / mx;
  }
  const double scale = ((double)H * (1.0 / __builtin_sqrt(2.0))) / (double)cam.fov_half_sin;
  D3<double> X{(double)x, {0.0, 1.0, 0.0}}, Y{(double)y, {0.0, 0.0, 1.0}};
  D3<double> xd = (X - (double)H / 2.0) / scale;
  D3<double> yd = (Y - (double)Wd / 2.0) / scale;
  D3<double> zd = SQRT((1.0 - xd * xd) - yd * yd);
  splat_adray_t o;
  for (int k = 0; k < 3; ++k) {
    D3<double> dk = ((xd * (double)cam.right[k] + yd * (double)cam.up[k]) + zd * (double)cam.dir[k]) + (double)jit[k];
    o.pos[k] = cam.pos[k];
    o.dir[k] = (float)dk.v;
    o.dddx[k] = (float)dk.d[1];
    o.dddy[k] = (float)dk.d[2];
    o.dpdl[k] = 0.f; o.dpdx[k] = 0.f; o.dpdy[k] = 0.f; o.dddl[k] = 0.f;
  }
  o.in_medium = 0;
  for (int k = 0; k < 7; ++k) o._pad0[k] = 0;
  o.ignore_tri = 1;
  o.dest = dest;
  o.lambda = lam;
  o.x = 0.f;
  o.y = 0.f;
  o.status = SPLAT_RAY_ACTIVE;
  o._pad1[0] = o._pad1[1] = o._pad1[2] = 0;
  return o;
}
__device__ __forceinline__ void raygen_one(const splat_pinhole_t& cam, int64_t q, splat_adray_t* out) {
  *out = raygen_ray(cam, q);
}

// The lights of one rays_from_lights call, passed by value: light i owns rays [base[i], base[i+1]).
struct LightSet {
  splat_rect_light_t light[SPLAT_MAX_LIGHTS];
  int64_t base[SPLAT_MAX_LIGHTS + 1];
  splat_light_rays_t p;
  int32_t n;
};
constexpr uint64_t RNG_LIGHT = 4ull << 32;  // cell offsets: stream RNG_LIGHT | 2 light (+1 for y)

// rect_light_ray (src/ray_generators.jl:66-94) for ray q of the set: Float32 in the reference's
// operation order, origin = (center + (dim1 * sx) / h2) + (dim2 * sy) / w2 with sx = (x + dx) - h2,
// and its ForwardDiff tangents d origin / dx = dim1 / h2, d origin / dy = dim2 / w2 (h2 = h / 2 and
// w2 = w / 2 are integer halves; h = 1 divides by zero exactly as the reference does).
__device__ __forceinline__ splat_adray_t light_ray(const LightSet& A, int64_t q) {
  int i = 0;
  while (i + 1 < A.n && q >= A.base[i + 1]) ++i;
  const splat_rect_light_t& L = A.light[i];
  const int64_t j = q - A.base[i];
  const int64_t h = L.res1 / A.p.upscale, w = L.res2 / A.p.upscale;
  const int64_t x = j % h + 1, y = j / h + 1;
  const float dx = A.p.jitter * u01(A.p.seed, RNG_LIGHT | (uint64_t)(2 * i), (uint64_t)j);
  const float dy = A.p.jitter * u01(A.p.seed, RNG_LIGHT | (uint64_t)(2 * i + 1), (uint64_t)j);
  const float X = (float)x + dx, Y = (float)y + dy;
  const float h2 = (float)(h / 2), w2 = (float)(w / 2);
  const float sx = X - h2, sy = Y - w2;
  float lam = A.p.lambda_min;
  if (A.p.lambda_mode == SPLAT_LAMBDA_UNIFORM)
    lam = FMA(u01(A.p.seed, RNG_LAMBDA | (uint64_t)(i + 1), (uint64_t)j), A.p.lambda_max - A.p.lambda_min,
              A.p.lambda_min);
  splat_adray_t o;
  for (int k = 0; k < 3; ++k) {
    o.pos[k] = (L.center[k] + (L.dim1[k] * sx) / h2) + (L.dim2[k] * sy) / w2;
    o.dpdl[k] = 0.f;
    o.dpdx[k] = L.dim1[k] / h2;
    o.dpdy[k] = L.dim2[k] / w2;
    o.dir[k] = L.dir[k];
    o.dddl[k] = 0.f; o.dddx[k] = 0.f; o.dddy[k] = 0.f;
  }
  o.in_medium = 0;
  for (int k = 0; k < 7; ++k) o._pad0[k] = 0;
  o.ignore_tri = 1;
  o.dest = 0;  // forward rays: the pixel is not known ahead of time
  o.lambda = lam;
  o.x = X;
  o.y = Y;
  o.status = SPLAT_RAY_ACTIVE;
  o._pad1[0] = o._pad1[1] = o._pad1[2] = 0;
  return o;
}

}  // namespace spl
