// oracle.cpp -- CPU restatement of SPLAT's per-frame tracing hot path.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product (splat_amd/, include/) links, loads or calls
// this file; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the
// checker. Every function below restates one reference function (cited file:line, paths relative
// to the reference checkout) and is instantiated twice:
//   * REAL = float  -- the FP32 spec. The GPU kernels must reproduce it bit for bit (hit indices,
//                      partition permutations, ray states); it is written from the same operation
//                      list as the kernels, with every fused multiply-add explicit (fma) and this
//                      file compiled with -ffp-contract=off, so the compiler adds none.
//   * REAL = double -- the FP64 tolerance spec for distances, tangents and the light map. Inputs
//                      are the same float32 arrays, widened exactly; constants are the float32
//                      constants widened, so both forms describe the same problem.
//
// The reference is Julia (CUDA.jl + ForwardDiff); it cannot run here (no Julia toolchain), so
// this restatement is pinned instead by the reference's own known-answer tests (test/geo.jl:11-29,
// src/atomic_argmin.jl:40-96), by closed forms of the reference formulas, by central finite
// differences of the FP64 value path against the FP64 dual tangents, and by brute force on tiny
// scenes (tests/test_oracle_*.py). Deliberate choices where the reference is unpinnable
// (ForwardDiff's exact FP32 rounding, @fastmath, unseeded CUDA.rand) are the explicit operation
// orders below; they are the spec.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <type_traits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/splat_hip.h"

namespace {

// ------------------------------------------------------------------------------------------
// scalar / dual arithmetic. A dual carries a value and three tangents (d/dlambda, d/dx, d/dy),
// replacing the reference's nine scalar ForwardDiff.derivative calls (src/tracer.jl:126-169)
// with one pass; each tangent is computed from the values and its own input tangents only, so
// the result per tangent is the same as three single-tangent passes.
// ------------------------------------------------------------------------------------------

inline float FMA(float a, float b, float c) { return std::fmaf(a, b, c); }
inline double FMA(double a, double b, double c) { return std::fma(a, b, c); }
inline float SQRT(float a) { return std::sqrt(a); }
inline double SQRT(double a) { return std::sqrt(a); }

template <class T>
struct D3 {
  T v;
  T d[3];
};

template <class T> inline T val(T a) { return a; }
template <class T> inline T val(const D3<T>& a) { return a.v; }

#define D3_LOOP for (int k = 0; k < 3; ++k)

template <class T> inline D3<T> operator+(const D3<T>& a, const D3<T>& b) {
  D3<T> r; r.v = a.v + b.v; D3_LOOP r.d[k] = a.d[k] + b.d[k]; return r;
}
template <class T> inline D3<T> operator+(const D3<T>& a, T b) {
  D3<T> r; r.v = a.v + b; D3_LOOP r.d[k] = a.d[k]; return r;
}
template <class T> inline D3<T> operator+(T a, const D3<T>& b) {
  D3<T> r; r.v = a + b.v; D3_LOOP r.d[k] = b.d[k]; return r;
}
template <class T> inline D3<T> operator-(const D3<T>& a, const D3<T>& b) {
  D3<T> r; r.v = a.v - b.v; D3_LOOP r.d[k] = a.d[k] - b.d[k]; return r;
}
template <class T> inline D3<T> operator-(const D3<T>& a, T b) {
  D3<T> r; r.v = a.v - b; D3_LOOP r.d[k] = a.d[k]; return r;
}
template <class T> inline D3<T> operator-(T a, const D3<T>& b) {
  D3<T> r; r.v = a - b.v; D3_LOOP r.d[k] = -b.d[k]; return r;
}
template <class T> inline D3<T> operator-(const D3<T>& a) {
  D3<T> r; r.v = -a.v; D3_LOOP r.d[k] = -a.d[k]; return r;
}
template <class T> inline D3<T> operator*(const D3<T>& a, const D3<T>& b) {
  D3<T> r; r.v = a.v * b.v; D3_LOOP r.d[k] = FMA(a.v, b.d[k], a.d[k] * b.v); return r;
}
template <class T> inline D3<T> operator*(const D3<T>& a, T s) {
  D3<T> r; r.v = a.v * s; D3_LOOP r.d[k] = a.d[k] * s; return r;
}
template <class T> inline D3<T> operator*(T s, const D3<T>& a) {
  D3<T> r; r.v = s * a.v; D3_LOOP r.d[k] = s * a.d[k]; return r;
}
template <class T> inline D3<T> operator/(const D3<T>& a, const D3<T>& b) {
  D3<T> r; r.v = a.v / b.v; D3_LOOP r.d[k] = FMA(-r.v, b.d[k], a.d[k]) / b.v; return r;
}
template <class T> inline D3<T> operator/(const D3<T>& a, T s) {
  D3<T> r; r.v = a.v / s; D3_LOOP r.d[k] = a.d[k] / s; return r;
}
template <class T> inline D3<T> operator/(T s, const D3<T>& b) {
  D3<T> r; r.v = s / b.v; D3_LOOP r.d[k] = (-r.v * b.d[k]) / b.v; return r;
}
template <class T> inline D3<T> SQRT(const D3<T>& a) {
  D3<T> r; r.v = SQRT(a.v); T two_v = r.v + r.v; D3_LOOP r.d[k] = a.d[k] / two_v; return r;
}
// fused multiply-add over every mix of scalar (s) and dual (D) operands
template <class T> inline D3<T> FMA(const D3<T>& a, const D3<T>& b, const D3<T>& c) {
  D3<T> r; r.v = FMA(a.v, b.v, c.v); D3_LOOP r.d[k] = FMA(a.v, b.d[k], FMA(a.d[k], b.v, c.d[k])); return r;
}
template <class T> inline D3<T> FMA(T a, const D3<T>& b, const D3<T>& c) {
  D3<T> r; r.v = FMA(a, b.v, c.v); D3_LOOP r.d[k] = FMA(a, b.d[k], c.d[k]); return r;
}
template <class T> inline D3<T> FMA(const D3<T>& a, T b, const D3<T>& c) {
  D3<T> r; r.v = FMA(a.v, b, c.v); D3_LOOP r.d[k] = FMA(a.d[k], b, c.d[k]); return r;
}
template <class T> inline D3<T> FMA(const D3<T>& a, const D3<T>& b, T c) {
  D3<T> r; r.v = FMA(a.v, b.v, c); D3_LOOP r.d[k] = FMA(a.v, b.d[k], a.d[k] * b.v); return r;
}
template <class T> inline D3<T> FMA(T a, T b, const D3<T>& c) {
  D3<T> r; r.v = FMA(a, b, c.v); D3_LOOP r.d[k] = c.d[k]; return r;
}
template <class T> inline D3<T> FMA(T a, const D3<T>& b, T c) {
  D3<T> r; r.v = FMA(a, b.v, c); D3_LOOP r.d[k] = a * b.d[k]; return r;
}
template <class T> inline D3<T> FMA(const D3<T>& a, T b, T c) {
  D3<T> r; r.v = FMA(a.v, b, c); D3_LOOP r.d[k] = a.d[k] * b; return r;
}

template <class S>
struct V3 {
  S x, y, z;
};

template <class A, class B> inline auto operator+(const V3<A>& a, const V3<B>& b) {
  return V3<decltype(a.x + b.x)>{a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class A, class B> inline auto operator-(const V3<A>& a, const V3<B>& b) {
  return V3<decltype(a.x - b.x)>{a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class S> inline V3<S> operator-(const V3<S>& a) { return V3<S>{-a.x, -a.y, -a.z}; }

// dot(a, b) = a.z*b.z + (a.y*b.y + a.x*b.x), fused (the spec's order)
template <class A, class B> inline auto dot3(const V3<A>& a, const V3<B>& b) {
  return FMA(a.z, b.z, FMA(a.y, b.y, a.x * b.x));
}
template <class A, class B> inline auto cross3(const V3<A>& a, const V3<B>& b) {
  using R = decltype(a.x * b.x);
  return V3<R>{FMA(a.y, b.z, -(a.z * b.y)), FMA(a.z, b.x, -(a.x * b.z)), FMA(a.x, b.y, -(a.y * b.x))};
}
// StaticArrays normalize: v / norm(v) with norm = sqrt(dot(v, v))
template <class S> inline V3<S> normalize3(const V3<S>& v) {
  S n = SQRT(dot3(v, v));
  return V3<S>{v.x / n, v.y / n, v.z / n};
}
template <class T> inline V3<T> vval(const V3<D3<T>>& a) { return V3<T>{a.x.v, a.y.v, a.z.v}; }

template <class T> inline V3<T> ld3(const float* p) { return V3<T>{T(p[0]), T(p[1]), T(p[2])}; }

// ------------------------------------------------------------------------------------------
// ray / triangle state in working precision
// ------------------------------------------------------------------------------------------

template <class T>
struct RayT {
  V3<T> pos, dp[3], dir, dd[3];  // dp[k], dd[k]: tangents wrt (lambda, x, y)
  bool in_medium;
  int64_t ignore_tri, dest;
  T lambda, x, y;
  uint8_t status;
};

template <class T> RayT<T> load_ray(const splat_adray_t& a) {
  RayT<T> r;
  r.pos = ld3<T>(a.pos); r.dp[0] = ld3<T>(a.dpdl); r.dp[1] = ld3<T>(a.dpdx); r.dp[2] = ld3<T>(a.dpdy);
  r.dir = ld3<T>(a.dir); r.dd[0] = ld3<T>(a.dddl); r.dd[1] = ld3<T>(a.dddx); r.dd[2] = ld3<T>(a.dddy);
  r.in_medium = a.in_medium != 0; r.ignore_tri = a.ignore_tri; r.dest = a.dest;
  r.lambda = T(a.lambda); r.x = T(a.x); r.y = T(a.y); r.status = a.status;
  return r;
}

inline void st3(float* p, const V3<float>& v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
inline void store_ray(splat_adray_t& a, const RayT<float>& r) {
  st3(a.pos, r.pos); st3(a.dpdl, r.dp[0]); st3(a.dpdx, r.dp[1]); st3(a.dpdy, r.dp[2]);
  st3(a.dir, r.dir); st3(a.dddl, r.dd[0]); st3(a.dddx, r.dd[1]); st3(a.dddy, r.dd[2]);
  a.in_medium = r.in_medium ? 1 : 0; a.ignore_tri = r.ignore_tri; a.dest = r.dest;
  a.lambda = r.lambda; a.x = r.x; a.y = r.y; a.status = r.status;
}

}  // namespace

// FP64 ray record handed back to the tests (numpy structured dtype ADRAY64 in oracle.py).
struct oracle_ray64 {
  double pos[3], dpdl[3], dpdx[3], dpdy[3], dir[3], dddl[3], dddx[3], dddy[3];
  int64_t ignore_tri, dest;
  double lambda, x, y;
  int32_t in_medium, status;
};
static_assert(sizeof(oracle_ray64) == 8 * 24 + 16 + 24 + 8, "oracle_ray64 layout");

namespace {

inline void st3d(double* p, const V3<double>& v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
inline void store_ray64(oracle_ray64& a, const RayT<double>& r) {
  st3d(a.pos, r.pos); st3d(a.dpdl, r.dp[0]); st3d(a.dpdx, r.dp[1]); st3d(a.dpdy, r.dp[2]);
  st3d(a.dir, r.dir); st3d(a.dddl, r.dd[0]); st3d(a.dddx, r.dd[1]); st3d(a.dddy, r.dd[2]);
  a.ignore_tri = r.ignore_tri; a.dest = r.dest; a.lambda = r.lambda; a.x = r.x; a.y = r.y;
  a.in_medium = r.in_medium ? 1 : 0; a.status = r.status;
}

template <class T>
struct FTriT {
  V3<T> v[10];  // n, a, b, c, n_a, n_b, n_c, uv_a, uv_b, uv_c
};
template <class T> FTriT<T> load_ftri(const splat_ftri_t& f) {
  FTriT<T> t;
  for (int i = 0; i < 10; ++i) t.v[i] = ld3<T>(f.v[i]);
  return t;
}

// ------------------------------------------------------------------------------------------
// hit test: distance_to_plane (src/geo.jl:433-443) + in_triangle (src/geo.jl:446-470) inside
// get_hit (src/tracer.jl:45-59). The per-triangle terms of Blackpawn's barycentric test are
// hoisted: u = ((d11*v0 - d01*v1)/den) . (p - a), v = ((d00*v1 - d01*v0)/den) . (p - a), which
// is the reference's u = (d11*d02 - d01*d12)*invDenom and v = (d00*d12 - d01*d02)*invDenom with
// the dot products regrouped. Hit iff u >= 0, v >= 0, u + v < 1, t > 0, i != ignore_tri.
// ------------------------------------------------------------------------------------------

template <class T>
struct TriPrep {
  V3<T> n;
  T nadot;
  V3<T> a, Uu, Vv;
};

template <class T> TriPrep<T> tri_prep(const V3<T>& n, const V3<T>& a, const V3<T>& b, const V3<T>& c) {
  TriPrep<T> P;
  V3<T> v0 = c - a, v1 = b - a;
  T d00 = dot3(v0, v0), d01 = dot3(v0, v1), d11 = dot3(v1, v1);
  T den = FMA(d00, d11, -(d01 * d01));
  T inv = T(1) / den;
  P.Uu = V3<T>{FMA(d11, v0.x, -(d01 * v1.x)) * inv, FMA(d11, v0.y, -(d01 * v1.y)) * inv,
               FMA(d11, v0.z, -(d01 * v1.z)) * inv};
  P.Vv = V3<T>{FMA(d00, v1.x, -(d01 * v0.x)) * inv, FMA(d00, v1.y, -(d01 * v0.y)) * inv,
               FMA(d00, v1.z, -(d01 * v0.z)) * inv};
  P.n = n;
  P.a = a;
  P.nadot = dot3(n, a);
  return P;
}

template <class T>
inline bool hit_test(const TriPrep<T>& P, const V3<T>& o, const V3<T>& d, int64_t ignore, int64_t idx, T* t_out) {
  T no = dot3(P.n, o), nd = dot3(P.n, d);
  T t = (P.nadot - no) / nd;
  V3<T> p{FMA(d.x, t, o.x), FMA(d.y, t, o.y), FMA(d.z, t, o.z)};
  V3<T> v2 = p - P.a;
  T u = dot3(P.Uu, v2), v = dot3(P.Vv, v2);
  *t_out = t;
  return (u >= T(0)) && (v >= T(0)) && (u + v < T(1)) && (t > T(0)) && (ignore != idx);
}

// next_hit! semantics (StableHitter, src/hitters.jl:24-30 with hit_argmin :16-22): lexicographic
// minimum of (t, index) over triangles 2..N (index 1 is the degenerate miss sentinel, skipped as
// in EH1-EH6, e.g. src/hitters.jl:55), ties to the lowest index, (Inf, 1) on a miss.
template <class T>
void next_hit_bf(const splat_adray_t* rays, int64_t R, const std::vector<TriPrep<T>>& preps, int32_t* hit_idx, T* t_out) {
  const int32_t N = (int32_t)preps.size();
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t r = 0; r < R; ++r) {
    V3<T> o = ld3<T>(rays[r].pos), d = ld3<T>(rays[r].dir);
    int64_t ig = rays[r].ignore_tri;
    T best = std::numeric_limits<T>::infinity();
    int32_t bi = 1;
    for (int32_t j = 1; j < N; ++j) {  // array slot j holds index j+1
      T t;
      if (hit_test(preps[j], o, d, ig, (int64_t)j + 1, &t) && t < best) { best = t; bi = j + 1; }
    }
    hit_idx[r] = bi;
    if (t_out) t_out[r] = best;
  }
}

template <class T> std::vector<TriPrep<T>> prep_ntris(const splat_ntri_t* tris, int32_t N) {
  std::vector<TriPrep<T>> P(N);
  for (int32_t j = 0; j < N; ++j)
    P[j] = tri_prep(ld3<T>(tris[j].v[0]), ld3<T>(tris[j].v[1]), ld3<T>(tris[j].v[2]), ld3<T>(tris[j].v[3]));
  return P;
}
template <class T> std::vector<TriPrep<T>> prep_ftris(const splat_ftri_t* tris, int32_t N) {
  std::vector<TriPrep<T>> P(N);
  for (int32_t j = 0; j < N; ++j)
    P[j] = tri_prep(ld3<T>(tris[j].v[0]), ld3<T>(tris[j].v[1]), ld3<T>(tris[j].v[2]), ld3<T>(tris[j].v[3]));
  return P;
}

// One-level BV culling (the DPBVHitter structure, src/hitters.jl:486-648) used only to make the
// CPU baseline fast: member boxes are padded so the cull is conservative and the result is the
// brute-force one above (checked in tests).
struct OracleBV {
  std::vector<double> lo, hi;        // 3 per BV
  std::vector<int64_t> start, count;  // into members
  std::vector<int32_t> members;       // 0-based array slots
};

OracleBV build_oracle_bv(const splat_ftri_t* tris, int32_t n_bv, const int64_t* counts, const int64_t* bv_tris, int64_t stride) {
  OracleBV B;
  B.lo.resize(3 * n_bv); B.hi.resize(3 * n_bv); B.start.resize(n_bv); B.count.resize(n_bv);
  for (int32_t b = 0; b < n_bv; ++b) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    B.start[b] = (int64_t)B.members.size();
    for (int64_t m = 0; m < counts[b]; ++m) {
      int64_t idx = bv_tris[b + m * (int64_t)n_bv];  // Julia column-major (n_bv x stride)
      if (idx <= 1) continue;                        // the degenerate sentinel never hits
      B.members.push_back((int32_t)(idx - 1));
      for (int v = 1; v <= 3; ++v)
        for (int k = 0; k < 3; ++k) {
          lo[k] = std::min(lo[k], (double)tris[idx - 1].v[v][k]);
          hi[k] = std::max(hi[k], (double)tris[idx - 1].v[v][k]);
        }
    }
    B.count[b] = (int64_t)B.members.size() - B.start[b];
    double ext = 0, mag = 0;
    for (int k = 0; k < 3; ++k) { ext = std::max(ext, hi[k] - lo[k]); mag = std::max(mag, std::max(std::fabs(lo[k]), std::fabs(hi[k]))); }
    double pad = 1e-3 * ext + 1e-5 * mag + 1e-6;
    for (int k = 0; k < 3; ++k) { B.lo[3 * b + k] = lo[k] - pad; B.hi[3 * b + k] = hi[k] + pad; }
  }
  (void)stride;
  return B;
}

template <class T>
void next_hit_bv(const splat_adray_t* rays, int64_t R, const std::vector<TriPrep<T>>& preps, const OracleBV& B, int32_t* hit_idx) {
  const int32_t nbv = (int32_t)B.count.size();
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t r = 0; r < R; ++r) {
    V3<T> o = ld3<T>(rays[r].pos), d = ld3<T>(rays[r].dir);
    int64_t ig = rays[r].ignore_tri;
    double od[3] = {rays[r].pos[0], rays[r].pos[1], rays[r].pos[2]};
    double inv[3];
    for (int k = 0; k < 3; ++k) {
      double dk = rays[r].dir[k];
      if (std::fabs(dk) < 1e-30) dk = dk < 0 ? -1e-30 : 1e-30;
      inv[k] = 1.0 / dk;
    }
    T best = std::numeric_limits<T>::infinity();
    int32_t bi = 1;
    for (int32_t b = 0; b < nbv; ++b) {
      double tmin = 0, tmax = INFINITY;
      for (int k = 0; k < 3; ++k) {
        double t1 = (B.lo[3 * b + k] - od[k]) * inv[k], t2 = (B.hi[3 * b + k] - od[k]) * inv[k];
        tmin = std::max(tmin, std::min(t1, t2));
        tmax = std::min(tmax, std::max(t1, t2));
      }
      if (!(tmin <= tmax) || tmin > (double)best * (1 + 1e-6)) continue;
      for (int64_t m = 0; m < B.count[b]; ++m) {
        int32_t j = B.members[B.start[b] + m];
        T t;
        if (hit_test(preps[j], o, d, ig, (int64_t)j + 1, &t) && (t < best || (t == best && j + 1 < bi))) { best = t; bi = j + 1; }
      }
    }
    hit_idx[r] = bi;
  }
}

// ------------------------------------------------------------------------------------------
// optics (src/geo.jl:194-258, :481-561; src/material.jl:1-13)
// ------------------------------------------------------------------------------------------

// reverse_uv (src/geo.jl:232-258): barycentric weights (u of a, v of b) of P in triangle t
template <class T, class S>
inline void reverse_uv(const FTriT<T>& t, const V3<S>& P, S* u_out, S* v_out) {
  const V3<T>& a = t.v[1];
  const V3<T>& b = t.v[2];
  const V3<T>& c = t.v[3];
  V3<T> v0v1 = b - a, v0v2 = c - a;
  V3<T> N = cross3(v0v1, v0v2);
  T denom = dot3(N, N);
  V3<T> edge1 = c - b;
  V3<S> vp1 = P - b;
  S u = dot3(N, cross3(edge1, vp1)) / denom;
  V3<T> edge2 = a - c;
  V3<S> vp2 = P - c;
  S v = dot3(N, cross3(edge2, vp2)) / denom;
  *u_out = u;
  *v_out = v;
}

// optical_normal(::FTri, pos) (src/geo.jl:202-212): normalize(n_a*u + n_b*v + n_c*(1-u-v))
template <class T, class S>
inline V3<S> optical_normal(const FTriT<T>& t, const V3<S>& P) {
  S u, v;
  reverse_uv(t, P, &u, &v);
  S w = (T(1) - u) - v;
  const V3<T>& na = t.v[4];
  const V3<T>& nb = t.v[5];
  const V3<T>& nc = t.v[6];
  V3<S> raw{FMA(nc.x, w, FMA(nb.x, v, na.x * u)), FMA(nc.y, w, FMA(nb.y, v, na.y * u)),
            FMA(nc.z, w, FMA(nb.z, v, na.z * u))};
  return normalize3(raw);
}

// reflect (src/geo.jl:481-486): v - 2 (n.v) n with n normalised
template <class S> inline V3<S> reflect(const V3<S>& v, const V3<S>& normal) {
  V3<S> nh = normalize3(normal);
  S s = dot3(nh, v);
  S s2 = s + s;
  return V3<S>{FMA(-s2, nh.x, v.x), FMA(-s2, nh.y, v.y), FMA(-s2, nh.z, v.z)};
}

// refract (src/geo.jl:548-561), as written -- including its sign: with c1 = n.v/|v| <= 0 it
// returns normalize(eta v + n (eta c1 - c2)), which is not Snell's law (see the closed-form test).
template <class T, class S, class E>
inline V3<S> refract(const V3<S>& v, const V3<S>& normal, const E& n1, const E& n2) {
  V3<S> nh = normalize3(normal);
  if (val(dot3(v, nh)) > T(0)) nh = -nh;
  S c1 = dot3(normalize3(v), nh);
  E eta = n1 / n2;
  S s1 = SQRT(FMA(-c1, c1, T(1)));
  S s2 = eta * s1;
  S c2 = SQRT(FMA(-s2, s2, T(1)));
  S k = FMA(eta, c1, -c2);
  V3<S> out{FMA(nh.x, k, v.x * eta), FMA(nh.y, k, v.y * eta), FMA(nh.z, k, v.z * eta)};
  return normalize3(out);
}

// can_refract (src/geo.jl:528-542). In Float32, 1 - 1f-9 rounds to exactly 1, so both tests are
// against 1; the FP64 form keeps that FP32 meaning.
template <class T> inline bool can_refract(const V3<T>& v, const V3<T>& normal, T n1, T n2) {
  V3<T> nh = normalize3(normal);
  if (dot3(nh, v) > T(0)) nh = -nh;
  T c1 = dot3(normalize3(v), nh);
  if (c1 * c1 > T(1)) return false;
  T s1 = SQRT(FMA(-c1, c1, T(1)));
  T s2 = (n1 / n2) * s1;
  return std::fabs(s2) < T(1);
}

// reflectance (src/geo.jl:488-502): the reference's s/p "average"
template <class T> inline T reflectance(const V3<T>& v, const V3<T>& normal, T n1, T n2) {
  V3<T> nh = normalize3(normal);
  if (dot3(nh, v) < T(0)) nh = -nh;
  T c1 = dot3(normalize3(v), nh);
  T s1 = SQRT(FMA(-c1, c1, T(1)));
  T e = n1 / n2;
  T s2 = SQRT(FMA(-(e * e), s1 * s1, T(1)));
  T h = (n1 + n2) * T(0.5f);
  T num = FMA(h, c1, -(h * s2));
  T den = FMA(n1, c1, n2 * s2);
  T q = num / den;
  return q * q;
}

// glass(lambda) as a dual in lambda: the reference's linear IOR (src/material.jl:1-9) or Cauchy
// n = A + B / lambda_um^2 (the configurations' glass).
template <class T> inline D3<T> glass_ior(const splat_optics_t& o, T lambda) {
  D3<T> L{lambda, {T(1), T(0), T(0)}};
  if (o.ior_model == SPLAT_IOR_CAUCHY) {
    D3<T> Lu = L / T(1000.0f);
    D3<T> L2 = Lu * Lu;
    return T(o.cauchy_a) + T(o.cauchy_b) / L2;
  }
  T m = (T(1.4f) - T(1.3f)) / (T(700.0f) - T(400.0f));
  return FMA(m, L - T(400.0f), T(1.3f));
}
template <class T> inline D3<T> air_ior() { return D3<T>{T(1), {T(0), T(0), T(0)}}; }

// get_hit((i, T), ::ADRay) (src/tracer.jl:61-81): the hit test above on the ray's value, then
// t as a dual: t = (n.a - n.p(l,x,y)) / (n.d(l,x,y)) with p, d the first-order expansions
// (src/geo.jl:398-410) -- dt_k = -(n.dp_k + t n.dd_k) / (n.d).
template <class T>
inline bool get_hit_ad(const RayT<T>& r, const FTriT<T>& F, int64_t i, D3<T>* tD) {
  TriPrep<T> P = tri_prep(F.v[0], F.v[1], F.v[2], F.v[3]);
  T tv;
  if (!hit_test(P, r.pos, r.dir, r.ignore_tri, i, &tv)) {
    T inf = std::numeric_limits<T>::infinity();
    *tD = D3<T>{inf, {inf, inf, inf}};
    return false;
  }
  V3<D3<T>> P0{{r.pos.x, {r.dp[0].x, r.dp[1].x, r.dp[2].x}},
               {r.pos.y, {r.dp[0].y, r.dp[1].y, r.dp[2].y}},
               {r.pos.z, {r.dp[0].z, r.dp[1].z, r.dp[2].z}}};
  V3<D3<T>> D0{{r.dir.x, {r.dd[0].x, r.dd[1].x, r.dd[2].x}},
               {r.dir.y, {r.dd[0].y, r.dd[1].y, r.dd[2].y}},
               {r.dir.z, {r.dd[0].z, r.dd[1].z, r.dd[2].z}}};
  D3<T> no = dot3(P.n, P0);
  D3<T> nd = dot3(P.n, D0);
  D3<T> num = P.nadot - no;
  *tD = num / nd;
  return true;
}

template <class T> inline V3<D3<T>> pos_dual(const RayT<T>& r) {
  return V3<D3<T>>{{r.pos.x, {r.dp[0].x, r.dp[1].x, r.dp[2].x}},
                   {r.pos.y, {r.dp[0].y, r.dp[1].y, r.dp[2].y}},
                   {r.pos.z, {r.dp[0].z, r.dp[1].z, r.dp[2].z}}};
}
template <class T> inline V3<D3<T>> dir_dual(const RayT<T>& r) {
  return V3<D3<T>>{{r.dir.x, {r.dd[0].x, r.dd[1].x, r.dd[2].x}},
                   {r.dir.y, {r.dd[0].y, r.dd[1].y, r.dd[2].y}},
                   {r.dir.z, {r.dd[0].z, r.dd[1].z, r.dd[2].z}}};
}
// next_p (src/tracer.jl:109-116) at the expansion point, as a dual: p + d t
template <class T> inline V3<D3<T>> next_p_dual(const V3<D3<T>>& P0, const V3<D3<T>>& D0, const D3<T>& t) {
  return V3<D3<T>>{FMA(D0.x, t, P0.x), FMA(D0.y, t, P0.y), FMA(D0.z, t, P0.z)};
}
template <class T> inline void split(const V3<D3<T>>& a, V3<T>* v, V3<T>* d3) {
  *v = vval(a);
  for (int k = 0; k < 3; ++k) d3[k] = V3<T>{a.x.d[k], a.y.d[k], a.z.d[k]};
}

// evolve_ray + handle_optics (src/tracer.jl:181-201, :118-179), one bounce of one ray.
template <class T>
RayT<T> evolve_one(const RayT<T>& r, int64_t i, const FTriT<T>& F, T rndm, int32_t first_diffuse, const splat_optics_t& opt) {
  if (r.status != SPLAT_RAY_ACTIVE) return r;
  D3<T> tD;
  bool hit = get_hit_ad(r, F, i, &tD);
  int64_t i2 = hit ? i : 1;
  RayT<T> out = r;
  if (i2 >= first_diffuse) { out.status = SPLAT_RAY_DIFFUSE; return out; }
  if (!hit) { out.status = SPLAT_RAY_INFINITY; return out; }

  D3<T> g = glass_ior<T>(opt, r.lambda), a = air_ior<T>();
  D3<T> n1 = r.in_medium ? g : a, n2 = r.in_medium ? a : g;

  V3<D3<T>> P0 = pos_dual(r), D0 = dir_dual(r);
  V3<D3<T>> P = next_p_dual(P0, D0, tD);
  V3<D3<T>> Nd = optical_normal(F, P);
  V3<T> Nv = vval(Nd);
  bool refracts = can_refract(r.dir, Nv, n1.v, n2.v) && (rndm > reflectance(r.dir, Nv, n1.v, n2.v));

  V3<D3<T>> Dout;
  if (refracts) {
    V3<D3<T>> Nr = Nd;
    if (opt.compat & SPLAT_COMPAT_REFRACT_NORMAL_LAMBDA) { Nr.x.d[0] = T(0); Nr.y.d[0] = T(0); Nr.z.d[0] = T(0); }
    Dout = refract<T>(D0, Nr, n1, n2);
  } else {
    Dout = reflect(D0, Nd);
  }
  split(P, &out.pos, out.dp);
  split(Dout, &out.dir, out.dd);

  if ((opt.compat & SPLAT_COMPAT_DY_AT_X) && r.x != r.y) {
    // the reference differentiates the direction wrt y at y = r.x (src/tracer.jl:138-141, :166-169)
    T delta = r.x - r.y;
    T z = T(0);
    V3<D3<T>> Py0{{FMA(r.dp[2].x, delta, r.pos.x), {z, z, r.dp[2].x}},
                  {FMA(r.dp[2].y, delta, r.pos.y), {z, z, r.dp[2].y}},
                  {FMA(r.dp[2].z, delta, r.pos.z), {z, z, r.dp[2].z}}};
    V3<D3<T>> Dy0{{FMA(r.dd[2].x, delta, r.dir.x), {z, z, r.dd[2].x}},
                  {FMA(r.dd[2].y, delta, r.dir.y), {z, z, r.dd[2].y}},
                  {FMA(r.dd[2].z, delta, r.dir.z), {z, z, r.dd[2].z}}};
    D3<T> sy{FMA(tD.d[2], delta, tD.v), {z, z, tD.d[2]}};
    V3<D3<T>> Py = next_p_dual(Py0, Dy0, sy);
    V3<D3<T>> Ny = optical_normal(F, Py);
    V3<D3<T>> Dy = refracts ? refract<T>(Dy0, Ny, n1.v, n2.v) : reflect(Dy0, Ny);
    out.dd[2] = V3<T>{Dy.x.d[2], Dy.y.d[2], Dy.z.d[2]};
  }
  out.in_medium = refracts ? !r.in_medium : r.in_medium;
  out.ignore_tri = i;
  out.status = SPLAT_RAY_ACTIVE;
  return out;
}

// final_evolution (src/ray_imagers.jl:4-37): DIFFUSE rays move to their hit on the diffuse
// triangle (value + position tangents); direction kept. A miss gives NaN, as the reference's
// next_p with t = Inf does.
template <class T> RayT<T> final_evolution_one(const RayT<T>& r, int64_t i, const FTriT<T>& F) {
  if (r.status != SPLAT_RAY_DIFFUSE) return r;
  D3<T> tD;
  bool hit = get_hit_ad(r, F, i, &tD);
  RayT<T> out = r;
  if (hit) {
    V3<D3<T>> P = next_p_dual(pos_dual(r), dir_dual(r), tD);
    split(P, &out.pos, out.dp);
  } else {
    T nan = std::numeric_limits<T>::quiet_NaN();
    out.pos = V3<T>{nan, nan, nan};
    for (int k = 0; k < 3; ++k) out.dp[k] = out.pos;
  }
  out.ignore_tri = hit ? i : 1;
  out.status = SPLAT_RAY_DIFFUSE;
  return out;
}

// piecewise-linear response (src/rgb_spectrum.jl:3-32, normalised as :116-130): on segment m
// (x_m <= l < x_{m+1}, the end segments extended), slope*(l - x_{m+1}) + y_{m+1}
template <class T> inline T curve_eval(const splat_lightmap_params_t& lm, int c, T lam) {
  int n = lm.curve_n[c];
  int m = 0;
  while (m < n - 2 && lam >= T(lm.curve_x[c][m + 1])) ++m;
  T x0 = T(lm.curve_x[c][m]), x1 = T(lm.curve_x[c][m + 1]);
  T y0 = T(lm.curve_y[c][m]), y1 = T(lm.curve_y[c][m + 1]);
  T slope = (y1 - y0) / (x1 - x0);
  return FMA(slope, lam - x1, y1);
}

// atomic_spectrum_kernel2 (src/ray_imagers.jl:78-113) for one ray: every splat sample (k, iy, ix)
// in the reference's thread order (CartesianIndices((dx, dy, lambda))), handed to
// fn(k, lam, U*W, V*H, I) with the texture coordinates scaled to texels (before ceil and clamp)
// and the cosine intensity; samples with a non-finite (u, v) are skipped, as there.
template <class T, class Fn>
void splat_samples(const RayT<T>& r, const FTriT<T>& F, const splat_lightmap_params_t& lm, Fn&& fn) {
  if (r.status != SPLAT_RAY_DIFFUSE) return;
  const int nl = lm.n_lambda > 0 ? lm.n_lambda : 1;
  const T W = T((float)lm.width), H = T((float)lm.height);
  for (int k = 0; k < nl; ++k) {
    T lam = lm.n_lambda > 0 ? T(lm.spectrum[k]) : r.lambda;
    for (int iy = 0; iy < lm.n_delta; ++iy) {
      for (int ix = 0; ix < lm.n_delta; ++ix) {
        T xs = r.x + T(lm.delta[ix]), ys = r.y + T(lm.delta[iy]);
        T dl = lam - r.lambda, dx = xs - r.x, dy = ys - r.y;
        V3<T> p{FMA(r.dp[2].x, dy, FMA(r.dp[1].x, dx, FMA(r.dp[0].x, dl, r.pos.x))),
                FMA(r.dp[2].y, dy, FMA(r.dp[1].y, dx, FMA(r.dp[0].y, dl, r.pos.y))),
                FMA(r.dp[2].z, dy, FMA(r.dp[1].z, dx, FMA(r.dp[0].z, dl, r.pos.z)))};
        V3<T> d{FMA(r.dd[2].x, dy, FMA(r.dd[1].x, dx, FMA(r.dd[0].x, dl, r.dir.x))),
                FMA(r.dd[2].y, dy, FMA(r.dd[1].y, dx, FMA(r.dd[0].y, dl, r.dir.y))),
                FMA(r.dd[2].z, dy, FMA(r.dd[1].z, dx, FMA(r.dd[0].z, dl, r.dir.z)))};
        // tex_uv (src/geo.jl:222-229)
        T bu, bv;
        reverse_uv(F, p, &bu, &bv);
        T w = (T(1) - bu) - bv;
        T U = FMA(F.v[9].x, w, FMA(F.v[8].x, bv, F.v[7].x * bu));
        T V = FMA(F.v[9].y, w, FMA(F.v[8].y, bv, F.v[7].y * bu));
        if (!std::isfinite(U) || !std::isfinite(V)) continue;
        // cosine_shading (src/geo.jl:283-286)
        V3<T> nrm = optical_normal(F, p);
        T I = std::fabs(dot3(nrm, normalize3(d)));
        fn(k, lam, U * W, V * H, I);
      }
    }
  }
}

// the splats of one ray accumulated into tex: texel (clamp(ceil(U W), 1, W), clamp(ceil(V H), 1, H))
template <class T, class A>
void splat_one(const RayT<T>& r, const FTriT<T>& F, const splat_lightmap_params_t& lm, A* tex) {
  const T W = T((float)lm.width), H = T((float)lm.height);
  const int64_t plane = (int64_t)lm.width * lm.height;
  splat_samples(r, F, lm, [&](int k, T lam, T uw, T vh, T I) {
    T fi = std::ceil(uw), fj = std::ceil(vh);
    fi = fi < T(1) ? T(1) : (fi > W ? W : fi);
    fj = fj < T(1) ? T(1) : (fj > H ? H : fj);
    int64_t ii = (int64_t)fi - 1, jj = (int64_t)fj - 1;
    int64_t base = ii + (int64_t)lm.width * jj;
    if (lm.weight_mode == SPLAT_WEIGHTS_IDENTITY) {
#pragma omp atomic
      tex[base + plane * k] += (A)I;
    } else if (lm.weight_mode == SPLAT_WEIGHTS_TABLE) {
      for (int c = 0; c < lm.channels; ++c) {
#pragma omp atomic
        tex[base + plane * c] += (A)(I * T(lm.weights[k * 3 + c]));
      }
    } else {
      for (int c = 0; c < lm.channels; ++c) {
#pragma omp atomic
        tex[base + plane * c] += (A)(I * curve_eval(lm, c, lam));
      }
    }
  });
}

// distance (in texels) from a scaled texture coordinate to the nearest texel edge that changes
// its clamped bin: the integers 1 .. n-1
template <class T> inline T edge_distance(T x, int32_t n) {
  if (n < 2) return std::numeric_limits<T>::infinity();
  T k = std::round(x);
  k = k < T(1) ? T(1) : (k > T(n - 1) ? T(n - 1) : k);
  return std::fabs(x - k);
}

// ------------------------------------------------------------------------------------------
// backward shading: continuum_shade! / shade_tex4_kernel (src/ray_imagers.jl:381-503)
// ------------------------------------------------------------------------------------------
// Bilinear sample of one float plane (width W along u, height H along v, u fastest) at 1-based
// coordinates (u, v) with clamp-to-edge: the reference reads its light map and skydome through
// CUDA.jl textures with linear interpolation, where an integer coordinate i hits texel i exactly
// and a fraction a blends texels floor(u) and floor(u) + 1. The weights are exact FP32 here
// (texture hardware rounds them to 8 fractional bits: a documented, bounded difference).
template <class T> inline T tex_bilinear(const float* plane, int64_t W, int64_t H, T u, T v) {
  T fu = std::floor(u), fv = std::floor(v);
  T a = u - fu, b = v - fv;
  fu = fu < T(0) ? T(0) : (fu > T(W) ? T(W) : fu);
  fv = fv < T(0) ? T(0) : (fv > T(H) ? T(H) : fv);
  int64_t i0 = (int64_t)fu, j0 = (int64_t)fv;
  int64_t i1 = i0 + 1 > W ? W : i0 + 1, j1 = j0 + 1 > H ? H : j0 + 1;
  i0 = i0 < 1 ? 1 : i0;
  j0 = j0 < 1 ? 1 : j0;
  i1 = i1 < 1 ? 1 : i1;
  j1 = j1 < 1 ? 1 : j1;
  T t00 = T(plane[(i0 - 1) + W * (j0 - 1)]), t10 = T(plane[(i1 - 1) + W * (j0 - 1)]);
  T t01 = T(plane[(i0 - 1) + W * (j1 - 1)]), t11 = T(plane[(i1 - 1) + W * (j1 - 1)]);
  T r0 = FMA(a, t10 - t00, t00), r1 = FMA(a, t11 - t01, t01);
  return FMA(b, r1 - r0, r0);
}

// One (ray, dx, dy) sample of shade_tex4_kernel (src/ray_imagers.jl:381-436) for a final-evolved
// ray r on triangle F: an INFINITY ray reads the skydome at its direction's polar angles (u from
// the polar angle, v from the azimuth), a DIFFUSE ray sums retina-weighted light-map values at
// its texture coordinate over the spectrum; channels clamped to [0, 1].
template <class T>
void shade_sample(const RayT<T>& r, const FTriT<T>& F, const splat_lightmap_params_t& lm, const float* tex,
                  const float* sky, int64_t SW, int64_t SH, int a, int b, T rgb[3]) {
  rgb[0] = rgb[1] = rgb[2] = T(0);
  const T pi = T((float)3.14159265358979323846);
  if (r.status == SPLAT_RAY_INFINITY && sky) {
    T xs = r.x + T(lm.delta[a]), ys = r.y + T(lm.delta[b]);
    T dx = xs - r.x, dy = ys - r.y;  // expand(adr, adr.lambda, x, y): no lambda step
    V3<T> d{FMA(r.dd[2].x, dy, FMA(r.dd[1].x, dx, r.dir.x)), FMA(r.dd[2].y, dy, FMA(r.dd[1].y, dx, r.dir.y)),
            FMA(r.dd[2].z, dy, FMA(r.dd[1].z, dx, r.dir.z))};
    T th = std::acos(d.z), ph = std::atan2(d.y, d.x);
    T v = ((ph / pi) + T(1)) * T(0.5), u = th / pi;
    u = u * T((float)SW);
    v = v * T((float)SH);
    if (std::isfinite(u) && std::isfinite(v))
      for (int c = 0; c < 3; ++c) rgb[c] = tex_bilinear<T>(sky + SW * SH * c, SW, SH, u, v);
  } else if (r.status == SPLAT_RAY_DIFFUSE) {
    const T W = T((float)lm.width), H = T((float)lm.height);
    for (int k = 0; k < lm.n_lambda; ++k) {
      T lam = T(lm.spectrum[k]);
      T xs = r.x + T(lm.delta[a]), ys = r.y + T(lm.delta[b]);
      T dl = lam - r.lambda, dx = xs - r.x, dy = ys - r.y;
      V3<T> p{FMA(r.dp[2].x, dy, FMA(r.dp[1].x, dx, FMA(r.dp[0].x, dl, r.pos.x))),
              FMA(r.dp[2].y, dy, FMA(r.dp[1].y, dx, FMA(r.dp[0].y, dl, r.pos.y))),
              FMA(r.dp[2].z, dy, FMA(r.dp[1].z, dx, FMA(r.dp[0].z, dl, r.pos.z)))};
      T bu, bv;
      reverse_uv(F, p, &bu, &bv);
      T w = (T(1) - bu) - bv;
      T U = FMA(F.v[9].x, w, FMA(F.v[8].x, bv, F.v[7].x * bu));
      T V = FMA(F.v[9].y, w, FMA(F.v[8].y, bv, F.v[7].y * bu));
      if (std::isnan(U) || std::isnan(V)) continue;
      T X = tex_bilinear<T>(tex + (int64_t)lm.width * lm.height * k, lm.width, lm.height, U * W, V * H);
      for (int c = 0; c < 3; ++c) rgb[c] = FMA(T(lm.weights[k * 3 + c]), X, rgb[c]);
    }
  }
  for (int c = 0; c < 3; ++c) rgb[c] = rgb[c] > T(1) ? T(1) : (rgb[c] < T(0) ? T(0) : rgb[c]);
}

// continuum_shade! (src/ray_imagers.jl:440-503) over rays in pixel order (ray q is pixel
// (q % h + 1, q / h + 1) of an h-row grid): final_evolution in place, then each ray's n_delta^2
// samples into the (h * n_delta)-row RGB image, interleaved, row index fastest.
template <class T, class O>
void shade_frame(splat_adray_t* rays, int64_t R, const int32_t* hit_idx, const std::vector<FTriT<T>>& F, int32_t N,
                 const splat_lightmap_params_t& lm, const float* tex, const float* sky, int64_t SW, int64_t SH,
                 int64_t h, O* rgb, bool store_rays) {
  const int nd = lm.n_delta;
  const int64_t img_h = h * nd;
#pragma omp parallel for schedule(static, 1024)
  for (int64_t q = 0; q < R; ++q) {
    int32_t i = hit_idx[q];
    if (i < 1 || i > N) i = 1;
    RayT<T> fe = final_evolution_one<T>(load_ray<T>(rays[q]), i, F[i - 1]);
    if (store_rays) {
      if constexpr (std::is_same<T, float>::value) store_ray(rays[q], fe);
    }
    const int64_t x = q % h, y = q / h;
    for (int b = 0; b < nd; ++b)
      for (int a = 0; a < nd; ++a) {
        T c3[3];
        shade_sample<T>(fe, F[i - 1], lm, tex, sky, SW, SH, a, b, c3);
        const int64_t X = a + nd * x, Y = b + nd * y;
        for (int c = 0; c < 3; ++c) rgb[(X + img_h * Y) * 3 + c] = (O)c3[c];
      }
  }
}

// ------------------------------------------------------------------------------------------
// seeded random numbers (the reference's CUDA.rand is unseeded, src/tracer.jl:219,
// src/ray_generators.jl:57): splitmix64's finaliser over (seed, stream, counter)
// ------------------------------------------------------------------------------------------
inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline float u01(uint64_t seed, uint64_t stream, uint64_t counter) {
  uint64_t h = mix64(mix64(seed ^ mix64(stream)) ^ counter);
  return (float)(h >> 40) * 0x1p-24f;
}
constexpr uint64_t RNG_RNDM = 1ull << 32, RNG_LAMBDA = 2ull << 32, RNG_JITTER = 3ull << 32;

template <class T>
inline T bounce_random(float force_rand, const float* rndm, int64_t r, uint64_t seed, int32_t bounce, int64_t dest) {
  if (!std::isnan(force_rand)) return T(force_rand);
  if (rndm) return T(rndm[r]);
  return T(u01(seed, RNG_RNDM | (uint32_t)bounce, (uint64_t)dest));
}

// camera_ray (src/ray_generators.jl:18-50) at a pinhole. The reference evaluates it in Float64
// (height * _COS_45 is Float64) and rounds the ADRay fields to Float32; so does this.
void raygen_one(const splat_pinhole_t& cam, int64_t q, splat_adray_t* out) {
  const int64_t H = cam.height, Wd = cam.width, L = cam.samples_per_pixel;
  int64_t P = q / L, l = q % L;
  int64_t x, y;  // 1-based row / column
  if (cam.order == SPLAT_ORDER_TILED8 && H % 8 == 0 && Wd % 8 == 0) {
    int64_t tile = P / 64, w = P % 64, tx = tile % (H / 8), ty = tile / (H / 8);
    x = tx * 8 + (w % 8) + 1;
    y = ty * 8 + (w / 8) + 1;
  } else {
    x = P % H + 1;
    y = P / H + 1;
  }
  int64_t dest = ((y - 1) * H + (x - 1)) * L + l + 1;
  float lam = cam.lambda_min;
  float span = cam.lambda_max - cam.lambda_min;
  if (cam.lambda_mode == SPLAT_LAMBDA_UNIFORM) {
    lam = std::fmaf(u01(cam.seed, RNG_LAMBDA, (uint64_t)dest), span, cam.lambda_min);
  } else if (cam.lambda_mode == SPLAT_LAMBDA_STRATIFIED) {
    float u = u01(cam.seed, RNG_LAMBDA, (uint64_t)dest);
    lam = std::fmaf(((float)l + u) / (float)L, span, cam.lambda_min);
  }
  float jit[3] = {0.f, 0.f, 0.f};
  if (cam.jitter != 0.f) {
    float mx = (float)std::max(H, Wd);
    for (int k = 0; k < 3; ++k) jit[k] = (u01(cam.seed, RNG_JITTER | (uint64_t)k, (uint64_t)dest) * cam.jitter) / mx;
  }
  double scale = ((double)H * (1.0 / std::sqrt(2.0))) / (double)cam.fov_half_sin;
  D3<double> X{(double)x, {0, 1, 0}}, Y{(double)y, {0, 0, 1}};
  D3<double> xd = (X - (double)H / 2.0) / scale;
  D3<double> yd = (Y - (double)Wd / 2.0) / scale;
  D3<double> zd = SQRT((1.0 - xd * xd) - yd * yd);
  D3<double> dd[3];
  for (int k = 0; k < 3; ++k)
    dd[k] = ((xd * (double)cam.right[k] + yd * (double)cam.up[k]) + zd * (double)cam.dir[k]) + (double)jit[k];
  std::memset(out, 0, sizeof(*out));
  for (int k = 0; k < 3; ++k) {
    out->pos[k] = cam.pos[k];
    out->dir[k] = (float)dd[k].v;
    out->dddx[k] = (float)dd[k].d[1];
    out->dddy[k] = (float)dd[k].d[2];
  }
  out->in_medium = 0;
  out->ignore_tri = 1;
  out->dest = dest;
  out->lambda = lam;
  out->x = 0.f;
  out->y = 0.f;
  out->status = SPLAT_RAY_ACTIVE;
}

// stable partition: ACTIVE first (partition!(rays, swap; by = status != ACTIVE),
// src/tracer.jl:263, src/partition.jl:122-132), both halves in original order.
int64_t stable_partition(splat_adray_t* rays, int64_t R, int32_t* carry) {
  std::vector<splat_adray_t> tmp(rays, rays + R);
  std::vector<int32_t> ctmp;
  if (carry) ctmp.assign(carry, carry + R);
  int64_t w = 0;
  for (int64_t r = 0; r < R; ++r)
    if (tmp[r].status == SPLAT_RAY_ACTIVE) { rays[w] = tmp[r]; if (carry) carry[w] = ctmp[r]; ++w; }
  int64_t count = w;
  for (int64_t r = 0; r < R; ++r)
    if (tmp[r].status != SPLAT_RAY_ACTIVE) { rays[w] = tmp[r]; if (carry) carry[w] = ctmp[r]; ++w; }
  return count;
}

template <class T> std::vector<FTriT<T>> load_ftris(const splat_ftri_t* tris, int32_t N) {
  std::vector<FTriT<T>> F(N);
  for (int32_t j = 0; j < N; ++j) F[j] = load_ftri<T>(tris[j]);
  return F;
}

void set_threads(int32_t threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
}

struct HitCtx {
  std::vector<TriPrep<float>> preps;
  bool use_bv = false;
  OracleBV bv;
};

void hits_f32(const HitCtx& H, const splat_adray_t* rays, int64_t R, int32_t* hit_idx) {
  if (H.use_bv) next_hit_bv<float>(rays, R, H.preps, H.bv, hit_idx);
  else next_hit_bf<float>(rays, R, H.preps, hit_idx, nullptr);
}

// run_evolution!(hitter, ::ExperimentalTracer; reorder = false) (src/tracer.jl:233-282)
int64_t run_evolution_f32(const HitCtx& H, const std::vector<FTriT<float>>& F, splat_adray_t* rays, int64_t R,
                          int32_t* hit_idx, const splat_run_params_t& p) {
  int64_t cap = R;
  for (int32_t iter = 1; iter <= p.depth; ++iter) {
    if (cap == 0) break;
    hits_f32(H, rays, cap, hit_idx);
#pragma omp parallel for schedule(static, 1024)
    for (int64_t r = 0; r < cap; ++r) {
      RayT<float> ray = load_ray<float>(rays[r]);
      int32_t i = hit_idx[r];
      float u = bounce_random<float>(p.force_rand, nullptr, r, p.seed, iter, ray.dest);
      RayT<float> o = evolve_one<float>(ray, i, F[i - 1], u, p.first_diffuse, p.optics);
      store_ray(rays[r], o);
    }
    cap = stable_partition(rays, R, nullptr);
  }
  hits_f32(H, rays, R, hit_idx);  // "needed to realign hit_idx" (src/tracer.jl:279-280)
  return cap;
}

}  // namespace

// ============================================================================================
// extern "C" surface for the tests (ctypes, oracle/oracle.py)
// ============================================================================================
extern "C" {

uint64_t oracle_mix64(uint64_t z) { return mix64(z); }
float oracle_u01(uint64_t seed, uint64_t stream, uint64_t counter) { return u01(seed, stream, counter); }

// monotonic_reinterpret / retreive_arg (src/atomic_argmin.jl:5-30)
uint32_t oracle_key32(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u > (1u << 31) ? ~u : (u | (1u << 31));
}
uint64_t oracle_key64(float f, uint32_t idx) { return ((uint64_t)oracle_key32(f) << 32) | idx; }
float oracle_key32_inv(uint32_t k) {
  uint32_t u = (k >> 31) == 0 ? ~k : (k ^ (1u << 31));
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
uint32_t oracle_key64_arg(uint64_t k) { return (uint32_t)(k ^ ((k >> 32) << 32)); }

// distance_to_plane (src/geo.jl:433-435) as the hit test computes it: (n.a - n.o) / (n.d)
double oracle_distance_to_plane_f64(const double* o, const double* d, const double* a, const double* n) {
  V3<double> O{o[0], o[1], o[2]}, Dv{d[0], d[1], d[2]}, A{a[0], a[1], a[2]}, Nn{n[0], n[1], n[2]};
  return (dot3(Nn, A) - dot3(Nn, O)) / dot3(Nn, Dv);
}
float oracle_distance_to_plane_f32(const float* o, const float* d, const float* a, const float* n) {
  V3<float> O{o[0], o[1], o[2]}, Dv{d[0], d[1], d[2]}, A{a[0], a[1], a[2]}, Nn{n[0], n[1], n[2]};
  return (dot3(Nn, A) - dot3(Nn, O)) / dot3(Nn, Dv);
}

// hit test of one ray against one Tri [n, a, b, c] (12 floats); returns hit, writes t
int32_t oracle_hit_test_f32(const float* o, const float* d, const float* tri12, int64_t ignore, int64_t idx, float* t) {
  TriPrep<float> P = tri_prep(ld3<float>(tri12), ld3<float>(tri12 + 3), ld3<float>(tri12 + 6), ld3<float>(tri12 + 9));
  return hit_test(P, ld3<float>(o), ld3<float>(d), ignore, idx, t) ? 1 : 0;
}
int32_t oracle_hit_test_f64(const double* o, const double* d, const double* tri12, int64_t ignore, int64_t idx, double* t) {
  auto L = [](const double* p) { return V3<double>{p[0], p[1], p[2]}; };
  TriPrep<double> P = tri_prep(L(tri12), L(tri12 + 3), L(tri12 + 6), L(tri12 + 9));
  return hit_test(P, L(o), L(d), ignore, idx, t) ? 1 : 0;
}

void oracle_next_hit_f32(const splat_adray_t* rays, int64_t R, const splat_ntri_t* tris, int32_t N, int32_t* hit_idx, float* t_out, int32_t threads) {
  set_threads(threads);
  next_hit_bf<float>(rays, R, prep_ntris<float>(tris, N), hit_idx, t_out);
}
void oracle_next_hit_f64(const splat_adray_t* rays, int64_t R, const splat_ntri_t* tris, int32_t N, int32_t* hit_idx, double* t_out, int32_t threads) {
  set_threads(threads);
  next_hit_bf<double>(rays, R, prep_ntris<double>(tris, N), hit_idx, t_out);
}
void oracle_next_hit_bv_f32(const splat_adray_t* rays, int64_t R, const splat_ftri_t* tris, int32_t N, int32_t n_bv,
                            const int64_t* counts, const int64_t* bv_tris, int64_t stride, int32_t* hit_idx, int32_t threads) {
  set_threads(threads);
  OracleBV B = build_oracle_bv(tris, n_bv, counts, bv_tris, stride);
  next_hit_bv<float>(rays, R, prep_ftris<float>(tris, N), B, hit_idx);
}

// optics primitives in FP64 (closed-form tests)
void oracle_reflect_f64(const double* v, const double* n, double* out) {
  V3<double> r = reflect(V3<double>{v[0], v[1], v[2]}, V3<double>{n[0], n[1], n[2]});
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void oracle_reflect_f32(const float* v, const float* n, float* out) {
  V3<float> r = reflect(ld3<float>(v), ld3<float>(n));
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void oracle_refract_f64(const double* v, const double* n, double n1, double n2, double* out) {
  V3<double> r = refract<double>(V3<double>{v[0], v[1], v[2]}, V3<double>{n[0], n[1], n[2]}, n1, n2);
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
int32_t oracle_can_refract_f64(const double* v, const double* n, double n1, double n2) {
  return can_refract(V3<double>{v[0], v[1], v[2]}, V3<double>{n[0], n[1], n[2]}, n1, n2) ? 1 : 0;
}
double oracle_reflectance_f64(const double* v, const double* n, double n1, double n2) {
  return reflectance(V3<double>{v[0], v[1], v[2]}, V3<double>{n[0], n[1], n[2]}, n1, n2);
}
void oracle_optical_normal_f64(const splat_ftri_t* tri, const double* p, double* out) {
  FTriT<double> F = load_ftri<double>(*tri);
  V3<double> r = optical_normal(F, V3<double>{p[0], p[1], p[2]});
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void oracle_reverse_uv_f64(const splat_ftri_t* tri, const double* p, double* uv) {
  FTriT<double> F = load_ftri<double>(*tri);
  double u, v;
  reverse_uv(F, V3<double>{p[0], p[1], p[2]}, &u, &v);
  uv[0] = u; uv[1] = v;
}
double oracle_glass_ior_f64(const splat_optics_t* o, double lambda, double* dn) {
  D3<double> g = glass_ior<double>(*o, lambda);
  if (dn) *dn = g.d[0];
  return g.v;
}

// evolve over R rays (rays in/out for f32; FP64 writes oracle_ray64 records)
void oracle_evolve_f32(splat_adray_t* rays, int64_t R, const int32_t* hit_idx, const splat_ftri_t* tris, int32_t N,
                       const float* rndm, float force_rand, uint64_t seed, int32_t bounce, int32_t first_diffuse,
                       const splat_optics_t* optics, int32_t threads) {
  set_threads(threads);
  std::vector<FTriT<float>> F = load_ftris<float>(tris, N);
#pragma omp parallel for schedule(static, 1024)
  for (int64_t r = 0; r < R; ++r) {
    RayT<float> ray = load_ray<float>(rays[r]);
    int32_t i = hit_idx[r];
    float u = bounce_random<float>(force_rand, rndm, r, seed, bounce, ray.dest);
    store_ray(rays[r], evolve_one<float>(ray, i, F[i - 1], u, first_diffuse, *optics));
  }
}
void oracle_evolve_f64(const splat_adray_t* rays, int64_t R, const int32_t* hit_idx, const splat_ftri_t* tris, int32_t N,
                       const float* rndm, float force_rand, uint64_t seed, int32_t bounce, int32_t first_diffuse,
                       const splat_optics_t* optics, oracle_ray64* out, int32_t threads) {
  set_threads(threads);
  std::vector<FTriT<double>> F = load_ftris<double>(tris, N);
#pragma omp parallel for schedule(static, 1024)
  for (int64_t r = 0; r < R; ++r) {
    RayT<double> ray = load_ray<double>(rays[r]);
    int32_t i = hit_idx[r];
    double u = bounce_random<double>(force_rand, rndm, r, seed, bounce, ray.dest);
    store_ray64(out[r], evolve_one<double>(ray, i, F[i - 1], u, first_diffuse, *optics));
  }
}
// FP64 evolve of FP64 ray records (finite-difference checks perturb the expansion point in FP64)
void oracle_evolve64_f64(const oracle_ray64* in, int64_t R, const int32_t* hit_idx, const splat_ftri_t* tris, int32_t N,
                         float force_rand, int32_t first_diffuse, const splat_optics_t* optics, oracle_ray64* out) {
  std::vector<FTriT<double>> F = load_ftris<double>(tris, N);
  for (int64_t r = 0; r < R; ++r) {
    const oracle_ray64& a = in[r];
    RayT<double> ray;
    auto L = [](const double* p) { return V3<double>{p[0], p[1], p[2]}; };
    ray.pos = L(a.pos); ray.dp[0] = L(a.dpdl); ray.dp[1] = L(a.dpdx); ray.dp[2] = L(a.dpdy);
    ray.dir = L(a.dir); ray.dd[0] = L(a.dddl); ray.dd[1] = L(a.dddx); ray.dd[2] = L(a.dddy);
    ray.in_medium = a.in_medium != 0; ray.ignore_tri = a.ignore_tri; ray.dest = a.dest;
    ray.lambda = a.lambda; ray.x = a.x; ray.y = a.y; ray.status = (uint8_t)a.status;
    int32_t i = hit_idx[r];
    store_ray64(out[r], evolve_one<double>(ray, i, F[i - 1], (double)force_rand, first_diffuse, *optics));
  }
}

int64_t oracle_partition(splat_adray_t* rays, int64_t R, int32_t* carry) { return stable_partition(rays, R, carry); }

void oracle_final_evolution_f32(splat_adray_t* rays, int64_t R, const int32_t* hit_idx, const splat_ftri_t* tris, int32_t N) {
  std::vector<FTriT<float>> F = load_ftris<float>(tris, N);
  for (int64_t r = 0; r < R; ++r) {
    RayT<float> ray = load_ray<float>(rays[r]);
    int32_t i = hit_idx[r];
    store_ray(rays[r], final_evolution_one<float>(ray, i, F[i - 1]));
  }
}
void oracle_final_evolution_f64(const splat_adray_t* rays, int64_t R, const int32_t* hit_idx, const splat_ftri_t* tris, int32_t N, oracle_ray64* out) {
  std::vector<FTriT<double>> F = load_ftris<double>(tris, N);
  for (int64_t r = 0; r < R; ++r) {
    RayT<double> ray = load_ray<double>(rays[r]);
    int32_t i = hit_idx[r];
    store_ray64(out[r], final_evolution_one<double>(ray, i, F[i - 1]));
  }
}

// continuum_light_map2! (src/ray_imagers.jl:142-165): final_evolution then splat, sequential
void oracle_light_map_f32(splat_adray_t* rays, int64_t R, const int32_t* hit_idx, const splat_ftri_t* tris, int32_t N,
                          const splat_lightmap_params_t* lm, float* tex) {
  std::vector<FTriT<float>> F = load_ftris<float>(tris, N);
  for (int64_t r = 0; r < R; ++r) {
    RayT<float> ray = load_ray<float>(rays[r]);
    int32_t i = hit_idx[r];
    RayT<float> fe = final_evolution_one<float>(ray, i, F[i - 1]);
    store_ray(rays[r], fe);
    splat_one<float, float>(fe, F[i - 1], *lm, tex);
  }
}
void oracle_light_map_f64(const splat_adray_t* rays, int64_t R, const int32_t* hit_idx, const splat_ftri_t* tris, int32_t N,
                          const splat_lightmap_params_t* lm, double* tex) {
  std::vector<FTriT<double>> F = load_ftris<double>(tris, N);
  for (int64_t r = 0; r < R; ++r) {
    RayT<double> ray = load_ray<double>(rays[r]);
    int32_t i = hit_idx[r];
    RayT<double> fe = final_evolution_one<double>(ray, i, F[i - 1]);
    splat_one<double, double>(fe, F[i - 1], *lm, tex);
  }
}

// Binning margins of each ray's splats in FP64: texel_margin[r] = the smallest distance (texels)
// from any of its samples' scaled texture coordinates to a bin edge; edge_margin[r] = the smallest
// barycentric coordinate of its final hit point on its diffuse triangle (negative outside; +inf
// for rays that are not DIFFUSE, or whose final hit misses). A sample within rounding of a texel
// edge (or a hit within rounding of a triangle edge) lands in a precision-dependent bin.
void oracle_splat_margins_f64(const splat_adray_t* rays, int64_t R, const int32_t* hit_idx, const splat_ftri_t* tris,
                              int32_t N, const splat_lightmap_params_t* lm, double* texel_margin, double* edge_margin) {
  std::vector<FTriT<double>> F = load_ftris<double>(tris, N);
  const double inf = std::numeric_limits<double>::infinity();
  for (int64_t r = 0; r < R; ++r) {
    texel_margin[r] = inf;
    edge_margin[r] = inf;
    RayT<double> ray = load_ray<double>(rays[r]);
    if (ray.status != SPLAT_RAY_DIFFUSE) continue;
    int32_t i = hit_idx[r];
    RayT<double> fe = final_evolution_one<double>(ray, i, F[i - 1]);
    if (fe.ignore_tri == i) {
      double bu, bv;
      reverse_uv(F[i - 1], fe.pos, &bu, &bv);
      edge_margin[r] = std::min(std::min(bu, bv), (1.0 - bu) - bv);
    }
    double m = inf;
    splat_samples(fe, F[i - 1], *lm, [&](int, double, double uw, double vh, double) {
      m = std::min(m, std::min(edge_distance(uw, lm->width), edge_distance(vh, lm->height)));
    });
    texel_margin[r] = m;
  }
}

void oracle_raygen(const splat_pinhole_t* cam, splat_adray_t* rays, int64_t R) {
  for (int64_t q = 0; q < R; ++q) raygen_one(*cam, q, &rays[q]);
}

// the rays at the given generation indices of a pinhole frame (a sample of a frame too large to
// generate whole on the host)
void oracle_raygen_at(const splat_pinhole_t* cam, const int64_t* q, int64_t n, splat_adray_t* rays) {
  for (int64_t k = 0; k < n; ++k) raygen_one(*cam, q[k], &rays[k]);
}

// rays_from_lights(lights, upscale) (src/ray_generators.jl:64-113): each light's rays in turn;
// within a light, ray_generator.(1:h, (1:w)') reshaped to a vector is column-major, rows fastest.
// rect_light_ray (:66-94) in Float32, in the reference's order: origin(x, y) = center +
// dim1 * (x - h/2) / (h/2) + dim2 * (y - w/2) / (w/2) (Julia evaluates a * b / c as (a * b) / c
// and a + b + c as (a + b) + c; h/2 is integer division), its ForwardDiff derivatives in x and y
// (dual partial dim1 * 1 divided by h/2), dest 0 and (x, y) = (x + dx, y + dy). rand(Float32)'s
// offsets are replaced by seeded ones: dx = jitter * u01(seed, 4<<32 | 2 light, j), dy from
// stream 2 light + 1; a uniform wavelength from stream RNG_LAMBDA | (light + 1).
// Returns the number of rays written.
int64_t oracle_raygen_lights(const splat_rect_light_t* lights, int32_t n_lights, const splat_light_rays_t* p,
                             splat_adray_t* rays) {
  const uint64_t RNG_LIGHT = 4ull << 32;
  int64_t out = 0;
  for (int32_t li = 0; li < n_lights; ++li) {
    const splat_rect_light_t& L = lights[li];
    const int64_t h = L.res1 / p->upscale, w = L.res2 / p->upscale;
    const float half_h = (float)(h / 2), half_w = (float)(w / 2);
    int64_t j = 0;
    for (int64_t y = 1; y <= w; ++y) {
      for (int64_t x = 1; x <= h; ++x, ++j) {
        const float ox = p->jitter * u01(p->seed, RNG_LIGHT | (uint64_t)(2 * li), (uint64_t)j);
        const float oy = p->jitter * u01(p->seed, RNG_LIGHT | (uint64_t)(2 * li + 1), (uint64_t)j);
        const float xf = (float)x + ox, yf = (float)y + oy;
        splat_adray_t r;
        std::memset(&r, 0, sizeof r);
        for (int k = 0; k < 3; ++k) {
          const float a = (L.dim1[k] * (xf - half_h)) / half_h;
          const float b = (L.dim2[k] * (yf - half_w)) / half_w;
          r.pos[k] = (L.center[k] + a) + b;
          r.dpdx[k] = (L.dim1[k] * 1.0f) / half_h;
          r.dpdy[k] = (L.dim2[k] * 1.0f) / half_w;
          r.dir[k] = L.dir[k];
        }
        r.in_medium = 0;
        r.ignore_tri = 1;
        r.dest = 0;
        r.lambda = p->lambda_min;
        if (p->lambda_mode == SPLAT_LAMBDA_UNIFORM)
          r.lambda = std::fmaf(u01(p->seed, RNG_LAMBDA | (uint64_t)(li + 1), (uint64_t)j),
                               p->lambda_max - p->lambda_min, p->lambda_min);
        r.x = xf;
        r.y = yf;
        r.status = SPLAT_RAY_ACTIVE;
        rays[out++] = r;
      }
    }
  }
  return out;
}

// run_evolution! with the brute-force (bv_counts == NULL) or the BV-culled hitter
int64_t oracle_run_evolution_f32(splat_adray_t* rays, int64_t R, const splat_ftri_t* tris, int32_t N, int32_t* hit_idx,
                                 const splat_run_params_t* params, int32_t n_bv, const int64_t* bv_counts,
                                 const int64_t* bv_tris, int64_t bv_stride, int32_t threads) {
  set_threads(threads);
  HitCtx H;
  H.preps = prep_ftris<float>(tris, N);
  if (bv_counts) { H.use_bv = true; H.bv = build_oracle_bv(tris, n_bv, bv_counts, bv_tris, bv_stride); }
  std::vector<FTriT<float>> F = load_ftris<float>(tris, N);
  return run_evolution_f32(H, F, rays, R, hit_idx, *params);
}

// a whole light-map frame: rays from the pinhole, run_evolution!, continuum_light_map2!
// (the per-frame half of trace!, src/tracer.jl:317-344). Returns the final live count.
int64_t oracle_trace_frame_f32(const splat_pinhole_t* cam, splat_adray_t* rays, int64_t R, const splat_ftri_t* tris,
                               int32_t N, int32_t* hit_idx, const splat_run_params_t* params,
                               const splat_lightmap_params_t* lm, float* tex, int32_t n_bv, const int64_t* bv_counts,
                               const int64_t* bv_tris, int64_t bv_stride, int32_t threads) {
  set_threads(threads);
#pragma omp parallel for schedule(static, 4096)
  for (int64_t q = 0; q < R; ++q) raygen_one(*cam, q, &rays[q]);
  HitCtx H;
  H.preps = prep_ftris<float>(tris, N);
  if (bv_counts) { H.use_bv = true; H.bv = build_oracle_bv(tris, n_bv, bv_counts, bv_tris, bv_stride); }
  std::vector<FTriT<float>> F = load_ftris<float>(tris, N);
  int64_t live = run_evolution_f32(H, F, rays, R, hit_idx, *params);
  std::memset(tex, 0, sizeof(float) * (size_t)lm->width * lm->height * lm->channels);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t r = 0; r < R; ++r) {
    RayT<float> ray = load_ray<float>(rays[r]);
    int32_t i = hit_idx[r];
    RayT<float> fe = final_evolution_one<float>(ray, i, F[i - 1]);
    store_ray(rays[r], fe);
    splat_one<float, float>(fe, F[i - 1], *lm, tex);
  }
  return live;
}

// oracle-subset frame for the CPU baseline: trace only the rays q = offset + k*stride of a
// pinhole frame (a bounded, image-wide sample). Returns the number of rays traced.
int64_t oracle_trace_sample_f32(const splat_pinhole_t* cam, int64_t R_full, int64_t stride, int64_t offset,
                                const splat_ftri_t* tris, int32_t N, const splat_run_params_t* params,
                                const splat_lightmap_params_t* lm, float* tex, int32_t n_bv, const int64_t* bv_counts,
                                const int64_t* bv_tris, int64_t bv_stride, int32_t threads) {
  set_threads(threads);
  int64_t R = (R_full - offset + stride - 1) / stride;
  std::vector<splat_adray_t> rays((size_t)R);
  std::vector<int32_t> hit_idx((size_t)R);
#pragma omp parallel for schedule(static, 4096)
  for (int64_t k = 0; k < R; ++k) raygen_one(*cam, offset + k * stride, &rays[k]);
  HitCtx H;
  H.preps = prep_ftris<float>(tris, N);
  if (bv_counts) { H.use_bv = true; H.bv = build_oracle_bv(tris, n_bv, bv_counts, bv_tris, bv_stride); }
  std::vector<FTriT<float>> F = load_ftris<float>(tris, N);
  run_evolution_f32(H, F, rays.data(), R, hit_idx.data(), *params);
  std::memset(tex, 0, sizeof(float) * (size_t)lm->width * lm->height * lm->channels);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t r = 0; r < R; ++r) {
    RayT<float> ray = load_ray<float>(rays[r]);
    int32_t i = hit_idx[r];
    RayT<float> fe = final_evolution_one<float>(ray, i, F[i - 1]);
    splat_one<float, float>(fe, F[i - 1], *lm, tex);
  }
  return R;
}

// backward shading of rays in pixel order (continuum_shade!, src/ray_imagers.jl:440-503)
void oracle_shade_f32(splat_adray_t* rays, int64_t R, const int32_t* hit_idx, const splat_ftri_t* tris, int32_t N,
                      const splat_lightmap_params_t* lm, const float* tex, const float* sky, int32_t sky_w,
                      int32_t sky_h, int64_t h, float* rgb, int32_t threads) {
  set_threads(threads);
  std::vector<FTriT<float>> F = load_ftris<float>(tris, N);
  shade_frame<float, float>(rays, R, hit_idx, F, N, *lm, tex, sky, sky_w, sky_h, h, rgb, true);
}
void oracle_shade_f64(const splat_adray_t* rays, int64_t R, const int32_t* hit_idx, const splat_ftri_t* tris, int32_t N,
                      const splat_lightmap_params_t* lm, const float* tex, const float* sky, int32_t sky_w,
                      int32_t sky_h, int64_t h, double* rgb, int32_t threads) {
  set_threads(threads);
  std::vector<FTriT<double>> F = load_ftris<double>(tris, N);
  std::vector<splat_adray_t> tmp(rays, rays + R);
  shade_frame<double, double>(tmp.data(), R, hit_idx, F, N, *lm, tex, sky, sky_w, sky_h, h, rgb, false);
}
double oracle_bilinear_f64(const float* plane, int64_t W, int64_t H, double u, double v) {
  return tex_bilinear<double>(plane, W, H, u, v);
}

// run_evolution!'s reorder (src/tracer.jl:270-278): ray and hit index to slot dest - 1
void oracle_reorder(splat_adray_t* rays, int64_t R, int32_t* hit_idx) {
  std::vector<splat_adray_t> t(rays, rays + R);
  std::vector<int32_t> h(hit_idx, hit_idx + R);
  for (int64_t r = 0; r < R; ++r) {
    int64_t d = t[r].dest;
    if (d >= 1 && d <= R) { rays[d - 1] = t[r]; hit_idx[d - 1] = h[r]; }
  }
}

}  // extern "C"
