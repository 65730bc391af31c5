// spec.hpp -- device arithmetic of the tracing hot path (gfx950).
//
// Every floating-point operation here is written out explicitly (fused multiply-adds via fma,
// the file compiled with -ffp-contract=off, correctly rounded / and sqrt), so a kernel's result
// is a fixed function of its inputs: the CPU restatement used by the tests reproduces it bit for
// bit. The formulas restate the reference's math (file:line cited per function); where the
// reference leaves rounding open (@fastmath, ForwardDiff internals) the operation order below is
// the definition.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/splat_hip.h"

#define SPL_HD __host__ __device__ __forceinline__

namespace spl {

SPL_HD float FMA(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
SPL_HD double FMA(double a, double b, double c) { return __builtin_fma(a, b, c); }
SPL_HD float SQRT(float a) { return __builtin_sqrtf(a); }
SPL_HD double SQRT(double a) { return __builtin_sqrt(a); }

// ---- dual numbers: value + tangents wrt (lambda, x, y) ------------------------------------
template <class T>
struct D3 {
  T v;
  T d[3];
};
template <class T> SPL_HD T val(T a) { return a; }
template <class T> SPL_HD T val(const D3<T>& a) { return a.v; }

#define SPL_K _Pragma("unroll") for (int k = 0; k < 3; ++k)

template <class T> SPL_HD D3<T> operator+(const D3<T>& a, const D3<T>& b) { D3<T> r; r.v = a.v + b.v; SPL_K r.d[k] = a.d[k] + b.d[k]; return r; }
template <class T> SPL_HD D3<T> operator+(const D3<T>& a, T b) { D3<T> r; r.v = a.v + b; SPL_K r.d[k] = a.d[k]; return r; }
template <class T> SPL_HD D3<T> operator+(T a, const D3<T>& b) { D3<T> r; r.v = a + b.v; SPL_K r.d[k] = b.d[k]; return r; }
template <class T> SPL_HD D3<T> operator-(const D3<T>& a, const D3<T>& b) { D3<T> r; r.v = a.v - b.v; SPL_K r.d[k] = a.d[k] - b.d[k]; return r; }
template <class T> SPL_HD D3<T> operator-(const D3<T>& a, T b) { D3<T> r; r.v = a.v - b; SPL_K r.d[k] = a.d[k]; return r; }
template <class T> SPL_HD D3<T> operator-(T a, const D3<T>& b) { D3<T> r; r.v = a - b.v; SPL_K r.d[k] = -b.d[k]; return r; }
template <class T> SPL_HD D3<T> operator-(const D3<T>& a) { D3<T> r; r.v = -a.v; SPL_K r.d[k] = -a.d[k]; return r; }
template <class T> SPL_HD D3<T> operator*(const D3<T>& a, const D3<T>& b) { D3<T> r; r.v = a.v * b.v; SPL_K r.d[k] = FMA(a.v, b.d[k], a.d[k] * b.v); return r; }
template <class T> SPL_HD D3<T> operator*(const D3<T>& a, T s) { D3<T> r; r.v = a.v * s; SPL_K r.d[k] = a.d[k] * s; return r; }
template <class T> SPL_HD D3<T> operator*(T s, const D3<T>& a) { D3<T> r; r.v = s * a.v; SPL_K r.d[k] = s * a.d[k]; return r; }
template <class T> SPL_HD D3<T> operator/(const D3<T>& a, const D3<T>& b) { D3<T> r; r.v = a.v / b.v; SPL_K r.d[k] = FMA(-r.v, b.d[k], a.d[k]) / b.v; return r; }
template <class T> SPL_HD D3<T> operator/(const D3<T>& a, T s) { D3<T> r; r.v = a.v / s; SPL_K r.d[k] = a.d[k] / s; return r; }
template <class T> SPL_HD D3<T> operator/(T s, const D3<T>& b) { D3<T> r; r.v = s / b.v; SPL_K r.d[k] = (-r.v * b.d[k]) / b.v; return r; }
template <class T> SPL_HD D3<T> SQRT(const D3<T>& a) { D3<T> r; r.v = SQRT(a.v); T tv = r.v + r.v; SPL_K r.d[k] = a.d[k] / tv; return r; }
template <class T> SPL_HD D3<T> FMA(const D3<T>& a, const D3<T>& b, const D3<T>& c) { D3<T> r; r.v = FMA(a.v, b.v, c.v); SPL_K r.d[k] = FMA(a.v, b.d[k], FMA(a.d[k], b.v, c.d[k])); return r; }
template <class T> SPL_HD D3<T> FMA(T a, const D3<T>& b, const D3<T>& c) { D3<T> r; r.v = FMA(a, b.v, c.v); SPL_K r.d[k] = FMA(a, b.d[k], c.d[k]); return r; }
template <class T> SPL_HD D3<T> FMA(const D3<T>& a, T b, const D3<T>& c) { D3<T> r; r.v = FMA(a.v, b, c.v); SPL_K r.d[k] = FMA(a.d[k], b, c.d[k]); return r; }
template <class T> SPL_HD D3<T> FMA(const D3<T>& a, const D3<T>& b, T c) { D3<T> r; r.v = FMA(a.v, b.v, c); SPL_K r.d[k] = FMA(a.v, b.d[k], a.d[k] * b.v); return r; }
template <class T> SPL_HD D3<T> FMA(T a, T b, const D3<T>& c) { D3<T> r; r.v = FMA(a, b, c.v); SPL_K r.d[k] = c.d[k]; return r; }
template <class T> SPL_HD D3<T> FMA(T a, const D3<T>& b, T c) { D3<T> r; r.v = FMA(a, b.v, c); SPL_K r.d[k] = a * b.d[k]; return r; }
template <class T> SPL_HD D3<T> FMA(const D3<T>& a, T b, T c) { D3<T> r; r.v = FMA(a.v, b, c); SPL_K r.d[k] = a.d[k] * b; return r; }

template <class S>
struct V3 {
  S x, y, z;
};
template <class A, class B> SPL_HD auto operator+(const V3<A>& a, const V3<B>& b) { return V3<decltype(a.x + b.x)>{a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class A, class B> SPL_HD auto operator-(const V3<A>& a, const V3<B>& b) { return V3<decltype(a.x - b.x)>{a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class S> SPL_HD V3<S> operator-(const V3<S>& a) { return V3<S>{-a.x, -a.y, -a.z}; }

// dot: z*z' + (y*y' + x*x'), fused
template <class A, class B> SPL_HD auto dot3(const V3<A>& a, const V3<B>& b) { return FMA(a.z, b.z, FMA(a.y, b.y, a.x * b.x)); }
template <class A, class B> SPL_HD auto cross3(const V3<A>& a, const V3<B>& b) {
  using R = decltype(a.x * b.x);
  return V3<R>{FMA(a.y, b.z, -(a.z * b.y)), FMA(a.z, b.x, -(a.x * b.z)), FMA(a.x, b.y, -(a.y * b.x))};
}
template <class S> SPL_HD V3<S> normalize3(const V3<S>& v) {
  S n = SQRT(dot3(v, v));
  return V3<S>{v.x / n, v.y / n, v.z / n};
}
template <class T> SPL_HD V3<T> vval(const V3<D3<T>>& a) { return V3<T>{a.x.v, a.y.v, a.z.v}; }
SPL_HD V3<float> ld3(const float* p) { return V3<float>{p[0], p[1], p[2]}; }
SPL_HD void st3(float* p, const V3<float>& v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

// ---- hit test (distance_to_plane src/geo.jl:433-443, in_triangle :446-470, get_hit
// src/tracer.jl:45-59), per-triangle terms hoisted into a 64-byte record ----------------------
//   q0 = (n, n.a)   q1 = (a, index bits)   q2 = (Uu, 0)   q3 = (Vv, 0)
// Uu = (d11 v0 - d01 v1)/den, Vv = (d00 v1 - d01 v0)/den, v0 = c - a, v1 = b - a,
// den = d00 d11 - d01^2: u = Uu.(p - a), v = Vv.(p - a) are Blackpawn's barycentrics.
struct TriRec {
  float4 q0, q1, q2, q3;
};
static_assert(sizeof(TriRec) == 64, "TriRec is one 64-byte scalar load");

SPL_HD TriRec tri_prep(const V3<float>& n, const V3<float>& a, const V3<float>& b, const V3<float>& c, int32_t idx) {
  V3<float> v0 = c - a, v1 = b - a;
  float d00 = dot3(v0, v0), d01 = dot3(v0, v1), d11 = dot3(v1, v1);
  float den = FMA(d00, d11, -(d01 * d01));
  float inv = 1.0f / den;
  TriRec t;
  t.q0 = make_float4(n.x, n.y, n.z, dot3(n, a));
  t.q1 = make_float4(a.x, a.y, a.z, __int_as_float(idx));
  t.q2 = make_float4(FMA(d11, v0.x, -(d01 * v1.x)) * inv, FMA(d11, v0.y, -(d01 * v1.y)) * inv,
                     FMA(d11, v0.z, -(d01 * v1.z)) * inv, 0.0f);
  t.q3 = make_float4(FMA(d00, v1.x, -(d01 * v0.x)) * inv, FMA(d00, v1.y, -(d01 * v0.y)) * inv,
                     FMA(d00, v1.z, -(d01 * v0.z)) * inv, 0.0f);
  return t;
}

// One ray against one record: returns the hit predicate and t.
SPL_HD bool hit_rec(const float4& q0, const float4& q1, const float4& q2, const float4& q3, float ox, float oy,
                    float oz, float dx, float dy, float dz, int64_t ignore, float& t_out) {
  float no = FMA(q0.z, oz, FMA(q0.y, oy, q0.x * ox));
  float nd = FMA(q0.z, dz, FMA(q0.y, dy, q0.x * dx));
  float t = (q0.w - no) / nd;
  float px = FMA(dx, t, ox), py = FMA(dy, t, oy), pz = FMA(dz, t, oz);
  float wx = px - q1.x, wy = py - q1.y, wz = pz - q1.z;
  float u = FMA(q2.z, wz, FMA(q2.y, wy, q2.x * wx));
  float v = FMA(q3.z, wz, FMA(q3.y, wy, q3.x * wx));
  t_out = t;
  int32_t idx = __float_as_int(q1.w);
  return (u >= 0.0f) & (v >= 0.0f) & (u + v < 1.0f) & (t > 0.0f) & (ignore != (int64_t)idx);
}

// packed (t, index) key of src/atomic_argmin.jl:5-20: order-preserving float map in the high
// word, index in the low word; the u64 minimum is the lexicographic (t, index) minimum.
SPL_HD uint64_t pack_key(float t, int32_t idx) {
  uint32_t u = __float_as_uint(t);
  uint32_t m = u > 0x80000000u ? ~u : (u | 0x80000000u);
  return ((uint64_t)m << 32) | (uint32_t)idx;
}
SPL_HD int32_t key_index(uint64_t k) { return (int32_t)(uint32_t)(k & 0xffffffffull); }

// ---- FTri helpers (src/geo.jl:18) --------------------------------------------------------
struct FTriF {
  V3<float> v[10];
};
SPL_HD FTriF load_ftri(const splat_ftri_t* f) {
  FTriF t;
#pragma unroll
  for (int i = 0; i < 10; ++i) t.v[i] = V3<float>{f->v[i][0], f->v[i][1], f->v[i][2]};
  return t;
}

// reverse_uv (src/geo.jl:232-258)
template <class S>
SPL_HD void reverse_uv(const FTriF& t, const V3<S>& P, S* u_out, S* v_out) {
  const V3<float>& a = t.v[1];
  const V3<float>& b = t.v[2];
  const V3<float>& c = t.v[3];
  V3<float> v0v1 = b - a, v0v2 = c - a;
  V3<float> N = cross3(v0v1, v0v2);
  float denom = dot3(N, N);
  V3<float> edge1 = c - b;
  V3<S> vp1 = P - b;
  S u = dot3(N, cross3(edge1, vp1)) / denom;
  V3<float> edge2 = a - c;
  V3<S> vp2 = P - c;
  S v = dot3(N, cross3(edge2, vp2)) / denom;
  *u_out = u;
  *v_out = v;
}

// optical_normal(::FTri, pos) (src/geo.jl:202-212)
template <class S>
SPL_HD V3<S> optical_normal(const FTriF& t, const V3<S>& P) {
  S u, v;
  reverse_uv(t, P, &u, &v);
  S w = (1.0f - u) - v;
  const V3<float>& na = t.v[4];
  const V3<float>& nb = t.v[5];
  const V3<float>& nc = t.v[6];
  V3<S> raw{FMA(nc.x, w, FMA(nb.x, v, na.x * u)), FMA(nc.y, w, FMA(nb.y, v, na.y * u)),
            FMA(nc.z, w, FMA(nb.z, v, na.z * u))};
  return normalize3(raw);
}

// reflect (src/geo.jl:481-486)
template <class S> SPL_HD V3<S> reflect(const V3<S>& v, const V3<S>& normal) {
  V3<S> nh = normalize3(normal);
  S s = dot3(nh, v);
  S s2 = s + s;
  return V3<S>{FMA(-s2, nh.x, v.x), FMA(-s2, nh.y, v.y), FMA(-s2, nh.z, v.z)};
}

// refract (src/geo.jl:548-561), as written (its c1 keeps the sign of n.v)
template <class S, class E>
SPL_HD V3<S> refract(const V3<S>& v, const V3<S>& normal, const E& n1, const E& n2) {
  V3<S> nh = normalize3(normal);
  if (val(dot3(v, nh)) > 0.0f) nh = -nh;
  S c1 = dot3(normalize3(v), nh);
  E eta = n1 / n2;
  S s1 = SQRT(FMA(-c1, c1, 1.0f));
  S s2 = eta * s1;
  S c2 = SQRT(FMA(-s2, s2, 1.0f));
  S k = FMA(eta, c1, -c2);
  V3<S> out{FMA(nh.x, k, v.x * eta), FMA(nh.y, k, v.y * eta), FMA(nh.z, k, v.z * eta)};
  return normalize3(out);
}

// can_refract (src/geo.jl:528-542); 1 - 1f-9 == 1 in Float32
SPL_HD bool can_refract(const V3<float>& v, const V3<float>& normal, float n1, float n2) {
  V3<float> nh = normalize3(normal);
  if (dot3(nh, v) > 0.0f) nh = -nh;
  float c1 = dot3(normalize3(v), nh);
  if (c1 * c1 > 1.0f) return false;
  float s1 = SQRT(FMA(-c1, c1, 1.0f));
  float s2 = (n1 / n2) * s1;
  return __builtin_fabsf(s2) < 1.0f;
}

// reflectance (src/geo.jl:488-502)
SPL_HD float reflectance(const V3<float>& v, const V3<float>& normal, float n1, float n2) {
  V3<float> nh = normalize3(normal);
  if (dot3(nh, v) < 0.0f) nh = -nh;
  float c1 = dot3(normalize3(v), nh);
  float s1 = SQRT(FMA(-c1, c1, 1.0f));
  float e = n1 / n2;
  float s2 = SQRT(FMA(-(e * e), s1 * s1, 1.0f));
  float h = (n1 + n2) * 0.5f;
  float num = FMA(h, c1, -(h * s2));
  float den = FMA(n1, c1, n2 * s2);
  float q = num / den;
  return q * q;
}

// glass(lambda) as a dual in lambda (src/material.jl:1-9 or Cauchy)
SPL_HD D3<float> glass_ior(const splat_optics_t& o, float lambda) {
  D3<float> L{lambda, {1.0f, 0.0f, 0.0f}};
  if (o.ior_model == SPLAT_IOR_CAUCHY) {
    D3<float> Lu = L / 1000.0f;
    D3<float> L2 = Lu * Lu;
    return o.cauchy_a + o.cauchy_b / L2;
  }
  float m = (1.4f - 1.3f) / (700.0f - 400.0f);
  return FMA(m, L - 400.0f, 1.3f);
}

// ---- seeded random numbers: splitmix64 finaliser over (seed, stream, counter) ---------------
SPL_HD uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
SPL_HD float u01(uint64_t seed, uint64_t stream, uint64_t counter) {
  uint64_t h = mix64(mix64(seed ^ mix64(stream)) ^ counter);
  return (float)(h >> 40) * 0x1p-24f;
}
constexpr uint64_t RNG_RNDM = 1ull << 32, RNG_LAMBDA = 2ull << 32, RNG_JITTER = 3ull << 32;

}  // namespace spl
