"""Pins of the CPU oracle: the reference's own known-answer tests, closed forms of the reference
formulas, and published constants. A plausible slip in the restatement (dropped term, wrong
sign, swapped operand) fails at least one of these."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import oracle as O
from splat_amd import abi

L = O.lib()


def v3(*x):
    return (C.c_double * 3)(*x)


def f3(*x):
    return (C.c_float * 3)(*x)


def out3():
    return (C.c_double * 3)()


def P(a):
    return C.cast(a, C.c_void_p)


# ---- packed (t, index) key: src/atomic_argmin.jl:5-30 and its property tests :32-96 -------------

def test_key_fixed_vectors():
    assert L.oracle_key64(math.inf, 1) == 0xFF800000_00000001
    assert L.oracle_key32(1.0) == 0xBF800000
    assert L.oracle_key32(0.0) == 0x80000000
    assert L.oracle_key32(-0.0) == 0x80000000
    assert L.oracle_key32(-math.inf) == 0x007FFFFF
    assert L.oracle_key64_arg(0xFF800000_00000001) == 1
    assert L.oracle_key64_arg(L.oracle_key64(2.5, 12345)) == 12345


def _test_array(rng):
    # the reference's test_array: rand * 2^i * 2 - 1 for i in -100:5:100, plus -Inf, 0, Inf
    parts = [(rng.random(1000, dtype=np.float32) * np.float32(2.0) ** i * 2 - 1).astype(np.float32)
             for i in range(-100, 101, 5)]
    return np.concatenate(parts + [np.array([-np.inf, 0.0, np.inf], dtype=np.float32)])


def test_key_order_preserving_and_reversible():
    rng = np.random.default_rng(0)
    a = _test_array(rng)
    keys = np.array([L.oracle_key32(float(x)) for x in a], dtype=np.uint64)
    order = np.argsort(keys, kind="stable")
    assert np.all(np.diff(a[order].astype(np.float64)) >= 0)  # test_comparison
    back = np.array([L.oracle_key32_inv(int(k)) for k in keys], dtype=np.float32)
    assert np.array_equal(back, a)  # test_reversibility
    # test_comparison2: the inverse map preserves order over random 32-bit keys
    ks = (rng.integers(0, 1 << 24, 20000) + rng.integers(0, 1 << 31, 20000)).astype(np.uint64)
    fs = np.array([L.oracle_key32_inv(int(k)) for k in ks], dtype=np.float64)
    ok = ~np.isnan(fs)
    o = np.argsort(ks[ok], kind="stable")
    assert np.all(np.diff(fs[ok][o]) >= 0)


def test_key64_is_lexicographic():
    rng = np.random.default_rng(1)
    t = rng.choice(np.array([0.5, 1.0, 1.0, 2.0, math.inf], dtype=np.float32), 500)
    i = rng.integers(1, 50, 500).astype(np.uint32)
    keys = [L.oracle_key64(float(a), int(b)) for a, b in zip(t, i)]
    by_key = sorted(range(500), key=lambda k: keys[k])
    by_tuple = sorted(range(500), key=lambda k: (float(t[k]), int(i[k])))
    assert [(t[k], i[k]) for k in by_key] == [(t[k], i[k]) for k in by_tuple]


# ---- distance_to_plane: test/geo.jl:24-29 --------------------------------------------------------

def test_distance_to_plane_kats():
    o, d, a, n = (1, 2, 4), (2, 3, 1), (-2, 0, -2), (8, 1, 0)
    t = L.oracle_distance_to_plane_f64(P(v3(*o)), P(v3(*d)), P(v3(*a)), P(v3(*n)))
    assert t == pytest.approx(-26 / 19, rel=1e-15)
    tn = L.oracle_distance_to_plane_f64(P(v3(*o)), P(v3(*d)), P(v3(*a)), P(v3(-8, -1, 0)))
    assert tn == t  # the normal's sign does not matter
    td = L.oracle_distance_to_plane_f64(P(v3(*o)), P(v3(-2, -3, -1)), P(v3(*a)), P(v3(*n)))
    assert td == -t  # a negated direction negates the distance
    t32 = L.oracle_distance_to_plane_f32(P(f3(*o)), P(f3(*d)), P(f3(*a)), P(f3(*n)))
    assert t32 == np.float32(-26 / 19)


# ---- hit test: in_triangle (src/geo.jl:446-470) inside get_hit (src/tracer.jl:45-59) ----------

TRI = [0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0]  # n = cross(a-b, b-c) = (0,0,1), a=(0,0,0), b=(1,0,0), c=(0,1,0)


def _hit(px, py, z=1.0, dz=-1.0, ignore=1, idx=2, tri=TRI, f32=False):
    if f32:
        t = C.c_float()
        h = L.oracle_hit_test_f32(P(f3(px, py, z)), P(f3(0, 0, dz)), P((C.c_float * 12)(*tri)), ignore, idx, C.byref(t))
    else:
        t = C.c_double()
        h = L.oracle_hit_test_f64(P(v3(px, py, z)), P(v3(0, 0, dz)), P((C.c_double * 12)(*tri)), ignore, idx, C.byref(t))
    return bool(h), t.value


@pytest.mark.parametrize("f32", [False, True])
def test_in_triangle_closed_forms(f32):
    assert _hit(0, 0, f32=f32) == (True, 1.0)  # vertex a: u = v = 0, inside
    assert not _hit(1, 0, f32=f32)[0]  # vertex b: u + v = 1, outside (strict)
    assert not _hit(0, 1, f32=f32)[0]  # vertex c
    assert _hit(1 / 3, 1 / 3, f32=f32)[0]  # centroid
    assert _hit(0.5, 0.0, f32=f32)[0]  # on edge ab: v = 0.5, u = 0
    assert not _hit(-1e-3, 0.5, f32=f32)[0]
    assert not _hit(0.6, 0.6, f32=f32)[0]
    assert _hit(0.25, 0.25, z=3.0, f32=f32) == (True, 3.0)
    assert not _hit(0.25, 0.25, z=-1.0, f32=f32)[0]  # t < 0
    assert not _hit(0.25, 0.25, z=0.0, f32=f32)[0]  # t = 0 is not > 0
    assert not _hit(0.25, 0.25, ignore=2, idx=2, f32=f32)[0]  # ignore_tri
    assert not _hit(0.25, 0.25, tri=[0] * 12, f32=f32)[0]  # the degenerate sentinel (NaN)


def _moller_trumbore(o, d, a, b, c):
    """Independent intersection (Moller-Trumbore, float64), returns (t, u, v) or None."""
    e1, e2 = b - a, c - a
    p = np.cross(d, e2)
    det = e1 @ p
    if abs(det) < 1e-14:
        return None
    s = o - a
    u = (s @ p) / det
    q = np.cross(s, e1)
    v = (d @ q) / det
    t = (e2 @ q) / det
    return t, u, v


def test_brute_force_hitter_vs_independent_intersection():
    from splat_amd import scene
    from tests import helpers as H

    sc = scene.assemble([scene.sphere_mesh(1, center=(0, 0, 0)), scene.sphere_mesh(1, center=(0.7, 0.2, 0.3), radius=0.6)],
                        [scene.ground_quad(-1.5, 3.0)], "tiny")
    pts = H.sample_surface_points(sc.ftris, 400, seed=2)
    rays = np.concatenate([H.rays_towards(pts, seed=3), H.random_rays(200, seed=4, n_tris=sc.n_tris)])
    hit64, t64 = O.next_hit(rays, sc.ntris(), "f64")
    hit32, t32 = O.next_hit(rays, sc.ntris(), "f32")
    V = sc.ftris["v"].astype(np.float64)
    robust = 0
    for r in range(rays.shape[0]):
        o, d = rays["pos"][r].astype(np.float64), rays["dir"][r].astype(np.float64)
        best, bi, margin = math.inf, 1, math.inf
        for j in range(1, sc.n_tris):
            res = _moller_trumbore(o, d, V[j, 1], V[j, 2], V[j, 3])
            if res is None:
                continue
            t, u, v = res
            m = min(abs(u), abs(v), abs(1 - u - v), abs(t))
            if j + 1 == rays["ignore_tri"][r]:
                continue
            if t > 0 and u >= 0 and v >= 0 and u + v <= 1:
                margin = min(margin, m)
                if t < best:
                    best, bi = t, j + 1
            else:
                margin = min(margin, m)
        if margin < 1e-6:
            continue  # too close to an edge for two algorithms to be compared
        robust += 1
        assert hit64[r] == bi and hit32[r] == bi, r
        if bi > 1:  # the oracle uses the stored Float32 face normal, not the exact vertex plane
            assert t64[r] == pytest.approx(best, rel=1e-6)
            assert t32[r] == pytest.approx(best, rel=1e-5)
    assert robust > 500


# ---- optics: reflect (test/geo.jl:11-22), refract / reflectance / can_refract closed forms ------

def _reflect(v, n):
    o = out3()
    L.oracle_reflect_f64(P(v3(*v)), P(v3(*n)), P(o))
    return np.array(o[:])


def test_reflect_kats():
    cases = [((-1, -1, 0), (0, 1, 0), (-1, 1, 0)), ((-1, -1, 0), (0, -1, 0), (-1, 1, 0)),
             ((-1, -1, 0), (1, 0, 0), (1, -1, 0)), ((-1, -1, 0), (-1, 0, 0), (1, -1, 0)),
             ((-4, -1, 0), (0, 1, 0), (-4, 1, 0)), ((-1, -4, 0), (0, 1, 0), (-1, 4, 0))]
    for v, n, want in cases:
        assert np.allclose(_reflect(v, n), want, atol=1e-15)
    d0, n, d1 = (-0.0993602, -0.990078, 0.0993602), (0.21059, 0.310929, -0.926809), (0.077896, -0.728365, -0.680746)
    assert np.linalg.norm(_reflect(d0, n) - np.array(d1)) < 1e-6
    o = (C.c_float * 3)()
    L.oracle_reflect_f32(P(f3(*d0)), P(f3(*n)), P(o))
    assert np.linalg.norm(np.array(o[:]) - np.array(d1)) < 1e-6


def _refract(v, n, n1, n2):
    o = out3()
    L.oracle_refract_f64(P(v3(*v)), P(v3(*n)), n1, n2, P(o))
    return np.array(o[:])


@pytest.mark.parametrize("theta_deg", [0.0, 10.0, 28.6, 40.0, 60.0])
@pytest.mark.parametrize("n1,n2", [(1.0, 1.5), (1.5, 1.0), (1.0, 1.0), (1.3, 1.4)])
def test_refract_closed_form(theta_deg, n1, n2):
    """refract(v, n, n1, n2) as written (src/geo.jl:548-561): for v = (sin t, 0, -cos t) onto the
    z = 0 plane, normalize(eta sin t, 0, -2 eta cos t - sqrt(1 - eta^2 sin^2 t))."""
    th = math.radians(theta_deg)
    eta = n1 / n2
    v = (math.sin(th), 0.0, -math.cos(th))
    if eta * math.sin(th) >= 1:
        # total internal reflection: the formula's square root of a negative number (callers
        # test can_refract first, src/tracer.jl:120)
        assert np.isnan(_refract(v, (0, 0, 1), n1, n2)).all()
        return
    want = np.array([eta * math.sin(th), 0.0, -2 * eta * math.cos(th) - math.sqrt(1 - (eta * math.sin(th)) ** 2)])
    want /= np.linalg.norm(want)
    assert np.allclose(_refract(v, (0, 0, 1), n1, n2), want, atol=1e-14)
    assert np.allclose(_refract(v, (0, 0, -3), n1, n2), want, atol=1e-14)  # either normal orientation


def test_refract_is_not_snell_reference_hazard():
    # equal indices: Snell keeps the ray straight; the reference formula bends 28.6 deg to 10.3 deg
    th = math.radians(28.6)
    r = _refract((math.sin(th), 0, -math.cos(th)), (0, 0, 1), 1.0, 1.0)
    assert math.degrees(math.atan2(r[0], -r[2])) == pytest.approx(10.3, abs=0.05)


def _reflectance(v, n, n1, n2):
    return L.oracle_reflectance_f64(P(v3(*v)), P(v3(*n)), n1, n2)


def test_reflectance_closed_forms():
    # normal incidence: ((h - h) / (n1 + n2))^2 = 0 for the reference's s/p "average"
    assert _reflectance((0, 0, -1), (0, 0, 1), 1.0, 1.5) == pytest.approx(0.0, abs=1e-15)
    # grazing (c1 = 0): ((n1 + n2) / (2 n2))^2
    assert _reflectance((1, 0, 0), (0, 0, 1), 1.0, 1.5) == pytest.approx((2.5 / 3.0) ** 2, rel=1e-14)
    assert _reflectance((1, 0, 0), (0, 0, -1), 1.2, 1.5) == pytest.approx((2.7 / 3.0) ** 2, rel=1e-14)


def test_can_refract_total_internal_reflection_threshold():
    crit = math.degrees(math.asin(1 / 1.5))
    for th, want in ((0.0, True), (crit - 0.5, True), (crit + 0.5, False), (80.0, False)):
        t = math.radians(th)
        got = L.oracle_can_refract_f64(P(v3(math.sin(t), 0, -math.cos(t))), P(v3(0, 0, 1)), 1.5, 1.0)
        assert bool(got) == want, th
    assert L.oracle_can_refract_f64(P(v3(0.7, 0, -0.7)), P(v3(0, 0, 1)), 1.0, 1.5) == 1  # into denser glass


def _ftri(a, b, c, na, nb, nc):
    f = np.zeros(1, dtype=abi.FTRI)
    a, b, c = (np.asarray(x, dtype=np.float32) for x in (a, b, c))
    f["v"][0, 0] = np.cross(a - b, b - c)
    f["v"][0, 1:4] = [a, b, c]
    f["v"][0, 4:7] = [na, nb, nc]
    f["v"][0, 7:10] = [(0, 0, 0), (1, 0, 0), (0, 1, 0)]
    return f


def test_optical_normal_and_reverse_uv_vertices():
    a, b, c = (0.0, 0.0, 0.0), (2.0, 0.0, 0.0), (0.0, 1.0, 0.5)
    na, nb, nc = (0.0, 0.0, 2.0), (0.0, 1.0, 1.0), (1.0, 0.0, 1.0)
    f = _ftri(a, b, c, na, nb, nc)
    fp = f.ctypes.data_as(C.c_void_p)
    for p, want_uv, want_n in ((a, (1, 0), na), (b, (0, 1), nb), (c, (0, 0), nc)):
        uv = (C.c_double * 2)()
        L.oracle_reverse_uv_f64(fp, P(v3(*p)), P(uv))
        assert np.allclose(uv[:], want_uv, atol=1e-12)
        o = out3()
        L.oracle_optical_normal_f64(fp, P(v3(*p)), P(o))
        assert np.allclose(o[:], np.array(want_n) / np.linalg.norm(want_n), atol=1e-12)
    cen = np.mean([a, b, c], axis=0)
    o = out3()
    L.oracle_optical_normal_f64(fp, P(v3(*cen)), P(o))
    s = np.sum([np.array(na), np.array(nb), np.array(nc)], axis=0)
    assert np.allclose(o[:], s / np.linalg.norm(s), atol=1e-12)


def test_glass_ior_closed_forms():
    lin = abi.optics(abi.IOR_REFERENCE_LINEAR)
    dn = C.c_double()
    assert L.oracle_glass_ior_f64(C.byref(lin), 400.0, C.byref(dn)) == pytest.approx(1.3, abs=1e-7)
    assert L.oracle_glass_ior_f64(C.byref(lin), 700.0, None) == pytest.approx(1.4, abs=1e-7)
    assert dn.value == pytest.approx(0.1 / 300, rel=1e-6)
    cau = abi.optics(abi.IOR_CAUCHY, 1.5, 0.004)
    assert L.oracle_glass_ior_f64(C.byref(cau), 500.0, C.byref(dn)) == pytest.approx(1.516, rel=1e-7)
    assert dn.value == pytest.approx(-2 * 0.004 / 0.5 ** 3 * 1e-3, rel=1e-7)


# ---- seeded random numbers: splitmix64's published first outputs from state 0 -------------------

def test_splitmix64_known_answers():
    # splitmix64 seeded with 0 yields 0xE220A8397B1DCDAF, then 0x6E789E6AA1B965F4 (state advances
    # by the golden gamma before each finalisation)
    assert L.oracle_mix64(0) == 0xE220A8397B1DCDAF
    assert L.oracle_mix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4
    u = [L.oracle_u01(7, 1 << 32, k) for k in range(20000)]
    assert 0.0 <= min(u) and max(u) < 1.0 and abs(np.mean(u) - 0.5) < 0.01


def test_upres_kats():
    assert np.allclose(abi.upres(4), [-0.375, -0.125, 0.125, 0.375])
    assert np.allclose(abi.upres(2), [-0.25, 0.25])
    assert abi.upres(1).tolist() == [0.0]
