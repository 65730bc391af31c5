"""Pins of the oracle's evolve (dual-number differentials), partition, splat and ray source by
finite differences, closed forms and independent computations (all FP64 unless stated)."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import oracle as O
from splat_amd import abi, scene
from tests import helpers as H

L = O.lib()


def _hit_rays(sc, n, seed, glass_only=True):
    hi = sc.first_diffuse if glass_only else sc.n_tris
    pts = H.sample_surface_points(sc.ftris[:hi], n, seed=seed)
    rays = H.rays_towards(pts, seed=seed + 1)
    rng = np.random.default_rng(seed + 2)
    for f in ("dpdl", "dpdx", "dpdy", "dddl", "dddx", "dddy"):
        rays[f] = (rng.standard_normal((n, 3)) * 1e-2).astype(np.float32)
    hit, _ = O.next_hit(rays, sc.ntris())
    keep = (hit > 1) & (hit < sc.first_diffuse)
    return rays[keep], hit[keep]


def _shift(r64, k, h):
    """Move the expansion point of FP64 rays along tangent k (0: lambda, 1: x, 2: y) by h."""
    out = r64.copy()
    dp, dd = ("dpdl", "dpdx", "dpdy")[k], ("dddl", "dddx", "dddy")[k]
    out["pos"] = r64["pos"] + r64[dp] * h
    out["dir"] = r64["dir"] + r64[dd] * h
    out[("lam", "x", "y")[k]] = r64[("lam", "x", "y")[k]] + h
    return out


@pytest.mark.parametrize("force_rand", [0.0, 1.0])
@pytest.mark.parametrize("ior", [abi.IOR_REFERENCE_LINEAR, abi.IOR_CAUCHY])
def test_evolve_tangents_are_derivatives(force_rand, ior):
    """With the reference quirks off (compat = 0) every output tangent of evolve is the derivative
    of the output value along the input's first-order expansion: central differences in FP64."""
    sc = scene.assemble([scene.sphere_mesh(2)], [scene.ground_quad(-1.5, 3.0)], "fd")
    rays, hit = _hit_rays(sc, 600, seed=10)
    opt = abi.optics(ior, compat=0)
    r64 = O.to64(rays)
    r64["in_medium"] = np.arange(r64.shape[0]) % 2  # both directions through the surface
    base = O.evolve64(r64, hit, sc.ftris, force_rand=force_rand, first_diffuse=sc.first_diffuse, opt=opt)
    ok = base["status"] == abi.RAY_ACTIVE
    checked = 0
    for k, h in ((0, 1e-3), (1, 1e-5), (2, 1e-5)):
        plus = O.evolve64(_shift(r64, k, h), hit, sc.ftris, force_rand=force_rand, first_diffuse=sc.first_diffuse, opt=opt)
        minus = O.evolve64(_shift(r64, k, -h), hit, sc.ftris, force_rand=force_rand, first_diffuse=sc.first_diffuse, opt=opt)
        same = ok & (plus["status"] == abi.RAY_ACTIVE) & (minus["status"] == abi.RAY_ACTIVE) & \
            (plus["in_medium"] == base["in_medium"]) & (minus["in_medium"] == base["in_medium"])
        for val, tan in (("pos", ("dpdl", "dpdx", "dpdy")[k]), ("dir", ("dddl", "dddx", "dddy")[k])):
            fd = (plus[val][same] - minus[val][same]) / (2 * h)
            an = base[tan][same]
            scale = np.maximum(np.linalg.norm(an, axis=1), 1e-3)
            err = np.linalg.norm(fd - an, axis=1) / scale
            assert np.median(err) < 1e-6 and np.quantile(err, 0.98) < 1e-3, (k, val, np.quantile(err, [0.5, 0.98]))
            checked += int(same.sum())
    assert checked > 1000


def test_position_tangent_closed_form():
    """dp' = dp + dd t + d dt and dt = -(n.dp + t n.dd) / (n.d): the new position's tangents."""
    sc = scene.assemble([scene.sphere_mesh(2)], [scene.ground_quad(-1.5, 3.0)], "cf")
    rays, hit = _hit_rays(sc, 400, seed=20)
    r64 = O.to64(rays)
    out = O.evolve64(r64, hit, sc.ftris, force_rand=1.0, first_diffuse=sc.first_diffuse, opt=abi.optics())
    n = sc.ftris["v"][hit - 1, 0].astype(np.float64)
    d = r64["dir"]
    t = np.einsum("ij,ij->i", out["pos"] - r64["pos"], d) / np.einsum("ij,ij->i", d, d)
    nd = np.einsum("ij,ij->i", n, d)
    for dp, dd in (("dpdl", "dddl"), ("dpdx", "dddx"), ("dpdy", "dddy")):
        dt = -(np.einsum("ij,ij->i", n, r64[dp]) + t * np.einsum("ij,ij->i", n, r64[dd])) / nd
        want = r64[dp] + r64[dd] * t[:, None] + d * dt[:, None]
        assert np.allclose(out[dp], want, rtol=1e-9, atol=1e-12)


def test_refract_lambda_tangent_holds_the_normal_fixed():
    """SPLAT_COMPAT_REFRACT_NORMAL_LAMBDA reproduces src/tracer.jl:130-133: the lambda tangent of
    the refracted direction is d/dl refract(d(l), N0, n1(l), n2(l)) with the hit normal N0 fixed."""
    sc = scene.assemble([scene.sphere_mesh(2)], [], "glass")
    rays, hit = _hit_rays(sc, 200, seed=30)
    opt = abi.optics(abi.IOR_CAUCHY)
    r64 = O.to64(rays)
    out = O.evolve64(r64, hit, sc.ftris, force_rand=1.0, first_diffuse=sc.first_diffuse, opt=opt)
    n_checked = 0
    h = 1e-3
    for r in range(r64.shape[0]):
        if out["status"][r] != abi.RAY_ACTIVE or out["in_medium"][r] == r64["in_medium"][r]:
            continue  # keep refracted rays only
        f = sc.ftris[hit[r] - 1: hit[r]]
        N0 = (C.c_double * 3)()
        L.oracle_optical_normal_f64(f.ctypes.data_as(C.c_void_p), (C.c_double * 3)(*out["pos"][r]), N0)
        res = []
        for s in (h, -h):
            lam = r64["lam"][r] + s
            g = L.oracle_glass_ior_f64(C.byref(opt), lam, None)
            n1, n2 = (g, 1.0) if r64["in_medium"][r] else (1.0, g)
            v = r64["dir"][r] + r64["dddl"][r] * s
            o = (C.c_double * 3)()
            L.oracle_refract_f64((C.c_double * 3)(*v), N0, n1, n2, o)
            res.append(np.array(o[:]))
        fd = (res[0] - res[1]) / (2 * h)
        assert np.allclose(out["dddl"][r], fd, rtol=1e-5, atol=1e-9)
        n_checked += 1
    assert n_checked > 50


def test_dy_at_x_quirk_matches_the_reference_closure():
    """SPLAT_COMPAT_DY_AT_X reproduces src/tracer.jl:138-141/:166-169: dd/dy is the derivative of
    y -> f(d(l0, x0, y), N(l0, x0, y)) at y = x0 (not y0), N from next_p's linear expansion."""
    sc = scene.assemble([scene.sphere_mesh(2)], [], "glass")
    rays, hit = _hit_rays(sc, 200, seed=40)
    rng = np.random.default_rng(41)
    rays["x"] = rng.uniform(-0.5, 0.5, rays.shape[0]).astype(np.float32)
    rays["y"] = rng.uniform(-0.5, 0.5, rays.shape[0]).astype(np.float32)
    opt = abi.optics(abi.IOR_CAUCHY)
    r64 = O.to64(rays)
    for fr in (0.0, 1.0):
        quirk = O.evolve64(r64, hit, sc.ftris, force_rand=fr, first_diffuse=sc.first_diffuse, opt=opt)
        plain = O.evolve64(r64, hit, sc.ftris, force_rand=fr, first_diffuse=sc.first_diffuse, opt=abi.optics(abi.IOR_CAUCHY, compat=abi.COMPAT_REFRACT_NORMAL_LAMBDA))
        n_checked = 0
        for r in range(r64.shape[0]):
            if quirk["status"][r] != abi.RAY_ACTIVE:
                continue
            d, p0 = r64["dir"][r], r64["pos"][r]
            t = (plain["pos"][r] - p0) @ d / (d @ d)
            dty = (plain["dpdy"][r] - r64["dpdy"][r] - r64["dddy"][r] * t) @ d / (d @ d)
            f = sc.ftris[hit[r] - 1: hit[r]]
            g = L.oracle_glass_ior_f64(C.byref(opt), r64["lam"][r], None)
            n1, n2 = (g, 1.0) if r64["in_medium"][r] else (1.0, g)
            refracts = quirk["in_medium"][r] != r64["in_medium"][r]
            res = []
            h = 1e-5
            for yy in (r64["x"][r] + h, r64["x"][r] - h):
                dy = yy - r64["y"][r]
                dv = d + r64["dddy"][r] * dy
                P = p0 + r64["dpdy"][r] * dy + dv * (t + dty * dy)
                N = (C.c_double * 3)()
                L.oracle_optical_normal_f64(f.ctypes.data_as(C.c_void_p), (C.c_double * 3)(*P), N)
                o = (C.c_double * 3)()
                if refracts:
                    L.oracle_refract_f64((C.c_double * 3)(*dv), N, n1, n2, o)
                else:
                    L.oracle_reflect_f64((C.c_double * 3)(*dv), N, o)
                res.append(np.array(o[:]))
            fd = (res[0] - res[1]) / (2 * h)
            assert np.allclose(quirk["dddy"][r], fd, rtol=1e-4, atol=1e-7), r
            n_checked += 1
        assert n_checked > 50


def test_partition_is_numpy_stable_partition():
    rng = np.random.default_rng(5)
    r = H.random_rays(3000, seed=6)
    r["status"] = rng.choice([0, 1, 2], 3000).astype(np.uint8)
    carry = rng.integers(0, 1000, 3000).astype(np.int32)
    out, n, c = O.partition(r, carry)
    act = r["status"] == abi.RAY_ACTIVE
    order = np.concatenate([np.nonzero(act)[0], np.nonzero(~act)[0]])
    assert n == act.sum()
    assert out.tobytes() == r[order].tobytes() and np.array_equal(c, carry[order])


def test_splat_single_ray_known_texel_and_cosine():
    """A DIFFUSE ray onto the ground quad lands in texel (ceil(u W), ceil(v H)) with
    |n . d/|d||; the light map's total is the sum of the cosines (flat ground, n = +z)."""
    half, z = 3.0, -1.5
    sc = scene.assemble([], [scene.ground_quad(z, half)], "ground")
    W, Hh = 64, 48
    lm = abi.lightmap_params(W, Hh, mode="mono")
    rng = np.random.default_rng(7)
    n = 400
    r = np.zeros(n, dtype=abi.ADRAY)
    target = np.stack([rng.uniform(-half, half, n), rng.uniform(-half, half, n), np.full(n, z)], axis=1)
    d = np.stack([rng.uniform(-0.5, 0.5, n), rng.uniform(-0.5, 0.5, n), -np.ones(n)], axis=1)
    r["pos"] = (target - 2.0 * d).astype(np.float32)
    r["dir"] = d.astype(np.float32)
    r["status"] = abi.RAY_DIFFUSE
    r["ignore_tri"] = 1
    r["lam"] = 550.0
    hit, _ = O.next_hit(r, sc.ntris())
    assert (hit > 1).all()
    tex64, _ = O.light_map(r, hit, sc.ftris, lm, precision="f64")
    dd = r["dir"].astype(np.float64)
    cosines = np.abs(dd[:, 2]) / np.linalg.norm(dd, axis=1)
    assert tex64.sum() == pytest.approx(cosines.sum(), rel=1e-12)
    for k in range(20):  # one ray at a time: exactly its texel
        t1, _ = O.light_map(r[k:k + 1], hit[k:k + 1], sc.ftris, lm, precision="f64")
        u = (target[k, 0] + half) / (2 * half)
        v = (target[k, 1] + half) / (2 * half)
        i, j = int(math.ceil(u * W)), int(math.ceil(v * Hh))
        assert t1[0, j - 1, i - 1] == pytest.approx(cosines[k], rel=1e-9)
        assert np.count_nonzero(t1) == 1


def test_pinhole_rays_closed_forms():
    cam = scene.pinhole((0, 0, 5), (0, 0, 0), (0, 1, 0), 60.0, 64, 48, lambda_mode=abi.LAMBDA_UNIFORM,
                        order=abi.ORDER_COLUMN_MAJOR, seed=3)
    rays = O.raygen(cam)
    assert np.allclose(np.linalg.norm(rays["dir"], axis=1), 1.0, atol=2e-7)  # unit without jitter
    # pixel (x, y) = (H/2, W/2) looks straight down the view axis; its d/dx is right/scale
    q = (48 // 2 - 1) * 64 + (64 // 2 - 1)
    scale = 64 * (1 / math.sqrt(2)) / cam.fov_half_sin
    assert np.allclose(rays["dir"][q], cam.dir[:], atol=1e-7)
    assert np.allclose(rays["dddx"][q], np.array(cam.right[:]) / scale, rtol=1e-6, atol=1e-9)
    assert np.allclose(rays["dddy"][q], np.array(cam.up[:]) / scale, rtol=1e-6, atol=1e-9)
    # d/dx matches the difference of neighbouring rows to the curvature's order (1/scale^2)
    q2 = q + 1
    fd = rays["dir"][q2].astype(np.float64) - rays["dir"][q].astype(np.float64)
    avg = 0.5 * (rays["dddx"][q].astype(np.float64) + rays["dddx"][q2].astype(np.float64))
    assert np.linalg.norm(fd - avg) < 2e-3 * np.linalg.norm(avg)
    assert (rays["lam"] >= 400).all() and (rays["lam"] < 700).all()
    assert np.array_equal(np.sort(rays["dest"]), np.arange(1, 64 * 48 + 1))
    assert (rays["status"] == abi.RAY_ACTIVE).all() and (rays["ignore_tri"] == 1).all()


def test_tiled_order_is_a_permutation_of_column_major():
    a = O.raygen(scene.camera_c1(seed=2, height=32, width=16, order=abi.ORDER_COLUMN_MAJOR))
    b = O.raygen(scene.camera_c1(seed=2, height=32, width=16, order=abi.ORDER_TILED8))
    assert a[np.argsort(a["dest"])].tobytes() == b[np.argsort(b["dest"])].tobytes()
    assert not np.array_equal(a["dest"], b["dest"])
