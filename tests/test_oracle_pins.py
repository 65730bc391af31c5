"""Independent pins of the oracle's light-map weighting, splat expansion and BV-culled CPU hitter.

Each expected value here comes from a route that shares no code with the oracle:
  * retina curves: the reference's own recursive `piecewise` (src/rgb_spectrum.jl:3-32) restated
    literally, checked against numpy's linear interpolation, a few table entries read off
    src/rgb_spectrum.jl:34-87 by hand, and the normalising sums over integer lambda 400..700
    (src/rgb_spectrum.jl:116-130);
  * splat expansion (src/ray_imagers.jl:91-109 via src/geo.jl:170-177, :398-410): the hit point's
    first-order expansion from a complex-step derivative of the exact ray-plane intersection,
    barycentrics from a linear solve, the normal and uv interpolated from them, texels by
    ceil(u w) (src/ray_imagers.jl:98-101) -- never the dual-number code of the oracle;
  * the BV-culled CPU hitter (the CPU baseline's): the brute-force oracle on the same rays.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from splat_amd import abi, bv, scene

# ---- retina curves ------------------------------------------------------------------------


def reference_piecewise(points):
    """src/rgb_spectrum.jl:3-32 restated: a binary tree of `lam < x_i` splits over 1-based
    helper(lo, hi, i) ranges; a two-point range is the line through points[lo+1] and points[hi]."""
    L = len(points)

    def helper(lo, hi, i):
        if hi - lo < 2:
            raise ValueError("bad")
        if hi - lo == 2:
            x1, y1 = points[hi - 1]
            x2, y2 = points[lo]
            return lambda lam: (y1 - y2) / (x1 - x2) * (lam - x1) + y1
        left = helper(lo, i, int(math.ceil((lo + i) / 2)))
        right = helper(i - 1, hi, int(math.ceil((i + hi) / 2)))
        splitter = points[i - 1][0]
        return lambda lam: left(lam) if lam < splitter else right(lam)

    return helper(0, L, int(math.ceil(L / 2)))


def interp_extrap(lam, xs, ys):
    """Piecewise-linear through (xs, ys), the end segments extended (numpy inside the range)."""
    lam = np.asarray(lam, dtype=np.float64)
    out = np.interp(lam, xs, ys)
    lo, hi = lam < xs[0], lam > xs[-1]
    out[lo] = ys[0] + (ys[1] - ys[0]) / (xs[1] - xs[0]) * (lam[lo] - xs[0])
    out[hi] = ys[-1] + (ys[-1] - ys[-2]) / (xs[-1] - xs[-2]) * (lam[hi] - xs[-1])
    return out


CHANNELS = ("red", "green", "blue")


def _table(name):
    pts = abi._RETINA[name]
    return np.array([p[0] for p in pts], dtype=np.float64), np.array([p[1] for p in pts], dtype=np.float64)


def _norm(name):
    xs, ys = _table(name)
    return float(np.interp(np.arange(400, 701, dtype=np.float64), xs, ys).sum())


def test_retina_table_entries_read_off_the_reference():
    # src/rgb_spectrum.jl:43, :58, :74, :36, :86, :64 (hand-read)
    assert (571, 0.798209) in abi._RETINA["red"] and (534, 0.924245) in abi._RETINA["green"]
    assert (463, 0.72815) in abi._RETINA["blue"] and (400, 0.0647036) in abi._RETINA["red"]
    assert (700, 0.0424846) in abi._RETINA["blue"] and (612, 0) in abi._RETINA["green"]
    assert [len(abi._RETINA[c]) for c in CHANNELS] == [15, 13, 18]
    for c in CHANNELS:
        xs, _ = _table(c)
        assert xs[0] == 399 and xs[-1] == 700 and (np.diff(xs) > 0).all()


def test_reference_piecewise_is_linear_interpolation_with_extended_ends():
    lams = np.concatenate([np.arange(380.0, 720.25, 0.25), [399.999, 400.001, 537.9999, 538.0001]])
    for c in CHANNELS:
        f = reference_piecewise(abi._RETINA[c])
        xs, ys = _table(c)
        ref = np.array([f(float(x)) for x in lams])
        assert np.allclose(ref, interp_extrap(lams, xs, ys), rtol=0, atol=1e-12), c
    # worked by hand: midpoints and a one-third point of segments of the table
    red, green, blue = (reference_piecewise(abi._RETINA[c]) for c in CHANNELS)
    assert red(410) == pytest.approx((0.0647036 + 0.147492) / 2, abs=1e-12)
    assert green(605) == pytest.approx(0.372516 + (0.158767 - 0.372516) * 5 / 9, abs=1e-12)
    assert blue(622.5) == pytest.approx(0.220637 + (0.225269 - 0.220637) * 2.5 / 15, abs=1e-12)
    assert red(398) == pytest.approx(-0.0647036, abs=1e-12)  # below the table: first segment extended
    assert green(720) == 0.0


def _down_rays(lams, W, H, half):
    """One DIFFUSE ray per wavelength, straight down onto the ground quad at the centre of its own
    texel of a W x H map (cosine exactly 1, so the texel holds the weight itself)."""
    n = len(lams)
    assert n <= W * H
    k = np.arange(n)
    i, j = k % W, k // W
    x = -half + (i + 0.5) / W * 2 * half
    y = -half + (j + 0.5) / H * 2 * half
    r = np.zeros(n, dtype=abi.ADRAY)
    r["pos"] = np.stack([x, y, np.full(n, 2.0)], axis=1).astype(np.float32)
    r["dir"] = np.array([0.0, 0.0, -1.0], dtype=np.float32)
    r["status"] = abi.RAY_DIFFUSE
    r["ignore_tri"] = 1
    r["lam"] = np.asarray(lams, dtype=np.float32)
    return r, i, j


@pytest.fixture(scope="module")
def ground():
    half = 2.0
    sc = scene.assemble([], [scene.ground_quad(0.0, half)], "ground")
    return sc, half


def _splat_weights(rays, i, j, sc, W, H, precision):
    hit, _ = O.next_hit(rays, sc.ntris())
    assert (hit > 1).all()
    lm = abi.lightmap_params(W, H)  # own-wavelength RGB: the retina curves
    tex, _ = O.light_map(rays, hit, sc.ftris, lm, precision=precision)
    return tex[:, j, i]  # (3, n)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_oracle_curve_weights_equal_the_normalised_reference_curves(ground, precision):
    sc, half = ground
    lams = np.concatenate([np.arange(400, 701), np.arange(400.5, 700, 7.0), [399.25, 452.75, 600.0625, 699.5]])
    rays, i, j = _down_rays(lams, 32, 32, half)
    got = _splat_weights(rays, i, j, sc, 32, 32, precision)
    lam32 = rays["lam"].astype(np.float64)
    for c, name in enumerate(CHANNELS):
        f = reference_piecewise(abi._RETINA[name])
        want = np.array([f(x) for x in lam32]) / _norm(name)
        tol = 2e-7 if precision == "f64" else 1e-6
        assert np.allclose(got[c], want, rtol=tol, atol=1e-9), (name, np.max(np.abs(got[c] - want)))


def test_curve_normalisation_is_the_sum_over_integer_400_to_700(ground):
    """Each normalised curve sums to 1 over lambda = 400, 401, ..., 700 (src/rgb_spectrum.jl:116-118)
    as the splat evaluates it. Red and blue are non-zero at both ends, so a sum over 401..700 or
    400..699 is off by more than 1e-4 and would fail; every curve is 0 at 399, so a sum from 399
    is the same number. Green is 0 at 400 and 700, where no such shift shows."""
    sc, half = ground
    rays, i, j = _down_rays(np.arange(400, 701), 32, 32, half)
    got = _splat_weights(rays, i, j, sc, 32, 32, "f64")
    for c in range(3):
        assert got[c].sum() == pytest.approx(1.0, abs=1e-6)
    for c in (0, 2):  # red, blue
        assert abs(got[c][1:].sum() - 1.0) > 1e-4 and abs(got[c][:-1].sum() - 1.0) > 1e-4


def test_retina_factor_table_equals_the_reference_curves():
    """retina_factor of _spectrum_datastructs (src/rgb_spectrum.jl:135-148) on the reference's
    13-wavelength grid, used by spectral light maps and the backward shading."""
    spec = np.arange(400, 701, 25, dtype=np.float32)
    rf = abi.retina_factor(spec)
    for c, name in enumerate(CHANNELS):
        f = reference_piecewise(abi._RETINA[name])
        want = np.array([f(float(x)) for x in spec]) / _norm(name)
        assert np.allclose(rf[:, c], want, rtol=1e-6, atol=1e-10)


# ---- splat expansion at non-zero tangents --------------------------------------------------

def _tilted_triangle():
    pos = np.array([[-2.0, -1.5, 0.3], [2.5, -1.0, -0.2], [0.0, 2.5, 0.1]])
    nrm = np.array([[0.1, 0.2, 1.0], [-0.3, 0.1, 1.0], [0.2, -0.25, 1.0]])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    uv = np.array([[0.1, 0.2, 0.0], [0.9, 0.15, 0.0], [0.4, 0.95, 0.0]])
    tri = scene.mesh_to_ftri(pos, np.array([[0, 1, 2]]), nrm, uv)
    return scene.assemble([], [tri], "tilted triangle")


def _tangent_rays(sc, n, seed):
    """DIFFUSE rays aimed at interior points of the triangle from ~2 units away, with random
    non-zero position and direction tangents in lambda, x and y."""
    rng = np.random.default_rng(seed)
    v = sc.ftris["v"][1].astype(np.float64)
    a, b, c = v[1], v[2], v[3]
    w = rng.uniform(0.2, 0.6, (n, 3))
    w /= w.sum(axis=1, keepdims=True)
    target = w[:, :1] * a + w[:, 1:2] * b + w[:, 2:] * c
    nf = np.cross(b - a, c - a)
    nf /= np.linalg.norm(nf)
    d = -nf + rng.uniform(-0.35, 0.35, (n, 3))
    r = np.zeros(n, dtype=abi.ADRAY)
    r["pos"] = (target - 2.0 * d).astype(np.float32)
    r["dir"] = d.astype(np.float32)
    r["dpdl"] = rng.uniform(-1.5e-3, 1.5e-3, (n, 3)).astype(np.float32)
    r["dpdx"] = rng.uniform(-0.3, 0.3, (n, 3)).astype(np.float32)
    r["dpdy"] = rng.uniform(-0.3, 0.3, (n, 3)).astype(np.float32)
    r["dddl"] = rng.uniform(-1e-3, 1e-3, (n, 3)).astype(np.float32)
    r["dddx"] = rng.uniform(-0.15, 0.15, (n, 3)).astype(np.float32)
    r["dddy"] = rng.uniform(-0.15, 0.15, (n, 3)).astype(np.float32)
    r["lam"] = rng.uniform(480, 620, n).astype(np.float32)
    r["x"] = rng.uniform(-3, 3, n).astype(np.float32)
    r["y"] = rng.uniform(-3, 3, n).astype(np.float32)
    r["status"] = abi.RAY_DIFFUSE
    r["ignore_tri"] = 1
    return r


def expected_map(rays, sc, lm, spectrum, delta, *, swap_pd=False, neg_lambda=False, drop_dpdx=False,
                 cos_without_dd=False):
    """The light map from first principles (module docstring); the keyword switches plant one
    mistake each, to show that the check tells them apart."""
    F = sc.ftris["v"][1].astype(np.float64)
    a, b, c = F[1], F[2], F[3]
    nv, uv = F[4:7], F[7:10, :2]
    nf = F[0]
    W, H = lm.width, lm.height
    tex = np.zeros((lm.channels, H, W))
    margin = np.inf
    f64 = lambda f: rays[f].astype(np.float64)  # noqa: E731
    P0, D0 = f64("pos"), f64("dir")
    dP = [f64("dpdl"), f64("dpdx"), f64("dpdy")]
    dD = [f64("dddl"), f64("dddx"), f64("dddy")]
    if swap_pd:
        dP, dD = dD, dP
    if drop_dpdx:
        dP[1] = np.zeros_like(dP[1])
    M = np.array([[a[0], b[0], c[0]], [a[1], b[1], c[1]], [a[2], b[2], c[2]], [1.0, 1.0, 1.0]])
    curves = [reference_piecewise(abi._RETINA[nm]) for nm in CHANNELS]
    norms = [_norm(nm) for nm in CHANNELS]
    for q in range(rays.shape[0]):
        def hit_point(s):  # exact intersection of the ray family at expansion offsets s (complex ok)
            o = P0[q] + sum(dP[k][q] * s[k] for k in range(3))
            d = D0[q] + sum(dD[k][q] * s[k] for k in range(3))
            t = np.dot(nf, a - o) / np.dot(nf, d)
            return o + t * d
        h = 1e-30
        J = np.stack([hit_point([1j * h if k == m else 0.0 for k in range(3)]).imag / h for m in range(3)], axis=1)
        p0 = hit_point([0.0, 0.0, 0.0]).real
        lam0 = float(rays["lam"][q])
        for kk, lam in enumerate(spectrum):
            for iy in range(len(delta)):
                for ix in range(len(delta)):
                    s = np.array([(lam0 - lam) if neg_lambda else (lam - lam0), delta[ix], delta[iy]])
                    p = p0 + J @ s
                    d = D0[q] + (0 if cos_without_dd else sum(dD[k][q] * s[k] for k in range(3)))
                    w = np.linalg.lstsq(M, np.append(p, 1.0), rcond=None)[0]  # barycentric (a, b, c)
                    U, V = w @ uv
                    n = w @ nv
                    n /= np.linalg.norm(n)
                    cos = abs(np.dot(n, d / np.linalg.norm(d)))
                    fu, fv = U * W, V * H
                    margin = min(margin, abs(fu - round(fu)), abs(fv - round(fv)))
                    i = min(max(math.ceil(fu), 1), W) - 1
                    j = min(max(math.ceil(fv), 1), H) - 1
                    if lm.weight_mode == abi.WEIGHTS_IDENTITY:
                        tex[kk, j, i] += cos
                    else:
                        for ch in range(3):
                            tex[ch, j, i] += cos * curves[ch](float(np.float32(lam))) / norms[ch]
    return tex, margin


SPLAT_CASES = [
    ("spectral", np.array([430.0, 505.0, 610.0], dtype=np.float32), 2),
    ("rgb", np.array([430.0, 505.0, 610.0], dtype=np.float32), 2),
    ("rgb", None, 3),
]


@pytest.mark.parametrize("mode,spectrum,n_delta", SPLAT_CASES)
def test_splat_expansion_at_nonzero_tangents(mode, spectrum, n_delta):
    sc = _tilted_triangle()
    rays = _tangent_rays(sc, 40, seed=3)
    hit, _ = O.next_hit(rays, sc.ntris())
    assert (hit == 2).all()
    lm = abi.lightmap_params(64, 64, n_delta=n_delta, spectrum=spectrum, mode=mode)
    got, _ = O.light_map(rays, hit, sc.ftris, lm, precision="f64")
    delta = abi.upres(n_delta).astype(np.float64)
    if spectrum is None:  # own wavelength: one lambda per ray, expanded at zero offset
        want = np.zeros_like(got)
        margin = np.inf
        for q in range(rays.shape[0]):
            wq, mq = expected_map(rays[q:q + 1], sc, lm, [float(rays["lam"][q])], delta)
            want += wq
            margin = min(margin, mq)
    else:
        want, margin = expected_map(rays, sc, lm, spectrum.astype(np.float64), delta)
    assert margin > 1e-6  # no sample within rounding of a texel edge
    assert np.count_nonzero(want) > rays.shape[0]  # the offsets spread each ray's splats
    scale = np.abs(want).max()
    assert np.abs(got - want).max() < 1e-6 * scale
    # the FP32 oracle (the GPU's arithmetic) agrees to FP32 rounding
    got32, _ = O.light_map(rays, hit, sc.ftris, lm, precision="f32")
    assert np.linalg.norm(got32 - got) / np.linalg.norm(got) < 1e-5
    # each planted mistake moves splats or changes their weights
    if spectrum is not None:
        for kw in ("swap_pd", "neg_lambda", "drop_dpdx", "cos_without_dd"):
            bad, _ = expected_map(rays, sc, lm, spectrum.astype(np.float64), delta, **{kw: True})
            assert np.abs(got - bad).max() > 1e-4 * scale, kw


# ---- the BV-culled CPU hitter -------------------------------------------------------------

def test_cpu_bv_hitter_equals_brute_force_on_the_1mp_scene():
    """The CPU baseline's hitter (flat BV culling over object-median leaves) returns the
    brute-force oracle's hit for primary rays and for rays leaving glass surfaces."""
    sc = scene.scene_c2(seed=2)
    cam = scene.camera_c2(seed=2, height=96, width=96, jitter=0.025)
    rays = O.raygen(cam)
    counts, mat, _ = bv.pack_bv_tris(bv.object_partition(sc.ftris, 7)[1], sc.n_tris)
    want, _ = O.next_hit(rays, sc.ntris())
    assert np.array_equal(O.next_hit_bv(rays, sc.ftris, counts, mat), want)
    # second bounce: rays start on glass, ignoring the triangle they left
    r2 = O.evolve(rays, want, sc.ftris, force_rand=1.0, first_diffuse=sc.first_diffuse)
    r2 = r2[r2["status"] == abi.RAY_ACTIVE]
    assert r2.shape[0] > 1000
    want2, _ = O.next_hit(r2, sc.ntris())
    assert np.array_equal(O.next_hit_bv(r2, sc.ftris, counts, mat), want2)
    assert (want2 > 1).any() and (want2 == 1).any()


# ---- FP32 bounce vs FP64, every well-conditioned ray ----------------------------------------

@pytest.mark.parametrize("which", ["c1", "small_c2"])
@pytest.mark.parametrize("force_rand", [1.0, 0.0])
def test_fp32_bounces_within_1e5_of_fp64_on_every_well_conditioned_ray(which, force_rand):
    """Bounces 1-3 of the FP32 oracle (the device arithmetic, bit for bit) against the FP64 oracle
    on the same inputs: every comparable ray whose FP64 conditioning kappa is <= 30 has all eight
    vectors (position, direction and their six tangents) within 1e-5 relative; at least 85% of
    the comparable rays of a bounce are that well conditioned."""
    from tests import helpers as H

    if which == "c1":
        sc, cam = scene.scene_c1(), scene.camera_c1(seed=11, height=80, width=80)
    else:
        sc, cam = H.small_c2(3), scene.camera_c2(seed=4, height=80, width=80)
    rays = O.raygen(cam)
    checked = 0
    for bounce in range(1, 4):
        hit, _ = O.next_hit(rays, sc.ntris())
        got = O.evolve(rays, hit, sc.ftris, force_rand=force_rand, first_diffuse=sc.first_diffuse)
        err, kappa, comparable = H.fp64_check(got, rays, hit, sc, force_rand)
        good = comparable & (kappa <= 30)
        assert (err[good] <= 1e-5).all(), (bounce, err[good].max())
        if comparable.sum() >= 500:
            assert good.sum() >= 0.85 * comparable.sum(), (bounce, good.sum(), comparable.sum())
        checked += int(good.sum())
        rays = got[got["status"] == abi.RAY_ACTIVE]
        if rays.shape[0] < 50:
            break
    assert checked > 1000


def test_raygen_at_picks_the_frames_rays():
    cam = scene.camera_c1(seed=3, height=40, width=24, spp=3, lambda_mode=abi.LAMBDA_STRATIFIED)
    q = np.array([7, 0, 40 * 24 * 3 - 1, 1000, 7])
    assert O.raygen_at(cam, q).tobytes() == O.raygen(cam)[q].tobytes()


# ---- FP32 light map vs FP64, away from bin edges --------------------------------------------

@pytest.mark.parametrize("jitter", [0.0, 0.025])
@pytest.mark.parametrize("mode,spectrum,n_delta", [
    ("rgb", None, 1), ("rgb", None, 2), ("mono", None, 3),
    ("spectral", np.arange(400, 701, 25, dtype=np.float32), 2),
    ("rgb", np.arange(400, 701, 50, dtype=np.float32), 1),
])
def test_fp32_light_map_within_1e4_of_fp64_away_from_bin_edges(jitter, mode, spectrum, n_delta):
    """Three bounces of glass, then final evolution and splat: the FP32 oracle's map (the device's
    arithmetic; the device adds only atomic accumulation order) against the FP64 oracle's, over the
    rays whose every splat is clear of a bin edge -- at least 98% of them, also with the unjittered
    pinhole, whose central column of rays lands on the edge u = 1/2 exactly by symmetry."""
    from tests import helpers as H

    sc = scene.scene_c1()
    cam = scene.camera_c1(seed=11, height=64, width=64, jitter=jitter)
    r, hit, _ = O.run_evolution(O.raygen(cam), sc.ftris, abi.run_params(3, sc.first_diffuse, force_rand=1.0))
    lm = abi.lightmap_params(96, 80, n_delta=n_delta, spectrum=spectrum, mode=mode)
    keep = H.unambiguous_rays(r, hit, sc.ftris, lm)
    assert keep.mean() > 0.98
    t32, _ = O.light_map(r[keep], hit[keep], sc.ftris, lm)
    t64, _ = O.light_map(r[keep], hit[keep], sc.ftris, lm, precision="f64")
    assert t64.sum() > 0 and H.rel_l2(t32, t64) < 1e-5
