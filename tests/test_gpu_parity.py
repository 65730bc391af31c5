"""HIP path vs the CPU oracle, through the C ABI (needs an MI355X).

Bars: hit indices, partition permutations and ray states bit-exact against the FP32 oracle;
light maps within 1e-4 relative L2 of the FP32 oracle (only the atomic accumulation order
differs); distances and tangents within 1e-5 relative of the FP64 oracle on well-conditioned rays.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from splat_amd import abi, api, bv, scene
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _sync():
    import torch

    torch.cuda.synchronize()


@pytest.fixture(scope="module")
def c1():
    sc = scene.scene_c1()
    cam = scene.camera_c1(seed=11, height=64, width=64)
    rays = O.raygen(cam)
    return sc, cam, rays


# ---- ray source ---------------------------------------------------------------------------

@pytest.mark.parametrize("order,lmode,spp,jitter", [
    (abi.ORDER_TILED8, abi.LAMBDA_UNIFORM, 1, 0.0),
    (abi.ORDER_COLUMN_MAJOR, abi.LAMBDA_FIXED, 1, 0.025),
    (abi.ORDER_TILED8, abi.LAMBDA_STRATIFIED, 4, 0.025),
    (abi.ORDER_TILED8, abi.LAMBDA_UNIFORM, 1, 0.025),  # 36 x 20 is not a multiple of 8: column-major fallback
])
def test_raygen_bitexact(order, lmode, spp, jitter):
    h, w = (36, 20) if jitter and lmode == abi.LAMBDA_UNIFORM else (32, 48)
    cam = scene.camera_c1(seed=5, height=h, width=w, spp=spp, order=order, lambda_mode=lmode, jitter=jitter)
    want = O.raygen(cam)
    got = api.from_device(api.rays_from_pinhole(cam), abi.ADRAY)
    assert got.tobytes() == want.tobytes()


def _lights():
    return [
        abi.rect_light((0.0, 0.0, 8.0), (0.0, 0.0, -1.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), 64, 48),
        abi.rect_light((-50.0, 50.0, 0.0), (1.0, -1.0, 0.0), (14.0, 14.0, 0.0), (0.0, 0.0, 20.0), 32, 32),
        abi.rect_light((1.5, -2.0, 3.0), (0.3, 0.4, -0.866), (0.7, 0.1, -0.2), (0.05, 0.9, 0.3), 37, 22),
    ]


@pytest.mark.parametrize("which,up,lmode,jitter", [
    ((0,), 1, abi.LAMBDA_FIXED, 1.0),
    ((0, 1), 2, abi.LAMBDA_FIXED, 1.0),
    ((2,), 1, abi.LAMBDA_UNIFORM, 0.5),
    ((2,), 2, abi.LAMBDA_FIXED, 1.0),  # one light: (37 / 2) x (22 / 2) cells, no slice check
    ((0, 1, 2), 1, abi.LAMBDA_UNIFORM, 1.0),
])
def test_rays_from_lights_bitexact(which, up, lmode, jitter):
    lights = [_lights()[i] for i in which]
    p = abi.light_rays(up, lambda_mode=lmode, lambda_min=420.0, lambda_max=680.0, jitter=jitter, seed=9)
    want = O.raygen_lights(lights, p)
    got = api.from_device(api.rays_from_lights(lights, p), abi.ADRAY)
    assert got.shape == want.shape
    assert got.tobytes() == want.tobytes()


def test_rays_from_lights_edges():
    import torch

    # no lights: nothing to generate
    assert api.rays_from_lights([], abi.light_rays()).shape[0] == 0
    # one row: the integer half height is 0 and the reference divides by it; the GPU gives the
    # same infinities and NaNs as the oracle
    L = abi.rect_light((0.0, 0.0, 1.0), (0.0, 0.0, -1.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), 1, 8)
    p = abi.light_rays(1, seed=4)
    want = O.raygen_lights([L], p)
    got = api.from_device(api.rays_from_lights([L], p), abi.ADRAY)
    for f in ("pos", "dpdx", "dpdy", "x", "y"):
        np.testing.assert_array_equal(got[f], want[f])  # NaN == NaN here
    # two lights whose grids do not divide by the upscale: rejected, nothing written
    bad = [_lights()[0], _lights()[2]]
    with pytest.raises(api.L.SplatError):
        api.rays_from_lights(bad, abi.light_rays(2))
    torch.cuda.synchronize()


# ---- hitters ------------------------------------------------------------------------------

def _bf_hits(rays_np, ntris_np, waves=0):
    import torch

    rays = api.to_device(rays_np)
    hitter = api.HIPBruteForceHitter(api.to_device(ntris_np), tune_waves=waves)
    tr = api.HIPTracer(rays)
    api.next_hit_(tr, hitter, rays)
    _sync()
    return tr.hit_idx.cpu().numpy()


@pytest.mark.parametrize("waves", [0, 1, 2, 4, 8, 16])
def test_next_hit_bf_bitexact(c1, waves):
    sc, cam, rays = c1
    want, _ = O.next_hit(rays, sc.ntris())
    got = _bf_hits(rays, sc.ntris(), waves)
    assert (got == want).all()
    assert (want >= sc.first_diffuse).sum() > 100 and ((want > 1) & (want < sc.first_diffuse)).sum() > 100


def test_next_hit_bf_random_rays_and_ignore(c1):
    sc = c1[0]
    pts = H.sample_surface_points(sc.ftris, 3000, seed=3)
    rays = np.concatenate([H.rays_towards(pts, seed=4), H.random_rays(2000, seed=5, n_tris=sc.n_tris)])
    # rays that start on a triangle and must ignore it
    rays[:500]["pos"] = pts[:500].astype(np.float32)
    want, _ = O.next_hit(rays, sc.ntris())
    rays[:500]["ignore_tri"] = want[:500]
    want, _ = O.next_hit(rays, sc.ntris())
    got = _bf_hits(rays, sc.ntris())
    assert (got == want).all()


def test_next_hit_ties_lowest_index():
    base = scene.sphere_mesh(1)
    sc = scene.assemble([base, base, base], [], "triplicate")  # every triangle three times
    pts = H.sample_surface_points(sc.ftris[: 1 + base.shape[0]], 1500, seed=8)
    rays = H.rays_towards(pts, seed=9)
    want, _ = O.next_hit(rays, sc.ntris())
    got = _bf_hits(rays, sc.ntris())
    assert (got == want).all()
    hit = want > 1
    assert hit.sum() > 500 and (want[hit] <= 1 + base.shape[0]).all()  # the first copy wins


def test_next_hit_empty_and_ragged(c1):
    sc, _, rays = c1
    for n in (0, 1, 63, 65, 1000):
        if n == 0:
            import torch

            rays_d = torch.empty((0, 136), dtype=torch.uint8, device="cuda")
            hitter = api.HIPBruteForceHitter(api.to_device(sc.ntris()))
            tr = api.HIPTracer(rays_d)
            api.next_hit_(tr, hitter, rays_d)
            continue
        want, _ = O.next_hit(rays[:n], sc.ntris())
        assert (_bf_hits(rays[:n], sc.ntris()) == want).all()


def test_next_hit_d_count_limits_rays(c1):
    import torch

    sc, _, rays = c1
    want, _ = O.next_hit(rays[:1000], sc.ntris())
    rays_d = api.to_device(rays)
    hitter = api.HIPBruteForceHitter(api.to_device(sc.ntris()))
    tr = api.HIPTracer(rays_d)
    tr.hit_idx.fill_(-7)
    cnt = torch.tensor([1000], dtype=torch.int64, device="cuda")
    api.next_hit_(tr, hitter, rays_d, d_count=cnt)
    got = tr.hit_idx.cpu().numpy()
    assert (got[:1000] == want).all() and (got[1000:] == -7).all()


@pytest.mark.parametrize("build,depth,group_size", [
    ("reference", 0, 0), ("reference", 2, -1), ("reference", 4, 0), ("reference", 6, 3), ("reference", 6, 0),
    ("object", 6, -1), ("object", 8, 0), ("object", 8, 5), ("object", 1, 0), ("object", 3, 0), ("object", 11, 0),
])
def test_next_hit_bv_equals_bf(build, depth, group_size):
    sc = H.small_c2(3)
    cam = scene.camera_c2(seed=3, height=64, width=64)
    rays = O.raygen(cam)
    pts = H.sample_surface_points(sc.ftris, 2000, seed=12)
    rays = np.concatenate([rays, H.rays_towards(pts, seed=13), H.random_rays(1000, seed=14, n_tris=sc.n_tris)])
    want, _ = O.next_hit(rays, sc.ntris())
    tris_d = api.to_device(sc.ftris)
    part = bv.bv_partition if build == "reference" else bv.object_partition
    bvs, mem = part(sc.ftris, depth)
    hitter = api.HIPBVHitter(tris_d, bvs, mem, group_size=group_size)
    rays_d = api.to_device(rays)
    tr = api.HIPTracer(rays_d)
    api.next_hit_(tr, hitter, rays_d)
    _sync()
    got = tr.hit_idx.cpu().numpy()
    assert (got == want).all()


@pytest.mark.parametrize("depth,group_size,walk,n_rays", [
    (3, 0, 0, None), (5, 0, 1, None), (5, 0, 2, None), (4, 4, 0, None), (4, -1, 0, None), (1, 0, 0, None),
    # at most 64 live 64-ray groups: the tree's leaves are scanned flat, unless walk has bit 4
    (5, 0, 0, 3000), (8, 0, 0, 4096), (1, 0, 0, 100), (5, 0, 4, 3000), (5, 0, 5, 3000), (5, 0, 6, 3000),
    # the wave walk takes two levels per step: even and odd depths below the splits' roots, four
    # or two children at the last step, deep trees with split roots (walk bit 4: few rays, 32 splits)
    (2, 0, 0, None), (4, 0, 0, None), (6, 0, 0, None), (7, 0, 0, None), (10, 0, 0, None),
    (6, 0, 4, 3000), (9, 0, 4, 3000), (10, 0, 4, 3000),
])
def test_next_hit_bv_top_list_equals_bf(depth, group_size, walk, n_rays):
    """Large triangles tested directly before the structure (tree_partition), every structure
    and tree walk, and the flat scan of launches with few rays: the brute-force result."""
    sc = H.small_c2(4)
    cam = scene.camera_c2(seed=4, height=64, width=64)
    rays = O.raygen(cam)
    pts = H.sample_surface_points(sc.ftris, 2000, seed=15)
    rays = np.concatenate([H.rays_towards(pts, seed=16), H.random_rays(1000, seed=17, n_tris=sc.n_tris), rays])
    rays = rays[:n_rays] if n_rays else rays
    want, _ = O.next_hit(rays, sc.ntris())
    tris_d = api.to_device(sc.ftris)
    bvs, mem, n_top = bv.tree_partition(sc.ftris, depth)
    assert n_top == 1 and mem[0].size == 2  # the ground quad
    hitter = api.HIPBVHitter(tris_d, bvs, mem, group_size=group_size, n_top=n_top)
    hitter.hs.tune_waves = walk
    rays_d = api.to_device(rays)
    tr = api.HIPTracer(rays_d)
    api.next_hit_(tr, hitter, rays_d)
    _sync()
    assert (tr.hit_idx.cpu().numpy() == want).all()


# ---- evolve ---------------------------------------------------------------------------------

def _evolve_gpu(rays_np, hit, ftris, **kw):
    import torch

    rays = api.to_device(rays_np)
    api.evolve_(rays, torch.from_numpy(hit).cuda(), api.to_device(ftris), **kw)
    _sync()
    return api.from_device(rays, abi.ADRAY)


@pytest.mark.parametrize("force_rand", [0.0, 1.0, None])
@pytest.mark.parametrize("ior", [abi.IOR_REFERENCE_LINEAR, abi.IOR_CAUCHY])
def test_evolve_bitexact_first_bounces(c1, force_rand, ior):
    sc, _, rays = c1
    opt = abi.optics(ior)
    r = rays
    for bounce in (1, 2, 3):
        hit, _ = O.next_hit(r, sc.ntris())
        fr = math.nan if force_rand is None else force_rand
        want = O.evolve(r, hit, sc.ftris, force_rand=fr, seed=21, bounce=bounce, first_diffuse=sc.first_diffuse, opt=opt)
        got = _evolve_gpu(r, hit, sc.ftris, first_diffuse=sc.first_diffuse, force_rand=force_rand, seed=21,
                          bounce=bounce, optics=opt)
        assert got.tobytes() == want.tobytes(), f"bounce {bounce}"
        r = want


def test_evolve_quirk_x_ne_y_bitexact():
    sc = scene.scene_c1()
    pts = H.sample_surface_points(sc.ftris[: sc.first_diffuse], 3000, seed=31)
    rays = H.rays_towards(pts, seed=32)
    rng = np.random.default_rng(33)
    rays["x"] = rng.uniform(-1, 1, rays.shape[0]).astype(np.float32)
    rays["y"] = rng.uniform(-1, 1, rays.shape[0]).astype(np.float32)
    for f in ("dpdx", "dpdy", "dddx", "dddy", "dddl"):
        rays[f] = (rng.standard_normal((rays.shape[0], 3)) * 1e-2).astype(np.float32)
    hit, _ = O.next_hit(rays, sc.ntris())
    for fr in (0.0, 1.0):
        want = O.evolve(rays, hit, sc.ftris, force_rand=fr, first_diffuse=sc.first_diffuse)
        got = _evolve_gpu(rays, hit, sc.ftris, first_diffuse=sc.first_diffuse, force_rand=fr)
        assert got.tobytes() == want.tobytes()


def test_evolve_rndm_array_and_inactive_untouched(c1):
    import torch

    sc, _, rays = c1
    r = rays.copy()
    r["status"][::3] = abi.RAY_DIFFUSE
    hit, _ = O.next_hit(r, sc.ntris())
    rndm = np.random.default_rng(4).uniform(0, 1, r.shape[0]).astype(np.float32)
    want = O.evolve(r, hit, sc.ftris, rndm=rndm, first_diffuse=sc.first_diffuse)
    rays_d = api.to_device(r)
    api.evolve_(rays_d, torch.from_numpy(hit).cuda(), api.to_device(sc.ftris), first_diffuse=sc.first_diffuse,
                rndm=torch.from_numpy(rndm).cuda())
    got = api.from_device(rays_d, abi.ADRAY)
    assert got.tobytes() == want.tobytes()
    assert got[::3].tobytes() == r[::3].tobytes()


@pytest.mark.parametrize("which", ["c1", "small_c2"])
@pytest.mark.parametrize("force_rand", [1.0, 0.0])
def test_evolve_matches_fp64_on_every_well_conditioned_ray(which, force_rand):
    """Bounces 1-3 on the GPU (FP32) vs the FP64 oracle on the same inputs: every comparable ray
    whose FP64 conditioning is kappa <= 30 (tests/helpers.py fp64_check) has position, direction
    and all six tangents within 1e-5 relative; at least 85% of a bounce's comparable rays qualify."""
    if which == "c1":
        sc, cam = scene.scene_c1(), scene.camera_c1(seed=11, height=80, width=80)
    else:
        sc, cam = H.small_c2(3), scene.camera_c2(seed=4, height=80, width=80)
    rays = O.raygen(cam)
    checked = 0
    for bounce in range(1, 4):
        hit, _ = O.next_hit(rays, sc.ntris())
        got = _evolve_gpu(rays, hit, sc.ftris, first_diffuse=sc.first_diffuse, force_rand=force_rand)
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


# ---- partition ------------------------------------------------------------------------------

@pytest.mark.parametrize("R", [0, 1, 255, 256, 257, 1000, 70000])
def test_partition_bitexact(R):
    rng = np.random.default_rng(R)
    r = H.random_rays(R, seed=R) if R else np.zeros(0, dtype=abi.ADRAY)
    if R:
        r["status"] = rng.choice([0, 1, 2], R, p=[0.3, 0.5, 0.2]).astype(np.uint8)
    want, n_want, _ = O.partition(r)
    import torch

    rays_d = api.to_device(r) if R else torch.empty((0, 136), dtype=torch.uint8, device="cuda")
    swap = torch.empty_like(rays_d)
    n = api.partition_(rays_d, swap)
    assert n == n_want
    assert api.from_device(rays_d, abi.ADRAY).tobytes() == want.tobytes()


# ---- light map ------------------------------------------------------------------------------

def _diffuse_rays(sc, cam, depth=3):
    rays = O.raygen(cam)
    p = abi.run_params(depth, sc.first_diffuse, force_rand=1.0)
    r, hit, _ = O.run_evolution(rays, sc.ftris, p)
    return r, hit


@pytest.mark.parametrize("mode,spectrum,n_delta", [
    ("rgb", None, 1), ("rgb", None, 2), ("mono", None, 3),
    ("spectral", np.arange(400, 701, 25, dtype=np.float32), 2),
    ("rgb", np.arange(400, 701, 50, dtype=np.float32), 1),
])
def test_light_map_vs_oracle(c1, mode, spectrum, n_delta):
    import torch

    sc, cam, _ = c1
    r, hit = _diffuse_rays(sc, cam)
    lm = abi.lightmap_params(96, 80, n_delta=n_delta, spectrum=spectrum, mode=mode)
    want_tex, want_rays = O.light_map(r, hit, sc.ftris, lm)
    rays_d = api.to_device(r)
    tr = api.HIPTracer(rays_d, N=n_delta)
    tr.hit_idx.copy_(torch.from_numpy(hit))
    tex = torch.zeros((lm.channels, lm.height, lm.width), dtype=torch.float32, device="cuda")
    api.continuum_light_map2_(tracer=tr, tris=api.to_device(sc.ftris), rays=rays_d, tex=tex, lm=lm)
    _sync()
    assert api.from_device(rays_d, abi.ADRAY).tobytes() == want_rays.tobytes()  # final_evolution is exact
    got = tex.cpu().numpy()
    assert want_tex.sum() > 0
    assert H.rel_l2(got, want_tex) < 1e-4
    # the tolerance spec: the FP64 oracle's map (FP64 final evolution, expansion and accumulation)
    # over the rays whose splats all lie clear of a bin edge (H.unambiguous_rays)
    keep = H.unambiguous_rays(r, hit, sc.ftris, lm)
    assert keep.mean() > 0.98
    rays_k = api.to_device(r[keep])
    tr_k = api.HIPTracer(rays_k, N=n_delta)
    tr_k.hit_idx.copy_(torch.from_numpy(np.ascontiguousarray(hit[keep])))
    tex_k = torch.zeros_like(tex)
    api.continuum_light_map2_(tracer=tr_k, tris=api.to_device(sc.ftris), rays=rays_k, tex=tex_k, lm=lm)
    _sync()
    tex64, _ = O.light_map(r[keep], hit[keep], sc.ftris, lm, precision="f64")
    assert H.rel_l2(tex_k.cpu().numpy(), tex64) < 1e-4


# ---- the fused bounce loop and the whole frame ---------------------------------------------

@pytest.mark.parametrize("hitter_kind", ["bf", "bv"])
@pytest.mark.parametrize("depth,force_rand", [(2, 1.0), (4, None), (0, 1.0), (5, 0.0)])
def test_run_evolution_bitexact(hitter_kind, depth, force_rand):
    sc = H.small_c2(2)
    cam = scene.camera_c2(seed=4, height=64, width=48)
    rays = O.raygen(cam)
    fr = math.nan if force_rand is None else force_rand
    p = abi.run_params(depth, sc.first_diffuse, fr, seed=99)
    want_r, want_h, want_live = O.run_evolution(rays, sc.ftris, p)
    tris_d = api.to_device(sc.ftris)
    hitter = api.HIPBVHitter(tris_d, host_tris=sc.ftris, depth=4) if hitter_kind == "bv" else \
        api.HIPBruteForceHitter(tris_d)
    rays_d = api.to_device(rays)
    tr = api.HIPTracer(rays_d)
    api.run_evolution_(hitter, tr, depth=depth, first_diffuse=sc.first_diffuse, tris=tris_d, rays=rays_d,
                       force_rand=force_rand, seed=99)
    _sync()
    assert int(tr.d_count.item()) == want_live
    assert (tr.hit_idx.cpu().numpy() == want_h).all()
    assert api.from_device(rays_d, abi.ADRAY).tobytes() == want_r.tobytes()


@pytest.mark.parametrize("size,mode,spectrum,n_delta", [
    (8, "spectral", np.arange(400, 701, 25, dtype=np.float32), 1),  # 8x8x13 window: always LDS-binned
    (16, "spectral", np.array([430.0, 550.0, 660.0], dtype=np.float32), 2),
    (6, "rgb", None, 3),
    (200, "spectral", np.arange(400, 701, 50, dtype=np.float32), 2),  # spread out: straight to the map
])
def test_fused_multi_sample_splat_vs_oracle(size, mode, spectrum, n_delta):
    """Frames whose DIFFUSE rays splat several samples each (wavelength planes, sub-pixel
    offsets), splatted inside the bounce pass: on small maps the tile's samples are summed in LDS
    bins before the global atomics, on large ones they go straight to the map; either way the map
    equals the oracle's to 1e-4 and the final rays are bit-exact."""
    sc = scene.scene_c1()
    cam = scene.camera_c1(seed=6, height=64, width=64, jitter=0.025)
    p = abi.run_params(3, sc.first_diffuse, math.nan, seed=31)
    lm = abi.lightmap_params(size, size, n_delta=n_delta, spectrum=spectrum, mode=mode)
    tris_d = api.to_device(sc.ftris)
    bvs, mem, n_top = bv.tree_partition(sc.ftris, 3)
    fr = api.Frame(cam, api.HIPBVHitter(tris_d, bvs, mem, n_top=n_top), tris_d, p, lm)
    fr.trace()
    _sync()
    r, h, _ = O.run_evolution(O.raygen(cam), sc.ftris, p)
    want_tex, want_r = O.light_map(r, h, sc.ftris, lm)
    assert (fr.hit_idx.cpu().numpy() == h).all()
    assert api.from_device(fr.rays, abi.ADRAY).tobytes() == want_r.tobytes()
    assert want_tex.sum() > 0 and H.rel_l2(fr.tex.cpu().numpy(), want_tex) < 1e-4


@pytest.mark.parametrize("up,hitter_kind", [(1, "bf"), (2, "bv")])
def test_trace_lights_vs_oracle(up, hitter_kind):
    """trace!'s forward pass from two RectLights (rays_from_lights -> run_evolution! ->
    continuum_light_map2!) in one call: rays and hits bit-exact, map at 1e-4 rel L2."""
    sc = scene.scene_c1()
    lights = [abi.rect_light((0.0, 0.0, 6.0), (0.0, 0.0, -1.0), (1.5, 0.0, 0.0), (0.0, 1.5, 0.0), 64 * up, 48 * up),
              abi.rect_light((4.0, 0.5, 2.0), (-0.8, 0.0, -0.6), (0.0, 1.2, 0.0), (0.6, 0.0, -0.8), 32 * up, 32 * up)]
    lp = abi.light_rays(up, lambda_mode=abi.LAMBDA_UNIFORM, lambda_min=400.0, lambda_max=700.0, seed=13)
    p = abi.run_params(3, sc.first_diffuse, math.nan, seed=17)
    lm = abi.lightmap_params(96, 96, n_delta=up)
    tris_d = api.to_device(sc.ftris)
    if hitter_kind == "bf":
        hitter = api.HIPBruteForceHitter(tris_d)
    else:
        bvs, mem, n_top = bv.tree_partition(sc.ftris, 3)
        hitter = api.HIPBVHitter(tris_d, bvs, mem, n_top=n_top)
    fr = api.LightFrame(lights, lp, hitter, tris_d, p, lm)
    fr.trace()
    _sync()
    want_r, want_h, _ = O.run_evolution(O.raygen_lights(lights, lp), sc.ftris, p)
    want_tex, want_r2 = O.light_map(want_r, want_h, sc.ftris, lm)
    assert fr.R == 64 * 48 + 32 * 32
    assert (fr.hit_idx.cpu().numpy() == want_h).all()
    # bit-exact but for the sign of NaNs: the light rays' x, y are cell coordinates, so the
    # reference's d/dy at r.x makes NaN tangents on some rays, on both sides
    assert H.same_rays(api.from_device(fr.rays, abi.ADRAY), want_r2)
    assert want_tex.sum() > 0 and H.rel_l2(fr.tex.cpu().numpy(), want_tex) < 1e-4


def test_trace_frame_graph_replays():
    """The plain frame has no host synchronisation and no memset nodes, so a caller can capture
    it in a HIP graph: every replay, with eager work and synchronisations between them, equals
    the eager frame."""
    import torch

    sc = H.small_c2(2)
    cam = scene.camera_c2(seed=6, height=96, width=96)
    p = abi.run_params(4, sc.first_diffuse, math.nan, seed=5)
    lm = abi.lightmap_params(128, 128)
    tris_d = api.to_device(sc.ftris)
    fr = api.Frame(cam, api.HIPBVHitter.build_on_device(tris_d, 5), tris_d, p, lm)
    fr.trace()
    _sync()
    rays0, hit0, tex0 = fr.rays.clone(), fr.hit_idx.clone(), fr.tex.cpu().numpy()
    side = torch.cuda.Stream()  # default priority
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        fr.trace()  # warm-up on the capture's side stream
    torch.cuda.current_stream().wait_stream(side)
    _sync()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        fr.trace()
    for _ in range(3):
        fr.rays.zero_()
        fr.hit_idx.fill_(-1)
        fr.tex.fill_(7.0)
        _sync()
        g.replay()
        _sync()
        assert torch.equal(fr.rays, rays0) and torch.equal(fr.hit_idx, hit0)
        assert H.rel_l2(fr.tex.cpu().numpy(), tex0) < 1e-4
        fr.trace()  # eager frames in between
        _sync()


def test_trace_frame_vs_oracle():
    sc = H.small_c2(2)
    cam = scene.camera_c2(seed=6, height=96, width=96)
    p = abi.run_params(4, sc.first_diffuse, math.nan, seed=5)
    lm = abi.lightmap_params(128, 128)
    tris_d = api.to_device(sc.ftris)
    bvs, mem = bv.bv_partition(sc.ftris, 4)
    hitter = api.HIPBVHitter(tris_d, bvs, mem)
    fr = api.Frame(cam, hitter, tris_d, p, lm)
    fr.trace()
    _sync()
    rays = O.raygen(cam)
    want_r, want_h, _ = O.run_evolution(rays, sc.ftris, p)
    want_tex, want_r2 = O.light_map(want_r, want_h, sc.ftris, lm)
    assert (fr.hit_idx.cpu().numpy() == want_h).all()
    assert api.from_device(fr.rays, abi.ADRAY).tobytes() == want_r2.tobytes()
    got = fr.tex.cpu().numpy()
    assert want_tex.sum() > 0 and H.rel_l2(got, want_tex) < 1e-4
    fr.trace()  # a second frame on the same buffers gives the same map (tex is cleared per frame)
    _sync()
    assert H.rel_l2(fr.tex.cpu().numpy(), want_tex) < 1e-4


@pytest.mark.parametrize("height", [768, 760])
def test_large_frame_vs_oracle(height):
    """1024 x 768 rays are 12,288 64-ray groups: the first bounce's hitter runs one item per group
    and stores its keys, so the frame start writes no miss keys first. 1024 x 760 stays below that
    (miss keys written, items merged). Both equal the oracle: hit indices and final rays bit for
    bit, the light map within 1e-4."""
    sc = H.small_c2(2)
    cam = scene.camera_c2(seed=9, height=height, width=1024, jitter=0.025)
    p = abi.run_params(4, sc.first_diffuse, math.nan, seed=9)
    lm = abi.lightmap_params(256, 256)
    tris_d = api.to_device(sc.ftris)
    bvs, mem = bv.bv_partition(sc.ftris, 4)
    fr = api.Frame(cam, api.HIPBVHitter(tris_d, bvs, mem), tris_d, p, lm)
    fr.trace()
    _sync()
    want_r, want_h, _ = O.run_evolution(O.raygen(cam), sc.ftris, p)
    want_tex, want_r2 = O.light_map(want_r, want_h, sc.ftris, lm)
    assert (fr.hit_idx.cpu().numpy() == want_h).all()
    assert api.from_device(fr.rays, abi.ADRAY).tobytes() == want_r2.tobytes()
    assert H.rel_l2(fr.tex.cpu().numpy(), want_tex) < 1e-4


@pytest.mark.parametrize("lam", [None, 534.0, 399.0, 350.0, 750.0, 612.5])
def test_fused_splat_curve_weights_exact_on_single_splat_texels(lam):
    """The fused bounce pass weights a single-sample splat by the retina curves through a
    per-workgroup segment table. Every texel that exactly one splat reaches (a single float atomic
    onto zero, so no summation order) must equal the oracle's bit for bit: own wavelengths drawn
    uniformly, and every ray at one wavelength -- on a breakpoint shared by two curves (534), on
    the first abscissa (399), below and above the curves (the end segments extended) and
    mid-segment. The rest of the map within the usual tolerance."""
    sc = H.small_c2(2)
    kw = {} if lam is None else dict(lambda_mode=abi.LAMBDA_FIXED, lambda_min=lam, lambda_max=lam)
    cam = scene.pinhole((0.2, 0.3, 6.0), (0.0, 0.0, -1.5), (0.0, 1.0, 0.0), 90.0, 32, 32, jitter=0.025, seed=12, **kw)
    p = abi.run_params(4, sc.first_diffuse, math.nan, seed=12)
    lm = abi.lightmap_params(384, 384)
    tris_d = api.to_device(sc.ftris)
    bvs, mem = bv.bv_partition(sc.ftris, 4)
    fr = api.Frame(cam, api.HIPBVHitter(tris_d, bvs, mem), tris_d, p, lm)
    fr.trace()
    _sync()
    got = fr.tex.cpu().numpy()
    want_r, want_h, _ = O.run_evolution(O.raygen(cam), sc.ftris, p)
    want_tex, _ = O.light_map(want_r, want_h, sc.ftris, lm)
    assert H.rel_l2(got, want_tex) < 1e-4
    # texels each diffuse ray reaches, one oracle splat at a time
    count = np.zeros(want_tex.shape[1:], dtype=np.int32)
    for r in np.flatnonzero(want_r["status"] == abi.RAY_DIFFUSE):
        one, _ = O.light_map(want_r[r:r + 1], want_h[r:r + 1], sc.ftris, lm)
        count += (one != 0).any(axis=0)
    single = count == 1
    assert single.sum() > 200
    assert np.array_equal(got[:, single], want_tex[:, single])


def test_frame_where_every_ray_finishes_at_the_first_bounce():
    """Only diffuse triangles: after the first bounce no ray stays ACTIVE, so the later bounces
    run with no live ray (the bounce pass, which sums the tile counts itself after the first
    bounce, then publishes the zero live count); the frame still equals the oracle's."""
    sc = scene.assemble([], [scene.ground_quad(-1.5, 3.0)], "ground only")
    cam = scene.camera_c2(seed=8, height=64, width=96, jitter=0.025)
    p = abi.run_params(4, sc.first_diffuse, math.nan, seed=8)
    lm = abi.lightmap_params(64, 64)
    want_r, want_h, _ = O.run_evolution(O.raygen(cam), sc.ftris, p)
    assert (want_h > 1).any() and (want_h == 1).any()  # some rays on the ground, some missing it
    want_tex, want_r2 = O.light_map(want_r, want_h, sc.ftris, lm)
    tris_d = api.to_device(sc.ftris)
    bvs, mem = bv.object_partition(sc.ftris, 1)
    for hitter in (api.HIPBVHitter(tris_d, bvs, mem), api.HIPBruteForceHitter(api.to_device(sc.ntris()))):
        fr = api.Frame(cam, hitter, tris_d, p, lm)
        fr.trace()
        _sync()
        assert (fr.hit_idx.cpu().numpy() == want_h).all()
        assert api.from_device(fr.rays, abi.ADRAY).tobytes() == want_r2.tobytes()
        assert H.rel_l2(fr.tex.cpu().numpy(), want_tex) < 1e-4


def test_light_map_curves_must_ascend():
    """The retina curves' abscissae must ascend: a frame with two swapped breakpoints is refused
    before anything runs (SPLAT_ERR_INVALID_ARG)."""
    from splat_amd import lib as SL

    sc = scene.scene_c1()
    tris_d = api.to_device(sc.ftris)
    hitter = api.HIPBruteForceHitter(api.to_device(sc.ntris()))
    cam = scene.camera_c1(seed=1, height=16, width=16)
    p = abi.run_params(2, sc.first_diffuse, math.nan, seed=1)
    lm = abi.lightmap_params(32, 32)
    lm.curve_x[2][5], lm.curve_x[2][6] = lm.curve_x[2][6], lm.curve_x[2][5]
    fr = api.Frame(cam, hitter, tris_d, p, lm)
    with pytest.raises(SL.SplatError) as err:
        fr.trace()
    assert err.value.code == -1


def test_hit_set_deeper_than_the_walk_supports_is_refused():
    """The tree walk keeps the pending nodes of at most 24 levels: a hit set that claims a deeper
    tree is refused before anything runs, by next_hit and by engine admission."""
    from splat_amd import lib as SL

    sc = H.small_c2(2)
    tris_d = api.to_device(sc.ftris)
    bvs, mem, n_top = bv.tree_partition(sc.ftris, 5)
    hitter = api.HIPBVHitter(tris_d, bvs, mem, group_size=0, n_top=n_top)
    hitter.hs.tree_levels = 26
    rays = api.to_device(O.raygen(scene.camera_c2(seed=4, height=16, width=16)))
    tr = api.HIPTracer(rays)
    with pytest.raises(SL.SplatError) as err:
        api.next_hit_(tr, hitter, rays)
    assert err.value.code == -1
    import torch

    cam = scene.camera_c2(seed=4, height=16, width=16)
    lm = abi.lightmap_params(16, 16)
    eng = api.Engine(1 << 26, max_slots=1)
    with pytest.raises(SL.SplatError) as err:
        eng.admit(cam=cam, hitter=hitter, tris=tris_d, params=abi.run_params(2, sc.first_diffuse, math.nan, seed=1),
                  lm=lm, tex=torch.zeros((lm.channels, 16, 16), dtype=torch.float32, device="cuda"), tag=1)
    assert err.value.code == -1
    eng.close()
