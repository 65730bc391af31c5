"""Backward pass on the GPU (needs an MI355X): reorder and continuum_shade! against the oracle,
and the whole trace! iteration (forward light map + backward image) against the oracle's
composition of the same steps."""
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


def _backward_rays(sc, h, w, depth, seed=5):
    cam = scene.view_camera(h, w, pos=(2.5, -3.0, 2.5), look_at=(0.0, 0.0, -0.5), seed=seed)
    p = abi.run_params(depth, sc.first_diffuse, math.nan, seed=seed)
    rays, hit, _ = O.run_evolution(O.raygen(cam), sc.ftris, p)
    return cam, p, rays, hit


def test_reorder_bitexact():
    sc = scene.scene_c1()
    cam, p, rays, hit = _backward_rays(sc, 40, 56, 3)
    want_r, want_h = O.reorder(rays, hit)
    assert (want_r["dest"] == np.arange(1, rays.shape[0] + 1)).all()
    rays_d = api.to_device(rays)
    tr = api.HIPTracer(rays_d)
    tr.hit_idx.copy_(__import__("torch").from_numpy(hit).cuda())
    api.reorder_(rays_d, tr)
    _sync()
    assert api.from_device(rays_d, abi.ADRAY).tobytes() == want_r.tobytes()
    assert (tr.hit_idx.cpu().numpy() == want_h).all()


@pytest.mark.parametrize("nd,with_sky", [(1, True), (2, True), (1, False)])
def test_shade_vs_oracle(nd, with_sky):
    import torch

    sc = H.small_c2(3)
    cam, p, rays, hit = _backward_rays(sc, 48, 40, 4)
    rays, hit = O.reorder(rays, hit)
    spectrum = np.arange(400.0, 701.0, 50.0, dtype=np.float32)
    lm = abi.lightmap_params(64, 48, n_delta=nd, spectrum=spectrum, mode="spectral")
    rng = np.random.default_rng(9)
    tex = rng.uniform(0.0, 2.0, size=(spectrum.size, 48, 64)).astype(np.float32)
    sky = scene.skydome(64, 32) if with_sky else None
    want, want_rays = O.shade(rays, hit, sc.ftris, lm, tex, sky, cam.height)
    want64, _ = O.shade(rays, hit, sc.ftris, lm, tex, sky, cam.height, precision="f64")
    rays_d = api.to_device(rays)
    tr = api.HIPTracer(rays_d, nd)
    tr.hit_idx.copy_(torch.from_numpy(hit).cuda())
    rgb = api.continuum_shade_(tracer=tr, tris=api.to_device(sc.ftris), rays=rays_d, tex=torch.from_numpy(tex).cuda(),
                               lm=lm, grid_h=cam.height,
                               skydome=None if sky is None else torch.from_numpy(sky).cuda())
    _sync()
    got = rgb.cpu().numpy()
    st = np.repeat(rays["status"], nd * nd)  # not in image order; only for counting
    assert got.shape == want.shape
    assert api.from_device(rays_d, abi.ADRAY).tobytes() == want_rays.tobytes()  # final_evolution, bit-exact
    # DIFFUSE pixels are the same FP32 operations: exact; sky pixels go through acos/atan2
    # (different libm on each side): within 2e-6; all within 1e-5 of the FP64 oracle
    assert np.abs(got - want).max() < 2e-6
    assert np.abs(got - want64).max() < 1e-5
    assert (got > 0).any() and st.size == got.shape[0]


@pytest.mark.parametrize("source", ["pinhole", "rect_lights"])
def test_trace_iteration_vs_oracle(source):
    """One trace! iteration: forward spectral light map (from the point-light pinhole, or from
    RectLights as trace! does), backward camera image."""
    import torch

    sc = scene.scene_c1()
    tris = api.to_device(sc.ftris)
    spectrum = np.arange(400.0, 701.0, 25.0, dtype=np.float32)
    cam = scene.view_camera(40, 48, pos=(2.5, -3.0, 2.5), look_at=(0.0, 0.0, -0.5), seed=4)
    sky = scene.skydome(64, 32)
    bvs, mem, n_top = bv.tree_partition(sc.ftris, 3)
    hitter = api.HIPBVHitter(tris, bvs, mem, n_top=n_top)
    if source == "pinhole":
        light = scene.camera_c1(seed=3, height=96, width=96, lambda_mode=abi.LAMBDA_UNIFORM)
        src = dict(light=light)
        up = 1
    else:
        lights = [abi.rect_light((0.0, 0.0, 6.0), (0.0, 0.0, -1.0), (1.5, 0.0, 0.0), (0.0, 1.5, 0.0), 192, 192)]
        up = 2
        lp = abi.light_rays(up, lambda_mode=abi.LAMBDA_UNIFORM, lambda_min=400.0, lambda_max=700.0, seed=3)
        src = dict(lights=lights, light_params=lp)
    rgb, tex = api.trace_(forward_hitter=hitter, backward_hitter=api.HIPBruteForceHitter(tris), cam=cam, tris=tris,
                          depth=3, first_diffuse=sc.first_diffuse, spectrum=spectrum, tex_width=128, tex_height=128,
                          skydome=torch.from_numpy(sky).cuda(), seed=21, **src)
    _sync()
    # the oracle's composition of the same steps
    lm = abi.lightmap_params(128, 128, n_delta=up, spectrum=spectrum, mode="spectral")
    p_f = abi.run_params(3, sc.first_diffuse, math.nan, seed=21)
    fwd = O.raygen(light) if source == "pinhole" else O.raygen_lights(lights, lp)
    r_f, h_f, _ = O.run_evolution(fwd, sc.ftris, p_f)
    want_tex, _ = O.light_map(r_f, h_f, sc.ftris, lm)
    assert H.rel_l2(tex.cpu().numpy(), want_tex) < 1e-4
    p_b = abi.run_params(3, sc.first_diffuse, math.nan, seed=22)
    r_b, h_b, _ = O.run_evolution(O.raygen(cam), sc.ftris, p_b)
    r_b, h_b = O.reorder(r_b, h_b)
    lm_b = abi.lightmap_params(128, 128, spectrum=spectrum, mode="spectral")  # backward upscale 1
    want_rgb, _ = O.shade(r_b, h_b, sc.ftris, lm_b, want_tex, sky, cam.height)
    got = rgb.cpu().numpy()
    assert want_rgb.max() > 0.05
    assert np.abs(got - want_rgb).max() < 1e-4  # the light map's atomic order, through bilinear reads


def test_render_accumulates_iterations():
    """iterations_per_frame (src/tracer.jl:311, :380-384): the frame is the running average of
    the seeded iterations, out <- out (f-1)/f + rgb/f."""
    import torch

    sc = scene.scene_c1()
    tris = api.to_device(sc.ftris)
    hitter = api.HIPBruteForceHitter(tris)
    kw = dict(forward_hitter=hitter, backward_hitter=hitter,
              cam=scene.view_camera(24, 32, pos=(2.5, -3.0, 2.5), look_at=(0.0, 0.0, -0.5), seed=4),
              light=scene.camera_c1(seed=3, height=64, width=64), tris=tris, depth=2, first_diffuse=sc.first_diffuse,
              spectrum=np.arange(400.0, 701.0, 50.0, dtype=np.float32), tex_width=64, tex_height=64,
              skydome=torch.from_numpy(scene.skydome(32, 16)).cuda(), seed=8)
    r = api.Renderer(**kw)
    frame = r.render(iterations=3).clone()
    singles = []
    for it in range(3):
        r._iteration(it)
        singles.append(r.rgb.clone())
    _sync()
    want = singles[0]
    for f in (2, 3):
        want = want * ((f - 1) / f) + singles[f - 1] * (1.0 / f)
    assert torch.allclose(frame, want, atol=1e-6)
    assert not torch.equal(singles[0], singles[1])  # the iterations differ (seeds)
