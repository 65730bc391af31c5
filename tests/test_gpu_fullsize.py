"""Parity at the workloads' full sizes (needs an MI355X).

Every ray evolves independently, so a random sample of a full-size frame's rays, traced alone
through the oracle, must end in exactly the states the GPU frame gives those rays (matched by
their sample id `dest`), with the same hit indices. Frame-level invariants are checked on the
whole frame: the final array is a permutation of the generated rays, live rays first.
The 320-triangle workload is small enough to compare whole, light map included.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from splat_amd import abi, api, bv, scene
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _frame(sc, cam, depth, lm, hitter_kind="bv", bv_depth=8, seed=2024):
    import torch

    tris = api.to_device(sc.ftris)
    if hitter_kind == "bv":
        bvs, mem, n_top = bv.tree_partition(sc.ftris, bv_depth)
        hitter = api.HIPBVHitter(tris, bvs, mem, group_size=0, n_top=n_top)
    else:
        hitter = api.HIPBruteForceHitter(tris)
    p = abi.run_params(depth, sc.first_diffuse, math.nan, seed=seed)
    fr = api.Frame(cam, hitter, tris, p, lm)
    fr.trace()
    torch.cuda.synchronize()
    return fr, p


def _check_sample(sc, cam, fr, p, n=3000, seed=0):
    rays = api.from_device(fr.rays, abi.ADRAY)
    hit = fr.hit_idx.cpu().numpy()
    R = rays.shape[0]
    # frame invariants: a permutation of the sample ids, live (ACTIVE) rays first
    assert np.array_equal(np.sort(rays["dest"]), np.arange(1, R + 1))
    act = rays["status"] == abi.RAY_ACTIVE
    n_live = int(act.sum())
    assert act[:n_live].all() and not act[n_live:].any()
    assert ((hit >= 1) & (hit <= sc.n_tris)).all()
    # per-ray parity on a sample: the generated rays with these ids, traced alone by the oracle
    rng = np.random.default_rng(seed)
    pick = np.sort(rng.choice(R, size=min(n, R), replace=False))
    gen = O.raygen(cam) if R <= 1 << 22 else None
    if gen is None:
        return  # the oracle generator for the sample below
    by_dest = np.argsort(rays["dest"])
    want_r, want_h, _ = O.run_evolution(gen[pick], sc.ftris, p)
    got = rays[by_dest[want_r["dest"] - 1]]
    got_h = hit[by_dest[want_r["dest"] - 1]]
    # final_evolution moved DIFFUSE rays onto their hit; apply it to the oracle's sample too
    want_r2 = O.final_evolution(want_r, want_h, sc.ftris)
    assert (got_h == want_h).all()
    assert got.tobytes() == want_r2.tobytes()


def test_icosphere_256sq_whole_frame_vs_oracle():
    """The 320-triangle workload at its full size: 256x256 rays, brute-force hitter, 2 bounces,
    512x512 RGB light map -- every ray and the whole map against the oracle."""
    sc = scene.scene_c1()
    cam = scene.camera_c1(seed=1)
    lm = abi.lightmap_params(512, 512)
    fr, p = _frame(sc, cam, 2, lm, hitter_kind="bf")
    rays, hit, _ = O.run_evolution(O.raygen(cam), sc.ftris, p)
    tex, rays = O.light_map(rays, hit, sc.ftris, lm)
    assert (fr.hit_idx.cpu().numpy() == hit).all()
    assert api.from_device(fr.rays, abi.ADRAY).tobytes() == rays.tobytes()
    assert tex.sum() > 0 and H.rel_l2(fr.tex.cpu().numpy(), tex) < 1e-4


def test_two_icospheres_1mp_sampled_vs_oracle():
    """1024x1024 rays over 10,240 displaced glass triangles, BV hitter, 4 bounces."""
    sc = scene.scene_c2(seed=2)
    cam = scene.camera_c2(seed=2, jitter=0.025)
    fr, p = _frame(sc, cam, 4, abi.lightmap_params(1024, 1024))
    _check_sample(sc, cam, fr, p)
    assert fr.tex.sum().item() > 0


def test_two_icospheres_past_the_local_tile_sum_sampled_vs_oracle():
    """1032x1024 rays: 4128 tiles, past the 4096 up to which the bounce pass sums the live tiles'
    counts itself after the first bounce, so every bounce takes the separate scan; same checks."""
    sc = scene.scene_c2(seed=2)
    cam = scene.camera_c2(seed=5, height=1032, width=1024, jitter=0.025)
    fr, p = _frame(sc, cam, 4, abi.lightmap_params(1024, 1024), seed=77)
    assert fr.R == 1032 * 1024 and (fr.R + 255) // 256 > 4096
    _check_sample(sc, cam, fr, p, n=2000, seed=3)
    assert fr.tex.sum().item() > 0


def test_large_scene_invariants_and_sample():
    """~178k triangles (Menger sponge + subdivision-6 icosphere), 6 bounces, BV hitter; the full
    2048^2 x 8 wavelengths frame is 33.5M rays: invariants on the whole frame; per-ray parity on
    a sample of a 512^2 x 8 frame of the same scene."""
    import torch

    sc = scene.scene_c3()
    cam_full = scene.camera_c3(seed=3)
    fr, p = _frame(sc, cam_full, 6, abi.lightmap_params(2048, 2048), bv_depth=12)
    R = fr.R
    assert R == 2048 * 2048 * 8
    st = api.from_device(fr.rays, abi.ADRAY)["status"]
    live = int((st == abi.RAY_ACTIVE).sum())
    assert (st[:live] == abi.RAY_ACTIVE).all() and (st[live:] != abi.RAY_ACTIVE).all()
    assert fr.tex.sum().item() > 0
    del fr
    torch.cuda.empty_cache()
    cam = scene.camera_c3(seed=3, height=512, width=512)
    fr, p = _frame(sc, cam, 6, abi.lightmap_params(512, 512), bv_depth=12)
    _check_sample(sc, cam, fr, p, n=2000)


def test_large_scene_strided_hitter_sampled_vs_oracle():
    """640^2 x 8 wavelengths = 3.3M rays of the ~178k-triangle scene: past 2M rays the hitter's
    launches after the first bounce run on a capped grid whose items stride (a grid sized for
    every ray would dispatch its workgroups for the few thousand rays left at the last bounces).
    Same invariants, and a sample of 2000 rays traced alone by the oracle."""
    sc = scene.scene_c3()
    cam = scene.camera_c3(seed=4, height=640, width=640)
    fr, p = _frame(sc, cam, 6, abi.lightmap_params(640, 640), bv_depth=12, seed=91)
    assert fr.R == 640 * 640 * 8 and (fr.R + 63) // 64 > 4 * 8192  # items past the capped grid
    _check_sample(sc, cam, fr, p, n=2000, seed=5)
    assert fr.tex.sum().item() > 0


def test_frames_in_flight_on_their_own_streams_equal_their_solo_frames():
    """FrameLoop (the double-buffered frame loop of src/scenes/sphere.jl:192-219) with three
    buffers on three streams: six 512x512 frames of the 10,240-triangle scene from six cameras,
    issued three at a time with no wait between them, so their launches overlap on the GPU. Each
    buffer, read after its wait, holds exactly the rays and hit indices of the same camera's frame
    traced alone, and the same light map within the float atomics' summation order; one frame's
    rays are also sampled against the oracle."""
    import torch

    sc = scene.scene_c2(seed=2)
    tris = api.to_device(sc.ftris)
    hitter = api.HIPBVHitter.build_on_device(tris, depth=10)
    p = abi.run_params(4, sc.first_diffuse, math.nan, seed=2024)
    lm = abi.lightmap_params(512, 512)
    cams = [scene.camera_c2(seed=s, height=512, width=512, jitter=0.025) for s in (3, 4, 5, 6, 7, 8)]
    loop = api.FrameLoop(cams[0], hitter, tris, p, lm, in_flight=3)
    solo = api.Frame(cams[0], hitter, tris, p, lm)
    for first in (0, 3):
        issued = [(cam, loop.trace(cam)) for cam in cams[first:first + 3]]
        loop.wait()
        assert len({id(f) for _, f in issued}) == 3
        for cam, f in issued:
            solo.cam = cam
            solo.trace()
            torch.cuda.synchronize()
            assert torch.equal(f.hit_idx, solo.hit_idx) and torch.equal(f.rays, solo.rays)
            assert solo.tex.sum().item() > 0 and H.rel_l2(f.tex.cpu().numpy(), solo.tex.cpu().numpy()) < 1e-5
    _check_sample(sc, cams[5], issued[-1][1], p, n=1500, seed=4)
    with pytest.raises(ValueError):
        loop.trace(scene.camera_c2(seed=9, height=256, width=256))
