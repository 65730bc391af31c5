"""The autotuner picks a launch shape whose hits equal every other candidate's (needs a GPU)."""
import json

import numpy as np
import pytest

from oracle import oracle as O
from splat_amd import api, autotune, scene
from tests import helpers as H

pytestmark = pytest.mark.gpu


def test_autotune_bf_and_bv(tmp_path):
    sc = H.small_c2(2)
    cam = scene.camera_c2(seed=8, height=64, width=64)
    rays_np = O.raygen(cam)
    want, _ = O.next_hit(rays_np, sc.ntris())
    rays = api.to_device(rays_np)
    path = str(tmp_path / "tune.json")
    bf = autotune.tune_bf(api.to_device(sc.ntris()), rays, cache_path=path)
    assert bf["waves"] in (1, 2, 4, 8, 16) and bf["ms"] > 0
    ch = autotune.tune_bv(api.to_device(sc.ftris), sc.ftris, rays, cache_path=path, group_sizes=(-1, 0, 4))
    tr = api.HIPTracer(rays)
    api.next_hit_(tr, ch["hitter"], rays)
    assert (tr.hit_idx.cpu().numpy() == want).all()
    cached = json.load(open(path))
    assert any(k.startswith("bf:") for k in cached) and any(k.startswith("bv:") for k in cached)
    again = autotune.tune_bv(api.to_device(sc.ftris), sc.ftris, rays, cache_path=path)
    assert (again["depth"], again["group_size"]) == (ch["depth"], ch["group_size"])


def test_autotune_device_bv_frames(tmp_path):
    """The device-built tree hitter tuned on whole frames: every candidate's hit indices agree,
    the chosen hitter traces the same frame as the untuned one, the choice is cached, and the
    walk mode survives a rebuild."""
    import math

    import torch

    from splat_amd import abi

    sc = H.small_c2(3)
    tris = api.to_device(sc.ftris)
    cam = scene.camera_c2(seed=9, height=96, width=96)
    p = abi.run_params(3, sc.first_diffuse, math.nan, seed=5)
    lm = abi.lightmap_params(64, 64)
    path = str(tmp_path / "tune.json")
    ch = autotune.tune_device_bv(tris, cam, p, lm, cache_path=path)
    assert ch["walk"] in (0, 1, 2) and len(ch["candidates"]) >= 6
    base = api.Frame(cam, api.HIPBVHitter.build_on_device(tris), tris, p, lm)
    base.trace()
    tuned = api.Frame(cam, ch["hitter"], tris, p, lm)
    tuned.trace()
    torch.cuda.synchronize()
    assert torch.equal(base.hit_idx, tuned.hit_idx) and torch.equal(base.rays, tuned.rays)
    again = autotune.tune_device_bv(tris, cam, p, lm, cache_path=path)
    assert (again["depth"], again["walk"], again["split"]) == (ch["depth"], ch["walk"], ch["split"])
    assert ch["split"] in autotune.SPLIT_TARGETS and len(ch["candidates"]) >= 8  # the split targets too
    assert "bvdev" in "".join(json.load(open(path)))
    ch["hitter"].rebuild(tris)
    assert ch["hitter"].hs.tune_waves == ch["walk"] | (ch["split"] << 8)


@pytest.mark.parametrize("split", [1, 2, 3, 4])
def test_split_targets_equal_default(split):
    """Every split target of the single-frame BV hitter (HitSet.tune_waves bits 8-10) traces the
    same frame as the default: a 128x128 frame is 256 64-ray groups, so each target gives the
    bounces different split factors (and 3072 / 6144 the first bounce's keys merged, not stored)."""
    import math

    import torch

    from splat_amd import abi

    sc = H.small_c2(2)
    tris = api.to_device(sc.ftris)
    cam = scene.camera_c2(seed=12, height=128, width=128)
    p = abi.run_params(4, sc.first_diffuse, math.nan, seed=8)
    lm = abi.lightmap_params(64, 64)
    base = api.Frame(cam, api.HIPBVHitter.build_on_device(tris, 6), tris, p, lm)
    base.trace()
    other = api.Frame(cam, api.HIPBVHitter.build_on_device(tris, 6, walk=split << 8), tris, p, lm)
    other.trace()
    torch.cuda.synchronize()
    assert torch.equal(base.hit_idx, other.hit_idx) and torch.equal(base.rays, other.rays)
    assert H.rel_l2(other.tex.cpu().numpy(), base.tex.cpu().numpy()) < 1e-4
