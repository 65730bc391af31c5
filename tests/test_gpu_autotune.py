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
