"""The oracle still reproduces the committed golden vectors of the light-map path
(tests/golden/make_golden.py); the GPU path reproduces them too (marked gpu)."""
import hashlib
import math
import os

import numpy as np
import pytest

from tests.golden import make_golden as G

GOLD = np.load(G.OUT)


def test_oracle_reproduces_golden():
    now = G.compute()
    for k in GOLD.files:
        assert np.array_equal(now[k], GOLD[k]), k


def test_golden_is_nontrivial():
    assert (GOLD["hit_primary"] > 1).mean() > 0.5
    assert GOLD["d4_seeded_tex"].sum() > 0 and GOLD["d2_refract_tex"].sum() > 0
    st = GOLD["d2_refract_status"]  # after 2 bounces the rays through the glass are still live
    assert (st == 0).sum() > 1000 and (st == 1).sum() > 500
    assert (GOLD["d4_seeded_status"] == 0).sum() > 4000


@pytest.mark.gpu
def test_gpu_reproduces_golden():
    import torch

    from splat_amd import abi, api, scene

    sc = scene.scene_c1()
    cam = scene.camera_c1(seed=11, height=64, width=64)
    tris = api.to_device(sc.ftris)
    hitter = api.HIPBruteForceHitter(tris)
    for name, depth, fr in (("d2_refract", 2, 1.0), ("d4_seeded", 4, math.nan)):
        p = abi.run_params(depth, sc.first_diffuse, fr, seed=77)
        lm = abi.lightmap_params(64, 64)
        fr_ = api.Frame(cam, hitter, tris, p, lm)
        fr_.trace()
        torch.cuda.synchronize()
        assert (fr_.hit_idx.cpu().numpy() == GOLD[f"{name}_hit"]).all()
        rays = api.from_device(fr_.rays, abi.ADRAY)
        assert np.frombuffer(hashlib.sha256(rays.tobytes()).digest(), dtype=np.uint8).tolist() == \
            GOLD[f"{name}_rays_sha256"].tolist()
        tex = fr_.tex.cpu().numpy().astype(np.float64)
        want = GOLD[f"{name}_tex"].astype(np.float64)
        assert np.linalg.norm(tex - want) / np.linalg.norm(want) < 1e-4
