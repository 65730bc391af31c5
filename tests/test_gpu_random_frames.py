"""Seeded random frames against the oracle (needs an MI355X).

Each case draws a scene of one to three displaced icospheres (subdivision 0-3, random centres,
radii and vertex noise) over a ground quad; a pinhole camera of odd size (mostly not a multiple of
8: the column-major fallback and ragged tiles), 1-2 samples per pixel, fixed / uniform /
stratified wavelengths, with or without jitter; both IOR models and every compatibility flag
set, depth 0-5, seeded or forced random numbers; a light map of RGB curve weights, one mono
plane, or spectral planes with sub-pixel offsets; and one of the hitters -- brute force, the
reference's one-level BVs (with or without groups), or the device-built tree (every walk mode).
The frame's hit indices and final rays must equal the oracle's bit for bit and its light map must
be within 1e-4 relative L2 (only the atomic accumulation order differs).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from splat_amd import abi, api, bv, scene

pytestmark = pytest.mark.gpu


def _case(seed: int):
    rng = np.random.default_rng(seed)
    spheres = []
    for _ in range(int(rng.integers(1, 4))):
        c = (float(rng.uniform(-1.5, 1.5)), float(rng.uniform(-1.5, 1.5)), float(rng.uniform(-0.5, 0.8)))
        spheres.append(scene.sphere_mesh(int(rng.integers(0, 4)), center=c, radius=float(rng.uniform(0.4, 1.0)),
                                         noise_sigma=float(rng.choice([0.0, 0.02, 0.04])), rng=rng))
    sc = scene.assemble(spheres, [scene.ground_quad(-1.5, 4.0)], f"random-{seed}")
    lmode = int(rng.choice([abi.LAMBDA_FIXED, abi.LAMBDA_UNIFORM, abi.LAMBDA_STRATIFIED]))
    lam_fixed = float(rng.uniform(380.0, 720.0))
    cam = scene.pinhole((float(rng.uniform(-1, 1)), float(rng.uniform(-1, 1)), float(rng.uniform(4.0, 7.0))),
                        (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), float(rng.uniform(60.0, 100.0)),
                        int(rng.integers(17, 65)), int(rng.integers(17, 65)), spp=int(rng.integers(1, 3)),
                        lambda_mode=lmode, lambda_min=lam_fixed if lmode == abi.LAMBDA_FIXED else 400.0,
                        lambda_max=lam_fixed if lmode == abi.LAMBDA_FIXED else 700.0,
                        jitter=float(rng.choice([0.0, 0.025])), seed=int(rng.integers(1, 1 << 30)),
                        order=int(rng.choice([abi.ORDER_TILED8, abi.ORDER_COLUMN_MAJOR])))
    opt = abi.optics(int(rng.choice([abi.IOR_REFERENCE_LINEAR, abi.IOR_CAUCHY])), float(rng.uniform(1.3, 1.7)),
                     float(rng.uniform(0.002, 0.006)), int(rng.integers(0, 4)))
    force = float(rng.uniform(0.0, 1.0)) if rng.uniform() < 0.3 else math.nan
    p = abi.run_params(int(rng.integers(0, 6)), sc.first_diffuse, force, seed=int(rng.integers(1, 1 << 30)),
                       opt=opt)
    mode = str(rng.choice(["rgb", "mono", "spectral"]))
    size = int(rng.integers(32, 129))
    if mode == "spectral":
        spectrum = np.sort(rng.uniform(400.0, 700.0, int(rng.integers(3, 9)))).astype(np.float32)
        lm = abi.lightmap_params(size, size, n_delta=int(rng.integers(1, 3)), spectrum=spectrum, mode="spectral")
    elif mode == "rgb":
        lm = abi.lightmap_params(size, size, n_delta=int(rng.integers(1, 3)))
    else:
        lm = abi.lightmap_params(size, size, mode="mono")
    kind = str(rng.choice(["bf", "reference", "tree"]))
    return rng, sc, cam, p, lm, kind


@pytest.mark.parametrize("seed", range(48))
def test_random_frame_vs_oracle(seed):
    rng, sc, cam, p, lm, kind = _case(seed)
    tris_d = api.to_device(sc.ftris)
    if kind == "bf":
        hitter = api.HIPBruteForceHitter(tris_d)
    elif kind == "reference":
        bvs, mem = bv.bv_partition(sc.ftris, int(rng.integers(0, 6)))
        hitter = api.HIPBVHitter(tris_d, bvs, mem, group_size=int(rng.choice([0, -1, 3])))
    else:
        hitter = api.HIPBVHitter.build_on_device(tris_d, int(rng.integers(1, 7)), walk=int(rng.integers(0, 3)))
    fr = api.Frame(cam, hitter, tris_d, p, lm)
    fr.trace()
    import torch

    torch.cuda.synchronize()
    want_r, want_h, _ = O.run_evolution(O.raygen(cam), sc.ftris, p)
    want_tex, want_r2 = O.light_map(want_r, want_h, sc.ftris, lm)
    what = f"seed {seed}: {kind}, {sc.n_tris} triangles, {fr.R} rays, depth {p.depth}"
    assert (fr.hit_idx.cpu().numpy() == want_h).all(), what
    assert api.from_device(fr.rays, abi.ADRAY).tobytes() == want_r2.tobytes(), what
    got = fr.tex.cpu().numpy()
    assert got.shape == want_tex.shape, what
    nb = float(np.linalg.norm(want_tex.astype(np.float64)))
    err = float(np.linalg.norm(got.astype(np.float64) - want_tex.astype(np.float64)))
    assert err <= 1e-4 * nb or err <= 1e-7, (what, err, nb)
