"""Backward shading oracle (continuum_shade!, src/ray_imagers.jl:381-503) pinned by closed forms:
bilinear texture reads, the skydome's polar mapping, the retina-weighted light-map sum, the
[0, 1] clamp, and the reorder permutation."""
import math

import numpy as np

from oracle import oracle as O
from splat_amd import abi


def test_bilinear_closed_forms():
    H, W = 5, 7
    j, i = np.meshgrid(np.arange(1, H + 1), np.arange(1, W + 1), indexing="ij")
    ramp = (3.0 * i + 0.5 * j).astype(np.float32)  # T[i, j] = 3 i + j / 2 (1-based)
    for u, v in [(1.0, 1.0), (2.25, 3.5), (6.9, 4.1), (4.0, 2.0), (1.5, 4.75)]:
        assert abs(O.bilinear(ramp, u, v) - (3.0 * u + 0.5 * v)) < 1e-9  # exact on a linear field
    # integer coordinates hit texels exactly; outside the texture the edge is held (clamp)
    assert O.bilinear(ramp, 3.0, 2.0) == ramp[1, 2]
    assert O.bilinear(ramp, -4.0, 2.0) == ramp[1, 0]
    assert O.bilinear(ramp, 99.0, 99.0) == ramp[H - 1, W - 1]
    assert abs(O.bilinear(ramp, 0.3, 1.0) - ramp[0, 0]) < 1e-12


def _ray(status, direction, x=0.0, y=0.0, lam=550.0):
    r = np.zeros(1, dtype=abi.ADRAY)
    r["dir"] = direction
    r["status"] = status
    r["x"], r["y"], r["lam"] = x, y, lam
    r["dest"] = 1
    r["ignore_tri"] = 1
    return r


def _tri(uv=((0.1, 0.2), (0.6, 0.2), (0.1, 0.7))):
    """A diffuse triangle in z = 0 with given vertex uv (FTri: n, v1, v2, v3, n1, n2, n3, uv1-3)."""
    t = np.zeros(2, dtype=abi.FTRI)  # [sentinel, triangle]
    v = t["v"][1]
    v[0] = (0, 0, 1)
    v[1], v[2], v[3] = (0, 0, 0), (1, 0, 0), (0, 1, 0)
    v[4] = v[5] = v[6] = (0, 0, 1)
    for k in range(3):
        v[7 + k] = (uv[k][0], uv[k][1], 0)
    t["v"][1] = v
    return t


def test_skydome_polar_mapping():
    SW, SH = 8, 6
    sky = np.zeros((3, SH, SW), dtype=np.float32)
    jj, ii = np.meshgrid(np.arange(1, SH + 1), np.arange(1, SW + 1), indexing="ij")
    sky[0] = ii  # red = u index, green = v index, blue = 0.25
    sky[1] = jj
    sky[2] = 0.25
    sky /= 10.0
    lm = abi.lightmap_params(4, 4, spectrum=[550.0], mode="spectral")
    tris = _tri()
    tex = np.zeros((1, 4, 4), dtype=np.float32)
    for d in [(1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.3, -0.4, 0.866), (-0.6, 0.1, -0.79)]:
        r = _ray(abi.RAY_INFINITY, d)
        rgb, _ = O.shade(r, np.array([1]), tris, lm, tex, sky, 1, precision="f64")
        th, ph = math.acos(d[2]), math.atan2(d[1], d[0])
        u, v = th / math.pi * SW, (ph / math.pi + 1) * 0.5 * SH
        want = [min(max(O.bilinear(sky[c], u, v), 0.0), 1.0) for c in range(3)]
        assert np.allclose(rgb[0], want, atol=1e-6), (d, rgb[0], want)


def test_diffuse_retina_sum_and_clamp():
    spectrum = [450.0, 550.0, 650.0]
    lm = abi.lightmap_params(16, 16, spectrum=spectrum, mode="spectral")
    rf = abi.retina_factor(spectrum)
    tris = _tri()
    # a DIFFUSE ray sitting on the triangle (final_evolution moves it to the hit: start above it)
    r = _ray(abi.RAY_DIFFUSE, (0.0, 0.0, -1.0))
    r["pos"] = (0.25, 0.25, 1.0)
    for level in (0.3, 40.0):
        tex = np.full((3, 16, 16), level, dtype=np.float32)
        rgb, fe = O.shade(r, np.array([2]), tris, lm, tex, None, 1, precision="f64")
        want = np.clip(level * rf.astype(np.float64).sum(axis=0), 0.0, 1.0)
        assert np.allclose(rgb[0], want, rtol=1e-6, atol=1e-9)
    # a linear light map: the value at the ray's texture coordinate (uv of the hit point)
    jj, ii = np.meshgrid(np.arange(1, 17), np.arange(1, 17), indexing="ij")
    tex = np.stack([(ii + 2.0 * jj).astype(np.float32)] * 3) * 1e-3
    rgb, fe = O.shade(r, np.array([2]), tris, lm, tex, None, 1, precision="f32")
    p = fe["pos"][0]
    assert np.allclose(p, (0.25, 0.25, 0.0), atol=1e-6)  # final_evolution moved it onto the triangle
    U = 0.1 + 0.5 * 0.25  # uv is affine in the (x, y) of the triangle
    V = 0.2 + 0.5 * 0.25
    X = (U * 16 + 2.0 * V * 16) * 1e-3
    assert np.allclose(rgb[0], X * rf.sum(axis=0), rtol=1e-5)


def test_active_rays_are_black_and_reorder_is_the_dest_permutation():
    lm = abi.lightmap_params(4, 4, spectrum=[550.0], mode="spectral")
    r = _ray(abi.RAY_ACTIVE, (0.0, 0.0, 1.0))
    rgb, _ = O.shade(r, np.array([1]), _tri(), lm, np.ones((1, 4, 4), np.float32), None, 1)
    assert (rgb == 0).all()
    rng = np.random.default_rng(3)
    R = 97
    rays = np.zeros(R, dtype=abi.ADRAY)
    perm = rng.permutation(R)
    rays["dest"] = perm + 1
    rays["lam"] = np.arange(R)
    hit = np.arange(R, dtype=np.int32) + 10
    out, h = O.reorder(rays, hit)
    assert (out["dest"] == np.arange(1, R + 1)).all()
    assert (out["lam"][perm] == np.arange(R)).all() and (h[perm] == hit).all()
