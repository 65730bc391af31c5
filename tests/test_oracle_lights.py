"""Pins the oracle's rays_from_lights restatement (src/ray_generators.jl:64-113).

The checks come from what the reference's definition fixes, not from its formula retyped:
  * ordering: ray j of a light is cell (j % h + 1, j // h + 1) (broadcast over (1:h, (1:w)')
    reshaped column-major), lights concatenated in order;
  * geometry: every origin lies in the plane through `center` spanned by dim1, dim2; with zero
    offsets the cell (h/2, w/2) sits exactly on `center`, and the grid covers the coefficient
    range [(1 - h/2)/(h/2), (h + 1 - h/2)/(h/2)) (the reference's grid is not centred);
  * derivatives: the origin is affine in (x, y), so the difference of neighbouring cells equals
    d origin/dx = dim1 / (h/2) (and likewise for y) up to rounding -- a finite difference;
  * offsets: x - floor(x) lies in [0, jitter] (Float32 rounding of x + dx), mean jitter / 2;
  * the fixed fields: dir = light.dir, dest 0, ignore_tri 1, ACTIVE, 550 nm by default.
"""
import numpy as np
import pytest

from oracle import oracle as O
from splat_amd import abi


def light(res1=8, res2=6, center=(0.5, -1.0, 8.0), dim1=(1.0, 0.0, 0.0), dim2=(0.0, 1.5, 0.25), d=(0.0, 0.0, -1.0)):
    return abi.rect_light(center, d, dim1, dim2, res1, res2)


def vec(rays, name):
    return np.asarray(rays[name], dtype=np.float64)


def test_order_and_fixed_fields():
    L = light()
    rays = O.raygen_lights([L], abi.light_rays(1, jitter=0.0))
    assert rays.shape[0] == 48
    j = np.arange(48)
    np.testing.assert_array_equal(rays["x"], (j % 8 + 1).astype(np.float32))
    np.testing.assert_array_equal(rays["y"], (j // 8 + 1).astype(np.float32))
    np.testing.assert_array_equal(rays["dir"], np.tile(np.float32([0.0, 0.0, -1.0]), (48, 1)))
    assert (rays["dest"] == 0).all() and (rays["ignore_tri"] == 1).all()
    assert (rays["status"] == abi.RAY_ACTIVE).all() and (rays["in_medium"] == 0).all()
    assert (rays["lam"] == np.float32(550.0)).all()
    for f in ("dpdl", "dddl", "dddx", "dddy"):
        assert (rays[f] == 0).all()


def test_centre_cell_and_plane():
    L = light(res1=8, res2=6)
    rays = O.raygen_lights([L], abi.light_rays(1, jitter=0.0))
    # cell (4, 3) = (h/2, w/2): both coefficients are exactly zero
    k = (3 - 1) * 8 + (4 - 1)
    np.testing.assert_array_equal(rays["pos"][k], np.float32([0.5, -1.0, 8.0]))
    nrm = np.cross(np.float64(L.dim1[:]), np.float64(L.dim2[:]))
    off = vec(rays, "pos") - np.float64(L.center[:])
    assert np.abs(off @ nrm).max() < 1e-5


def test_grid_covers_reference_range():
    # coefficient of dim1 = (x - h/2) / (h/2) for x in 1..h: min (1 - 4)/4, max (8 - 4)/4
    L = light(res1=8, res2=6, dim1=(2.0, 0.0, 0.0), dim2=(0.0, 0.0, 0.0), center=(0.0, 0.0, 0.0))
    rays = O.raygen_lights([L], abi.light_rays(1, jitter=0.0))
    cx = vec(rays, "pos")[:, 0] / 2.0
    assert cx.min() == pytest.approx(-0.75) and cx.max() == pytest.approx(1.0)
    # with offsets the coefficient stays inside [(1 - h/2)/(h/2), (h + 1 - h/2)/(h/2))
    rays = O.raygen_lights([L], abi.light_rays(1, jitter=1.0, seed=7))
    cx = vec(rays, "pos")[:, 0] / 2.0
    assert cx.min() >= -0.75 and cx.max() <= 1.25


def test_neighbour_difference_is_the_derivative():
    L = light(res1=16, res2=10, dim1=(0.3, -0.7, 0.2), dim2=(1.1, 0.4, -0.5))
    rays = O.raygen_lights([L], abi.light_rays(1, jitter=0.0))
    pos = vec(rays, "pos").reshape(10, 16, 3)  # [y-1, x-1]
    fdx = pos[:, 1:, :] - pos[:, :-1, :]
    fdy = pos[1:, :, :] - pos[:-1, :, :]
    dpdx = vec(rays, "dpdx").reshape(10, 16, 3)
    dpdy = vec(rays, "dpdy").reshape(10, 16, 3)
    np.testing.assert_allclose(fdx, dpdx[:, 1:, :], atol=2e-6)
    np.testing.assert_allclose(fdy, dpdy[1:, :, :], atol=2e-6)
    # the tangents are the rectangle's half extents over the integer half sizes
    np.testing.assert_allclose(dpdx[0, 0], np.float64(L.dim1[:]) / 8, rtol=1e-7)
    np.testing.assert_allclose(dpdy[0, 0], np.float64(L.dim2[:]) / 5, rtol=1e-7)


def test_offsets_in_cell():
    L = light(res1=64, res2=64)
    rays = O.raygen_lights([L], abi.light_rays(1, jitter=0.5, seed=3))
    fx = vec(rays, "x") - np.floor(vec(rays, "x"))
    fy = vec(rays, "y") - np.floor(vec(rays, "y"))
    # x + dx is rounded to Float32, so an offset just below 0.5 may round up onto it
    assert fx.min() >= 0 and fx.max() <= 0.5 and fy.min() >= 0 and fy.max() <= 0.5
    assert abs(fx.mean() - 0.25) < 0.01 and abs(fy.mean() - 0.25) < 0.01
    # x and y offsets are independent streams
    assert abs(np.corrcoef(fx, fy)[0, 1]) < 0.05


def test_upscale_and_concatenation():
    A = light(res1=8, res2=8)
    B = light(res1=4, res2=12, center=(-3.0, 2.0, 5.0), d=(0.6, 0.0, -0.8))
    p = abi.light_rays(2, jitter=0.0)
    both = O.raygen_lights([A, B], p)
    assert both.shape[0] == 16 + 12
    np.testing.assert_array_equal(both[:16], O.raygen_lights([A], p))
    np.testing.assert_array_equal(both[16:], O.raygen_lights([B], p))
    # upscale 2: a 4 x 4 grid, so the step in x is dim1 / 2
    np.testing.assert_allclose(vec(both[:16], "dpdx")[0], np.float64(A.dim1[:]) / 2)


def test_uniform_lambda():
    rays = O.raygen_lights([light(res1=32, res2=32)],
                           abi.light_rays(1, lambda_mode=abi.LAMBDA_UNIFORM, lambda_min=400.0, lambda_max=700.0, seed=5))
    lam = vec(rays, "lam")
    assert lam.min() >= 400.0 and lam.max() < 700.0 and abs(lam.mean() - 550.0) < 10.0


def test_seed_changes_offsets_only():
    L = light()
    a = O.raygen_lights([L], abi.light_rays(1, seed=1))
    b = O.raygen_lights([L], abi.light_rays(1, seed=2))
    assert not np.array_equal(a["x"], b["x"])
    np.testing.assert_array_equal(np.floor(a["x"]), np.floor(b["x"]))
    np.testing.assert_array_equal(a["dpdx"], b["dpdx"])


def test_odd_rows_use_integer_half():
    # h = 7: h/2 is 3 in integer division (not 3.5), so the step in x is dim1 / 3 and the
    # cell x = 3 sits on the centre
    L = light(res1=7, res2=4, dim1=(1.0, 0.0, 0.0), dim2=(0.0, 1.0, 0.0), center=(0.0, 0.0, 0.0))
    rays = O.raygen_lights([L], abi.light_rays(1, jitter=0.0))
    np.testing.assert_allclose(vec(rays, "dpdx")[0], [1.0 / 3.0, 0.0, 0.0], rtol=1e-7)
    k = (2 - 1) * 7 + (3 - 1)
    assert rays["pos"][k][0] == 0.0 and rays["pos"][k][1] == 0.0
