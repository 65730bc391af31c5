"""The GPU light map against the independent pins of tests/test_oracle_pins.py (needs an MI355X):
retina-curve weights at single splats, and splats expanded at non-zero tangents with sub-pixel
offsets and several wavelengths, against values built from first principles and the FP64 oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from splat_amd import abi, api, scene
from tests import helpers as H
from tests.test_oracle_pins import (CHANNELS, SPLAT_CASES, _down_rays, _norm, _tangent_rays, _tilted_triangle,
                                    expected_map, reference_piecewise)

pytestmark = pytest.mark.gpu


def _gpu_light_map(rays, hit, sc, lm, n_delta=1):
    import torch

    rays_d = api.to_device(rays)
    tr = api.HIPTracer(rays_d, N=n_delta)
    tr.hit_idx.copy_(torch.from_numpy(np.ascontiguousarray(hit, dtype=np.int32)))
    tex = torch.zeros((lm.channels, lm.height, lm.width), dtype=torch.float32, device="cuda")
    api.continuum_light_map2_(tracer=tr, tris=api.to_device(sc.ftris), rays=rays_d, tex=tex, lm=lm)
    torch.cuda.synchronize()
    return tex.cpu().numpy()


def test_gpu_curve_weights_are_the_reference_curves():
    half = 2.0
    sc = scene.assemble([], [scene.ground_quad(0.0, half)], "ground")
    lams = np.concatenate([np.arange(400, 701), np.arange(400.5, 700, 7.0), [399.25, 452.75, 600.0625, 699.5]])
    rays, i, j = _down_rays(lams, 32, 32, half)
    hit, _ = O.next_hit(rays, sc.ntris())
    lm = abi.lightmap_params(32, 32)
    got = _gpu_light_map(rays, hit, sc, lm)
    want32, _ = O.light_map(rays, hit, sc.ftris, lm)
    assert got.tobytes() == want32.tobytes()  # one splat per texel: no accumulation order
    for c, name in enumerate(CHANNELS):
        f = reference_piecewise(abi._RETINA[name])
        want = np.array([f(float(x)) for x in rays["lam"]]) / _norm(name)
        assert np.allclose(got[c, j, i], want, rtol=1e-6, atol=1e-9), name
    assert abs(got[:, j[:301], i[:301]].sum(axis=1) - 1.0).max() < 1e-5  # the 400..700 sums


@pytest.mark.parametrize("mode,spectrum,n_delta", SPLAT_CASES)
def test_gpu_splat_expansion_at_nonzero_tangents(mode, spectrum, n_delta):
    sc = _tilted_triangle()
    rays = _tangent_rays(sc, 40, seed=3)
    hit, _ = O.next_hit(rays, sc.ntris())
    lm = abi.lightmap_params(64, 64, n_delta=n_delta, spectrum=spectrum, mode=mode)
    got = _gpu_light_map(rays, hit, sc, lm, n_delta)
    delta = abi.upres(n_delta).astype(np.float64)
    if spectrum is None:
        want = np.zeros(got.shape)
        for q in range(rays.shape[0]):
            want += expected_map(rays[q:q + 1], sc, lm, [float(rays["lam"][q])], delta)[0]
    else:
        want = expected_map(rays, sc, lm, spectrum.astype(np.float64), delta)[0]
    assert np.abs(got - want).max() < 1e-5 * np.abs(want).max()
    tex64, _ = O.light_map(rays, hit, sc.ftris, lm, precision="f64")
    assert H.rel_l2(got, tex64) < 1e-5
