"""The BV structure built on the device (splat_hitset_build_bv) against the host build.

Bars: the device build's member records, leaf boxes, ranges and tree nodes are byte-identical to
the host bv.tree_partition + splat_hitset_prepare_bv of the same depth (the large-triangle list
leading, possibly empty); hits through a device-built hitter equal the brute-force hitter's;
a rebuild in place for moved triangles equals a fresh host build.
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


def _bytes(hitter, ptr, nbytes):
    off = ptr - hitter.ws.data_ptr()
    assert 0 <= off and off + nbytes <= hitter.ws.numel()
    return hitter.ws[off:off + nbytes].cpu().numpy().tobytes()


def _host_hitter(tris_d, ftris, depth):
    bvs, mem, n_top = bv.tree_partition(ftris, depth)
    if n_top == 0:  # the device build always leads with the (here empty) large-triangle list
        mem = [np.zeros(0, dtype=np.int64)] + mem
        bvs = np.concatenate([np.zeros(1, dtype=abi.BVBOX), bvs])
    return api.HIPBVHitter(tris_d, bvs, mem, n_top=1), len(mem) - 1


def _menger_scene():
    cubes, edge = scene.menger_cubes(2)  # 400 cubes on a grid: many equal centroid coordinates
    return scene.assemble([scene.cube_mesh(cubes, edge)], [scene.ground_quad(-1.5, 3.0)], "menger2+quad")


def _scenes():
    return {"c1": scene.scene_c1(), "small_c2": H.small_c2(2), "menger2": _menger_scene()}


@pytest.mark.parametrize("name,depth", [("c1", 5), ("small_c2", 6), ("small_c2", 8), ("menger2", 7), ("menger2", 9)])
def test_device_build_equals_host_build(name, depth):
    sc = _scenes()[name]
    tris_d = api.to_device(sc.ftris)
    host, n_leaves = _host_hitter(tris_d, sc.ftris, depth)
    assert n_leaves == 2 ** depth, "pick a depth without empty leaves for the byte comparison"
    dev = api.HIPBVHitter.build_on_device(tris_d, depth)
    _sync()
    a, b = host.hs, dev.hs
    for f in ("kind", "n_tris", "n_bv", "total_members", "tree_levels", "top_bvs", "top_recs", "max_members"):
        assert getattr(a, f) == getattr(b, f), f
    n, n_bv = a.total_members, a.n_bv
    assert _bytes(host, a.d_recs, n * 64) == _bytes(dev, b.d_recs, n * 64)
    assert _bytes(host, a.d_boxes, n_bv * 32) == _bytes(dev, b.d_boxes, n_bv * 32)
    assert _bytes(host, a.d_ranges, n_bv * 8) == _bytes(dev, b.d_ranges, n_bv * 8)
    nodes = 2 << (a.tree_levels - 1)  # heap nodes 1 .. 2^(levels+1) - 1, 32 bytes each
    assert _bytes(host, a.d_groups + 32, (nodes - 1) * 32) == _bytes(dev, b.d_groups + 32, (nodes - 1) * 32)


@pytest.mark.parametrize("name,depth,n_pts", [("c1", 9, 3000), ("small_c2", 4, 3000), ("menger2", 12, 3000),
                                              ("c1", 9, 6000), ("menger2", 12, 6000), ("menger2", 15, 6000)])
def test_device_built_hits_equal_brute_force(name, depth, n_pts):
    """Deep builds leave empty leaves (c1 at depth 9: 320 triangles over 512 leaves). 3000 rays
    (47 groups of 64) take the flat scan of few live rays, 6000 the tree walk."""
    import torch

    sc = _scenes()[name]
    tris_d = api.to_device(sc.ftris)
    dev = api.HIPBVHitter.build_on_device(tris_d, depth)
    pts = H.sample_surface_points(sc.ftris, n_pts, seed=depth)
    rays_np = H.rays_towards(pts, seed=depth + 1)
    rays = api.to_device(rays_np)
    tr = api.HIPTracer(rays)
    api.next_hit_(tr, dev, rays)
    bf = api.HIPBruteForceHitter(api.to_device(sc.ntris()))
    tr2 = api.HIPTracer(rays)
    api.next_hit_(tr2, bf, rays)
    _sync()
    assert torch.equal(tr.hit_idx, tr2.hit_idx)
    assert (tr.hit_idx > 1).sum().item() > 1000


def test_rebuild_in_place_for_moved_triangles():
    """An animated scene rebuilds every frame (src/scenes/splash_scene.jl:41): rebuild() on the
    same workspace equals a fresh host build of the moved scene, and the frame traced with it
    equals the oracle's."""
    sc = H.small_c2(2)
    tris_d = api.to_device(sc.ftris)
    dev = api.HIPBVHitter.build_on_device(tris_d, 6)
    moved = scene.rotate_ftris(sc.ftris, axis=(0.0, 0.0, 1.0), angle=0.7)
    tris_m = api.to_device(moved)
    dev.rebuild(tris_m)
    host, _ = _host_hitter(tris_m, moved, 6)
    _sync()
    n, n_bv = host.hs.total_members, host.hs.n_bv
    assert dev.hs.top_recs == host.hs.top_recs
    assert _bytes(host, host.hs.d_recs, n * 64) == _bytes(dev, dev.hs.d_recs, n * 64)
    assert _bytes(host, host.hs.d_ranges, n_bv * 8) == _bytes(dev, dev.hs.d_ranges, n_bv * 8)
    cam = scene.camera_c2(seed=6, height=64, width=64)
    p = abi.run_params(3, sc.first_diffuse, math.nan, seed=5)
    lm = abi.lightmap_params(64, 64)
    fr = api.Frame(cam, dev, tris_m, p, lm)
    fr.trace()
    _sync()
    want_r, want_h, _ = O.run_evolution(O.raygen(cam), moved, p)
    want_tex, want_r2 = O.light_map(want_r, want_h, moved, lm)
    assert (fr.hit_idx.cpu().numpy() == want_h).all()
    assert api.from_device(fr.rays, abi.ADRAY).tobytes() == want_r2.tobytes()
    assert H.rel_l2(fr.tex.cpu().numpy(), want_tex) < 1e-4
