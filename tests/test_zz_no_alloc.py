"""splat_next_hit_noalloc (the entry point that allocates nothing) against splat_next_hit, through
the C ABI: the same hit indices, bit for bit, on consecutive calls that share the library's
device scratch -- the second call smaller than the first, and one limited by a device live count
-- and refusal of ray counts above its fixed capacity."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from splat_amd import abi, api, bv, scene
from splat_amd import lib as L

pytestmark = pytest.mark.gpu


def _hits(fn_noalloc: bool, rays_d, hitter, d_count=None):
    import torch

    R = rays_d.shape[0]
    out = torch.full((R,), -7, dtype=torch.int32, device="cuda")
    lib = L.lib()
    if fn_noalloc:
        code = lib.splat_next_hit_noalloc(L.ptr(rays_d), R, L.ptr(d_count), L.ref(hitter.hs), L.ptr(out),
                                          L.stream_handle())
    else:
        ws = torch.empty(lib.splat_next_hit_ws_bytes(R), dtype=torch.uint8, device="cuda")
        code = lib.splat_next_hit(L.ptr(rays_d), R, L.ptr(d_count), L.ref(hitter.hs), L.ptr(out), L.ptr(ws),
                                  ws.numel(), L.stream_handle())
    L.check(code, "next_hit")
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.fixture(scope="module")
def c1_bv():
    sc = scene.scene_c1()
    tris_d = api.to_device(sc.ftris)
    bvs, mem, n_top = bv.tree_partition(sc.ftris, 4)
    return sc, api.HIPBVHitter(tris_d, bvs, mem, n_top=n_top), api.HIPBruteForceHitter(api.to_device(sc.ntris()))


def test_noalloc_equals_next_hit_on_consecutive_calls(c1_bv):
    import torch

    sc, hitter, _ = c1_bv
    big = O.raygen(scene.camera_c1(seed=21, height=192, width=160))  # 30,720 rays
    small = O.raygen(scene.camera_c1(seed=22, height=40, width=24))  # 960 rays, other rays
    for rays in (big, small):
        print(f"noalloc BV hitter: {rays.size} rays", flush=True)
        rays_d = api.to_device(rays)
        want = _hits(False, rays_d, hitter)
        got = _hits(True, rays_d, hitter)
        assert np.array_equal(got, want)
        ref, _ = O.next_hit(rays, sc.ntris())
        assert np.array_equal(got, ref)
    # a device live count: only the first 700 rays are hit, the rest of hit_idx is untouched
    rays_d = api.to_device(big)
    live = torch.tensor([700], dtype=torch.int64, device="cuda")
    want = _hits(False, rays_d, hitter, live)
    got = _hits(True, rays_d, hitter, live)
    assert np.array_equal(got, want) and (got[700:] == -7).all() and (got[:700] > 0).all()


def test_noalloc_brute_force_needs_no_scratch(c1_bv):
    _, _, bf = c1_bv
    rays = O.raygen(scene.camera_c1(seed=23, height=48, width=40))
    rays_d = api.to_device(rays)
    assert np.array_equal(_hits(True, rays_d, bf), _hits(False, rays_d, bf))


def test_noalloc_refuses_more_rays_than_its_scratch(c1_bv):
    _, hitter, _ = c1_bv
    lib = L.lib()
    cap = abi.NOALLOC_MAX_RAYS
    code = lib.splat_next_hit_noalloc(C.c_void_p(8), cap + 1, None, L.ref(hitter.hs), C.c_void_p(8), None)
    assert code == abi.SPLAT_ERR_CAPACITY
