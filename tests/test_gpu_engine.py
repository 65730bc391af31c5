"""The multi-frame engine (batching, continuous admission, HBM budget) against the plain
single-frame path and the oracle, through the C ABI (needs an MI355X). Each frame's rays and hit
indices must be bit-identical to tracing it alone; light maps within 1e-4 relative L2 (only the
atomic accumulation order differs)."""
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


@pytest.fixture(scope="module")
def scenes():
    out = []
    for k, (sc, depth) in enumerate(((scene.scene_c1(), 3), (H.small_c2(2), 4), (H.small_c2(3), 5))):
        tris = api.to_device(sc.ftris)
        if k == 1:  # ground triangles tested directly, tree over the rest
            bvs, mem, n_top = bv.tree_partition(sc.ftris, depth)
        else:
            (bvs, mem), n_top = bv.object_partition(sc.ftris, depth), 0
        out.append((sc, tris, api.HIPBVHitter(tris, bvs, mem, group_size=0, n_top=n_top)))
    return out


def _frames(n, seed):
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(n):
        si = int(rng.integers(0, 3))
        h, w = [(32, 32), (64, 48), (48, 64), (96, 96), (40, 24)][int(rng.integers(0, 5))]
        depth = int(rng.integers(0, 6))
        fr = [math.nan, 1.0, 0.0][int(rng.integers(0, 3))]
        mode = ["rgb", "mono"][int(rng.integers(0, 2))]
        specs.append(dict(si=si, h=h, w=w, depth=depth, fr=fr, seed=int(rng.integers(1, 1 << 30)), mode=mode,
                          nd=int(rng.integers(1, 3)), lmw=int(rng.integers(40, 120)), lmh=int(rng.integers(40, 120))))
    return specs


def _make(spec, scenes):
    import torch

    sc, tris, hitter = scenes[spec["si"]]
    cam = scene.camera_c2(seed=spec["seed"], height=spec["h"], width=spec["w"], jitter=0.025)
    p = abi.run_params(spec["depth"], sc.first_diffuse, spec["fr"], seed=spec["seed"])
    lm = abi.lightmap_params(spec["lmw"], spec["lmh"], n_delta=spec["nd"], mode=spec["mode"])
    R = spec["h"] * spec["w"]
    out = dict(sc=sc, tris=tris, hitter=hitter, cam=cam, p=p, lm=lm,
               tex=torch.full((lm.channels, lm.height, lm.width), 7.0, dtype=torch.float32, device="cuda"),
               rays=torch.empty((R, 136), dtype=torch.uint8, device="cuda"),
               hit=torch.empty(R, dtype=torch.int32, device="cuda"))
    return out


def _plain(f):
    fr = api.Frame(f["cam"], f["hitter"], f["tris"], f["p"], f["lm"])
    fr.trace()
    _sync()
    return fr.tex.cpu().numpy(), api.from_device(fr.rays, abi.ADRAY), fr.hit_idx.cpu().numpy()


def _check(f, want):
    tex, rays, hit = want
    assert (f["hit"].cpu().numpy() == hit).all()
    assert api.from_device(f["rays"], abi.ADRAY).tobytes() == rays.tobytes()
    got = f["tex"].cpu().numpy()
    if np.abs(tex).sum() == 0:
        assert np.abs(got).sum() == 0
    else:
        assert H.rel_l2(got, tex) < 1e-4


def _admit(eng, f, tag):
    return eng.admit(cam=f["cam"], hitter=f["hitter"], tris=f["tris"], params=f["p"], lm=f["lm"], tex=f["tex"],
                     tag=tag, rays_out=f["rays"], hit_out=f["hit"])


@pytest.mark.parametrize("use_graph,use_queue", [(True, True), (False, True), (True, False)])
def test_engine_batch_equals_plain(scenes, use_graph, use_queue):
    """Graph-replayed or stream-launched steps; the hitter as one persistent launch taking its items
    from the device work queues, or one wave per item over a host-sized grid."""
    specs = _frames(12, seed=1)
    frames = [_make(s, scenes) for s in specs]
    want = [_plain(f) for f in frames]
    eng = api.Engine(1 << 30, max_slots=16, use_graph=use_graph, use_queue=use_queue)
    for k, f in enumerate(frames):
        assert _admit(eng, f, 100 + k)
    done, steps = [], 0
    while len(done) < len(frames):
        eng.step()
        steps += 1
        done += eng.poll(wait=True)
        assert steps <= 20
    assert steps == max(s["depth"] for s in specs) + 1
    assert sorted(done) == [100 + k for k in range(len(frames))]
    for f, w in zip(frames, want):
        _check(f, w)
    eng.close()


def test_engine_many_slots(scenes):
    """More slots than one wave's 64 lanes: a work item's slot lookup counts the prefixes 64 at a
    time, so items of slots past 64 take the lookup's later passes; every frame still equals its
    single-frame trace."""
    specs = _frames(80, seed=7)
    frames = [_make(s, scenes) for s in specs]
    want = [_plain(f) for f in frames]
    eng = api.Engine(1 << 30, max_slots=150)
    for k, f in enumerate(frames):
        assert _admit(eng, f, 500 + k)
    done, steps = [], 0
    while len(done) < len(frames):
        eng.step()
        steps += 1
        done += eng.poll(wait=True)
        assert steps <= 20
    assert sorted(done) == [500 + k for k in range(len(frames))]
    for f, w in zip(frames, want):
        _check(f, w)
    eng.close()


@pytest.mark.parametrize("use_queue", [True, False])
def test_engine_many_flat_tails_beside_a_split_walk(scenes, use_queue):
    """188 frames of 64 ray groups each (each small enough to scan the leaves flat) in one step with
    a 700-group frame whose tree walk is split 32 ways: 22,400 walk items plus 12,032 flat items,
    more than the hitter grid held before the flat-item bound counted the tail frames' groups (and
    more than the queue hitter's waves: each takes several). Every frame still equals its
    single-frame trace."""
    rng = np.random.default_rng(11)
    specs = [dict(si=1, h=224, w=200, depth=3, fr=math.nan, seed=901, mode="rgb", nd=1, lmw=96, lmh=80)]
    for k in range(188):
        specs.append(dict(si=int(rng.integers(0, 3)), h=64, w=64, depth=int(rng.integers(1, 3)), fr=math.nan,
                          seed=int(rng.integers(1, 1 << 30)), mode="mono", nd=1, lmw=48, lmh=40))
    frames = [_make(s, scenes) for s in specs]
    want = [_plain(f) for f in frames]
    eng = api.Engine(1 << 30, max_slots=256, use_queue=use_queue)
    for k, f in enumerate(frames):
        assert _admit(eng, f, 700 + k)
    done, steps = [], 0
    while len(done) < len(frames):
        eng.step()
        steps += 1
        done += eng.poll(wait=True)
        assert steps <= 10
    assert sorted(done) == [700 + k for k in range(len(frames))]
    for f, w in zip(frames, want):
        _check(f, w)
    eng.close()


def test_engine_hitter_mode_switch_between_steps(scenes):
    """The hitter mode switched between the steps of one engine (graphs are kept per mode): the
    188 small frames and the split walk of the test above, the first step on the work queues
    (more items than the persistent grid's waves), the next without, then back; every frame equals
    its single-frame trace."""
    rng = np.random.default_rng(13)
    specs = [dict(si=1, h=224, w=200, depth=3, fr=math.nan, seed=902, mode="rgb", nd=1, lmw=96, lmh=80)]
    for k in range(188):
        specs.append(dict(si=int(rng.integers(0, 3)), h=64, w=64, depth=int(rng.integers(1, 4)), fr=math.nan,
                          seed=int(rng.integers(1, 1 << 30)), mode="mono", nd=1, lmw=48, lmh=40))
    frames = [_make(s, scenes) for s in specs]
    want = [_plain(f) for f in frames]
    eng = api.Engine(1 << 30, max_slots=256, use_queue=True)
    for k, f in enumerate(frames):
        assert _admit(eng, f, 900 + k)
    done, steps = [], 0
    while len(done) < len(frames):
        eng.set_use_queue(steps % 2 == 0)
        eng.step()
        steps += 1
        done += eng.poll(wait=True)
        assert steps <= 10
    assert steps >= 3 and sorted(done) == [900 + k for k in range(len(frames))]
    for f, w in zip(frames, want):
        _check(f, w)
    eng.close()


def test_engine_continuous_admission_reuses_slots(scenes):
    """Frames arrive one per step while others are mid-flight; 3 slots serve 10 frames."""
    specs = _frames(10, seed=2)
    frames = [_make(s, scenes) for s in specs]
    want = [_plain(f) for f in frames]
    eng = api.Engine(1 << 30, max_slots=3)
    queue = list(range(len(frames)))
    done = []
    for _ in range(200):
        if queue and _admit(eng, frames[queue[0]], queue[0]):
            queue.pop(0)
        if eng.runnable:
            eng.step()
        done += eng.poll(wait=eng.runnable == 0)  # nothing left to launch: wait for completions
        if not queue and eng.in_flight == 0:
            break
    done += eng.poll(wait=True)
    assert sorted(done) == list(range(len(frames)))
    for f, w in zip(frames, want):
        _check(f, w)
    eng.close()


def test_engine_hbm_budget_admission(scenes):
    specs = [dict(si=1, h=64, w=64, depth=3, fr=1.0, seed=s, mode="rgb", nd=1, lmw=64, lmh=64) for s in (3, 4, 5)]
    frames = [_make(s, scenes) for s in specs]
    probe = api.Engine(1 << 20, max_slots=4)
    need = probe.frame_bytes(frames[0]["cam"])
    probe.close()
    eng = api.Engine(2 * need + 4096, max_slots=4)
    assert _admit(eng, frames[0], 0) and _admit(eng, frames[1], 1)
    assert not _admit(eng, frames[2], 2)  # over budget: refused, nothing changed
    assert eng.in_flight == 2
    while eng.in_flight:
        eng.step()
        eng.poll(wait=True)
    assert _admit(eng, frames[2], 2)  # space came back
    while eng.in_flight:
        eng.step()
        eng.poll(wait=True)
    for f in frames:
        _check(f, _plain(f))
    eng.close()


def test_engine_frame_matches_oracle(scenes):
    sc, tris, hitter = scenes[0]
    spec = dict(si=0, h=48, w=40, depth=3, fr=math.nan, seed=9, mode="rgb", nd=2, lmw=64, lmh=48)
    f = _make(spec, scenes)
    eng = api.Engine(1 << 28, max_slots=2)
    assert _admit(eng, f, 1)
    while eng.in_flight:
        eng.step()
        eng.poll(wait=True)
    rays, hit, _ = O.run_evolution(O.raygen(f["cam"]), sc.ftris, f["p"])
    tex, rays = O.light_map(rays, hit, sc.ftris, f["lm"])
    assert (f["hit"].cpu().numpy() == hit).all()
    assert api.from_device(f["rays"], abi.ADRAY).tobytes() == rays.tobytes()
    assert tex.sum() > 0 and H.rel_l2(f["tex"].cpu().numpy(), tex) < 1e-4
    eng.close()


def test_engine_admission_checks_the_light_map(scenes):
    """Admission applies the C ABI's light-map check: curve point counts within SPLAT_MAX_CURVE
    and ascending abscissae (the kernels read the curves by those counts)."""
    from splat_amd import lib as SL

    f = _make(dict(si=1, h=32, w=32, depth=2, fr=1.0, seed=3, mode="rgb", nd=1, lmw=40, lmh=40), scenes)
    eng = api.Engine(1 << 28, max_slots=2)
    lm_bad_n = abi.lightmap_params(40, 40)
    lm_bad_n.curve_n[0] = abi.MAX_CURVE + 1
    lm_bad_x = abi.lightmap_params(40, 40)
    lm_bad_x.curve_x[1][3], lm_bad_x.curve_x[1][4] = lm_bad_x.curve_x[1][4], lm_bad_x.curve_x[1][3]
    for lm in (lm_bad_n, lm_bad_x):
        with pytest.raises(SL.SplatError) as err:
            eng.admit(cam=f["cam"], hitter=f["hitter"], tris=f["tris"], params=f["p"], lm=lm, tex=f["tex"], tag=1)
        assert err.value.code == -1  # SPLAT_ERR_INVALID_ARG
    assert _admit(eng, f, 2)  # a valid request still goes in
    while eng.in_flight:
        eng.step()
        eng.poll(wait=True)
    eng.close()


def test_light_map_only_frames_equal_full_frames(scenes):
    """Frames that keep only their light map (no final rays or hit indices: plain keep_rays=False,
    engine requests without outputs) give the light maps of full frames; the engine finishes
    them with their last bounce, one step before frames that keep their rays."""
    import torch

    specs = _frames(10, seed=5)
    for s in specs:
        s["depth"] = max(s["depth"], 1)
    frames = [_make(s, scenes) for s in specs]
    want = [_plain(f) for f in frames]
    for f, w in zip(frames, want):  # plain, map only
        fr = api.Frame(f["cam"], f["hitter"], f["tris"], f["p"], f["lm"], keep_rays=False)
        fr.trace()
        _sync()
        assert fr.rays is None and H.rel_l2(fr.tex.cpu().numpy(), w[0]) < 1e-4
    eng = api.Engine(1 << 30, max_slots=16)
    for k, f in enumerate(frames):
        assert eng.admit(cam=f["cam"], hitter=f["hitter"], tris=f["tris"], params=f["p"], lm=f["lm"], tex=f["tex"],
                         tag=300 + k)
    done, steps = [], 0
    while len(done) < len(frames):
        eng.step()
        steps += 1
        done += eng.poll(wait=True)
        assert steps <= 20
    assert steps == max(s["depth"] for s in specs)
    for f, w in zip(frames, want):
        got = f["tex"].cpu().numpy()
        assert (np.abs(w[0]).sum() == 0 and np.abs(got).sum() == 0) or H.rel_l2(got, w[0]) < 1e-4
    eng.close()
    torch.cuda.empty_cache()
