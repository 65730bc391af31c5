"""The serving loop's host logic (admission, waiting on the budget, completion bookkeeping) over a
stand-in engine with the real engine's admission rule: a frame is admitted while the arena has
room for it. No GPU: the stand-in completes a frame `depth + 1` steps after its admission."""
import math
from types import SimpleNamespace

import pytest

from splat_amd import abi, scene, serve


class StandInEngine:
    def __init__(self, arena_bytes: int, bytes_per_ray: int = 100):
        self.free = arena_bytes
        self.bpr = bytes_per_ray
        self.live = {}  # tag -> [steps left, bytes]
        self.done = []
        self.admitted = []

    def frame_bytes(self, cam, with_outputs=False):
        return cam.height * cam.width * cam.samples_per_pixel * self.bpr

    def synchronize(self):
        pass

    def admit(self, *, cam, hitter, tris, params, lm, tex, tag, **kw):
        need = self.frame_bytes(cam)
        if need > self.free:
            return False
        self.free -= need
        self.live[tag] = [params.depth + 1, need]
        self.admitted.append(tag)
        return True

    @property
    def in_flight(self):
        return len(self.live) + len(self.done)

    @property
    def runnable(self):
        return len(self.live)

    @property
    def arena_free(self):
        return self.free

    def step(self):
        for tag in list(self.live):
            self.live[tag][0] -= 1
            if self.live[tag][0] == 0:
                self.free += self.live.pop(tag)[1]
                self.done.append(tag)

    def sync_steps(self, keep):
        pass

    def poll(self, wait=False):
        out, self.done = self.done, []
        return out


def _requests(sizes, arrivals=None):
    sc = SimpleNamespace(hitter=None, tris=None)
    reqs = []
    for k, res in enumerate(sizes):
        cam = scene.pinhole((0, 0, 5), (0, 0, 0), (0, 1, 0), 60.0, res, res)
        reqs.append(serve.Request(k, 0, cam, abi.run_params(2, 3, math.nan), abi.lightmap_params(8, 8),
                                  arrival=0.0 if arrivals is None else arrivals[k], tex=object()))
    return [sc], reqs


@pytest.mark.timeout(30)
def test_requests_wait_for_the_budget_and_all_complete():
    scenes, reqs = _requests([32, 32, 32, 16, 32])
    eng = StandInEngine(arena_bytes=2 * 32 * 32 * 100)  # two 32x32 frames at a time
    st = serve.run_requests(eng, scenes, reqs, realtime=False)
    assert sorted(eng.admitted) == list(range(5)) and eng.admitted == sorted(eng.admitted)  # oldest first
    assert st.refused > 0 and st.max_in_flight == 2
    assert all(math.isfinite(rq.done) for rq in reqs) and st.latencies_ms.size == 5


@pytest.mark.timeout(30)
def test_a_request_larger_than_the_arena_is_an_error_not_a_wait():
    scenes, reqs = _requests([16, 64, 16])
    eng = StandInEngine(arena_bytes=48 * 48 * 100)
    with pytest.raises(ValueError, match="request 1"):
        serve.run_requests(eng, scenes, reqs, realtime=False)


def test_frame_loop_needs_a_buffer():
    """api.FrameLoop refuses fewer than one buffer before it allocates anything."""
    from splat_amd import api

    for n in (0, -2):
        with pytest.raises(ValueError, match="in_flight"):
            api.FrameLoop(scene.camera_c1(), None, None, None, None, in_flight=n)
