"""Render-request serving over the multi-frame engine, and the batched workload.

Scenes (FTri arrays + prepared BV hit sets) stay resident on the device. Requests -- a scene, a
pinhole (resolution, samples per pixel, seed), bounces and a light map -- arrive over time; the
loop admits every arrived request, oldest first, while the engine's HBM arena and slot table have
room (SPLAT_ERR_BUDGET otherwise: it waits), and keeps stepping the engine, so requests that
arrive while others are on the GPU join at the next step and share every launch with them
(continuous batching at bounce granularity). Latency = completion - arrival, per request.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import abi, api, scene


@dataclass
class ResidentScene:
    sc: scene.Scene
    tris: object  # device FTri tensor
    hitter: api.HIPBVHitter
    name: str = ""

    tuned: dict | None = None  # the autotuner's choice (depth, walk, candidates), when tuned

    @staticmethod
    def build(sc: scene.Scene, depth: int | None = None, device: str = "cuda", tune: bool = True,
              tune_cache: str | None = None) -> "ResidentScene":
        """Uploads the triangles and builds the tree hitter's BVs on the device (0.5-1 ms for
        10k-178k triangles; the host tree_partition takes 60-1000 ms). With tune (and no fixed
        depth) the leaf depth and walk mode come from the autotuner, timed on 512^2 frames of a
        light above the scene (autotune.tune_device_bv; cached per scene and frame shape)."""
        from . import autotune

        tris = api.to_device(sc.ftris, device)
        if depth is not None or not tune:
            return ResidentScene(sc, tris, api.HIPBVHitter.build_on_device(tris, depth), sc.name)
        cam = lights_cam(1, 512, 1, abi.LAMBDA_UNIFORM)
        p = abi.run_params(4, sc.first_diffuse, math.nan, seed=1)
        # (the engine sets its split factors for all its frames at once: no split target to tune)
        ch = autotune.tune_device_bv(tris, cam, p, abi.lightmap_params(512, 512), cache_path=tune_cache, targets=(0,))
        tuned = {k: v for k, v in ch.items() if k != "hitter"}
        return ResidentScene(sc, tris, ch["hitter"], sc.name, tuned)


@dataclass
class Request:
    tag: int
    scene: int
    cam: abi.Pinhole
    params: abi.RunParams
    lm: abi.LightMapParams
    arrival: float = 0.0  # seconds after the start
    tex: object = None
    rays_out: object = None  # optional device (R, 136) uint8: the frame's final rays
    hit_out: object = None   # optional device int32[R]: their hit indices
    admitted: float = math.nan
    done: float = math.nan

    @property
    def rays(self) -> int:
        return self.cam.height * self.cam.width * self.cam.samples_per_pixel


@dataclass
class ServeStats:
    wall_s: float
    rays: int
    latencies_ms: np.ndarray
    steps: int
    max_in_flight: int
    refused: int = 0
    extra: dict = field(default_factory=dict)

    @property
    def mrays_per_s(self) -> float:
        return self.rays / self.wall_s / 1e6

    def summary(self) -> dict:
        lat = self.latencies_ms
        return {"mrays_per_s": round(self.mrays_per_s, 3), "wall_s": round(self.wall_s, 3), "requests": int(lat.size),
                "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                "steps": self.steps, "max_in_flight": self.max_in_flight, "budget_waits": self.refused, **self.extra}


def run_requests(engine: api.Engine, scenes: list[ResidentScene], requests: list[Request], *, realtime: bool = True,
                 max_pending_steps: int = 2) -> ServeStats:
    """Serve `requests` (sorted by arrival). realtime=False admits everything at t = 0 (a batch)."""
    import torch

    for rq in requests:
        if rq.tex is None:
            rq.tex = torch.empty((rq.lm.channels, rq.lm.height, rq.lm.width), dtype=torch.float32, device="cuda")
    by_tag = {rq.tag: rq for rq in requests}
    queue = list(requests)
    engine.synchronize()
    t0 = time.perf_counter()
    now = lambda: time.perf_counter() - t0  # noqa: E731
    steps = refused = max_in_flight = 0
    finished = 0
    while finished < len(requests):
        t = now()
        while queue and (not realtime or queue[0].arrival <= t):
            rq = queue[0]
            sc = scenes[rq.scene]
            if not engine.admit(cam=rq.cam, hitter=sc.hitter, tris=sc.tris, params=rq.params, lm=rq.lm, tex=rq.tex,
                                tag=rq.tag, rays_out=rq.rays_out, hit_out=rq.hit_out):
                if not engine.in_flight:
                    # an empty engine refused it: the frame is larger than the whole arena, and
                    # waiting would never admit it
                    raise ValueError(f"request {rq.tag} ({rq.rays} rays) needs {engine.frame_bytes(rq.cam)} bytes of "
                                     f"engine arena; the empty engine has {engine.arena_free}")
                refused += 1
                break
            rq.admitted = t
            queue.pop(0)
        max_in_flight = max(max_in_flight, engine.in_flight)
        if engine.runnable:
            engine.step()
            steps += 1
            engine.sync_steps(max_pending_steps)
        elif queue and not engine.in_flight:  # idle until the next arrival
            time.sleep(max(0.0, min(queue[0].arrival - now(), 0.002)))
        # with nothing left to launch, block on the completions still in flight
        tags = engine.poll(wait=engine.in_flight > 0 and engine.runnable == 0)
        tn = now()
        for tg in tags:
            by_tag[tg].done = tn
            finished += 1
    wall = now()
    lat = np.array([(rq.done - rq.arrival) * 1e3 for rq in requests])
    return ServeStats(wall, sum(rq.rays for rq in requests), lat, steps, max_in_flight, refused)


def poisson_arrivals(n: int, rate_per_s: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_per_s, n))


def lights_cam(seed: int, res: int, spp: int, lambda_mode: int, scene_extent: float = 2.0) -> abi.Pinhole:
    """A seeded point light / pinhole above the scene, looking at the ground."""
    rng = np.random.default_rng(seed)
    pos = (float(rng.uniform(-0.5, 0.5)), float(rng.uniform(-0.5, 0.5)), float(rng.uniform(4.0, 7.0) * scene_extent / 2))
    look = (float(rng.uniform(-0.3, 0.3)), float(rng.uniform(-0.3, 0.3)), -1.5)
    return scene.pinhole(pos, look, (0.0, 1.0, 0.0), 90.0, res, res, spp=spp, lambda_mode=lambda_mode,
                         jitter=0.025, seed=seed)


def batched_workload(n_frames: int = 64, seed: int = 4, depth: int = 4, tune_cache: str | None = None):
    """64 frames, resolution from {256^2, 512^2, 1024^2}, each over an icosphere of subdivision
    2-5 (320-20,480 glass triangles) above a diffuse plane; pose, light and wavelengths from the
    seed. Returns (scenes, requests)."""
    scs = [scene.assemble([scene.sphere_mesh(k, radius=1.0)], [scene.ground_quad(-1.5, 4.0)], f"icosphere{k}+plane")
           for k in (2, 3, 4, 5)]
    res_scenes = [ResidentScene.build(sc, tune_cache=tune_cache) for sc in scs]
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(n_frames):
        si = int(rng.integers(0, 4))
        res = int(rng.choice([256, 512, 1024]))
        cam = lights_cam(int(rng.integers(1, 1 << 30)), res, 1, abi.LAMBDA_UNIFORM)
        p = abi.run_params(depth, scs[si].first_diffuse, math.nan, seed=int(rng.integers(1, 1 << 30)))
        reqs.append(Request(k, si, cam, p, abi.lightmap_params(res, res)))
    return res_scenes, reqs


def serving_scenes(tune_cache: str | None = None):
    """8 resident scenes from 320 to ~178k triangles."""
    ground = lambda h=4.0: scene.ground_quad(-1.5, h)  # noqa: E731
    scs = [scene.assemble([scene.sphere_mesh(k)], [ground()], f"icosphere{k}+plane") for k in (2, 3, 4, 5, 6)]
    scs.append(scene.scene_c2(seed=2))
    cen, edge = scene.menger_cubes(3, center=(0.0, 0.0, -0.2), size=2.4)
    scs.append(scene.assemble([scene.sphere_mesh(4, center=(0.0, 0.0, 2.0), radius=0.8)],
                              [scene.cube_mesh(cen, edge), ground(5.0)], "menger3+icosphere4+plane"))
    scs.append(scene.scene_c3())
    return [ResidentScene.build(sc, tune_cache=tune_cache) for sc in scs]


def serving_requests(n: int, rate_per_s: float, n_scenes: int, scenes_fd: list[int], seed: int = 5,
                     res_range=(256, 2048), spp_range=(1, 16), depth: int = 4):
    """Poisson arrivals; resolution uniform over res_range (multiples of 8), 1-16 stratified
    wavelengths per pixel, a uniformly drawn resident scene."""
    rng = np.random.default_rng(seed)
    arr = poisson_arrivals(n, rate_per_s, seed)
    reqs = []
    for k in range(n):
        si = int(rng.integers(0, n_scenes))
        res = int(rng.integers(res_range[0] // 8, res_range[1] // 8 + 1)) * 8
        spp = int(rng.integers(spp_range[0], spp_range[1] + 1))
        cam = lights_cam(int(rng.integers(1, 1 << 30)), res, spp,
                         abi.LAMBDA_STRATIFIED if spp > 1 else abi.LAMBDA_UNIFORM)
        p = abi.run_params(depth, scenes_fd[si], math.nan, seed=int(rng.integers(1, 1 << 30)))
        lm_res = min(res, 1024)
        reqs.append(Request(k, si, cam, p, abi.lightmap_params(lm_res, lm_res), arrival=float(arr[k])))
    return reqs
