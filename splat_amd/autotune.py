"""Autotuner for the hitter launch shapes.

Candidates are timed with HIP events on the caller's real scene and rays (through the C ABI),
the fastest is kept, and choices are cached in memory and, optionally, in a JSON file keyed by
(kernel, log2 ray count, log2 triangle count), so a serving process tunes each shape once.

Tuned knobs:
  brute force -- waves per workgroup W that split the triangle list (1, 2, 4, 8, 16);
  BV          -- leaf count of the object-median build (2^d; large triangles apart, tested
                 directly) and the top structure over the leaves (binary tree, none, or a coarse
                 level of fixed-size groups);
  device BV   -- for the tree built on the device (resident and served scenes): leaf count 2^d
                 and the walk mode (waves walk together, per wave by direction octant, or per
                 lane), then, for the best of those, the work items a launch's split factor
                 aims for (64-ray groups x splits: 3072 ... 49152), all timed on whole frames of
                 a representative camera, all bounces included.
All candidates return identical hit indices (the tuner asserts it on the tuning rays).
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from . import api, bv

_cache: dict[str, dict] = {}


def _key(kind: str, R: int, N: int) -> str:
    return f"{kind}:r{int(math.log2(max(R, 1)))}:n{int(math.log2(max(N, 1)))}"


def _load(path: str | None) -> None:
    if path and os.path.exists(path):
        with open(path) as f:
            _cache.update(json.load(f))


def _save(path: str | None) -> None:
    if path:
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(_cache, f, indent=1, sort_keys=True)
        os.replace(tmp, path)


def _time_hits(hitter, rays, reps: int = 5) -> tuple[float, np.ndarray]:
    import torch

    tr = api.HIPTracer(rays)
    api.next_hit_(tr, hitter, rays)  # warm-up
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        api.next_hit_(tr, hitter, rays)
        b.record()
        torch.cuda.synchronize()
        ts.append(a.elapsed_time(b))
    return float(np.median(ts)), tr.hit_idx.cpu().numpy()


def tune_bf(n_tris_dev, rays, cache_path: str | None = None) -> dict:
    """Best W for the brute-force hitter on these rays; returns {"waves": W, "ms": t}."""
    _load(cache_path)
    key = _key("bf", int(rays.shape[0]), int(n_tris_dev.shape[0]))
    if key in _cache:
        return _cache[key]
    best, ref = None, None
    for w in (1, 2, 4, 8, 16):
        h = api.HIPBruteForceHitter(n_tris_dev, tune_waves=w)
        ms, hits = _time_hits(h, rays)
        if ref is None:
            ref = hits
        elif not np.array_equal(ref, hits):
            raise RuntimeError(f"brute-force candidate W={w} changed the hit indices")
        if best is None or ms < best["ms"]:
            best = {"waves": w, "ms": ms}
    _cache[key] = best
    _save(cache_path)
    return best


def tune_bv(tris_dev, host_tris: np.ndarray, rays, cache_path: str | None = None,
            depths: tuple[int, ...] | None = None, group_sizes: tuple[int, ...] = (0, -1, 16)) -> dict:
    """Best (leaf depth, group size) for the BV hitter on these rays; returns a dict with
    "depth", "group_size", "ms" and builds the winning hitter under "hitter" (not cached)."""
    _load(cache_path)
    N = int(host_tris.shape[0])
    key = _key("bv", int(rays.shape[0]), N)
    if depths is None:
        d0 = bv.tree_depth(N)
        depths = tuple(d for d in (d0 - 2, d0 - 1, d0, d0 + 1) if d >= 1)
    if key not in _cache:
        best, ref = None, None
        for d in depths:
            bvs, mem, n_top = bv.tree_partition(host_tris, d)
            for g in group_sizes:
                h = api.HIPBVHitter(tris_dev, bvs, mem, group_size=g, n_top=n_top)
                ms, hits = _time_hits(h, rays)
                if ref is None:
                    ref = hits
                elif not np.array_equal(ref, hits):
                    raise RuntimeError(f"BV candidate depth={d} group={g} changed the hit indices")
                if best is None or ms < best["ms"]:
                    best = {"depth": d, "group_size": g, "ms": ms}
        _cache[key] = best
        _save(cache_path)
    choice = dict(_cache[key])
    bvs, mem, n_top = bv.tree_partition(host_tris, choice["depth"])
    choice["hitter"] = api.HIPBVHitter(tris_dev, bvs, mem, group_size=choice["group_size"], n_top=n_top)
    return choice


# split-target codes of HitSet.tune_waves bits 8-10 (items the split factor aims for)
SPLIT_TARGETS = {0: 12288, 1: 3072, 2: 6144, 3: 24576, 4: 49152}


def tune_device_bv(tris_dev, cam, params, lm, cache_path: str | None = None, depths: tuple[int, ...] | None = None,
                   walks: tuple[int, ...] = (0, 1, 2), targets: tuple[int, ...] = (0, 2, 3),
                   reps: int = 3, frames_per_rep: int = 4) -> dict:
    """Best (leaf depth, walk mode) of the device-built tree hitter for frames like `cam` over this
    scene, then the best split target (code of SPLIT_TARGETS) for that pair: each candidate traces
    whole frames (pinhole rays, params.depth bounces, light map) and the fastest median frame time
    wins. Returns {"depth", "walk", "split", "ms", "candidates"} and the winning hitter under
    "hitter" (not cached). Every candidate's final hit indices must agree."""
    import torch

    _load(cache_path)
    N = int(tris_dev.shape[0])
    R = cam.height * cam.width * cam.samples_per_pixel
    key = _key(f"bvdev2:d{params.depth}", R, N)
    if key not in _cache:
        if depths is None:
            d0 = bv.tree_depth(N - 1)
            depths = tuple(d for d in (d0 - 2, d0 - 1, d0, d0 + 1) if 1 <= d <= 20)
        frame = None
        best, ref, tried = None, None, []

        def run(h, d, w, c):
            nonlocal frame, best, ref
            h.walk = w | (c << 8)
            h.hs.tune_waves = h.walk
            if frame is None:
                frame = api.Frame(cam, h, tris_dev, params, lm, device=tris_dev.device)
            frame.hitter = h
            frame.trace()  # warm-up
            torch.cuda.synchronize()
            ts = []
            for _ in range(reps):  # frames back to back, as a frame loop issues them
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record()
                for _ in range(frames_per_rep):
                    frame.trace()
                b.record()
                torch.cuda.synchronize()
                ts.append(a.elapsed_time(b) / frames_per_rep)
            hits = frame.hit_idx.cpu().numpy()
            if ref is None:
                ref = hits
            elif not np.array_equal(ref, hits):
                raise RuntimeError(f"device BV candidate depth={d} walk={w} split={c} changed the hit indices")
            ms = float(np.median(ts))
            tried.append([d, w, c, round(ms, 4)])
            if best is None or ms < best["ms"]:
                best = {"depth": d, "walk": w, "split": c, "ms": ms}

        for d in depths:
            h = api.HIPBVHitter.build_on_device(tris_dev, d)
            for w in walks:
                run(h, d, w, 0)
            del h
        d, w = best["depth"], best["walk"]
        h = api.HIPBVHitter.build_on_device(tris_dev, d)
        for c in targets:
            if c != 0:
                run(h, d, w, c)
        del h
        best["candidates"] = tried
        _cache[key] = best
        _save(cache_path)
    choice = dict(_cache[key])
    choice["hitter"] = api.HIPBVHitter.build_on_device(tris_dev, choice["depth"],
                                                      walk=choice["walk"] | (choice.get("split", 0) << 8))
    return choice
