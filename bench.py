#!/usr/bin/env python3
"""Light-map frame benchmark on one MI355X.

Workload: two displaced icospheres (subdivision 4 each, 10,240 glass triangles, radial noise
sigma 0.02) over a diffuse plane; 1024 x 1024 rays from a pinhole at the point light, one seeded
wavelength per ray, Cauchy glass (A 1.5, B 0.004 um^2), bounding-volume hitter (ground triangles
tested directly, tree over 2^10 object-median leaves of the rest), 4
bounces with dual-number differentials, splat into a 1024 x 1024 RGB float32 light map.
A step is one whole frame: clear map -> pinhole rays -> bounce loop (hit, evolve, stable
compaction) -> realigning hit -> final evolution + splat, all device-resident, one frame per
launch sequence ("plain" path).

Prints one JSON line (the last line of stdout).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s + p50/p99 frame latency, served-batched vs plain HIP, 1x MI355X"
FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 vector peak
HBM_PEAK_GBS = 8000.0
FLOP_PER_TRI_TEST = 40  # distance_to_plane + in_triangle + p = o + t d, per-triangle terms hoisted
FLOP_PER_BOX_TEST = 12  # slab test with precomputed reciprocals
# VALU issue peak: each of a CU's 4 SIMDs issues one wave64 vector instruction every 4 cycles (16
# lanes a cycle); MI355X's published peak engine clock
PEAK_CLOCK_HZ = 2.4e9


class Hip:
    """The few raw HIP event calls needed to time kernels inside the C ABI's launch sequence."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.h.hipEventDestroy.argtypes = [C.c_void_p]

    def events(self, n):
        out = (C.c_void_p * n)()
        for i in range(n):
            e = C.c_void_p()
            if self.h.hipEventCreate(C.byref(e)) != 0:
                raise RuntimeError("hipEventCreate failed")
            out[i] = e
        return out

    def elapsed_ms(self, a, b) -> float:
        ms = C.c_float()
        if self.h.hipEventElapsedTime(C.byref(ms), a, b) != 0:
            raise RuntimeError("hipEventElapsedTime failed")
        return float(ms.value)

    def destroy(self, evs):
        for e in evs:
            self.h.hipEventDestroy(e)


def build_workload(device="cuda", tune=True, tune_cache=None):
    """The 1 MP frame; its BV tree is built on the device with the leaf depth and walk mode the
    autotuner picks on this frame (whole frames timed per candidate)."""
    from splat_amd import abi, api, autotune, scene

    sc = scene.scene_c2(seed=2)
    cam = scene.camera_c2(seed=2, jitter=0.025)
    tris_d = api.to_device(sc.ftris, device)
    params = abi.run_params(4, sc.first_diffuse, math.nan, seed=2024)
    lm = abi.lightmap_params(1024, 1024)
    if tune:
        ch = autotune.tune_device_bv(tris_d, cam, params, lm, cache_path=tune_cache)
        hitter, tuned = ch["hitter"], {k: v for k, v in ch.items() if k != "hitter"}
    else:
        hitter, tuned = api.HIPBVHitter.build_on_device(tris_d), None
    frame = api.Frame(cam, hitter, tris_d, params, lm, device)
    return sc, cam, tuned, hitter, params, lm, frame


def _profile_json(name):
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def hitter_traffic(depth):
    """HBM bytes per hitter launch on this workload at this leaf depth, from the committed
    rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by what the same
    counters read for a device copy of known size); None when absent for that depth."""
    t = _profile_json("k_hit_bv_c2_traffic.json")
    e = t and t["by_leaf_depth"].get(str(depth))
    return (float(e["hbm_bytes_per_launch"]), t.get("source")) if e else (None, None)


def hitter_pmc_flops(depth):
    """FP32 FLOP per hitter launch on this workload at this leaf depth from the committed
    rocprofv3 --pmc summary: SQ_INSTS_VALU_FLOPS_FP32 per dispatch times the factor the same
    counter reads for the calibration kernel of known FLOP count (bench.py --flop-probe); None
    when absent for that depth."""
    t = _profile_json("k_hit_bv_c2_counters.json")
    e = t and t["by_leaf_depth"].get(str(depth))
    return (float(e["valu_flops_fp32_per_dispatch"]) * float(t["flop_per_count"]), t.get("source")) if e else (None, None)


def valu_counts(depth):
    """SQ_INSTS_VALU (wave64 vector instructions) per dispatch of the 1 MP frame's hitter launches
    (bounces 1-4 and the realigning hit) and fused bounce passes (bounces 1-4) at this leaf depth,
    from the committed rocprofv3 --pmc summary; None when absent for that depth."""
    t = _profile_json("valu_c2_counts.json")
    e = t and t["by_leaf_depth"].get(str(depth))
    return dict(e, source=t.get("source")) if e else None


def valu_roofline(kernel, insts, ms, n_cu):
    """Achieved vector-instruction issue rate of `kernel` (instructions per launch over event-timed
    milliseconds per launch) against the chip's issue peak."""
    peak = n_cu * PEAK_CLOCK_HZ / 1e9  # G wave-instructions/s
    rates = [i / (t * 1e-3) / 1e9 for i, t in zip(insts, ms)]
    tot = sum(insts) / (sum(ms) * 1e-3) / 1e9
    return {"kernel": kernel, "bound": "valu-issue", "unit": "G wave64 VALU instructions/s", "peak": round(peak, 1),
            "achieved": round(tot, 1), "frac": round(tot / peak, 4),
            "per_launch": [{"insts": int(i), "ms": round(t, 4), "achieved": round(r, 1), "frac": round(r / peak, 4)}
                           for i, t, r in zip(insts, ms, rates)]}


def live_counts(cam, hitter, tris, params) -> list[int]:
    """Live (ACTIVE) rays before each bounce of the frame, A_0 = R, ..., A_depth, from the same
    hitter / evolve / partition operators stepped one bounce at a time (identical rays and
    random numbers to the fused loop)."""
    from splat_amd import api

    rays = api.rays_from_pinhole(cam)
    tr = api.HIPTracer(rays)
    out = [int(rays.shape[0])]
    for k in range(1, params.depth + 1):
        api.next_hit_(tr, hitter, rays)
        api.evolve_(rays, tr.hit_idx, tris, first_diffuse=params.first_diffuse, seed=params.seed, bounce=k,
                    optics=params.optics)
        out.append(api.partition_(rays, tr.ray_swap, tr))
    return out


def host_info() -> dict:
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def bf_frame_hitter(tris, cam, tune_cache=None):
    """The brute-force hitter of the 320-triangle frame, with the waves per 64-ray group the
    autotuner picks on the frame's first-bounce rays (cached per ray and triangle count)."""
    from splat_amd import api, autotune

    choice = autotune.tune_bf(tris, api.rays_from_pinhole(cam), cache_path=tune_cache)
    return api.HIPBruteForceHitter(tris, tune_waves=choice["waves"]), choice


def small_frame(budget_s: float = 3.0, tune_cache=None):
    """The 320-triangle workload (icosphere subdivision 2 over a plane, 256x256 rays, one seeded
    wavelength each, brute-force hitter, 2 bounces, 512x512 RGB map): GPU plain-path rate and the
    oracle's CPU rate on whole frames."""
    import numpy as np
    import torch

    from oracle import oracle as O
    from splat_amd import abi, api, scene

    sc = scene.scene_c1()
    cam = scene.camera_c1(seed=1)
    tris = api.to_device(sc.ftris)
    p = abi.run_params(2, sc.first_diffuse, math.nan, seed=2024)
    lm = abi.lightmap_params(512, 512)
    hitter, tuned = bf_frame_hitter(tris, cam, tune_cache)
    fr = api.Frame(cam, hitter, tris, p, lm)
    R = fr.R
    for _ in range(10):
        fr.trace()
    torch.cuda.synchronize()
    K = 200
    t = time.perf_counter()
    for _ in range(K):
        fr.trace()
    torch.cuda.synchronize()
    gpu = R * K / (time.perf_counter() - t) / 1e6
    # the brute-force hitter's launches timed with events on the frame's stream, and its
    # algorithmic tests (every live ray x every triangle) from a counting frame
    hip = Hip()
    D, M = p.depth, 50
    per = 4 * D + 2
    hev = hip.events(per * M)
    for m in range(M):
        p.hit_events = C.addressof(hev) + m * per * C.sizeof(C.c_void_p)
        p.n_hit_events = per
        fr.trace()
    torch.cuda.synchronize()
    launch_ms = [hip.elapsed_ms(hev[m * per + 2 * j], hev[m * per + 2 * j + 1]) for m in range(M) for j in range(D + 1)]
    p.hit_events, p.n_hit_events = None, 0
    hip.destroy(hev)
    stats = torch.zeros(5, dtype=torch.int64, device="cuda")
    hitter.hs.d_stats = stats.data_ptr()
    fr.trace()
    torch.cuda.synchronize()
    hitter.hs.d_stats = None
    _, _, alg, exe, _ = (int(x) for x in stats.cpu().tolist())
    avg_s = float(np.mean(launch_ms)) / 1e3
    flop = FLOP_PER_TRI_TEST * alg / (D + 1)
    bf = {"kernel": "k_hit_bf (brute-force hitter)", "bound": "valu", "unit": "TFLOP/s", "peak": FP32_PEAK_TFLOPS,
          "achieved": round(flop / avg_s / 1e12, 3), "frac": round(flop / avg_s / 1e12 / FP32_PEAK_TFLOPS, 4),
          "frac_of_unpacked_peak": round(flop / avg_s / 1e12 / (FP32_PEAK_TFLOPS / 2), 4),
          "flop_per_launch": flop, "flop_per_launch_executed": FLOP_PER_TRI_TEST * exe / (D + 1),
          "avg_launch_ms": round(avg_s * 1e3, 4), "launches_per_frame": D + 1,
          "tri_tests_per_frame": alg}
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    n, reps = 0, 0
    t = time.perf_counter()
    while reps == 0 or (time.perf_counter() - t < budget_s and reps < 50):
        n += O.trace_sample(cam, 1, 0, sc.ftris, p, lm, None, threads)[1]
        reps += 1
    cpu = n / (time.perf_counter() - t) / 1e6
    return {"workload": "320-triangle icosphere + plane, 256x256 rays, brute force, 2 bounces, 512x512x3 map",
            "gpu_mrays_per_s": round(gpu, 2), "cpu_mrays_per_s": round(cpu, 3), "cpu_threads": threads,
            "cpu_frames": reps, "gpu_ms_per_frame": round(R / gpu / 1e3, 4), "roofline_bf_hitter": bf,
            "autotuned": {"waves": tuned["waves"]}}


def bf_sweep(reps: int = 5):
    """The reference's brute-force hitter sweep (src/scenes/perf_test.jl:11-36): R in {256^2, 512^2,
    768^2, 1024^2} rays from the origin in uniformly random directions (components in [0, 1)),
    N in {4^0 .. 4^8} random triangles (vertices in [0, 1)^3, after the sentinel); the median of
    `reps` next_hit! calls, each R x N ray-triangle tests of 40 FLOP."""
    import numpy as np
    import torch

    from splat_amd import abi, api

    rng = np.random.default_rng(36)
    out = []
    for side in (256, 512, 768, 1024):
        R = side * side
        rays = np.zeros(R, dtype=abi.ADRAY)
        rays["dir"] = rng.uniform(0, 1, (R, 3)).astype(np.float32)
        rays["ignore_tri"] = 1
        rays["status"] = abi.RAY_ACTIVE
        rays_d = api.to_device(rays)
        tr = api.HIPTracer(rays_d)
        for e in range(0, 17, 2):
            N = 1 << e
            t = np.zeros(N + 1, dtype=abi.NTRI)
            t["idx"] = np.arange(1, N + 2, dtype=np.int32)
            v = rng.uniform(0, 1, (N, 3, 3)).astype(np.float32)
            t["v"][1:, 1:] = v
            t["v"][1:, 0] = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]).astype(np.float32)
            hitter = api.HIPBruteForceHitter(api.to_device(t))
            for _ in range(2):
                api.next_hit_(tr, hitter, rays_d)
            torch.cuda.synchronize()
            ms = []
            for _ in range(reps):
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record()
                api.next_hit_(tr, hitter, rays_d)
                b.record()
                torch.cuda.synchronize()
                ms.append(a.elapsed_time(b))
            m = float(np.median(ms))
            tflops = FLOP_PER_TRI_TEST * R * N / (m / 1e3) / 1e12
            out.append([R, N, round(m, 4), round(tflops, 2), round(tflops / FP32_PEAK_TFLOPS, 4)])
    return {"columns": ["rays", "triangles", "median_ms", "TFLOP/s", "frac_of_157.3"], "rows": out}


def splat_operator(sc, cam, hitter, tris, params, lm, reps: int = 20):
    """continuum_light_map2! as the standalone operator (splat_light_map: final evolution of the
    DIFFUSE rays, then the atomic splat) on the 1 MP frame's rays after run_evolution!; algorithmic
    bytes: per diffuse ray its record's 109 used bytes and its 4-byte hit index read, and the 52
    bytes final_evolution changes written back (position and its three tangents, the hit index:
    the reference assigns the evolved rays back, src/ray_imagers.jl:155), each touched triangle's
    120 bytes once, the map read and written once."""
    import numpy as np
    import torch

    from splat_amd import abi, api

    rays0 = api.rays_from_pinhole(cam)
    tr = api.HIPTracer(rays0)
    api.run_evolution_(hitter, tr, depth=params.depth, first_diffuse=params.first_diffuse, tris=tris, rays=rays0,
                       seed=params.seed, optics=params.optics)
    torch.cuda.synchronize()
    host = api.from_device(rays0, abi.ADRAY)
    dif = host["status"] == abi.RAY_DIFFUSE
    n_dif = int(dif.sum())
    touched = int(np.unique(tr.hit_idx.cpu().numpy()[dif]).size)
    tex = torch.zeros((lm.channels, lm.height, lm.width), dtype=torch.float32, device="cuda")
    rays = torch.empty_like(rays0)
    ms = []
    for k in range(reps + 2):
        rays.copy_(rays0)
        tex.zero_()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        api.continuum_light_map2_(tracer=tr, tris=tris, rays=rays, tex=tex, lm=lm)
        b.record()
        torch.cuda.synchronize()
        if k >= 2:
            ms.append(a.elapsed_time(b))
    m = float(np.median(ms))
    nbytes = n_dif * (109 + 4 + 52) + touched * 120 + 2 * lm.width * lm.height * lm.channels * 4
    return {"kernel": "k_light_map (continuum_light_map2!: final evolution + atomic splat, standalone)",
            "bound": "hbm", "achieved": round(nbytes / (m / 1e3) / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(nbytes / (m / 1e3) / 1e9 / HBM_PEAK_GBS, 4), "bytes": nbytes, "median_ms": round(m, 4),
            "diffuse_rays": n_dif, "splats_per_s_G": round(n_dif / (m / 1e3) / 1e9, 2)}


def partition_operator(cam, hitter, tris, params, reps: int = 20):
    """partition! as the standalone operator (splat_partition_active: per-tile ACTIVE counts, scan,
    stable scatter into the swap buffer, copy back) on the 1 MP frame's rays after the first
    bounce's evolve (ACTIVE, DIFFUSE and escaped rays mixed); algorithmic bytes: each 136-byte
    record read and written once plus its status byte, R x (2 x 136 + 1). partition! also copies
    the result back into `vals` (src/partition.jl:124-131 does the same), another read and write of
    every record: `with_copy_back` counts those bytes too."""
    import numpy as np
    import torch

    from splat_amd import abi, api

    rays0 = api.rays_from_pinhole(cam)
    tr = api.HIPTracer(rays0)
    api.next_hit_(tr, hitter, rays0)
    api.evolve_(rays0, tr.hit_idx, tris, first_diffuse=params.first_diffuse, seed=params.seed, bounce=1,
                optics=params.optics)
    torch.cuda.synchronize()
    R = int(rays0.shape[0])
    active = int((api.from_device(rays0, abi.ADRAY)["status"] == abi.RAY_ACTIVE).sum())
    rays = torch.empty_like(rays0)
    ms = []
    for k in range(reps + 2):
        rays.copy_(rays0)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        api.partition_(rays, tr.ray_swap, tr, host_count=False)
        b.record()
        torch.cuda.synchronize()
        if k >= 2:
            ms.append(a.elapsed_time(b))
    m = float(np.median(ms))
    nbytes = R * (2 * abi.ADRAY.itemsize + 1)
    return {"kernel": "k_count_active + k_scan_tiles + k_scatter (partition!, standalone)", "bound": "hbm",
            "achieved": round(nbytes / (m / 1e3) / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(nbytes / (m / 1e3) / 1e9 / HBM_PEAK_GBS, 4), "bytes": nbytes, "median_ms": round(m, 4),
            "rays": R, "active": active,
            "with_copy_back": {"bytes": 2 * nbytes - R, "achieved": round((2 * nbytes - R) / (m / 1e3) / 1e9, 1),
                               "frac": round((2 * nbytes - R) / (m / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}}


def cpu_baseline(sc, cam, params, lm, budget_s: float = 12.0):
    """The oracle's FP32 C++ restatement of the same frame (BV-culled hitter, OpenMP), timed on
    this host on an image-wide strided sample of the frame's rays."""
    from oracle import oracle as O
    from splat_amd import bv

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    # the CPU port walks the leaves flat (the reference's one-level BV hitter): ~sqrt(N) leaves
    # balance its box and triangle tests
    depth = max(1, int(round(0.5 * math.log2(max(sc.n_tris, 2)))))
    counts, mat, stride_m = bv.pack_bv_tris(bv.object_partition(sc.ftris, depth)[1], sc.n_tris)
    R_full = cam.height * cam.width * cam.samples_per_pixel
    t = time.perf_counter()
    _, n = O.trace_sample(cam, 64, 0, sc.ftris, params, lm, (counts, mat, stride_m), threads)
    per_ray = (time.perf_counter() - t) / max(n, 1)
    # a strided sample of about budget_s seconds; a frame cheaper than that is traced repeatedly
    stride = max(1, int(math.ceil(R_full * per_ray / budget_s)))
    n, reps = 0, 0
    t = time.perf_counter()
    while reps == 0 or (stride == 1 and time.perf_counter() - t < budget_s and reps < 100):
        n += O.trace_sample(cam, stride, 0, sc.ftris, params, lm, (counts, mat, stride_m), threads)[1]
        reps += 1
    dt = time.perf_counter() - t
    what = f"every {stride}th ray of the {cam.width}x{cam.height} frame" if stride > 1 else \
        f"the whole {cam.width}x{cam.height} frame x {reps}"
    return {"value": round(n / dt / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port", "host": host_info(),
            "sample": f"{what} ({n} rays, {dt:.1f} s): oracle FP32 C++ frame (flat BV culling over "
                      f"2^{depth} object-median leaves, {params.depth} bounces, splat), OpenMP {threads} threads"}


def batched_vs_plain(reps: int = 3, tune_cache=None):
    """64 mixed frames (256^2-1024^2 rays, icosphere subdiv 2-5 + plane, 4 bounces): all in one
    engine batch (every frame's work items in shared launches, one graph replay per step) versus
    the same frames traced one frame per launch sequence. Both keep each frame's light map only
    (the served product): no final rays or hit indices on either side. Also the batch with each
    step's hitter one persistent launch fed by device work queues (the engine's use_queue mode)."""
    import torch

    from splat_amd import api, serve

    scenes, reqs = serve.batched_workload(64, tune_cache=tune_cache)
    rays = sum(r.rays for r in reqs)
    frames = [api.Frame(r.cam, scenes[r.scene].hitter, scenes[r.scene].tris, r.params, r.lm, keep_rays=False)
              for r in reqs]
    for f in frames:  # warm-up
        f.trace()
    torch.cuda.synchronize()
    plain = []
    for _ in range(reps):
        t = time.perf_counter()
        for f in frames:
            f.trace()
        torch.cuda.synchronize()
        plain.append(time.perf_counter() - t)
    del frames
    arena = int(sum(r.rays for r in reqs) * 440 * 1.05) + (1 << 28)

    def batch_s(use_queue):
        eng = api.Engine(arena, max_slots=64, use_queue=use_queue)
        out = []
        for rep in range(reps + 1):
            for r in reqs:
                r.done = float("nan")
            st = serve.run_requests(eng, scenes, reqs, realtime=False, max_pending_steps=8)
            if rep:
                out.append(st.wall_s)
        eng.close()
        return min(out), st.steps

    tb, steps = batch_s(False)  # the default: one hitter wave per item over a host-sized grid
    tq, _ = batch_s(True)  # each step's hitter one persistent launch over the device work queues
    tp = min(plain)
    return {"frames": len(reqs), "rays": rays, "plain_mrays_per_s": round(rays / tp / 1e6, 3),
            "autotuned": [{k: s.tuned.get(k) for k in ("depth", "walk", "split")} if s.tuned else None for s in scenes],
            "batched_mrays_per_s": round(rays / tb / 1e6, 3), "batched_vs_plain": round(tp / tb, 3),
            "plain_s": round(tp, 4), "batched_s": round(tb, 4), "steps_per_batch": steps,
            "batched_queue_mrays_per_s": round(rays / tq / 1e6, 3),
            "batched_queue_vs_plain": round(tp / tq, 3)}


def _sweep_point(st, reqs, rate, extra=None):
    import numpy as np

    out = st.summary()
    offered = sum(rq.rays for rq in reqs) / reqs[-1].arrival / 1e6
    out.update({"arrival_rate_per_s": rate, "offered_mrays_per_s": round(offered, 1),
                "achieved_vs_offered": round(out["mrays_per_s"] / offered, 3)})
    small = np.array([(rq.done - rq.arrival) * 1e3 for rq in reqs if rq.rays <= (1 << 20)])
    if extra == "small" and small.size:
        out["p99_ms_requests_upto_1M_rays"] = round(float(np.percentile(small, 99)), 3)
    return out


def served(n_requests: int = 120, rates=(40.0, 150.0, 300.0, 600.0, 1200.0), tune_cache=None):
    """Poisson arrivals over 8 resident scenes (320 to ~178k triangles), resolution uniform over
    256^2-2048^2, 1-16 stratified wavelengths per pixel, 4 bounces, admitted against an HBM budget
    and continuously batched into shared steps; one run per arrival rate, from below to past
    saturation (achieved below offered)."""
    from splat_amd import api, serve

    scenes = serve.serving_scenes(tune_cache)
    fd = [s.sc.first_diffuse for s in scenes]
    eng = api.Engine(96 << 30, max_slots=64)
    serve.run_requests(eng, scenes, serve.serving_requests(4, rates[0], len(scenes), fd), realtime=False)  # warm-up
    sweep = []
    for rate in rates:
        reqs = serve.serving_requests(n_requests, rate, len(scenes), fd)
        st = serve.run_requests(eng, scenes, reqs, realtime=True)
        sweep.append(_sweep_point(st, reqs, rate, "small"))
    eng.close()
    out = dict(sweep[0])
    out.update({"hbm_budget_gib": 96, "triangles": [s.sc.n_tris for s in scenes], "sweep": sweep,
                "autotuned": [{k: s.tuned.get(k) for k in ("depth", "walk", "split")} if s.tuned else None for s in scenes]})
    return out


def served_1mp(n_requests: int = 300, rates=(400.0, 1500.0, 2500.0, 3500.0, 5000.0), tune_cache=None):
    """The latency case: Poisson arrivals of 1 MP requests (1024^2 rays, one seeded wavelength per
    ray, 4 bounces, 1024^2 RGB light map) over the 10,240-triangle two-icosphere scene, each with
    its own pose and light, continuously batched; p50/p99 = completion - arrival. One run per
    arrival rate, from below to past saturation; `p99_over_33ms_from_rate` is the lowest rate
    whose p99 exceeds 33 ms (None: none of them)."""
    import math

    from splat_amd import abi, api, scene, serve

    sc = scene.scene_c2(seed=2)
    rs = serve.ResidentScene.build(sc, tune_cache=tune_cache)

    def make(rate):
        arr = serve.poisson_arrivals(n_requests, rate, seed=7)
        reqs = []
        for k in range(n_requests):
            cam = serve.lights_cam(100 + k, 1024, 1, abi.LAMBDA_UNIFORM)
            p = abi.run_params(4, sc.first_diffuse, math.nan, seed=1000 + k)
            reqs.append(serve.Request(k, 0, cam, p, abi.lightmap_params(1024, 1024), arrival=float(arr[k])))
        return reqs

    eng = api.Engine(16 << 30, max_slots=64)
    serve.run_requests(eng, [rs], make(rates[0])[:8], realtime=False)  # warm-up
    sweep = []
    for rate in rates:
        reqs = make(rate)
        st = serve.run_requests(eng, [rs], reqs, realtime=True)
        sweep.append(_sweep_point(st, reqs, rate))
    eng.close()
    out = dict(sweep[0])
    over = [pt["arrival_rate_per_s"] for pt in sweep if pt["p99_ms"] > 33.0]
    out.update({"rays_per_request": 1 << 20, "triangles": sc.n_tris, "sweep": sweep,
                "p99_over_33ms_from_rate": over[0] if over else None})
    return out


def large_frame(frames: int = 3):
    """The large-scene workload: a level-3 Menger sponge (96,000 triangles) and a subdivision-6
    icosphere (81,920) over a plane, 2048^2 pixels x 8 stratified wavelengths (33.5M rays), BV
    tree built on the device, 6 bounces, 2048^2 RGB light map; whole frames timed with events."""
    import math

    import numpy as np
    import torch

    from splat_amd import abi, api, scene

    sc = scene.scene_c3()
    cam = scene.camera_c3(seed=3)
    tris = api.to_device(sc.ftris)
    hitter = api.HIPBVHitter.build_on_device(tris)
    p = abi.run_params(6, sc.first_diffuse, math.nan, seed=2024)
    fr = api.Frame(cam, hitter, tris, p, abi.lightmap_params(2048, 2048))
    fr.trace()
    torch.cuda.synchronize()
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(frames + 1)]
    evs[0].record()
    for i in range(frames):
        fr.trace()
        evs[i + 1].record()
    torch.cuda.synchronize()
    ms = [evs[i].elapsed_time(evs[i + 1]) for i in range(frames)]
    out = {"workload": "Menger sponge level 3 + icosphere subdiv 6 + plane (177,923 triangles), 2048^2 x 8 lambda "
                       "(33.5M rays), BV tree over 2^%d leaves built on the device, 6 bounces, 2048^2x3 map" % hitter.depth,
           "frames": frames, "ms_per_frame": [round(x, 3) for x in ms],
           "mrays_per_s": round(fr.R / (float(np.mean(ms)) / 1e3) / 1e6, 1)}
    del fr, hitter, tris
    torch.cuda.empty_cache()
    return out


def bv_build(reps: int = 10):
    """The tree hitter's BV structure built on the device (splat_hitset_build_bv, one stream
    synchronisation per build) vs the host build (bv.tree_partition + splat_hitset_prepare_bv),
    on the 10k- and 178k-triangle scenes: what an animated scene pays per frame."""
    import numpy as np
    import torch

    from splat_amd import api, bv, scene

    out = {}
    for name, sc in (("10k_tris", scene.scene_c2(seed=2)), ("178k_tris", scene.scene_c3())):
        tris = api.to_device(sc.ftris)
        h = api.HIPBVHitter.build_on_device(tris)
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            h.rebuild(tris)
            ts.append(time.perf_counter() - t)
        t = time.perf_counter()
        bvs, mem, n_top = bv.tree_partition(sc.ftris, h.depth)
        api.HIPBVHitter(tris, bvs, mem, n_top=n_top)
        torch.cuda.synchronize()
        host = time.perf_counter() - t
        out[name] = {"triangles": sc.n_tris, "leaves": 2 ** h.depth, "device_ms": round(float(np.median(ts)) * 1e3, 3),
                     "host_ms": round(host * 1e3, 1)}
        del h, tris
    return out


def render_frame(frames: int = 20):
    """One trace! iteration per step at 1 MP on the 10,240-triangle scene: forward light map from
    1024^2 light rays (uniform wavelengths) into 13 spectral planes of 1024^2 (400-700 nm every
    25 nm, the reference's grid), 4 bounces; backward image from 1024^2 camera rays, 4 bounces,
    reorder, skydome + retina-weighted light-map shading. Both passes use the tree BV hitter."""
    import numpy as np
    import torch

    from splat_amd import abi, api, bv, scene

    sc = scene.scene_c2(seed=2)
    tris = api.to_device(sc.ftris)
    bvs, mem, n_top = bv.tree_partition(sc.ftris)
    hitter = api.HIPBVHitter(tris, bvs, mem, n_top=n_top)
    light = scene.camera_c2(seed=2, jitter=0.025)
    cam = scene.view_camera(1024, 1024, pos=(3.5, -4.5, 3.0), look_at=(0.0, 0.0, -0.5), seed=12)
    sky = torch.from_numpy(scene.skydome(512, 256)).cuda()
    r = api.Renderer(forward_hitter=hitter, backward_hitter=hitter, cam=cam, light=light, tris=tris, depth=4,
                     first_diffuse=sc.first_diffuse, spectrum=np.arange(400.0, 701.0, 25.0, dtype=np.float32),
                     tex_width=1024, tex_height=1024, skydome=sky, seed=5)
    for _ in range(3):
        r.render()
    torch.cuda.synchronize()
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(frames + 1)]
    evs[0].record()
    for i in range(frames):
        r.render()
        evs[i + 1].record()
    torch.cuda.synchronize()
    ms = np.array([evs[i].elapsed_time(evs[i + 1]) for i in range(frames)])
    return {"workload": "trace! iteration: 1024^2 light rays -> 13 x 1024^2 spectral light map; 1024^2 camera rays "
                        "-> reorder -> shaded RGB image; 4 bounces each, 10,243 triangles",
            "frames": frames, "p50_ms": round(float(np.percentile(ms, 50)), 3),
            "p99_ms": round(float(np.percentile(ms, 99)), 3),
            "mrays_per_s": round(2 * (1 << 20) / float(np.mean(ms)) / 1e3, 1)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-bv-build", action="store_true")
    ap.add_argument("--no-bf-sweep", action="store_true")
    ap.add_argument("--bf-sweep-only", action="store_true",
                    help="only the brute-force hitter sweep (for a kernel profile); prints its table")
    ap.add_argument("--profile", action="store_true", help="frames only: no instrumentation or CPU baseline")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only a 256 MiB device-to-device copy, repeated: calibrates HBM byte counters")
    ap.add_argument("--flop-probe", action="store_true",
                    help="only a kernel of known FP32 operation count, repeated: calibrates the VALU FLOP counters")
    ap.add_argument("--latency-frames", type=int, default=200, help="frames in the latency sample (p50/p99)")
    ap.add_argument("--tune-cache", default=None,
                    help="autotuner JSON cache (a profiling run given the bench's cache traces no tuning frames)")
    ap.add_argument("--profile-bf", action="store_true",
                    help="only the 320-triangle brute-force frame, --steps times (for a kernel profile)")
    ap.add_argument("--profile-batched", action="store_true",
                    help="only the 64-frame batch through the engine, --steps times (for a kernel profile)")
    ap.add_argument("--profile-operators", action="store_true",
                    help="only the standalone partition! and continuum_light_map2! on the 1 MP frame's rays, "
                         "--steps timed calls each (for a kernel profile)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path (--gpus must be 1)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from splat_amd import abi

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 2
    if args.copy_probe:
        n = 256 << 20
        a = torch.ones(n // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(max(args.steps, 1)):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n, "copy_probe_bytes_written": n, "copies": max(args.steps, 1)}))
        return 0
    if args.flop_probe:
        P = C.CDLL(os.path.join(ROOT, "splat_amd", "libsplat_probe.so"))
        P.splat_probe_fma.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p]
        P.splat_probe_fma.restype = C.c_int32
        P.splat_probe_fma_flops.argtypes = [C.c_int64, C.c_int32]
        P.splat_probe_fma_flops.restype = C.c_int64
        n, iters = 1 << 24, 1024
        out = torch.empty(n, dtype=torch.float32, device="cuda")
        for _ in range(max(args.steps, 1)):
            if P.splat_probe_fma(out.data_ptr(), n, iters, None) != 0:
                raise RuntimeError("probe launch failed")
        torch.cuda.synchronize()
        print(json.dumps({"probe": "k_probe_fma", "threads": n, "iters": iters,
                          "flops_per_launch": int(P.splat_probe_fma_flops(n, iters)), "launches": max(args.steps, 1)}))
        return 0
    from splat_amd import api

    if args.bf_sweep_only:
        print(json.dumps(bf_sweep()))
        return 0
    if args.profile_bf:
        from splat_amd import scene

        sc = scene.scene_c1()
        tris = api.to_device(sc.ftris)
        cam1 = scene.camera_c1(seed=1)
        fr = api.Frame(cam1, bf_frame_hitter(tris, cam1, args.tune_cache)[0], tris,
                       abi.run_params(2, sc.first_diffuse, math.nan, seed=2024), abi.lightmap_params(512, 512))
        for _ in range(max(args.steps, 1)):
            fr.trace()
        torch.cuda.synchronize()
        print(json.dumps({"frames": max(args.steps, 1), "workload": "320-triangle brute-force frame"}))
        return 0
    if args.profile_batched:
        from splat_amd import serve

        scenes, reqs = serve.batched_workload(64, tune_cache=args.tune_cache)
        arena = int(sum(r.rays for r in reqs) * 440 * 1.05) + (1 << 28)
        eng = api.Engine(arena, max_slots=64)
        walls = []
        for _ in range(max(args.steps, 1)):
            for r in reqs:
                r.done = float("nan")
            walls.append(serve.run_requests(eng, scenes, reqs, realtime=False, max_pending_steps=8).wall_s)
        eng.close()
        print(json.dumps({"batches": len(walls), "batched_s": walls}))
        return 0
    sc, cam, tuned, hitter, params, lm, frame = build_workload(tune_cache=args.tune_cache)
    R = frame.R
    if args.profile_operators:
        reps = max(args.steps, 1)
        print(json.dumps({"partition_operator": partition_operator(cam, hitter, frame.tris, params, reps=reps),
                          "splat_operator": splat_operator(sc, cam, hitter, frame.tris, params, lm, reps=reps)}))
        return 0
    for _ in range(args.warmup):
        frame.trace()
    torch.cuda.synchronize()

    K = args.steps
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(K + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evs[0].record()
    for i in range(K):
        frame.trace()
        evs[i + 1].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    value = R * K / wall / 1e6
    if args.profile:
        print(json.dumps({"frames": K, "ms_per_step": wall / K * 1e3}))
        return 0
    # frame latency over a fixed sample of frames (not --steps): each frame's device time between
    # events recorded before and after it on its stream, frames issued back to back
    NL = max(args.latency_frames, 1)
    lev = [torch.cuda.Event(enable_timing=True) for _ in range(NL + 1)]
    lev[0].record()
    for i in range(NL):
        frame.trace()
        lev[i + 1].record()
    torch.cuda.synchronize()
    frame_ms = np.array([lev[i].elapsed_time(lev[i + 1]) for i in range(NL)])
    # the same frames keeping only their light map (the served product; no final rays or hit
    # indices, so no retired-ray writes and no realigning hit), timed the same way
    mfr = api.Frame(cam, hitter, frame.tris, params, lm, keep_rays=False)
    for _ in range(args.warmup):
        mfr.trace()
    torch.cuda.synchronize()
    t0m = time.perf_counter()
    for _ in range(K):
        mfr.trace()
    torch.cuda.synchronize()
    wall_map = time.perf_counter() - t0m
    del mfr
    # the same frames two in flight (api.FrameLoop: each buffer traced on its own stream, as the
    # reference's double-buffered scene loop runs frames A and B as @async tasks), K frames;
    # reported beside the headline, which keeps one frame per launch sequence. (More buffers gain
    # only when the runtime puts their streams on distinct hardware queues, which it chooses.)
    in_flight = {}
    for nf in (2,):
        loop = api.FrameLoop(cam, hitter, frame.tris, params, lm, in_flight=nf)
        for _ in range(max(args.warmup, nf)):
            loop.trace()
        loop.wait()
        t0f = time.perf_counter()
        for _ in range(K):
            loop.trace()
        loop.wait()
        wall_f = time.perf_counter() - t0f
        in_flight[str(nf)] = {"mrays_per_s": round(R * K / wall_f / 1e6, 1),
                              "ms_per_frame": round(wall_f / K * 1e3, 4),
                              "vs_one_in_flight": round(wall / wall_f, 3)}
        del loop

    # ---- instrumented frames: hitter and fused bounce-pass launches timed with events recorded
    # on the frame's own stream around each launch ----
    hip = Hip()
    D = params.depth
    n_launch = D + 1
    M = 20
    per = 4 * D + 2
    hev = hip.events(per * M)
    for m in range(M):
        params.hit_events = C.addressof(hev) + m * per * C.sizeof(C.c_void_p)
        params.n_hit_events = per
        frame.trace()
    torch.cuda.synchronize()
    ms = lambda m, i: hip.elapsed_ms(hev[m * per + i], hev[m * per + i + 1])  # noqa: E731
    launch_ms = [ms(m, 2 * j) for m in range(M) for j in range(n_launch)]
    pass_ms = [[ms(m, 2 * n_launch + 2 * j) for j in range(D)] for m in range(M)]
    params.hit_events, params.n_hit_events = None, 0
    hip.destroy(hev)
    avg_launch_s = float(np.mean(launch_ms)) / 1e3
    hitter_frame_ms = float(np.sum(launch_ms)) / M

    # ---- algorithmic bytes of the fused bounce passes (evolve + splat + stable partition): per
    # live ray its key (8 B), the ray read (136 B; not at the first bounce, which generates its
    # rays) and written (136 B) once, the hit index of each retiring ray (4 B); per frame the light
    # map read and written once and every triangle gathered once ----
    live = live_counts(cam, hitter, frame.tris, params)  # A_0 = R, A_1, ..., A_D
    fin = api.from_device(frame.rays, abi.ADRAY)
    n_diffuse = int((fin["status"] == abi.RAY_DIFFUSE).sum())
    tex_bytes = lm.width * lm.height * lm.channels * 4
    pass_bytes = sum((144 if k == 0 else 280) * live[k] + 4 * (live[k] - live[k + 1]) for k in range(D)) \
        + 2 * tex_bytes + sc.n_tris * 120
    pass_s = float(np.mean([sum(x) for x in pass_ms])) / 1e3

    # ---- algorithmic work of the hitter (one counting frame, untimed) ----
    stats = torch.zeros(5, dtype=torch.int64, device="cuda")
    hitter.hs.d_stats = stats.data_ptr()
    frame.trace()
    torch.cuda.synchronize()
    hitter.hs.d_stats = None
    box_tests, entered, alg_tests, exec_tests, box_exec = (int(x) for x in stats.cpu().tolist())
    # algorithmic: each ray's own path (boxes of the nodes it entered, members of the leaves it
    # entered); executed: every test the wavefronts ran, 64 lanes each
    flop_per_launch = (FLOP_PER_BOX_TEST * box_tests + FLOP_PER_TRI_TEST * alg_tests) / n_launch
    flop_exec = (FLOP_PER_BOX_TEST * box_exec + FLOP_PER_TRI_TEST * exec_tests) / n_launch
    achieved = flop_per_launch / avg_launch_s / 1e12
    traffic, traffic_src = hitter_traffic(hitter.depth)
    pmc_flop, pmc_src = hitter_pmc_flops(hitter.depth)
    # vector-instruction issue of the hitter and the fused bounce pass: counted instructions per
    # launch (committed PMC summary) over the event-timed launches of this run
    vc = valu_counts(hitter.depth)
    n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    valu_roofs = None
    if vc and len(vc["k_hit_bv"]) == n_launch and len(vc["k_bounce"]) == D:
        hit_ms = [float(np.mean(launch_ms[j::n_launch])) for j in range(n_launch)]  # launch j of every frame
        pass_ms_j = [float(np.mean([x[j] for x in pass_ms])) for j in range(D)]
        valu_roofs = {"hitter": valu_roofline("k_hit_bv", vc["k_hit_bv"], hit_ms, n_cu),
                      "bounce_pass": valu_roofline("k_bounce", vc["k_bounce"], pass_ms_j, n_cu),
                      "source": vc.get("source")}

    splat_op = splat_operator(sc, cam, hitter, frame.tris, params, lm)
    part_op = partition_operator(cam, hitter, frame.tris, params)
    del frame
    torch.cuda.empty_cache()
    sweep = None if args.no_bf_sweep else bf_sweep()
    batched = None if args.no_batched else batched_vs_plain(tune_cache=args.tune_cache)
    torch.cuda.empty_cache()
    serve_out = None if args.no_served else served(tune_cache=args.tune_cache)
    torch.cuda.empty_cache()
    serve_1mp = None if args.no_served else served_1mp(tune_cache=args.tune_cache)
    torch.cuda.empty_cache()
    render = None if args.no_served else render_frame()
    bvb = None if args.no_bv_build else bv_build()
    large = None if args.no_served else large_frame()
    cpu = None if args.no_cpu_baseline else cpu_baseline(sc, cam, params, lm)
    small = None if args.no_cpu_baseline else small_frame(tune_cache=args.tune_cache)
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(wall / K * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded procedural scene and pinhole rays (no dataset)",
        "config": {
            "workload": "1 MP light-map frame: 2 displaced icospheres (subdiv 4, 10,240 glass tris) + diffuse plane, "
                        "BV hitter (the 2 ground triangles tested directly, binary tree over 2^%d object-median leaves of "
                        "the rest, built on the device), 4 bounces, dual-number differentials, 1024x1024x3 f32 "
                        "light map; plain path (one frame per launch sequence)" % hitter.depth,
            "autotuned": tuned,
            "rays_per_frame": R,
            "triangles": sc.n_tris,
            "bounces": params.depth,
            "path": "plain",
        },
        "ray_segments_per_s_M": round(sum(live[:D]) * K / wall / 1e6, 3),
        "p50_ms": round(float(np.percentile(frame_ms, 50)), 4),
        "p99_ms": round(float(np.percentile(frame_ms, 99)), 4),
        "latency_frames": int(frame_ms.size),
        "value_map_only": round(R * K / wall_map / 1e6, 3),
        "ms_per_step_map_only": round(wall_map / K * 1e3, 4),
        "frames_in_flight": in_flight,
        "roofline": {
            "kernel": "k_hit_bv (bounding-volume hitter)",
            "bound": "valu",  # per-ray FP32 vector ALU work (no matrix instructions); peak: FP32 vector
            "pipe": "fp32 VALU (per-ray scalar FP32 box and triangle tests; no matrix instructions)",
            "achieved": round(achieved, 3),
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
            # the FP32 peak counts packed FMA; these kernels are built without packed FP32 ops
            "frac_of_unpacked_peak": round(achieved / (FP32_PEAK_TFLOPS / 2), 4),
            "traffic": traffic,
            "traffic_source": traffic_src,
            "flop_per_launch": flop_per_launch,
            "flop_per_launch_executed": flop_exec,
            "executed_tflops": round(flop_exec / avg_launch_s / 1e12, 3),
            "pmc_flop_per_launch": pmc_flop,
            "pmc_source": pmc_src,
            "avg_launch_ms": round(avg_launch_s * 1e3, 4),
        },
        "rooflines": {
            "bounce_pass": {"kernel": "k_bounce (evolve + splat + stable partition, fused)", "bound": "hbm",
                            "achieved": round(pass_bytes / pass_s / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(pass_bytes / pass_s / 1e9 / HBM_PEAK_GBS, 4), "bytes_per_frame": pass_bytes,
                            "ms_per_frame": round(pass_s * 1e3, 4), "live_before_bounce": live,
                            "diffuse_rays": n_diffuse},
            "splat_operator": splat_op,
            "partition_operator": part_op,
            "valu_issue": valu_roofs,
        },
        "hitter": {
            "ms_per_frame": round(hitter_frame_ms, 4),
            "box_tests_alg_per_frame": box_tests,
            "box_tests_exec_per_frame": box_exec,
            "tri_tests_alg_per_frame": alg_tests,
            "tri_tests_exec_per_frame": exec_tests,
            "simd_efficiency_tri": round(alg_tests / max(exec_tests, 1), 4),
            "simd_efficiency_box": round(box_tests / max(box_exec, 1), 4),
        },
        "batched": batched,
        "served": serve_out,
        "served_1mp": serve_1mp,
        "render": render,
        "bv_build": bvb,
        "bf_sweep": sweep,
        "large_scene": large,
        "cpu_baseline": cpu,
        "small_frame": small,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
