"""Host-side mirror of the reference's operator interface for the per-frame hot path.

The reference plugs algorithms together through Julia multiple dispatch on hitter/tracer types
(src/alg_types.jl): `next_hit!(tracer, hitter, rays, n_tris)` (src/hitters.jl), the
`evolve_ray` broadcast (src/tracer.jl:181-201), `partition!(rays, swap; by)`
(src/partition.jl:122-132), `run_evolution!(hitter, tracer; ...)` (src/tracer.jl:233-282) and
`continuum_light_map2!(; tracer, tris, rays, ...)` (src/ray_imagers.jl:142-165). The same names
(with `!` spelled `_`), the same argument meanings and in-place semantics are kept here; every
call goes through the C ABI of libsplat_hip.so. Device arrays are torch tensors holding the
reference's isbits records as bytes: rays (R, 136) uint8, n_tris (N, 52), tris (N, 120).
Indices stay 1-based Julia indices. Errors raise SplatError (never a silent CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import abi, bv as bvmod
from . import lib as L


def _torch():
    import torch

    return torch


def to_device(a: np.ndarray, device: str = "cuda"):
    """Structured numpy array -> (n, itemsize) uint8 device tensor (byte-identical)."""
    torch = _torch()
    a = np.ascontiguousarray(a)
    raw = a.view(np.uint8).reshape(a.shape[0], a.dtype.itemsize)
    return torch.from_numpy(raw.copy()).to(device)


def from_device(t, dtype) -> np.ndarray:
    raw = t.detach().to("cpu").contiguous().numpy()
    return raw.reshape(-1).view(dtype).copy()


def _nrows(t) -> int:
    return int(t.shape[0])


def _check_rows(t, itemsize: int, what: str) -> None:
    torch = _torch()
    if t.dtype != torch.uint8 or t.dim() != 2 or t.shape[1] != itemsize or not t.is_contiguous() or not t.is_cuda:
        raise ValueError(f"{what} must be a contiguous (n, {itemsize}) uint8 CUDA tensor")


def _ws(nbytes: int, device="cuda"):
    torch = _torch()
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)


# ---- hitters (src/alg_types.jl:8-90) --------------------------------------------------------

class HIPBruteForceHitter:
    """Brute-force hitter: the results of StableHitter and ExperimentalHitter{,3,6}
    (src/hitters.jl:24-30, :34-82, :168-310) -- lexicographic (t, index) minimum, ties to the
    lowest index. Holds the prepared 64-byte triangle records as its scratch."""

    def __init__(self, n_tris, tune_waves: int = 0):
        L.require_gpu()
        lib = L.lib()
        self.hs = abi.HitSet()
        if n_tris.shape[1] == abi.NTRI.itemsize:
            _check_rows(n_tris, abi.NTRI.itemsize, "n_tris")
            self.ws = _ws(lib.splat_hitset_bf_bytes(_nrows(n_tris)))
            L.check(lib.splat_hitset_prepare_bf(L.ptr(n_tris), _nrows(n_tris), L.ptr(self.ws), self.ws.numel(),
                                                L.ref(self.hs), L.stream_handle()), "prepare_bf")
        else:
            _check_rows(n_tris, abi.FTRI.itemsize, "tris")
            self.ws = _ws(lib.splat_hitset_bf_bytes(_nrows(n_tris)))
            L.check(lib.splat_hitset_prepare_bf_ftri(L.ptr(n_tris), _nrows(n_tris), L.ptr(self.ws), self.ws.numel(),
                                                     L.ref(self.hs), L.stream_handle()), "prepare_bf_ftri")
        self.hs.tune_waves = tune_waves


class HIPBVHitter:
    """One-level bounding-volume hitter: DPBVHitter(A, ray_length, tris, bvs, memberships)
    (src/alg_types.jl:81-90, src/hitters.jl:615-648). Memberships are packed as pack_bv_tris
    does (src/bv.jl:126-136); the result equals the brute-force hitter's. The first n_top
    memberships are tested by every ray directly (bv.tree_partition puts large triangles there)."""

    def __init__(self, tris, bvs=None, memberships=None, depth: int = 6, host_tris: np.ndarray | None = None,
                 group_size: int = 0, n_top: int = 0):
        L.require_gpu()
        _check_rows(tris, abi.FTRI.itemsize, "tris")
        if memberships is None:
            if host_tris is None:
                host_tris = from_device(tris, abi.FTRI)
            bvs, memberships = bvmod.bv_partition(host_tris, depth)
        self.bvs = bvs
        counts, mat, stride = bvmod.pack_bv_tris(memberships, _nrows(tris))
        self.counts, self.mat, self.stride = counts, np.asfortranarray(mat), stride
        lib = L.lib()
        total = int(counts.sum())
        self.ws = _ws(lib.splat_hitset_bv_bytes(counts.shape[0], total))
        self.hs = abi.HitSet()
        L.check(lib.splat_hitset_prepare_bv(L.ptr(tris), _nrows(tris), counts.shape[0],
                                            counts.ctypes.data_as(C.c_void_p),
                                            self.mat.ctypes.data_as(C.c_void_p), stride, group_size, n_top, L.ptr(self.ws),
                                            self.ws.numel(), L.ref(self.hs), L.stream_handle()), "prepare_bv")

    @classmethod
    def build_on_device(cls, tris, depth: int | None = None, big_factor: float = 16.0, max_big: int = 32,
                        walk: int = 0):
        """The tree hitter's BVs built on the device from the triangles alone
        (splat_hitset_build_bv): the memberships bv.tree_partition(depth, big_factor, max_big)
        makes on the host, with the large-triangle list always present (possibly empty) and
        empty leaves kept. depth defaults to bv.tree_depth of all the triangles. `walk` is the
        tree walk mode (HitSet.tune_waves: 0 waves together, 1 per wave by direction octant, 2 per
        lane; + 4 never the flat tail scan; + 256 x the split-target code of autotune.SPLIT_TARGETS),
        kept across rebuilds."""
        L.require_gpu()
        _check_rows(tris, abi.FTRI.itemsize, "tris")
        self = cls.__new__(cls)
        self.bvs = self.counts = self.mat = self.stride = None
        N = _nrows(tris)
        self.depth = int(depth) if depth is not None else bvmod.tree_depth(N - 1)
        self.big_factor, self.max_big, self.walk = float(big_factor), int(max_big), int(walk)
        nbytes = L.lib().splat_hitset_build_bv_bytes(N, self.depth)
        if nbytes == 0:
            raise ValueError(f"device BV build: unsupported size (N={N}, depth={self.depth})")
        self.ws = _ws(nbytes, tris.device)
        self.hs = abi.HitSet()
        self.rebuild(tris)
        return self

    def rebuild(self, tris) -> None:
        """Rebuilds a device-built hit set in place for moved triangles (same count): an animated
        scene's per-frame bv_partition (src/scenes/splash_scene.jl:41)."""
        if getattr(self, "depth", None) is None:
            raise ValueError("rebuild needs a hitter made by build_on_device")
        L.check(L.lib().splat_hitset_build_bv(L.ptr(tris), _nrows(tris), self.depth, self.big_factor, self.max_big,
                                              L.ptr(self.ws), self.ws.numel(), L.ref(self.hs), L.stream_handle()),
                "build_bv")
        self.hs.tune_waves = self.walk


class HIPTracer:
    """ExperimentalTracer(A, rays, N) (src/alg_types.jl:101-115): owns hit_idx (Int32, 1-based),
    rndm, ray_swap and the sub-pixel offsets delta = upres(N); plus the device workspace of the
    fused bounce loop."""

    def __init__(self, rays, N: int = 1):
        torch = _torch()
        _check_rows(rays, abi.ADRAY.itemsize, "rays")
        R = _nrows(rays)
        self.hit_idx = torch.ones(R, dtype=torch.int32, device=rays.device)
        self.rndm = torch.empty(R, dtype=torch.float32, device=rays.device)
        self.ray_swap = torch.empty_like(rays)
        self.delta = abi.upres(N)
        self.N = N
        self._run_ws = None
        self._part_ws = None
        self._hit_ws = None
        self.d_count = torch.zeros(1, dtype=torch.int64, device=rays.device)

    def run_ws(self, R: int):
        if self._run_ws is None or self._run_ws.numel() < L.lib().splat_run_ws_bytes(R):
            self._run_ws = _ws(L.lib().splat_run_ws_bytes(R))
        return self._run_ws

    def hit_ws(self, R: int):
        if self._hit_ws is None or self._hit_ws.numel() < L.lib().splat_next_hit_ws_bytes(R):
            self._hit_ws = _ws(L.lib().splat_next_hit_ws_bytes(R))
        return self._hit_ws

    def part_ws(self, R: int):
        if self._part_ws is None or self._part_ws.numel() < L.lib().splat_partition_ws_bytes(R):
            self._part_ws = _ws(L.lib().splat_partition_ws_bytes(R))
        return self._part_ws


# ---- operators ----------------------------------------------------------------------------------

def next_hit_(tracer: HIPTracer, hitter, rays, n_tris=None, d_count=None) -> None:
    """next_hit!(tracer, hitter, rays, n_tris): tracer.hit_idx[r] <- nearest hit index (1 = miss)."""
    _check_rows(rays, abi.ADRAY.itemsize, "rays")
    R = _nrows(rays)
    ws = tracer.hit_ws(R)
    L.check(L.lib().splat_next_hit(L.ptr(rays), R, L.ptr(d_count), L.ref(hitter.hs), L.ptr(tracer.hit_idx),
                                   L.ptr(ws), ws.numel(), L.stream_handle()), "next_hit")


def evolve_(rays, hit_idx, tris, *, first_diffuse: int, rndm=None, force_rand=None, seed: int = 0, bounce: int = 1,
            optics: abi.Optics | None = None, d_count=None) -> None:
    """rays .= evolve_ray.(rays, hit_idx, tris[hit_idx], rndm, first_diffuse) (src/tracer.jl:227, :261)."""
    _check_rows(rays, abi.ADRAY.itemsize, "rays")
    _check_rows(tris, abi.FTRI.itemsize, "tris")
    opt = optics if optics is not None else abi.optics()
    fr = math.nan if force_rand is None else float(force_rand)
    L.check(L.lib().splat_evolve(L.ptr(rays), _nrows(rays), L.ptr(d_count), L.ptr(hit_idx), L.ptr(tris), _nrows(tris),
                                 L.ptr(rndm), fr, seed, bounce, first_diffuse, L.ref(opt), L.stream_handle()), "evolve")


def partition_(vals, swap, tracer: HIPTracer | None = None, host_count: bool = True):
    """partition!(vals, swap; by = status != ACTIVE) -> number of ACTIVE rays, now first (stable)."""
    _check_rows(vals, abi.ADRAY.itemsize, "vals")
    torch = _torch()
    R = _nrows(vals)
    ws = tracer.part_ws(R) if tracer is not None else _ws(L.lib().splat_partition_ws_bytes(R))
    d = tracer.d_count if tracer is not None else torch.zeros(1, dtype=torch.int64, device=vals.device)
    h = C.c_int64(0)
    L.check(L.lib().splat_partition_active(L.ptr(vals), L.ptr(swap), R, L.ptr(d), C.byref(h) if host_count else None,
                                           L.ptr(ws), ws.numel(), L.stream_handle()), "partition")
    return int(h.value) if host_count else d


def run_evolution_(hitter, tracer: HIPTracer, *, depth: int, first_diffuse: int, tris, rays, n_tris=None,
                   force_rand=None, seed: int = 0, optics: abi.Optics | None = None, reorder: bool = False,
                   d_live=None) -> None:
    """run_evolution!(hitter, ::ExperimentalTracer; depth, first_diffuse, n_tris, tris, rays,
    reorder, force_rand) (src/tracer.jl:233-282), with every count kept on the device."""
    if reorder:
        raise NotImplementedError("reorder=true belongs to the backward camera pass, not the light-map path")
    _check_rows(rays, abi.ADRAY.itemsize, "rays")
    _check_rows(tris, abi.FTRI.itemsize, "tris")
    R = _nrows(rays)
    p = abi.run_params(depth, first_diffuse, math.nan if force_rand is None else float(force_rand), seed, optics)
    ws = tracer.run_ws(R)
    L.check(L.lib().splat_run_evolution(L.ptr(rays), R, L.ref(hitter.hs), L.ptr(tris), _nrows(tris),
                                        L.ptr(tracer.hit_idx), L.ref(p), L.ptr(ws), ws.numel(),
                                        L.ptr(d_live if d_live is not None else tracer.d_count), L.stream_handle()),
            "run_evolution")


def continuum_light_map2_(*, tracer: HIPTracer, tris, rays, tex, spectrum=None, mode: str = "rgb", lm=None,
                          d_count=None, **_ignored) -> None:
    """continuum_light_map2!(; tracer, tris, rays, spectrum, tex, ...) (src/ray_imagers.jl:142-165):
    final_evolution of DIFFUSE rays, then the atomic splat into tex (C, H, W) float32, which is
    accumulated into (the reference initialises it with tex_f())."""
    torch = _torch()
    _check_rows(rays, abi.ADRAY.itemsize, "rays")
    if tex.dtype != torch.float32 or tex.dim() != 3 or not tex.is_contiguous():
        raise ValueError("tex must be a contiguous (channels, height, width) float32 tensor")
    if lm is None:
        lm = abi.lightmap_params(int(tex.shape[2]), int(tex.shape[1]), n_delta=tracer.N, spectrum=spectrum, mode=mode)
    if lm.channels != tex.shape[0] or lm.height != tex.shape[1] or lm.width != tex.shape[2]:
        raise ValueError("tex shape does not match the light-map parameters")
    L.check(L.lib().splat_light_map(L.ptr(rays), _nrows(rays), L.ptr(d_count), L.ptr(tracer.hit_idx), L.ptr(tris),
                                    _nrows(tris), L.ref(lm), L.ptr(tex), L.stream_handle()), "light_map")


def rays_from_pinhole(cam: abi.Pinhole, device: str = "cuda"):
    """Seeded pinhole ray source (camera_ray convention, src/ray_generators.jl:18-62)."""
    torch = _torch()
    L.require_gpu()
    R = cam.height * cam.width * cam.samples_per_pixel
    rays = torch.empty((R, abi.ADRAY.itemsize), dtype=torch.uint8, device=device)
    L.check(L.lib().splat_raygen_pinhole(L.ref(cam), L.ptr(rays), R, L.stream_handle()), "raygen")
    return rays


def rays_from_lights(lights, params: abi.LightRays | None = None, device: str = "cuda"):
    """rays_from_lights(lights, upscale) (src/ray_generators.jl:64-113) on the device: `lights` is
    a sequence of abi.RectLight; each emits (res1/upscale) x (res2/upscale) rays, column-major,
    light after light. The reference's unseeded rand(Float32) cell offsets come from `params.seed`."""
    torch = _torch()
    L.require_gpu()
    params = params if params is not None else abi.light_rays()
    arr = (abi.RectLight * max(len(lights), 1))(*lights)
    R = L.lib().splat_lights_ray_count(arr, len(lights), params.upscale)
    if R < 0:
        L.check(int(R), "rays_from_lights")
    rays = torch.empty((R, abi.ADRAY.itemsize), dtype=torch.uint8, device=device)
    L.check(L.lib().splat_raygen_lights(arr, len(lights), L.ref(params), L.ptr(rays), R, L.stream_handle()),
            "rays_from_lights")
    return rays


def reorder_(rays, tracer: HIPTracer) -> None:
    """run_evolution!(...; reorder = true)'s reorder (src/tracer.jl:270-278): rays and
    tracer.hit_idx back to pixel order (slot dest - 1)."""
    _check_rows(rays, abi.ADRAY.itemsize, "rays")
    R = _nrows(rays)
    ws = _ws(L.lib().splat_reorder_ws_bytes(R), rays.device)
    L.check(L.lib().splat_reorder_by_dest(L.ptr(rays), R, L.ptr(tracer.hit_idx), L.ptr(ws), ws.numel(),
                                          L.stream_handle()), "reorder")


def continuum_shade_(*, tracer: HIPTracer, tris, rays, tex, lm: abi.LightMapParams, grid_h: int, skydome=None,
                     rgb=None):
    """continuum_shade!(::ExperimentalImager2; tracer, RGB, tris, rays, spectrum, retina_factor,
    tex, skydome) (src/ray_imagers.jl:440-503): rays in pixel order on a grid_h-row grid ->
    RGB image (grid_h*n_delta x R/grid_h*n_delta pixels, row index fastest) as an (n_pixels, 3)
    float32 tensor. tex: the spectral light map (lm.n_lambda planes); lm.weights: the retina
    factors; skydome: (3, sky_h, sky_w) float32 or None."""
    torch = _torch()
    _check_rows(rays, abi.ADRAY.itemsize, "rays")
    _check_rows(tris, abi.FTRI.itemsize, "tris")
    R = _nrows(rays)
    nd = lm.n_delta
    if rgb is None:
        rgb = torch.empty((R * nd * nd, 3), dtype=torch.float32, device=rays.device)
    sw = sh = 0
    if skydome is not None:
        sh, sw = int(skydome.shape[1]), int(skydome.shape[2])
    L.check(L.lib().splat_shade(L.ptr(rays), R, L.ptr(tracer.hit_idx), L.ptr(tris), _nrows(tris), L.ref(lm),
                                L.ptr(tex), L.ptr(skydome), sw, sh, grid_h, L.ptr(rgb), L.stream_handle()), "shade")
    return rgb


class Renderer:
    """trace! (src/tracer.jl:291-385) with its buffers kept between iterations: the forward light
    map -- rays from the light pinhole, run_evolution!, continuum_light_map2! into one plane per
    spectrum entry -- then the backward image -- camera rays (lambda 550, one per upscale x
    upscale pixel block), run_evolution!(reorder = true), continuum_shade! with upscale^2 sub-pixel
    offsets -- into `rgb` ((h*upscale) x (w*upscale) pixels x 3, row index fastest). The two
    passes are independent until shading, so the backward bounce loop runs on a second stream
    beside the forward pass (the reference runs the forward pass as an @async task,
    src/tracer.jl:317); iterations > 1 average several seeded iterations into `out`
    (iterations_per_frame, src/tracer.jl:311, :380-384)."""

    def __init__(self, *, forward_hitter, backward_hitter, cam: abi.Pinhole, tris, depth: int, first_diffuse: int,
                 spectrum, tex_width: int, tex_height: int, light: abi.Pinhole | None = None, lights=None,
                 light_params: abi.LightRays | None = None, skydome=None, seed: int = 0, force_rand=None,
                 upscale: int = 1):
        """The forward rays come from `lights` (abi.RectLight list, rays_from_lights with
        `light_params`, as trace! does) or from the point-light pinhole `light`."""
        torch = _torch()
        self.cam, self.tris, self.depth, self.first_diffuse = cam, tris, depth, first_diffuse
        self.backward_hitter, self.skydome, self.seed, self.force_rand = backward_hitter, skydome, seed, force_rand
        fr_p = abi.run_params(depth, first_diffuse, math.nan if force_rand is None else float(force_rand), seed=seed)
        if (light is None) == (lights is None):
            raise ValueError("give exactly one forward ray source: light (pinhole) or lights (RectLights)")
        if lights is not None:
            lp = light_params if light_params is not None else abi.light_rays()
            lm = abi.lightmap_params(tex_width, tex_height, n_delta=lp.upscale, spectrum=spectrum, mode="spectral")
            self.forward = LightFrame(lights, lp, forward_hitter, tris, fr_p, lm, device=tris.device)
            self.forward_source = lp
        else:
            lm = abi.lightmap_params(tex_width, tex_height, spectrum=spectrum, mode="spectral")
            self.forward = Frame(light, forward_hitter, tris, fr_p, lm, device=tris.device)
            self.forward_source = light
        self.shade_lm = abi.lightmap_params(tex_width, tex_height, n_delta=upscale, spectrum=spectrum, mode="spectral")
        R = cam.height * cam.width * cam.samples_per_pixel
        self.rays = torch.empty((R, abi.ADRAY.itemsize), dtype=torch.uint8, device=tris.device)
        self.tracer = HIPTracer(self.rays, upscale)
        self.rgb = torch.empty((R * upscale * upscale, 3), dtype=torch.float32, device=tris.device)
        self.out = torch.empty_like(self.rgb)
        self.reorder_ws = _ws(L.lib().splat_reorder_ws_bytes(R), tris.device)
        self.side = torch.cuda.Stream(device=tris.device)  # default priority

    @property
    def tex(self):
        return self.forward.tex

    def _iteration(self, it: int):
        torch = _torch()
        lib = L.lib()
        main = torch.cuda.current_stream()
        R = _nrows(self.rays)
        self.side.wait_stream(main)
        with torch.cuda.stream(self.side):  # backward bounce loop beside the forward pass
            s = L.stream_handle(self.side)
            self.cam.seed = self.cam_seed + it
            L.check(lib.splat_raygen_pinhole(L.ref(self.cam), L.ptr(self.rays), R, s), "raygen")
            run_evolution_(self.backward_hitter, self.tracer, depth=self.depth, first_diffuse=self.first_diffuse,
                           tris=self.tris, rays=self.rays, force_rand=self.force_rand, seed=self.seed + 2 * it + 1)
            L.check(lib.splat_reorder_by_dest(L.ptr(self.rays), R, L.ptr(self.tracer.hit_idx),
                                              L.ptr(self.reorder_ws), self.reorder_ws.numel(), s), "reorder")
        self.forward.params.seed = self.seed + 2 * it
        self.forward_source.seed = self.light_seed + it
        self.forward.trace()
        main.wait_stream(self.side)
        continuum_shade_(tracer=self.tracer, tris=self.tris, rays=self.rays, tex=self.forward.tex, lm=self.shade_lm,
                         grid_h=self.cam.height, skydome=self.skydome, rgb=self.rgb)

    def render(self, iterations: int = 1):
        """One frame: `iterations` seeded iterations averaged (out <- out (f-1)/f + rgb/f)."""
        self.cam_seed = getattr(self, "cam_seed", self.cam.seed)
        self.light_seed = getattr(self, "light_seed", self.forward_source.seed)
        if iterations == 1:
            self._iteration(0)
            return self.rgb
        for f in range(1, iterations + 1):
            self._iteration(f - 1)
            if f == 1:
                self.out.copy_(self.rgb)
            else:
                self.out.mul_((f - 1) / f).add_(self.rgb, alpha=1.0 / f)
        return self.out


def trace_(**kw):
    """One trace! iteration (Renderer.render); returns (rgb, light map)."""
    r = Renderer(**kw)
    rgb = r.render()
    return rgb, r.tex


class Frame:
    """Device-resident state of one light-map frame: rays, tracer arrays, workspace, light map.
    `trace(...)` runs pinhole -> run_evolution! -> continuum_light_map2! with no host sync.
    keep_rays=False keeps only the light map (rays and hit_idx are None): the final rays are not
    written and the realigning hit is skipped, as nothing reads them."""

    def __init__(self, cam: abi.Pinhole, hitter, tris, params: abi.RunParams, lm: abi.LightMapParams,
                 device: str = "cuda", keep_rays: bool = True):
        torch = _torch()
        L.require_gpu()
        self.cam, self.hitter, self.tris, self.params, self.lm = cam, hitter, tris, params, lm
        self.R = cam.height * cam.width * cam.samples_per_pixel
        self.rays = torch.empty((self.R, abi.ADRAY.itemsize), dtype=torch.uint8, device=device) if keep_rays else None
        self.hit_idx = torch.empty(self.R, dtype=torch.int32, device=device) if keep_rays else None
        self.tex = torch.zeros((lm.channels, lm.height, lm.width), dtype=torch.float32, device=device)
        self.ws = _ws(L.lib().splat_run_ws_bytes(self.R))

    def trace(self, stream=None) -> None:
        L.check(L.lib().splat_trace_frame(L.ref(self.cam), L.ptr(self.rays), self.R, L.ref(self.hitter.hs),
                                          L.ptr(self.tris), _nrows(self.tris), L.ptr(self.hit_idx), L.ref(self.params),
                                          L.ref(self.lm), L.ptr(self.tex), L.ptr(self.ws), self.ws.numel(),
                                          L.stream_handle(stream)), "trace_frame")


class FrameLoop:
    """The double-buffered frame loop of a scene script (src/scenes/sphere.jl:192-219, where frames
    A and B are @async tasks, each on its task's own stream, so that one frame's launches run
    while the other's finish), for `in_flight` buffers: each buffer is a Frame with its own stream.
    `trace(cam)` issues the next frame on the next buffer's stream, after that buffer's previous
    frame (the stream orders them), and returns the buffer without waiting; `wait(frame)` blocks
    until that buffer's frames are done (`wait()`: every buffer's). A buffer's tex, rays and
    hit_idx are those of its last frame: read them after `wait` and before the loop comes round to
    that buffer again. A frame's launches fill the idle tails of the others' (a hitter launch ends
    with a few long walks on an otherwise idle chip)."""

    def __init__(self, cam: abi.Pinhole, hitter, tris, params: abi.RunParams, lm: abi.LightMapParams,
                 in_flight: int = 2, device: str = "cuda", keep_rays: bool = True):
        torch = _torch()
        if in_flight < 1:
            raise ValueError(f"in_flight must be at least 1, not {in_flight}")
        self.frames = [Frame(cam, hitter, tris, params, lm, device, keep_rays) for _ in range(in_flight)]
        self.streams = [torch.cuda.Stream(device=device) for _ in range(in_flight)]  # default priority
        here = torch.cuda.current_stream(device)
        for f, s in zip(self.frames, self.streams):
            s.wait_stream(here)  # the buffers were allocated (the maps zeroed) on this stream
            for t in (f.rays, f.hit_idx, f.tex, f.ws):
                if t is not None:
                    t.record_stream(s)  # freed only after the work queued on s when they go
        self.next = 0

    def trace(self, cam: abi.Pinhole | None = None) -> Frame:
        i = self.next
        self.next = (i + 1) % len(self.frames)
        f = self.frames[i]
        if cam is not None:
            if cam.height * cam.width * cam.samples_per_pixel != f.R:
                raise ValueError("a FrameLoop's cameras keep the ray count its buffers were sized for")
            f.cam = cam  # passed by value at launch: the caller may change it after trace returns
        f.trace(self.streams[i])
        return f

    def wait(self, frame: Frame | None = None) -> None:
        for f, s in zip(self.frames, self.streams):
            if frame is None or f is frame:
                s.synchronize()


class LightFrame:
    """The forward pass of trace! from rectangular lights (src/tracer.jl:316-341): rays_from_lights
    -> run_evolution! -> continuum_light_map2!, one call, no host sync."""

    def __init__(self, lights, lp: abi.LightRays, hitter, tris, params: abi.RunParams, lm: abi.LightMapParams,
                 device: str = "cuda"):
        torch = _torch()
        L.require_gpu()
        self.lights = (abi.RectLight * max(len(lights), 1))(*lights)
        self.n_lights, self.lp = len(lights), lp
        self.hitter, self.tris, self.params, self.lm = hitter, tris, params, lm
        self.R = int(L.lib().splat_lights_ray_count(self.lights, self.n_lights, lp.upscale))
        if self.R < 0:
            L.check(self.R, "rays_from_lights")
        self.rays = torch.empty((self.R, abi.ADRAY.itemsize), dtype=torch.uint8, device=device)
        self.hit_idx = torch.empty(self.R, dtype=torch.int32, device=device)
        self.tex = torch.zeros((lm.channels, lm.height, lm.width), dtype=torch.float32, device=device)
        self.ws = _ws(L.lib().splat_run_ws_bytes(self.R), device)

    def trace(self, stream=None) -> None:
        L.check(L.lib().splat_trace_lights(self.lights, self.n_lights, L.ref(self.lp), L.ptr(self.rays), self.R,
                                           L.ref(self.hitter.hs), L.ptr(self.tris), _nrows(self.tris),
                                           L.ptr(self.hit_idx), L.ref(self.params), L.ref(self.lm), L.ptr(self.tex),
                                           L.ptr(self.ws), self.ws.numel(), L.stream_handle(stream)), "trace_lights")


class Engine:
    """Multi-frame engine (batching and continuous serving) over splat_engine_* of the C ABI:
    frames of different scenes, sizes and depths share every launch; a step advances each
    admitted frame one stage and replays one captured HIP graph."""

    def __init__(self, arena_bytes: int, max_slots: int = 256, use_graph: bool = True, use_queue: bool = False):
        L.require_gpu()
        self._lib = L.lib()
        h = C.c_void_p()
        L.check(self._lib.splat_engine_create(int(arena_bytes), int(max_slots), C.byref(h)), "engine_create")
        self.h = h
        L.check(self._lib.splat_engine_use_graph(self.h, 1 if use_graph else 0), "engine_use_graph")
        L.check(self._lib.splat_engine_use_queue(self.h, 1 if use_queue else 0), "engine_use_queue")
        self._keep = {}  # tag -> objects that must outlive the frame

    def set_use_queue(self, on: bool) -> None:
        """Selects the hitter of the following steps: one persistent launch over the device work
        queues (True) or one wave per item (False); the results are the same."""
        L.check(self._lib.splat_engine_use_queue(self.h, 1 if on else 0), "engine_use_queue")

    def frame_bytes(self, cam: abi.Pinhole, with_outputs: bool = False) -> int:
        return int(self._lib.splat_engine_frame_bytes(L.ref(cam), 1 if with_outputs else 0))

    def admit(self, *, cam: abi.Pinhole, hitter, tris, params: abi.RunParams, lm: abi.LightMapParams, tex, tag: int,
              rays_out=None, hit_out=None, stream=None) -> bool:
        """Queues a frame; False when the arena or the slot table is full (SPLAT_ERR_BUDGET)."""
        rq = abi.FrameRequest()
        rq.cam = cam
        rq.hs = hitter.hs
        rq.tris = tris.data_ptr()
        rq.n_tris = _nrows(tris)
        rq.depth = params.depth
        rq.first_diffuse = params.first_diffuse
        rq.force_rand = params.force_rand
        rq.seed = params.seed
        rq.optics = params.optics
        rq.tex = tex.data_ptr()
        rq.rays_out = None if rays_out is None else rays_out.data_ptr()
        rq.hit_out = None if hit_out is None else hit_out.data_ptr()
        rq.tag = int(tag)
        rq.lm = lm
        # admission enqueues nothing (the next step uploads it), so no stream is looked up here
        code = self._lib.splat_engine_admit(self.h, L.ref(rq), None if stream is None else L.stream_handle(stream))
        if code == -6:
            return False
        L.check(code, "engine_admit")
        self._keep[int(tag)] = (hitter, tris, tex, rays_out, hit_out)
        return True

    def step(self, stream=None) -> None:
        L.check(self._lib.splat_engine_step(self.h, L.stream_handle(stream)), "engine_step")

    def poll(self, wait: bool = False, max_tags: int = 4096) -> list[int]:
        if getattr(self, "_tag_buf", None) is None or len(self._tag_buf) < max_tags:
            self._tag_buf = (C.c_int64 * max_tags)()  # reused: polled once per step
            self._n_buf = C.c_int32()
        buf, n = self._tag_buf, self._n_buf
        L.check(self._lib.splat_engine_poll(self.h, 1 if wait else 0, C.cast(buf, C.c_void_p), max_tags, C.byref(n)),
                "engine_poll")
        tags = [int(buf[i]) for i in range(n.value)]
        for t in tags:
            self._keep.pop(t, None)
        return tags

    def synchronize(self) -> None:
        """Wait for all device work (the engine's steps run on the caller's streams)."""
        _torch().cuda.synchronize()

    def sync_steps(self, keep: int) -> None:
        """Block until at most `keep` launched steps are still running."""
        L.check(self._lib.splat_engine_sync_steps(self.h, int(keep)), "engine_sync_steps")

    @property
    def in_flight(self) -> int:
        return int(self._lib.splat_engine_in_flight(self.h))

    @property
    def runnable(self) -> int:
        """Frames the next step advances; 0 with frames in flight: only completions are pending."""
        return int(self._lib.splat_engine_runnable(self.h))

    @property
    def arena_free(self) -> int:
        return int(self._lib.splat_engine_arena_free(self.h))

    def close(self) -> None:
        if self.h:
            self._lib.splat_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
