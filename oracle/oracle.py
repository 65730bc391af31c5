"""ctypes binding of the CPU oracle (oracle/liboracle.so, built from oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker. The product package never imports it.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

from splat_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "liboracle.so"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        i32, i64, u64, f32, f64 = C.c_int32, C.c_int64, C.c_uint64, C.c_float, C.c_double
        sig = {
            "oracle_mix64": (u64, [u64]),
            "oracle_shade_f32": (None, [P, i64, P, P, i32, P, P, P, i32, i32, i64, P, i32]),
            "oracle_shade_f64": (None, [P, i64, P, P, i32, P, P, P, i32, i32, i64, P, i32]),
            "oracle_bilinear_f64": (f64, [P, i64, i64, f64, f64]),
            "oracle_reorder": (None, [P, i64, P]),
            "oracle_u01": (f32, [u64, u64, u64]),
            "oracle_key32": (C.c_uint32, [f32]),
            "oracle_key64": (u64, [f32, C.c_uint32]),
            "oracle_key32_inv": (f32, [C.c_uint32]),
            "oracle_key64_arg": (C.c_uint32, [u64]),
            "oracle_distance_to_plane_f64": (f64, [P, P, P, P]),
            "oracle_distance_to_plane_f32": (f32, [P, P, P, P]),
            "oracle_hit_test_f32": (i32, [P, P, P, i64, i64, P]),
            "oracle_hit_test_f64": (i32, [P, P, P, i64, i64, P]),
            "oracle_next_hit_f32": (None, [P, i64, P, i32, P, P, i32]),
            "oracle_next_hit_f64": (None, [P, i64, P, i32, P, P, i32]),
            "oracle_next_hit_bv_f32": (None, [P, i64, P, i32, i32, P, P, i64, P, i32]),
            "oracle_reflect_f64": (None, [P, P, P]),
            "oracle_reflect_f32": (None, [P, P, P]),
            "oracle_refract_f64": (None, [P, P, f64, f64, P]),
            "oracle_can_refract_f64": (i32, [P, P, f64, f64]),
            "oracle_reflectance_f64": (f64, [P, P, f64, f64]),
            "oracle_optical_normal_f64": (None, [P, P, P]),
            "oracle_reverse_uv_f64": (None, [P, P, P]),
            "oracle_glass_ior_f64": (f64, [P, f64, P]),
            "oracle_evolve_f32": (None, [P, i64, P, P, i32, P, f32, u64, i32, i32, P, i32]),
            "oracle_evolve_f64": (None, [P, i64, P, P, i32, P, f32, u64, i32, i32, P, P, i32]),
            "oracle_evolve64_f64": (None, [P, i64, P, P, i32, f32, i32, P, P]),
            "oracle_partition": (i64, [P, i64, P]),
            "oracle_final_evolution_f32": (None, [P, i64, P, P, i32]),
            "oracle_final_evolution_f64": (None, [P, i64, P, P, i32, P]),
            "oracle_light_map_f32": (None, [P, i64, P, P, i32, P, P]),
            "oracle_light_map_f64": (None, [P, i64, P, P, i32, P, P]),
            "oracle_raygen": (None, [P, P, i64]),
            "oracle_raygen_at": (None, [P, P, i64, P]),
            "oracle_splat_margins_f64": (None, [P, i64, P, P, i32, P, P, P]),
            "oracle_raygen_lights": (i64, [P, i32, P, P]),
            "oracle_run_evolution_f32": (i64, [P, i64, P, i32, P, P, i32, P, P, i64, i32]),
            "oracle_trace_frame_f32": (i64, [P, P, i64, P, i32, P, P, P, P, i32, P, P, i64, i32]),
            "oracle_trace_sample_f32": (i64, [P, i64, i64, i64, P, i32, P, P, P, i32, P, P, i64, i32]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _ref(s):
    return C.c_void_p(C.addressof(s))


# ---- wrappers -----------------------------------------------------------------------------

def next_hit(rays, ntris, precision="f32", threads=0):
    rays = np.ascontiguousarray(rays)
    ntris = np.ascontiguousarray(ntris)
    hit = np.empty(rays.shape[0], dtype=np.int32)
    if precision == "f32":
        t = np.empty(rays.shape[0], dtype=np.float32)
        lib().oracle_next_hit_f32(_p(rays), rays.shape[0], _p(ntris), ntris.shape[0], _p(hit), _p(t), threads)
    else:
        t = np.empty(rays.shape[0], dtype=np.float64)
        lib().oracle_next_hit_f64(_p(rays), rays.shape[0], _p(ntris), ntris.shape[0], _p(hit), _p(t), threads)
    return hit, t


def next_hit_bv(rays, ftris, counts, bv_tris, threads=0):
    rays = np.ascontiguousarray(rays)
    hit = np.empty(rays.shape[0], dtype=np.int32)
    mat = np.asfortranarray(bv_tris, dtype=np.int64)
    lib().oracle_next_hit_bv_f32(_p(rays), rays.shape[0], _p(ftris), ftris.shape[0], counts.shape[0],
                                 _p(counts), mat.ctypes.data_as(C.c_void_p), mat.shape[1], _p(hit), threads)
    return hit


def evolve(rays, hit_idx, ftris, *, force_rand=math.nan, rndm=None, seed=0, bounce=1, first_diffuse,
           opt=None, precision="f32", threads=0):
    opt = opt if opt is not None else abi.optics()
    rays = np.array(rays, copy=True)
    hit_idx = np.ascontiguousarray(hit_idx, dtype=np.int32)
    rndm = None if rndm is None else np.ascontiguousarray(rndm, dtype=np.float32)
    if precision == "f32":
        lib().oracle_evolve_f32(_p(rays), rays.shape[0], _p(hit_idx), _p(ftris), ftris.shape[0], _p(rndm),
                                force_rand, seed, bounce, first_diffuse, _ref(opt), threads)
        return rays
    out = np.zeros(rays.shape[0], dtype=abi.ADRAY64)
    lib().oracle_evolve_f64(_p(rays), rays.shape[0], _p(hit_idx), _p(ftris), ftris.shape[0], _p(rndm),
                            force_rand, seed, bounce, first_diffuse, _ref(opt), _p(out), threads)
    return out


def evolve64(rays64, hit_idx, ftris, *, force_rand, first_diffuse, opt=None):
    opt = opt if opt is not None else abi.optics()
    rays64 = np.ascontiguousarray(rays64)
    out = np.zeros(rays64.shape[0], dtype=abi.ADRAY64)
    hit_idx = np.ascontiguousarray(hit_idx, dtype=np.int32)
    lib().oracle_evolve64_f64(_p(rays64), rays64.shape[0], _p(hit_idx), _p(ftris), ftris.shape[0], force_rand,
                              first_diffuse, _ref(opt), _p(out))
    return out


def to64(rays):
    out = np.zeros(rays.shape[0], dtype=abi.ADRAY64)
    for f in abi.VEC_FIELDS + ["ignore_tri", "dest", "lam", "x", "y"]:
        out[f] = rays[f]
    out["in_medium"] = rays["in_medium"]
    out["status"] = rays["status"]
    return out


def partition(rays, carry=None):
    rays = np.array(rays, copy=True)
    c = None if carry is None else np.array(carry, dtype=np.int32, copy=True)
    n = lib().oracle_partition(_p(rays), rays.shape[0], _p(c))
    return rays, int(n), c


def final_evolution(rays, hit_idx, ftris, precision="f32"):
    hit_idx = np.ascontiguousarray(hit_idx, dtype=np.int32)
    if precision == "f32":
        rays = np.array(rays, copy=True)
        lib().oracle_final_evolution_f32(_p(rays), rays.shape[0], _p(hit_idx), _p(ftris), ftris.shape[0])
        return rays
    out = np.zeros(rays.shape[0], dtype=abi.ADRAY64)
    lib().oracle_final_evolution_f64(_p(np.ascontiguousarray(rays)), rays.shape[0], _p(hit_idx), _p(ftris),
                                     ftris.shape[0], _p(out))
    return out


def light_map(rays, hit_idx, ftris, lm, precision="f32"):
    hit_idx = np.ascontiguousarray(hit_idx, dtype=np.int32)
    shape = (lm.channels, lm.height, lm.width)
    if precision == "f32":
        rays = np.array(rays, copy=True)
        tex = np.zeros(shape, dtype=np.float32)
        lib().oracle_light_map_f32(_p(rays), rays.shape[0], _p(hit_idx), _p(ftris), ftris.shape[0], _ref(lm), _p(tex))
        return tex, rays
    tex = np.zeros(shape, dtype=np.float64)
    lib().oracle_light_map_f64(_p(np.ascontiguousarray(rays)), rays.shape[0], _p(hit_idx), _p(ftris),
                               ftris.shape[0], _ref(lm), _p(tex))
    return tex, None


def splat_margins(rays, hit_idx, ftris, lm):
    """FP64 binning margins of each ray's splats: (texel_margin, edge_margin) arrays, see
    oracle_splat_margins_f64."""
    rays = np.ascontiguousarray(rays)
    hit_idx = np.ascontiguousarray(hit_idx, dtype=np.int32)
    tm = np.empty(rays.shape[0], dtype=np.float64)
    em = np.empty(rays.shape[0], dtype=np.float64)
    lib().oracle_splat_margins_f64(_p(rays), rays.shape[0], _p(hit_idx), _p(ftris), ftris.shape[0], _ref(lm), _p(tm),
                                   _p(em))
    return tm, em


def raygen(cam):
    R = cam.height * cam.width * cam.samples_per_pixel
    rays = np.zeros(R, dtype=abi.ADRAY)
    lib().oracle_raygen(_ref(cam), _p(rays), R)
    return rays


def raygen_at(cam, q):
    """The pinhole frame's rays at generation indices q (0-based), in that order."""
    q = np.ascontiguousarray(q, dtype=np.int64)
    rays = np.zeros(q.shape[0], dtype=abi.ADRAY)
    lib().oracle_raygen_at(_ref(cam), _p(q), q.shape[0], _p(rays))
    return rays


def raygen_lights(lights, params):
    """rays_from_lights over a sequence of abi.RectLight with abi.LightRays parameters."""
    arr = (abi.RectLight * max(len(lights), 1))(*lights)
    up = params.upscale
    R = sum((L.res1 // up) * (L.res2 // up) for L in lights)
    rays = np.zeros(R, dtype=abi.ADRAY)
    n = lib().oracle_raygen_lights(arr, len(lights), _ref(params), _p(rays))
    assert n == R
    return rays


def run_evolution(rays, ftris, params, bv=None, threads=0):
    rays = np.array(rays, copy=True)
    hit = np.zeros(rays.shape[0], dtype=np.int32)
    if bv is None:
        live = lib().oracle_run_evolution_f32(_p(rays), rays.shape[0], _p(ftris), ftris.shape[0], _p(hit),
                                              _ref(params), 0, None, None, 0, threads)
    else:
        counts, mat, stride = bv
        mat = np.asfortranarray(mat, dtype=np.int64)
        live = lib().oracle_run_evolution_f32(_p(rays), rays.shape[0], _p(ftris), ftris.shape[0], _p(hit),
                                              _ref(params), counts.shape[0], _p(counts),
                                              mat.ctypes.data_as(C.c_void_p), stride, threads)
    return rays, hit, int(live)


def trace_sample(cam, stride, offset, ftris, params, lm, bv=None, threads=0):
    R_full = cam.height * cam.width * cam.samples_per_pixel
    tex = np.zeros((lm.channels, lm.height, lm.width), dtype=np.float32)
    if bv is None:
        n = lib().oracle_trace_sample_f32(_ref(cam), R_full, stride, offset, _p(ftris), ftris.shape[0], _ref(params),
                                          _ref(lm), _p(tex), 0, None, None, 0, threads)
    else:
        counts, mat, st = bv
        mat = np.asfortranarray(mat, dtype=np.int64)
        n = lib().oracle_trace_sample_f32(_ref(cam), R_full, stride, offset, _p(ftris), ftris.shape[0], _ref(params),
                                          _ref(lm), _p(tex), counts.shape[0], _p(counts),
                                          mat.ctypes.data_as(C.c_void_p), st, threads)
    return tex, int(n)


def shade(rays, hit_idx, ftris, lm, tex, sky, h, precision="f32", threads=0):
    """continuum_shade! (src/ray_imagers.jl:440-503) of rays in pixel order on an h-row grid:
    returns (rgb image as (h*nd, w*nd) pixel rows x 3, row index fastest; the final-evolved rays
    (f32 only))."""
    hit_idx = np.ascontiguousarray(hit_idx, dtype=np.int32)
    R = rays.shape[0]
    nd = lm.n_delta
    tex = np.ascontiguousarray(tex, dtype=np.float32)
    if sky is None:
        sky_p, sw, sh = None, 0, 0
    else:
        sky = np.ascontiguousarray(sky, dtype=np.float32)
        sky_p, sw, sh = _p(sky), sky.shape[2], sky.shape[1]
    if precision == "f32":
        rays = np.array(rays, copy=True)
        rgb = np.zeros((R * nd * nd, 3), dtype=np.float32)
        lib().oracle_shade_f32(_p(rays), R, _p(hit_idx), _p(ftris), ftris.shape[0], _ref(lm), _p(tex), sky_p, sw, sh,
                               h, _p(rgb), threads)
        return rgb, rays
    rgb = np.zeros((R * nd * nd, 3), dtype=np.float64)
    lib().oracle_shade_f64(_p(np.ascontiguousarray(rays)), R, _p(hit_idx), _p(ftris), ftris.shape[0], _ref(lm),
                           _p(tex), sky_p, sw, sh, h, _p(rgb), threads)
    return rgb, None


def bilinear(plane, u, v):
    plane = np.ascontiguousarray(plane, dtype=np.float32)
    return lib().oracle_bilinear_f64(_p(plane), plane.shape[1], plane.shape[0], float(u), float(v))


def reorder(rays, hit_idx):
    """run_evolution!'s reorder (src/tracer.jl:270-278): each ray and hit index to slot dest - 1."""
    rays = np.array(rays, copy=True)
    hit_idx = np.array(hit_idx, dtype=np.int32, copy=True)
    lib().oracle_reorder(_p(rays), rays.shape[0], _p(hit_idx))
    return rays, hit_idx
