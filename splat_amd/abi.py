"""ctypes mirrors of the C ABI structs in include/splat_hip.h and numpy dtypes of the Julia
isbits layouts the ABI shares with the reference (src/geo.jl:12-113, src/tracer.jl:313).

Both the product wrapper (splat_amd.lib) and the test-side oracle binding use these, so the
field layout is written down exactly once on the Python side.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

# ---- status / enums (include/splat_hip.h) ------------------------------------------------
SPLAT_OK = 0
SPLAT_ERR_INVALID_ARG, SPLAT_ERR_WORKSPACE, SPLAT_ERR_HIP, SPLAT_ERR_UNSUPPORTED = -1, -2, -3, -4
SPLAT_ERR_CAPACITY, SPLAT_ERR_BUDGET = -5, -6
NOALLOC_MAX_RAYS = 1 << 21  # splat_next_hit_noalloc's scratch capacity (rays)
RAY_DIFFUSE, RAY_ACTIVE, RAY_INFINITY = 0, 1, 2
IOR_REFERENCE_LINEAR, IOR_CAUCHY = 0, 1
COMPAT_DY_AT_X, COMPAT_REFRACT_NORMAL_LAMBDA = 1, 2
COMPAT_REFERENCE = COMPAT_DY_AT_X | COMPAT_REFRACT_NORMAL_LAMBDA
HITTER_BF, HITTER_BV = 0, 1
MAX_LAMBDA, MAX_DELTA, MAX_CURVE = 64, 8, 24
WEIGHTS_IDENTITY, WEIGHTS_TABLE, WEIGHTS_CURVES = 0, 1, 2
ORDER_COLUMN_MAJOR, ORDER_TILED8 = 0, 1
LAMBDA_FIXED, LAMBDA_UNIFORM, LAMBDA_STRATIFIED = 0, 1, 2

# ---- numpy layouts ------------------------------------------------------------------------
# ADRay (src/geo.jl:60-76): 136 bytes
ADRAY = np.dtype(
    {
        "names": ["pos", "dpdl", "dpdx", "dpdy", "dir", "dddl", "dddx", "dddy",
                  "in_medium", "_pad0", "ignore_tri", "dest", "lam", "x", "y", "status", "_pad1"],
        "formats": [("<f4", 3)] * 8 + ["u1", ("u1", 7), "<i8", "<i8", "<f4", "<f4", "<f4", "u1", ("u1", 3)],
        "offsets": [0, 12, 24, 36, 48, 60, 72, 84, 96, 97, 104, 112, 120, 124, 128, 132, 133],
        "itemsize": 136,
    }
)
# Tuple{Int32,Tri} (src/tracer.jl:313): 52 bytes
NTRI = np.dtype({"names": ["idx", "v"], "formats": ["<i4", ("<f4", (4, 3))], "offsets": [0, 4], "itemsize": 52})
# FTri (src/geo.jl:18): 120 bytes
FTRI = np.dtype({"names": ["v"], "formats": [("<f4", (10, 3))], "offsets": [0], "itemsize": 120})
BVBOX = np.dtype({"names": ["min", "max"], "formats": [("<f4", 3), ("<f4", 3)], "offsets": [0, 12], "itemsize": 24})
SPHERE = np.dtype({"names": ["origin", "radius"], "formats": [("<f4", 3), "<f4"], "offsets": [0, 12], "itemsize": 16})
# FP64 ray record returned by the oracle (oracle_ray64 in oracle/oracle.cpp)
ADRAY64 = np.dtype(
    {
        "names": ["pos", "dpdl", "dpdx", "dpdy", "dir", "dddl", "dddx", "dddy",
                  "ignore_tri", "dest", "lam", "x", "y", "in_medium", "status"],
        "formats": [("<f8", 3)] * 8 + ["<i8", "<i8", "<f8", "<f8", "<f8", "<i4", "<i4"],
        "offsets": [0, 24, 48, 72, 96, 120, 144, 168, 192, 200, 208, 216, 224, 232, 236],
        "itemsize": 240,
    }
)
VEC_FIELDS = ["pos", "dpdl", "dpdx", "dpdy", "dir", "dddl", "dddx", "dddy"]


# ---- ctypes structs -----------------------------------------------------------------------
class Optics(C.Structure):
    _fields_ = [("ior_model", C.c_int32), ("cauchy_a", C.c_float), ("cauchy_b", C.c_float), ("compat", C.c_int32)]


class HitSet(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("n_tris", C.c_int32), ("n_bv", C.c_int32), ("max_members", C.c_int32),
        ("total_members", C.c_int64), ("d_recs", C.c_void_p), ("d_boxes", C.c_void_p),
        ("d_ranges", C.c_void_p), ("d_groups", C.c_void_p), ("group_size", C.c_int32), ("n_groups", C.c_int32),
        ("tune_waves", C.c_int32), ("tree_levels", C.c_int32), ("top_bvs", C.c_int32), ("top_recs", C.c_int32),
        ("d_stats", C.c_void_p),
    ]


class LightMapParams(C.Structure):
    _fields_ = [
        ("n_lambda", C.c_int32), ("n_delta", C.c_int32), ("channels", C.c_int32), ("weight_mode", C.c_int32),
        ("width", C.c_int32), ("height", C.c_int32),
        ("spectrum", C.c_float * MAX_LAMBDA), ("delta", C.c_float * MAX_DELTA),
        ("weights", C.c_float * (MAX_LAMBDA * 3)),
        ("curve_n", C.c_int32 * 3), ("curve_x", (C.c_float * MAX_CURVE) * 3), ("curve_y", (C.c_float * MAX_CURVE) * 3),
    ]


class Pinhole(C.Structure):
    _fields_ = [
        ("pos", C.c_float * 3), ("dir", C.c_float * 3), ("up", C.c_float * 3), ("right", C.c_float * 3),
        ("fov_half_sin", C.c_float), ("height", C.c_int32), ("width", C.c_int32),
        ("samples_per_pixel", C.c_int32), ("order", C.c_int32), ("lambda_mode", C.c_int32),
        ("lambda_min", C.c_float), ("lambda_max", C.c_float), ("jitter", C.c_float), ("_pad", C.c_int32),
        ("seed", C.c_uint64),
    ]


MAX_LIGHTS = 16


class RectLight(C.Structure):
    """RectLight (src/ray_generators.jl:4-11): four SVector{3,Float32} then two Int, 64 bytes."""
    _fields_ = [
        ("center", C.c_float * 3), ("dir", C.c_float * 3), ("dim1", C.c_float * 3), ("dim2", C.c_float * 3),
        ("res1", C.c_int64), ("res2", C.c_int64),
    ]


class LightRays(C.Structure):
    _fields_ = [
        ("upscale", C.c_int32), ("lambda_mode", C.c_int32), ("lambda_min", C.c_float), ("lambda_max", C.c_float),
        ("jitter", C.c_float), ("_pad", C.c_int32), ("seed", C.c_uint64),
    ]


def rect_light(center, dir, dim1, dim2, res1: int, res2: int) -> RectLight:
    L = RectLight()
    for k in range(3):
        L.center[k], L.dir[k], L.dim1[k], L.dim2[k] = center[k], dir[k], dim1[k], dim2[k]
    L.res1, L.res2 = res1, res2
    return L


def light_rays(upscale: int = 1, *, lambda_mode: int = LAMBDA_FIXED, lambda_min: float = 550.0,
               lambda_max: float = 550.0, jitter: float = 1.0, seed: int = 0) -> LightRays:
    """Defaults follow rays_from_lights: 550 nm, cell offsets U[0, 1) (src/ray_generators.jl:59, :75-76)."""
    return LightRays(upscale, lambda_mode, lambda_min, lambda_max, jitter, 0, seed)


class RunParams(C.Structure):
    _fields_ = [
        ("depth", C.c_int32), ("first_diffuse", C.c_int32), ("force_rand", C.c_float), ("_pad", C.c_int32),
        ("seed", C.c_uint64), ("optics", Optics), ("hit_events", C.c_void_p), ("n_hit_events", C.c_int32),
        ("_pad2", C.c_int32),
    ]


class FrameRequest(C.Structure):
    _fields_ = [
        ("cam", Pinhole), ("hs", HitSet), ("tris", C.c_void_p), ("n_tris", C.c_int32), ("depth", C.c_int32),
        ("first_diffuse", C.c_int32), ("force_rand", C.c_float), ("seed", C.c_uint64), ("optics", Optics),
        ("tex", C.c_void_p), ("rays_out", C.c_void_p), ("hit_out", C.c_void_p), ("tag", C.c_int64),
        ("lm", LightMapParams),
    ]


PARAM_STRUCTS = [Optics, HitSet, LightMapParams, Pinhole, RunParams, FrameRequest]
assert C.sizeof(Pinhole) == 96 and C.sizeof(RunParams) == 56 and C.sizeof(HitSet) == 88
assert C.sizeof(RectLight) == 64 and C.sizeof(LightRays) == 32


def optics(model: int = IOR_CAUCHY, a: float = 1.5, b: float = 0.004, compat: int = COMPAT_REFERENCE) -> Optics:
    return Optics(model, a, b, compat)


def run_params(depth: int, first_diffuse: int, force_rand: float = math.nan, seed: int = 0,
               opt: Optics | None = None) -> RunParams:
    return RunParams(depth, first_diffuse, force_rand, 0, seed, opt if opt is not None else optics(), None, 0, 0)


def upres(n: int) -> np.ndarray:
    """Sub-pixel offsets (src/alg_types.jl:108): (k/N) - 0.5 - 1/(2N) for k = 1..N, in Float32."""
    k = np.arange(1, n + 1, dtype=np.float32)
    return (k / np.float32(n) - np.float32(0.5)) - (np.float32(1.0) / (np.float32(2.0) * np.float32(n)))


# Piecewise-linear retina response curves of src/rgb_spectrum.jl:34-87 (the same table; it is
# the reference's data, kept as is), evaluated as its `piecewise` does (src/rgb_spectrum.jl:3-32:
# segment m for x_m <= lambda < x_{m+1}, the end segments extended). Normalised by the sum over
# integer lambda 400..700 (src/rgb_spectrum.jl:116-130).
_RETINA = {
    "red": [(399, 0), (400, 0.0647036), (420, 0.147492), (438, 0.195683), (460, 0.162951), (466, 0),
            (534, 0), (538, 0.243127), (571, 0.798209), (593, 0.980253), (602, 0.999977), (610, 0.98481),
            (660, 0.375838), (680, 0.210109), (700, 0.11005)],
    "green": [(399, 0), (462, 0), (467, 0.165356), (474, 0.280228), (502, 0.587139), (521, 0.829692),
              (534, 0.924245), (551, 0.910167), (572, 0.777366), (600, 0.372516), (609, 0.158767), (612, 0),
              (700, 0)],
    "blue": [(399, 0), (413, 0.23889), (418, 0.291473), (437, 0.437376), (456, 0.590128), (463, 0.72815),
             (476, 0.581812), (494, 0.452254), (515, 0.407615), (528, 0.299756), (536, 0), (608, 0),
             (611, 0.0407256), (614, 0.154091), (620, 0.220637), (635, 0.225269), (667, 0.116114),
             (700, 0.0424846)],
}


def _piecewise64(points, lam: float) -> float:
    n = len(points)
    m = 0
    while m < n - 2 and lam >= points[m + 1][0]:
        m += 1
    (x0, y0), (x1, y1) = points[m], points[m + 1]
    return (y1 - y0) / (x1 - x0) * (lam - x1) + y1


def retina_curves() -> list[tuple[np.ndarray, np.ndarray]]:
    out = []
    for name in ("red", "green", "blue"):
        pts = _RETINA[name]
        norm = sum(_piecewise64(pts, float(x)) for x in range(400, 701))
        xs = np.array([p[0] for p in pts], dtype=np.float32)
        ys = np.array([p[1] / norm for p in pts], dtype=np.float32)
        out.append((xs, ys))
    return out


def retina_factor(spectrum) -> np.ndarray:
    """retina_factor of _spectrum_datastructs (src/rgb_spectrum.jl:135-148): the normalised
    retina responses (red, green, blue) at each wavelength, float32, shape (n_lambda, 3)."""
    out = np.zeros((len(spectrum), 3), dtype=np.float32)
    for c, name in enumerate(("red", "green", "blue")):
        pts = _RETINA[name]
        norm = sum(_piecewise64(pts, float(x)) for x in range(400, 701))
        for k, lam in enumerate(spectrum):
            out[k, c] = np.float32(_piecewise64(pts, float(lam)) / norm)
    return out


def lightmap_params(width: int, height: int, *, n_delta: int = 1, spectrum=None, mode: str = "rgb") -> LightMapParams:
    """Build splat parameters.

    spectrum=None  -> own-lambda mode (each ray splats at its own wavelength).
    mode "rgb"     -> 3 channels weighted by the retina curves at the splat's lambda.
    mode "spectral"-> one plane per spectrum entry, weight 1 (the reference's tex[i,j,k]); the
                      weights table holds the retina factors, which backward shading reads.
    mode "mono"    -> one plane, weight 1 (own-lambda only).
    """
    p = LightMapParams()
    p.width, p.height = width, height
    if not 1 <= n_delta <= MAX_DELTA:
        raise ValueError("n_delta out of range")
    p.n_delta = n_delta
    d = upres(n_delta)
    for i, v in enumerate(d):
        p.delta[i] = float(v)
    if spectrum is None:
        p.n_lambda = 0
    else:
        spectrum = np.asarray(spectrum, dtype=np.float32)
        if not 1 <= spectrum.size <= MAX_LAMBDA:
            raise ValueError("spectrum size out of range")
        p.n_lambda = spectrum.size
        for i, v in enumerate(spectrum):
            p.spectrum[i] = float(v)
    if mode == "rgb":
        p.channels = 3
        p.weight_mode = WEIGHTS_CURVES
        for c, (xs, ys) in enumerate(retina_curves()):
            p.curve_n[c] = xs.size
            for i in range(xs.size):
                p.curve_x[c][i] = float(xs[i])
                p.curve_y[c][i] = float(ys[i])
    elif mode == "spectral":
        if spectrum is None:
            raise ValueError("spectral mode needs a spectrum grid")
        p.channels = p.n_lambda
        p.weight_mode = WEIGHTS_IDENTITY
        rf = retina_factor(spectrum)
        for k in range(spectrum.size):
            for c in range(3):
                p.weights[k * 3 + c] = float(rf[k, c])
    elif mode == "mono":
        if spectrum is not None:
            raise ValueError("mono mode is own-lambda only")
        p.channels = 1
        p.weight_mode = WEIGHTS_IDENTITY
    else:
        raise ValueError(f"unknown light-map mode {mode!r}")
    return p
