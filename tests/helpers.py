"""Shared fixtures-by-function for the tests: seeded scenes, rays and comparison helpers."""
from __future__ import annotations

import numpy as np

from splat_amd import abi, scene


def random_rays(n: int, seed: int, *, origin_scale=3.0, n_tris: int = 10, status=abi.RAY_ACTIVE,
                x_eq_y: bool = True) -> np.ndarray:
    """Rays with random origins/directions and random tangents (ignore_tri drawn from 1..n_tris)."""
    rng = np.random.default_rng(seed)
    r = np.zeros(n, dtype=abi.ADRAY)
    r["pos"] = (rng.uniform(-origin_scale, origin_scale, (n, 3))).astype(np.float32)
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r["dir"] = d.astype(np.float32)
    for f in ("dpdl", "dpdx", "dpdy", "dddl", "dddx", "dddy"):
        r[f] = (rng.standard_normal((n, 3)) * 1e-3).astype(np.float32)
    r["ignore_tri"] = rng.integers(1, n_tris + 1, n)
    r["dest"] = np.arange(1, n + 1)
    r["lam"] = rng.uniform(400, 700, n).astype(np.float32)
    if x_eq_y:
        r["x"] = 0.0
        r["y"] = 0.0
    else:
        r["x"] = rng.uniform(-2, 2, n).astype(np.float32)
        r["y"] = rng.uniform(-2, 2, n).astype(np.float32)
    r["status"] = status
    r["in_medium"] = 0
    return r


def rays_towards(points: np.ndarray, seed: int, spread: float = 2.0) -> np.ndarray:
    """Rays from random origins aimed at the given points (so a good fraction hits)."""
    rng = np.random.default_rng(seed)
    n = points.shape[0]
    r = random_rays(n, seed)
    o = points + rng.standard_normal((n, 3)) * spread
    d = points - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r["pos"] = o.astype(np.float32)
    r["dir"] = d.astype(np.float32)
    r["ignore_tri"] = 1
    return r


def sample_surface_points(ftris: np.ndarray, n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    idx = rng.integers(1, ftris.shape[0], n)
    a, b, c = ftris["v"][idx, 1], ftris["v"][idx, 2], ftris["v"][idx, 3]
    u = rng.uniform(0, 1, n)
    v = rng.uniform(0, 1, n)
    flip = u + v > 1
    u[flip], v[flip] = 1 - u[flip], 1 - v[flip]
    return a + (b - a) * u[:, None] + (c - a) * v[:, None]


def small_c2(subdiv: int = 2) -> scene.Scene:
    """The two-displaced-icospheres workload at a smaller subdivision (parity sizes)."""
    import numpy as _np

    rng = _np.random.default_rng(7)
    s1 = scene.sphere_mesh(subdiv, center=(-1.15, 0.0, 0.0), noise_sigma=0.02, rng=rng)
    s2 = scene.sphere_mesh(subdiv, center=(1.15, 0.1, 0.2), noise_sigma=0.02, rng=rng)
    return scene.assemble([s1, s2], [scene.ground_quad(-1.5, 4.0)], "small-c2")


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    nb = float(np.linalg.norm(b.astype(np.float64)))
    return float(np.linalg.norm(a.astype(np.float64) - b.astype(np.float64))) / max(nb, 1e-300)


def canonical_nans(rays: np.ndarray) -> np.ndarray:
    """A copy of ADRay records with every float NaN replaced by one quiet NaN. The reference's
    arithmetic can make NaN tangents (its d/dy taken at r.x, src/tracer.jl:138-141); x86 and the
    GPU make default NaNs of opposite sign, and only the sign and payload may differ."""
    out = np.array(rays, copy=True)
    for name in out.dtype.names:
        f = out[name]
        if f.dtype.kind == "f":
            f[np.isnan(f)] = np.float32(np.nan)
    return out


def same_rays(a: np.ndarray, b: np.ndarray) -> bool:
    """Bit-exact equality of ray records, any NaN equal to any NaN (canonical_nans)."""
    return a.shape == b.shape and canonical_nans(a).tobytes() == canonical_nans(b).tobytes()
