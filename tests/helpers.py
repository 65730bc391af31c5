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


def fp64_check(got32: np.ndarray, rays: np.ndarray, hit: np.ndarray, sc, force_rand: float, *, eps: float = 1e-7,
               draws: int = 6, seed: int = 1):
    """Per-ray comparison of one FP32 bounce (`got32`, ADRay records) with the FP64 oracle on the
    same inputs, and each ray's conditioning.

    kappa: the largest relative change of any output vector per unit relative change of the inputs,
    over `draws` random input perturbations of relative size <= eps (vector fields, lambda, x, y)
    through the FP64 oracle; a ray is well conditioned when kappa <= 30 (FP32 rounding, ~6e-8 per
    operation, then moves it well under 1e-5). Rays whose discrete outcome (status, in_medium)
    differs between FP32, FP64 or any perturbation are not compared. Vector norms are floored at
    1e-3 (tangents near zero carry no relative precision).

    Returns (err, kappa, comparable) arrays over the rays."""
    from oracle import oracle as O

    r64 = O.to64(rays)
    w0 = O.evolve64(r64, hit, sc.ftris, force_rand=force_rand, first_diffuse=sc.first_diffuse)
    same = (got32["status"] == w0["status"]) & (got32["in_medium"] == w0["in_medium"]) & \
           (got32["status"] == abi.RAY_ACTIVE)
    err = np.zeros(rays.shape[0])
    for f in abi.VEC_FIELDS:
        nb = np.maximum(np.linalg.norm(w0[f], axis=1), 1e-3)
        err = np.maximum(err, np.linalg.norm(got32[f].astype(np.float64) - w0[f], axis=1) / nb)
    rng = np.random.default_rng(seed)
    kappa = np.zeros(rays.shape[0])
    for _ in range(draws):
        pert = r64.copy()
        for f in abi.VEC_FIELDS + ["lam", "x", "y"]:
            pert[f] = r64[f] * (1 + eps * rng.uniform(-1.0, 1.0, r64[f].shape))
        wp = O.evolve64(pert, hit, sc.ftris, force_rand=force_rand, first_diffuse=sc.first_diffuse)
        same &= (wp["status"] == w0["status"]) & (wp["in_medium"] == w0["in_medium"])
        for f in abi.VEC_FIELDS:
            nb = np.maximum(np.linalg.norm(w0[f], axis=1), 1e-3)
            kappa = np.maximum(kappa, np.linalg.norm(wp[f] - w0[f], axis=1) / nb / eps)
    return err, kappa, same


def unambiguous_rays(rays: np.ndarray, hit: np.ndarray, ftris: np.ndarray, lm, *, texel_tol: float = 2e-4,
                     edge_tol: float = 1e-6) -> np.ndarray:
    """Mask of the rays whose light-map bins do not depend on the precision: every splat sample at
    least `texel_tol` texels from a bin edge and the final hit point at least `edge_tol` (barycentric)
    inside its triangle, by the FP64 oracle. Binning (ceil of u*w) is discontinuous, so a sample
    within rounding of an edge lands in a neighbouring texel in FP32 and FP64 alike at random; the
    FP64 comparison of a light map is over the other rays."""
    from oracle import oracle as O

    tm, em = O.splat_margins(rays, hit, ftris, lm)
    return (tm > texel_tol) & (np.abs(em) > edge_tol)
