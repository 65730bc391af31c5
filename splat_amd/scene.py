"""Procedural, seeded scenes for the light-map workloads.

The reference builds its scenes from OBJ meshes plus a prepended degenerate triangle
(src/scenes/sphere.jl:22-39, mesh_to_FTri src/geo.jl:331-341). These generators produce the same
FTri layout -- [face normal cross(v1-v2, v2-v3), a, b, c, n_a, n_b, n_c, uv_a, uv_b, uv_c] -- with
triangle 1 the all-zero sentinel, glass triangles next and diffuse triangles last, so
`first_diffuse` is the 1-based index of the first diffuse triangle (src/tracer.jl:186).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import abi


@dataclass
class Scene:
    ftris: np.ndarray  # FTRI records, ftris[0] is the degenerate sentinel
    first_diffuse: int  # 1-based
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def n_tris(self) -> int:
        return int(self.ftris.shape[0])

    def ntris(self) -> np.ndarray:
        """n_tris = tuple.(1:N, tri_from_ftri.(tris)) (src/tracer.jl:313)."""
        out = np.zeros(self.n_tris, dtype=abi.NTRI)
        out["idx"] = np.arange(1, self.n_tris + 1, dtype=np.int32)
        out["v"] = self.ftris["v"][:, :4, :]
        return out


def icosphere(subdiv: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit icosphere: 20 * 4**subdiv faces, float64 vertices on the unit sphere."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
             (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    V = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    F = faces
    for _ in range(subdiv):
        cache: dict[tuple[int, int], int] = {}

        def mid(a: int, b: int) -> int:
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = V[a] + V[b]
                V.append(m / np.linalg.norm(m))
                cache[key] = len(V) - 1
            return cache[key]

        nf = []
        for a, b, c in F:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        F = nf
    return np.array(V), np.array(F, dtype=np.int64)


def _spherical_uv(p: np.ndarray) -> np.ndarray:
    u = np.arctan2(p[:, 1], p[:, 0]) / (2 * math.pi) + 0.5
    v = np.arccos(np.clip(p[:, 2], -1.0, 1.0)) / math.pi
    return np.stack([u, v, np.zeros_like(u)], axis=1)


def mesh_to_ftri(pos: np.ndarray, faces: np.ndarray, normals: np.ndarray, uvs: np.ndarray) -> np.ndarray:
    """mesh_to_FTri (src/geo.jl:331-341) over indexed arrays; positions in Float32 first, the face
    normal is cross(v1 - v2, v2 - v3) of the Float32 vertices."""
    p32 = pos.astype(np.float32)
    v1, v2, v3 = p32[faces[:, 0]], p32[faces[:, 1]], p32[faces[:, 2]]
    n = np.cross(v1 - v2, v2 - v3).astype(np.float32)
    out = np.zeros(faces.shape[0], dtype=abi.FTRI)
    out["v"][:, 0] = n
    out["v"][:, 1], out["v"][:, 2], out["v"][:, 3] = v1, v2, v3
    n32 = normals.astype(np.float32)
    out["v"][:, 4], out["v"][:, 5], out["v"][:, 6] = n32[faces[:, 0]], n32[faces[:, 1]], n32[faces[:, 2]]
    uv32 = uvs.astype(np.float32)
    out["v"][:, 7], out["v"][:, 8], out["v"][:, 9] = uv32[faces[:, 0]], uv32[faces[:, 1]], uv32[faces[:, 2]]
    return out


def sphere_mesh(subdiv: int, center=(0.0, 0.0, 0.0), radius: float = 1.0, noise_sigma: float = 0.0,
                rng: np.random.Generator | None = None) -> np.ndarray:
    """Icosphere with optional radial vertex noise r * (1 + sigma * N(0,1)); vertex normals are the
    (radial) unit directions, uv spherical."""
    V, F = icosphere(subdiv)
    r = np.full(V.shape[0], radius)
    if noise_sigma > 0:
        rng = rng if rng is not None else np.random.default_rng(0)
        r = radius * (1.0 + noise_sigma * rng.standard_normal(V.shape[0]))
    pos = V * r[:, None] + np.asarray(center, dtype=np.float64)
    return mesh_to_ftri(pos, F, V, _spherical_uv(V))


def ground_quad(z: float, half: float, cx: float = 0.0, cy: float = 0.0) -> np.ndarray:
    """Two-triangle diffuse square (cf. square(), src/procedural_assets.jl:76-83) spanning
    [cx-half, cx+half] x [cy-half, cy+half] at height z, uv = the square's own [0,1]^2."""
    corners = np.array([[cx - half, cy - half, z], [cx + half, cy - half, z], [cx - half, cy + half, z],
                        [cx + half, cy + half, z]])
    uv = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], dtype=np.float64)
    nrm = np.tile([0.0, 0.0, 1.0], (4, 1))
    faces = np.array([[0, 1, 2], [1, 3, 2]])
    return mesh_to_ftri(corners, faces, nrm, uv)


def menger_cubes(level: int, center=(0.0, 0.0, 0.0), size: float = 2.0) -> np.ndarray:
    """Centres of the 20**level sub-cubes of a Menger sponge of edge `size`."""
    cubes = np.array([center], dtype=np.float64)
    s = size
    for _ in range(level):
        s /= 3.0
        offs = np.array([(i, j, k) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)
                         if (i == 0) + (j == 0) + (k == 0) < 2], dtype=np.float64)
        cubes = (cubes[:, None, :] + offs[None, :, :] * s).reshape(-1, 3)
    return cubes, s


def cube_mesh(centers: np.ndarray, edge: float) -> np.ndarray:
    """12 triangles per axis-aligned cube (internal faces kept, as a triangle soup)."""
    h = edge / 2.0
    corner = np.array([[x, y, z] for x in (-h, h) for y in (-h, h) for z in (-h, h)])
    quads = [(0, 1, 3, 2, (-1, 0, 0)), (4, 6, 7, 5, (1, 0, 0)), (0, 4, 5, 1, (0, -1, 0)),
             (2, 3, 7, 6, (0, 1, 0)), (0, 2, 6, 4, (0, 0, -1)), (1, 5, 7, 3, (0, 0, 1))]
    tri_list, nrm_list = [], []
    for a, b, c, d, n in quads:
        tri_list += [(a, b, c), (a, c, d)]
        nrm_list += [n, n]
    tri = np.array(tri_list)
    nrm = np.array(nrm_list, dtype=np.float64)
    nc = centers.shape[0]
    pos = (centers[:, None, :] + corner[None, :, :]).reshape(-1, 3)
    faces = (tri[None, :, :] + 8 * np.arange(nc)[:, None, None]).reshape(-1, 3)
    normals_v = np.repeat(np.repeat(nrm[None, :, :], nc, axis=0).reshape(-1, 3), 3, axis=0)
    # per-face flat normals and planar uv: unroll to per-corner vertices
    flat_pos = pos[faces.reshape(-1)]
    flat_faces = np.arange(flat_pos.shape[0]).reshape(-1, 3)
    uv = np.stack([flat_pos[:, 0] - flat_pos[:, 0].min(), flat_pos[:, 1] - flat_pos[:, 1].min(),
                   np.zeros(flat_pos.shape[0])], axis=1)
    uv /= max(1e-9, float(uv.max()))
    return mesh_to_ftri(flat_pos, flat_faces, normals_v, uv)


def rotate_ftris(ftris: np.ndarray, axis, angle: float) -> np.ndarray:
    """rotate(t::FTri, R) (src/geo.jl:139-150) with R the rotation by `angle` about `axis`
    (Rodrigues): the face normal, vertices and vertex normals turn, the uvs stay."""
    k = np.asarray(axis, dtype=np.float64)
    k = k / np.linalg.norm(k)
    K = np.array([[0.0, -k[2], k[1]], [k[2], 0.0, -k[0]], [-k[1], k[0], 0.0]])
    R = np.eye(3) + math.sin(angle) * K + (1.0 - math.cos(angle)) * (K @ K)
    out = np.array(ftris, copy=True)
    v = out["v"].astype(np.float64)
    v[:, :7, :] = v[:, :7, :] @ R.T
    out["v"] = v.astype(np.float32)
    return out


def assemble(glass: list[np.ndarray], diffuse: list[np.ndarray], name: str, **meta) -> Scene:
    """[degenerate] + glass + diffuse, first_diffuse = 1 + 1 + n_glass (cf. src/scenes/sphere.jl:36-37)."""
    zero = np.zeros(1, dtype=abi.FTRI)
    g = np.concatenate(glass) if glass else np.zeros(0, dtype=abi.FTRI)
    d = np.concatenate(diffuse) if diffuse else np.zeros(0, dtype=abi.FTRI)
    ftris = np.concatenate([zero, g, d])
    return Scene(ftris=ftris, first_diffuse=2 + g.shape[0], name=name, meta=meta)


# ---- camera (get_camera, src/geo.jl:183-192) at the point light -----------------------------

def pinhole(pos, look_at, up, fov_deg: float, height: int, width: int, *, spp: int = 1,
            lambda_mode: int = abi.LAMBDA_UNIFORM, lambda_min: float = 400.0, lambda_max: float = 700.0,
            jitter: float = 0.0, seed: int = 0, order: int = abi.ORDER_TILED8) -> abi.Pinhole:
    pos = np.asarray(pos, dtype=np.float64)
    d = np.asarray(look_at, dtype=np.float64) - pos
    d /= np.linalg.norm(d)
    right = np.cross(d, np.asarray(up, dtype=np.float64))
    right /= np.linalg.norm(right)
    upv = np.cross(right, d)
    upv /= np.linalg.norm(upv)
    cam = abi.Pinhole()
    for i in range(3):
        cam.pos[i], cam.dir[i], cam.up[i], cam.right[i] = float(pos[i]), float(d[i]), float(upv[i]), float(right[i])
    cam.fov_half_sin = float(np.float32(math.sin(math.radians(fov_deg) / 2.0)))
    cam.height, cam.width, cam.samples_per_pixel = height, width, spp
    cam.order, cam.lambda_mode = order, lambda_mode
    cam.lambda_min, cam.lambda_max, cam.jitter = lambda_min, lambda_max, jitter
    cam.seed = seed
    return cam


# ---- the light-map workloads ----------------------------------------------------------------

def scene_c1() -> Scene:
    """Icosphere subdiv 2 (320 glass triangles, r = 1) above a 2-triangle diffuse ground quad."""
    return assemble([sphere_mesh(2)], [ground_quad(-1.5, 3.0)], "icosphere2+quad")


def camera_c1(seed: int = 1, height: int = 256, width: int = 256, **kw) -> abi.Pinhole:
    return pinhole((0.3, 0.2, 5.0), (0.0, 0.0, -1.5), (0.0, 1.0, 0.0), 60.0, height, width, seed=seed, **kw)


def scene_c2(seed: int = 2) -> Scene:
    """Two displaced icospheres (subdiv 4 each, 10,240 glass triangles, radial noise sigma 0.02)
    over a diffuse plane."""
    rng = np.random.default_rng(seed)
    s1 = sphere_mesh(4, center=(-1.15, 0.0, 0.0), noise_sigma=0.02, rng=rng)
    s2 = sphere_mesh(4, center=(1.15, 0.1, 0.2), noise_sigma=0.02, rng=rng)
    return assemble([s1, s2], [ground_quad(-1.5, 4.0)], "2x displaced icosphere4+plane", seed=seed)


def camera_c2(seed: int = 2, height: int = 1024, width: int = 1024, **kw) -> abi.Pinhole:
    return pinhole((0.2, 0.3, 6.0), (0.0, 0.0, -1.5), (0.0, 1.0, 0.0), 90.0, height, width, seed=seed, **kw)


def scene_c3() -> Scene:
    """Level-3 Menger sponge (8,000 cubes, 96,000 triangles, diffuse) plus a subdiv-6 glass
    icosphere (81,920 triangles) over a diffuse plane."""
    centers, edge = menger_cubes(3, center=(0.0, 0.0, -0.2), size=2.4)
    sponge = cube_mesh(centers, edge)
    ball = sphere_mesh(6, center=(0.0, 0.0, 2.2), radius=0.9)
    return assemble([ball], [sponge, ground_quad(-1.5, 5.0)], "menger3+icosphere6+plane")


def camera_c3(seed: int = 3, height: int = 2048, width: int = 2048, spp: int = 8, **kw) -> abi.Pinhole:
    return pinhole((0.3, 0.1, 7.0), (0.0, 0.0, -1.5), (0.0, 1.0, 0.0), 90.0, height, width, spp=spp,
                   lambda_mode=abi.LAMBDA_STRATIFIED, seed=seed, **kw)


def view_camera(height: int, width: int, *, pos=(3.5, -4.0, 3.0), look_at=(0.0, 0.0, -0.5), fov_deg: float = 60.0,
                seed: int = 11) -> abi.Pinhole:
    """The viewer's camera of the backward pass (camera_ray at lambda 550 nm, src/tracer.jl:347-352,
    src/ray_generators.jl:52-59): pixel order, one ray per pixel, seeded 0.025-pixel jitter."""
    return pinhole(pos, look_at, (0.0, 0.0, 1.0), fov_deg, height, width, lambda_mode=abi.LAMBDA_FIXED,
                   lambda_min=550.0, lambda_max=550.0, jitter=0.025, seed=seed, order=abi.ORDER_COLUMN_MAJOR)


def skydome(width: int = 256, height: int = 128) -> np.ndarray:
    """A procedural sky (3, height, width) float32: blue zenith to pale horizon along u (the polar
    angle), a warm band around one azimuth along v."""
    u = (np.arange(width, dtype=np.float64) + 0.5) / width
    v = (np.arange(height, dtype=np.float64) + 0.5) / height
    uu, vv = np.meshgrid(u, v)
    band = np.exp(-((vv - 0.3) / 0.08) ** 2)
    sky = np.stack([0.25 + 0.6 * uu + 0.3 * band, 0.45 + 0.4 * uu + 0.15 * band, 0.9 - 0.2 * uu])
    return np.clip(sky, 0.0, 1.0).astype(np.float32)
