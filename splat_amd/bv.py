"""One-level bounding-volume build on the host: bv_partition / spatial_partition / pack_bv_tris
(src/bv.jl:68-136).

Semantics kept: median split of triangle centroids on axis `depth % 3`, straddling triangles
placed in both halves, 2**depth leaves, boxes = member vertex bounds, memberships as 1-based
indices. Deliberate differences: the leaf order is deterministic (the reference shuffles it with
an unseeded RNG, src/bv.jl:89), and the degenerate sentinel triangle 1 is left out of every leaf
(the reference includes it, stretching boxes to the origin; it can never be hit). Neither changes
any hit result: the BV hitter returns the brute-force answer for any membership that covers the
triangles.
"""
from __future__ import annotations

import math

import numpy as np

from . import abi


def spatial_partition(verts: np.ndarray, idx: np.ndarray, depth: int) -> list[np.ndarray]:
    """verts: (N, 3, 3) triangle vertices; idx: array of 0-based slots in this node."""
    if depth < 0:
        raise ValueError("negative recursion")
    if depth == 0 or idx.size == 0:
        return [idx]
    dim = depth % 3
    coords = verts[idx, :, dim]
    mid = np.median(coords.mean(axis=1))
    left = idx[coords.min(axis=1) <= mid]
    right = idx[coords.max(axis=1) >= mid]
    return spatial_partition(verts, left, depth - 1) + spatial_partition(verts, right, depth - 1)


def bv_partition(ftris: np.ndarray, depth: int = 6):
    """Returns (bvs: BVBOX array, memberships: list of 1-based int64 index arrays)."""
    verts = ftris["v"][:, 1:4, :].astype(np.float64)
    slots = np.arange(1, ftris.shape[0], dtype=np.int64)  # skip the degenerate slot 0
    parts = [p for p in spatial_partition(verts, slots, depth) if p.size > 0]
    bvs = np.zeros(len(parts), dtype=abi.BVBOX)
    members = []
    for k, p in enumerate(parts):
        v = ftris["v"][p][:, 1:4, :].reshape(-1, 3)
        bvs[k]["min"] = v.min(axis=0)
        bvs[k]["max"] = v.max(axis=0)
        members.append(np.sort(p) + 1)
    return bvs, members


def tree_depth(n_tris: int, per_leaf: float = 8.0) -> int:
    """Leaf depth of the object-median build for about `per_leaf` triangles per leaf: the size
    at which the tree hitter's node tests and member tests balanced on the 10k- and 178k-triangle
    scenes (2^10 and 2^14 leaves)."""
    return max(1, min(20, int(round(math.log2(max(n_tris, 2) / per_leaf)))))


def centroids(ftris: np.ndarray) -> np.ndarray:
    """((a + b) + c) / 3 per triangle in float64, -0 folded into +0 (the device build's keys)."""
    v = ftris["v"][:, 1:4, :].astype(np.float64)
    return ((v[:, 0] + v[:, 1]) + v[:, 2]) / 3.0 + 0.0


SAH_MAX = 4096  # nodes of at most this many triangles choose their split axis by surface area


def box_area(lo: np.ndarray, hi: np.ndarray) -> float:
    d = np.maximum(hi - lo, 0.0)
    return 2.0 * ((d[0] * d[1] + d[1] * d[2]) + d[2] * d[0])


def object_partition(ftris: np.ndarray, depth: int = 8):
    """Object-median build: each triangle (except the sentinel) goes to exactly one leaf by its
    centroid, split at the median, 2**depth leaves in recursion order (siblings adjacent). A
    node's triangles are ordered by (centroid on the split axis, index) and the first half goes
    left, so equal centroids split by index. The split axis: for a node of 2..SAH_MAX triangles
    the one whose halves have the least surface-area cost (area of the half's triangle bounds x
    its count, summed over both halves; the lower axis on ties), above that the centroids' widest
    axis. Every leaf sits at the same depth (the hitter's heap-indexed tree). Boxes may overlap
    but nothing is duplicated, so a large triangle (a ground quad) enlarges only its own leaf
    instead of every leaf it straddles, which is what makes the reference's straddler-duplicating
    split (bv_partition) cull poorly. Any covering membership gives the same hit results; this
    one is the fast build for the workloads here, and the one splat_hitset_build_bv makes on the
    device. Returns (bvs, memberships) like bv_partition."""
    cent = centroids(ftris)
    v = ftris["v"][:, 1:4, :].astype(np.float64)
    tlo, thi = v.min(axis=1), v.max(axis=1)

    def rec(idx: np.ndarray, d: int) -> list[np.ndarray]:
        if d == 0 or idx.size <= 1:
            return [idx]
        h = idx.size // 2
        if idx.size > SAH_MAX:
            c = cent[idx]
            ax = int(np.argmax(c.max(axis=0) - c.min(axis=0)))
            o = idx[np.lexsort((idx, c[:, ax]))]
        else:
            o, best = None, None
            for ax in range(3):
                oa = idx[np.lexsort((idx, cent[idx, ax]))]
                L, R = oa[:h], oa[h:]
                cost = box_area(tlo[L].min(0), thi[L].max(0)) * L.size + box_area(tlo[R].min(0), thi[R].max(0)) * R.size
                if best is None or cost < best:
                    o, best = oa, cost
        return rec(o[:h], d - 1) + rec(o[h:], d - 1)

    parts = [p for p in rec(np.arange(1, ftris.shape[0], dtype=np.int64), depth) if p.size > 0]
    bvs = np.zeros(len(parts), dtype=abi.BVBOX)
    members = []
    for k, p in enumerate(parts):
        v = ftris["v"][p][:, 1:4, :].reshape(-1, 3)
        bvs[k]["min"] = v.min(axis=0)
        bvs[k]["max"] = v.max(axis=0)
        members.append(np.sort(p) + 1)
    return bvs, members


def tree_partition(ftris: np.ndarray, depth: int | None = None, big_factor: float = 16.0, max_big: int = 32):
    """The build for the tree hitter: triangles whose bounding-box diagonal exceeds big_factor x
    the median (a ground plane; at most max_big of the largest) go into one leading membership
    that every ray tests directly; the rest are split object-median into 2**depth leaves
    (depth = tree_depth of their count when None). A large triangle inside the tree would
    inflate its leaf's box and every box above it, so every ray would enter them. Returns
    (bvs, memberships, n_top) with n_top = 1 when such triangles exist, else 0."""
    v = ftris["v"][1:, 1:4, :].astype(np.float64)
    e = v.max(axis=1) - v.min(axis=1)
    diag = np.sqrt((e[:, 0] * e[:, 0] + e[:, 1] * e[:, 1]) + e[:, 2] * e[:, 2])
    med = float(np.median(diag)) if diag.size else 0.0
    big = np.nonzero(diag > big_factor * med)[0] if med > 0 else np.zeros(0, dtype=np.int64)
    if big.size > max_big:  # the largest first, equal diagonals by index
        big = big[np.lexsort((big, -diag[big]))[:max_big]]
    big = np.sort(big) + 2  # 1-based triangle indices (row 0 is the sentinel)
    rest = np.setdiff1d(np.arange(2, ftris.shape[0] + 1, dtype=np.int64), big)
    if depth is None:
        depth = tree_depth(rest.size)
    sub = ftris[np.concatenate([[0], rest - 1])]  # sentinel + the rest, re-indexed 2..
    bvs_r, mem_r = object_partition(sub, depth)
    mem_r = [rest[m - 2] for m in mem_r]
    if big.size == 0:
        return bvs_r, mem_r, 0
    vb = ftris["v"][big - 1][:, 1:4, :].reshape(-1, 3)
    top = np.zeros(1, dtype=abi.BVBOX)
    top[0]["min"] = vb.min(axis=0)
    top[0]["max"] = vb.max(axis=0)
    return np.concatenate([top, bvs_r]), [big] + mem_r, 1


def pack_bv_tris(memberships: list[np.ndarray], n_tris: int, max_overcount_factor: float = 4.0):
    """pack_bv_tris (src/bv.jl:126-136): dense n_bv x stride matrix of 1-based indices, Julia
    column-major (element (k, m) at k + m*n_bv), plus the per-BV counts. The stride grows to the
    largest membership if the reference's ceil(N/n_bv*4) would overflow."""
    n_bv = len(memberships)
    counts = np.array([m.size for m in memberships], dtype=np.int64)
    stride = max(int(np.ceil(n_tris / max(n_bv, 1) * max_overcount_factor)), int(counts.max(initial=1)))
    mat = np.ones((n_bv, stride), dtype=np.int64, order="F")
    for k, m in enumerate(memberships):
        mat[k, : m.size] = np.sort(m)
    return counts, mat, stride
