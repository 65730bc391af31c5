"""The host BV builders' defined choices, which the device build reproduces exactly:
equal centroids split by index, equal diagonals enter the large-triangle list by index,
-0 and +0 centroids are one key, every triangle lands in exactly one leaf."""
import numpy as np

from splat_amd import abi, bv


def _tris(verts):
    """Triangles (sentinel first) from an (n, 3, 3) vertex array."""
    out = np.zeros(len(verts) + 1, dtype=abi.FTRI)
    out["v"][1:, 1:4, :] = np.asarray(verts, dtype=np.float32)
    return out


def test_equal_centroids_split_by_index():
    tri = [[0, 0, 0], [1, 0, 0], [0, 1, 0]]
    ftris = _tris([tri] * 8)
    _, mem = bv.object_partition(ftris, 1)
    assert [m.tolist() for m in mem] == [[2, 3, 4, 5], [6, 7, 8, 9]]
    _, mem = bv.object_partition(ftris, 2)
    assert [m.tolist() for m in mem] == [[2, 3], [4, 5], [6, 7], [8, 9]]


def test_split_axis_then_index():
    # centroids spread along y only; x ties everywhere: the split follows y
    verts = [[[0, 3 * k, 0], [1, 3 * k, 0], [0, 3 * k + 1, 0]] for k in (3, 0, 2, 1)]
    _, mem = bv.object_partition(_tris(verts), 1)
    assert [m.tolist() for m in mem] == [[3, 5], [2, 4]]


def test_surface_area_axis_beats_widest_centroid_axis():
    # two thin strips (flat in z, long in x) at y = 0 and y = 1; centroids spread 2 along x but
    # only 1 along y. Splitting by x gives two 4 x 1 boxes (cost 2 x 8 x 2 = 32), by y two 6 x
    # 0.01 slivers (cost 0.48): the surface-area rule splits by y, the widest centroid axis
    # would split by x.
    def strip(x, y):
        return [[x - 2.0, y, 0.0], [x + 2.0, y, 0.0], [x, y + 0.01, 0.0]]
    verts = [strip(0.0, 0.0), strip(2.0, 0.0), strip(0.0, 1.0), strip(2.0, 1.0)]
    _, mem = bv.object_partition(_tris(verts), 1)
    assert [m.tolist() for m in mem] == [[2, 3], [4, 5]]
    assert bv.box_area(np.array([0.0, 0, 0]), np.array([4.0, 1, 0])) == 8.0


def test_large_triangle_cap_by_index():
    small = [[[k, 0, 0], [k + 0.1, 0, 0], [k, 0.1, 0]] for k in range(100)]
    huge = [[[0, 0, -1], [50, 0, -1], [0, 50, -1]]] * 40
    ftris = _tris(small + huge)
    _, mem, n_top = bv.tree_partition(ftris, 3, max_big=32)
    assert n_top == 1
    assert mem[0].tolist() == list(range(102, 102 + 32))  # the first 32 of the 40 equal ones
    rest = np.sort(np.concatenate(mem[1:]))
    assert rest.tolist() == list(range(2, 102)) + list(range(134, 142))


def test_signed_zero_centroids_are_one_key():
    tri = np.array([[[-0.0, 0, 0], [-0.0, 1, 0], [-0.0, 0, 1]]], dtype=np.float32)
    c = bv.centroids(_tris(tri))
    assert not np.signbit(c[1, 0])


def test_every_triangle_in_one_leaf():
    rng = np.random.default_rng(3)
    verts = rng.normal(size=(1000, 3, 3))
    ftris = _tris(verts)
    for depth in (1, 4, 7, 11):
        _, mem, n_top = bv.tree_partition(ftris, depth)
        allm = np.sort(np.concatenate(mem))
        assert allm.tolist() == list(range(2, 1002))
        sizes = [m.size for m in mem[n_top:]]
        assert max(sizes) - min(sizes) <= 1  # object-median halves
