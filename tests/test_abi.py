"""The C-ABI library loads and exports every entry point include/splat_hip.h declares, and the
Python mirrors of its structs have the C sizes (no GPU compute needed)."""
import ctypes as C
import os
import re

import numpy as np

from splat_amd import abi
from splat_amd import lib as L

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "splat_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(splat_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = L.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_wrapper_covers_the_header():
    assert set(declared_functions()) <= set(L.SIGNATURES)


def test_layout_sizes_match_julia_isbits():
    out = (C.c_int64 * 6)()
    assert L.lib().splat_layout_sizes(C.cast(out, C.c_void_p)) == 0
    assert list(out) == [136, 48, 52, 120, 24, 16]
    assert [abi.ADRAY.itemsize, abi.NTRI.itemsize, abi.FTRI.itemsize, abi.BVBOX.itemsize, abi.SPHERE.itemsize] == \
        [136, 52, 120, 24, 16]


def test_param_struct_sizes_match_ctypes():
    out = (C.c_int64 * 6)()
    assert L.lib().splat_param_sizes(C.cast(out, C.c_void_p)) == 0
    assert list(out) == [C.sizeof(t) for t in abi.PARAM_STRUCTS]


def test_strerror_and_version():
    lib = L.lib()
    assert lib.splat_abi_version() == 1
    assert lib.splat_strerror(0) == b"ok"
    assert lib.splat_strerror(-2) == b"workspace too small"


def test_workspace_sizes_are_monotone():
    lib = L.lib()
    assert lib.splat_run_ws_bytes(1000) >= 1000 * (2 * 136 + 4)
    assert lib.splat_run_ws_bytes(2000) > lib.splat_run_ws_bytes(1000)
    assert lib.splat_partition_ws_bytes(0) > 0
    assert lib.splat_hitset_bf_bytes(10) >= 640
    assert lib.splat_hitset_bv_bytes(4, 100) >= 100 * 64 + 4 * 32


def test_argument_errors_without_a_gpu():
    """Invalid arguments are rejected before anything is launched."""
    lib = L.lib()
    assert lib.splat_next_hit(None, -1, None, None, None, None, 0, None) == -1
    assert lib.splat_partition_active(None, None, -5, None, None, None, 0, None) == -1
    hs = abi.HitSet()
    assert lib.splat_next_hit(None, 0, None, L.ref(hs), None, None, 0, None) == -1  # no records
    assert lib.splat_next_hit_ws_bytes(1000) >= 8 * 1000 and lib.splat_next_hit_ws_bytes(-1) == 0
    assert lib.splat_next_hit_noalloc(None, -1, None, None, None, None) == abi.SPLAT_ERR_INVALID_ARG
    assert lib.splat_next_hit_noalloc(C.c_void_p(8), abi.NOALLOC_MAX_RAYS + 1, None, L.ref(hs), C.c_void_p(8),
                                      None) == abi.SPLAT_ERR_CAPACITY
    p = abi.run_params(65, 2)
    assert lib.splat_run_evolution(C.c_void_p(8), 10, L.ref(hs), C.c_void_p(8), 3, C.c_void_p(8), L.ref(p),
                                   C.c_void_p(8), 1 << 30, None, None) == -1
    lm = abi.lightmap_params(8, 8)
    lm.n_delta = 99
    assert lib.splat_light_map(C.c_void_p(8), 1, None, C.c_void_p(8), C.c_void_p(8), 3, L.ref(lm), C.c_void_p(8),
                               None) == -4
    assert lib.splat_engine_use_queue(None, 1) == abi.SPLAT_ERR_INVALID_ARG
    assert lib.splat_engine_use_graph(None, 1) == abi.SPLAT_ERR_INVALID_ARG


def test_light_ray_counts_without_a_gpu():
    """rays_from_lights sizing (src/ray_generators.jl:64-113): (res1/up) x (res2/up) per light; with
    several lights each must fill its res1*res2/up^2 slice; bad counts are rejected."""
    lib = L.lib()
    A = abi.rect_light((0, 0, 8), (0, 0, -1), (1, 0, 0), (0, 1, 0), 64, 48)
    B = abi.rect_light((0, 0, 8), (0, 0, -1), (1, 0, 0), (0, 1, 0), 37, 22)
    two = (abi.RectLight * 2)(A, B)
    assert lib.splat_lights_ray_count(two, 1, 1) == 64 * 48
    assert lib.splat_lights_ray_count(two, 1, 4) == 16 * 12
    assert lib.splat_lights_ray_count(two, 2, 1) == 64 * 48 + 37 * 22
    assert lib.splat_lights_ray_count(C.byref(two, C.sizeof(abi.RectLight)), 1, 2) == 18 * 11  # one light: no slice
    assert lib.splat_lights_ray_count(two, 2, 2) == -1  # 37 * 22 / 4 = 203 != 18 * 11
    assert lib.splat_lights_ray_count(two, 1, 0) == -1
    assert lib.splat_lights_ray_count(None, 0, 1) == 0
    assert lib.splat_lights_ray_count(two, abi.MAX_LIGHTS + 1, 1) == -1
    p = abi.light_rays(1, lambda_mode=abi.LAMBDA_STRATIFIED)
    assert lib.splat_raygen_lights(two, 1, L.ref(p), C.c_void_p(8), 64 * 48, None) == -1  # no strata for lights
    p = abi.light_rays(1)
    assert lib.splat_raygen_lights(two, 1, L.ref(p), C.c_void_p(8), 100, None) == -1  # R must equal the count


def test_retina_curves_normalised():
    for xs, ys in abi.retina_curves():
        assert xs[0] == 399 and xs[-1] == 700 and np.all(np.diff(xs) > 0)
        assert ys.max() > 0
    d = abi.upres(4)
    assert np.allclose(d, [-0.375, -0.125, 0.125, 0.375])
    assert abi.upres(1)[0] == 0.0


def _fake_bv_hitset(n_bv: int, **kw) -> abi.HitSet:
    hs = abi.HitSet()
    hs.kind, hs.n_tris, hs.n_bv, hs.total_members, hs.max_members = abi.HITTER_BV, 40, n_bv, 60, 6
    hs.d_recs = hs.d_boxes = hs.d_ranges = hs.d_groups = 256  # never dereferenced: refused or sized out first
    for k, v in kw.items():
        setattr(hs, k, v)
    return hs


def test_hit_set_structure_must_cover_its_bvs_without_a_gpu():
    """A BV hit set whose tree or group level cannot reach every BV below its top list (the
    kernels visit only those), or whose top list claims more records than the set holds, is
    refused (SPLAT_ERR_INVALID_ARG); a covering one passes the check and stops at the workspace
    size (SPLAT_ERR_WORKSPACE), before anything is launched."""
    lib = L.lib()
    p = abi.run_params(2, 3)

    def code(hs):
        return lib.splat_run_evolution(C.c_void_p(8), 10, L.ref(hs), C.c_void_p(8), 40, C.c_void_p(8), L.ref(p),
                                       C.c_void_p(8), 0, None, None)

    # tree over 2^(levels-1) leaves: 8 BVs below a top list of 2 need 8 leaves (levels 4)
    assert code(_fake_bv_hitset(10, tree_levels=4, top_bvs=2, top_recs=5)) == abi.SPLAT_ERR_WORKSPACE
    assert code(_fake_bv_hitset(10, tree_levels=3, top_bvs=2, top_recs=5)) == abi.SPLAT_ERR_INVALID_ARG
    assert code(_fake_bv_hitset(10, tree_levels=4, top_bvs=0)) == abi.SPLAT_ERR_INVALID_ARG  # 10 > 8 leaves
    # groups: n_groups * group_size must cover the BVs below the top list
    assert code(_fake_bv_hitset(10, group_size=4, n_groups=2, top_bvs=2, top_recs=5)) == abi.SPLAT_ERR_WORKSPACE
    assert code(_fake_bv_hitset(10, group_size=4, n_groups=1, top_bvs=2, top_recs=5)) == abi.SPLAT_ERR_INVALID_ARG
    # the top list's records lie inside the set's records
    assert code(_fake_bv_hitset(10, tree_levels=4, top_bvs=2, top_recs=61)) == abi.SPLAT_ERR_INVALID_ARG
    assert code(_fake_bv_hitset(10, tree_levels=4, top_bvs=2, top_recs=60)) == abi.SPLAT_ERR_WORKSPACE
    # flat (no structure over the BVs): nothing to cover
    assert code(_fake_bv_hitset(12)) == abi.SPLAT_ERR_WORKSPACE


def test_repeated_curve_breakpoints_are_refused_without_a_gpu():
    """A repeated first or last breakpoint would give a zero-width end segment, whose slope divides
    by zero for every wavelength outside the table: refused (SPLAT_ERR_INVALID_ARG)."""
    lib = L.lib()
    for c, j in ((0, 0), (1, -1), (2, 4)):
        lm = abi.lightmap_params(8, 8)
        n = lm.curve_n[c]
        j = j % (n - 1)
        lm.curve_x[c][j + 1] = lm.curve_x[c][j]
        assert lib.splat_light_map(C.c_void_p(8), 1, None, C.c_void_p(8), C.c_void_p(8), 3, L.ref(lm), C.c_void_p(8),
                                   None) == abi.SPLAT_ERR_INVALID_ARG
