"""ctypes loader of the gfx950 library (splat_amd/libsplat_hip.so, C ABI include/splat_hip.h).

There is no fallback: if the library is missing it is built; if it cannot be loaded, or no GPU
is visible when a compute entry point is called, this raises. Device memory and streams come
from PyTorch (tensors' data_ptr(), torch.cuda.current_stream()); PyTorch is only plumbing here.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsplat_hip.so")
CSRC = os.path.join(HERE, "csrc")


class SplatError(RuntimeError):
    def __init__(self, code: int, where: str):
        msg = _lib.splat_strerror(code).decode() if _lib is not None else str(code)
        super().__init__(f"{where}: {msg} ({code})")
        self.code = code


_lib = None


def build(force: bool = False) -> str:
    """Compile the HIP sources for gfx950 (hipcc cross-compiles without a GPU)."""
    if force or not os.path.exists(LIB_PATH):
        args = ["make", "-s", "-C", CSRC]
        if force:
            args.append("-B")
        subprocess.run(args, check=True)
    return LIB_PATH


P, i32, i64, u64, f32, sz = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_float, C.c_size_t

SIGNATURES = {
    "splat_strerror": (C.c_char_p, [i32]),
    "splat_abi_version": (i32, []),
    "splat_layout_sizes": (i32, [P]),
    "splat_device_count": (i32, []),
    "splat_hitset_bf_bytes": (sz, [i32]),
    "splat_hitset_prepare_bf": (i32, [P, i32, P, sz, P, P]),
    "splat_hitset_prepare_bf_ftri": (i32, [P, i32, P, sz, P, P]),
    "splat_hitset_bv_bytes": (sz, [i32, i64]),
    "splat_hitset_prepare_bv": (i32, [P, i32, i32, P, P, i64, i32, i32, P, sz, P, P]),
    "splat_hitset_build_bv_bytes": (sz, [i32, i32]),
    "splat_hitset_build_bv": (i32, [P, i32, i32, f32, i32, P, sz, P, P]),
    "splat_next_hit_ws_bytes": (sz, [i64]),
    "splat_next_hit": (i32, [P, i64, P, P, P, P, sz, P]),
    "splat_next_hit_bf": (i32, [P, i64, P, i32, P, P, sz, P]),
    "splat_next_hit_noalloc": (i32, [P, i64, P, P, P, P]),
    "splat_evolve": (i32, [P, i64, P, P, P, i32, P, f32, u64, i32, i32, P, P]),
    "splat_partition_ws_bytes": (sz, [i64]),
    "splat_partition_active": (i32, [P, P, i64, P, P, P, sz, P]),
    "splat_light_map": (i32, [P, i64, P, P, P, i32, P, P, P]),
    "splat_raygen_pinhole": (i32, [P, P, i64, P]),
    "splat_lights_ray_count": (i64, [P, i32, i32]),
    "splat_raygen_lights": (i32, [P, i32, P, P, i64, P]),
    "splat_fill_f32": (i32, [P, i64, f32, P]),
    "splat_run_ws_bytes": (sz, [i64]),
    "splat_run_evolution": (i32, [P, i64, P, P, i32, P, P, P, sz, P, P]),
    "splat_trace_frame": (i32, [P, P, i64, P, P, i32, P, P, P, P, P, sz, P]),
    "splat_trace_lights": (i32, [P, i32, P, P, i64, P, P, i32, P, P, P, P, P, sz, P]),
    "splat_param_sizes": (i32, [P]),
    "splat_reorder_ws_bytes": (sz, [i64]),
    "splat_reorder_by_dest": (i32, [P, i64, P, P, sz, P]),
    "splat_shade": (i32, [P, i64, P, P, i32, P, P, P, i32, i32, i64, P, P]),
    "splat_engine_frame_bytes": (sz, [P, i32]),
    "splat_engine_create": (i32, [i64, i32, P]),
    "splat_engine_destroy": (i32, [P]),
    "splat_engine_admit": (i32, [P, P, P]),
    "splat_engine_step": (i32, [P, P]),
    "splat_engine_poll": (i32, [P, i32, P, i32, P]),
    "splat_engine_in_flight": (i32, [P]),
    "splat_engine_runnable": (i32, [P]),
    "splat_engine_sync_steps": (i32, [P, i32]),
    "splat_engine_arena_free": (i64, [P]),
    "splat_engine_use_graph": (i32, [P, i32]),
    "splat_engine_use_queue": (i32, [P, i32]),
}


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        if L.splat_abi_version() != 1:
            raise RuntimeError("libsplat_hip.so ABI version mismatch")
        _lib = L
    return _lib


def check(code: int, where: str) -> None:
    if code != abi.SPLAT_OK:
        raise SplatError(code, where)


def ref(s) -> C.c_void_p:
    """Address of a ctypes struct (the caller keeps `s` alive across the call)."""
    return C.c_void_p(C.addressof(s))


def ptr(t) -> C.c_void_p | None:
    """Device pointer of a torch tensor (None passes NULL)."""
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


def stream_handle(stream=None) -> C.c_void_p:
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def require_gpu() -> None:
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("splat_amd: no GPU visible; the HIP path has no CPU fallback")
