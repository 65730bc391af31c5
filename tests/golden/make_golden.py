"""Writes tests/golden/golden_c1.npz: small fixed vectors of the light-map path on the 320-triangle
icosphere workload (64 x 64 pinhole rays), produced by the oracle (oracle/oracle.cpp).

They pin the oracle against drift between rounds (tests/test_golden.py) and give the GPU tests
a fixed target independent of the live oracle. Regenerate only on a deliberate spec change:
    python tests/golden/make_golden.py
"""
import hashlib
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from splat_amd import abi, scene  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden_c1.npz")


def compute():
    sc = scene.scene_c1()
    cam = scene.camera_c1(seed=11, height=64, width=64)
    rays = O.raygen(cam)
    hit0, t0 = O.next_hit(rays, sc.ntris())
    out = {"rays_sha256": np.frombuffer(hashlib.sha256(rays.tobytes()).digest(), dtype=np.uint8),
           "hit_primary": hit0, "t_primary": t0}
    for name, depth, fr in (("d2_refract", 2, 1.0), ("d4_seeded", 4, math.nan)):
        p = abi.run_params(depth, sc.first_diffuse, fr, seed=77)
        r, h, live = O.run_evolution(rays, sc.ftris, p)
        lm = abi.lightmap_params(64, 64)
        tex, r2 = O.light_map(r, h, sc.ftris, lm)
        out[f"{name}_hit"] = h
        out[f"{name}_live"] = np.array([live])
        out[f"{name}_rays_sha256"] = np.frombuffer(hashlib.sha256(r2.tobytes()).digest(), dtype=np.uint8)
        out[f"{name}_status"] = r2["status"].copy()
        out[f"{name}_tex"] = tex
    return out


if __name__ == "__main__":
    np.savez_compressed(OUT, **compute())
    print("wrote", OUT, os.path.getsize(OUT), "bytes")
