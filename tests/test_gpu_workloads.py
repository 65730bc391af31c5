"""The batched and served workloads at their full sizes (needs an MI355X).

  * 64 frames in shared engine steps: resolutions from {256^2, 512^2, 1024^2}, icospheres of
    subdivision 2-5 over a plane, 4 bounces (the batching workload);
  * served requests over the 8 resident scenes (320 to ~178k triangles): one request per scene at
    2048^2 pixels x 16 stratified wavelengths (67.1M rays each) plus mixed 256^2-2048^2 x 1-16
    requests, continuously admitted against an engine arena that holds at most two of the large
    frames at a time (the serving workload).

Every frame's final rays and hit indices are compared bit for bit with the same frame traced alone
(one frame per launch sequence), its light map within 1e-4 relative L2 (only the atomic
accumulation order differs), all on the device. On a sample of frames, rays drawn at random are
traced alone by the oracle and must end in exactly the GPU's states (matched by sample id).
"""
import numpy as np
import pytest

from oracle import oracle as O
from splat_amd import abi, api, bv, serve
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _outputs(rq):
    import torch

    rq.rays_out = torch.empty((rq.rays, abi.ADRAY.itemsize), dtype=torch.uint8, device="cuda")
    rq.hit_out = torch.empty(rq.rays, dtype=torch.int32, device="cuda")
    rq.tex = torch.full((rq.lm.channels, rq.lm.height, rq.lm.width), 3.0, dtype=torch.float32, device="cuda")


def _equals_solo(rq, scenes):
    """The request's engine outputs against the same frame traced alone."""
    import torch

    rs = scenes[rq.scene]
    fr = api.Frame(rq.cam, rs.hitter, rs.tris, rq.params, rq.lm)
    fr.trace()
    torch.cuda.synchronize()
    assert torch.equal(rq.hit_out, fr.hit_idx), rq.tag
    assert torch.equal(rq.rays_out, fr.rays), rq.tag
    a, b = rq.tex.double(), fr.tex.double()
    nb = float(torch.linalg.vector_norm(b))
    assert nb > 0 and float(torch.linalg.vector_norm(a - b)) <= 1e-4 * nb, rq.tag
    del fr


def _oracle_sample(rq, rs, n, seed):
    """n rays of the request drawn at random, traced alone by the oracle (BV-culled CPU hitter,
    equal to brute force), against the GPU's final states found by sample id."""
    import torch

    rng = np.random.default_rng(seed)
    q = np.sort(rng.choice(rq.rays, size=min(n, rq.rays), replace=False))
    gen = O.raygen_at(rq.cam, q)
    depth = max(1, int(round(0.5 * np.log2(max(rs.sc.n_tris, 2)))))
    counts, mat, stride = bv.pack_bv_tris(bv.object_partition(rs.sc.ftris, depth)[1], rs.sc.n_tris)
    want_r, want_h, _ = O.run_evolution(gen, rs.sc.ftris, rq.params, bv=(counts, mat, stride))
    want_r = O.final_evolution(want_r, want_h, rs.sc.ftris)
    dest = rq.rays_out[:, 112:120].contiguous().view(torch.int64).flatten()
    inv = torch.empty(rq.rays, dtype=torch.int64, device="cuda")
    inv[dest - 1] = torch.arange(rq.rays, dtype=torch.int64, device="cuda")
    idx = inv[torch.from_numpy(want_r["dest"].astype(np.int64) - 1).cuda()]
    got = rq.rays_out[idx].cpu().numpy().reshape(-1).view(abi.ADRAY)
    got_h = rq.hit_out[idx].cpu().numpy()
    assert np.array_equal(got_h, want_h), rq.tag
    assert H.same_rays(got, want_r), rq.tag


def test_batched_64_frames_equal_their_solo_traces():
    import torch

    scenes, reqs = serve.batched_workload(64)
    assert len(reqs) == 64 and {rq.cam.height for rq in reqs} == {256, 512, 1024}
    assert {scenes[rq.scene].sc.n_tris for rq in reqs} == {323, 1283, 5123, 20483}
    for rq in reqs:
        _outputs(rq)
    arena = int(sum(rq.rays for rq in reqs) * 440 * 1.05) + (1 << 28)
    eng = api.Engine(arena, max_slots=64)
    st = serve.run_requests(eng, scenes, reqs, realtime=False, max_pending_steps=8)
    eng.close()
    print(f"batched: {len(reqs)} frames, {st.rays} rays, {st.steps} steps", flush=True)
    assert st.steps == 5  # every frame admitted at once: 4 bounces + the final stage
    for rq in reqs:
        _equals_solo(rq, scenes)
    for k, rq in enumerate(sorted(reqs, key=lambda r: (r.scene, -r.rays))[::16]):
        _oracle_sample(rq, scenes[rq.scene], 1000, seed=k)
    torch.cuda.empty_cache()


@pytest.mark.timeout(900)
def test_served_requests_over_8_resident_scenes_equal_their_solo_traces():
    import torch

    scenes = serve.serving_scenes()
    tris = [rs.sc.n_tris for rs in scenes]
    assert len(scenes) == 8 and min(tris) == 323 and max(tris) == 177923
    fd = [rs.sc.first_diffuse for rs in scenes]
    big = serve.serving_requests(8, 100.0, 8, fd, seed=21, res_range=(2048, 2048), spp_range=(16, 16))
    for k, rq in enumerate(big):  # one large request per scene, arrivals interleaved with the mix
        rq.scene = k
        rq.params.first_diffuse = fd[k]
    mixed = serve.serving_requests(8, 100.0, 8, fd, seed=22)
    reqs = []
    for a, b in zip(big, mixed):
        reqs += [a, b]
    for k, rq in enumerate(reqs):
        rq.tag = k
        _outputs(rq)
    assert all(rq.rays == 2048 * 2048 * 16 for rq in big)
    eng = api.Engine(48 << 30, max_slots=64)
    assert 2 * eng.frame_bytes(big[0].cam) <= (48 << 30) < 3 * eng.frame_bytes(big[0].cam)
    st = serve.run_requests(eng, scenes, reqs, realtime=False)
    eng.close()
    print(f"served: {len(reqs)} requests, {st.rays} rays, {st.steps} steps, {st.refused} budget waits, "
          f"max {st.max_in_flight} in flight", flush=True)
    assert st.refused > 0 and st.max_in_flight >= 2  # admission waited on the budget, frames overlapped
    for rq in reqs:
        _equals_solo(rq, scenes)
        torch.cuda.empty_cache()
    for k, rq in enumerate(big[::3] + [big[7]]):  # the 178k-triangle scene included
        _oracle_sample(rq, scenes[rq.scene], 600, seed=100 + k)
