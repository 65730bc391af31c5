"""bench.py's host-side helpers (no GPU): the committed counter summaries it reads resolve for
the leaf depths the autotuner picks on the 1 MP frame, and the issue roofline is consistent."""
import bench


def test_counter_summaries_resolve_for_tuned_depths():
    for depth in (10, 11):
        traffic, src = bench.hitter_traffic(depth)
        assert traffic and traffic > 1e6 and src.startswith("profiles/")
        flop, src = bench.hitter_pmc_flops(depth)
        assert flop and flop > 1e6 and src.startswith("profiles/")
        vc = bench.valu_counts(depth)
        assert len(vc["k_hit_bv"]) == 5 and len(vc["k_bounce"]) == 4
    assert bench.hitter_traffic(3) == (None, None) and bench.valu_counts(3) is None


def test_valu_roofline():
    r = bench.valu_roofline("k", [614_400_000, 307_200_000], [1.0, 1.0], 256)
    assert r["peak"] == 614.4  # 256 CUs x 2.4 GHz wave64 instructions per s
    assert [p["frac"] for p in r["per_launch"]] == [1.0, 0.5] and r["frac"] == 0.75
