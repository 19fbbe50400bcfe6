"""The autotuner (no GPU): its candidate lattices -- the base set is the one for
2^20 agents, and other sizes scale every axis by the cube root of the agent
ratio, so the agents per cell stay the same --, the tie margin of its choices
and its launch shapes."""
from gpusteer_amd import autotune


def test_base_set_at_one_million():
    assert autotune.scaled_candidates(1 << 20) == list(autotune.BOIDS_CANDIDATES)


def test_eight_times_the_agents_doubles_every_axis():
    for (c, rf), (b, rb) in zip(autotune.scaled_candidates(1 << 23), autotune.BOIDS_CANDIDATES):
        assert rf == rb and c == tuple(2 * v for v in b)


def test_small_flocks_keep_whole_distinct_lattices():
    cands = autotune.scaled_candidates(4096)
    assert len(cands) == len(set(cands))
    assert all(v >= 8 and v % 8 == 0 for c, _ in cands for v in c)


def test_lattices_above_the_limit_are_left_out():
    cands = autotune.scaled_candidates(1 << 30)
    assert cands and all(c[0] * c[1] * c[2] <= autotune.MAX_CELLS for c, _ in cands)
    assert len(cands) < len(autotune.BOIDS_CANDIDATES)


def test_pick_keeps_the_earlier_candidate_within_the_margin():
    assert autotune.pick([1.00, 0.99]) == 0
    assert autotune.pick([1.00, 0.97]) == 1
    assert autotune.pick([0.5, 0.4, 0.395]) == 1
    assert autotune.pick([0.3]) == 0


def test_launch_shapes_are_the_library_set_default_first():
    assert autotune.SHAPES[0] == (256, 1)
    assert len(set(autotune.SHAPES)) == len(autotune.SHAPES)
    assert all(b in (128, 256, 512) and a in (1, 2) for b, a in autotune.SHAPES)
