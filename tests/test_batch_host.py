"""Host-side pieces of the batched workload that need no device."""
from __future__ import annotations

from coffee_sim_amd.batch import pour_specs


def test_scene_set_draws_are_seeded_and_in_range():
    specs = pour_specs(64, seed=7)
    assert specs == pour_specs(64, seed=7)
    assert specs != pour_specs(64, seed=8)
    assert {s.side for s in specs} == {48, 64, 80, 96}
    assert all(20_000 <= s.particles <= 400_000 for s in specs)
    assert all(30.0 <= s.kettle_angle_deg <= 60.0 for s in specs)
    assert all(500 <= s.pour_per_substep <= 3000 for s in specs)
    # log-uniform: about half the scenes below the geometric mean of the range
    below = sum(s.particles < (20_000 * 400_000) ** 0.5 for s in specs)
    assert 20 <= below <= 44
