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


def test_autotuner_cache_hit_needs_no_device(tmp_path):
    """A cached choice for the same scene set and device is returned as stored
    (keys: grid side -> P2G layout, and the batch configuration fields)."""
    import json

    from coffee_sim_amd import autotune

    specs = pour_specs(8, seed=3)
    key = autotune.specs_key(specs, "gpu0")
    assert key != autotune.specs_key(pour_specs(8, seed=4), "gpu0")
    cfg = {"sweep_groups": 512, "residual_groups": 1024, "check_batches": 4, "use_graphs": 1}
    path = tmp_path / "tune.json"
    path.write_text(json.dumps({key: {"p2g_tiled": {"48": True, "96": False}, "config": cfg}}))
    got = autotune.tune(None, specs, cache=path)  # no context: a hit never touches the device
    assert got == {"p2g_tiled": {48: True, 96: False}, "config": cfg}
    assert autotune.config_from(got).sweep_groups == 512
