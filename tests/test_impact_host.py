"""The impact driver's phase split (vumat-pd_amd/impact.py) on a constructed log: the intact
phase ends at the first logged chunk whose mean damage index reaches 1%. No GPU work."""
import impact


def rec(step, k, secs, damage, plastic=10):
    return dict(step=step, chunk_steps=k, chunk_seconds=secs, mean_damage_index=damage,
                plastic_points=plastic)


def test_phase_split():
    log = [rec(500, 500, 1.0, 0.0), rec(1000, 500, 1.0, 0.004, 20), rec(1500, 500, 0.5, 0.02),
           rec(1700, 200, 0.1, 0.9, 0)]
    p = impact.phases(log, 1000)
    assert p["intact"]["steps"] == 1000 and p["intact"]["last_step"] == 1000
    assert p["intact"]["particle_steps_per_s"] == 1000 * 1000 / 2.0
    assert p["intact"]["plastic_points_at_end"] == 20
    assert p["fragmenting"]["steps"] == 700 and p["fragmenting"]["last_step"] == 1700
    assert abs(p["fragmenting"]["particle_steps_per_s"] - 1000 * 700 / 0.6) < 1e-6


def test_phase_split_never_damaged():
    p = impact.phases([rec(500, 500, 1.0, 0.0)], 10)
    assert p["intact"]["steps"] == 500
    assert p["fragmenting"]["steps"] == 0 and p["fragmenting"]["particle_steps_per_s"] is None


def test_chunk_ends():
    # Every 100 steps over the first 500 (where the damage front forms), then every 500.
    assert impact.chunk_ends(1700, 500, 100, 500) == [100, 200, 300, 400, 500, 1000, 1500, 1700]
    assert impact.chunk_ends(250, 500, 100, 500) == [100, 200, 250]
    assert impact.chunk_ends(1000, 500) == [500, 1000]
    assert impact.chunk_ends(1000, 400, 150, 500) == [150, 300, 450, 500, 900, 1000]
