"""The time-step study's comparison metrics and base-case checks (vumat-pd_amd/study.py) on
constructed inputs whose answers are known in closed form. No GPU work."""
import math

import pytest

import study


def history(times, stress):
    return [{"time_ms": repr(t), "average_axial_stress_MPa": repr(s)} for t, s in zip(times, stress)]


def fields(points, u, eqps):
    out = {}
    for k, ((x, y, z), (ux, uy, uz), e) in enumerate(zip(points, u, eqps)):
        out[str(k)] = {"x0_mm": repr(x), "y0_mm": repr(y), "z0_mm": repr(z), "ux_mm": repr(ux),
                       "uy_mm": repr(uy), "uz_mm": repr(uz), "eqps": repr(e)}
    return out


POINTS = [(0.0, 0.0, 0.0), (2.0, 0.0, 0.0), (4.0, 2.0, 0.0)]


def run(dt, times, stress, u, eqps):
    return {"dt_ms": dt, "history": history(times, stress), "fields": fields(POINTS, u, eqps)}


def test_identical_runs_compare_to_zero():
    a = run(2e-5, [0.0, 0.1, 0.2], [0.0, 50.0, 100.0], [(0.1, 0, 0)] * 3, [0.0, 0.01, 0.02])
    c = study.compare(a, a, 2.0)
    assert c["common_history_times"] == 3
    assert c["stress_history_nrmse_percent"] == 0.0
    assert c["displacement_rmse_percent_of_amplitude"] == 0.0
    assert c["max_eqps_relative_difference_percent"] == 0.0
    assert c["coarser_peak_eqps_reference_xyz_mm"] == [4.0, 2.0, 0.0]


def test_known_offsets():
    # The fine run has extra output times (only shared times count), its stress is 3 MPa
    # higher everywhere, every displacement component 0.01 mm larger, peak ε̄ᵖ 0.025.
    coarse = run(4e-5, [0.0, 0.1, 0.2], [0.0, 60.0, -120.0], [(0.1, 0.0, 0.0)] * 3,
                 [0.0, 0.01, 0.02])
    fine = run(2e-5, [0.0, 0.05, 0.1, 0.15, 0.2], [3.0, 7.0, 63.0, 9.0, -117.0],
               [(0.11, 0.01, 0.01)] * 3, [0.0, 0.025, 0.02])
    c = study.compare(coarse, fine, -2.0)
    assert c["common_history_times"] == 3
    assert math.isclose(c["stress_history_nrmse_percent"], 100 * 3.0 / 120.0)
    assert math.isclose(c["displacement_rmse_percent_of_amplitude"], 100 * 0.01 / 2.0)
    assert math.isclose(c["max_eqps_relative_difference_percent"], 100 * 0.005 / 0.025)
    assert c["finer_peak_eqps_reference_xyz_mm"] == [2.0, 0.0, 0.0]


def test_comparison_refusals():
    a = run(4e-5, [0.0, 0.1], [0.0, 10.0], [(0, 0, 0)] * 3, [0, 0, 0])
    b = run(2e-5, [0.05, 0.1], [1.0, 10.0], [(0, 0, 0)] * 3, [0, 0, 0])
    with pytest.raises(ValueError, match="Fewer than two common history times"):
        study.compare(a, b, 1.0)
    z = run(4e-5, [0.0, 0.1], [0.0, 0.0], [(0, 0, 0)] * 3, [0, 0, 0])
    with pytest.raises(ValueError, match="Zero reference stress history"):
        study.compare(z, z, 1.0)
    moved = run(4e-5, [0.0, 0.1], [0.0, 10.0], [(0, 0, 0)] * 3, [0, 0, 0])
    moved["fields"]["1"]["x0_mm"] = "2.5"
    with pytest.raises(ValueError, match="geometry changed"):
        study.compare(a, moved, 1.0)


def base_case(**over):
    case = {"type": "vumat_pd_boundary_case", "materials": [{}],
            "time_integration": {"steps": 200, "dt": 2.5e-5, "totalTime": 200 * 2.5e-5},
            "dynamics": {"outputEvery": 10}, "boundary_conditions": {"amplitude": 2.0}}
    for key, value in over.items():
        section, name = key.split("__")
        case[section][name] = value
    return case


def test_base_case_checks():
    assert study.check_case(base_case()) is None
    assert "even" in study.check_case(base_case(time_integration__steps=201))
    assert "even" in study.check_case(base_case(dynamics__outputEvery=3))
    assert "totalTime" in study.check_case(base_case(time_integration__totalTime=1.0))
    assert "amplitude" in study.check_case(base_case(boundary_conditions__amplitude=0.0))
