"""The time-step study's comparison metrics and base-case checks (vumat-pd_amd/study.py) on
constructed inputs whose answers are known in closed form. No GPU work.

The metrics are those of benchmarks/tensile3d/time_step_study.py:29-66: stress-history NRMSE
at the output times both runs share, relative to the fine run's peak |stress| there;
displacement RMSE per component relative to |amplitude|; peak-ε̄ᵖ difference relative to the
fine peak; each run's first ε̄ᵖ maximum in file order."""
import math

import numpy as np
import pytest

import study

POINTS = [(0.0, 0.0, 0.0), (2.0, 0.0, 0.0), (4.0, 2.0, 0.0)]


def run(dt, times, stress, u, eqps, points=POINTS, ids=None):
    n = len(points)
    ids = [str(k) for k in range(n)] if ids is None else ids
    return study.VariantRun(dt_ms=dt, time=np.array(times, float), stress=np.array(stress, float),
                            finite=np.ones(len(times), bool), particle=np.array(ids, dtype=object),
                            x0=np.array(points, float), u=np.array(u, float),
                            eqps=np.array(eqps, float), axial=np.zeros(n))


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
    # higher everywhere (its peak |stress| at the shared times is 117), every displacement
    # component 0.01 mm larger, peak ε̄ᵖ 0.025.
    coarse = run(4e-5, [0.0, 0.1, 0.2], [0.0, 60.0, -120.0], [(0.1, 0.0, 0.0)] * 3,
                 [0.0, 0.01, 0.02])
    fine = run(2e-5, [0.0, 0.05, 0.1, 0.15, 0.2], [3.0, 7.0, 63.0, 9.0, -117.0],
               [(0.11, 0.01, 0.01)] * 3, [0.0, 0.025, 0.02])
    c = study.compare(coarse, fine, -2.0)
    assert c["common_history_times"] == 3
    assert math.isclose(c["stress_history_nrmse_percent"], 100 * 3.0 / 117.0)  # fine peak
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
    moved = run(4e-5, [0.0, 0.1], [0.0, 10.0], [(0, 0, 0)] * 3, [0, 0, 0],
                points=[(0.0, 0.0, 0.0), (2.5, 0.0, 0.0), (4.0, 2.0, 0.0)])
    with pytest.raises(ValueError, match="geometry changed"):
        study.compare(a, moved, 1.0)
    other = run(4e-5, [0.0, 0.1], [0.0, 10.0], [(0, 0, 0)] * 3, [0, 0, 0], ids=["0", "1", "7"])
    with pytest.raises(ValueError, match="points differ"):
        study.compare(a, other, 1.0)


def test_rows_paired_by_particle_and_times_by_value():
    # The fine snapshot lists the particles in another order; the time 0.30000000000000004
    # (0.1 + 0.2) is the same output time as 0.3. The peak location is the first maximum in
    # the coarse file's order.
    coarse = run(4e-5, [0.1, 0.3], [10.0, 20.0], [(1, 0, 0), (2, 0, 0), (3, 0, 0)],
                 [0.5, 0.1, 0.5])
    fine = run(2e-5, [0.1, 0.1 + 0.2], [10.0, 24.0], [(3, 0, 0), (1, 0, 0), (2, 0, 0)],
               [0.5, 0.5, 0.1], points=[POINTS[2], POINTS[0], POINTS[1]], ids=["2", "0", "1"])
    c = study.compare(coarse, fine, 1.0)
    assert c["common_history_times"] == 2
    assert math.isclose(c["stress_history_nrmse_percent"], 100 * math.sqrt(8.0) / 24.0)
    assert c["displacement_rmse_percent_of_amplitude"] == 0.0
    assert c["coarser_peak_eqps_reference_xyz_mm"] == list(POINTS[0])
    assert c["finer_peak_eqps_reference_xyz_mm"] == list(POINTS[2])


def base_case(**over):
    case = {"type": "vumat_pd_boundary_case", "materials": [{}],
            "time_integration": {"steps": 200, "dt": 2.5e-5, "totalTime": 200 * 2.5e-5},
            "dynamics": {"outputEvery": 10}, "boundary_conditions": {"amplitude": 2.0}}
    for key, value in over.items():
        section, name = key.split("__")
        case[section][name] = value
    return case


def test_base_case_checks():
    assert study.case_problem(base_case()) is None
    assert "even" in study.case_problem(base_case(time_integration__steps=201))
    assert "even" in study.case_problem(base_case(dynamics__outputEvery=3))
    assert "exceed" in study.case_problem(base_case(dynamics__outputEvery=400))
    assert "positive" in study.case_problem(base_case(time_integration__dt=-1.0))
    assert "totalTime" in study.case_problem(base_case(time_integration__totalTime=1.0))
    assert "amplitude" in study.case_problem(base_case(boundary_conditions__amplitude=0.0))
    two = base_case()
    two["materials"].append({})
    assert "one material" in study.case_problem(two)


def test_variant_cases_scale_time_and_cadence():
    case = base_case()
    case.update(name="bar", geometry={}, output={"directory": "x"})
    case["materials"][0]["materialConfigPath"] = "m.json"
    case["dynamics"]["snapshotEvery"] = 50
    for label, factor in study.VARIANTS:
        v = study.variant_case(case, label, factor, "/abs/m.json", "/out")
        ti = v["time_integration"]
        assert ti["steps"] == round(200 / factor) and ti["dt"] == 2.5e-5 * factor
        assert ti["totalTime"] == 200 * 2.5e-5
        assert v["dynamics"]["outputEvery"] == round(10 / factor)
        assert v["dynamics"]["snapshotEvery"] == ti["steps"]
        assert v["output"]["directory"] == "/out/" + label and v["name"] == "bar_" + label
    assert case["time_integration"]["steps"] == 200  # the base case is left as it was
