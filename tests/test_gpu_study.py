"""The time-step study (vumat-pd_amd/study.py), its three time steps batched on the GPU,
against the same three runs of the reference's own runner (oracle/_ref, on this host's CPU):
the study completes with the reference study's keys, each variant's outputs equal those of the
same case file run alone on the GPU byte for byte, and the coarse/fine comparisons agree with
those computed from the reference runs' outputs."""
import json
import os
import sys

import pytest

import case_files as cf

pytestmark = pytest.mark.gpu


def banner(msg):
    print("[gpu-case]", msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def vpd():
    import vpd as V
    if V.device_count() == 0:
        pytest.fail("no HIP device visible to the GPU tests")
    return V


KEYS = {"status", "scope", "base_case_sha256", "material_sha256", "source_sha256",
        "binary_sha256", "runs", "comparisons", "continuum_screen"}


def test_time_step_study_vs_reference_runs(vpd, ref, tmp_path):
    banner("time-step study on the GPU runner vs reference runs")
    import study
    case = cf.write_case(str(tmp_path / "in"), name="ts", nx=21, ny=7, nz=3, steps=240,
                         amplitude=0.6, rise_time=0.006, output_every=10, snapshot_every=240)
    report = study.run_study(case, tmp_path / "study")
    assert report["status"] == "completed"
    assert set(report) == KEYS
    on_disk = json.loads((tmp_path / "study" / "study.json").read_text())
    assert on_disk == json.loads(json.dumps(report))
    assert [r["label"] for r in report["runs"]] == ["dt_x2", "dt_x1", "dt_x0p5"]
    assert [r["steps"] for r in report["runs"]] == [120, 240, 480]
    assert all(r["returncode"] == 0 and r["result_sha256"] for r in report["runs"])
    assert (tmp_path / "study" / "batch.log").is_file()
    # Batched == alone, and the same three runs through the reference's runner.
    results = []
    for r in report["runs"]:
        case_file = tmp_path / "study" / (r["label"] + ".json")
        sel = json.loads(case_file.read_text())
        sel["output"]["directory"] = str(tmp_path / "alone" / r["label"])
        alone = tmp_path / ("alone_" + r["label"] + ".json")
        alone.write_text(json.dumps(sel))
        assert vpd.run_case_json(str(alone)) == 0
        for f in ("timeseries.csv", "field_snapshots.csv"):
            assert (tmp_path / "study" / r["label"] / f).read_bytes() == \
                (tmp_path / "alone" / r["label"] / f).read_bytes(), (r["label"], f)
        sel["output"]["directory"] = str(tmp_path / "ref" / r["label"])
        path = tmp_path / ("ref_" + r["label"] + ".json")
        path.write_text(json.dumps(sel))
        assert ref.ref().ref_run_case_json(os.fsencode(str(path)),
                                           os.fsencode(ref.REF_DIR)) == 0
        results.append(study.VariantRun.load(sel["output"]["directory"], r["dt_ms"]))
    expect = [study.compare(results[i], results[i + 1], 0.6) for i in range(2)]
    for got, want in zip(report["comparisons"], expect):
        assert got["common_history_times"] == want["common_history_times"]
        assert got["coarser_peak_eqps_reference_xyz_mm"] == want["coarser_peak_eqps_reference_xyz_mm"]
        for key in ("stress_history_nrmse_percent", "displacement_rmse_percent_of_amplitude",
                    "max_eqps_relative_difference_percent"):
            assert abs(got[key] - want[key]) <= 1e-8 * max(1.0, abs(want[key])), key
    print(json.dumps(report["comparisons"], indent=1))
