"""The reference's seven tensile PD cases (benchmarks/tensile3d/cases/, the set
benchmarks/tensile3d/run_cases.py:95-96 runs as a process pool) as one batch on the GPU
(vpd_run_cases_json: every case a specimen of one particle block, its half-cosine ramp as a
band table, per-specimen run diagnostics).

* Each case's timeseries.csv and field_snapshots.csv equal, byte for byte, those of the same
  case run alone (vpd_run_case_json), and its metadata.json equals it apart from the measured
  runtime: the batch changes nothing a case computes.
* The archived reference runs of pd_coarse, pd_shared and pd_delta2 (tests/golden/, another OS,
  compiler and libm) are matched within 1e-9 of each column's scale.
"""
import json
import os
import sys

import numpy as np
import pytest

import case_files as cf
from conftest import ROOT

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(ROOT, "tests", "golden")
CASES = ["pd_coarse", "pd_shared", "pd_m1p5", "pd_delta4", "pd_delta2", "pd_dt50", "pd_dt12p5"]
MATERIAL = os.path.join(GOLDEN, "pd_coarse", "johnson_cook.json")


def banner(msg):
    print("[gpu-case]", msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def vpd():
    import vpd as V
    if V.device_count() == 0:
        pytest.fail("no HIP device visible to the GPU tests")
    return V


def write_cases(directory, names):
    paths = []
    for name in names:
        with open(os.path.join(GOLDEN, "cases", name + ".json")) as f:
            case = json.load(f)
        case["materials"][0]["materialConfigPath"] = MATERIAL
        case["output"]["directory"] = os.path.join(directory, name)
        path = os.path.join(directory, name + ".json")
        with open(path, "w") as f:
            json.dump(case, f, indent=1)
        paths.append(path)
    return paths


def without_runtime(meta):
    for m in meta["materials"]:
        m.pop("runtimeSeconds")
    return meta


def column_errors(a, b, header):
    scale = np.maximum(np.abs(b).max(0), 1e-30)
    return dict(zip(header, np.abs(a - b).max(0) / scale))


@pytest.fixture(scope="module")
def batch_run(vpd, tmp_path_factory):
    base = tmp_path_factory.mktemp("cases")
    banner("seven tensile cases as one batch")
    os.makedirs(base / "batch")
    paths = write_cases(str(base / "batch"), CASES)
    assert vpd.run_cases_json(paths) == 0  # every archived case passed its checks
    return base


@pytest.mark.parametrize("name", CASES)
def test_batched_case_equals_single_run(vpd, batch_run, name):
    banner("%s: batched == alone" % name)
    single = str(batch_run / "single")
    os.makedirs(single, exist_ok=True)
    path = write_cases(single, [name])[0]
    assert vpd.run_case_json(path) == 0
    a, b = os.path.join(str(batch_run / "batch"), name), os.path.join(single, name)
    for f in ("timeseries.csv", "field_snapshots.csv"):
        with open(os.path.join(a, f), "rb") as fa, open(os.path.join(b, f), "rb") as fb:
            assert fa.read() == fb.read(), f
    ma = without_runtime(json.load(open(os.path.join(a, "metadata.json"))))
    mb = without_runtime(json.load(open(os.path.join(b, "metadata.json"))))
    assert ma == mb


@pytest.mark.parametrize("name", ["pd_coarse", "pd_shared", "pd_delta2"])
def test_batched_case_vs_archived_reference_run(vpd, batch_run, name):
    banner("%s: batched vs the archived reference outputs" % name)
    out = os.path.join(str(batch_run / "batch"), name)
    arch = os.path.join(GOLDEN, name)
    ha, ra = cf.read_csv(os.path.join(out, "timeseries.csv"))
    hb, rb = cf.read_csv(os.path.join(arch, "timeseries.csv.gz"))
    assert ha == hb and [r[:2] for r in ra] == [r[:2] for r in rb]
    errs = column_errors(cf.numeric(ra, 2), cf.numeric(rb, 2), ha[2:])
    print("%s timeseries max rel err:" % name,
          {k: float("%.2e" % v) for k, v in errs.items()})
    for col in ("time_ms", "load_displacement_mm", "load_velocity_mm_per_ms",
                "engineering_strain"):
        assert errs[col] <= 1e-15, col
    for col, e in errs.items():
        if col not in ("net_internal_force_N", "finite"):
            assert e <= 1e-9, (col, e)
    hs, rs = cf.read_csv(os.path.join(out, "field_snapshots.csv"))
    hg, rg = cf.read_csv(os.path.join(arch, "field_snapshots.csv.gz"))
    assert hs == hg and [r[:4] for r in rs] == [r[:4] for r in rg]
    serrs = column_errors(cf.numeric(rs, 4), cf.numeric(rg, 4), hs[4:])
    print("%s snapshot max rel err:" % name, {k: float("%.2e" % v) for k, v in serrs.items()})
    for col, e in serrs.items():
        assert e <= 1e-9, (col, e)
    meta = json.load(open(os.path.join(out, "metadata.json")))
    gold = json.load(open(os.path.join(arch, "metadata.json")))
    assert meta["status"] == gold["status"] == "pass"
    for key in ("maxEqps", "finalReaction_N", "finalNominalStress_MPa", "finalEnergyError",
                "earlyElasticSlope_MPa"):
        va, vb = meta["materials"][0][key], gold["materials"][0][key]
        assert abs(va - vb) <= 1e-9 * abs(vb), (key, va, vb)
