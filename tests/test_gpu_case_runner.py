"""The GPU case runner (vpd_run_case_json) against the reference's own runner.

* Small boundary-value cases run through oracle/_ref (the reference's runBoundaryValueCase
  compiled from its sources, on this host's CPU) and through the GPU runner: the same CSV
  headers, rows, step and particle columns; the load-path columns (time, ramp, engineering
  strain) bit for bit; every other column within the stated tolerance of its scale; the same
  metadata.json apart from the measured runtime.
* The archived pd_coarse run of the reference (tests/golden/pd_coarse, another OS, compiler
  and libm): the full 4800-step case on the GPU against the archived outputs.
* The fused run diagnostics change no state bit and do not depend on the launch tuning.
"""
import json
import os
import re
import sys

import numpy as np
import pytest

import case_files as cf
from conftest import JC_PROPS, ROOT

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(ROOT, "tests", "golden", "pd_coarse")

TS_EXACT = ("time_ms", "load_displacement_mm", "load_velocity_mm_per_ms", "engineering_strain")


def banner(msg):
    print("[gpu-case]", msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def vpd():
    import vpd as V
    if V.device_count() == 0:
        pytest.fail("no HIP device visible to the GPU tests")
    return V


def run_reference(orc, path):
    return orc.ref().ref_run_case_json(os.fsencode(path), os.fsencode(orc.REF_DIR))


def column_errors(a, b, header, floor):
    """max |a − b| / max(|b|, floor) per column."""
    scale = np.maximum(np.abs(b).max(0), floor)
    return dict(zip(header, (np.abs(a - b).max(0) / scale)))


def compare_timeseries(path_a, path_b, tol, exact=TS_EXACT):
    ha, ra = cf.read_csv(path_a)
    hb, rb = cf.read_csv(path_b)
    assert ha == hb
    assert [r[:2] for r in ra] == [r[:2] for r in rb]
    a, b = cf.numeric(ra, 2), cf.numeric(rb, 2)
    names = ha[2:]
    assert np.array_equal(a[:, -1], b[:, -1])  # finite flag
    for name in exact:
        k = names.index(name)
        assert np.array_equal(a[:, k], b[:, k]), name
    # Net internal force is rounding noise around zero: compare it to the reaction scale.
    k_net, k_rea = names.index("net_internal_force_N"), names.index("reaction_force_N")
    force_scale = np.abs(b[:, k_rea]).max()
    assert np.abs(a[:, k_net]).max() <= 1e-9 * force_scale
    errs = column_errors(a, b, names, 1e-30)
    for name, e in errs.items():
        if name in exact or name in ("net_internal_force_N", "finite"):
            continue
        assert e <= tol, (name, e)
    return errs


def compare_snapshots(path_a, path_b, tol):
    ha, ra = cf.read_csv(path_a)
    hb, rb = cf.read_csv(path_b)
    assert ha == hb
    assert [r[:4] for r in ra] == [r[:4] for r in rb]
    a, b = cf.numeric(ra, 4), cf.numeric(rb, 4)
    names = ha[4:]
    for name in ("x0_mm", "y0_mm", "z0_mm"):
        k = names.index(name)
        assert np.array_equal(a[:, k], b[:, k])
    errs = column_errors(a, b, names, 1e-30)
    for name, e in errs.items():
        assert e <= tol, (name, e)
    return errs


def compare_metadata(path_a, path_b, tol):
    ta, tb = open(path_a).read(), open(path_b).read()
    number = re.compile(r"-?\d+(\.\d+)?([eE][-+]?\d+)?")
    assert number.sub("#", ta) == number.sub("#", tb)  # same text apart from the numbers
    ma, mb = json.loads(ta), json.loads(tb)
    assert ma["status"] == mb["status"]
    for key in ma:
        if key != "materials":
            assert ma[key] == mb[key], key
    for xa, xb in zip(ma["materials"], mb["materials"]):
        for key, va in xa.items():
            vb = xb[key]
            if key == "runtimeSeconds":
                continue
            if isinstance(va, float) and isinstance(vb, float):
                assert abs(va - vb) <= tol * max(abs(vb), 1e-30) or abs(va - vb) <= 1e-9, key
            else:
                assert va == vb, key


CASES = [
    # name, write_case arguments, tolerance (relative to each column's scale). Measured on
    # MI355X: ≤ 5e-14 for every column but the relative energy error (≤ 4e-13, a ratio of
    # nearly cancelling sums).
    ("elastic_plastic", dict(nx=21, ny=7, nz=3, steps=240, amplitude=0.6, rise_time=0.004,
                             output_every=10, snapshot_every=120), 1e-11),
    ("y_loading_indicator", dict(nx=7, ny=19, nz=3, steps=160, amplitude=-0.4, rise_time=0.003,
                                 component="y", output_every=8, snapshot_every=80,
                                 internalIndicator={"indices": [1, 4], "reduction": "max"}),
     1e-11),
    ("critical_stretch", dict(nx=21, ny=7, nz=3, steps=120, amplitude=2.0, rise_time=0.004,
                              output_every=6, snapshot_every=60,
                              damage={"model": "critical_stretch", "initiation": 0.004,
                                      "failure": 0.0, "deletionEnabled": False}), 1e-10),
]


@pytest.mark.parametrize("name,kw,tol", CASES, ids=[c[0] for c in CASES])
def test_case_runner_vs_reference_build(vpd, ref, tmp_path, name, kw, tol):
    banner("case runner vs reference build: %s" % name)
    gpu_case = cf.write_case(str(tmp_path / "gpu"), name=name, **kw)
    ref_case = cf.write_case(str(tmp_path / "ref"), name=name, **kw)
    rc_gpu = vpd.run_case_json(gpu_case, ref.REF_DIR)
    rc_ref = run_reference(ref, ref_case)
    assert rc_gpu == rc_ref
    dg, dr = cf.output_dir(gpu_case), cf.output_dir(ref_case)
    errs = compare_timeseries(os.path.join(dg, "timeseries.csv"),
                              os.path.join(dr, "timeseries.csv"), tol)
    serrs = compare_snapshots(os.path.join(dg, "field_snapshots.csv"),
                              os.path.join(dr, "field_snapshots.csv"), tol)
    print("timeseries max rel err:", {k: float("%.2e" % v) for k, v in errs.items()})
    print("snapshot max rel err:", {k: float("%.2e" % v) for k, v in serrs.items()})
    compare_metadata(os.path.join(dg, "metadata.json"), os.path.join(dr, "metadata.json"),
                     max(tol, 1e-9))


def test_pd_coarse_vs_archived_reference_run(vpd, tmp_path):
    banner("pd_coarse 4800 steps vs the archived reference outputs")
    with open(os.path.join(GOLDEN, "case.json")) as f:
        case = json.load(f)
    case["materials"][0]["materialConfigPath"] = os.path.join(GOLDEN, "johnson_cook.json")
    case["output"]["directory"] = str(tmp_path / "out")
    path = str(tmp_path / "pd_coarse.json")
    with open(path, "w") as f:
        json.dump(case, f)
    assert vpd.run_case_json(path, "") == 0  # the archived run passed its acceptance checks
    out = case["output"]["directory"]
    ha, ra = cf.read_csv(os.path.join(out, "timeseries.csv"))
    hb, rb = cf.read_csv(os.path.join(GOLDEN, "timeseries.csv.gz"))
    assert ha == hb and [r[:2] for r in ra] == [r[:2] for r in rb]
    a, b = cf.numeric(ra, 2), cf.numeric(rb, 2)
    names = ha[2:]
    errs = column_errors(a, b, names, 1e-30)
    print("pd_coarse timeseries max rel err:", {k: float("%.2e" % v) for k, v in errs.items()})
    for name in ("time_ms", "load_displacement_mm", "load_velocity_mm_per_ms",
                 "engineering_strain"):
        assert errs[name] <= 1e-15, name
    # Measured on MI355X: ≤ 3e-13 per column (2.8e-12 for the relative energy error) after
    # 4800 steps against a run on another OS, compiler and libm.
    for name in ("reaction_force_N", "nominal_stress_MPa", "average_axial_strain",
                 "average_axial_stress_MPa", "max_eqps", "kinetic_energy_Nmm",
                 "internal_energy_Nmm", "external_work_Nmm", "lateral_strain",
                 "relative_energy_error"):
        assert errs[name] <= 1e-9, (name, errs[name])
    k = names.index("net_internal_force_N")
    assert np.abs(a[:, k]).max() <= 1e-6  # netForceTolerance of the case
    hs, rs = cf.read_csv(os.path.join(out, "field_snapshots.csv"))
    hg, rg = cf.read_csv(os.path.join(GOLDEN, "field_snapshots.csv.gz"))
    assert hs == hg and [r[:4] for r in rs] == [r[:4] for r in rg]
    serrs = column_errors(cf.numeric(rs, 4), cf.numeric(rg, 4), hs[4:], 1e-30)
    print("pd_coarse snapshot max rel err:", {k: float("%.2e" % v) for k, v in serrs.items()})
    for name, e in serrs.items():  # measured ≤ 4e-11 (vz)
        assert e <= 1e-9, (name, e)
    meta = json.load(open(os.path.join(out, "metadata.json")))
    gold = json.load(open(os.path.join(GOLDEN, "metadata.json")))
    for key in ("maxEqps", "finalReaction_N", "finalNominalStress_MPa", "finalEnergyError",
                "earlyElasticSlope_MPa"):
        va, vb = meta["materials"][0][key], gold["materials"][0][key]
        assert abs(va - vb) <= 1e-9 * abs(vb), (key, va, vb)


def diag_bar(vpd, **knobs):
    kw = dict(nx=30, ny=8, nz=8, spacing=1.0, horizon=3.015, props=JC_PROPS, dt=8.5377e-5,
              density=7.85e-3, velocity=60.0, damage=(0.003, 0.0))
    bar = vpd.Bar(**kw)
    if knobs:
        bar.ctx.set_tuning(**knobs)
    return bar


def test_diagnostics_change_no_state_and_ignore_tuning(vpd):
    banner("run diagnostics: state unchanged, tuning-independent records")
    plain = diag_bar(vpd)
    plain.step(80)
    recs = []
    for knobs in ({}, dict(force_rounds=3, defgrad_rounds=2, force_lds=50000)):
        bar = diag_bar(vpd, **knobs)
        w = bar.vol * (bar.x[:, 0] > 3.5) * (bar.x[:, 0] < 25.5)
        lat = (bar.x[:, 1] == bar.x[:, 1].min()) * 1 + (bar.x[:, 1] == bar.x[:, 1].max()) * 2
        bar.ctx.enable_diagnostics(0, 1, w, lat.astype(np.uint8), 1, capacity=100)
        bar.step(80)
        recs.append(bar.ctx.fetch_diagnostics())
        for name in ("u", "v", "a", "stressN", "statev", "damage"):
            assert np.array_equal(bar.get(name), plain.get(name)), name
    assert recs[0].shape == (80, vpd.DIAG_STRIDE)
    assert np.array_equal(recs[0], recs[1])
    r = recs[0]
    assert r[:, vpd.DIAG_SLOTS.index("kinetic")].max() > 0
    assert r[:, vpd.DIAG_SLOTS.index("max_eqps")].max() > 0
    assert (r[:, vpd.DIAG_SLOTS.index("nonfinite_material")] == 0).all()
    # Global balance: Σ f·V is rounding noise next to the end-band reaction.
    net = np.abs(r[:, 6:9]).max()
    assert net <= 1e-9 * np.abs(r[:, vpd.DIAG_SLOTS.index("reaction")]).max()


def test_diagnostics_buffer_full_is_an_error(vpd):
    banner("run diagnostics: full record buffer")
    bar = diag_bar(vpd)
    bar.ctx.enable_diagnostics(0, 1, capacity=3)
    bar.step(3)
    with pytest.raises(vpd.VpdError, match="buffer is full"):
        bar.step(1)
    assert bar.ctx.fetch_diagnostics().shape[0] == 3
    bar.step(1)
    assert bar.ctx.fetch_diagnostics().shape[0] == 1
