"""The affine demo-case host on the GPU (vpd_run_case_json for vumat_pd_demo_case files)
against the reference's own runner (oracle/_ref, its DemoCaseRunner and the Fortran elastic and
Johnson-Cook VUMATs compiled from their sources).

* The reference's affine CTests (examples/affine_bar_elastic.json, affine_shear_elastic.json,
  kept under tests/golden/affine): same exit status, same CSV headers and rows, every column
  within the 6 significant digits the runner prints (rounding-noise columns — off-axis
  stresses, net forces, F errors — below the case tolerances on both sides), the same
  metadata.json.
* The same host with the Johnson-Cook material into the plastic range, and with
  critical-stretch damage.
"""
import json
import os
import re
import sys

import numpy as np
import pytest

import case_files as cf

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "affine")


def banner(msg):
    print("[gpu-case]", msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def vpd():
    import vpd as V
    if V.device_count() == 0:
        pytest.fail("no HIP device visible to the GPU tests")
    return V


def example(name, directory, out):
    case = json.load(open(os.path.join(GOLDEN, name + ".json")))
    case["materials"][0]["materialConfigPath"] = os.path.join(GOLDEN, "elastic.json")
    case["output"]["directory"] = out
    path = os.path.join(directory, os.path.basename(out) + ".json")
    with open(path, "w") as f:
        json.dump(case, f)
    return path


# Columns that are rounding noise around zero in these cases: both sides must stay below the
# case's own tolerance for them (F error, net force), or below 1e-9 of the stress scale.
NOISE = {"max_f_error": 1e-12, "net_force_x": 1e-7, "net_force_y": 1e-7, "net_force_z": 1e-7,
         "net_force_norm": 1e-7}


def compare_csv(path_a, path_b, keys, stress_cols):
    a, b = cf.read_csv(path_a), cf.read_csv(path_b)
    assert a[0] == b[0]
    ha = a[0]
    ra, rb = a[1], b[1]
    assert [r[:keys] for r in ra] == [r[:keys] for r in rb]
    xa, xb = cf.numeric(ra, keys), cf.numeric(rb, keys)
    scale = max(np.abs(xb[:, [ha.index(c) - keys for c in stress_cols]]).max(), 1.0)
    for k, name in enumerate(ha[keys:]):
        col_a, col_b = xa[:, k], xb[:, k]
        floor = NOISE.get(name, 1e-9 * scale)
        big = np.abs(col_b) > floor
        if big.any():
            rel = np.abs(col_a[big] - col_b[big]) / np.abs(col_b[big])
            assert rel.max() <= 2e-5, (name, rel.max())
        small = ~big
        assert (np.abs(col_a[small]) <= max(floor, 1e-5 * np.abs(col_b).max())).all(), name


def compare_metadata(path_a, path_b):
    ta, tb = open(path_a).read(), open(path_b).read()
    number = re.compile(r"-?\d+(\.\d+)?([eE][-+]?\d+)?")
    assert number.sub("#", ta) == number.sub("#", tb)
    ma, mb = json.loads(ta), json.loads(tb)
    for key in ("name", "status", "scope", "particles", "steps", "allValuesFinite",
                "tolerances", "notes", "checks"):
        assert ma[key] == mb[key], key
    for key in ("maxReferenceStressError", "maxJcRawStressError", "maxDamage"):
        assert abs(ma[key] - mb[key]) <= 1e-8 * max(abs(mb[key]), 1e-30) + 1e-12, key
    tol = mb["tolerances"]
    for m in (ma, mb):
        assert m["maxDeformationGradientError"] < tol["deformationGradientAbsolute"]
        assert m["maxNetForceNorm"] < tol["netForceAbsolute"]


def run_pair(vpd, ref, gpu_path, ref_path):
    rc_gpu = vpd.run_case_json(gpu_path, ref.REF_DIR)
    rc_ref = ref.ref().ref_run_case_json(os.fsencode(ref_path), os.fsencode(ref.REF_DIR))
    assert rc_gpu == rc_ref
    return rc_gpu


STRESS_COLS = ("avg_raw_stress_xx", "avg_raw_stress_xy", "elastic_reference_stress_xx",
               "elastic_reference_stress_xy")
FIELD_STRESS = ("raw_stress_xx", "force_x")


@pytest.mark.parametrize("name", ["affine_bar_elastic", "affine_shear_elastic"])
def test_affine_examples_vs_reference(vpd, ref, tmp_path, name):
    banner("affine demo case %s vs the reference runner" % name)
    g = example(name, str(tmp_path), str(tmp_path / ("gpu_" + name)))
    r = example(name, str(tmp_path), str(tmp_path / ("ref_" + name)))
    assert run_pair(vpd, ref, g, r) == 0
    dg, dr = str(tmp_path / ("gpu_" + name)), str(tmp_path / ("ref_" + name))
    compare_csv(os.path.join(dg, "timeseries.csv"), os.path.join(dr, "timeseries.csv"), 2,
                STRESS_COLS)
    compare_csv(os.path.join(dg, "field_final.csv"), os.path.join(dr, "field_final.csv"), 2,
                FIELD_STRESS)
    compare_metadata(os.path.join(dg, "metadata.json"), os.path.join(dr, "metadata.json"))
    meta = json.load(open(os.path.join(dg, "metadata.json")))
    # The closed form the CTest checks: |σ − σ_elastic| within 0.002 MPa at every point.
    assert meta["maxReferenceStressError"] < 0.002


@pytest.mark.parametrize("kind", ["jc_plastic", "jc_damage"])
def test_demo_host_johnson_cook_vs_reference(vpd, ref, tmp_path, kind):
    banner("affine demo case, Johnson-Cook %s, vs the reference runner" % kind)
    kw = dict(nx=7, ny=4, nz=4, spacing=1.0, horizon=2.015, steps=40, dt=1e-5,
              max_strain=0.02, material=cf.JC_FILE, reference=False,
              equivalentPlasticStrainIndex=1, internalIndicator={"indices": [1], "reduction": "max"})
    if kind == "jc_damage":
        kw["damage"] = {"model": "critical_stretch", "initiation": 0.012, "failure": 0.0,
                        "deletionEnabled": False}
    g = cf.write_demo_case(str(tmp_path / "gpu"), name=kind, **kw)
    r = cf.write_demo_case(str(tmp_path / "ref"), name=kind, **kw)
    rc = run_pair(vpd, ref, g, r)
    dg, dr = cf.output_dir(g), cf.output_dir(r)
    compare_csv(os.path.join(dg, "timeseries.csv"), os.path.join(dr, "timeseries.csv"), 2,
                STRESS_COLS)
    compare_csv(os.path.join(dg, "field_final.csv"), os.path.join(dr, "field_final.csv"), 2,
                FIELD_STRESS)
    ha, rows = cf.read_csv(os.path.join(dg, "timeseries.csv"))
    x = cf.numeric(rows, 2)
    ind = x[:, ha.index("max_internal_indicator") - 2]
    assert ind.max() > 0.001  # plastic flow reached
    if kind == "jc_damage":
        assert x[:, ha.index("max_damage") - 2].max() > 0  # bonds broke
    print("exit", rc)
