"""Host side of the GPU case runner, on the CPU: the case/material JSON validation ends in the
reference's error messages and exit status before any device work, region selection matches
the oracle, and (where oracle/_ref is built) each rejected file is rejected by the reference's
own runner with the same message."""
import json
import os

import numpy as np
import pytest

import case_files as cf


def run_both(vpd, path, capfd, ref=None):
    rc = vpd.run_case_json(path, "")
    err = capfd.readouterr().err
    out = (rc, err.strip())
    if ref is not None:
        rrc = ref.ref().ref_run_case_json(os.fsencode(path), b"")
        rerr = capfd.readouterr().err
        assert (rrc, rerr.strip()) == out
    return out


def mutate(path, fn):
    with open(path) as f:
        case = json.load(f)
    fn(case)
    with open(path, "w") as f:
        json.dump(case, f)


BAD_CASES = [
    ("wrong type", lambda c: c.update(type="vumat_pd_nonsense"),
     "Unsupported case type: vumat_pd_nonsense"),
    ("schema", lambda c: c.update(schemaVersion="2.0"),
     "Unsupported VUMAT-PD case schemaVersion: 2.0"),
    ("no materials", lambda c: c.pop("materials"), "Demo case requires a materials array."),
    ("total time", lambda c: c["time_integration"].update(totalTime=1.0),
     "time_integration.totalTime must equal steps * dt."),
    ("units", lambda c: c["units"].update(length="m"),
     "Boundary-value cases currently require the documented mm-ms-N-MPa unit system."),
    ("thin", lambda c: c["geometry"].update(nz=2),
     "PD boundary cases require nz>=3 and horizon/spacing>=1.5."),
    ("loading", lambda c: c["boundary_conditions"].update(component="y"),
     "Regional boundary case supports left/right-x or bottom/top-y loading."),
    ("rigid", lambda c: c["boundary_conditions"].update(rigidBodySuppression=False),
     "Boundary-value case requires regional smooth loading and rigid-body suppression."),
    ("dynamics", lambda c: c["dynamics"].update(outputEvery=0),
     "Invalid boundary-value dynamics configuration."),
    ("integer", lambda c: c["geometry"].update(nx=2.5),
     "Expected representable integer field: nx"),
    ("eqps index", lambda c: c["materials"][0].update(equivalentPlasticStrainIndex=6),
     "equivalentPlasticStrainIndex is outside nstatev for material: JC"),
    ("indicator", lambda c: c["materials"][0].update(internalIndicator={"reduction": "mean"}),
     "internalIndicator.reduction must be min or max."),
    ("observation", lambda c: c.update(observation={"axialMin": 3.0}),
     "Observation bounds must coincide with reference nodal planes."),
]


@pytest.mark.parametrize("name,fn,message", BAD_CASES, ids=[b[0] for b in BAD_CASES])
def test_rejected_cases_match_reference(tmp_path, capfd, orc, name, fn, message):
    import vpd
    path = cf.write_case(str(tmp_path), steps=4)
    mutate(path, fn)
    ref = orc if orc.ref_available() else None
    rc, err = run_both(vpd, path, capfd, ref)
    assert rc == 1
    assert err == "VUMAT-PD demo case failed: " + message


@pytest.mark.parametrize("text", ['{"name": "x", "steps": 1,}', '{"name": "x" "steps": 1}',
                                  '{"name": "x", "steps": 01.}', '{"name": "\\q"}',
                                  '{"name": "x"} trailing', '[1, 2', '{"name": 1e999}'])
def test_malformed_json_matches_reference(tmp_path, capfd, orc, text):
    import vpd
    bad = tmp_path / "bad.json"
    bad.write_text(text)
    rc, err = run_both(vpd, str(bad), capfd, orc if orc.ref_available() else None)
    assert rc == 1 and err.startswith("VUMAT-PD demo case failed: ")


def test_missing_file(tmp_path, capfd):
    import vpd
    assert vpd.run_case_json(str(tmp_path / "none.json")) == 1
    assert "Failed to open JSON file" in capfd.readouterr().err


def test_non_jc_material_is_refused(tmp_path, capfd):
    import vpd
    elastic = dict(cf.JC_FILE, name="Elastic", cmname="ELASTIC", plugin="elastic_vumat_plugin",
                   nstatev=1, props=[210000.0, 0.3])
    path = cf.write_case(str(tmp_path), steps=4, material=elastic,
                         equivalentPlasticStrainIndex=-1)
    assert vpd.run_case_json(path) == 1
    assert "Only the Johnson-Cook VUMAT runs on the device" in capfd.readouterr().err


@pytest.mark.parametrize("side,width,planes", [("left", 2.0, 3), ("right", 1e-9, 1),
                                               ("bottom", 4.0, 5), ("top", 0.5, 1),
                                               ("left", 2.0 - 1e-12, 3)])
def test_select_boundary_region_closed_form(orc, side, width, planes):
    # On a 9×5×3 lattice of spacing 1 shifted to x ∈ [3, 11], a band of width w holds the
    # ⌊w⌋+1 lattice planes nearest the side (w within the 1e-10 tolerance of an integer
    # counts as that integer); the oracle agrees.
    import vpd
    x, _ = vpd.make_patch(9, 5, 3, 1.0)
    x = x + np.array([3.0, 0.0, 0.0])
    mask = vpd.select_boundary_region(x, side, width)
    comp = 0 if side in ("left", "right") else 1
    vals = np.unique(x[:, comp])
    chosen = vals[:planes] if side in ("left", "bottom") else vals[-planes:]
    assert np.array_equal(mask, np.isin(x[:, comp], chosen))
    ref = np.zeros(x.shape[0], np.uint8)
    code = {"left": 0, "right": 1, "bottom": 2, "top": 3}[side]
    assert orc.lib().orc_select_region(x, x.shape[0], code, width, ref) == 0
    assert np.array_equal(mask, ref.astype(bool))


def test_select_boundary_region_errors():
    import vpd
    x, _ = vpd.make_patch(3, 3, 3, 1.0)
    with pytest.raises(vpd.VpdError):
        vpd.select_boundary_region(x, "front", 1.0)
    with pytest.raises(vpd.VpdError):
        vpd.select_boundary_region(x, "left", 0.0)


DEMO_BAD = [
    ("mode", lambda c: c["boundary_conditions"].update(mode="biaxial"),
     "Unsupported boundary_conditions type or mode."),
    ("bc type", lambda c: c["boundary_conditions"].update(type="regional_smooth_displacement"),
     "Unsupported boundary_conditions type or mode."),
    ("kinematics", lambda c: c["materials"][0].update(kinematicsMode="legacy_small_strain"),
     "Unsupported kinematicsMode for material: M"),
    ("no materials", lambda c: c.update(materials=[]),
     "Demo case requires at least one material."),
    ("zero strain", lambda c: c["boundary_conditions"].update(maxStrain=0.0),
     "Affine displacement boundary condition requires maxStrain."),
]


@pytest.mark.parametrize("name,fn,message", DEMO_BAD, ids=[b[0] for b in DEMO_BAD])
def test_rejected_demo_cases_match_reference(tmp_path, capfd, orc, name, fn, message):
    """Affine demo cases (src/host/DemoCaseRunner.cpp) end in the reference's messages; a zero
    maxStrain fails after the output files and their headers exist, as there."""
    import vpd
    path = cf.write_demo_case(str(tmp_path))
    mutate(path, fn)
    ref = orc if orc.ref_available() else None
    rc, err = run_both(vpd, path, capfd, ref)
    assert rc == 1
    assert err == "VUMAT-PD demo case failed: " + message
    if name == "zero strain":
        out = cf.output_dir(path)
        header = open(os.path.join(out, "timeseries.csv")).readline()
        assert header.startswith("material,step,time,applied_strain,max_f_error")
        assert open(os.path.join(out, "field_final.csv")).readline().startswith(
            "material,particle,x0,y0,z0")


def test_material_point_cases_are_refused(tmp_path, capfd):
    import vpd
    path = cf.write_demo_case(str(tmp_path))
    mutate(path, lambda c: c.update(type="vumat_material_point_case"))
    rc = vpd.run_case_json(path, "")
    err = capfd.readouterr().err.strip()
    assert rc == 1 and "vumat_material_point_case" in err



def test_command_line_runner(tmp_path):
    # vpd_pd_run: the reference's usage (src/apps/main_pd_case_runner.cpp:15-27), its default
    # case when called without arguments, one case per call, several cases as one batch; every
    # failure here happens while reading the files, before any device work.
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "vumat-pd_amd", "build", "vpd_pd_run")
    if not os.path.exists(exe):
        pytest.skip("vumat-pd_amd/build/vpd_pd_run is not built")
    r = subprocess.run([exe, "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.startswith("Usage: vpd_pd_run [case.json")
    assert "default: examples/affine_bar_elastic.json" in r.stdout
    r = subprocess.run([exe], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 1
    assert r.stderr.strip() == ("VUMAT-PD demo case failed: Failed to open JSON file: "
                                "examples/affine_bar_elastic.json")
    missing = str(tmp_path / "missing.json")
    for args in ([missing], [missing, str(tmp_path / "other.json")]):
        r = subprocess.run([exe] + args, capture_output=True, text=True)
        assert r.returncode == 1
        assert r.stderr.strip() == "VUMAT-PD demo case failed: Failed to open JSON file: " + missing
