"""Writers of small case and material JSON files in the reference's schema
(src/config/DemoConfig.cpp, src/config/MaterialConfigJson.cpp) for the case-runner tests,
and readers of the runner's CSV outputs."""
import csv
import gzip
import json
import os

import numpy as np

# configs/johnson_cook.json of the reference: E, ν, A, B, C, n, m, ε̇₀, T_ref, T_melt, ρ, β, c_p,
# workToHeat.
JC_FILE = {
    "schemaVersion": "1.0", "name": "JC", "cmname": "JOHNSONCOOK", "plugin": "jc_vumat_plugin",
    "symbol": "vumat_adapter_call", "ndir": 3, "nshr": 3, "nstatev": 6, "nfieldv": 1,
    "props": [210000.0, 0.30, 350.0, 500.0, 0.025, 0.36, 1.0, 1.0, 293.0, 1793.0, 7800.0, 0.9,
              500.0, 1.0],
    "density": 7800.0,
}


def write_case(directory, name="case", nx=21, ny=7, nz=3, spacing=2.0, horizon=6.0, steps=200,
               dt=2.5e-5, amplitude=2.0, rise_time=0.12, layers=2, output_every=10,
               snapshot_every=100, damage=None, component="x", material=None,
               observation=(2.0, None), plane_strain=False, **material_entry):
    """Writes <directory>/<name>.json and its material file; returns the case path. Output goes
    to <directory>/<name>_out."""
    os.makedirs(directory, exist_ok=True)
    mat_path = os.path.join(directory, name + "_material.json")
    with open(mat_path, "w") as f:
        json.dump(material or JC_FILE, f)
    support, load = ("left", "right") if component == "x" else ("bottom", "top")
    case = {
        "schemaVersion": "1.0", "name": name, "type": "vumat_pd_boundary_case",
        "classification": "verification", "scope": "test",
        "units": {"length": "mm", "time": "ms", "force": "N", "stress": "MPa",
                  "massDensity": "N*ms^2/mm^4", "sourceDensity": "kg/m^3"},
        "geometry": {"type": "patch", "nx": nx, "ny": ny, "nz": nz, "spacing": spacing,
                     "horizon": horizon},
        "time_integration": {"steps": steps, "dt": dt, "totalTime": steps * dt},
        "boundary_conditions": {"type": "regional_smooth_displacement",
                                "mode": "boundary_value", "supportRegion": support,
                                "loadRegion": load, "component": component,
                                "regionLayers": layers, "amplitude": amplitude,
                                "riseTime": rise_time, "planeStrain": plane_strain,
                                "rigidBodySuppression": True},
        "dynamics": {"density": 0.0078, "energyToleranceRelative": 0.15,
                     "netForceTolerance": 1e-6, "outputEvery": output_every,
                     "snapshotEvery": snapshot_every, "minimumPlasticStrain": 0.0,
                     "kineticToInternalLimit": 1e6},
        "output": {"directory": os.path.join(directory, name + "_out")},
        "materials": [dict({"name": "JC", "materialConfigPath": mat_path,
                            "reference": True, "elasticE": 210000, "elasticNu": 0.3,
                            "referenceTemperature": 293,
                            "equivalentPlasticStrainIndex": 1}, **material_entry)],
    }
    lo, hi = observation
    obs = {}
    if lo is not None:
        obs["axialMin"] = lo
    if hi is not None:
        obs["axialMax"] = hi
    if obs:
        case["observation"] = obs
    if damage is not None:
        case["damage"] = damage
    path = os.path.join(directory, name + ".json")
    with open(path, "w") as f:
        json.dump(case, f, indent=1)
    return path


ELASTIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "affine",
                            "elastic.json")


def write_demo_case(directory, name="demo", nx=5, ny=3, nz=3, spacing=1.0, horizon=1.75,
                    steps=20, dt=1e-6, mode="uniaxial_xx", max_strain=1e-4, material=None,
                    damage=None, tolerances=None, reference=True, **material_entry):
    """Writes an affine demo case (the schema of examples/affine_bar_elastic.json) and returns
    its path; material None uses the elastic fixture, a dict is written as the material file."""
    os.makedirs(directory, exist_ok=True)
    if material is None:
        mat_path = ELASTIC_FILE
    else:
        mat_path = os.path.join(directory, name + "_material.json")
        with open(mat_path, "w") as f:
            json.dump(material, f)
    case = {
        "schemaVersion": "1.0", "name": name, "type": "vumat_pd_demo_case",
        "scope": "test affine case",
        "geometry": {"type": "patch", "nx": nx, "ny": ny, "nz": nz, "spacing": spacing,
                     "horizon": horizon},
        "time_integration": {"steps": steps, "dt": dt, "totalTime": steps * dt},
        "boundary_conditions": {"type": "affine_displacement", "mode": mode,
                                "maxStrain": max_strain},
        "tolerances": tolerances or {"deformationGradientAbsolute": 1e-12,
                                     "referenceStressAbsolute": 0.002,
                                     "netForceAbsolute": 1e-7},
        "output": {"directory": os.path.join(directory, name + "_out")},
        "materials": [dict({"name": "M", "materialConfigPath": mat_path,
                            "kinematicsMode": "corotational", "reference": reference,
                            "elasticE": 210000.0, "elasticNu": 0.3,
                            "referenceTemperature": 293.0}, **material_entry)],
        "checks": ["affine deformation-gradient recovery", "global force balance"],
        "damage": damage or {"model": "none", "initiation": 1e30, "failure": 1e30,
                             "deletionEnabled": False},
    }
    path = os.path.join(directory, name + ".json")
    with open(path, "w") as f:
        json.dump(case, f, indent=1)
    return path


def output_dir(case_path):
    with open(case_path) as f:
        return json.load(f)["output"]["directory"]


def read_csv(path):
    """(header, rows of strings) of a CSV file, gzip-aware."""
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "rt", newline="") as f:
        rows = list(csv.reader(f))
    return rows[0], rows[1:]


def numeric(rows, first):
    """Columns first.. of the rows as float64 (the runner writes nan/inf as printf does)."""
    return np.array([[float(v) for v in r[first:]] for r in rows], dtype=np.float64)
