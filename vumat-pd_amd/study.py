"""Fixed-space PD time-step study on the GPU case runner, written to a new directory.

The reference's study (benchmarks/tensile3d/time_step_study.py) runs one regional PD case at
dt×2, dt×1 and dt×0.5 through its CPU runner and writes `study.json`; this is the same study
with `vpd_pd_run` (the GPU runner) in place of `vumat_pd_run`. The output directory holds the
same files (base-case.json, material.json, one case file, log and output folder per run,
study.json) and study.json has the same keys: status, scope, the SHA-256 of the base case,
the material, the sources and the binaries, the continuum wave-speed screen, one record per
run (with its diagnostics and output hashes) and the two coarse/fine comparisons
(time_step_study.py:28-66).

    python study.py --case path/to/pd_coarse.json --output NEW_DIR [--repo DIR]

Relative material paths in the case resolve from --repo (the reference checkout the case
file belongs to), as the reference resolves them from its own root.
"""
from __future__ import annotations

import argparse
import copy
import csv
import hashlib
import json
import math
import os
import subprocess
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
RUNNER = HERE / "build" / "vpd_pd_run"
LIBRARY = HERE / "libvpd.so"
SCOPE = ("Fixed-space PD time-step study on the GPU runner. This measures temporal "
         "sensitivity, not a spectral stability limit or validation of physical heating. The "
         "material configuration is kept byte-identical.")
VARIANTS = (("dt_x2", 2), ("dt_x1", 1), ("dt_x0p5", 0.5))


def sha256(path):
    return hashlib.sha256(Path(path).read_bytes()).hexdigest()


def csv_rows(path):
    with open(path, newline="", encoding="utf-8") as f:
        return list(csv.DictReader(f))


def same_time(a, b):
    return math.isclose(float(a), float(b), rel_tol=1e-10, abs_tol=1e-14)


def compare(coarse, fine, amplitude):
    """Coarse vs fine run: stress-history NRMSE over the output times both runs share (no
    interpolation), final-displacement RMSE relative to the load amplitude, peak-ε̄ᵖ
    difference and the peak locations (the metrics of time_step_study.py:28-66)."""
    pairs = []
    for row in coarse["history"]:
        hit = next((r for r in fine["history"] if same_time(r["time_ms"], row["time_ms"])), None)
        if hit is not None:
            pairs.append((row, hit))
    if len(pairs) < 2:
        raise ValueError("Fewer than two common history times")
    stress = [(float(a["average_axial_stress_MPa"]), float(b["average_axial_stress_MPa"]))
              for a, b in pairs]
    peak = max(abs(a) for a, _ in stress)
    if peak == 0:
        raise ValueError("Zero reference stress history")
    nrmse = 100 * math.sqrt(sum((a - b) ** 2 for a, b in stress) / len(stress)) / peak
    left, right = coarse["fields"], fine["fields"]
    if left.keys() != right.keys():
        raise ValueError("Final field points differ")
    ref_keys, u_keys = ("x0_mm", "y0_mm", "z0_mm"), ("ux_mm", "uy_mm", "uz_mm")
    sq = 0.0
    for point, a in left.items():
        b = right[point]
        if any(float(a[k]) != float(b[k]) for k in ref_keys):
            raise ValueError("Reference geometry changed during time refinement")
        sq += sum((float(a[k]) - float(b[k])) ** 2 for k in u_keys)
    top_a = max(left.values(), key=lambda r: float(r["eqps"]))
    top_b = max(right.values(), key=lambda r: float(r["eqps"]))
    e_a, e_b = float(top_a["eqps"]), float(top_b["eqps"])
    return {
        "coarser_dt_ms": coarse["dt_ms"], "finer_dt_ms": fine["dt_ms"],
        "common_history_times": len(pairs),
        "stress_history_nrmse_percent": nrmse,
        "displacement_rmse_percent_of_amplitude":
            100 * math.sqrt(sq / (3 * len(left))) / abs(amplitude),
        "max_eqps_relative_difference_percent": 100 * abs(e_a - e_b) / e_b if e_b else None,
        "coarser_peak_eqps_reference_xyz_mm": [float(top_a[k]) for k in ref_keys],
        "finer_peak_eqps_reference_xyz_mm": [float(top_b[k]) for k in ref_keys],
    }


def check_case(case):
    """The base-case requirements of the reference study; returns an error text or None."""
    ti, dyn = case["time_integration"], case["dynamics"]
    steps, every, dt = ti["steps"], dyn["outputEvery"], ti["dt"]
    if case["type"] != "vumat_pd_boundary_case" or len(case["materials"]) != 1:
        return "Requires positive threads and a single-material regional PD case"
    if (not isinstance(steps, int) or steps <= 0 or steps % 2 or not isinstance(every, int)
            or every < 2 or every % 2 or not math.isfinite(dt) or dt <= 0 or steps < every):
        return "Base steps and outputEvery must be positive even integers; dt must be positive"
    if not math.isclose(steps * dt, ti["totalTime"], rel_tol=1e-12):
        return "Base totalTime must equal steps * dt"
    amp = case["boundary_conditions"]["amplitude"]
    if not math.isfinite(amp) or amp == 0:
        return "Nonzero finite displacement amplitude is required"
    return None


def source_files():
    out = []
    for folder in (ROOT / "include", HERE / "csrc", HERE / "include"):
        out += [p for p in sorted(folder.rglob("*")) if p.is_file()]
    return out


def run_study(case_path, output, repo=None, runner=RUNNER):
    case_path = Path(case_path).resolve()
    repo = Path(repo).resolve() if repo else case_path.parent
    case = json.loads(case_path.read_text())
    problem = check_case(case)
    if problem:
        raise ValueError(problem)
    material = (repo / case["materials"][0]["materialConfigPath"]).resolve()
    if not runner.is_file() or not material.is_file():
        raise FileNotFoundError("Build the runner and supply an existing material configuration first")
    output = Path(output).resolve()
    output.mkdir(parents=True, exist_ok=False)  # never mix with an earlier study
    (output / "base-case.json").write_bytes(case_path.read_bytes())
    (output / "material.json").write_bytes(material.read_bytes())
    binaries = [runner, LIBRARY]
    report = {"status": "running", "scope": SCOPE, "base_case_sha256": sha256(case_path),
              "material_sha256": sha256(material),
              "source_sha256": {p.relative_to(ROOT).as_posix(): sha256(p) for p in source_files()},
              "binary_sha256": {p.name: sha256(p) for p in binaries}, "runs": [],
              "comparisons": []}
    ti = case["time_integration"]
    steps, dt, every = ti["steps"], ti["dt"], case["dynamics"]["outputEvery"]
    duration = steps * dt
    entry = case["materials"][0]
    E, nu = entry.get("elasticE", 0), entry.get("elasticNu", 0)
    rho = case["dynamics"]["density"]
    if E > 0 and rho > 0 and -1 < nu < 0.5:
        c = math.sqrt(E * (1 - nu) / (rho * (1 + nu) * (1 - 2 * nu)))
        h_over_c = case["geometry"]["spacing"] / c
        report["continuum_screen"] = {
            "wave_speed_mm_per_ms": c, "h_over_cp_ms": h_over_c,
            "base_dt_over_h_over_cp": dt / h_over_c,
            "scope": "Continuum screening estimate using configured E and nu; NOT a PD stability bound"}
    study_json = output / "study.json"

    def save():
        study_json.write_text(json.dumps(report, indent=2, allow_nan=False) + "\n",
                              encoding="utf-8")

    save()
    results = []
    try:
        for label, factor in VARIANTS:
            sel = copy.deepcopy(case)
            t = sel["time_integration"]
            t.update(dt=dt * factor, steps=round(steps / factor), totalTime=duration)
            sel["name"] = case["name"] + "_" + label
            sel["dynamics"].update(outputEvery=round(every / factor), snapshotEvery=t["steps"])
            sel["materials"][0]["materialConfigPath"] = material.as_posix()
            sel["output"]["directory"] = (output / label).as_posix()
            case_file = output / (label + ".json")
            case_file.write_text(json.dumps(sel, indent=2) + "\n")
            print("START %s: dt=%r ms" % (label, t["dt"]), flush=True)
            t0 = time.monotonic()
            with open(output / (label + ".log"), "w") as log:
                rc = subprocess.run([str(runner), str(case_file)], cwd=repo, stdout=log,
                                    stderr=subprocess.STDOUT).returncode
            run = {"label": label, "dt_ms": t["dt"], "steps": t["steps"], "returncode": rc,
                   "threads": 1, "device": "gpu", "elapsed_seconds": time.monotonic() - t0}
            report["runs"].append(run)
            save()
            if rc:
                raise RuntimeError("%s failed; inspect its log" % label)
            dest = output / label
            meta = json.loads((dest / "metadata.json").read_text())
            history = csv_rows(dest / "timeseries.csv")
            rows = csv_rows(dest / "field_snapshots.csv")
            complete = (meta["status"] == "pass" and len(rows) == meta["particles"] and
                        same_time(history[-1]["time_ms"], duration) and
                        all(r["finite"] == "1" for r in history) and
                        all(math.isfinite(float(r[k])) for r in rows
                            for k in ("ux_mm", "uy_mm", "uz_mm", "eqps", "axial_stress_MPa")))
            if not complete:
                raise RuntimeError("%s did not produce complete finite results" % label)
            run["diagnostics"] = meta["materials"][0]
            run["result_sha256"] = {p.name: sha256(p) for p in sorted(dest.iterdir())
                                    if p.is_file()}
            results.append({"dt_ms": t["dt"], "history": history,
                            "fields": {r["particle"]: r for r in rows}})
            print("END %s: passed in %.1fs" % (label, run["elapsed_seconds"]), flush=True)
        amp = case["boundary_conditions"]["amplitude"]
        report["comparisons"] = [compare(results[i], results[i + 1], amp) for i in range(2)]
        if (sha256(material) != report["material_sha256"] or
                any(sha256(p) != report["binary_sha256"][p.name] for p in binaries)):
            raise RuntimeError("Material or executable changed during study")
        report["status"] = "completed"
    except Exception as err:
        report.update(status="failed", error=str(err))
        raise
    finally:
        save()
    return report


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--case", type=Path, required=True)
    ap.add_argument("--output", type=Path, required=True, help="Must not already exist")
    ap.add_argument("--repo", type=Path, default=None,
                    help="root that relative paths in the case resolve from")
    a = ap.parse_args(argv)
    try:
        report = run_study(a.case, a.output, a.repo)
    except ValueError as err:
        ap.error(str(err))
    print(json.dumps(report["comparisons"], indent=2))


if __name__ == "__main__":
    main()
