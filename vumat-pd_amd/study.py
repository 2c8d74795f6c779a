"""Fixed-space PD time-step study: one regional boundary case at three time steps, run as one
batch on the GPU and compared coarse against fine.

The study is that of benchmarks/tensile3d/time_step_study.py: the base case at dt×2, dt×1 and
dt×½ over the same duration, with the output cadence scaled so that every variant writes at
the same physical times. Here the three variants are three specimens of one particle block
(vpd_run_cases_json): one process, one context, one step loop over the slowest variant's
steps, each specimen stopping at its own last step. Each variant's folder holds the
reference's outputs (timeseries.csv, field_snapshots.csv, metadata.json), byte-equal to what
the variant gives when run alone.

study.json carries the reference study's record: status, scope, SHA-256 of the base case, the
material, this library's sources and binary, the continuum wave-speed screen, one entry per
variant and the two coarse→fine comparisons (stress-history NRMSE at shared output times,
final-displacement RMSE over the load amplitude, peak-ε̄ᵖ difference and where each peak is).

    python study.py --case path/to/pd_coarse.json --output NEW_DIR [--repo DIR]

Relative material paths in the case resolve from --repo (the checkout the case file belongs
to), as the reference resolves them from its own root.
"""
from __future__ import annotations

import argparse
import csv
import hashlib
import json
import math
import os
import sys
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
SCOPE = ("Fixed-space PD time-step study, the three time steps batched on one GPU. It measures "
         "temporal sensitivity, not a spectral stability limit or a validation of physical "
         "heating; the material configuration is kept byte-identical.")
# (label, dt factor); the factors are exact binary fractions, so steps/factor and
# outputEvery/factor are exact integers for even base values.
VARIANTS = (("dt_x2", 2.0), ("dt_x1", 1.0), ("dt_x0p5", 0.5))
TIME_RTOL, TIME_ATOL = 1e-10, 1e-14  # two output rows are the same physical time


def file_sha256(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 20), b""):
            h.update(chunk)
    return h.hexdigest()


def case_problem(case):
    """Why the base case cannot be studied, or None. The three variants need one material, an
    even positive step count and an even output cadence of at least 2 (halving them stays
    integral), totalTime = steps·dt, and a finite non-zero load amplitude (the displacement
    error is quoted relative to it)."""
    if case.get("type") != "vumat_pd_boundary_case":
        return "The study takes a vumat_pd_boundary_case"
    if len(case.get("materials", ())) != 1:
        return "The study takes a case with exactly one material"
    ti, every = case["time_integration"], case["dynamics"]["outputEvery"]
    steps, dt = ti["steps"], ti["dt"]
    if not (isinstance(dt, (int, float)) and math.isfinite(dt) and dt > 0):
        return "dt must be positive and finite"
    for name, value, least in (("steps", steps, 2), ("outputEvery", every, 2)):
        if not isinstance(value, int) or value < least or value % 2:
            return "%s must be an even integer of at least %d" % (name, least)
    if every > steps:
        return "outputEvery cannot exceed steps"
    if not math.isclose(steps * dt, ti["totalTime"], rel_tol=1e-12):
        return "totalTime must equal steps * dt"
    amp = case["boundary_conditions"]["amplitude"]
    if not (isinstance(amp, (int, float)) and math.isfinite(amp) and amp != 0):
        return "The load amplitude must be finite and non-zero"
    return None


def variant_case(case, label, factor, material, folder):
    """The base case at dt·factor over the same duration, its output cadence scaled with it
    and one field snapshot at the end."""
    ti = case["time_integration"]
    steps = int(ti["steps"] / factor)
    out = json.loads(json.dumps(case))  # a deep copy of the JSON tree
    out["name"] = "%s_%s" % (case["name"], label)
    out["time_integration"].update(dt=ti["dt"] * factor, steps=steps,
                                   totalTime=ti["steps"] * ti["dt"])
    out["dynamics"].update(outputEvery=int(case["dynamics"]["outputEvery"] / factor),
                           snapshotEvery=steps)
    out["materials"][0]["materialConfigPath"] = Path(material).as_posix()
    out["output"]["directory"] = (Path(folder) / label).as_posix()
    return out


def read_columns(path):
    """A CSV file as {column: array of its text}."""
    with open(path, newline="", encoding="utf-8") as f:
        reader = csv.reader(f)
        header = next(reader)
        rows = list(reader)
    table = np.array(rows, dtype=object).reshape(len(rows), len(header))
    return {name: table[:, k] for k, name in enumerate(header)}


@dataclass
class VariantRun:
    """What the comparisons read from one variant's output folder."""
    dt_ms: float
    time: np.ndarray     # output times of timeseries.csv (ms)
    stress: np.ndarray   # average axial stress at those times (MPa)
    finite: np.ndarray   # the row's finite flag
    particle: np.ndarray # field_snapshots.csv particle ids (text), in file order
    x0: np.ndarray       # (n, 3) reference positions (mm)
    u: np.ndarray        # (n, 3) displacements (mm)
    eqps: np.ndarray     # (n,)
    axial: np.ndarray    # (n,) axial stress (MPa)

    @classmethod
    def load(cls, folder, dt_ms):
        ts = read_columns(Path(folder) / "timeseries.csv")
        fs = read_columns(Path(folder) / "field_snapshots.csv")
        num = lambda col: col.astype(float)
        return cls(dt_ms=dt_ms, time=num(ts["time_ms"]), stress=num(ts["average_axial_stress_MPa"]),
                   finite=ts["finite"] == "1", particle=fs["particle"],
                   x0=np.stack([num(fs[k]) for k in ("x0_mm", "y0_mm", "z0_mm")], 1),
                   u=np.stack([num(fs[k]) for k in ("ux_mm", "uy_mm", "uz_mm")], 1),
                   eqps=num(fs["eqps"]), axial=num(fs["axial_stress_MPa"]))

    def complete(self, particles, duration):
        """Every particle in the snapshot, the history reaching the end time, all finite."""
        last_ok = self.time.size > 0 and math.isclose(self.time[-1], duration, rel_tol=TIME_RTOL)
        fields = np.concatenate([self.u.ravel(), self.eqps, self.axial])
        return (self.particle.size == particles and last_ok and bool(self.finite.all()) and
                bool(np.isfinite(fields).all()))


def shared_times(coarse, fine):
    """For each coarse output time, the first fine row at the same physical time (symmetric
    relative tolerance TIME_RTOL, absolute TIME_ATOL), or -1. Rows are never interpolated."""
    a, b = coarse.time[:, None], fine.time[None, :]
    same = np.abs(a - b) <= np.maximum(TIME_RTOL * np.maximum(np.abs(a), np.abs(b)), TIME_ATOL)
    return np.where(same.any(1), same.argmax(1), -1)


def compare(coarse, fine, amplitude):
    """One coarse→fine comparison of the reference study (time_step_study.py:29-66 defines
    the metrics): the RMS of the stress-history difference at the shared output times over
    the fine run's peak |stress| there; the RMS of the final displacement difference per
    component over |amplitude|; the peak-ε̄ᵖ difference relative to the fine peak; and the
    reference position of each run's first ε̄ᵖ maximum in file order."""
    match = shared_times(coarse, fine)
    rows = np.nonzero(match >= 0)[0]
    if rows.size < 2:
        raise ValueError("Fewer than two common history times")
    s_c, s_f = coarse.stress[rows], fine.stress[match[rows]]
    peak = np.abs(s_f).max()
    if peak == 0:
        raise ValueError("Zero reference stress history")
    # Pair the snapshot rows by particle id (the two files may list them in any order).
    oc, of = np.argsort(coarse.particle, kind="stable"), np.argsort(fine.particle, kind="stable")
    if coarse.particle.size != fine.particle.size or \
            not np.array_equal(coarse.particle[oc], fine.particle[of]):
        raise ValueError("Final field points differ")
    if not np.array_equal(coarse.x0[oc], fine.x0[of]):
        raise ValueError("Reference geometry changed during time refinement")
    du = coarse.u[oc] - fine.u[of]
    kc, kf = int(np.argmax(coarse.eqps)), int(np.argmax(fine.eqps))
    e_c, e_f = float(coarse.eqps[kc]), float(fine.eqps[kf])
    return {
        "coarser_dt_ms": coarse.dt_ms, "finer_dt_ms": fine.dt_ms,
        "common_history_times": int(rows.size),
        "stress_history_nrmse_percent": float(100 * np.sqrt(np.mean((s_c - s_f) ** 2)) / peak),
        "displacement_rmse_percent_of_amplitude":
            float(100 * np.sqrt(np.sum(du * du) / du.size) / abs(amplitude)),
        "max_eqps_relative_difference_percent": 100 * abs(e_c - e_f) / e_f if e_f else None,
        "coarser_peak_eqps_reference_xyz_mm": [float(v) for v in coarse.x0[kc]],
        "finer_peak_eqps_reference_xyz_mm": [float(v) for v in fine.x0[kf]],
    }


def continuum_screen(case):
    """The P-wave speed of the configured E, ν and ρ and the base dt against h/c_p; None when
    the material entry gives no usable E, ν."""
    entry = case["materials"][0]
    E, nu = entry.get("elasticE", 0), entry.get("elasticNu", 0)
    rho = case["dynamics"]["density"]
    if not (E > 0 and rho > 0 and -1 < nu < 0.5):
        return None
    c = math.sqrt(E * (1 - nu) / (rho * (1 + nu) * (1 - 2 * nu)))
    h_over_c = case["geometry"]["spacing"] / c
    return {"wave_speed_mm_per_ms": c, "h_over_cp_ms": h_over_c,
            "base_dt_over_h_over_cp": case["time_integration"]["dt"] / h_over_c,
            "scope": "Continuum screening estimate using configured E and nu; NOT a PD "
                     "stability bound"}


class _StdoutTo:
    """Sends this process's stdout and stderr (the library prints to them) to a file."""

    def __init__(self, path):
        self.path = path

    def __enter__(self):
        sys.stdout.flush()
        sys.stderr.flush()
        self.log = open(self.path, "w")
        self.saved = [os.dup(1), os.dup(2)]
        os.dup2(self.log.fileno(), 1)
        os.dup2(self.log.fileno(), 2)
        return self

    def __exit__(self, *exc):
        for fd, saved in zip((1, 2), self.saved):
            os.dup2(saved, fd)
            os.close(saved)
        self.log.close()
        return False


def library_files():
    lib = HERE / "libvpd.so"
    sources = [p for folder in (ROOT / "include", HERE / "csrc", HERE / "include")
               for p in sorted(folder.rglob("*")) if p.is_file()]
    return lib, sources


def run_study(case_path, output, repo=None):
    import vpd  # the GPU library; raises if it is not built

    case_path = Path(case_path).resolve()
    repo = Path(repo).resolve() if repo else case_path.parent
    case = json.loads(case_path.read_text())
    problem = case_problem(case)
    if problem:
        raise ValueError(problem)
    material = (repo / case["materials"][0]["materialConfigPath"]).resolve()
    lib, sources = library_files()
    if not material.is_file():
        raise FileNotFoundError("No material configuration at %s" % material)
    output = Path(output).resolve()
    output.mkdir(parents=True, exist_ok=False)  # a study never reuses a folder
    (output / "base-case.json").write_bytes(case_path.read_bytes())
    (output / "material.json").write_bytes(material.read_bytes())
    report = {"status": "running", "scope": SCOPE, "base_case_sha256": file_sha256(case_path),
              "material_sha256": file_sha256(material),
              "source_sha256": {p.relative_to(ROOT).as_posix(): file_sha256(p) for p in sources},
              "binary_sha256": {lib.name: file_sha256(lib)}, "runs": [], "comparisons": []}
    screen = continuum_screen(case)
    if screen:
        report["continuum_screen"] = screen

    def save():
        (output / "study.json").write_text(json.dumps(report, indent=2, allow_nan=False) + "\n",
                                           encoding="utf-8")

    duration = case["time_integration"]["steps"] * case["time_integration"]["dt"]
    variants = [variant_case(case, label, f, material, output) for label, f in VARIANTS]
    paths = []
    for (label, _), v in zip(VARIANTS, variants):
        paths.append(output / (label + ".json"))
        paths[-1].write_text(json.dumps(v, indent=2) + "\n")
    save()
    try:
        print("START batch of %s" % ", ".join(label for label, _ in VARIANTS), flush=True)
        t0 = time.monotonic()
        cwd = os.getcwd()
        os.chdir(repo)
        try:
            with _StdoutTo(output / "batch.log"):
                status = vpd.run_cases_json([str(p) for p in paths])
        finally:
            os.chdir(cwd)
        elapsed = time.monotonic() - t0
        runs = []
        for (label, _), v in zip(VARIANTS, variants):
            ti = v["time_integration"]
            report["runs"].append({"label": label, "dt_ms": ti["dt"], "steps": ti["steps"],
                                   "returncode": status, "threads": 1, "device": "gpu",
                                   "batch": [lb for lb, _ in VARIANTS],
                                   "elapsed_seconds": elapsed})
        save()
        if status:
            raise RuntimeError("The batch failed (status %d); inspect batch.log" % status)
        for (label, _), run, v in zip(VARIANTS, report["runs"], variants):
            dest = output / label
            meta = json.loads((dest / "metadata.json").read_text())
            got = VariantRun.load(dest, v["time_integration"]["dt"])
            if meta["status"] != "pass" or not got.complete(meta["particles"], duration):
                raise RuntimeError("%s did not produce complete finite results" % label)
            run["diagnostics"] = meta["materials"][0]
            run["result_sha256"] = {p.name: file_sha256(p) for p in sorted(dest.iterdir())
                                    if p.is_file()}
            runs.append(got)
        print("END batch: %.1fs" % elapsed, flush=True)
        amp = case["boundary_conditions"]["amplitude"]
        report["comparisons"] = [compare(runs[k], runs[k + 1], amp) for k in range(2)]
        if file_sha256(material) != report["material_sha256"] or \
                file_sha256(lib) != report["binary_sha256"][lib.name]:
            raise RuntimeError("The material or the library changed during the study")
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
    ap.add_argument("--output", type=Path, required=True, help="must not exist yet")
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
