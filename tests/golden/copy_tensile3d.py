"""Copies the reference's tensile-benchmark data used as fixtures by the case-runner tests:
the seven PD case files of benchmarks/tensile3d/cases/ (the set run_cases.py runs) into
tests/golden/cases/, and the archived outputs of pd_shared and pd_delta2
(benchmarks/tensile3d/raw/<case>/: timeseries.csv.gz, field_snapshots.csv.gz, metadata.json)
with their case files into tests/golden/<case>/. Data only; run from the repository root
where the reference is present:

    python tests/golden/copy_tensile3d.py [/root/reference]
"""
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = ["pd_coarse", "pd_shared", "pd_m1p5", "pd_delta4", "pd_delta2", "pd_dt50", "pd_dt12p5"]
ARCHIVED = ["pd_shared", "pd_delta2"]


def main(ref):
    bench = os.path.join(ref, "benchmarks", "tensile3d")
    os.makedirs(os.path.join(HERE, "cases"), exist_ok=True)
    for name in CASES:
        shutil.copyfile(os.path.join(bench, "cases", name + ".json"),
                        os.path.join(HERE, "cases", name + ".json"))
    for name in ARCHIVED:
        dest = os.path.join(HERE, name)
        os.makedirs(dest, exist_ok=True)
        shutil.copyfile(os.path.join(bench, "cases", name + ".json"),
                        os.path.join(dest, "case.json"))
        for f in ("timeseries.csv.gz", "field_snapshots.csv.gz", "metadata.json"):
            shutil.copyfile(os.path.join(bench, "raw", name, f), os.path.join(dest, f))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
