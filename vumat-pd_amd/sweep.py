"""Grid / horizon / time-step sensitivity sweep of the Johnson-Cook tensile bar, batched on one
GPU. The reference runs such studies as one process per case (benchmarks/tensile3d/
run_cases.py:95-96, a thread pool of `vumat_pd_run` subprocesses); here every specimen of
the sweep sits in one particle block and each pass is one launch over all of them
(vpd.SpecimenBatch).

The default sweep is 48 bars: Δx ∈ {0.2, 0.3, 0.5, 1.0} mm × δ ∈ {2.015, 3.015, 4.015}Δx
× dt ∈ {0.25, 0.5, 0.75, 0.9}·CFL on the 40×10×10 mm bar of the 500k benchmark (200×50×50
at 0.2 mm, 40×10×10 at 1 mm), end-face velocity 1 mm/ms (1 m/s), critical-stretch damage
s_c = 0.02. (The three time-step fractions 0.25, 0.5, 0.9 give 36 bars; 0.75 completes the
48.) Every specimen runs to the same end time, `end_steps` steps of the finest time
step (0.2 mm, 0.25·CFL), so its step count is end_time / dt.

    python sweep.py [--end-steps 2000] [--steps K] [--json out.json]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import vpd  # noqa: E402

# Johnson-Cook A=350 MPa, B=275 MPa, n=0.36, C=0.022, m=1, ε̇₀ = 1/s, T_room 293 K, T_melt
# 1793 K, E = 200 GPa, ν = 0.3, ρ = 7850 kg/m³ (the tensile-bar material).
JC_PROPS = [200000.0, 0.3, 350.0, 275.0, 0.022, 0.36, 1.0, 1.0e-3, 293.0, 1793.0, 7850.0, 0.9,
            500.0, 1.0]
RHO = 7.85e-3  # N·ms²/mm⁴
SPACINGS = (0.2, 0.3, 0.5, 1.0)
HORIZON_FACTORS = (2.015, 3.015, 4.015)
CFL_FRACTIONS = (0.25, 0.5, 0.75, 0.9)


def cfl_dt(h, fraction, props=JC_PROPS, rho=RHO):
    """fraction · Δx / c_p with the P-wave speed c_p = sqrt(E(1−ν)/(ρ(1+ν)(1−2ν)))."""
    E, nu = props[0], props[1]
    return fraction * h / math.sqrt(E * (1 - nu) / (rho * (1 + nu) * (1 - 2 * nu)))


def specimens(end_steps=2000, length=40.0, width=10.0, spacings=SPACINGS,
              horizon_factors=HORIZON_FACTORS, cfl_fractions=CFL_FRACTIONS, velocity=1.0):
    """The sweep's specimens, largest first; each runs to end_steps · dt_finest."""
    end_time = end_steps * cfl_dt(min(spacings), min(cfl_fractions))
    out = []
    for h in spacings:
        nx, ny = int(round(length / h)), int(round(width / h))
        for f in horizon_factors:
            for c in cfl_fractions:
                dt = cfl_dt(h, c)
                steps = max(1, int(round(end_time / dt)))
                sp = vpd.Specimen(nx, ny, ny, h, f, dt, velocity, steps)
                sp.label = "dx%.1f_d%.3f_cfl%.2f" % (h, f, c)
                out.append(sp)
    return out


def summaries(batch):
    """End-of-run observables per specimen: peak ε̄ᵖ, peak and mean damage index (broken
    fraction of a particle's bonds) and the axial displacement of the loaded band."""
    statev, dmg, u = batch.split("statev"), batch.split("damage_index"), batch.split("u")
    out = []
    for s in range(len(batch.specimens)):
        band = (batch.codes[batch.first[s]:batch.first[s + 1]] & 3) == 2
        out.append(dict(max_eqps=float(statev[s][:, 1].max()),
                        max_damage_index=float(dmg[s].max()),
                        mean_damage_index=float(dmg[s].mean()),
                        load_displacement_mm=float(u[s][band, 0].mean()) if band.any() else 0.0))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--end-steps", type=int, default=2000)
    ap.add_argument("--steps", type=int, default=None,
                    help="batch steps to take (default: until every specimen has finished)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--json", default=None)
    a = ap.parse_args(argv)
    sps = specimens(a.end_steps)
    t0 = time.perf_counter()
    batch = vpd.SpecimenBatch(sps, JC_PROPS, RHO, damage=(0.02, 0.0))
    setup = time.perf_counter() - t0
    tuned = None if a.no_autotune else batch.ctx.autotune()
    batch.reset()
    k = batch.max_steps if a.steps is None else a.steps
    t0 = time.perf_counter()
    batch.step(k)
    batch.ctx.synchronize()
    run = time.perf_counter() - t0
    work = sum(sp.n * min(k, sp.steps) for sp in sps)
    out = dict(specimens=len(sps), particles=batch.n, bonds=int(batch.rowptr[-1]), steps=k,
               setup_s=setup, run_s=run, particle_steps_per_s=work / run, tuning=tuned,
               results=[dict(label=sp.label, n=sp.n, steps=sp.steps, dt=sp.dt, **r)
                        for sp, r in zip(sps, summaries(batch))])
    text = json.dumps(out, indent=1)
    if a.json:
        with open(a.json, "w") as f:
            f.write(text)
    print(text)


if __name__ == "__main__":
    main()
