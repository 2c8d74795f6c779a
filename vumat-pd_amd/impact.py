"""High-rate, impact-like tension of a voided Johnson-Cook bar on one GPU.

400×70×70 lattice at Δx = 0.2 mm (1.96M sites) with 5% of the sites removed at random
(seeded voids, vpd.voided_patch), δ = 3.015Δx, end-face velocity 20 mm/ms (20 m/s),
dt = 0.5·CFL, JC with adiabatic heating (Taylor-Quinney β = 0.9, c_p = 500 J/(kg·K), the
MPa → J/m³ factor 1e6 of the reference's configs/johnson_cook_heat_units.json as
workToHeat), critical-stretch damage s_c = 0.02, the learned return map (network seed,
safeguarded Newton finish to |r| ≤ 1e-12·q, bisection fallback), 10000 steps.

Every `log_every` steps (every `early_log_every` over the first `early_steps`, where the
damage front forms) one unfused step keeps the fields, and the log records the plastic
points of that step, their mean return-map iterations, how many needed a bisection fallback
inside the Newton solve, peak ε̄ᵖ and temperature rise, the mean damage index (the broken
fraction of each particle's bonds), the particles whose bonds are all broken (those
`deletionEnabled` would flag, PdSimpleDamage.cpp:81-84), and the chunk's own throughput.
The summary splits the run at the first logged step whose mean damage index reaches 1%:
the intact phase before it (every bond pass sums nearly every bond) and the fragmenting
phase after it (the passes skip more and more broken bonds), each with its throughput.

    python impact.py [--steps 10000] [--log-every 500] [--early-log-every 100] [--early-steps 500]
                     [--return-map learned|newton|bisection]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import vpd  # noqa: E402

# The tensile-bar Johnson-Cook material with heating in consistent units: E, ν, A, B, C, n,
# m, ε̇₀ (1/ms), T_ref, T_melt, ρ, β, c_p, workToHeat.
JC_HEAT_PROPS = [200000.0, 0.3, 350.0, 275.0, 0.022, 0.36, 1.0, 1.0e-3, 293.0, 1793.0, 7850.0,
                 0.9, 500.0, 1.0e6]
RHO = 7.85e-3  # N·ms²/mm⁴
MODES = {"bisection": vpd.RETURN_MAP_BISECTION, "newton": vpd.RETURN_MAP_NEWTON,
         "learned": vpd.RETURN_MAP_LEARNED}


def cfl_dt(h, fraction=0.5, props=JC_HEAT_PROPS, rho=RHO):
    E, nu = props[0], props[1]
    return fraction * h / math.sqrt(E * (1 - nu) / (rho * (1 + nu) * (1 - 2 * nu)))


def build(nx=400, ny=70, nz=70, spacing=0.2, void_fraction=0.05, seed=20240607, velocity=20.0,
          damage=(0.02, 0.0), return_map="learned", net=None, rise=0.0):
    points = vpd.voided_patch(nx, ny, nz, spacing, void_fraction, seed)
    mode = MODES[return_map]
    if mode == vpd.RETURN_MAP_LEARNED and net is None:
        net = vpd.load_warm_start()
    return vpd.Bar(nx, ny, nz, spacing, 3.015 * spacing, JC_HEAT_PROPS, cfl_dt(spacing), RHO,
                   velocity=velocity, damage=damage, return_map=mode, warm_start=net,
                   points=points, velocity_rise=rise)


def log_record(bar, step, tref=293.0):
    """Observables of the step just taken (an unfused step that kept its fields)."""
    it = bar.get("iterations")
    fb = bar.get("fallbacks")
    statev = bar.get("statev")
    plastic = statev[:, 4] > 0  # Δε̄ᵖ of this step
    dmg = bar.get("damage_index")
    return dict(step=step, plastic_points=int(plastic.sum()),
                mean_iterations=float(it[plastic].mean()) if plastic.any() else 0.0,
                fallback_points=int((fb > 0).sum()),
                fallback_rate=float((fb[plastic] > 0).mean()) if plastic.any() else 0.0,
                max_eqps=float(statev[:, 1].max()),
                max_temperature_rise_K=float(statev[:, 2].max() - tref),
                fully_broken_particles=int((dmg >= 1.0).sum()),
                mean_damage_index=float(dmg.mean()))


def chunk_ends(steps, log_every, early_every=0, early_steps=0):
    """The logged steps: every early_every steps up to early_steps (the bar's intact start,
    where the damage front forms within a few hundred steps), then every log_every steps,
    and the last step."""
    ends, done = [], 0
    while done < steps:
        if early_every and done < early_steps:
            done = min(done + early_every, early_steps, steps)
        else:
            done = min(done + log_every, steps)
        ends.append(done)
    return ends


def run(bar, steps, log_every, early_every=0, early_steps=0):
    """`steps` steps in chunks ending at chunk_ends() (fused inside a chunk, its last step
    unfused and logged). Returns (seconds, log); logging downloads are not timed."""
    log = []
    seconds = 0.0
    done = 0
    for end in chunk_ends(steps, log_every, early_every, early_steps):
        k = end - done
        bar.ctx.synchronize()
        t0 = time.perf_counter()
        bar.step(k, want_fields=True)
        bar.ctx.synchronize()
        dt = time.perf_counter() - t0
        seconds += dt
        done += k
        rec = log_record(bar, done)
        rec.update(chunk_steps=k, chunk_seconds=dt, chunk_particle_steps_per_s=bar.n * k / dt)
        log.append(rec)
    return seconds, log


INTACT_DAMAGE = 0.01  # mean damage index that ends the intact phase


def phases(log, n):
    """Throughput of the intact phase (chunks ending below INTACT_DAMAGE mean damage index)
    and of the rest of the run."""
    out = {}
    split = next((k for k, r in enumerate(log) if r["mean_damage_index"] >= INTACT_DAMAGE),
                 len(log))
    for name, part in (("intact", log[:split]), ("fragmenting", log[split:])):
        steps = sum(r["chunk_steps"] for r in part)
        secs = sum(r["chunk_seconds"] for r in part)
        out[name] = dict(steps=steps, last_step=part[-1]["step"] if part else None,
                         seconds=secs, particle_steps_per_s=n * steps / secs if secs else None,
                         plastic_points_at_end=part[-1]["plastic_points"] if part else None)
    out["split_mean_damage_index"] = INTACT_DAMAGE
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--log-every", type=int, default=500)
    ap.add_argument("--early-log-every", type=int, default=100,
                    help="log cadence over the first --early-steps steps (0: none)")
    ap.add_argument("--early-steps", type=int, default=500)
    ap.add_argument("--return-map", choices=sorted(MODES), default="learned")
    ap.add_argument("--nx", type=int, default=400)
    ap.add_argument("--ny", type=int, default=70)
    ap.add_argument("--rise", type=float, default=0.0,
                    help="time (ms) over which the band accelerates to the end velocity")
    ap.add_argument("--json", default=None)
    a = ap.parse_args(argv)
    t0 = time.perf_counter()
    bar = build(a.nx, a.ny, a.ny, return_map=a.return_map, rise=a.rise)
    setup = time.perf_counter() - t0
    tuned = bar.ctx.autotune(3)
    bar.reset()
    seconds, log = run(bar, a.steps, a.log_every, a.early_log_every, a.early_steps)
    plastic = sum(r["plastic_points"] for r in log)
    out = dict(workload="voided_jc_bar_%dx%dx%d" % (a.nx, a.ny, a.ny), particles=bar.n,
               velocity_rise_ms=a.rise,
               bonds=bar.ctx.nnz, steps=a.steps, return_map=a.return_map, setup_s=setup,
               seconds=seconds, particle_steps_per_s=bar.n * a.steps / seconds,
               launch_tuning=tuned, phases=phases(log, bar.n),
               logged_fallback_rate=(sum(r["fallback_points"] for r in log) / plastic
                                     if plastic else 0.0),
               log=log)
    text = json.dumps(out, indent=1)
    if a.json:
        with open(a.json, "w") as f:
            f.write(text)
    print(text)


if __name__ == "__main__":
    main()
