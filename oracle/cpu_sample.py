"""ORACLE — test infrastructure only: the CPU-baseline sample of the benchmark.

Times the reference's own OpenMP increment (oracle/_ref, the reference's classes driven in
the order of BoundaryValueCaseRunner.cpp:207-229; the restated oracle when _ref is absent)
on this host, on two samples:

  * the 30×8×8 JC bar (Δx = 1 mm, δ = 3.015Δx, dt = 0.5·CFL, damage off) for its whole
    2000-step run, elastic and plastic phases;
  * the 500k JC bar (200×50×50, Δx = 0.2 mm, damage s_c = 0.02) over a window of steps that
    starts from a state the GPU run reached (an .npz of the fields a step reads), so that
    the window holds the plastic phase without the hours the CPU would need to get there.

The reference runner only drives the end band with its half-cosine ramp
(BoundaryConditions.cpp:70-79). The ramp used here matches the velocity-driven band's
displacement and velocity at the window's first step (tan(θ/2) = θ with θ = πt/T_rise), so
the band continues the GPU run's motion; a step does the same work whichever values the band
is given.

Run as a child process by bench.py so that the OpenMP environment (OMP_NUM_THREADS,
OMP_PROC_BIND=close, OMP_DYNAMIC=FALSE) is in place before the runtime starts. Prints one JSON
object.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import oracle as O  # noqa: E402

JC_PROPS = [200000.0, 0.3, 350.0, 275.0, 0.022, 0.36, 1.0, 1.0e-3, 293.0, 1793.0, 7850.0, 0.9,
            500.0, 1.0]
RHO = 7.85e-3


def cfl_dt(h, fraction=0.5):
    E, nu = JC_PROPS[0], JC_PROPS[1]
    return fraction * h / math.sqrt(E * (1 - nu) / (RHO * (1 + nu) * (1 - 2 * nu)))


def cpu_info():
    """Model name, sockets, physical cores and threads of this host (lscpu), and the CPUs this
    process may run on."""
    info = {"cpus_allowed": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        out = ""
    fields = {}
    for line in out.splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            fields[k.strip()] = v.strip()
    info["model"] = fields.get("Model name")
    try:
        sockets = int(fields["Socket(s)"])
        per_socket = int(fields["Core(s) per socket"])
        info.update(sockets=sockets, physical_cores=sockets * per_socket,
                    threads_per_core=int(fields.get("Thread(s) per core", "1")),
                    logical_cpus=int(fields.get("CPU(s)", "0")))
    except (KeyError, ValueError):
        pass
    if info["model"] is None and os.path.exists("/proc/cpuinfo"):
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    info["model"] = line.split(":", 1)[1].strip()
                    break
    return info


def matched_ramp(velocity, t):
    """(amplitude, rise time) of the half-cosine ramp ½A(1 − cos(πt/T)) whose displacement and
    velocity at time t equal v·t and v: θ = πt/T solves tan(θ/2) = θ, A = 2vt/(1 − cos θ)."""
    if t <= 0.0:
        return 0.0, 1.0
    lo, hi = 1.0, 3.0  # tan(θ/2) − θ changes sign once in (1, 3)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if math.tan(0.5 * mid) - mid < 0.0:
            lo = mid
        else:
            hi = mid
    theta = 0.5 * (lo + hi)
    return 2.0 * velocity * t / (1.0 - math.cos(theta)), math.pi * t / theta


def make_bar(kind, **kw):
    if kind == "reference":
        kw.pop("velocity", None)
        return O.RefBar(**kw)
    return O.OracleBar(**kw)


def small_bar(kind, steps):
    """The 30×8×8 bar, whole run: end band to v·t_end = 1 mm/ms × steps·dt."""
    dt = cfl_dt(1.0)
    t_end = steps * dt
    kw = dict(nx=30, ny=8, nz=8, spacing=1.0, horizon=3.015, props=JC_PROPS, dt=dt,
              density=RHO, amplitude=1.0 * t_end, rise_time=t_end)
    if kind != "reference":
        kw["velocity"] = 1.0
    bar = make_bar(kind, **kw)
    t0 = time.perf_counter()
    bar.step(steps)
    sec = time.perf_counter() - t0
    eqps = bar.get("statev")[:, 1]
    n = 30 * 8 * 8
    return {"workload": "jc_tensile_bar_30x8x8", "particles": n, "steps": "1-%d" % steps,
            "seconds": sec, "value": n * steps / sec, "max_eqps": float(eqps.max())}


def large_window(kind, state_path, window, velocity):
    """The 500k bar over `window` steps from the state in `state_path`."""
    st = np.load(state_path)
    first = int(st["step"]) + 1
    dt = cfl_dt(0.2)
    amp, rise = matched_ramp(velocity, first * dt)
    kw = dict(nx=200, ny=50, nz=50, spacing=0.2, horizon=3.015 * 0.2, props=JC_PROPS, dt=dt,
              density=RHO, amplitude=amp, rise_time=rise, damage=(0.02, 0.0))
    if kind != "reference":
        kw["velocity"] = velocity
    t0 = time.perf_counter()
    bar = make_bar(kind, **kw)
    setup = time.perf_counter() - t0
    for name in ("u", "v", "a", "fN", "vN", "rN", "stressN", "statev", "temp", "eint", "einel"):
        bar.set(name, st[name])
    broken = st["broken"]
    if broken.size:
        d = np.zeros(bar.nnz)
        d[broken] = 1.0
        bar.set("damage", d)
    bar.set_step(first - 1)
    plastic_before = int((st["statev"][:, 4] > 0.0).sum())
    t0 = time.perf_counter()
    bar.step(window)
    sec = time.perf_counter() - t0
    state = bar.get("statev")
    n = state.shape[0]
    return {"workload": "jc_tensile_bar_500k", "particles": n,
            "steps": "%d-%d" % (first, first + window - 1), "seconds": sec,
            "value": n * window / sec, "setup_seconds": setup,
            "plastic_points_at_start": plastic_before,
            "plastic_points_last_step": int((state[:, 4] > 0.0).sum()),
            "max_eqps": float(state[:, 1].max())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--state", help=".npz of the 500k bar's state (fields, step, broken)")
    ap.add_argument("--window", type=int, default=10)
    ap.add_argument("--small-steps", type=int, default=2000)
    ap.add_argument("--velocity", type=float, default=1.0)
    args = ap.parse_args()
    kind = "reference" if O.ref_available() else "port"
    if kind == "port":
        O.build()
    out = {"kind": kind, "cpu": cpu_info(), "threads": int(O.lib().orc_num_threads()),
           "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND",
                                                  "OMP_DYNAMIC", "OMP_PLACES")}}
    if args.small_steps:
        out["small"] = small_bar(kind, args.small_steps)
    if args.state:
        out["large"] = large_window(kind, args.state, args.window, args.velocity)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
