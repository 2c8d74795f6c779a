"""Benchmark: particle-steps/s of the explicit correspondence-PD increment with the
Johnson-Cook VUMAT on one MI355X, on the 500k-particle JC tensile bar (200×50×50 lattice,
Δx = 0.2 mm, δ = 3.015Δx, FP64, critical-stretch damage s_c = 0.02, end-face velocity
1 mm/ms, dt = 0.5·Δx/c_p), with the dominant kernel's HBM roofline and the reference's CPU
path timed on this host beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--from-step S] [--chunk C]
                    [--no-cpu-baseline]

Prints one JSON line. A step is one fused increment (F + damage bond pass, material point
update, force bond pass + velocity-Verlet) over all 500k particles, inputs resident in HBM.
The timed window is steps S .. S+K−1 (default 4001-5000, the plastic tail of the bar's
5000-step run: yielding starts near step 3500, and by step 4000 some 75k points take a
plastic increment per step); the bar is advanced to step S−1−W untimed and runs W warmup
steps first. The same window is timed with the learned return map on a second context of
the bar from the same state, the two alternating chunks of --chunk steps so that drift over
the run reaches both alike; a second window from step 1 (elastic) is reported beside it.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "vumat-pd_amd"))

# Johnson-Cook A=350 MPa, B=275 MPa, n=0.36, C=0.022, m=1, ε̇₀ = 1/s = 1e-3/ms, T_room 293 K,
# T_melt 1793 K, E = 200 GPa, ν = 0.3, ρ = 7850 kg/m³, β = 0.9, c_p = 500, workToHeat = 1.
JC_PROPS = [200000.0, 0.3, 350.0, 275.0, 0.022, 0.36, 1.0, 1.0e-3, 293.0, 1793.0, 7850.0, 0.9,
            500.0, 1.0]
RHO = 7.85e-3  # N·ms²/mm⁴
HBM_PEAK = 8.0e12


def cfl_dt(h, fraction=0.5, props=JC_PROPS, rho=RHO):
    E, nu = props[0], props[1]
    cp = math.sqrt(E * (1 - nu) / (rho * (1 + nu) * (1 - 2 * nu)))
    return fraction * h / cp


WORKLOAD = dict(name="jc_tensile_bar_500k", nx=200, ny=50, nz=50, spacing=0.2,
                horizon=3.015 * 0.2, velocity=1.0, damage=(0.02, 0.0))


def bytes_per_particle(kernel, bbar, damage):
    """Algorithmic HBM bytes per particle and launch: every FP64 field the kernel must read
    or write once, int32 CSR indices once, bit-packed bond flags (one read, one write)."""
    flags = bbar / 8.0 if damage else 0.0
    if kernel == "defgrad_bond_pass":
        # X 24 + u 24 + V 8 + K⁻¹ 72 + rowptr 4 + col 4b̄ + F 72 out (+ flags r/w)
        return 24 + 24 + 8 + 72 + 4 + 4 * bbar + 72 + 2 * flags
    if kernel == "material_jc":
        # F 72 + (F_n, V, R) 216 r + 216 w + σ̄ 48+48 + statev(6) 48+48 + T 8 w
        # + energies 16+16 + K⁻¹ 72 + A 72 out
        return 72 + 432 + 96 + 96 + 8 + 32 + 72 + 72
    if kernel == "force_bond_pass":
        # X 24 + V 8 + A 72 + rowptr 4 + col 4b̄ + u, v, a 72 r + 72 w + code 1 (+ flags r)
        return 24 + 8 + 72 + 4 + 4 * bbar + 144 + 1 + flags
    if kernel == "predict":
        return 144 + 1
    return 0.0


def step_bytes_per_particle(bbar, damage):
    """The whole-step model: 816 persistent state + 136 (X, V both passes, K⁻¹) + 144 (A)
    + 24 (u re-read) + 8 (rowptr) + 8b̄ (col, both passes) + b̄/4 (flags r/w)."""
    return 816 + 136 + 144 + 24 + 8 + 8 * bbar + (bbar / 4.0 if damage else 0.0)


def bond_pass_device_kernel(kernel, tuning):
    """The device kernel (as rocprofv3 names it) that runs bond pass `kernel` in the step's
    form (damage on; the force pass integrating, with the next prediction fused) under the
    launch tuning `tuning`."""
    p = "defgrad" if kernel == "defgrad_bond_pass" else "force"
    window, form = tuning[p + "_window"], tuning[p + "_window_plan"]
    head, args = ("k_defgrad", "true") if p == "defgrad" else ("k_force", "true, true, true")
    if window == 0:
        return "k_defgrad<true>" if p == "defgrad" else "k_force<true, true, true, false>"
    if form == 0:
        return "%s_w<%s, %d>" % (head, args, window)
    if form == 3:
        return "%s_g<%s, %d>" % (head, args, window)
    if form == 4:
        return "%s_m<%s, %d>" % (head, args, window)
    return "%s_p<%s, %d, %s>" % (head, args, window, "true" if form == 2 else "false")


def measured_traffic(kernel, workload, device_kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary of this workload (calibrated against a device copy), or None when there is none
    for the device kernel that runs it now."""
    path = os.path.join(ROOT, "profiles", "traffic_%s_%s.json" % (kernel, workload))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    if (data.get("kernel") != kernel or data.get("workload") != workload
            or data.get("device_kernel") != device_kernel):
        return None
    return data["hbm_bytes_per_launch"]


def cpu_baseline(state, window, small_steps, w):
    """The reference's own OpenMP increment (oracle/_ref; the restated oracle when the
    reference build is absent) on this host's cores, run by oracle/cpu_sample.py in a child
    process with OMP_PROC_BIND=close and OMP_DYNAMIC=FALSE set before its OpenMP runtime
    starts: the 500k bar over `window` steps from `state` (the GPU run's state just before its
    timed window, so the CPU sample covers the same plastic steps), and the 30×8×8 bar for its
    whole `small_steps`-step run."""
    import subprocess
    import tempfile

    import numpy as np

    threads = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
               OMP_DYNAMIC="FALSE")
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "state.npz")
        np.savez(path, **state)
        cmd = [sys.executable, os.path.join(ROOT, "oracle", "cpu_sample.py"), "--state", path,
               "--window", str(window), "--small-steps", str(small_steps),
               "--velocity", str(w["velocity"])]
        run = subprocess.run(cmd, env=env, capture_output=True, text=True)
    if run.returncode != 0:
        raise SystemExit("CPU baseline failed:\n" + run.stderr[-2000:])
    res = json.loads(run.stdout.strip().splitlines()[-1])
    large, small = res["large"], res.get("small")
    cpu = res["cpu"]
    sample = ("steps %s of the %d-particle bar (the GPU run's timed window, plastic phase: %d "
              "points plastic in the first step) from the GPU state, setup untimed; %d OpenMP "
              "threads, OMP_PROC_BIND=close, on %s (%s socket(s), %s physical cores, %d CPUs "
              "allowed)" % (large["steps"], large["particles"], large["plastic_points_at_start"],
                            res["threads"], cpu.get("model"), cpu.get("sockets"),
                            cpu.get("physical_cores"), cpu.get("cpus_allowed")))
    out = dict(value=large["value"], unit="particle-steps/s", cores=res["threads"],
               kind=res["kind"], sample=sample, seconds=large["seconds"], cpu=cpu,
               omp=res["omp"], window=large)
    if small:
        out["small_bar"] = small
    return out


# Kernel means come from HIP events on every TIMING_STRIDE-th step of the timed region: events
# on every kernel of every step cost about 4% of the 500k bar's step time.
TIMING_STRIDE = 10


def timed_window(bar, steps):
    """`steps` timed steps from the bar's current state. Returns (seconds, per-kernel HIP-event
    times of every TIMING_STRIDE-th step)."""
    bar.ctx.set_timing(True, TIMING_STRIDE)
    bar.ctx.synchronize()
    t0 = time.perf_counter()
    bar.step(steps)
    bar.ctx.synchronize()
    elapsed = time.perf_counter() - t0
    kt = bar.ctx.collect_timing()
    bar.ctx.set_timing(False)
    return elapsed, kt


def timed_interleaved(bars, steps, chunk):
    """`steps` timed steps of each bar in `bars`, in alternating chunks of `chunk` steps
    (bar 0's chunk, bar 1's chunk, ...), so that clock and temperature drift over the window
    reaches every bar alike. Returns per bar (seconds, summed per-kernel HIP-event times)."""
    out = [[0.0, {}] for _ in bars]
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        for b, acc in zip(bars, out):
            elapsed, kt = timed_window(b, k)
            acc[0] += elapsed
            for name, (ms, launches) in kt.items():
                tot = acc[1].setdefault(name, [0.0, 0])
                tot[0] += ms
                tot[1] += launches
        done += k
    return [(e, {k: tuple(v) for k, v in kt.items()}) for e, kt in out]


def timed_run(bar, steps, warmup):
    """Warmup steps, reset to the state before step 1, then `steps` timed steps from step 1.
    Returns (seconds, per-kernel HIP-event times of every TIMING_STRIDE-th step)."""
    if warmup:
        bar.step(warmup)
        bar.ctx.synchronize()
    bar.reset()
    return timed_window(bar, steps)


def window_result(bar, first, steps, elapsed, kt, n):
    st = bar.get("statev")
    return {"first_step": first, "last_step": first + steps - 1, "value": n * steps / elapsed,
            "ms_per_step": elapsed / steps * 1e3, "max_eqps": float(st[:, 1].max()),
            "plastic_points_last_step": int((st[:, 4] > 0.0).sum()),
            "broken_bonds": int(bar.get("damage").sum()),
            "kernels_ms_per_launch": {k: v[0] / v[1] for k, v in kt.items()}}


def return_map_study(points, seed=4321):
    """The return-map solvers on `points` plastic material points logged from seed-fixed plain
    runs of the 30×8×8 JC bar (the warm start's training workload, a held-out seed): kernel
    time, mean iterations and the largest stress deviation from the FP64 tight bisection
    (|r| ≤ 1e-14·q), per solver. The learned rows run seeding and solve in one kernel
    (vpd_return_map_learned_device), so their time includes the network's inference."""
    import numpy as np
    import vpd
    import warm_start as ws

    rec = ws.log_pairs(points, seed=seed, verbose=False)
    m = len(rec["y"])
    inc = np.ascontiguousarray(rec["inc"].T)
    stress = np.ascontiguousarray(rec["stress"].T)
    state = np.ascontiguousarray(rec["state"].T)
    z = np.zeros(m)
    dt = ws.cfl_dt(1.0)
    props = ws.JC_PROPS
    exact = vpd.return_map(vpd.RETURN_MAP_BISECTION_TIGHT, props, dt, inc, stress, state, z, z,
                           rtol=1e-14)
    scale = np.abs(exact["stress"]).max()
    net = vpd.load_warm_start()
    first = state[4] == 0.0  # first yield (or reloading): state 5 carries no seed
    no5 = state.copy()
    no5[4] = 0.0
    out = {"points": m, "source": "plastic points of 30x8x8 JC bar plain runs, seed %d" % seed,
           "first_yield_points": int(first.sum())}
    for name, mode in (("bisection_parity", vpd.RETURN_MAP_BISECTION),
                       ("bisection_tight", vpd.RETURN_MAP_BISECTION_TIGHT),
                       ("newton_state5_seed", vpd.RETURN_MAP_NEWTON),
                       ("newton_learned_seed", vpd.RETURN_MAP_LEARNED),
                       ("newton_network_everywhere", -1)):
        best = None
        for _ in range(3):  # the fastest of three launches
            if mode in (vpd.RETURN_MAP_LEARNED, -1):
                r = vpd.return_map_learned(net, props, dt, inc, stress,
                                           state if mode >= 0 else no5, z, z, want_seeds=False)
                total = r["ms"]
            else:
                r = vpd.return_map(mode, props, dt, inc, stress, state, z, z)
                total = r["ms"]
            if best is None or total < best[0]:
                best = (total, r)
        total, r = best
        row = {"ms": total, "mean_iterations": float(r["iterations"].mean()),
               "mean_iterations_first_yield": float(r["iterations"][first].mean())
               if first.any() else None,
               "max_stress_err_rel": float(np.abs(r["stress"] - exact["stress"]).max() / scale),
               "max_eqps_inc_err": float(np.abs(r["state"][4] - exact["state"][4]).max()),
               "newton_fallbacks_per_point": float(r["fallbacks"].mean())}
        if mode in (vpd.RETURN_MAP_LEARNED, -1):
            row["seeds_needing_network"] = int((~(state[4] > 0.0) if mode >= 0
                                                else np.ones(m, bool)).sum())
            seeds = vpd.return_map_learned(net, props, dt, inc, stress,
                                           state if mode >= 0 else no5, z, z)["seed"]
            row["seed_median_rel_err"] = float(np.median(
                np.abs(seeds - exact["state"][4]) / np.maximum(exact["state"][4], 1e-30)))
        out[name] = row
    learned = out["newton_learned_seed"]["ms"]
    out["speedup_learned_vs_tight_bisection"] = out["bisection_tight"]["ms"] / learned
    out["speedup_learned_vs_parity_bisection"] = out["bisection_parity"]["ms"] / learned
    out["speedup_learned_vs_newton_state5"] = out["newton_state5_seed"]["ms"] / learned
    return out


def sweep_measure(steps, warmup, end_steps=2000):
    """The 48-bar sensitivity sweep (Δx × δ/Δx × dt/CFL, 4k-500k particles per bar): `steps`
    batch steps of all specimens in one particle block (vpd.SpecimenBatch), against the same
    specimens stepped one after another as separate single-bar contexts (the reference's
    process-per-case sweep, on the GPU). Both autotuned; setup untimed."""
    import vpd
    import sweep

    sps = sweep.specimens(end_steps)
    if steps > min(sp.steps for sp in sps):
        raise SystemExit("--steps exceeds the shortest specimen's step count")
    cache = {}
    print("[bench] building the %d-specimen batch" % len(sps), file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    batch = vpd.SpecimenBatch(sps, JC_PROPS, RHO, damage=(0.02, 0.0), neighbor_cache=cache)
    setup = time.perf_counter() - t0
    tuned = batch.ctx.autotune(3)
    batch.reset()
    batch.step(warmup)
    batch.ctx.synchronize()
    batch.reset()
    batch.ctx.set_timing(True, TIMING_STRIDE)
    t0 = time.perf_counter()
    batch.step(steps)
    batch.ctx.synchronize()
    t_batch = time.perf_counter() - t0
    kt = batch.ctx.collect_timing()
    batch.ctx.set_timing(False)
    batch.reset()  # the same steps as replays of one captured HIP graph of a step
    t0 = time.perf_counter()
    batch.run(steps)
    t_graph = time.perf_counter() - t0
    n_total, nnz_total = batch.n, int(batch.rowptr[-1])
    del batch
    t_seq, per = 0.0, []
    for sp in sps:
        rowptr, col = cache[(sp.nx, sp.ny, sp.nz, sp.spacing, sp.horizon)]
        bar = vpd.Bar(props=JC_PROPS, density=RHO, damage=(0.02, 0.0), rowptr=rowptr, col=col,
                      **sp.bar_kwargs())
        bar.ctx.autotune(3)
        elapsed, _ = timed_run(bar, steps, warmup)
        t_seq += elapsed
        per.append(dict(label=sp.label, n=sp.n, ms_per_step=elapsed / steps * 1e3))
        bar.ctx.close()
    work = n_total * steps
    return {"metric": "particle-steps/s (48-bar JC sweep, one block)", "specimens": len(sps),
            "particles": n_total, "bonds": nnz_total, "steps": steps, "warmup": warmup,
            "batched": {"value": work / t_batch, "ms_per_step": t_batch / steps * 1e3,
                        "launch_tuning": tuned,
                        "kernels_ms_per_launch": {k: v[0] / v[1] for k, v in kt.items()},
                        "graph_value": work / t_graph, "graph_ms_per_step": t_graph / steps * 1e3},
            "one_at_a_time": {"value": work / t_seq, "ms_per_step": t_seq / steps * 1e3,
                              "per_specimen": per},
            "speedup": t_seq / t_batch, "setup_s": setup}


def small_bar_graph(steps=2000, tuning=None):
    """The CPU-sized JC bar (30×8×8, Δx = 1 mm, δ = 3.015Δx, dt = 0.5·CFL, 1 mm/ms, damage
    off) for `steps` steps: host-issued fused steps (vpd_step) against the same bar as a
    one-specimen batch replaying one captured HIP graph per step. `tuning`: launch knobs to
    set instead of autotuning."""
    import vpd

    dt = cfl_dt(1.0)
    bar = vpd.Bar(30, 8, 8, 1.0, 3.015, JC_PROPS, dt, RHO, velocity=1.0)
    if tuning:
        bar.ctx.set_tuning(**tuning)
    else:
        bar.autotune_in_steps(3)
    bar.step(20)
    bar.reset()
    bar.ctx.synchronize()
    t0 = time.perf_counter()
    bar.step(steps)
    bar.ctx.synchronize()
    elapsed = time.perf_counter() - t0
    u_eager = bar.get("u")
    batch = vpd.SpecimenBatch([vpd.Specimen(30, 8, 8, 1.0, 3.015, dt, 1.0, steps)], JC_PROPS,
                              RHO)
    batch.ctx.set_tuning(**bar.ctx.get_tuning())
    batch.run(20)
    batch.reset()
    t0 = time.perf_counter()
    batch.run()
    t_graph = time.perf_counter() - t0
    if not (batch.get("u") == u_eager).all():
        raise SystemExit("graph-replayed bar differs from the eager bar")
    return {"workload": "jc_tensile_bar_30x8x8", "particles": bar.n, "steps": steps,
            "eager_ms_per_step": elapsed / steps * 1e3,
            "graph_ms_per_step": t_graph / steps * 1e3,
            "eager_value": bar.n * steps / elapsed, "graph_value": bar.n * steps / t_graph,
            "launch_tuning": bar.ctx.get_tuning()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--from-step", type=int, default=4001,
                    help="first timed step: the bar is advanced untimed to step from-step − 1 − "
                         "warmup, then runs `warmup` steps and times `steps` steps (4001 times "
                         "the plastic phase of the 5000-step run; 1 the start)")
    ap.add_argument("--chunk", type=int, default=100,
                    help="steps per chunk of the timed window; the plain and the learned "
                         "return map alternate chunks")
    ap.add_argument("--no-step1-window", action="store_true",
                    help="skip the second timed window from step 1")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-window", type=int, default=10,
                    help="500k-bar steps the CPU baseline times (from the first timed step)")
    ap.add_argument("--cpu-small-steps", type=int, default=2000,
                    help="steps of the 30x8x8 bar the CPU baseline also times (0: none)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tuning", default=None,
                    help="launch knobs as JSON (the launch_tuning of an earlier run) instead of "
                         "autotuning: a profiled run then holds the step's launches only")
    ap.add_argument("--no-warm-start", action="store_true")
    ap.add_argument("--return-map-points", type=int, default=1000000)
    ap.add_argument("--sweep", action="store_true",
                    help="measure the 48-bar batched sweep instead (JSON on stdout)")
    ap.add_argument("--small-bar", action="store_true",
                    help="measure only the 30x8x8 bar's 2000-step run, eager and graph-replayed "
                         "(JSON on stdout; --tuning fixes its launch knobs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this path is not sharded across GPUs. "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.small_bar:
        print(json.dumps(small_bar_graph(tuning=json.loads(args.tuning) if args.tuning
                                         else None)))
        return
    if args.sweep:
        out = sweep_measure(min(args.steps, 100), min(args.warmup, 5))
        out["small_bar"] = small_bar_graph()
        print(json.dumps(out))
        return
    if args.from_step < 1:
        raise SystemExit("--from-step must be at least 1")

    import vpd

    w = WORKLOAD
    dt = cfl_dt(w["spacing"])
    print("[bench] building %s on the GPU" % w["name"], file=sys.stderr, flush=True)
    bar = vpd.Bar(w["nx"], w["ny"], w["nz"], w["spacing"], w["horizon"], JC_PROPS, dt, RHO,
                  velocity=w["velocity"], damage=w["damage"])
    n, nnz = bar.n, bar.ctx.nnz
    bbar = nnz / n
    if args.tuning:
        bar.ctx.set_tuning(**json.loads(args.tuning))
        tuned = bar.ctx.get_tuning()
    else:
        tuned = bar.ctx.get_tuning() if args.no_autotune else bar.autotune_in_steps(3)
    print("[bench] launch tuning %s" % tuned, file=sys.stderr, flush=True)

    # The timed window: steps from_step .. from_step + steps − 1, after an untimed advance and
    # `warmup` untimed steps. The learned return map runs on a second context of the same bar
    # from the same snapshot with the same launch tuning; the two time alternating chunks of
    # the window (timed_interleaved).
    first = args.from_step
    warm = min(args.warmup, first - 1)
    bar.reset()
    advance = first - 1 - warm
    if advance:
        print("[bench] advancing to step %d untimed" % advance, file=sys.stderr, flush=True)
        bar.step(advance)
    bars = [bar]
    if not args.no_warm_start:
        snap = bar.snapshot()
        learned = vpd.Bar(w["nx"], w["ny"], w["nz"], w["spacing"], w["horizon"], JC_PROPS, dt,
                          RHO, velocity=w["velocity"], damage=w["damage"],
                          return_map=vpd.RETURN_MAP_LEARNED, warm_start=vpd.load_warm_start())
        learned.ctx.set_tuning(**tuned)
        learned.restore(snap)
        del snap
        bars.append(learned)
    if warm:
        for b in bars:
            b.step(warm)
    cpu_state = bar.snapshot() if not args.no_cpu_baseline else None
    print("[bench] timing %d steps from step %d, %s in alternating chunks of %d"
          % (args.steps, first, "plain and learned" if len(bars) > 1 else "plain",
             args.chunk), file=sys.stderr, flush=True)
    timed = timed_interleaved(bars, args.steps, args.chunk)
    elapsed, kt = timed[0]
    u = bar.get("u")
    if not np.isfinite(u).all():
        raise SystemExit("non-finite displacement after the timed run")
    win = window_result(bar, first, args.steps, elapsed, kt, n)

    value = n * args.steps / elapsed
    dom = max((k for k in kt if k != "neighbor_build"), key=lambda k: kt[k][0])
    avg_s = kt[dom][0] / kt[dom][1] / 1e3
    per_launch = bytes_per_particle(dom, bbar, w["damage"] is not None) * n
    achieved = per_launch / avg_s
    step_bps = step_bytes_per_particle(bbar, w["damage"] is not None)
    dev_kernel = bond_pass_device_kernel(dom, tuned) if dom.endswith("bond_pass") else None
    traffic = measured_traffic(dom, w["name"], dev_kernel)
    out = {
        "metric": "particle-steps/s (500k JC bar, \u03b4=3.015\u0394x) GPU plain vs warm-start; %HBM peak",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (deterministic lattice, JC material of the benchmark config)",
        "config": {"workload": w["name"], "particles": n, "bonds": nnz,
                   "grid": [w["nx"], w["ny"], w["nz"]], "spacing_mm": w["spacing"],
                   "horizon_over_spacing": 3.015, "dt_ms": dt,
                   "end_velocity_mm_per_ms": w["velocity"], "damage": "critical_stretch s_c=0.02",
                   "return_map": "bisection (parity)", "steps_from": first,
                   "launch_tuning": tuned},
        "roofline": {"bound": "hbm", "kernel": dom, "device_kernel": dev_kernel,
                     "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "traffic_source": None if traffic is None else
                     "committed rocprofv3 FETCH_SIZE+WRITE_SIZE of this device kernel "
                     "(profiles/traffic_%s_%s.json), not measured in this run" % (dom, w["name"]),
                     "algorithmic_bytes_per_launch": per_launch, "avg_launch_ms": avg_s * 1e3},
        "step_model": {"bytes_per_particle_step": step_bps,
                       "hbm_fraction": step_bps * value / HBM_PEAK},
        "kernels_ms_per_launch": win["kernels_ms_per_launch"],
        "window": {k: v for k, v in win.items() if k != "kernels_ms_per_launch"},
        "result": {"max_eqps": win["max_eqps"], "broken_bonds": win["broken_bonds"]},
    }
    if not args.no_warm_start:
        # The same window with the learned return map (compacted: Newton on the plastic
        # points only, the network on those without a positive state 5).
        ws_elapsed, ws_kt = timed[1]
        learned = bars.pop()
        ws = window_result(learned, first, args.steps, ws_elapsed, ws_kt, n)
        ws["unit"] = "particle-steps/s"
        ws["over_plain"] = ws["value"] / value
        out["warm_start"] = ws
        learned.ctx.close()
        del learned
    if not args.no_step1_window and first > 1:
        print("[bench] timing %d plain steps from step 1" % args.steps, file=sys.stderr,
              flush=True)
        s1_elapsed, s1_kt = timed_run(bar, args.steps, args.warmup)
        out["window_step1"] = window_result(bar, 1, args.steps, s1_elapsed, s1_kt, n)
    if not args.no_warm_start and args.return_map_points:
        print("[bench] return-map study on %d plastic points" % args.return_map_points,
              file=sys.stderr, flush=True)
        out["return_map"] = return_map_study(args.return_map_points)
    if not args.no_cpu_baseline:
        del bar
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        cb = cpu_baseline(cpu_state, args.cpu_window, args.cpu_small_steps, w)
        cb["gpu_over_cpu"] = value / cb["value"]
        out["cpu_baseline"] = cb
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
