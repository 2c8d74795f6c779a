"""Training pipeline of the learned return-map warm start (no reference counterpart).

1. Log (input, Δε̄ᵖ) pairs from seed-fixed plain GPU runs of the 30×8×8 JC tensile bar
   (Δx = 1 mm, δ = 3.015Δx, dt = 0.5·CFL, end velocity drawn per run from a seeded
   generator): for every plastic point of every step, the network input is the strain
   increment the return map received (6 Voigt components), the equivalent plastic strain and
   the temperature before the update, and the target is the converged Δε̄ᵖ.
2. Fit the input/output normalisation, then train the 8 → 64 → 64 → 64 → 1 network on the GPU
   with Adam: the forward and backward products as bf16 MFMA with FP32 accumulation (the
   default) or the FP32 kernels (--form fp32).
3. Save θ and the normalisation (an .npz of FP32 arrays).

    python vumat-pd_amd/warm_start.py --pairs 2000000 --out vumat-pd_amd/warm_start_jc.npz
"""
from __future__ import annotations

import argparse
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import vpd  # noqa: E402

JC_PROPS = [200000.0, 0.3, 350.0, 275.0, 0.022, 0.36, 1.0, 1.0e-3, 293.0, 1793.0, 7850.0, 0.9,
            500.0, 1.0]
RHO = 7.85e-3


def cfl_dt(h, fraction=0.5, props=JC_PROPS, rho=RHO):
    E, nu = props[0], props[1]
    return fraction * h / math.sqrt(E * (1 - nu) / (rho * (1 + nu) * (1 - 2 * nu)))


def initial_params(seed=0):
    """Glorot-uniform weights, zero biases, in the flat θ layout."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((8, 64), (64, 64), (64, 64), (64, 1)):
        lim = math.sqrt(6.0 / (fan_in + fan_out))
        parts.append(rng.uniform(-lim, lim, fan_out * fan_in))
        parts.append(np.zeros(fan_out))
    return np.concatenate(parts).astype(np.float32)


VOIGT_OF_TENSOR = [0, 4, 8, 5, 2, 1]  # Voigt [11,22,33,23,13,12] entries of a row-major 3×3


def log_pairs(target_pairs, seed=1234, steps=2000, nx=30, ny=8, nz=8, spacing=1.0,
              return_map=vpd.RETURN_MAP_BISECTION, runs=8, verbose=True):
    """Runs `runs` bars with end velocities drawn from U[0.5, 4] mm/ms and keeps an even,
    seeded random share of each run's plastic points, target_pairs in all. Returns a dict
    of per-point records: x (m, 8) raw network inputs, y (m,) converged Δε̄ᵖ, and the full
    return-map inputs inc (m, 6), stress (m, 6) (σ̄_n, Voigt), state (m, 6) (before the
    update), for evaluation."""
    rng = np.random.default_rng(seed)
    out = {k: [] for k in ("x", "y", "inc", "stress", "state")}
    total = 0
    dt = cfl_dt(spacing)
    for run in range(runs):
        v = float(rng.uniform(0.5, 4.0))
        rec = {k: [] for k in out}
        bar = vpd.Bar(nx, ny, nz, spacing, 3.015 * spacing, JC_PROPS, dt, RHO, velocity=v,
                      return_map=return_map)
        for _ in range(steps):
            before = bar.get("statev")
            stress = bar.get("stressN")[:, VOIGT_OF_TENSOR]
            bar.step(1, fuse=False, want_fields=True)
            after = bar.get("statev")
            plastic = after[:, 4] > 0.0
            if plastic.any():
                inc = bar.get("strain_inc")[plastic]
                rec["x"].append(np.concatenate([inc, before[plastic, 1:3]], 1))
                rec["y"].append(after[plastic, 4])
                rec["inc"].append(inc)
                rec["stress"].append(stress[plastic])
                rec["state"].append(before[plastic])
        bar.ctx.close()
        got = sum(len(y) for y in rec["y"])
        want = min(got, (target_pairs - total) // (runs - run))
        keep = np.sort(rng.choice(got, size=want, replace=False)) if got else []
        for k in out:
            if got:
                out[k].append(np.concatenate(rec[k])[keep])
        total += want
        if verbose:
            print("[warm-start] run %d v=%.3f mm/ms: %d plastic points, kept %d (%d so far)"
                  % (run, v, got, want, total), flush=True)
    return {k: np.concatenate(v) for k, v in out.items()}


def evaluate(net, records, dt=None, limit=200000):
    """Return-map solvers on logged plastic points: mean iterations, kernel time and the
    largest stress deviation from the tight bisection (the accuracy reference), per mode."""
    dt = cfl_dt(1.0) if dt is None else dt
    m = min(limit, len(records["y"]))
    inc = np.ascontiguousarray(records["inc"][:m].T)
    stress = np.ascontiguousarray(records["stress"][:m].T)
    state = np.ascontiguousarray(records["state"][:m].T)
    z = np.zeros(m)
    seed = np.zeros(m)
    theta, norm = net.get_params()
    norm = norm.astype(np.float64)
    xn = (records["x"][:m] - norm[:8]) / norm[8:16]  # FP64, as the seed kernels do
    seed = np.maximum(net.predict(xn).astype(np.float64) * norm[17] + norm[16], 0.0)
    out = {}
    exact = vpd.return_map(vpd.RETURN_MAP_BISECTION_TIGHT, JC_PROPS, dt, inc, stress, state, z,
                           z, rtol=1e-14)
    scale = np.abs(exact["stress"]).max()
    for name, mode, sd in (("bisection", vpd.RETURN_MAP_BISECTION, None),
                           ("bisection_tight", vpd.RETURN_MAP_BISECTION_TIGHT, None),
                           ("newton", vpd.RETURN_MAP_NEWTON, None),
                           ("learned", vpd.RETURN_MAP_LEARNED, seed)):
        r = vpd.return_map(mode, JC_PROPS, dt, inc, stress, state, z, z, seed=sd)
        out[name] = dict(mean_iterations=float(r["iterations"].mean()), ms=r["ms"],
                         max_stress_err_rel=float(np.abs(r["stress"] - exact["stress"]).max()
                                                  / scale),
                         fallbacks=float(r["fallbacks"].mean()))
    out["seed_rel_err_median"] = float(np.median(np.abs(seed - records["y"][:m]) /
                                                 np.maximum(records["y"][:m], 1e-30)))
    return out


def normalisation(x, y):
    """Input mean/std per feature and output mean/std: 18 FP32 values. An input's std is
    floored at 1e-3 of its mean's magnitude, so a feature that barely moves in the training
    runs (the temperature, heated by micro-kelvins) maps to ~0 instead of amplifying its
    noise, and the network stays bounded where it does move (adiabatic heating)."""
    mx, sx = x.mean(0), x.std(0)
    sx = np.maximum(sx, 1e-3 * np.abs(mx))
    sx = np.where(sx > 0, sx, 1.0)
    return np.concatenate([mx, sx, [y.mean(), max(y.std(), 1e-30)]]).astype(np.float32)


def train(x, y, steps=20000, batch=4096, lr=1e-3, seed=7, verbose=True, form="mfma"):
    norm = normalisation(x, y)
    xn = ((x - norm[:8]) / norm[8:16]).astype(np.float32)
    yn = ((y - norm[16]) / norm[17]).astype(np.float32)
    net = vpd.WarmStart(batch)
    net.set_training_form(net.TRAIN_MFMA if form == "mfma" else net.TRAIN_FP32)
    net.set_params(initial_params(seed), norm)
    net.set_dataset(xn, yn, seed=seed)
    t0 = time.perf_counter()
    chunk = max(1, steps // 10)
    loss = float("nan")
    for s in range(0, steps, chunk):
        # Step decay of the learning rate over the last half of training.
        rate = lr if s < steps // 2 else lr * 0.1
        loss = net.train(min(chunk, steps - s), batch, rate)
        if verbose:
            print("[warm-start] step %d loss %.4e" % (s + chunk, loss), flush=True)
    net.train_seconds = time.perf_counter() - t0
    if verbose:
        print("[warm-start] trained %d steps (%s) in %.2f s" % (steps, form, net.train_seconds))
    return net, norm, loss


def save(path, net):
    theta, norm = net.get_params()
    np.savez(path, theta=theta, norm=norm)


def load(path, max_batch=4096):
    data = np.load(path)
    net = vpd.WarmStart(max_batch)
    net.set_params(data["theta"], data["norm"])
    return net


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=2000000)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--out", default=os.path.join(HERE, "warm_start_jc.npz"))
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--eval-seed", type=int, default=98765)
    ap.add_argument("--form", choices=("mfma", "fp32"), default="mfma")
    a = ap.parse_args()
    rec = log_pairs(a.pairs, seed=a.seed)
    net, norm, loss = train(rec["x"], rec["y"], steps=a.steps, form=a.form)
    save(a.out, net)
    print("[warm-start] saved %s (final batch loss %.4e, %d pairs)" % (a.out, loss, len(rec["y"])))
    held = log_pairs(200000, seed=a.eval_seed, verbose=False)
    import json
    print("[warm-start] held-out evaluation " + json.dumps(evaluate(net, held)), flush=True)


if __name__ == "__main__":
    main()
