"""Parity and invariants at the benchmark configurations' full sizes.

* The 500k JC bar (200×50×50, Δx = 0.2 mm, δ = 3.015Δx, critical-stretch damage) in its
  plastic phase: the GPU bar is advanced to step 4000 under the half-cosine ramp of
  BoundaryConditions.cpp:70-79, its state is handed to the reference's own classes
  (oracle/_ref) and both take steps 4001-4003 (BoundaryValueCaseRunner.cpp:207-229):
  displacement, stress, ε̄ᵖ within 1e-9 of their scale per step, broken-bond sets identical.
* The 30×8×8 JC bar for its whole 2000-step run against the oracle: within 1e-6 over the run.
* The 48-bar sensitivity sweep as one batched block: every bar equals the same bar run alone,
  bit for bit.
* The voided 400×70×70 bar (1.86M particles): global balance Σ f·V = 0 to rounding, two runs
  bit-identical, direct and group-window gathers bit-identical.
"""
import math
import sys

import numpy as np
import pytest

from conftest import JC_PROPS

pytestmark = pytest.mark.gpu

RHO = 7.85e-3


def banner(msg):
    print("[gpu-case]", msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def vpd():
    import vpd as V
    if V.device_count() == 0:
        pytest.fail("no HIP device visible to the GPU tests")
    return V


def cfl_dt(h, fraction=0.5):
    E, nu = JC_PROPS[0], JC_PROPS[1]
    return fraction * h / math.sqrt(E * (1 - nu) / (RHO * (1 + nu) * (1 - 2 * nu)))


def intact_bond_limit(bar):
    return bar.ctx.nnz // 100


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_500k_plastic_steps_vs_reference(vpd, ref):
    banner("500k bar: steps 4001-4003 vs the reference's classes from the GPU state")
    dt = cfl_dt(0.2)
    kw = dict(nx=200, ny=50, nz=50, spacing=0.2, horizon=3.015 * 0.2, props=JC_PROPS, dt=dt,
              density=RHO, amplitude=0.1, rise_time=5000 * dt, damage=(0.02, 0.0))
    bar = vpd.Bar(**kw)
    bar.step(4000)
    snap = bar.snapshot()
    plastic = int((snap["statev"][:, 4] > 0).sum())
    print("plastic points at step 4000: %d, broken bonds %d" % (plastic, snap["broken"].size))
    assert plastic > 10000
    assert snap["broken"].size < intact_bond_limit(bar)  # intact: not the fragmented regime
    r = ref.RefBar(**kw)
    for name in bar.STATE_FIELDS:
        r.set(name, snap[name])
    d = np.zeros(r.nnz)
    d[snap["broken"]] = 1.0
    r.set("damage", d)
    r.set_step(4000)
    for step in range(4001, 4004):
        bar.step(1)
        r.step(1)
        for name, tol in (("u", 1e-9), ("v", 1e-9), ("stressN", 1e-9)):
            err = rel(bar.get(name), r.get(name))
            assert err <= tol, (step, name, err)
        g, h = bar.get("statev"), r.get("statev")
        assert rel(g[:, 1], h[:, 1]) <= 1e-9, step
        assert np.array_equal(bar.get("damage"), r.get("damage")), step
    print("broken bonds at step 4003: %d" % int(bar.get("damage").sum()))


def test_30x8x8_full_run_vs_oracle(vpd, orc):
    banner("30x8x8 bar: 2000 steps vs the oracle")
    dt = cfl_dt(1.0)
    kw = dict(nx=30, ny=8, nz=8, spacing=1.0, horizon=3.015, props=JC_PROPS, dt=dt,
              density=RHO, amplitude=0.0, rise_time=1.0, velocity=1.0)
    g = vpd.Bar(**kw)
    o = orc.OracleBar(**kw)
    worst = {}
    for chunk in range(4):
        g.step(500)
        o.step(500)
        for name, ref in (("u", o.u), ("v", o.v), ("stressN", o.mat["stressN"]),
                          ("eqps", o.mat["statev"][:, 1])):
            got = g.get("statev")[:, 1] if name == "eqps" else g.get(name)
            worst[name] = max(worst.get(name, 0.0), rel(got, ref))
    print("max relative error over the run:", worst)
    assert o.mat["statev"][:, 1].max() > 0  # the run reaches the plastic phase
    for name, err in worst.items():
        assert err <= 1e-6, (name, err)


def test_sweep_48_batched_equals_alone(vpd):
    banner("48-bar sweep: batched == each bar alone, 3 steps")
    import sweep
    sps = sweep.specimens(2000)
    assert len(sps) == 48
    cache = {}
    batch = vpd.SpecimenBatch(sps, JC_PROPS, RHO, damage=(0.02, 0.0), neighbor_cache=cache)
    batch.step(3, fuse=True)
    fields = ("u", "v", "stressN", "statev")
    got = {name: batch.split(name) for name in fields}
    for s, sp in enumerate(sps):
        rowptr, col = cache[(sp.nx, sp.ny, sp.nz, sp.spacing, sp.horizon)]
        bar = vpd.Bar(props=JC_PROPS, density=RHO, damage=(0.02, 0.0), rowptr=rowptr, col=col,
                      **sp.bar_kwargs())
        bar.step(3)
        for name in fields:
            assert np.array_equal(got[name][s], bar.get(name)), (sp.label, name)
        bar.ctx.close()


def test_voided_2m_bar_invariants(vpd):
    banner("voided 400x70x70 bar: balance, determinism, gather forms")
    import impact
    kw = dict(return_map="bisection", damage=(0.004, 0.0))
    a = impact.build(**kw)
    n = a.n
    assert 1_800_000 < n < 1_900_000
    a.ctx.set_tuning(defgrad_window=0, force_window=0)  # direct gathers
    a.step(6, want_fields=True)
    f = a.get("force")
    fv = f * a.vol[:, None]
    assert np.abs(fv.sum(0)).max() <= 1e-9 * np.abs(fv).sum()
    ua, da = a.get("u"), a.get("damage")
    assert np.isfinite(ua).all()
    del a
    for window in (512, 640):
        b = impact.build(**kw)
        b.ctx.set_tuning(defgrad_window=window, force_window=window, defgrad_window_plan=3,
                         force_window_plan=3, defgrad_window_bonds=0, force_window_bonds=0)
        b.step(6, want_fields=True)
        assert np.array_equal(b.get("u"), ua), window
        assert np.array_equal(b.get("damage"), da), window
        del b
