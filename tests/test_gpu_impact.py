"""The high-rate tension case on a voided lattice with adiabatic heating, on the GPU.

* A voided 30×8×8 bar (5% of the sites removed, seeded), 20 mm/ms, JC with heating in
  consistent units (workToHeat = 1e6), critical-stretch damage: the GPU's parity path
  against the oracle on the same point set, broken-bond sets identical, every state field
  within 1e-9 relative per step.
* The learned return map on the same bar against the FP64 tight bisection (both finish at
  |r| ≤ 1e-12·q), with the logged fallback rate.
"""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

JC_HEAT = [200000.0, 0.3, 350.0, 275.0, 0.022, 0.36, 1.0, 1.0e-3, 293.0, 1793.0, 7850.0, 0.9,
           500.0, 1.0e6]
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
    E, nu = JC_HEAT[0], JC_HEAT[1]
    return fraction * h / math.sqrt(E * (1 - nu) / (RHO * (1 + nu) * (1 - 2 * nu)))


def bar_kw(vpd):
    points = vpd.voided_patch(30, 8, 8, 1.0, 0.05, 11)
    return dict(nx=30, ny=8, nz=8, spacing=1.0, horizon=3.015, props=JC_HEAT, dt=cfl_dt(1.0),
                density=RHO, velocity=20.0, damage=(0.003, 0.0), points=points)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_voided_heated_bar_vs_oracle(vpd, orc):
    banner("voided heated bar vs oracle, 80 steps")
    kw = bar_kw(vpd)
    g = vpd.Bar(**kw)
    o = orc.OracleBar(amplitude=0.0, rise_time=1.0, **kw)
    rp, col = g.ctx.get_neighbors()
    assert np.array_equal(rp, o.rowptr) and np.array_equal(col, o.col)
    for _ in range(4):
        o.step(20)
        g.step(20, want_fields=True)
        assert np.array_equal(g.get("damage"), o.damage)
        for name, ov in (("u", o.u), ("v", o.v), ("stressN", o.mat["stressN"]),
                         ("temp", o.mat["temp"])):
            assert rel(g.get(name), ov) <= 1e-9 * g.step_index, name
        assert rel(g.get("statev")[:, :5], o.mat["statev"][:, :5]) <= 1e-9 * g.step_index
    rise = o.mat["statev"][:, 2].max() - 293.0
    print("max temperature rise %.4f K, broken bonds %d" % (rise, int(o.damage.sum())))
    assert rise > 0.01          # adiabatic heating is active
    assert o.damage.sum() > 0   # and so is damage


def test_voided_learned_vs_tight(vpd):
    banner("voided heated bar: learned return map vs tight bisection")
    import impact
    kw = bar_kw(vpd)
    net = vpd.load_warm_start()
    tight = vpd.Bar(return_map=vpd.RETURN_MAP_BISECTION_TIGHT, **kw)
    learned = vpd.Bar(return_map=vpd.RETURN_MAP_LEARNED, warm_start=net, **kw)
    logs = []
    for _ in range(4):
        tight.step(20, want_fields=True)
        learned.step(20, want_fields=True)
        logs.append(impact.log_record(learned, learned.step_index))
        for name in ("u", "stressN"):
            assert rel(learned.get(name), tight.get(name)) <= 1e-8, name
        assert rel(learned.get("statev")[:, 1], tight.get("statev")[:, 1]) <= 1e-8
    print(logs)
    it_l = learned.get("iterations")
    it_t = tight.get("iterations")
    plastic = learned.get("statev")[:, 4] > 0
    assert plastic.sum() > 20
    assert it_l[plastic].mean() < it_t[plastic].mean() / 3
    assert all(0.0 <= r["fallback_rate"] <= 1.0 for r in logs)
