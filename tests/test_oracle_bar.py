"""The restated step loop against the reference's own classes driven in the order of
BoundaryValueCaseRunner.cpp:207-229: bitwise equal state after plastic steps, with and
without critical-stretch damage, on the 30×8×8 JC bar (δ = 3.015Δx)."""
import numpy as np
import pytest

from conftest import JC_PROPS

FIELDS = ["u", "v", "a", "sigma", "force"]
MAT = ["fN", "vN", "rN", "stressN", "statev", "temp", "eint", "einel"]


@pytest.mark.parametrize("damage", [None, (0.002, 0.0)])
def test_bar_bitwise_vs_reference(ref, damage):
    kw = dict(nx=30, ny=8, nz=8, spacing=1.0, horizon=3.015, props=JC_PROPS, dt=8.5377e-5,
              density=7.85e-3, amplitude=2.0, rise_time=0.01, damage=damage)
    o = ref.OracleBar(**kw)
    r = ref.RefBar(**kw)
    rp, col = r.neighbors()
    assert np.array_equal(rp, o.rowptr) and np.array_equal(col, o.col)
    for _ in range(2):
        o.step(25)
        r.step(25)
        for name in FIELDS:
            assert np.array_equal(getattr(o, name), r.get(name)), name
        for name in MAT:
            assert np.array_equal(o.mat[name], r.get(name)), name
        if damage is not None:
            assert np.array_equal(o.damage, r.get("damage"))
    assert o.mat["statev"][:, 1].max() > 0.01  # the bar has yielded
    if damage is not None:
        assert o.damage.sum() > 0  # and bonds have broken
