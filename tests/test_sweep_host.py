"""The sweep and impact drivers' case definitions (no GPU work): the 48-bar grid of
vumat-pd_amd/sweep.py and the voided lattice of the high-rate case."""
import math

import numpy as np

import sweep


def test_sweep_grid():
    sps = sweep.specimens(end_steps=2000)
    assert len(sps) == 48
    assert len({sp.label for sp in sps}) == 48
    sizes = sorted({sp.n for sp in sps})
    assert sizes == [40 * 10 * 10, 80 * 20 * 20, 133 * 33 * 33, 200 * 50 * 50]
    # one end time for all: steps · dt equal to that of the finest (0.2 mm, 0.25·CFL) bar
    finest = max(sps, key=lambda sp: sp.steps)
    assert finest.steps == 2000 and finest.spacing == 0.2
    t_end = finest.steps * finest.dt
    for sp in sps:
        assert abs(sp.steps * sp.dt - t_end) <= 0.5 * sp.dt + 1e-15
        assert math.isclose(sp.horizon / sp.spacing, round(sp.horizon / sp.spacing, 3))
    # dt = fraction · Δx / c_p with the P-wave speed of the bar material
    E, nu, rho = sweep.JC_PROPS[0], sweep.JC_PROPS[1], sweep.RHO
    cp = math.sqrt(E * (1 - nu) / (rho * (1 + nu) * (1 - 2 * nu)))
    assert math.isclose(sweep.cfl_dt(0.5, 0.9), 0.9 * 0.5 / cp)


def test_voided_case_points_are_a_subset():
    import vpd
    x, vol = vpd.voided_patch(40, 7, 7, 0.2, 0.05, 20240607)
    xf, _ = vpd.make_patch(40, 7, 7, 0.2)
    assert len(x) == len(xf) - round(0.05 * len(xf))
    assert np.isin(x.view([("", x.dtype)] * 3), xf.view([("", xf.dtype)] * 3)).all()
