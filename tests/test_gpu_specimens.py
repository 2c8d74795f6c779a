"""Batched specimens (vpd_set_specimens / vpd_step_specimens) on the GPU.

The batch replaces the reference's process-per-case sweep (benchmarks/tensile3d/
run_cases.py:95-96): specimens of different lattice, horizon, time step, end velocity and
step count share one particle block and one launch per pass. Each specimen's end state must
equal the same specimen run alone (vpd.Bar through vpd_step) bit for bit — the same rows in
the same order, the same per-particle dt and load band, the same arithmetic — and one
specimen is checked against the oracle as well, at the per-step bound of the bar tests.
"""
import math
import sys

import numpy as np
import pytest

from conftest import JC_PROPS

pytestmark = pytest.mark.gpu

RHO = 7.85e-3
FIELDS = ("u", "v", "a", "stressN", "statev", "fN", "rN", "eint", "einel", "temp", "kinv")


def banner(msg):
    print("[gpu-case]", msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def vpd():
    import vpd as V
    if V.device_count() == 0:
        pytest.fail("no HIP device visible to the GPU tests")
    return V


def cfl_dt(h, fraction):
    E, nu = JC_PROPS[0], JC_PROPS[1]
    return fraction * h / math.sqrt(E * (1 - nu) / (RHO * (1 + nu) * (1 - 2 * nu)))


def specimens(vpd):
    S = vpd.Specimen
    return [
        # lattice, Δx, δ/Δx, dt, end velocity (mm/ms), steps
        S(30, 8, 8, 1.0, 3.015, cfl_dt(1.0, 0.5), 60.0, 70),
        S(20, 6, 6, 0.5, 2.015, cfl_dt(0.5, 0.9), 30.0, 45),
        S(16, 5, 5, 0.7, 4.015, cfl_dt(0.7, 0.25), -40.0, 90),   # compression
        S(12, 4, 4, 1.0, 3.015, cfl_dt(1.0, 0.5), 80.0, 20),     # finishes early
        S(9, 9, 3, 0.3, 3.015, cfl_dt(0.3, 0.5), 25.0, 55, component=1),  # loaded along y
    ]


def alone(vpd, sp, damage, return_map=0, steps=None):
    bar = vpd.Bar(props=JC_PROPS, density=RHO, damage=damage, return_map=return_map,
                  **sp.bar_kwargs())
    bar.step(sp.steps if steps is None else steps)
    return bar


def assert_same(batch, s, bar, what):
    for name in FIELDS + ("damage",) if bar.ctx.nnz else FIELDS:
        got, want = batch.get(name, s), bar.get(name)
        assert got.shape == want.shape, (what, name)
        assert np.array_equal(got, want), (what, name, np.abs(got - want).max())


@pytest.mark.parametrize("return_map,windows", [(0, 0), (2, 0), (0, 80), (2, 128), (0, 96),
                                                (0, 384), (2, 16)],
                         ids=["bisection", "newton", "bisection-windowed",
                              "newton-windowed-limited", "bisection-planned-prefetch",
                              "bisection-group-windows", "newton-shared-rows"])
def test_batch_equals_specimens_alone(vpd, return_map, windows):
    banner("batched specimens vs each alone, return map %d, windows %d" % (return_map, windows))
    damage = (0.003, 0.0)
    sps = specimens(vpd)
    batch = vpd.SpecimenBatch(sps, JC_PROPS, RHO, damage=damage, return_map=return_map)
    take = {128: 7, 96: 6}.get(windows, 0)
    form = {96: 2, 384: 3, 16: 4}.get(windows, 0)
    batch.ctx.set_tuning(defgrad_window=windows, force_window=windows,
                         defgrad_window_bonds=take, force_window_bonds=take,
                         defgrad_window_plan=form, force_window_plan=form)
    assert batch.n == sum(sp.n for sp in sps)
    batch.step()
    assert batch.step_index == max(sp.steps for sp in sps)
    broke = 0
    for s, sp in enumerate(sps):
        bar = alone(vpd, sp, damage, return_map)
        assert np.array_equal(batch.x[batch.first[s]:batch.first[s + 1]], bar.x)
        assert_same(batch, s, bar, "specimen %d" % s)
        broke += int(bar.get("damage").sum() > 0)
        assert bar.get("statev")[:, 1].max() > 0  # every specimen reached the plastic range
        bar.ctx.close()
    assert broke >= 2  # the damage path is exercised


def test_batch_in_chunks_and_reset(vpd):
    banner("batched specimens: stepped in chunks, then reset and rerun")
    sps = specimens(vpd)[:4]
    batch = vpd.SpecimenBatch(sps, JC_PROPS, RHO, damage=(0.003, 0.0))
    batch.step(7)           # unfused at 7, fused chain resumes at 8
    batch.step(13)          # specimen 3 (20 steps) finishes inside this chunk
    batch.step()
    first = {s: {n: batch.get(n, s) for n in ("u", "statev", "damage")} for s in range(4)}
    for s, sp in enumerate(sps):
        bar = alone(vpd, sp, (0.003, 0.0))
        assert_same(batch, s, bar, "chunked %d" % s)
        bar.ctx.close()
    batch.reset()
    batch.step()
    for s in range(4):
        for n, v in first[s].items():
            assert np.array_equal(batch.get(n, s), v), (s, n)


def test_batch_specimen_vs_oracle(vpd, orc):
    banner("one batched specimen vs the oracle")
    sps = specimens(vpd)[:2]
    batch = vpd.SpecimenBatch(sps, JC_PROPS, RHO, damage=(0.003, 0.0))
    batch.step()
    sp = sps[1]
    o = orc.OracleBar(nx=sp.nx, ny=sp.ny, nz=sp.nz, spacing=sp.spacing, horizon=sp.horizon,
                      props=JC_PROPS, dt=sp.dt, density=RHO, amplitude=0.0, rise_time=1.0,
                      damage=(0.003, 0.0), velocity=sp.velocity)
    o.step(sp.steps)
    assert np.array_equal(batch.get("damage", 1), o.damage)
    for name, ov in (("u", o.u), ("v", o.v), ("stressN", o.mat["stressN"]),
                     ("statev", o.mat["statev"])):
        gv = batch.get(name, 1)
        if name == "statev":
            gv, ov = gv[:, :5], ov[:, :5]
        err = np.abs(gv - ov).max() / max(np.abs(ov).max(), 1e-300)
        assert err <= 1e-9 * sp.steps, (name, err)


def test_batch_errors(vpd):
    banner("batched specimens: refused uses")
    sp = specimens(vpd)[3]
    bar = vpd.Bar(props=JC_PROPS, density=RHO, **sp.bar_kwargs())
    n = bar.n
    with pytest.raises(vpd.VpdError, match="bonded"):  # one bar cut in two shares bonds
        bar.ctx.set_specimens([0, n // 2, n], [sp.dt] * 2, [1.0] * 2, [5, 5])
    with pytest.raises(vpd.VpdError, match="cover the particle block"):
        bar.ctx.set_specimens([0, n - 1], [sp.dt], [1.0], [5])
    with pytest.raises(vpd.VpdError, match="positive"):
        bar.ctx.set_specimens([0, n], [0.0], [1.0], [5])
    with pytest.raises(vpd.VpdError, match="at least one step"):
        bar.ctx.set_specimens([0, n], [sp.dt], [1.0], [0])
    with pytest.raises(vpd.VpdError, match="No specimen table"):
        bar.ctx.step_specimens(1)
    bar.ctx.set_specimens([0, n], [sp.dt], [sp.velocity], [5])
    with pytest.raises(vpd.VpdError, match="vpd_step_specimens"):
        bar.step(1)
    # Diagnostics are laid out for the specimen table in place when they are enabled: the
    # table cannot change under them.
    bar.ctx.enable_diagnostics(0, 1)
    with pytest.raises(vpd.VpdError, match="before enabling run diagnostics"):
        bar.ctx.set_specimens([0, n], [sp.dt], [sp.velocity], [5])
    with pytest.raises(vpd.VpdError, match="Disable run diagnostics"):
        bar.ctx.set_specimens([], [], [], [])
    bar.ctx.enable_diagnostics(None)
    bar.ctx.step_specimens(1, fuse_next=True)
    with pytest.raises(vpd.VpdError, match="prediction was made for step 2"):
        bar.ctx.step_specimens(3)
    for k in range(2, 6):
        bar.ctx.step_specimens(k, fuse_next=True)  # the last step is never followed by one
    u_batch = bar.get("u")
    bar.ctx.set_specimens([], [], [], [])  # back to one specimen: the scalar API works again
    ref = vpd.Bar(props=JC_PROPS, density=RHO, **sp.bar_kwargs())
    ref.step(5)
    assert np.array_equal(u_batch, ref.get("u"))


@pytest.mark.parametrize("return_map,planned", [(0, False), (2, False), (3, False), (0, True),
                                                (0, "group"), (0, "rows")],
                         ids=["bisection", "newton", "learned", "bisection-planned",
                              "bisection-group-windows", "bisection-shared-rows"])
def test_graph_replay_equals_eager(vpd, return_map, planned):
    """Every replay of the captured step against eager steps, with eager steps and host
    synchronisations between the replays (the graph replays with planned windows in the
    last case, the eager steps gather directly)."""
    banner("batched specimens: graph-replayed steps vs eager, return map %d, planned %s"
           % (return_map, planned))
    sps = specimens(vpd)[:4]
    net = vpd.load_warm_start() if return_map == vpd.RETURN_MAP_LEARNED else None
    kw = dict(damage=(0.003, 0.0), return_map=return_map, warm_start=net)
    eager = vpd.SpecimenBatch(sps, JC_PROPS, RHO, **kw)
    graph = vpd.SpecimenBatch(sps, JC_PROPS, RHO, **kw)
    if planned == "group":
        graph.ctx.set_tuning(defgrad_window=512, force_window=256, defgrad_window_plan=3,
                             force_window_plan=3)
    elif planned == "rows":
        graph.ctx.set_tuning(defgrad_window=64, force_window=32, defgrad_window_plan=4,
                             force_window_plan=4)
    elif planned:
        graph.ctx.set_tuning(defgrad_window=80, force_window=64, defgrad_window_bonds=7,
                             force_window_bonds=8, defgrad_window_plan=2, force_window_plan=1)
    # Segments: replayed by the graph (True) or stepped eagerly (False), each ending with an
    # unfused eager step on both sides so that the states can be compared.
    for replay, k in ((True, 9), (False, 4), (True, 1), (True, 12), (False, 2), (True, 30)):
        eager.step(k + 1)
        if replay:
            graph.run(k)
            graph.step(1)
        else:
            graph.step(k + 1)
        graph.ctx.synchronize()
        for name in ("u", "v", "a", "stressN", "statev", "damage"):
            assert np.array_equal(graph.get(name), eager.get(name)), (replay, k, name)
    graph.run()   # the rest, to every specimen's last step
    eager.step()
    assert graph.step_index == eager.step_index == max(sp.steps for sp in sps)
    for name in ("u", "v", "a", "stressN", "statev", "damage", "fN", "rN"):
        assert np.array_equal(graph.get(name), eager.get(name)), name


def test_velocity_band_closed_form(vpd):
    """The velocity-driven band on the GPU against its closed form after k fused steps: the
    band's axial u = v·(k·dt), v and a = 0 exactly, the support band at zero; and, for a batch
    of specimens, each specimen's band at its own dt, velocity and last step."""
    banner("velocity band closed form")
    v, dt, k = 1.5, 8.5377e-5, 30
    bar = vpd.Bar(16, 4, 4, 1.0, 3.015, JC_PROPS, dt, RHO, velocity=v)
    bar.step(k)
    codes = vpd.unpack_codes(bar.codes)[:, 0]
    band, support = codes == 2, codes == 1
    u, vel, acc = bar.get("u"), bar.get("v"), bar.get("a")
    assert np.array_equal(u[band, 0], np.full(band.sum(), v * (k * dt)))
    assert np.array_equal(vel[band, 0], np.full(band.sum(), v))
    assert not acc[band, 0].any()
    assert not u[support, 0].any() and not vel[support, 0].any()
    sps = specimens(vpd)
    batch = vpd.SpecimenBatch(sps, JC_PROPS, RHO)
    batch.step()
    for s, sp in enumerate(sps):
        ax = sp.component
        c = vpd.unpack_codes(batch.codes)[batch.first[s]:batch.first[s + 1], ax]
        us, vs = batch.get("u", s)[:, ax], batch.get("v", s)[:, ax]
        assert np.array_equal(us[c == 2], np.full((c == 2).sum(), sp.velocity * (sp.steps * sp.dt)))
        assert np.array_equal(vs[c == 2], np.full((c == 2).sum(), sp.velocity))
