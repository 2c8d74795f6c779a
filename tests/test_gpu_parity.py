"""GPU parity of the HIP path against the oracle (and, where oracle/_ref is built, the
reference itself), through the C ABI. Integer work (neighbour CSR, broken-bond sets) must be
bit-exact; FP64 fields are compared at 1e-12 relative for single kernels (only sin/pow/log
may differ from libm, by about an ulp) and at 1e-9 relative per step over short runs."""
import sys

import numpy as np
import pytest

from conftest import JC_PROPS

pytestmark = pytest.mark.gpu

REL_KERNEL = 1e-12
REL_STEP = 1e-9


def banner(msg):
    print("[gpu-case]", msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def vpd():
    import vpd as V
    if V.device_count() == 0:
        pytest.fail("no HIP device visible to the GPU tests")
    return V


def rel_err(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / scale


def ctx_for(vpd, x, vol, horizon=3.015, props=JC_PROPS, damage=None, rowptr=None, col=None):
    c = vpd.Context(horizon, 7.85e-3, props, damage=damage, char_length=1.0)
    c.set_particles(x, vol)
    if rowptr is None:
        c.build_neighbors()
    else:
        c.set_neighbors(rowptr, col)
    return c


def lattice(n, h, shift):
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    return np.stack([shift + i.ravel() * h, shift + j.ravel() * h, shift + k.ravel() * h], 1)


@pytest.mark.parametrize("h", [2.0 / 3.0, 1.0, 2.0])
@pytest.mark.parametrize("shift", [0.0, -123.456, 10000.125])
def test_neighbors_lattice_bitwise(vpd, orc, h, shift):
    banner(f"neighbors lattice h={h} shift={shift}")
    x = lattice(9, h, shift)
    c = ctx_for(vpd, x, np.ones(len(x)), horizon=3 * h)
    rp, col = c.get_neighbors()
    rpo, colo = orc.build_neighbors(x, 3 * h)
    assert np.array_equal(rp, rpo) and np.array_equal(col, colo)


@pytest.mark.parametrize("seed", [0, 1])
def test_neighbors_jittered_bitwise(vpd, orc, seed):
    banner(f"neighbors jittered seed={seed}")
    rng = np.random.default_rng(seed)
    x, vol = orc.make_patch(20, 11, 7, 0.7)
    x = x + rng.uniform(-0.2, 0.2, x.shape)
    x = x[rng.permutation(len(x))]
    c = ctx_for(vpd, x, vol, horizon=3.015 * 0.7)
    rp, col = c.get_neighbors()
    rpo, colo = orc.build_neighbors(x, 3.015 * 0.7)
    assert np.array_equal(rp, rpo) and np.array_equal(col, colo)


def test_neighbors_edge_cases(vpd, orc):
    banner("neighbors edge cases")
    # An isolated point, a coincident pair and a far cluster: empty and ragged rows.
    x = np.array([[0.0, 0, 0], [100.0, 0, 0], [100.0, 0, 0], [100.5, 0, 0], [-50, 3, 1]])
    c = ctx_for(vpd, x, np.ones(5), horizon=1.0)
    rp, col = c.get_neighbors()
    rpo, colo = orc.build_neighbors(x, 1.0)
    assert np.array_equal(rp, rpo) and np.array_equal(col, colo)
    assert rp[1] == 0
    bad = x.copy()
    bad[2, 1] = np.inf
    c2 = vpd.Context(1.0, 1.0, JC_PROPS)
    c2.set_particles(bad, np.ones(5))
    with pytest.raises(vpd.VpdError, match="Nonfinite"):
        c2.build_neighbors()


def test_neighbors_full_size_closed_form(vpd):
    # The 500k-particle bar (200×50×50, h = 0.2, δ = 3.015h): every row's degree equals the
    # number of lattice offsets with 0 < |o|² ≤ 9 that stay inside the grid, rows are
    # strictly ascending, and each entry is such an offset.
    banner("neighbors 500k closed form")
    nx, ny, nz, h = 200, 50, 50, 0.2
    x, vol = vpd.make_patch(nx, ny, nz, h)
    c = ctx_for(vpd, x, vol, horizon=3.015 * h)
    rp, col = c.get_neighbors()
    g = np.stack(np.unravel_index(np.arange(len(x)), (nx, ny, nz)), 1)
    offs = np.array([(a, b, d) for a in range(-3, 4) for b in range(-3, 4) for d in range(-3, 4)
                     if 0 < a * a + b * b + d * d <= 9])
    assert len(offs) == 122
    deg = np.zeros(len(x), np.int64)
    for o in offs:
        ok = np.all((g + o >= 0) & (g + o < [nx, ny, nz]), 1)
        deg += ok
    assert np.array_equal(np.diff(rp), deg)
    assert rp[-1] == 57894248
    row = np.repeat(np.arange(len(x)), np.diff(rp))
    same_row = row[1:] == row[:-1]
    assert np.all(col[1:][same_row] > col[:-1][same_row])
    d = g[col] - g[row]
    assert np.all((d * d).sum(1) <= 9) and np.all((d * d).sum(1) > 0)


def bar_setup(orc, nx=12, ny=6, nz=5, h=1.0, horizon=3.015):
    x, vol = orc.make_patch(nx, ny, nz, h)
    rp, col = orc.build_neighbors(x, horizon)
    return x, vol, rp, col


def test_shape_and_defgrad_bitwise(vpd, orc):
    banner("shape tensor + deformation gradient")
    x, vol, rp, col = bar_setup(orc)
    c = ctx_for(vpd, x, vol, rowptr=rp, col=col)
    c.initialize_shape()
    kinv = c.download("kinv")
    assert np.array_equal(kinv, orc.shape_inverse(x, vol, rp, col))
    rng = np.random.default_rng(3)
    u = 0.01 * rng.standard_normal(x.shape)
    c.upload("u", u)
    c.compute_deformation_gradient()
    F = c.download("F")
    assert np.array_equal(F, orc.deformation_gradient(x, u, vol, rp, col, kinv))


def test_singular_shape_raises(vpd):
    banner("singular shape")
    x = np.array([[0.0, 0, 0], [1.0, 0, 0], [2.0, 0, 0]])
    c = ctx_for(vpd, x, np.ones(3), horizon=2.5)
    with pytest.raises(vpd.VpdError, match="Singular tensor in PD shape tensor"):
        c.initialize_shape()


def random_jc_block(n, seed):
    rng = np.random.default_rng(seed)
    inc = rng.standard_normal((6, n)) * 2e-3
    stress = rng.standard_normal((6, n)) * 150.0
    state = np.zeros((6, n))
    state[0] = (rng.random(n) > 0.1)
    state[1] = rng.random(n) * 0.05
    state[2] = 293.0 + rng.random(n) * 50
    state[3] = rng.random(n) * 400.0
    state[4] = rng.random(n) * 1e-3
    state[5] = rng.integers(1, 50, n)
    return inc, stress, state


def test_vumat_device_abi_vs_oracle(vpd, orc):
    banner("device VUMAT ABI")
    n = 5000
    inc, stress, state = random_jc_block(n, 9)
    temp, e0, e1 = np.full(n, 300.0), np.zeros(n), np.zeros(n)
    dt = 8.5377e-5
    ref = orc.vumat_jc(JC_PROPS, dt, inc, stress, state, temp, e0, e1)
    got = vpd.vumat_device(JC_PROPS, dt, inc, stress, state, temp, e0, e1)
    plastic = ref["state"][4] > 0
    assert plastic.sum() > 1000
    assert rel_err(got["stress"], ref["stress"]) <= REL_KERNEL
    assert rel_err(got["state"][:5], ref["state"][:5]) <= REL_KERNEL
    assert rel_err(got["eint"], ref["eint"]) <= REL_KERNEL
    assert rel_err(got["einel"], ref["einel"]) <= REL_KERNEL
    # The bisection path (iteration counts on state 6) agrees point by point.
    assert (got["state"][5] == ref["state"][5]).mean() >= 0.999
    assert np.array_equal(got["temp"], got["state"][2])


def test_material_update_vs_oracle(vpd, orc):
    banner("material update (FT + JC + rotation + Piola)")
    x, vol, rp, col = bar_setup(orc)
    c = ctx_for(vpd, x, vol, rowptr=rp, col=col)
    c.initialize_shape()
    c.initialize_material()
    n = len(x)
    rng = np.random.default_rng(5)
    st = orc.material_init(c.download("fN"), 6, 293.0)
    dt = 1e-4
    for k in range(4):
        F = np.eye(3).ravel() + (k + 1) * 2e-3 * rng.standard_normal((n, 9))
        sig_o = orc.material_update(F, st, dt, np.asarray(JC_PROPS))
        c.upload("F", F)
        c.update_material(dt)
        assert rel_err(c.download("sigma"), sig_o) <= REL_KERNEL * 10
        for name, key in [("fN", "fN"), ("vN", "vN"), ("rN", "rN"), ("stressN", "stressN"),
                          ("eint", "eint"), ("einel", "einel"), ("temp", "temp")]:
            assert rel_err(c.download(name), st[key]) <= REL_KERNEL * 10, name
        assert rel_err(c.download("statev")[:, :5], st["statev"][:, :5]) <= REL_KERNEL * 10
        assert rel_err(c.download("strain"), orc.green_lagrange(F)) <= 1e-15 * 10
    assert (st["statev"][:, 4] > 0).sum() > 10


def run_pair(vpd, orc, steps, damage=None, chunks=4, nx=30, ny=8, nz=8, amplitude=2.0,
             rise=0.01):
    kw = dict(nx=nx, ny=ny, nz=nz, spacing=1.0, horizon=3.015, props=JC_PROPS, dt=8.5377e-5,
              density=7.85e-3, amplitude=amplitude, rise_time=rise, damage=damage)
    o = orc.OracleBar(**kw)
    g = vpd.Bar(**kw)
    rp, col = g.ctx.get_neighbors()
    assert np.array_equal(rp, o.rowptr) and np.array_equal(col, o.col)
    assert np.array_equal(g.get("kinv"), o.kinv)
    worst = 0.0
    for _ in range(chunks):
        o.step(steps)
        g.step(steps, want_fields=True)
        for name, ov in [("u", o.u), ("v", o.v), ("sigma", o.sigma), ("force", o.force),
                         ("stressN", o.mat["stressN"]), ("rN", o.mat["rN"]),
                         ("statev", o.mat["statev"][:, :5])]:
            gv = g.get(name)
            if name == "statev":
                gv = gv[:, :5]
            worst = max(worst, rel_err(gv, ov))
        if damage is not None:
            assert np.array_equal(g.get("damage"), o.damage)
    return o, g, worst


def test_bar_steps_vs_oracle(vpd, orc):
    banner("bar 30x8x8, 100 fused steps vs oracle")
    o, g, worst = run_pair(vpd, orc, 25)
    print("worst relative difference", worst)
    assert worst <= REL_STEP * 100  # 100 steps, ≤ 1e-9 per step
    assert o.mat["statev"][:, 1].max() > 0.01


def test_bar_damage_vs_oracle(vpd, orc):
    banner("bar 30x8x8 with critical-stretch damage vs oracle")
    o, g, worst = run_pair(vpd, orc, 25, damage=(0.002, 0.0))
    assert worst <= REL_STEP * 100
    assert o.damage.sum() > 0


def test_bar_vs_reference_build(vpd, ref):
    banner("bar vs oracle/_ref (reference classes)")
    kw = dict(nx=30, ny=8, nz=8, spacing=1.0, horizon=3.015, props=JC_PROPS, dt=8.5377e-5,
              density=7.85e-3, amplitude=2.0, rise_time=0.01, damage=(0.002, 0.0))
    r = ref.RefBar(**kw)
    g = vpd.Bar(**kw)
    for _ in range(2):
        r.step(30)
        g.step(30, want_fields=True)
        for name in ["u", "v", "sigma", "stressN"]:
            assert rel_err(g.get(name), r.get(name)) <= REL_STEP * 60, name
        assert np.array_equal(g.get("damage"), r.get("damage"))


def test_defgrad_fields_after_swapped_steps(vpd):
    # A plain step leaves F_{n+1} where F_n was (the two buffers swap instead of the material
    # stage copying): the current F and the state's F_n read alike. A later F pass, or an
    # uploaded F, goes to the other buffer and leaves F_n as it was.
    banner("deformation-gradient fields across swapped steps")
    b = vpd.Bar(nx=10, ny=4, nz=4, spacing=1.0, horizon=2.015, props=JC_PROPS, dt=8.5377e-5,
                density=7.85e-3, velocity=300.0)
    b.step(5)
    F, FN = b.get("F"), b.get("fN")
    assert np.array_equal(F, FN)
    assert np.abs(FN - np.eye(3).reshape(9)[None, :]).max() > 0  # deformed
    b.ctx.compute_deformation_gradient()
    assert np.array_equal(b.get("fN"), FN)
    assert np.array_equal(b.get("F"), F)  # the same u as the step's F pass
    up = FN * 1.001
    b.ctx.upload("F", up)
    assert np.array_equal(b.get("F"), up) and np.array_equal(b.get("fN"), FN)
    b.step(3)
    assert np.array_equal(b.get("F"), b.get("fN"))
    # A material update right after a step (no new F) updates from the latest F, as before
    # the buffers swapped: a zero increment, so the stress stays where the step left it (up
    # to the return map's tolerance; the stale F_n would undo a whole step's increment, about
    # E·2.5e-3 here).
    c = vpd.Bar(nx=10, ny=4, nz=4, spacing=1.0, horizon=2.015, props=JC_PROPS, dt=8.5377e-5,
                density=7.85e-3, velocity=300.0)
    c.step(4)
    sig, fn = c.get("stressN"), c.get("fN")
    c.ctx.update_material(8.5377e-5)
    assert np.array_equal(c.get("fN"), fn)
    assert np.abs(c.get("stressN") - sig).max() <= 1e-4 * np.abs(sig).max()


def test_fused_equals_unfused(vpd, orc):
    banner("fused predict == separate predict")
    kw = dict(nx=16, ny=6, nz=6, spacing=1.0, horizon=3.015, props=JC_PROPS, dt=8.5377e-5,
              density=7.85e-3, amplitude=1.0, rise_time=0.01)
    a = vpd.Bar(**kw)
    b = vpd.Bar(**kw)
    a.step(40, fuse=True)
    b.step(40, fuse=False)
    for name in ["u", "v", "a", "stressN", "statev", "rN", "vN"]:
        assert np.array_equal(a.get(name), b.get(name)), name


@pytest.mark.parametrize("knobs", [
    dict(defgrad_rounds=3, force_rounds=5),
    dict(defgrad_lds=36408, force_lds=64000, defgrad_rounds=2, force_rounds=2),
    dict(defgrad_window=64, force_window=80),
    dict(defgrad_window=128, force_window=96, defgrad_rounds=2, force_rounds=3,
         force_lds=64000),
    dict(defgrad_window=128, force_window=160, defgrad_window_bonds=8, force_window_bonds=7),
    dict(defgrad_window=64, force_window=80, defgrad_window_bonds=1, force_window_bonds=3),
    dict(defgrad_window=128, force_window=80, defgrad_window_bonds=6, force_window_bonds=6,
         defgrad_window_plan=1, force_window_plan=1),
    dict(defgrad_window=64, force_window=96, defgrad_window_bonds=8, force_window_bonds=5,
         defgrad_window_plan=2, force_window_plan=2, defgrad_rounds=3, force_rounds=2),
    dict(defgrad_window=80, force_window=128, defgrad_window_bonds=1, force_window_bonds=2,
         defgrad_window_plan=2, force_window_plan=1, force_lds=64000),
    dict(defgrad_window=512, force_window=480, defgrad_window_plan=3, force_window_plan=3,
         defgrad_rounds=2, force_rounds=3),
    dict(defgrad_window=256, force_window=256, defgrad_window_plan=3, force_window_plan=3,
         defgrad_lds=64000, force_lds=50000),
    dict(defgrad_window=640, force_window=512, defgrad_window_plan=3, force_window_plan=3),
    dict(defgrad_window=16, force_window=16, defgrad_window_plan=4, force_window_plan=4),
    dict(defgrad_window=32, force_window=512, defgrad_window_plan=4, force_window_plan=3,
         defgrad_lds=64000),
    dict(defgrad_window=80, force_window=64, defgrad_window_bonds=7, defgrad_window_plan=2,
         force_window_plan=4, force_lds=36408),
    dict(defgrad_window=64, force_window=32, defgrad_window_plan=4, force_window_plan=4),
])
def test_launch_tuning_is_bitwise_neutral(vpd, knobs):
    # Launch knobs change how the bond passes are scheduled, never what they sum or in which
    # order: every field equals the default launch's bit for bit, damage included.
    banner("launch tuning %s" % knobs)
    kw = dict(nx=40, ny=10, nz=10, spacing=1.0, horizon=3.015, props=JC_PROPS, dt=8.5377e-5,
              density=7.85e-3, velocity=400.0, damage=(0.002, 0.0))
    a = vpd.Bar(**kw)
    b = vpd.Bar(**kw)
    b.ctx.set_tuning(**knobs)
    assert all(b.ctx.get_tuning()[k] == v for k, v in knobs.items())
    a.step(60)
    b.step(60)
    assert a.get("damage").sum() > 0
    for name in ["u", "v", "a", "stressN", "statev", "damage"]:
        assert np.array_equal(a.get(name), b.get(name)), name


@pytest.mark.parametrize("voids,take,plan", [(0.0, 0, 0), (0.08, 0, 0), (0.08, 5, 0), (0.0, 2, 0),
                                            (0.08, 6, 1), (0.08, 7, 2), (0.0, 8, 2), (0.08, 0, 3),
                                            (0.0, 0, 3), (0.08, 0, 4)])
def test_windowed_gathers_on_irregular_rows(vpd, voids, take, plan):
    # Voids make every row irregular: the windows then cover fewer bonds per step, and the
    # results must still equal the direct gathers bit for bit, with or without a limit on the
    # bonds a particle takes from one window, walked or replayed from a window plan.
    banner("windowed gathers, void fraction %g, bonds per window %d, plan form %d"
           % (voids, take, plan))
    points = vpd.voided_patch(24, 9, 7, 1.0, voids, 5)
    kw = dict(nx=24, ny=9, nz=7, spacing=1.0, horizon=3.015, props=JC_PROPS, dt=8.5377e-5,
              density=7.85e-3, velocity=300.0, damage=(0.002, 0.0), points=points)
    a = vpd.Bar(**kw)
    b = vpd.Bar(**kw)
    window = {3: 256, 4: 16}.get(plan, 64)
    b.ctx.set_tuning(defgrad_window=window, force_window=window, defgrad_window_bonds=take,
                     force_window_bonds=take, defgrad_window_plan=plan, force_window_plan=plan)
    a.step(50, want_fields=True)
    b.step(50, want_fields=True)
    assert a.get("damage").sum() > 0
    for name in ["u", "v", "a", "stressN", "statev", "damage", "force", "F"]:
        assert np.array_equal(a.get(name), b.get(name)), name


FORMS = {
    "walked": dict(defgrad_window=64, force_window=64),
    "planned-1": dict(defgrad_window=64, force_window=64, defgrad_window_bonds=1,
                      force_window_bonds=1, defgrad_window_plan=1, force_window_plan=1),
    "planned-prefetch": dict(defgrad_window=80, force_window=96, defgrad_window_bonds=3,
                             force_window_bonds=8, defgrad_window_plan=2, force_window_plan=2),
    "group": dict(defgrad_window=384, force_window=256, defgrad_window_plan=3,
                  force_window_plan=3),
    "rows": dict(defgrad_window=16, force_window=16, defgrad_window_plan=4, force_window_plan=4),
    "rows-64": dict(defgrad_window=64, force_window=32, defgrad_window_plan=4,
                    force_window_plan=4),
}


@pytest.mark.parametrize("shape", [(5, 3, 3), (11, 3, 2), (13, 4, 3)],
                         ids=["45-particles", "66-particles", "156-particles"])
def test_gather_forms_on_small_blocks(vpd, shape):
    # Blocks smaller than a workgroup, and ones whose last wavefront is nearly empty: every
    # gather form equals the direct gathers bit for bit, damage included.
    banner("gather forms on a %dx%dx%d block" % shape)
    kw = dict(spacing=1.0, horizon=2.015, props=JC_PROPS, dt=8.5377e-5, density=7.85e-3,
              velocity=300.0, damage=(0.002, 0.0))
    ref = vpd.Bar(*shape, **kw)
    ref.step(40, want_fields=True)
    assert ref.get("damage").sum() > 0
    for name, knobs in FORMS.items():
        b = vpd.Bar(*shape, **kw)
        b.ctx.set_tuning(**knobs)
        b.step(40, want_fields=True)
        for field in ["u", "v", "stressN", "statev", "damage", "force", "F"]:
            assert np.array_equal(b.get(field), ref.get(field)), (name, field)


def test_gather_forms_on_unsorted_rows(vpd):
    # Bond lists supplied with every row shuffled: each lane still sums its row in the order
    # given, so every gather form equals the direct gathers bit for bit (a window takes a
    # bond only while its neighbour lies inside the window, also when it lies below the
    # window's start).
    banner("gather forms on shuffled bond lists")
    kw = dict(nx=14, ny=6, nz=5, spacing=1.0, horizon=3.015, props=JC_PROPS, dt=8.5377e-5,
              density=7.85e-3, velocity=300.0, damage=(0.002, 0.0))
    rowptr, col = vpd.Bar(**kw).ctx.get_neighbors()
    col = col.copy()
    rng = np.random.default_rng(11)
    for i in range(len(rowptr) - 1):
        rng.shuffle(col[rowptr[i]:rowptr[i + 1]])
    assert any(np.any(np.diff(col[rowptr[i]:rowptr[i + 1]]) < 0) for i in range(len(rowptr) - 1))
    runs = {}
    for name, knobs in [("direct", {})] + list(FORMS.items()):
        b = vpd.Bar(**kw)
        b.ctx.set_neighbors(rowptr, col)
        b.ctx.initialize_shape()
        b.reset()
        b.ctx.set_tuning(**knobs)
        b.step(30, want_fields=True)
        runs[name] = b
    ref = runs.pop("direct")
    assert ref.get("damage").sum() > 0
    for name, b in runs.items():
        for field in ["u", "v", "stressN", "statev", "damage", "force", "F"]:
            assert np.array_equal(b.get(field), ref.get(field)), (name, field)


def test_plan_follows_knob_and_neighbor_changes(vpd):
    # The window plan is rebuilt when the knobs change between steps and when the bond lists
    # are replaced: the run stays equal to one with direct gathers.
    banner("window plan rebuilt after knob and bond-list changes")
    kw = dict(nx=16, ny=6, nz=5, spacing=1.0, horizon=3.015, props=JC_PROPS, dt=8.5377e-5,
              density=7.85e-3, velocity=300.0, damage=(0.002, 0.0))
    ref, b = vpd.Bar(**kw), vpd.Bar(**kw)
    b.ctx.set_tuning(**FORMS["planned-1"])
    for knobs in (dict(), dict(force_window_bonds=5, defgrad_window=128, defgrad_window_bonds=8),
                  dict(force_window_plan=2, defgrad_window_plan=2),
                  dict(force_window=512, force_window_bonds=0, force_window_plan=3)):
        b.ctx.set_tuning(**knobs)
        ref.step(8)
        b.step(8)
        for field in ["u", "v", "stressN", "damage"]:
            assert np.array_equal(b.get(field), ref.get(field)), (knobs, field)
    # Bonds broken from the host (the flags and the per-particle marks both come from the
    # upload): the planned force pass must skip them as the direct one does.
    d = np.zeros(ref.ctx.nnz)
    d[::7] = 1.0
    for bar in (ref, b):
        bar.reset()
        bar.ctx.upload("damage", d)
        bar.step(6)
    for field in ["u", "v", "stressN", "damage"]:
        assert np.array_equal(b.get(field), ref.get(field)), ("uploaded damage", field)
    # New bond lists (a smaller horizon's) on the planned context, then a fresh start.
    small = vpd.Bar(**dict(kw, horizon=2.015))
    rowptr, col = small.ctx.get_neighbors()
    b.ctx.set_neighbors(rowptr, col)
    b.ctx.initialize_shape()
    b.reset()
    small.step(20)
    b.step(20)
    for field in ["u", "v", "stressN", "statev", "damage"]:
        assert np.array_equal(b.get(field), small.get(field)), field


def test_sampled_kernel_timing(vpd):
    # Timing stride k: the kernels of every k-th step carry events, counted from the first
    # step after the call.
    b = vpd.Bar(nx=8, ny=4, nz=4, spacing=1.0, horizon=2.015, props=JC_PROPS, dt=8.5377e-5,
                density=7.85e-3, velocity=1.0)
    with pytest.raises(vpd.VpdError, match="timing stride"):
        b.ctx.set_timing(True, -2)
    for stride, timed in ((1, 9), (3, 3), (4, 3)):
        b.ctx.set_timing(True, stride)
        b.step(9)
        kt = b.ctx.collect_timing()
        for name in ("defgrad_bond_pass", "material_jc", "force_bond_pass"):
            assert kt[name][1] == timed, (stride, name, kt[name])
    b.ctx.set_timing(False)
    b.step(2)
    assert b.ctx.collect_timing() == {}


def test_window_knob_limits(vpd):
    b = vpd.Bar(nx=6, ny=4, nz=4, spacing=1.0, horizon=3.015, props=JC_PROPS, dt=8.5377e-5,
                density=7.85e-3, velocity=1.0)
    for bad in (48, 100, 400):
        with pytest.raises(vpd.VpdError, match="window"):
            b.ctx.set_tuning(force_window=bad)
    b.ctx.set_tuning(force_window=0, defgrad_window=160)
    assert b.ctx.get_tuning()["defgrad_window"] == 160
    for bad in (-1, 1025):
        with pytest.raises(vpd.VpdError, match="Bonds per lane and window"):
            b.ctx.set_tuning(force_window_bonds=bad)
    b.ctx.set_tuning(force_window_bonds=1024, defgrad_window_bonds=6)
    t = b.ctx.get_tuning()
    assert (t["force_window_bonds"], t["defgrad_window_bonds"]) == (1024, 6)
    for bad in (-1, 5):
        with pytest.raises(vpd.VpdError, match="Window form"):
            b.ctx.set_tuning(force_window_plan=bad)
    # A planned form needs a planned window size and 1..8 bonds per window: the step refuses
    # other combinations, and runs once they are valid.
    for knobs in (dict(force_window=160, force_window_bonds=8), dict(force_window=80,
                                                                       force_window_bonds=0),
                  dict(force_window=80, force_window_bonds=9)):
        b.ctx.set_tuning(force_window_plan=2, **knobs)
        with pytest.raises(vpd.VpdError, match="Planned windows need"):
            b.step(1)
    b.ctx.set_tuning(force_window=80, force_window_bonds=8, force_window_plan=2)
    b.step(1)
    assert b.ctx.get_tuning()["force_window_plan"] == 2
    # Group windows need a group window size and no bond limit; walked windows a walked size.
    for knobs, msg in ((dict(force_window=256, force_window_bonds=8), "Group windows need"),
                       (dict(force_window=128, force_window_bonds=0), "Group windows need")):
        b.ctx.set_tuning(force_window_plan=3, **knobs)
        with pytest.raises(vpd.VpdError, match=msg):
            b.step(1)
    b.ctx.set_tuning(force_window=640, force_window_bonds=0, force_window_plan=3)
    b.step(1)  # 7 planes of 640 pairs: 70 KB of the workgroup's LDS
    b.ctx.set_tuning(force_window=384, force_window_bonds=0, force_window_plan=3,
                     defgrad_window=640, defgrad_window_bonds=0, defgrad_window_plan=3)
    b.step(1)
    # Shared rows go only with 16, 32 or 64 lanes per particle and no bond limit, and windows
    # 16 and 32 only with them.
    for knobs in (dict(force_window=16, force_window_plan=3), dict(force_window=32,
                                                                    force_window_plan=0),
                  dict(force_window=80, force_window_plan=4),
                  dict(force_window=16, force_window_plan=4, force_window_bonds=3)):
        b.ctx.set_tuning(**dict(dict(force_window_bonds=0), **knobs))
        with pytest.raises(vpd.VpdError, match="Shared rows"):
            b.step(1)
    for lanes in (16, 32, 64):
        b.ctx.set_tuning(force_window=lanes, force_window_bonds=0, force_window_plan=4)
        b.step(1)
    b.ctx.set_tuning(defgrad_window=160, defgrad_window_plan=0)
    b.ctx.set_tuning(force_window=384, force_window_plan=0)
    with pytest.raises(vpd.VpdError, match="Walked windows need"):
        b.step(1)
    b.ctx.set_tuning(force_window_plan=1, force_window_bonds=4)
    with pytest.raises(vpd.VpdError, match="Planned windows need"):
        b.step(1)
    b.ctx.set_tuning(force_window=0)
    b.step(1)


@pytest.mark.parametrize("mode,shape,plastic,rtol",
                         [(0, (30, 8, 8), 500, 1e-12), (1, (30, 8, 8), 500, 1e-12),
                          (0, (40, 16, 16), 4200, 1e-12), (1, (30, 8, 8), 500, 1e-300)],
                         ids=["parity-bisection", "tight-bisection", "long-list",
                              "iteration-cap"])
def test_grouped_plastic_list_equals_serial(vpd, mode, shape, plastic, rtol):
    # Small blocks run the bisections' plastic list on lane groups (several iterations per
    # round) while it is short, and on one lane per point past 4096 points (the long-list
    # case crosses over); with run diagnostics the material stage runs every point on one
    # lane. The two must agree bit for bit, iteration counts (state 6) included. With
    # rtol = 1e-300 the tight bisection never meets its tolerance and every plastic point
    # runs to the 200-iteration cap.
    banner("plastic list on lane groups vs one lane per point, return map %d, %s, rtol %g"
           % (mode, shape, rtol))
    kw = dict(spacing=1.0, horizon=3.015, props=JC_PROPS, dt=8.5377e-5, density=7.85e-3,
              velocity=60.0, damage=(0.003, 0.0), return_map=mode, rtol=rtol)
    grouped, serial = vpd.Bar(*shape, **kw), vpd.Bar(*shape, **kw)
    serial.ctx.enable_diagnostics(0, 1, capacity=200)
    for k in (40, 80):
        grouped.step(k)
        serial.step(k)
        for name in ("u", "v", "a", "stressN", "statev", "damage"):
            assert np.array_equal(grouped.get(name), serial.get(name)), (k, name)
        print("step %d: %d plastic points" % (grouped.step_index,
                                              (grouped.get("statev")[:, 4] > 0).sum()))
    st = grouped.get("statev")
    assert (st[:, 4] > 0).sum() > plastic and st[:, 5].max() > 100
    if rtol < 1e-100:  # state 6 sums the iterations: 200 per plastic update
        assert st[:, 5].max() >= 200 * 20


def test_autotune_and_in_step_confirmation(vpd):
    # The autotuner leaves the state as it was and sets its pick; its re-timed candidates per
    # pass come fastest first and include the pick. The in-step confirmation restores the
    # state bit for bit and sets one of those candidates; the run then equals an untuned one.
    banner("autotune, candidates and in-step confirmation on a damaged bar")
    kw = dict(nx=24, ny=8, nz=8, spacing=1.0, horizon=3.015, props=JC_PROPS, dt=8.5377e-5,
              density=7.85e-3, velocity=300.0, damage=(0.002, 0.0))
    ref, b = vpd.Bar(**kw), vpd.Bar(**kw)
    assert b.ctx.autotune_candidates(0) == [] and b.ctx.autotune_candidates(1) == []
    ref.step(15)
    b.step(15)
    assert b.get("damage").sum() > 0
    before = b.snapshot()
    pick = b.ctx.autotune(2)
    assert b.ctx.get_tuning() == pick
    for bond_pass, prefix in ((0, "defgrad_"), (1, "force_")):
        cands = b.ctx.autotune_candidates(bond_pass)
        assert 1 <= len(cands) <= 4
        assert [ms for _, ms in cands] == sorted(ms for _, ms in cands)
        keys = [k for k in pick if k.startswith(prefix)]
        assert any(all(c[k] == pick[k] for k in keys) for c, _ in cands)
        other = [k for k in pick if not k.startswith(prefix)]
        assert all(all(c[k] == pick[k] for k in other) for c, _ in cands)
    assert len(b.ctx.autotune_candidates(1, 1)) == 1
    with pytest.raises(vpd.VpdError, match="Bond pass"):
        b.ctx.autotune_candidates(2)
    kept = b.autotune_in_steps(reps=2, steps=5, warmup=2, rounds=1)
    assert b.ctx.get_tuning() == kept
    assert b.step_index == 15
    after = b.snapshot()
    for k in before:
        assert np.array_equal(before[k], after[k]), k
    ref.step(20)
    b.step(20)
    for name in ["u", "v", "stressN", "statev", "damage"]:
        assert np.array_equal(b.get(name), ref.get(name)), name


@pytest.mark.parametrize("take,plan,gwin", [(0, 0, 0), (8, 0, 0), (6, 1, 0), (7, 2, 0),
                                            (0, 3, (512, 480)), (0, 3, (480, 480)),
                                            (0, 3, (640, 640))])
def test_full_size_windowed_equals_direct(vpd, take, plan, gwin):
    banner("500k bar: windowed gathers (bonds per window %d, plan form %d, group windows %s) "
           "== direct, 12 steps with damage" % (take, plan, gwin))
    kw = dict(nx=200, ny=50, nz=50, spacing=0.2, horizon=0.603, props=JC_PROPS, dt=1.7075e-5,
              density=7.85e-3, velocity=60.0, damage=(0.0005, 0.0))
    a = vpd.Bar(**kw)
    b = vpd.Bar(**kw)
    window = gwin[0] if plan == 3 else 80 if take == 0 else 128
    b.ctx.set_tuning(defgrad_window=window, force_window=gwin[1] if plan == 3 else window,
                     defgrad_rounds=4, force_rounds=4, defgrad_window_bonds=take,
                     force_window_bonds=take, defgrad_window_plan=plan, force_window_plan=plan)
    a.step(12)
    b.step(12)
    assert a.get("damage").sum() > 0
    for name in ["u", "v", "stressN", "damage"]:
        assert np.array_equal(a.get(name), b.get(name)), name


def test_full_size_balance_and_determinism(vpd):
    # 500k-particle bar, damage on: Σ f_i V_i = 0 up to rounding (pairwise cancellation),
    # two runs bit-identical, state finite.
    banner("500k bar: balance + determinism")
    kw = dict(nx=200, ny=50, nz=50, spacing=0.2, horizon=0.603, props=JC_PROPS, dt=1.7075e-5,
              density=7.85e-3, velocity=1.0, damage=(0.02, 0.0))
    a = vpd.Bar(**kw)
    a.step(20, want_fields=True)
    f = a.get("force")
    vol = a.vol
    net = np.abs((f * vol[:, None]).sum(0)).max()
    assert net <= 1e-9 * np.abs(f * vol[:, None]).sum()
    u1 = a.get("u")
    assert np.isfinite(u1).all()
    del a
    b = vpd.Bar(**kw)
    b.step(20)
    assert np.array_equal(b.get("u"), u1)


@pytest.mark.parametrize("failure", [0.0, 0.01, 1.5])
def test_damage_near_threshold_bitwise(vpd, orc, failure):
    # Bonds stretched to within 1e-13 … 1e-6 of the critical ratio (and of the compression
    # threshold) on both sides: the GPU's filtered test must break exactly the oracle's set.
    banner(f"damage near threshold failure={failure}")
    rng = np.random.default_rng(21)
    x, vol = orc.make_patch(10, 6, 5, 1.0)
    rp, col = orc.build_neighbors(x, 1.5)
    n = len(x)
    sc = 0.02
    u = np.zeros((n, 3))
    # Give every particle a displacement that scales its bonds to (1+sc)(1+η) along x.
    eta = rng.choice([-1e-6, -1e-10, -1e-13, 0.0, 1e-13, 1e-10, 1e-6], n)
    u[:, 0] = x[:, 0] * sc * (1 + eta)
    u[:, 1] = -x[:, 1] * failure * (1 + eta) * 0.5
    c = ctx_for(vpd, x, vol, horizon=1.5, damage=(sc, failure), rowptr=rp, col=col)
    c.initialize_shape()
    c.initialize_dynamics()
    c.upload("u", u)
    c.compute_deformation_gradient()
    got = c.download("damage")
    ref = np.zeros(len(col))
    orc.damage_update(x, u, vol, rp, col, sc, failure, ref)
    assert np.array_equal(got, ref)
    assert 0 < got.sum() < len(col)
