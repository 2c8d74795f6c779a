"""GPU checks of the return-map solvers and of the learned warm start.

* Return-map solvers on identical inputs: the parity bisection equals the oracle; the tight
  bisection, Newton and learned-seed Newton all land on the FP64 oracle solved to
  |r| < 1e-14·q (stress within 1e-10 relative).
* The warm-start network: GPU gradients (FP32 kernels and bf16 MFMA) against the FP64 oracle
  network, bf16 MFMA inference against the FP64 forward pass, deterministic training that
  lowers the loss in both forms.
"""
import sys

import numpy as np
import pytest

from conftest import JC_PROPS

pytestmark = pytest.mark.gpu


def banner(msg):
    print("[gpu-case]", msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def vpd():
    import vpd as V
    if V.device_count() == 0:
        pytest.fail("no HIP device visible to the GPU tests")
    return V


def plastic_block(n, seed):
    """Points on the yield surface with a loading increment: state initialised, random
    hardening history, trial stress beyond the current yield."""
    rng = np.random.default_rng(seed)
    inc = rng.standard_normal((6, n)) * 1e-3
    inc[:3] += 1e-3
    stress = rng.standard_normal((6, n)) * 50.0
    state = np.zeros((6, n))
    state[0] = 1.0
    state[1] = rng.random(n) * 0.05
    state[2] = 293.0 + rng.random(n) * 100
    state[3] = 300.0 + rng.random(n) * 200.0
    state[4] = rng.random(n) * 2e-4
    state[5] = 1.0
    return inc, stress, state


def test_return_map_modes(vpd, orc):
    banner("return-map solvers vs oracle")
    n = 20000
    dt = 8.5377e-5
    inc, stress, state = plastic_block(n, 4)
    z = np.zeros(n)
    ref = orc.vumat_jc(JC_PROPS, dt, inc, stress, state, np.full(n, 293.0), z, z)
    tight = orc.vumat_jc(JC_PROPS, dt, inc, stress, state, np.full(n, 293.0), z, z,
                         rtol=1e-14, max_iter=400)
    plastic = tight["state"][4] > 0
    assert plastic.mean() > 0.5
    scale = np.abs(tight["stress"]).max()
    runs = {}
    for mode, seed in ((0, None), (1, None), (2, None), (3, tight["state"][4].copy())):
        r = vpd.return_map(mode, JC_PROPS, dt, inc, stress, state, z, z, seed=seed)
        runs[mode] = r
    # Parity mode is the reference algorithm.
    assert np.abs(runs[0]["stress"] - ref["stress"]).max() <= 1e-12 * scale
    assert (runs[0]["iterations"] == ref["state"][5] - state[5]).mean() >= 0.999
    # Fast modes converge to the tight solution.
    for mode in (1, 2, 3):
        err = np.abs(runs[mode]["stress"] - tight["stress"]).max() / scale
        assert err <= 1e-10, (mode, err)
        dep = np.abs(runs[mode]["state"][4] - tight["state"][4]).max()
        assert dep <= 1e-12, (mode, dep)
    it = {m: runs[m]["iterations"][plastic].mean() for m in runs}
    print("mean iterations on plastic points:", it)
    assert it[2] < it[1]  # Newton beats bisection at equal accuracy
    assert it[3] <= 3.0   # an exact seed needs only the acceptance test


def test_screen_hardware_error_bounds(vpd):
    """The FP32 yield screen's error budget (vpd_math.hpp) assumes the hardware log2 errs by
    at most kScreenLog2Hw (absolute, every float in [0.5, 1]) and the hardware exp2 by at most
    kScreenExp2Hw (relative, every float in [0, 1]); measured exhaustively on this device."""
    import os
    import re
    banner("screen: hardware log2/exp2 error bounds")
    src = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                            "vumat-pd_amd", "csrc", "vpd_math.hpp")).read()
    bound = {k: float(re.search(r"%s = ([0-9.e+-]+);" % k, src).group(1))
             for k in ("kScreenLog2Hw", "kScreenExp2Hw")}
    log2_err, exp2_err = vpd.screen_accuracy()
    print("log2 %.4e (bound %.1e), exp2 %.4e (bound %.1e)"
          % (log2_err, bound["kScreenLog2Hw"], exp2_err, bound["kScreenExp2Hw"]))
    # Not exact: the FP32 results are rounded, so each function errs by a fraction of an ulp.
    assert 0.0 < log2_err <= bound["kScreenLog2Hw"]
    assert 0.0 < exp2_err <= bound["kScreenExp2Hw"]


@pytest.mark.parametrize("case", ["plastic", "first_yield", "heated", "high_rate", "tiny",
                                  "reference_props", "wide"])
def test_screened_bisection_equals_unscreened(vpd, case):
    """The parity bisection decides most iterations from an FP32 estimate with an error bound
    and evaluates the FP64 yield only where that cannot decide: every output (stress, state
    including the iteration count in state 6, energies, iterations) must equal the bisection
    that evaluates every iteration, bit for bit."""
    from conftest import JC_PROPS_REFERENCE
    banner("screened bisection, %s" % case)
    n = 200000
    rng = np.random.default_rng(sum(map(ord, case)))
    inc, stress, state = plastic_block(n, 11)
    props, dt = JC_PROPS, 8.5377e-5
    if case == "first_yield":
        state[1] = 0.0
        state[3] = 0.0
        state[4] = 0.0
    elif case == "heated":
        state[2] = 293.0 + rng.random(n) * 1499.0
    elif case == "high_rate":
        inc *= 20.0
        dt = 1e-6
    elif case == "tiny":
        inc *= 1e-3
        state[3] = 0.0
    elif case == "reference_props":
        props = JC_PROPS_REFERENCE
        state[1] = rng.random(n) * 2.0
    elif case == "wide":  # increments over five decades, any history and temperature
        n = 1000000
        inc, stress, state = plastic_block(n, 12)
        inc *= 10.0 ** rng.uniform(-4.0, 1.0, n)
        state[1] = 10.0 ** rng.uniform(-6.0, 0.5, n)
        state[2] = 293.0 + rng.random(n) * 1400.0
        state[4] = np.where(rng.random(n) < 0.3, 0.0, state[4])
        dt = 1e-5
    eint = rng.random(n)
    einel = rng.random(n)
    a = vpd.return_map(0, props, dt, inc, stress, state, eint, einel)
    b = vpd.return_map(4, props, dt, inc, stress, state, eint, einel)
    plastic = (b["state"][4] > 0).mean()
    print("%s: plastic %.3f, mean iterations %.2f, screened %.3f ms, every-iteration %.3f ms"
          % (case, plastic, b["iterations"].mean(), a["ms"], b["ms"]))
    # The same bisection on lane groups (several iterations per round).
    c = vpd.return_map(vpd.RETURN_MAP_BISECTION_GROUPED, props, dt, inc, stress, state, eint,
                       einel)
    print("%s: lane groups %.3f ms, max iterations %d" % (case, c["ms"], b["iterations"].max()))
    for k in ("stress", "state", "eint", "einel", "iterations"):
        assert np.array_equal(a[k], b[k]), k
        assert np.array_equal(c[k], b[k]), ("lane groups", k)
    assert case == "tiny" or plastic > 0.1


def test_return_map_elastic_points_unchanged(vpd, orc):
    banner("return-map elastic points")
    n = 5000
    rng = np.random.default_rng(8)
    inc = rng.standard_normal((6, n)) * 1e-6
    stress = rng.standard_normal((6, n)) * 10.0
    state = np.zeros((6, n))
    state[0] = 1.0
    state[2] = 293.0
    state[3] = 350.0
    z = np.zeros(n)
    ref = orc.vumat_jc(JC_PROPS, 1e-4, inc, stress, state, np.full(n, 293.0), z, z)
    for mode in (2, 3):
        r = vpd.return_map(mode, JC_PROPS, 1e-4, inc, stress, state, z, z, seed=np.zeros(n))
        assert np.array_equal(r["state"][4], ref["state"][4])  # all zero
        assert np.abs(r["stress"] - ref["stress"]).max() <= 1e-12 * np.abs(ref["stress"]).max()


# The parameter blocks of the flat vector (W1, b1, W2, b2, W3, b3, W4, b4).
MLP_BLOCKS = (("W1", 0, 512), ("b1", 512, 576), ("W2", 576, 4672), ("b2", 4672, 4736),
              ("W3", 4736, 8832), ("b3", 8832, 8896), ("W4", 8896, 8960), ("b4", 8960, 8961))


@pytest.mark.parametrize("form", ["FP32", "MFMA"])
def test_mlp_gradient_vs_fp64_oracle(vpd, form):
    banner("warm-start network gradient, %s" % form)
    import mlp
    rng = np.random.default_rng(0)
    theta = (mlp.init(5) + 0.02 * rng.standard_normal(mlp.n_params())).astype(np.float32)
    x = rng.standard_normal((1000, 8)).astype(np.float32)  # ragged: not a multiple of 64
    t = rng.standard_normal(1000).astype(np.float32)
    net = vpd.WarmStart(1024)
    net.set_params(theta)
    net.set_training_form(getattr(net, "TRAIN_" + form))
    loss, g = net.loss_and_gradient(x, t)
    lref, gref = mlp.loss_and_grad(theta.astype(np.float64), x.astype(np.float64),
                                   t.astype(np.float64))
    cos = float(g @ gref / (np.linalg.norm(g) * np.linalg.norm(gref)))
    blocks = {name: float(np.linalg.norm(g[a:b] - gref[a:b]) / np.linalg.norm(gref[a:b]))
              for name, a, b in MLP_BLOCKS}
    print("%s: loss rel err %.2e, cosine %.8f, block rel errs %s"
          % (form, abs(loss - lref) / lref, cos, {k: "%.1e" % v for k, v in blocks.items()}))
    if form == "FP32":
        assert abs(loss - lref) <= 1e-5 * lref
        assert np.abs(g - gref).max() <= 1e-4 * np.abs(gref).max()
        # FP32 against FP64: a pre-activation within rounding of a ReLU kink may fall on the
        # other side, which moves a few components by more than rounding.
        assert cos > 1 - 1e-6
        return
    # bf16 operands (8-bit significands) in every product: against the FP64 network the
    # gradient moves by up to ~10% in the small bias blocks (the forward pass's bf16 error
    # shifts every dy alike), so the FP64 comparison is a loose sanity bound. The tight check
    # is the oracle's statement of the same bf16 pipeline (the same operands rounded at the
    # same places, FP64 sums): only the FP32 accumulation order and the rare activation that
    # rounds to the other bf16 neighbour differ. A transposed or misplaced operand tile, a
    # wrong mask or a dropped bias term gives errors of order 1 there.
    assert abs(loss - lref) <= 1e-2 * lref
    assert cos > 0.995
    lb, gb = mlp.loss_and_grad_bf16(theta.astype(np.float64), x, t.astype(np.float64))
    tight = {name: float(np.linalg.norm(g[a:b] - gb[a:b]) / np.linalg.norm(gb[a:b]))
             for name, a, b in MLP_BLOCKS}
    print("MFMA vs bf16-pipeline oracle: loss rel err %.2e, block rel errs %s"
          % (abs(loss - lb) / lb, {k: "%.1e" % v for k, v in tight.items()}))
    assert abs(loss - lb) <= 1e-5 * lb
    for name, err in tight.items():
        assert err <= 1e-3, (name, err)


def test_mlp_mfma_inference_vs_fp64(vpd):
    banner("warm-start bf16 MFMA inference")
    import mlp
    rng = np.random.default_rng(1)
    theta = (mlp.init(7) + 0.05 * rng.standard_normal(mlp.n_params())).astype(np.float32)
    net = vpd.WarmStart(256)
    net.set_params(theta)
    x = rng.standard_normal((1003, 8)).astype(np.float32)  # ragged: not a multiple of 16
    y = net.predict(x)
    yref, _ = mlp.forward(theta.astype(np.float64), x.astype(np.float64))
    # bf16 operands (8-bit mantissa) through three layers: ~1e-2 of the output scale.
    assert np.abs(y - yref).max() <= 3e-2 * max(1.0, np.abs(yref).max())
    assert np.corrcoef(y, yref)[0, 1] > 0.999


@pytest.mark.parametrize("form", ["FP32", "MFMA"])
def test_mlp_training_deterministic_and_descends(vpd, form):
    banner("warm-start training, %s" % form)
    import mlp
    rng = np.random.default_rng(2)
    x = rng.standard_normal((20000, 8)).astype(np.float32)
    t = (np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2]).astype(np.float32)
    losses = []
    params = []
    for _ in range(2):
        net = vpd.WarmStart(1024)
        net.set_training_form(getattr(net, "TRAIN_" + form))
        net.set_params(mlp.init(3).astype(np.float32))
        net.set_dataset(x, t, seed=11)
        l0, _ = net.loss_and_gradient(x[:1024], t[:1024])
        losses.append((l0, net.train(600, 1024, 3e-3)))
        params.append(net.get_params()[0])
    assert losses[0][1] < 0.25 * losses[0][0]
    assert np.array_equal(params[0], params[1])  # a training run is deterministic


def plastic_bar(vpd, mode, net=None, diag=False):
    kw = dict(nx=30, ny=8, nz=8, spacing=1.0, horizon=3.015, props=JC_PROPS, dt=8.5377e-5,
              density=7.85e-3, velocity=40.0, return_map=mode, warm_start=net)
    bar = vpd.Bar(**kw)
    if diag:
        bar.ctx.enable_diagnostics(0, 1, capacity=200)
    bar.step(150, want_fields=True)
    return bar


def any_network(vpd):
    import mlp
    net = vpd.WarmStart(4096)
    norm = np.zeros(18, np.float32)
    norm[8:16] = 1e-3
    norm[16], norm[17] = 1e-4, 1e-4
    net.set_params(mlp.init(9).astype(np.float32), norm)
    return net


@pytest.mark.parametrize("mode", ["BISECTION", "BISECTION_TIGHT", "NEWTON"])
def test_material_compacted_equals_fused(vpd, mode):
    # Without run diagnostics the material stage is compacted (k_material_screen finishes
    # the points whose return map ends at its first iteration, reusing that iteration's
    # yield, and lists the rest for k_material_return_list); with them every point goes
    # through the one fused kernel. Both give the same bits.
    banner("%s return map: compacted == fused" % mode.lower())
    m = getattr(vpd, "RETURN_MAP_" + mode)
    a = plastic_bar(vpd, m, diag=True)
    b = plastic_bar(vpd, m)
    assert (b.get("statev")[:, 4] > 0).sum() > 100  # plastic flow on this step
    for name in ("u", "v", "a", "stressN", "statev", "temp", "eint", "einel", "damage"):
        assert np.array_equal(a.get(name), b.get(name)), name
    assert np.array_equal(a.get("iterations"), b.get("iterations"))


def test_learned_step_compacted_equals_split(vpd):
    # The learned update runs compacted (screen → network and Newton on the plastic list)
    # unless run diagnostics are on, which take the uncompacted split path. Both give the
    # same bits: every point's arithmetic is the same, whatever its slot in the list.
    banner("learned return map: compacted == split")
    net = any_network(vpd)
    a = plastic_bar(vpd, vpd.RETURN_MAP_LEARNED, net, diag=True)
    b = plastic_bar(vpd, vpd.RETURN_MAP_LEARNED, net)
    assert (b.get("statev")[:, 4] > 0).sum() > 100  # plastic flow on this step
    for name in ("u", "v", "a", "stressN", "statev", "temp", "eint", "einel"):
        assert np.array_equal(a.get(name), b.get(name)), name
    assert np.array_equal(a.get("iterations"), b.get("iterations"))
    # Against the state-5-seeded Newton: the same converged states within the tolerance.
    c = plastic_bar(vpd, vpd.RETURN_MAP_NEWTON)
    for name in ("stressN", "u"):
        x, y = b.get(name), c.get(name)
        assert np.abs(x - y).max() <= 1e-9 * np.abs(y).max(), name
    # State 1-5 (state 6 counts iterations, which differ with the seed).
    x, y = b.get("statev")[:, :5], c.get("statev")[:, :5]
    assert (np.abs(x - y) <= 1e-9 * np.abs(y).max(0)).all()


def test_trained_warm_start_gain(vpd):
    # The shipped network (warm_start_jc.npz, trained on the GPU from plain runs of the
    # 30×8×8 JC bar) on plastic points of held-out runs (other velocities): every solver lands
    # on the tight FP64 solution. The learned mode seeds from state 5 where it is positive and
    # from the network where it is zero (first yield); there the network's seed needs fewer
    # iterations than Newton started from the bracket midpoint, so the learned mode needs no
    # more iterations than the state-5 Newton overall, and far fewer than bisection.
    banner("trained warm start: gain on held-out plastic points")
    import warm_start as ws
    net = vpd.load_warm_start()
    rec = ws.log_pairs(40000, seed=777, runs=2, steps=1500, verbose=False)
    m = len(rec["y"])
    assert m >= 20000
    inc = np.ascontiguousarray(rec["inc"].T)
    stress = np.ascontiguousarray(rec["stress"].T)
    state = np.ascontiguousarray(rec["state"].T)
    first = state[4] == 0.0
    no5 = state.copy()
    no5[4] = 0.0  # every point takes the network's seed (state 5 is not read otherwise)
    z = np.zeros(m)
    dt = ws.cfl_dt(1.0)
    exact = vpd.return_map(vpd.RETURN_MAP_BISECTION_TIGHT, ws.JC_PROPS, dt, inc, stress, state,
                           z, z, rtol=1e-14)
    scale = np.abs(exact["stress"]).max()
    runs = {
        "learned": vpd.return_map_learned(net, ws.JC_PROPS, dt, inc, stress, state, z, z),
        "network_everywhere": vpd.return_map_learned(net, ws.JC_PROPS, dt, inc, stress, no5, z,
                                                     z),
        "no_seed": vpd.return_map(vpd.RETURN_MAP_LEARNED, ws.JC_PROPS, dt, inc, stress, state,
                                  z, z, seed=np.zeros(m)),
        "state5_seed": vpd.return_map(vpd.RETURN_MAP_NEWTON, ws.JC_PROPS, dt, inc, stress, state,
                                      z, z),
        "bisection_tight": vpd.return_map(vpd.RETURN_MAP_BISECTION_TIGHT, ws.JC_PROPS, dt, inc,
                                          stress, state, z, z),
    }
    it = {k: float(r["iterations"].mean()) for k, r in runs.items()}
    it_first = {k: float(r["iterations"][first].mean()) for k, r in runs.items()}
    net_err = np.median(np.abs(runs["network_everywhere"]["seed"] - exact["state"][4]) /
                        exact["state"][4])
    print("first-yield points %d of %d; mean iterations:" % (first.sum(), m), it,
          "on first-yield points:", it_first, "network median seed error %.3g" % net_err)
    for k, r in runs.items():
        assert np.abs(r["stress"] - exact["stress"]).max() <= 1e-10 * scale, k
    assert first.sum() > 0
    assert net_err < 0.05
    # The learned seeds are state 5 where it is positive, the network's elsewhere.
    assert np.array_equal(runs["learned"]["seed"][~first], state[4][~first])
    assert np.array_equal(runs["learned"]["seed"][first], runs["network_everywhere"]["seed"][first])
    assert it_first["learned"] < 0.8 * it_first["state5_seed"]
    assert it["learned"] <= it["state5_seed"]
    assert it["network_everywhere"] < 0.8 * it["no_seed"]
    assert it["learned"] < 0.25 * it["bisection_tight"]


def test_fused_learned_return_map_equals_two_pass(vpd):
    # vpd_return_map_learned_device seeds the block in one launch (each workgroup runs the
    # network on its own points without a positive state 5), then solves. It must give the
    # seeds and every output of the two-call form (vpd_mlp_seed_device, then mode 3 with those
    # seeds), bit for bit, on a block mixing continuing and first-yield points, ragged at the
    # end; and the same outputs without seed_out, where only the network's seeds are written
    # and the solve takes state 5 itself.
    banner("learned return map: fused seeding == two-pass")
    net = vpd.load_warm_start()
    n = 100003
    inc, stress, state = plastic_block(n, 21)
    rng = np.random.default_rng(5)
    state[4] = np.where(rng.random(n) < 0.4, 0.0, state[4])  # 40% without a state-5 seed
    z = np.zeros(n)
    dt = 8.5377e-5
    a = vpd.return_map_learned(net, JC_PROPS, dt, inc, stress, state, z, z)
    b = vpd.return_map_learned(net, JC_PROPS, dt, inc, stress, state, z, z, fused=False)
    c = vpd.return_map_learned(net, JC_PROPS, dt, inc, stress, state, z, z, want_seeds=False)
    for k in ("seed", "stress", "state", "eint", "einel", "iterations", "fallbacks"):
        assert np.array_equal(a[k], b[k]), k
        assert k == "seed" or np.array_equal(a[k], c[k]), ("without seed_out", k)
    assert "seed" not in c
    first = ~(state[4] > 0.0)
    assert np.array_equal(a["seed"][~first], state[4][~first])
    assert (a["seed"][first] > 0).mean() > 0.5
