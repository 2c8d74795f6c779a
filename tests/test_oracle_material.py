"""Pins the restated oracle's material side: Flanagan-Taylor kinematics
(tests/corotational_kinematics_test.cpp closed forms), the Johnson-Cook VUMAT
(elastic trial, init branch, residual bound, ε=0 yield, heat-unit ratio of
tests/jc_thermal_units_test.cpp) and the VUMAT-PD block coupling, each also bitwise
against the reference build in oracle/_ref where it exists."""
import numpy as np
import pytest

from conftest import JC_PROPS, JC_PROPS_REFERENCE

I9 = np.eye(3).ravel()


def one(t):
    return np.asarray(t, float).reshape(1, 9)


def test_rigid_rotation(orc):
    a = 1e-7
    f1 = one([np.cos(a), -np.sin(a), 0, np.sin(a), np.cos(a), 0, 0, 0, 1])
    k = orc.corotational(one(I9), f1, one(I9), one(I9), 1.0)
    assert np.abs(k["Dbar"]).max() <= 1e-13
    assert np.abs(k["V"] - I9).max() <= 1e-13


def test_extension_and_green_lagrange(orc):
    f1 = one([1.01, 0, 0, 0, 1, 0, 0, 0, 1])
    k = orc.corotational(one(I9), f1, one(I9), one(I9), 0.01)
    assert abs(k["Dbar"][0, 0] - 1.0) <= 1e-12 and abs(k["V"][0, 0] - 1.01) <= 1e-12
    assert abs(orc.green_lagrange(f1)[0, 0] - 0.01005) <= 1e-12


def test_stress_rotation(orc):
    r = one([0, -1, 0, 1, 0, 0, 0, 0, 1])
    s = orc.rotate_to_spatial(one([10, 0, 0, 0, 2, 0, 0, 0, 1]), r)
    assert abs(s[0, 0] - 2) <= 1e-12 and abs(s[0, 4] - 10) <= 1e-12


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("angle", [1e-7, 0.01, 0.1, 0.7, -0.4])
def test_finite_spin_exponential(orc, axis, angle):
    # corotational_kinematics_test.cpp:79-100: F = I + W isolates exp(Ω).
    a, b = (axis + 1) % 3, (axis + 2) % 3
    f = np.eye(3)
    f[a, b], f[b, a] = -angle, angle
    exp = np.eye(3)
    exp[a, a] = exp[b, b] = np.cos(angle)
    exp[a, b], exp[b, a] = -np.sin(angle), np.sin(angle)
    k = orc.corotational(one(I9), one(f), one(I9), one(I9), 1.0)
    r = k["R"].reshape(3, 3)
    assert np.abs(r - exp).max() <= 2e-14
    assert np.abs(r.T @ r - np.eye(3)).max() <= 3e-14


def test_accumulated_rotation(orc):
    # corotational_kinematics_test.cpp:102-137
    rot, prev = one(I9), one(I9)
    for step in range(1, 1001):
        ang = 0.01 * step
        f = one([np.cos(ang), -np.sin(ang), 0, np.sin(ang), np.cos(ang), 0, 0, 0, 1])
        rot = orc.corotational(prev, f, one(I9), rot, 1.0)["R"]
        prev = f
    ang = 1000 * np.sin(0.01)
    exp = [np.cos(ang), -np.sin(ang), 0, np.sin(ang), np.cos(ang), 0, 0, 0, 1]
    assert np.abs(rot[0] - exp).max() <= 2e-12


def random_kinematics(n, seed):
    rng = np.random.default_rng(seed)
    fN = I9 + 0.05 * rng.standard_normal((n, 9))
    f1 = fN + 0.01 * rng.standard_normal((n, 9))
    vN = I9 + 0.02 * rng.standard_normal((n, 9))
    vN = 0.5 * (vN + vN[:, [0, 3, 6, 1, 4, 7, 2, 5, 8]])
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    rN = np.tile((q * np.sign(np.linalg.det(q))).ravel(), (n, 1))
    return fN, f1, vN, rN


def test_corotational_bitwise_vs_reference(ref):
    args = random_kinematics(200, 3)
    a = ref.corotational(*args, 1e-4)
    b = ref.ref_corotational(*args, 1e-4)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def jc_block(orc, props, dt, dstrain, stress, state, temp=None):
    nb = dstrain.shape[1]
    temp = np.full(nb, 293.0) if temp is None else temp
    return orc.vumat_jc(props, dt, dstrain, stress, state, temp, np.zeros(nb), np.zeros(nb))


def test_jc_init_branch_and_elastic_trial(orc):
    p = JC_PROPS
    E, nu = p[0], p[1]
    mu, lam = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    de = np.array([[1e-4], [-2e-5], [3e-5], [1e-5], [-4e-5], [2e-5]])
    r = jc_block(orc, p, 1e-4, de, np.zeros((6, 1)), np.zeros((6, 1)))
    tr = de[:3, 0].sum()
    exp = np.concatenate([lam * tr + 2 * mu * de[:3, 0], 2 * mu * de[3:, 0]])
    assert np.abs(r["stress"][:, 0] - exp).max() <= 1e-9
    assert list(r["state"][:, 0]) == [1.0, 0.0, 293.0, 0.0, 0.0, 1.0]
    assert r["temp"][0] == 293.0 and r["eint"][0] == 0.0


def test_jc_plastic_residual_and_yield(orc):
    p = JC_PROPS
    E, nu, A, B, Cc, n = p[0], p[1], p[2], p[3], p[4], p[5]
    mu = E / (2 * (1 + nu))
    state = np.array([[1.0], [0.0], [293.0], [0.0], [0.0], [1.0]])
    de = np.array([[2e-3], [-1e-3], [-1e-3], [0.0], [0.0], [0.0]])
    dt = 1e-4
    r = jc_block(orc, p, dt, de, np.zeros((6, 1)), state)
    s = r["stress"][:, 0]
    pr = s[:3].mean()
    dev = np.concatenate([s[:3] - pr, s[3:]])
    q = np.sqrt(1.5 * ((dev[:3] ** 2).sum() + 2 * (dev[3:] ** 2).sum()))
    dep = r["state"][4, 0]
    assert dep > 0 and r["state"][1, 0] == dep
    y = (A + B * dep ** n) * (1 + Cc * np.log(dep / dt / p[7]))
    assert abs(q - y) < E * 1e-9 * 1.0001 + 1e-9  # converged: |q_new − Y| < tol
    assert r["state"][5, 0] > 2  # iterations accumulated on state 6
    assert r["einel"][0] > 0 and r["state"][2, 0] > 293.0
    # Yield at ε̄ = 0 and zero rate: Y0 = A(1 − θ^m) (effectiveRate = rate0). Uniaxial
    # Δε = e e₁ from zero stress has q_trial = 2μe, so e just under / over Y0/(2μ) must
    # exit elastic at iteration 1 / go plastic.
    T = 543.0
    y0 = A * (1 - ((T - 293.0) / 1500.0) ** p[6])
    st = np.array([[1.0], [0.0], [T], [0.0], [0.0], [1.0]])
    for factor, plastic in ((1 - 1e-6, False), (1 + 1e-3, True)):
        de = np.zeros((6, 1))
        de[0, 0] = factor * y0 / (2 * mu)
        r0 = jc_block(orc, p, dt, de, np.zeros((6, 1)), st)
        assert (r0["state"][4, 0] > 0) == plastic
        if not plastic:
            assert r0["state"][5, 0] == 2.0 and r0["state"][3, 0] == y0


def test_jc_heat_units(orc):
    # tests/jc_thermal_units_test.cpp:7-42 on the reference props: workToHeat 1 vs 1e6.
    outs = []
    for w2h in (1.0, 1e6):
        p = list(JC_PROPS_REFERENCE)
        p[13] = w2h
        init = jc_block(orc, p, 2.5e-5, np.zeros((6, 1)), np.zeros((6, 1)), np.zeros((6, 1)))
        de = np.zeros((6, 1))
        de[0, 0] = 400 * 2.5e-5
        outs.append(jc_block(orc, p, 2.5e-5, de, init["stress"], init["state"], init["temp"]))
    a, b = outs
    assert np.abs(a["stress"] - b["stress"]).max() <= 1e-10
    da, db = a["temp"][0] - 293, b["temp"][0] - 293
    assert da > 0 and abs(db / (1e6 * da) - 1) <= 1e-6


def test_jc_rejects_dimensions(orc):
    with pytest.raises(RuntimeError):
        orc.vumat_jc(JC_PROPS[:13], 1e-4, np.zeros((6, 1)), np.zeros((6, 1)), np.zeros((6, 1)),
                     np.zeros(1), np.zeros(1), np.zeros(1))


def random_jc_inputs(n, seed, props):
    """A mix of first calls, elastic and plastic points with warm-start history."""
    rng = np.random.default_rng(seed)
    E = props[0]
    strain_rate = rng.standard_normal((n, 9)) * 20.0
    stress = rng.standard_normal((n, 9)) * 150.0
    stress = 0.5 * (stress + stress[:, [0, 3, 6, 1, 4, 7, 2, 5, 8]])
    state = np.zeros((n, 6))
    state[:, 0] = (rng.random(n) > 0.1).astype(float)
    state[:, 1] = rng.random(n) * 0.05
    state[:, 2] = 293.0 + rng.random(n) * 50
    state[:, 3] = rng.random(n) * 400.0
    state[:, 4] = rng.random(n) * 1e-3
    state[:, 5] = rng.integers(1, 50, n)
    del E
    return strain_rate, stress, state


@pytest.mark.parametrize("props", [JC_PROPS, JC_PROPS_REFERENCE])
def test_jc_bitwise_vs_reference_fortran(ref, props):
    # The restated C++ JC against the reference's Fortran example compiled by amdflang,
    # both behind VumatMaterialAdapter::updateBlock (padding, chunking, packing).
    sr, s, st = random_jc_inputs(600, 5, props)
    n = sr.shape[0]
    t, e0, e1 = np.full(n, 293.0), np.zeros(n), np.zeros(n)
    a = ref.ref_vumat_block(props, 1e-4, sr, s, st, t, e0, e1, plugin=ref.REF_PLUGIN)
    b = ref.ref_vumat_block(props, 1e-4, sr, s, st, t, e0, e1, plugin=ref.ORACLE_PLUGIN)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    assert (a["state"][:, 4] > 0).sum() > 100  # many plastic points


def test_material_update_bitwise_vs_reference_block_path(ref):
    # orc_material_update (FT + packing + JC + rotation) against the reference bar after a
    # few steps is covered in test_oracle_bar; here the first material call on a random
    # F field exercises the init branch through the 256-wide padding of 300 points.
    fN, f1, vN, rN = random_kinematics(300, 11)
    st = ref.material_init(fN, 6, 293.0)
    sig = ref.material_update(f1, st, 1e-4, np.asarray(JC_PROPS))
    assert np.isfinite(sig).all() and np.all(st["statev"][:, 0] == 1.0)
