"""Pins the restated oracle's PD pieces: neighbour search, shape tensor, deformation
gradient, force gather, damage and integrator. Closed forms follow the reference's own
tests (tests/neighbor_lattice_regression.cpp, tests/neighbor_search_test.cpp,
tests/correspondence_sampling_test.cpp, tests/outer_damage_test.cpp); bitwise checks run
against the reference build in oracle/_ref where it exists."""
import numpy as np
import pytest


def lattice(n, h, shift):
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    return np.stack([shift + i.ravel() * h, shift + j.ravel() * h, shift + k.ravel() * h], 1)


@pytest.mark.parametrize("h", [2.0 / 3.0, 1.0, 2.0])
@pytest.mark.parametrize("shift", [0.0, -123.456, 10000.125])
def test_neighbors_integer_lattice(orc, h, shift):
    # tests/neighbor_lattice_regression.cpp:5-31: both searches equal 0 < d^2 <= 9.
    n = 9
    x = lattice(n, h, shift)
    idx = np.arange(n ** 3)
    g = np.stack(np.unravel_index(idx, (n, n, n)), 1)
    d2 = ((g[:, None, :] - g[None, :, :]) ** 2).sum(-1)
    for brute in (False, True):
        rowptr, col = orc.build_neighbors(x, 3 * h, brute=brute)
        for p in range(n ** 3):
            expected = np.nonzero((d2[p] > 0) & (d2[p] <= 9))[0]
            assert np.array_equal(col[rowptr[p]:rowptr[p + 1]], expected)


def test_neighbors_cell_equals_brute(orc):
    # tests/neighbor_search_test.cpp:7-25
    x, _ = orc.make_patch(17, 9, 3, 1.25)
    a = orc.build_neighbors(x, 3 * 1.25)
    b = orc.build_neighbors(x, 3 * 1.25, brute=True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_neighbors_rejects(orc):
    x, _ = orc.make_patch(3, 3, 3, 1.0)
    with pytest.raises(RuntimeError, match="horizon must be positive"):
        orc.build_neighbors(x, 0.0)
    x[4, 1] = np.nan
    with pytest.raises(RuntimeError, match="Nonfinite"):
        orc.build_neighbors(x, 2.0)


def test_patch_volume_and_affine_completeness(orc):
    # tests/correspondence_sampling_test.cpp:7-46
    x, vol = orc.make_patch(9, 9, 9, 1.0)
    assert abs(vol.sum() - 512.0) <= 1e-12
    rowptr, col = orc.build_neighbors(x, 2.0)
    kinv = orc.shape_inverse(x, vol, rowptr, col)
    u = np.stack([0.02 * x[:, 0] + 0.03 * x[:, 1], -0.01 * x[:, 1], 0.04 * x[:, 2]], 1)
    F = orc.deformation_gradient(x, u, vol, rowptr, col, kinv)
    expected = np.array([1.02, 0.03, 0, 0, 0.99, 0, 0, 0, 1.04])
    assert np.abs(F - expected).max() <= 1e-12
    g = np.stack(np.unravel_index(np.arange(729), (9, 9, 9)), 1)
    sign = np.where(g.sum(1) % 2 == 1, 1.0, -1.0)
    u = np.zeros_like(x)
    u[:, 0] = 0.01 * sign
    F = orc.deformation_gradient(x, u, vol, rowptr, col, kinv)
    inner = np.all((g >= 3) & (g <= 5), 1)
    assert np.abs(F[inner] - np.eye(3).ravel()).max() <= 1e-12


def test_shape_inverse_closed_form(orc):
    # δ = 1.5h on a cubic lattice: 6 face + 12 edge neighbours, so an interior point has
    # K = Σ V ξ⊗ξ = 10 V h² I and K⁻¹ = I / (10 V h²).
    x, vol = orc.make_patch(5, 5, 5, 2.0)
    rowptr, col = orc.build_neighbors(x, 3.0)
    kinv = orc.shape_inverse(x, vol, rowptr, col)
    centre = (2 * 5 + 2) * 5 + 2
    assert rowptr[centre + 1] - rowptr[centre] == 18
    assert np.abs(kinv[centre] - np.eye(3).ravel() / (10 * 8.0 * 4.0)).max() <= 1e-18
    # A neighbourhood on one line is singular; the reference throws.
    xs = np.array([[0.0, 0, 0], [1.0, 0, 0], [2.0, 0, 0]])
    rp, cl = orc.build_neighbors(xs, 2.5)
    with pytest.raises(RuntimeError, match="Singular"):
        orc.shape_inverse(xs, np.ones(3), rp, cl)


def test_force_balance_and_uniform_stress(orc):
    x, vol = orc.make_patch(8, 6, 5, 1.0)
    rowptr, col = orc.build_neighbors(x, 3.015)
    kinv = orc.shape_inverse(x, vol, rowptr, col)
    rng = np.random.default_rng(1)
    n = x.shape[0]
    F = np.tile(np.eye(3).ravel(), (n, 1)) + 1e-3 * rng.standard_normal((n, 9))
    s = rng.standard_normal((n, 9))
    s = 0.5 * (s + s[:, [0, 3, 6, 1, 4, 7, 2, 5, 8]])
    f = orc.force_density(x, vol, rowptr, col, kinv, F, s)
    # Directed pairs cancel: Σ f_i V_i = 0 (BoundaryValueCaseRunner.cpp:251-252).
    assert np.abs((f * vol[:, None]).sum(0)).max() <= 1e-9 * np.abs(f * vol[:, None]).sum()
    # Uniform stress and F = I: a point whose neighbours all have full neighbourhoods
    # (two horizons clear of the half-volume boundary layer) feels no force.
    x, vol = orc.make_patch(9, 9, 9, 1.0)
    rowptr, col = orc.build_neighbors(x, 1.5)
    kinv = orc.shape_inverse(x, vol, rowptr, col)
    n = x.shape[0]
    F = np.tile(np.eye(3).ravel(), (n, 1))
    s = np.tile(np.array([100.0, 5, 3, 5, -20, 7, 3, 7, 40]), (n, 1))
    f = orc.force_density(x, vol, rowptr, col, kinv, F, s)
    g = np.stack(np.unravel_index(np.arange(n), (9, 9, 9)), 1)
    deep = np.all((g >= 3) & (g <= 5), 1)
    assert np.abs(f[deep]).max() <= 1e-12 * 100
    assert np.abs(f[~deep]).max() > 1.0


def test_damage_two_particles(orc):
    # tests/outer_damage_test.cpp:61-91
    x = np.array([[0.0, 0, 0], [1.0, 0, 0]])
    vol = np.array([1.0, 2.0])
    rowptr = np.array([0, 1, 2], np.int32)
    col = np.array([1, 0], np.int32)
    dmg = np.zeros(2)
    u = np.array([[0.0, 0, 0], [0.10, 0, 0]])
    idx = orc.damage_update(x, u, vol, rowptr, col, 0.05, 0.0, dmg, want_index=True)
    assert list(dmg) == [1.0, 1.0] and list(idx) == [1.0, 1.0]
    u[1] = 0.0
    orc.damage_update(x, u, vol, rowptr, col, 0.05, 0.0, dmg)
    assert list(dmg) == [1.0, 1.0]  # no healing
    dmg = np.zeros(2)
    u[1] = [-0.03, 0, 0]
    orc.damage_update(x, u, vol, rowptr, col, 0.05, 0.02, dmg)
    assert dmg[0] == 1.0
    dmg = np.zeros(2)
    orc.damage_update(x, u, vol, rowptr, col, 0.05, 0.0, dmg)
    assert dmg[0] == 0.0  # failure = 0 disarms the compression threshold
    with pytest.raises(RuntimeError, match="initiation"):
        orc.damage_update(x, u, vol, rowptr, col, 0.0, 0.0, dmg)


def test_integrator_constant_acceleration(orc):
    n = 4
    u = np.zeros((n, 3))
    v = np.full((n, 3), 2.0)
    a = np.full((n, 3), 3.0)
    orc.lib().orc_predict(u, v, a, n, 0.5)
    assert np.all(u == 0.5 * 2.0 + 0.5 * 0.25 * 3.0)
    f = np.full((n, 3), 6.0)
    orc.lib().orc_correct(v, a, f, n, 0.5, 2.0)
    assert np.all(a == 3.0) and np.all(v == 2.0 + 0.25 * 6.0)


def test_constraint_codes_and_ramp(orc):
    x, vol = orc.make_patch(10, 5, 3, 1.0)
    codes, anchor, pin = orc.constraint_codes(x, 0, 2, 1.0)
    nxband = 2 * 5 * 3
    assert (codes[:, 0] == 1).sum() == nxband and (codes[:, 0] == 2).sum() == nxband
    # Closest support particle to the centroid: the inner support layer (x = 1).
    assert x[anchor, 0] == 1.0 and abs(x[anchor, 1]) < 1e-12 and abs(x[anchor, 2]) < 1e-12
    assert codes[anchor, 1] == 1 and codes[anchor, 2] == 1 and codes[pin, 2] == 1
    assert x[pin, 1] == x[:, 1].max() and x[pin, 0] == x[anchor, 0]
    r = orc.ramp(2.0, 0.12, 0.06)
    assert abs(r[0] - 1.0) < 1e-15 and abs(r[1] - np.pi / 0.12) < 1e-12 and abs(r[2]) < 1e-9
    assert list(orc.ramp(2.0, 0.12, 0.2)) == [2.0, 0.0, 0.0]


def test_pd_pieces_bitwise_vs_reference(ref):
    o = ref
    x, vol = o.make_patch(12, 7, 5, 0.5)
    h = 3.015 * 0.5
    for brute in (False, True):
        a = o.build_neighbors(x, h, brute=brute)
        b = o.ref_build_neighbors(x, h, brute=brute)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    rowptr, col = a
    kinv = o.shape_inverse(x, vol, rowptr, col)
    assert np.array_equal(kinv, o.ref_shape_inverse(x, vol, rowptr, col, h))
    rng = np.random.default_rng(7)
    u = 0.01 * rng.standard_normal(x.shape)
    F = o.deformation_gradient(x, u, vol, rowptr, col, kinv)
    assert np.array_equal(F, o.ref_deformation_gradient(x, u, vol, rowptr, col, kinv, h))
    s = rng.standard_normal((x.shape[0], 9))
    dmg = (rng.random(len(col)) < 0.1).astype(float)
    f = o.force_density(x, vol, rowptr, col, kinv, F, s, dmg)
    assert np.array_equal(f, o.ref_force_density(x, vol, rowptr, col, kinv, F, s, h, dmg))
