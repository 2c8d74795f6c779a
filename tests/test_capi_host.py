"""CPU-side checks of the product library: libvpd.so loads, exports every entry point that
include/vpd.h declares, and its host helpers (lattice, boundary bands, constraint codes,
ramp) agree with the oracle. No GPU work is issued here."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "vpd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vpd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header(orc):
    import vpd
    lib = vpd.lib()
    names = declared_symbols()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
    assert lib.vpd_abi_version() == 3


@pytest.mark.parametrize("dims", [(30, 8, 8, 1.0), (11, 5, 3, 0.2), (2, 2, 2, 1.0)])
def test_patch_matches_oracle(orc, dims):
    import vpd
    x, v = vpd.make_patch(*dims)
    xo, vo = orc.make_patch(*dims)
    assert np.array_equal(x, xo) and np.array_equal(v, vo)


@pytest.mark.parametrize("component", [0, 1])
@pytest.mark.parametrize("plane", [False, True])
@pytest.mark.parametrize("layers", [1, 3])
def test_constraint_codes_match_oracle(orc, component, plane, layers):
    import vpd
    x, _ = vpd.make_patch(13, 9, 4, 0.5)
    codes, anchor, pin = vpd.regional_constraint_codes(x, component, layers, 0.5, plane)
    c3, a, p = orc.constraint_codes(x, component, layers, 0.5, plane)
    assert np.array_equal(vpd.unpack_codes(codes), c3)
    assert (anchor, pin) == (a, p)


def test_voided_patch(orc):
    """Seeded site removal keeps the survivors' lattice order and volumes; the boundary bands
    of the voided set agree with the oracle's."""
    import vpd
    x, vol = vpd.voided_patch(30, 8, 8, 1.0, 0.05, 11)
    xf, vf = vpd.make_patch(30, 8, 8, 1.0)
    assert len(x) == 1920 - 96
    keys = {tuple(p): k for k, p in enumerate(xf)}
    idx = np.array([keys[tuple(p)] for p in x])
    assert (np.diff(idx) > 0).all() and np.array_equal(vf[idx], vol)
    assert np.array_equal(vpd.voided_patch(30, 8, 8, 1.0, 0.05, 11)[0], x)
    assert not np.array_equal(vpd.voided_patch(30, 8, 8, 1.0, 0.05, 12)[0], x)
    codes, anchor, pin = vpd.regional_constraint_codes(x, 0, 3, 1.0)
    c3, a, p = orc.constraint_codes(x, 0, 3, 1.0)
    assert np.array_equal(vpd.unpack_codes(codes), c3) and (anchor, pin) == (a, p)


def test_ramp_matches_oracle(orc):
    import vpd
    for t in (0.0, 1e-5, 0.03, 0.0599, 0.06, 0.2):
        assert np.array_equal(vpd.smooth_ramp(2.0, 0.06, t), orc.ramp(2.0, 0.06, t))


def test_create_without_device_fails_loudly():
    import vpd
    if vpd.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(vpd.VpdError, match="No HIP device"):
        vpd.Context(1.0, 1.0, [1.0] * 14)


def test_config_rejections_precede_device():
    import vpd
    with pytest.raises(vpd.VpdError, match="horizon > 0"):
        vpd.Context(0.0, 1.0, [1.0] * 14)
    with pytest.raises(vpd.VpdError, match="density > 0"):
        vpd.Context(1.0, -1.0, [1.0] * 14)
    with pytest.raises(vpd.VpdError, match="14 props and 6 states"):
        vpd.Context(1.0, 1.0, [1.0] * 13)
    with pytest.raises(vpd.VpdError, match="initiation > 0"):
        vpd.Context(1.0, 1.0, [1.0] * 14, damage=(0.0, 0.0))


def test_no_alloc_neighbors_refuse_before_any_device_work():
    # Refusals of the allocation-free neighbour build are decided on the host, before the
    # pointers are used.
    import ctypes
    import vpd
    L = vpd.lib()
    dummy = ctypes.c_void_p(1)
    nnz = ctypes.c_int64()
    assert L.vpd_build_neighbors_noalloc(dummy, (1 << 21) + 1, 1.0, dummy, dummy, 10,
                                         ctypes.byref(nnz), None) == 1
    assert b"at most 2097152" in L.vpd_last_error(None)
    assert L.vpd_build_neighbors_noalloc(dummy, 10, 0.0, dummy, dummy, 10,
                                         ctypes.byref(nnz), None) == 1
    assert L.vpd_build_neighbors_noalloc(None, 10, 1.0, dummy, dummy, 10,
                                         ctypes.byref(nnz), None) == 1


def test_device_vumat_rejections_precede_device():
    # vpd_vumat_adapter_call_device refuses what the Fortran routine STOPs on
    # (JohnsonCook.f90:52-53) and what VumatMaterialAdapter::updateBlock refuses
    # (VumatMaterialAdapter.cpp:150-151), before any device pointer is used.
    import ctypes as C
    import math
    import vpd
    L = vpd.lib()

    def call(nb=4, ndir=3, nshr=3, nstatev=6, nfieldv=1, nprops=14, dt=1e-4, strain=True):
        ints = [C.c_int(v) for v in (nb, ndir, nshr, nstatev, nfieldv, nprops, 0)]
        dbl = [C.c_double(v) for v in (0.0, 0.0, dt)]
        p = C.c_void_p(16)
        arrays = [p] * 22
        if not strain:
            arrays[4] = None  # strainInc
        return L.vpd_vumat_adapter_call_device(*[C.byref(v) for v in ints],
                                               *[C.byref(d) for d in dbl],
                                               b"JOHNSONCOOK".ljust(80), *arrays)

    def err():
        return L.vpd_last_error(None).decode()

    assert call(ndir=2) == 1 and err() == "JohnsonCook VUMAT requires ndir=3,nshr=3"
    assert call(nshr=2) == 1 and err() == "JohnsonCook VUMAT requires ndir=3,nshr=3"
    assert call(nstatev=5) == 1 and err() == "JohnsonCook VUMAT requires 14 props and 6 states"
    assert call(nprops=13) == 1 and err() == "JohnsonCook VUMAT requires 14 props and 6 states"
    for dt in (math.nan, 0.0, -1e-4, math.inf):
        assert call(dt=dt) == 1 and err() == "VUMAT update requires dt > 0."
    assert call(strain=False) == 1 and err() == "Null VUMAT array."
    assert call(nb=0) == 0  # an empty block is a no-op


def test_velocity_band_closed_form(orc):
    """The velocity-driven end band (u = v·t, v, a = 0 at t = step·dt) against its closed form
    after k steps of the oracle bar: the band's axial displacement, velocity and acceleration,
    the support band held at zero, and the interior actually moving."""
    from conftest import JC_PROPS
    v, dt, k = 1.5, 8.5377e-5, 30
    bar = orc.OracleBar(nx=16, ny=4, nz=4, spacing=1.0, horizon=3.015, props=JC_PROPS, dt=dt,
                        density=7.85e-3, amplitude=0.0, rise_time=1.0, velocity=v)
    bar.step(k)
    codes = bar.codes[:, 0]
    band, support = codes == 2, codes == 1
    assert band.sum() == support.sum() == 3 * 4 * 4
    t = k * dt
    assert np.array_equal(bar.u[band, 0], np.full(band.sum(), v * t))
    assert np.array_equal(bar.v[band, 0], np.full(band.sum(), v))
    assert np.array_equal(bar.a[band, 0], np.zeros(band.sum()))
    assert not bar.u[support, 0].any() and not bar.v[support, 0].any()
    inner = ~(band | support)
    assert 0.0 < np.abs(bar.u[inner, 0]).max() < v * t
