"""The VUMAT plugin ABI on device arrays (vpd_vumat_adapter_call_device), pinned as the
reference pins its bridge (tests/interface_contract_test.cpp:24-112 with the
tests/fixtures/ContractProbe.f90 plugin): literal per-component expectations over blocks of
1, 255, 256, 257 and 513 points, two steps with the history fed back, and no write outside the
block.

The Johnson-Cook routine (materials/examples/JohnsonCook.f90:74-89) copies stretchOld,
defgradOld, fieldOld and the state variables past the sixth into their *New arrays before its
update; with every component of every point distinct, a transposed, shifted or permuted
column-major layout ([k + nblock*c]) breaks one of those literal copies. The stress, the six
JC states, the temperature and the energies are compared with the oracle's restatement of the
Fortran routine on the same arrays (1e-12 relative; the device pow/log differ from libm by an
ulp at most). Rejections of malformed calls are host checks (tests/test_capi_host.py).
"""
import ctypes as C

import numpy as np
import pytest

import vpd
from conftest import JC_PROPS

pytestmark = pytest.mark.gpu

GUARD = 12345.678


class Dev:
    """A device copy of a column-major block with one guard element past its end."""

    def __init__(self, host):
        host = np.ascontiguousarray(host, dtype=np.float64)
        self.shape = host.shape
        buf = np.concatenate([host.ravel(), [GUARD]])
        self.buf = vpd.DeviceBuffer(buf)

    @property
    def ptr(self):
        return self.buf.ptr

    def read(self):
        a = self.buf.read()
        return a[:-1].reshape(self.shape), a[-1]


def block(nb, nstatev, nfieldv, seed):
    """Distinct values in every component of every point; a mix of initialised-plastic,
    elastic and first-call (state 1 = 0) points."""
    rng = np.random.default_rng(seed)
    k = np.arange(nb)
    inc = np.empty((6, nb))
    for c in range(6):
        inc[c] = (1.0 + 0.1 * c) * 1e-3 * (1.0 + 0.001 * k) * (1 if c < 3 else 0.5)
    inc[:, k % 3 == 1] *= 1e-3  # small increments: elastic
    stress = np.stack([(10.0 * (c + 1) + 0.01 * k) * (-1) ** c for c in range(6)])
    state = np.zeros((nstatev, nb))
    state[0] = (k % 7 != 3).astype(float)  # every 7th point takes the init branch
    state[1] = 0.01 + 1e-5 * k
    state[2] = 293.0 + 0.1 * k
    state[3] = 360.0 + 0.01 * k
    state[4] = 1e-4 * (1 + 0.001 * k)
    state[5] = 1.0 + k
    for c in range(6, nstatev):
        state[c] = 1000.0 * c + k  # carried unchanged
    arrays = dict(
        coord=np.stack([k + 0.25 * d for d in range(3)]).astype(float),
        charlen=2.0 + 0.001 * k,
        density=np.full(nb, 7850.0),
        strain=inc,
        spin=np.stack([0.1 * (d + 1) + 1e-4 * k for d in range(3)]),
        temp=300.0 + 0.5 * k,
        stretch=np.stack([1.0 + 0.01 * c + 1e-6 * k for c in range(6)]),
        defgrad=np.stack([10.0 * (c + 1) + 1e-3 * k for c in range(9)]),
        field=np.stack([-(c + 1) - 1e-3 * k for c in range(max(1, nfieldv))]),
        stress=stress, state=state,
        eint=10.0 + 1e-3 * k, einel=20.0 + 1e-3 * k)
    return arrays


def call(props, dt, a, nstatev, nfieldv, step_time, total_time):
    nb = a["strain"].shape[1]
    ins = {name: Dev(v) for name, v in a.items()}
    props_d = vpd.DeviceBuffer(np.asarray(props, float))
    outs = dict(temp=Dev(np.full(nb, -1.0)), stretch=Dev(np.full((6, nb), -2.0)),
                defgrad=Dev(np.full((9, nb), -3.0)),
                field=Dev(np.full((max(1, nfieldv), nb), -4.0)),
                stress=Dev(np.full((6, nb), -5.0)), state=Dev(np.full((nstatev, nb), -6.0)),
                eint=Dev(np.full(nb, -7.0)), einel=Dev(np.full(nb, -8.0)))
    ints = [C.c_int(v) for v in (nb, 3, 3, nstatev, nfieldv, len(props), 0)]
    dbl = [C.c_double(v) for v in (step_time, total_time, dt)]
    i, o = ins, outs
    rc = vpd.lib().vpd_vumat_adapter_call_device(
        *[C.byref(v) for v in ints], *[C.byref(d) for d in dbl], b"JOHNSONCOOK".ljust(80),
        i["coord"].ptr, i["charlen"].ptr, props_d.ptr, i["density"].ptr, i["strain"].ptr,
        i["spin"].ptr, i["temp"].ptr, i["stretch"].ptr, i["defgrad"].ptr, i["field"].ptr,
        i["stress"].ptr, i["state"].ptr, i["eint"].ptr, i["einel"].ptr, o["temp"].ptr,
        o["stretch"].ptr, o["defgrad"].ptr, o["field"].ptr, o["stress"].ptr, o["state"].ptr,
        o["eint"].ptr, o["einel"].ptr)
    assert rc == 0, vpd.lib().vpd_last_error(None)
    # Inputs untouched, guards intact.
    for name, d in ins.items():
        got, guard = d.read()
        assert np.array_equal(got, np.asarray(a[name], float).reshape(got.shape)), name
        assert guard == GUARD, name
    res = {}
    for name, d in o.items():
        got, guard = d.read()
        assert guard == GUARD, "write past the block: " + name
        res[name] = got
    return res


@pytest.mark.parametrize("nb", [1, 255, 256, 257, 513])
def test_device_vumat_abi_literals(orc, nb):
    print("[gpu-case] device VUMAT ABI, nblock %d" % nb, flush=True)
    nstatev, nfieldv = 9, 2
    dt = 8.5377e-5
    a = block(nb, nstatev, nfieldv, nb)
    for step in range(2):
        out = call(JC_PROPS, dt, a, nstatev, nfieldv, step * dt, 2.0 + step * dt)
        # Literal copies (JohnsonCook.f90:84-89): every component of every point.
        assert np.array_equal(out["stretch"], a["stretch"])
        assert np.array_equal(out["defgrad"], a["defgrad"])
        assert np.array_equal(out["field"], a["field"])
        assert np.array_equal(out["state"][6:], a["state"][6:])
        # The JC update itself against the oracle's restatement of the Fortran routine.
        ref = orc.vumat_jc(JC_PROPS, dt, a["strain"], a["stress"], a["state"][:6],
                           a["temp"], a["eint"], a["einel"])
        scale = max(np.abs(ref["stress"]).max(), 1.0)
        assert np.abs(out["stress"] - ref["stress"]).max() <= 1e-12 * scale
        for c in range(6):
            r = ref["state"][c]
            assert np.abs(out["state"][c] - r).max() <= 1e-12 * max(np.abs(r).max(), 1.0), c
        for name in ("temp", "eint", "einel"):
            r = ref[name]
            assert np.abs(out[name] - r).max() <= 1e-12 * max(np.abs(r).max(), 1.0), name
        init = a["state"][0] == 0.0
        # Init branch (JohnsonCook.f90:92-102): trial stress, state (1, 0, T_ref, 0, 0, 1).
        if init.any():
            assert np.array_equal(out["state"][:6, init].T,
                                  np.tile([1.0, 0.0, 293.0, 0.0, 0.0, 1.0], (init.sum(), 1)))
            assert (out["temp"][init] == 293.0).all()
        # Feed the history back, as the reference's two-step probe does.
        a = dict(a, stress=out["stress"], state=out["state"], temp=out["temp"],
                 eint=out["eint"], einel=out["einel"])
