"""The allocation-free neighbour build (vpd_build_neighbors_noalloc) against the allocating one
(vpd_build_neighbors), through the C ABI: identical CSR rows, bit for bit, on two calls in a
row, the second with a smaller input (the shared work memory keeps the first call's contents
and must not leak into the second)."""
import ctypes as C

import numpy as np
import pytest

import vpd
from conftest import JC_PROPS

pytestmark = pytest.mark.gpu


def lattice(nx, ny, nz, h, jitter, seed):
    x, vol = vpd.make_patch(nx, ny, nz, h)
    if jitter:
        x = x + np.random.default_rng(seed).uniform(-jitter * h, jitter * h, x.shape)
    return x, vol


def allocating(x, vol, horizon):
    ctx = vpd.Context(horizon, 7.85e-3, JC_PROPS)
    ctx.set_particles(x, vol)
    ctx.build_neighbors()
    rp, col = ctx.get_neighbors()
    ctx.close()
    return rp, col


def no_alloc(x, horizon, capacity=None):
    """Returns (status, nnz, rowptr, col) of one call; capacity None sizes col from a first
    call with no col array."""
    L = vpd.lib()
    n = x.shape[0]
    xd = vpd.DeviceBuffer(np.ascontiguousarray(x.T))  # planes x[d*n + i]
    rd = vpd.DeviceBuffer(np.zeros(n + 1, np.int32), np.int32)
    nnz = C.c_int64(-1)
    if capacity is None:
        st = L.vpd_build_neighbors_noalloc(xd.ptr, n, horizon, rd.ptr, None, 0, C.byref(nnz), None)
        assert st == 1 and nnz.value > 0, (st, nnz.value)  # VPD_ERR_ARGUMENT, count reported
        assert b"capacity too small" in L.vpd_last_error(None)
        capacity = nnz.value
    cd = vpd.DeviceBuffer(np.full(max(capacity, 1), -7, np.int32), np.int32)
    st = L.vpd_build_neighbors_noalloc(xd.ptr, n, horizon, rd.ptr, cd.ptr, capacity,
                                       C.byref(nnz), None)
    return st, nnz.value, rd.read(), cd.read()


def test_no_alloc_neighbors_match_allocating_build():
    cases = [((40, 12, 12, 0.5, 0.1, 1), 3.015 * 0.5),   # 5760 jittered points
             ((20, 6, 6, 1.0, 0.0, 0), 2.015)]            # then a smaller lattice
    for (dims, horizon) in cases:
        x, vol = lattice(*dims)
        print("no-alloc neighbours: %d particles, horizon %g" % (len(x), horizon), flush=True)
        rp, col = allocating(x, vol, horizon)
        st, nnz, rp2, col2 = no_alloc(x, horizon)
        assert st == 0, vpd.lib().vpd_last_error(None)
        assert nnz == len(col) == rp[-1]
        assert np.array_equal(rp2, rp)
        assert np.array_equal(col2[:nnz], col)


def test_no_alloc_refuses_and_reports():
    L = vpd.lib()
    x, _ = lattice(6, 5, 4, 1.0, 0.0, 0)
    st, nnz, rp, _ = no_alloc(x, 1.5, capacity=3)  # too small: count reported, col untouched
    assert st == 1 and nnz > 3 and rp[-1] == nnz
    dummy = C.c_void_p(1)
    nnz = C.c_int64()
    assert L.vpd_build_neighbors_noalloc(dummy, (1 << 21) + 1, 1.0, dummy, dummy, 10,
                                         C.byref(nnz), None) == 1
    assert b"at most 2097152" in L.vpd_last_error(None)
