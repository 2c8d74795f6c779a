"""ORACLE — test infrastructure only.

numpy/ctypes front end of the CPU restatement (oracle/liboracle.so, built from
oracle/vpd_oracle.cpp and oracle/jc_vumat.cpp) and of the reference itself
(oracle/_ref/libvpd_ref.so + oracle/_ref/jc_vumat_plugin.so, built from the reference
sources by `make -C oracle ref`). Only tests, the smoke check and the CPU-baseline leg
of the benchmark import this module; the product never does.

Layouts are the reference's: points as (n, 3) rows, tensors as (n, 9) row-major rows,
neighbours as CSR (rowptr int32[n+1], col int32[nnz]) with rows sorted ascending.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")
REF_LIB = os.path.join(REF_DIR, "libvpd_ref.so")
REF_PLUGIN = os.path.join(REF_DIR, "jc_vumat_plugin.so")
ORACLE_PLUGIN = os.path.join(HERE, "libjc_vumat_oracle.so")

_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_I = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_U8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_L = C.c_long
_lib = None
_ref = None


def build(ref: bool = False) -> None:
    """Compile the oracle (and, where the reference sources exist, oracle/_ref)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if ref and os.path.isdir(os.environ.get("VPD_REFERENCE", "/root/reference")):
        subprocess.run(["make", "-s", "-C", HERE, "ref",
                        "REF=" + os.environ.get("VPD_REFERENCE", "/root/reference")], check=True)


def _nullable(dtype):
    base = np.ctypeslib.ndpointer(dtype=dtype, flags="C_CONTIGUOUS")

    def from_param(cls, obj):
        if obj is None:
            return None
        return base.from_param(obj)

    return type("Nullable" + base.__name__, (base,), {"from_param": classmethod(from_param)})


_DN = _nullable(np.float64)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_last_error.restype = C.c_char_p
        L.orc_num_threads.restype = C.c_int
        L.orc_build_neighbors.argtypes = [_D, _L, C.c_double, _I, _nullable(np.int32), _L]
        L.orc_build_neighbors.restype = _L
        L.orc_build_neighbors_brute.argtypes = [_D, _L, C.c_double, _I, _nullable(np.int32), _L]
        L.orc_build_neighbors_brute.restype = _L
        L.orc_shape_inverse.argtypes = [_D, _D, _L, _I, _I, _D]
        L.orc_deformation_gradient.argtypes = [_D, _D, _D, _L, _I, _I, _D, _D]
        L.orc_force_density.argtypes = [_D, _D, _L, _I, _I, _D, _D, _D, _DN, _D, _DN]
        L.orc_corotational.argtypes = [_D, _D, _D, _D, C.c_double, _L, _DN, _DN, _DN, _DN, _DN]
        L.orc_rotate_to_spatial.argtypes = [_D, _D, _L, _D]
        L.orc_green_lagrange.argtypes = [_D, _L, _D]
        L.orc_material_init.argtypes = [_D, _L, C.c_int, _DN, C.c_double] + [_D] * 9
        L.orc_material_update.argtypes = [_D, _L, C.c_double, C.c_int, C.c_int, _D] + [_D] * 10
        L.orc_damage_update.argtypes = [_D, _D, _D, _L, _I, _I, C.c_double, C.c_double, _D, _DN]
        L.orc_predict.argtypes = [_D, _D, _D, _L, C.c_double]
        L.orc_predict.restype = None
        L.orc_correct.argtypes = [_D, _D, _D, _L, C.c_double, C.c_double]
        L.orc_correct.restype = None
        L.orc_make_patch.argtypes = [C.c_int] * 3 + [C.c_double, _D, _D]
        L.orc_make_patch.restype = None
        L.orc_select_region.argtypes = [_D, _L, C.c_int, C.c_double, _U8]
        L.orc_constraint_codes.argtypes = [_D, _L, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                           C.c_int, _U8, C.POINTER(C.c_long), C.POINTER(C.c_long)]
        L.orc_ramp.argtypes = [C.c_double, C.c_double, C.c_double, _D]
        L.orc_ramp.restype = None
        L.orc_apply_constraints.argtypes = [_U8, _L, _D, _D, _D, _D]
        L.orc_apply_constraints.restype = None
        L.orc_step.argtypes = ([_D, _D, _L, _I, _I, _D, _U8, _D, C.c_double, C.c_double, C.c_int,
                                C.c_int, _D, C.c_double, C.c_double] + [_D] * 12 + [_DN] +
                               [_D] * 3)
        L.orc_vumat_jc.argtypes = [C.c_int] * 6 + [_D] * 11 + [_D] * 8
        L.orc_vumat_jc_tol.argtypes = [C.c_int] * 6 + [_D] * 11 + [_D] * 8 + [C.c_double,
                                                                            C.c_int]
        _lib = L
    return _lib


def _check(rc, getter):
    if rc != 0:
        raise RuntimeError(getter().decode())


def _err():
    return lib().orc_last_error()


def f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


# ----------------------------------------------------------------------------- oracle
def build_neighbors(x, horizon, brute=False):
    x = f64(x)
    n = x.shape[0]
    rowptr = np.zeros(n + 1, np.int32)
    fn = lib().orc_build_neighbors_brute if brute else lib().orc_build_neighbors
    nnz = fn(x, n, horizon, rowptr, None, 0)
    if nnz < 0:
        raise RuntimeError(_err().decode())
    col = np.zeros(max(nnz, 1), np.int32)
    fn(x, n, horizon, rowptr, col, nnz)
    return rowptr, col[:nnz]


def shape_inverse(x, vol, rowptr, col):
    n = x.shape[0]
    out = np.zeros((n, 9))
    _check(lib().orc_shape_inverse(f64(x), f64(vol), n, rowptr, col, out), _err)
    return out


def deformation_gradient(x, u, vol, rowptr, col, kinv):
    n = x.shape[0]
    out = np.zeros((n, 9))
    _check(lib().orc_deformation_gradient(f64(x), f64(u), f64(vol), n, rowptr, col, f64(kinv),
                                          out), _err)
    return out


def force_density(x, vol, rowptr, col, kinv, F, sigma, bond_damage=None, want_map=False):
    n = x.shape[0]
    out = np.zeros((n, 3))
    amap = np.zeros((n, 9)) if want_map else None
    _check(lib().orc_force_density(f64(x), f64(vol), n, rowptr, col, f64(kinv), f64(F), f64(sigma),
                                   None if bond_damage is None else f64(bond_damage), out, amap),
           _err)
    return (out, amap) if want_map else out


def corotational(fN, fNP1, vN, rN, dt):
    n = fN.shape[0]
    outs = [np.zeros((n, 9)) for _ in range(5)]
    _check(lib().orc_corotational(f64(fN), f64(fNP1), f64(vN), f64(rN), dt, n, *outs), _err)
    return dict(zip(["L", "Dbar", "relSpin", "R", "V"], outs))


def rotate_to_spatial(s, r):
    out = np.zeros((s.shape[0], 9))
    lib().orc_rotate_to_spatial(f64(s), f64(r), s.shape[0], out)
    return out


def green_lagrange(F):
    out = np.zeros((F.shape[0], 9))
    lib().orc_green_lagrange(f64(F), F.shape[0], out)
    return out


def vumat_jc(props, dt, strain_inc, stress_old, state_old, temp_old, eint_old, einel_old,
             nstatev=6, rtol=0.0, max_iter=50):
    """JohnsonCook.f90 on a column-major block: strain_inc/stress_old (6, nb), state (nstatev, nb).
    rtol > 0 runs the bisection to |r| < rtol·q_trial within max_iter iterations."""
    nb = strain_inc.shape[1]
    z6, z9, z1 = np.zeros((6, nb)), np.zeros((9, nb)), np.zeros((1, nb))
    out = dict(temp=np.zeros(nb), stretch=np.zeros((6, nb)), defgrad=np.zeros((9, nb)),
               field=np.zeros((1, nb)), stress=np.zeros((6, nb)), state=np.zeros((nstatev, nb)),
               eint=np.zeros(nb), einel=np.zeros(nb))
    rc = lib().orc_vumat_jc_tol(nb, 3, 3, nstatev, 1, len(props), f64([dt]), f64(props),
                                f64(strain_inc), f64(temp_old), z6, z9, z1, f64(stress_old),
                                f64(state_old), f64(eint_old), f64(einel_old), out["temp"],
                                out["stretch"], out["defgrad"], out["field"], out["stress"],
                                out["state"], out["eint"], out["einel"], rtol, max_iter)
    if rc:
        raise RuntimeError("JohnsonCook VUMAT requires ndir=3,nshr=3, 14 props and 6 states")
    return out


def make_patch(nx, ny, nz, h):
    n = nx * ny * nz
    x, vol = np.zeros((n, 3)), np.zeros(n)
    lib().orc_make_patch(nx, ny, nz, h, x, vol)
    return x, vol


def constraint_codes(x, load_comp, region_layers, spacing, plane_strain=False):
    n = x.shape[0]
    codes = np.zeros((n, 3), np.uint8)
    anchor, pin = C.c_long(), C.c_long()
    sup, ld = (0, 1) if load_comp == 0 else (2, 3)
    _check(lib().orc_constraint_codes(f64(x), n, load_comp, sup, ld, region_layers, spacing,
                                      int(plane_strain), codes, C.byref(anchor), C.byref(pin)),
           _err)
    return codes, anchor.value, pin.value


def ramp(amplitude, rise_time, t):
    out = np.zeros(3)
    lib().orc_ramp(amplitude, rise_time, t, out)
    return out


MATERIAL_FIELDS = ["fN", "vN", "rN", "strainN", "stressN", "statev", "temp", "eint", "einel"]


def material_init(F0, nstatev, tref, initial_state=None):
    n = F0.shape[0]
    st = dict(fN=np.zeros((n, 9)), vN=np.zeros((n, 9)), rN=np.zeros((n, 9)),
              strainN=np.zeros((n, 9)), stressN=np.zeros((n, 9)),
              statev=np.zeros((n, nstatev)), temp=np.zeros(n), eint=np.zeros(n),
              einel=np.zeros(n))
    _check(lib().orc_material_init(f64(F0), n, nstatev,
                                   None if initial_state is None else f64(initial_state), tref,
                                   *[st[k] for k in MATERIAL_FIELDS]), _err)
    return st


def material_update(F, st, dt, props):
    """VumatPdCorrespondenceMaterial::update; st is updated in place. Returns spatial σ."""
    n = F.shape[0]
    sig = np.zeros((n, 9))
    _check(lib().orc_material_update(f64(F), n, dt, st["statev"].shape[1], len(props), f64(props),
                                     *[st[k] for k in MATERIAL_FIELDS], sig), _err)
    return sig


def damage_update(x, u, vol, rowptr, col, initiation, failure, damage, want_index=False):
    n = x.shape[0]
    idx = np.zeros(n) if want_index else None
    _check(lib().orc_damage_update(f64(x), f64(u), f64(vol), n, rowptr, col, initiation, failure,
                                   damage, idx), _err)
    return idx


class OracleBar:
    """The reference bar run (BoundaryValueCaseRunner.cpp:142-229) on the restated oracle."""

    def __init__(self, nx, ny, nz, spacing, horizon, props, dt, density, amplitude, rise_time,
                 component=0, region_layers=3, plane_strain=False, nstatev=6, tref=293.0,
                 initial_state=None, damage=None, velocity=None, rowptr=None, col=None,
                 points=None):
        self.x, self.vol = make_patch(nx, ny, nz, spacing) if points is None else points
        self.n = self.x.shape[0]
        if rowptr is None:
            rowptr, col = build_neighbors(self.x, horizon)
        self.rowptr, self.col = rowptr, col
        self.nnz = len(col)
        self.kinv = shape_inverse(self.x, self.vol, rowptr, col)
        self.codes, self.anchor, self.pin = constraint_codes(self.x, component, region_layers,
                                                             spacing, plane_strain)
        self.props = f64(props)
        self.dt, self.density = dt, density
        self.amplitude, self.rise_time, self.velocity = amplitude, rise_time, velocity
        self.u = np.zeros((self.n, 3))
        self.v = np.zeros((self.n, 3))
        self.a = np.zeros((self.n, 3))
        F0 = deformation_gradient(self.x, self.u, self.vol, rowptr, col, self.kinv)
        self.mat = material_init(F0, nstatev, tref, initial_state)
        self.damage_cfg = damage
        self.damage = np.zeros(len(col)) if damage is not None else None
        self.F = np.zeros((self.n, 9))
        self.sigma = np.zeros((self.n, 9))
        self.force = np.zeros((self.n, 3))
        self.step_index = 0
        lib().orc_apply_constraints(self.codes, self.n, self.ramp(0.0), self.u, self.v, self.a)

    def ramp(self, t):
        if self.velocity is not None:
            return f64([self.velocity * t, self.velocity, 0.0])
        return ramp(self.amplitude, self.rise_time, t)

    def get(self, name):
        """State field `name` (RefBar's names)."""
        if name in ("u", "v", "a"):
            return getattr(self, name).copy()
        if name == "damage":
            return self.damage.copy()
        return self.mat[name].copy()

    def set(self, name, arr):
        """Overwrites state field `name`; F_n also resets E_n = greenLagrangeStrain(F_n)."""
        a = np.array(arr, dtype=np.float64)
        if name in ("u", "v", "a"):
            getattr(self, name)[...] = a.reshape(self.n, 3)
        elif name == "damage":
            self.damage[...] = a
        else:
            self.mat[name][...] = a.reshape(self.mat[name].shape)
            if name == "fN":
                self.mat["strainN"][...] = green_lagrange(self.mat["fN"])

    def set_step(self, step):
        self.step_index = int(step)

    def step(self, k=1):
        for _ in range(k):
            self.step_index += 1
            t = self.step_index * self.dt
            ini, fail = self.damage_cfg if self.damage_cfg is not None else (0.0, 0.0)
            m = self.mat
            _check(lib().orc_step(self.x, self.vol, self.n, self.rowptr, self.col, self.kinv,
                                  self.codes, self.ramp(t), self.dt, self.density,
                                  m["statev"].shape[1], len(self.props), self.props, ini, fail,
                                  self.u, self.v, self.a, *[m[k] for k in MATERIAL_FIELDS],
                                  self.damage, self.F, self.sigma, self.force), _err)


# ----------------------------------------------------------------------------- reference
def ref_available() -> bool:
    return os.path.exists(REF_LIB) and os.path.exists(REF_PLUGIN)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError("oracle/_ref is not built (needs the reference sources)")
        R = C.CDLL(REF_LIB)
        R.ref_last_error.restype = C.c_char_p
        R.ref_build_neighbors.argtypes = [_D, _L, C.c_double, C.c_int, _I, _nullable(np.int32), _L]
        R.ref_build_neighbors.restype = _L
        R.ref_shape_inverse.argtypes = [_D, _D, _L, _I, _I, C.c_double, _D]
        R.ref_deformation_gradient.argtypes = [_D, _D, _D, _L, _I, _I, C.c_double, _D, _D]
        R.ref_force_density.argtypes = [_D, _D, _L, _I, _I, C.c_double, _D, _D, _D, _DN, _D]
        R.ref_corotational.argtypes = [_D, _D, _D, _D, C.c_double, _L] + [_D] * 5
        R.ref_vumat_block.argtypes = ([C.c_char_p, C.c_int, C.c_int, _D, C.c_double, _L,
                                       C.c_double] + [_D] * 11)
        R.ref_case_create.restype = C.c_void_p
        R.ref_case_create.argtypes = ([C.c_int] * 3 + [C.c_double] * 2 + [C.c_int] * 2 +
                                      [C.c_double] * 2 + [C.c_int, C.c_char_p, C.c_int, C.c_int,
                                                          _D, C.c_double, C.c_double, _DN,
                                                          C.c_double, C.c_int, C.c_double,
                                                          C.c_double, C.c_double])
        R.ref_case_destroy.argtypes = [C.c_void_p]
        R.ref_case_destroy.restype = None
        R.ref_case_size.argtypes = [C.c_void_p, C.POINTER(C.c_long)]
        R.ref_case_size.restype = _L
        R.ref_case_step.argtypes = [C.c_void_p, C.c_int]
        R.ref_case_get.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        R.ref_case_set.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        R.ref_case_set_step.argtypes = [C.c_void_p, C.c_int]
        R.ref_case_set_step.restype = None
        R.ref_run_case_json.argtypes = [C.c_char_p, C.c_char_p]
        _ref = R
    return _ref


def _rerr():
    return ref().ref_last_error()


def ref_build_neighbors(x, horizon, brute=False):
    x = f64(x)
    n = x.shape[0]
    rowptr = np.zeros(n + 1, np.int32)
    nnz = ref().ref_build_neighbors(x, n, horizon, int(brute), rowptr, None, 0)
    if nnz < 0:
        raise RuntimeError(_rerr().decode())
    col = np.zeros(max(nnz, 1), np.int32)
    ref().ref_build_neighbors(x, n, horizon, int(brute), rowptr, col, nnz)
    return rowptr, col[:nnz]


def ref_shape_inverse(x, vol, rowptr, col, horizon):
    out = np.zeros((x.shape[0], 9))
    _check(ref().ref_shape_inverse(f64(x), f64(vol), x.shape[0], rowptr, col, horizon, out), _rerr)
    return out


def ref_deformation_gradient(x, u, vol, rowptr, col, kinv, horizon):
    out = np.zeros((x.shape[0], 9))
    _check(ref().ref_deformation_gradient(f64(x), f64(u), f64(vol), x.shape[0], rowptr, col,
                                          horizon, f64(kinv), out), _rerr)
    return out


def ref_force_density(x, vol, rowptr, col, kinv, F, sigma, horizon, bond_damage=None):
    out = np.zeros((x.shape[0], 3))
    _check(ref().ref_force_density(f64(x), f64(vol), x.shape[0], rowptr, col, horizon, f64(kinv),
                                   f64(F), f64(sigma),
                                   None if bond_damage is None else f64(bond_damage), out), _rerr)
    return out


def ref_corotational(fN, fNP1, vN, rN, dt):
    n = fN.shape[0]
    outs = [np.zeros((n, 9)) for _ in range(5)]
    _check(ref().ref_corotational(f64(fN), f64(fNP1), f64(vN), f64(rN), dt, n, *outs), _rerr)
    return dict(zip(["L", "Dbar", "relSpin", "R", "V"], outs))


def ref_vumat_block(props, dt, strain_rate_full, stress_full, state_old, temp, eint, einel,
                    plugin=None, material_density=7850.0):
    n = strain_rate_full.shape[0]
    nst = state_old.shape[1]
    out = dict(stress=np.zeros((n, 9)), state=np.zeros((n, nst)), temp=np.zeros(n),
               eint=np.zeros(n), einel=np.zeros(n))
    _check(ref().ref_vumat_block((plugin or REF_PLUGIN).encode(), nst, len(props), f64(props),
                                 material_density, n, dt, f64(strain_rate_full), f64(stress_full),
                                 f64(state_old), f64(temp), f64(eint), f64(einel), out["stress"],
                                 out["state"], out["temp"], out["eint"], out["einel"]), _rerr)
    return out


REF_FIELDS = {"x": (0, 3), "vol": (1, 1), "u": (2, 3), "v": (3, 3), "a": (4, 3), "fN": (5, 9),
              "vN": (6, 9), "rN": (7, 9), "stressN": (8, 9), "statev": (9, None),
              "temp": (10, 1), "eint": (11, 1), "einel": (12, 1), "sigma": (13, 9),
              "strain": (14, 9), "force": (15, 3), "damage": (16, "nnz"),
              "damage_index": (17, 1), "kinv": (18, 9)}


class RefBar:
    """The reference's own classes driven in the order of BoundaryValueCaseRunner.cpp:207-229."""

    def __init__(self, nx, ny, nz, spacing, horizon, props, dt, density, amplitude, rise_time,
                 component=0, region_layers=3, plane_strain=False, nstatev=6, tref=293.0,
                 initial_state=None, damage=None, material_density=7850.0, plugin=None):
        ini, fail = damage if damage is not None else (0.0, 1.0)
        self.h = ref().ref_case_create(nx, ny, nz, spacing, horizon, component, region_layers,
                                       amplitude, rise_time, int(plane_strain),
                                       (plugin or REF_PLUGIN).encode(), nstatev, len(props),
                                       f64(props), material_density, tref,
                                       None if initial_state is None else f64(initial_state),
                                       density, 1 if damage is not None else 0, ini, fail, dt)
        if not self.h:
            raise RuntimeError(_rerr().decode())
        nnz = C.c_long()
        self.n = ref().ref_case_size(self.h, C.byref(nnz))
        self.nnz = nnz.value
        self.nstatev = nstatev

    def __del__(self):
        if getattr(self, "h", None):
            ref().ref_case_destroy(self.h)
            self.h = None

    def step(self, k=1):
        _check(ref().ref_case_step(self.h, k), _rerr)

    def get(self, name):
        fid, w = REF_FIELDS[name]
        if w == "nnz":
            out = np.zeros(self.nnz)
        elif w is None:
            out = np.zeros((self.n, self.nstatev))
        elif w == 1:
            out = np.zeros(self.n)
        else:
            out = np.zeros((self.n, w))
        if ref().ref_case_get(self.h, fid, out.ctypes.data_as(C.c_void_p)) != 0:
            raise RuntimeError("field %s not available yet" % name)
        return out

    SETTABLE = ("u", "v", "a", "fN", "vN", "rN", "stressN", "statev", "temp", "eint", "einel",
                "damage")

    def set(self, name, arr):
        """Overwrites state field `name` (a field a step reads; shapes as `get` returns)."""
        if name not in self.SETTABLE:
            raise ValueError("field %s is not settable" % name)
        fid, _ = REF_FIELDS[name]
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.size != self.get(name).size:
            raise ValueError("field %s: wrong size %d" % (name, a.size))
        _check(ref().ref_case_set(self.h, fid, a.ctypes.data_as(C.c_void_p)), _rerr)

    def set_step(self, step):
        """The number of completed steps: the next `step` call runs step + 1."""
        ref().ref_case_set_step(self.h, int(step))

    def neighbors(self):
        rowptr = np.zeros(self.n + 1, np.int32)
        col = np.zeros(self.nnz, np.int32)
        ref().ref_case_get(self.h, 19, rowptr.ctypes.data_as(C.c_void_p))
        ref().ref_case_get(self.h, 20, col.ctypes.data_as(C.c_void_p))
        return rowptr, col
