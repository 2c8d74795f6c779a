"""Python front end of libvpd.so (include/vpd.h) — the MI355X explicit correspondence-PD
increment with the Johnson-Cook VUMAT.

The library is loaded from this directory and is required: there is no CPU fallback, and
every entry point raises VpdError when the HIP library is missing or a call fails.
Arrays use the reference's host layouts: points (n, 3), tensors (n, 9) row-major,
neighbours CSR int32 with ascending rows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvpd.so")

_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_I32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_U8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")

FIELD = dict(x=0, volume=1, u=2, v=3, a=4, fN=5, vN=6, rN=7, stressN=8, statev=9, temp=10,
             eint=11, einel=12, sigma=13, strain=14, force=15, damage=16, damage_index=17,
             kinv=18, F=19, iterations=20, strain_inc=21, fallbacks=22)
DAMAGE_NONE, DAMAGE_CRITICAL_STRETCH = 0, 1
MATERIAL_JOHNSON_COOK, MATERIAL_ELASTIC = 0, 1
RETURN_MAP_BISECTION, RETURN_MAP_BISECTION_TIGHT, RETURN_MAP_NEWTON, RETURN_MAP_LEARNED = 0, 1, 2, 3
RETURN_MAP_BISECTION_UNSCREENED = 4  # vpd_return_map_device only
RETURN_MAP_BISECTION_GROUPED = 5  # vpd_return_map_device only: mode 0 on lane groups
MLP_PARAMS, MLP_NORM = 8961, 18


class VpdError(RuntimeError):
    pass


class Config(C.Structure):
    _fields_ = [("horizon", C.c_double), ("density", C.c_double), ("nprops", C.c_int),
                ("props", C.POINTER(C.c_double)), ("nstatev", C.c_int),
                ("initial_state", C.POINTER(C.c_double)), ("reference_temperature", C.c_double),
                ("char_length", C.c_double), ("material_density", C.c_double),
                ("damage_mode", C.c_int), ("damage_initiation", C.c_double),
                ("damage_failure", C.c_double), ("return_map", C.c_int),
                ("material_model", C.c_int)]


class StepArgs(C.Structure):
    _fields_ = [("dt", C.c_double), ("time", C.c_double), ("step_time", C.c_double),
                ("ramp", C.c_double * 3), ("fuse_next", C.c_int), ("ramp_next", C.c_double * 3),
                ("want_fields", C.c_int)]


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char_p), ("total_ms", C.c_double), ("launches", C.c_int64)]


class DiagConfig(C.Structure):
    _fields_ = [("load_axis", C.c_int), ("lateral_axis", C.c_int),
                ("observation_weight", C.c_void_p), ("lateral_sets", C.c_void_p),
                ("eqps_index", C.c_int), ("indicator_count", C.c_int),
                ("indicator_indices", C.c_int * 8), ("indicator_min", C.c_int),
                ("indicator_one_minus", C.c_int), ("capacity", C.c_int)]


# Slots of a run-diagnostics record (enum vpd_diag_slot).
DIAG_SLOTS = ("internal_energy", "observed_stress", "observed_strain", "max_eqps",
              "max_indicator", "nonfinite_material", "net_force_x", "net_force_y", "net_force_z",
              "reaction", "kinetic", "lateral_neg", "lateral_pos", "nonfinite_force")
DIAG_STRIDE = 16


_lib = None


def build() -> None:
    """Compile libvpd.so for gfx950 (hipcc cross-compiles without a GPU)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise VpdError("libvpd.so is not built: run `make -C vumat-pd_amd` "
                           "(or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.vpd_last_error.restype = C.c_char_p
        L.vpd_last_error.argtypes = [P]
        L.vpd_create.argtypes = [C.POINTER(Config), C.POINTER(P)]
        L.vpd_destroy.argtypes = [P]
        L.vpd_destroy.restype = None
        L.vpd_device_count.argtypes = [C.POINTER(C.c_int)]
        L.vpd_set_particles.argtypes = [P, _D, _D, C.c_int64]
        L.vpd_build_neighbors.argtypes = [P, C.POINTER(C.c_int64)]
        L.vpd_set_neighbors.argtypes = [P, _I32, _I32, C.c_int64]
        L.vpd_get_neighbors.argtypes = [P, _I32, _I32]
        L.vpd_set_constraints.argtypes = [P, C.c_void_p]
        for f in ("vpd_initialize_shape", "vpd_initialize_material", "vpd_initialize_dynamics",
                  "vpd_compute_deformation_gradient", "vpd_synchronize"):
            getattr(L, f).argtypes = [P]
        L.vpd_upload.argtypes = [P, C.c_int, C.c_void_p]
        L.vpd_download.argtypes = [P, C.c_int, C.c_void_p]
        L.vpd_predict.argtypes = [P, C.c_double, _D]
        L.vpd_apply_constraints.argtypes = [P, _D]
        L.vpd_update_material.argtypes = [P, C.c_double, C.c_double, C.c_double]
        L.vpd_compute_force_and_correct.argtypes = [P, C.c_double, _D]
        L.vpd_step.argtypes = [P, C.POINTER(StepArgs)]
        L.vpd_set_specimens.argtypes = [P, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p]
        L.vpd_step_specimens.argtypes = [P, C.c_int32, C.c_int, C.c_int]
        L.vpd_run_specimens.argtypes = [P, C.c_int32, C.c_int32]
        L.vpd_set_timing.argtypes = [P, C.c_int]
        L.vpd_set_tuning.argtypes = [P, C.c_int, C.c_int]
        L.vpd_get_tuning.argtypes = [P, C.c_int, C.POINTER(C.c_int)]
        L.vpd_autotune.argtypes = [P, C.c_int, C.POINTER(C.c_int)]
        L.vpd_autotune_candidates.argtypes = [P, C.c_int, C.c_int, C.POINTER(C.c_int),
                                              C.POINTER(C.c_float), C.POINTER(C.c_int)]
        L.vpd_collect_timing.argtypes = [P, C.POINTER(KernelTime), C.c_int,
                                         C.POINTER(C.c_int)]
        L.vpd_make_patch.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _D, _D]
        L.vpd_regional_constraint_codes.argtypes = [_D, C.c_int64, C.c_int, C.c_int, C.c_double,
                                                    C.c_int, _U8, C.POINTER(C.c_int64),
                                                    C.POINTER(C.c_int64)]
        L.vpd_smooth_ramp.argtypes = [C.c_double, C.c_double, C.c_double, _D]
        L.vpd_smooth_ramp.restype = None
        L.vpd_select_boundary_region.argtypes = [_D, C.c_int64, C.c_char_p, C.c_double, _U8]
        L.vpd_select_boundary_region.restype = C.c_int64
        L.vpd_run_case_json.argtypes = [C.c_char_p, C.c_char_p]
        L.vpd_run_cases_json.argtypes = [C.POINTER(C.c_char_p), C.c_int, C.c_char_p]
        L.vpd_set_specimen_bands.argtypes = [P, C.c_void_p, C.c_void_p]
        L.vpd_enable_diagnostics.argtypes = [P, C.POINTER(DiagConfig)]
        L.vpd_fetch_diagnostics.argtypes = [P, _D, C.c_int, C.POINTER(C.c_int)]
        L.vpd_set_return_map.argtypes = [P, C.c_int, C.c_double]
        L.vpd_attach_warm_start.argtypes = [P, P]
        _F = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
        L.vpd_mlp_create.argtypes = [C.c_int, C.POINTER(P)]
        L.vpd_mlp_destroy.argtypes = [P]
        L.vpd_mlp_destroy.restype = None
        L.vpd_mlp_last_error.argtypes = [P]
        L.vpd_mlp_last_error.restype = C.c_char_p
        L.vpd_mlp_set_params.argtypes = [P, C.c_void_p, C.c_void_p]
        L.vpd_mlp_get_params.argtypes = [P, _F, _F]
        L.vpd_mlp_loss_and_gradient.argtypes = [P, _F, _F, C.c_int, C.POINTER(C.c_float), _F]
        L.vpd_mlp_set_dataset.argtypes = [P, _F, _F, C.c_int64, C.c_uint64]
        L.vpd_mlp_train.argtypes = [P, C.c_int, C.c_int, C.c_float, C.POINTER(C.c_float)]
        L.vpd_mlp_set_training_form.argtypes = [P, C.c_int]
        L.vpd_mlp_predict.argtypes = [P, _F, C.c_int64, _F]
        L.vpd_mlp_seed_device.argtypes = [P, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                          C.c_int64, C.c_void_p, C.POINTER(C.c_double)]
        L.vpd_return_map_device.argtypes = ([C.c_int, C.c_int64, C.c_int, _D, C.c_int, C.c_double,
                                             C.c_double] + [C.c_void_p] * 12 +
                                            [C.POINTER(C.c_double)])
        L.vpd_return_map_learned_device.argtypes = ([P, C.c_int64, C.c_int, _D, C.c_int,
                                                     C.c_double, C.c_double] + [C.c_void_p] * 12
                                                    + [C.POINTER(C.c_double)])
        L.vpd_build_neighbors_noalloc.argtypes = [C.c_void_p, C.c_int64, C.c_double, C.c_void_p,
                                                  C.c_void_p, C.c_int64,
                                                  C.POINTER(C.c_int64), C.c_void_p]
        L.vpd_device_malloc.argtypes = [C.c_uint64, C.POINTER(C.c_void_p)]
        L.vpd_device_free.argtypes = [C.c_void_p]
        L.vpd_memcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_int]
        L.vpd_vumat_adapter_call_device.argtypes = ([C.POINTER(C.c_int)] * 7 +
                                                    [C.POINTER(C.c_double)] * 3 + [C.c_char_p] +
                                                    [C.c_void_p] * 22)
        _lib = L
    return _lib


def device_count() -> int:
    n = C.c_int(0)
    lib().vpd_device_count(C.byref(n))
    return n.value


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


# ------------------------------------------------------------------ host helpers
def make_patch(nx, ny, nz, spacing):
    n = nx * ny * nz
    x, vol = np.zeros((n, 3)), np.zeros(n)
    if lib().vpd_make_patch(nx, ny, nz, spacing, x, vol):
        raise VpdError("Invalid patch dimensions.")
    return x, vol


def voided_patch(nx, ny, nz, spacing, fraction, seed):
    """The lattice of make_patch with round(fraction·n) sites removed at random (numpy
    default_rng(seed), without replacement): the voided bar of the high-rate tension case. The
    reference builds full patches only (src/host/Geometry.cpp:10-34); kept particles keep
    their lattice volumes and order."""
    x, vol = make_patch(nx, ny, nz, spacing)
    n = x.shape[0]
    keep = np.ones(n, bool)
    keep[np.random.default_rng(seed).choice(n, size=int(round(fraction * n)), replace=False)] = False
    return np.ascontiguousarray(x[keep]), np.ascontiguousarray(vol[keep])


def regional_constraint_codes(x, component, region_layers, spacing, plane_strain=False):
    n = x.shape[0]
    codes = np.zeros(n, np.uint8)
    anchor, pin = C.c_int64(), C.c_int64()
    if lib().vpd_regional_constraint_codes(_f64(x), n, component, region_layers, spacing,
                                           int(plane_strain), codes, C.byref(anchor),
                                           C.byref(pin)):
        raise VpdError("Boundary-region selection returned no particles.")
    return codes, anchor.value, pin.value


def unpack_codes(codes):
    """Packed per-particle codes -> (n, 3) array of 0 free / 1 zero / 2 ramp."""
    c = np.asarray(codes, np.uint8)
    return np.stack([(c >> (2 * d)) & 3 for d in range(3)], 1)


def pack_codes(codes3):
    c = np.asarray(codes3, np.uint8)
    return (c[:, 0] | (c[:, 1] << 2) | (c[:, 2] << 4)).astype(np.uint8)


def smooth_ramp(amplitude, rise_time, t):
    out = np.zeros(3)
    lib().vpd_smooth_ramp(amplitude, rise_time, t, out)
    return out


def select_boundary_region(x, side, width):
    """selectBoundaryRegion (ParticleSets.cpp:12-50) as a boolean mask."""
    x = _f64(x)
    mask = np.zeros(x.shape[0], np.uint8)
    count = lib().vpd_select_boundary_region(x, x.shape[0], side.encode(), width, mask)
    if count < 0:
        raise VpdError("boundary-region selection failed for side %r, width %r" % (side, width))
    return mask.astype(bool)


def run_cases_json(case_paths, plugin_directory=""):
    """vpd_run_cases_json: several boundary cases as one batch on the GPU; returns the exit
    status (0 all pass, 2 a material failed, 1 error)."""
    arr = (C.c_char_p * len(case_paths))(*[str(p).encode() for p in case_paths])
    return lib().vpd_run_cases_json(arr, len(case_paths), str(plugin_directory).encode())


def run_case_json(case_path, plugin_directory=""):
    """runDemoCaseFromJson on the GPU for a vumat_pd_boundary_case file: 0 pass, 2 fail,
    1 error (message on stderr)."""
    return lib().vpd_run_case_json(os.fsencode(case_path), os.fsencode(plugin_directory))


# ------------------------------------------------------------------ context
class Context:
    """One particle block on the GPU (PdCorrespondenceSkeleton + VumatPdCorrespondenceMaterial
    + PdSimpleDamage + ExplicitTimeIntegrator state behind one HIP stream)."""

    def __init__(self, horizon, density, props, nstatev=6, initial_state=None,
                 reference_temperature=293.0, char_length=1.0, material_density=7850.0,
                 damage=None, return_map=RETURN_MAP_BISECTION, material_model=0):
        self._props = _f64(props)
        self._init = None if initial_state is None else _f64(initial_state)
        cfg = Config()
        cfg.horizon = horizon
        cfg.density = density
        cfg.nprops = len(self._props)
        cfg.props = self._props.ctypes.data_as(C.POINTER(C.c_double))
        cfg.nstatev = nstatev
        cfg.initial_state = (None if self._init is None else
                             self._init.ctypes.data_as(C.POINTER(C.c_double)))
        cfg.reference_temperature = reference_temperature
        cfg.char_length = char_length
        cfg.material_density = material_density
        if damage is not None:
            cfg.damage_mode = DAMAGE_CRITICAL_STRETCH
            cfg.damage_initiation, cfg.damage_failure = damage
        cfg.return_map = return_map
        cfg.material_model = material_model
        self.h = C.c_void_p()
        self.n = 0
        self.nnz = 0
        self.nstatev = nstatev
        self.damage = damage is not None
        if lib().vpd_create(C.byref(cfg), C.byref(self.h)):
            raise VpdError(lib().vpd_last_error(None).decode())

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().vpd_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc:
            raise VpdError(lib().vpd_last_error(self.h).decode())

    def set_particles(self, x, volume):
        x = _f64(x)
        self.n = x.shape[0]
        self._check(lib().vpd_set_particles(self.h, x, _f64(volume), self.n))

    def build_neighbors(self):
        nnz = C.c_int64()
        self._check(lib().vpd_build_neighbors(self.h, C.byref(nnz)))
        self.nnz = nnz.value
        return self.nnz

    def set_neighbors(self, rowptr, col):
        rowptr = np.ascontiguousarray(rowptr, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        if col.size == 0:
            col = np.zeros(1, np.int32)
        self.nnz = int(rowptr[-1])
        self._check(lib().vpd_set_neighbors(self.h, rowptr, col, self.nnz))

    def get_neighbors(self):
        rowptr = np.zeros(self.n + 1, np.int32)
        col = np.zeros(max(self.nnz, 1), np.int32)
        self._check(lib().vpd_get_neighbors(self.h, rowptr, col))
        return rowptr, col[:self.nnz]

    def set_constraints(self, codes):
        codes = np.ascontiguousarray(codes, np.uint8)
        self._check(lib().vpd_set_constraints(self.h, codes.ctypes.data_as(C.c_void_p)))

    def initialize_shape(self):
        self._check(lib().vpd_initialize_shape(self.h))

    def initialize_material(self):
        self._check(lib().vpd_initialize_material(self.h))

    def initialize_dynamics(self):
        self._check(lib().vpd_initialize_dynamics(self.h))

    def _shape(self, name):
        if name == "damage":
            return (self.nnz,), np.float64
        if name in ("iterations", "fallbacks"):
            return (self.n,), np.int32
        if name == "strain_inc":
            return (self.n, 6), np.float64
        w = {"x": 3, "u": 3, "v": 3, "a": 3, "force": 3, "volume": 1, "temp": 1, "eint": 1,
             "einel": 1, "damage_index": 1, "statev": self.nstatev}.get(name, 9)
        return ((self.n,) if w == 1 else (self.n, w)), np.float64

    def upload(self, name, arr):
        shape, dt = self._shape(name)
        a = np.ascontiguousarray(arr, dtype=dt).reshape(shape)
        self._check(lib().vpd_upload(self.h, FIELD[name], a.ctypes.data_as(C.c_void_p)))

    def download(self, name):
        shape, dt = self._shape(name)
        out = np.zeros(shape, dt)
        self._check(lib().vpd_download(self.h, FIELD[name], out.ctypes.data_as(C.c_void_p)))
        return out

    def predict(self, dt, ramp):
        self._check(lib().vpd_predict(self.h, dt, _f64(ramp)))

    def apply_constraints(self, ramp):
        self._check(lib().vpd_apply_constraints(self.h, _f64(ramp)))

    def compute_deformation_gradient(self):
        self._check(lib().vpd_compute_deformation_gradient(self.h))

    def update_material(self, dt, step_time=0.0, total_time=0.0):
        self._check(lib().vpd_update_material(self.h, dt, step_time, total_time))

    def force_and_correct(self, dt, ramp):
        self._check(lib().vpd_compute_force_and_correct(self.h, dt, _f64(ramp)))

    def step(self, dt, time, step_time, ramp, ramp_next=None, want_fields=False):
        a = StepArgs()
        a.dt, a.time, a.step_time = dt, time, step_time
        a.ramp[:] = list(ramp)
        if ramp_next is not None:
            a.fuse_next = 1
            a.ramp_next[:] = list(ramp_next)
        a.want_fields = int(want_fields)
        self._check(lib().vpd_step(self.h, C.byref(a)))

    def synchronize(self):
        self._check(lib().vpd_synchronize(self.h))

    def set_specimens(self, first, dt, velocity, last_step):
        """Batched specimens: particles [first[s], first[s+1]) with their own dt, end
        velocity and step count (vpd_set_specimens). Empty arrays return to one specimen."""
        first = np.ascontiguousarray(first, np.int64)
        dt, velocity = _f64(dt), _f64(velocity)
        last_step = np.ascontiguousarray(last_step, np.int32)
        count = len(dt)
        self._specimen_tables = (first, dt, velocity, last_step)  # alive during the call
        ptr = lambda a: a.ctypes.data_as(C.c_void_p)
        self._check(lib().vpd_set_specimens(self.h, count, ptr(first), ptr(dt), ptr(velocity),
                                            ptr(last_step)))

    def step_specimens(self, step, fuse_next=True, want_fields=False):
        self._check(lib().vpd_step_specimens(self.h, int(step), int(fuse_next),
                                             int(want_fields)))

    def run_specimens(self, first_step, count):
        """`count` fused steps from `first_step` as replays of one captured HIP graph."""
        self._check(lib().vpd_run_specimens(self.h, int(first_step), int(count)))

    def set_return_map(self, mode, rtol=1e-12, warm_start=None):
        self._check(lib().vpd_set_return_map(self.h, mode, rtol))
        if warm_start is not None:
            self._warm = warm_start  # keep the network alive while attached
            self._check(lib().vpd_attach_warm_start(self.h, warm_start.h))

    TUNE_KEYS = ("defgrad_lds", "force_lds", "defgrad_rounds", "force_rounds", "defgrad_window",
                 "force_window", "defgrad_window_bonds", "force_window_bonds",
                 "defgrad_window_plan", "force_window_plan")

    def autotune(self, reps=3):
        out = (C.c_int * len(self.TUNE_KEYS))()
        self._check(lib().vpd_autotune(self.h, reps, out))
        return dict(zip(self.TUNE_KEYS, list(out)))

    def autotune_candidates(self, bond_pass, max_count=4):
        """The settings of bond pass `bond_pass` (0 F, 1 force) the last autotune timed again in
        its final stage, fastest first: [(tuning dict, mean trial ms)] (vpd_autotune_candidates)."""
        n = len(self.TUNE_KEYS)
        knobs = (C.c_int * (n * max_count))()
        ms = (C.c_float * max_count)()
        cnt = C.c_int()
        self._check(lib().vpd_autotune_candidates(self.h, bond_pass, max_count, knobs, ms,
                                                  C.byref(cnt)))
        return [(dict(zip(self.TUNE_KEYS, list(knobs[r * n:(r + 1) * n]))), ms[r])
                for r in range(cnt.value)]

    def set_tuning(self, **knobs):
        for k, v in knobs.items():
            self._check(lib().vpd_set_tuning(self.h, self.TUNE_KEYS.index(k), int(v)))

    def get_tuning(self):
        out = {}
        for i, k in enumerate(self.TUNE_KEYS):
            v = C.c_int()
            self._check(lib().vpd_get_tuning(self.h, i, C.byref(v)))
            out[k] = v.value
        return out

    def set_timing(self, on=True, stride=1):
        """Per-kernel HIP-event timing of every `stride`-th step (and every other launch)."""
        self._check(lib().vpd_set_timing(self.h, int(stride) if on else 0))

    def enable_diagnostics(self, load_axis=0, lateral_axis=1, observation_weight=None,
                           lateral_sets=None, eqps_index=1, indicator=(), indicator_min=False,
                           indicator_one_minus=False, capacity=1024):
        """Per-step run diagnostics (vpd_enable_diagnostics); None disables."""
        if load_axis is None:
            self._check(lib().vpd_enable_diagnostics(self.h, None))
            return
        cfg = DiagConfig()
        cfg.load_axis, cfg.lateral_axis = load_axis, lateral_axis
        keep = []
        if observation_weight is not None:
            w = _f64(observation_weight)
            keep.append(w)
            cfg.observation_weight = w.ctypes.data
        if lateral_sets is not None:
            ls = np.ascontiguousarray(lateral_sets, np.uint8)
            keep.append(ls)
            cfg.lateral_sets = ls.ctypes.data
        cfg.eqps_index = eqps_index
        cfg.indicator_count = len(indicator)
        for k, v in enumerate(indicator):
            cfg.indicator_indices[k] = v
        cfg.indicator_min, cfg.indicator_one_minus = int(indicator_min), int(indicator_one_minus)
        cfg.capacity = capacity
        self._diag_capacity = capacity
        self._check(lib().vpd_enable_diagnostics(self.h, C.byref(cfg)))

    def fetch_diagnostics(self):
        """The pending step records as an array (steps, DIAG_STRIDE)."""
        buf = np.zeros((self._diag_capacity, DIAG_STRIDE))
        cnt = C.c_int()
        self._check(lib().vpd_fetch_diagnostics(self.h, buf, self._diag_capacity, C.byref(cnt)))
        return buf[:cnt.value].copy()

    def collect_timing(self):
        buf = (KernelTime * 16)()
        cnt = C.c_int()
        self._check(lib().vpd_collect_timing(self.h, buf, 16, C.byref(cnt)))
        return {buf[i].name.decode(): (buf[i].total_ms, buf[i].launches) for i in range(cnt.value)}


class DeviceBuffer:
    """A device allocation made through libvpd's own HIP runtime."""

    def __init__(self, host, dtype=np.float64):
        self.host = np.ascontiguousarray(host, dtype=dtype)
        self.ptr = C.c_void_p()
        if lib().vpd_device_malloc(self.host.nbytes, C.byref(self.ptr)):
            raise VpdError(lib().vpd_last_error(None).decode())
        if lib().vpd_memcpy(self.ptr, self.host.ctypes.data_as(C.c_void_p), self.host.nbytes, 1):
            raise VpdError(lib().vpd_last_error(None).decode())

    def read(self):
        out = np.empty_like(self.host)
        if lib().vpd_memcpy(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes, 0):
            raise VpdError(lib().vpd_last_error(None).decode())
        return out

    def __del__(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            lib().vpd_device_free(self.ptr)
            self.ptr = C.c_void_p()


def vumat_device(props, dt, strain_inc, stress_old, state_old, temp_old, eint_old, einel_old,
                 nfieldv=1):
    """vpd_vumat_adapter_call_device on device copies of column-major numpy blocks:
    strain_inc/stress_old (6, nb), state_old (nstatev, nb), the rest (nb,). Returns the
    outputs read back (temp, stretch, defgrad, field, stress, state, eint, einel)."""
    nb = strain_inc.shape[1]
    nst = state_old.shape[0]
    nf = max(1, nfieldv)
    D = DeviceBuffer
    ins = dict(coord=D(np.zeros((3, nb))), charlen=D(np.ones(nb)), props=D(props),
               density=D(np.zeros(nb)), strain=D(strain_inc), spin=D(np.zeros((3, nb))),
               temp=D(temp_old), stretch=D(np.zeros((6, nb))), defgrad=D(np.zeros((9, nb))),
               field=D(np.zeros((nf, nb))), stress=D(stress_old), state=D(state_old),
               eint=D(eint_old), einel=D(einel_old))
    out = dict(temp=D(np.zeros(nb)), stretch=D(np.zeros((6, nb))), defgrad=D(np.zeros((9, nb))),
               field=D(np.zeros((nf, nb))), stress=D(np.zeros((6, nb))),
               state=D(np.zeros((nst, nb))), eint=D(np.zeros(nb)), einel=D(np.zeros(nb)))
    ints = [C.c_int(v) for v in (nb, 3, 3, nst, nfieldv, len(props), 0)]
    dbl = [C.c_double(v) for v in (0.0, 0.0, dt)]
    i = ins
    rc = lib().vpd_vumat_adapter_call_device(
        *[C.byref(v) for v in ints], *[C.byref(d) for d in dbl], b"JOHNSONCOOK".ljust(80),
        i["coord"].ptr, i["charlen"].ptr, i["props"].ptr, i["density"].ptr, i["strain"].ptr,
        i["spin"].ptr, i["temp"].ptr, i["stretch"].ptr, i["defgrad"].ptr, i["field"].ptr,
        i["stress"].ptr, i["state"].ptr, i["eint"].ptr, i["einel"].ptr, out["temp"].ptr,
        out["stretch"].ptr, out["defgrad"].ptr, out["field"].ptr, out["stress"].ptr,
        out["state"].ptr, out["eint"].ptr, out["einel"].ptr)
    if rc:
        raise VpdError(lib().vpd_last_error(None).decode())
    return {k: v.read() for k, v in out.items()}


class WarmStart:
    """The return-map warm-start network (vpd_mlp_*): FP32 training on the GPU, bf16 MFMA
    inference. theta has the 8961-value layout W1, b1, W2, b2, W3, b3, W4, b4."""

    def __init__(self, max_batch=4096):
        self.h = C.c_void_p()
        self.max_batch = max_batch
        if lib().vpd_mlp_create(max_batch, C.byref(self.h)):
            raise VpdError(lib().vpd_mlp_last_error(None).decode())

    def _check(self, rc):
        if rc:
            raise VpdError(lib().vpd_mlp_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().vpd_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        self.close()

    def set_params(self, theta=None, norm=None):
        t = None if theta is None else np.ascontiguousarray(theta, np.float32)
        nm = None if norm is None else np.ascontiguousarray(norm, np.float32)
        self._check(lib().vpd_mlp_set_params(
            self.h, None if t is None else t.ctypes.data_as(C.c_void_p),
            None if nm is None else nm.ctypes.data_as(C.c_void_p)))

    def get_params(self):
        theta = np.zeros(MLP_PARAMS, np.float32)
        norm = np.zeros(MLP_NORM, np.float32)
        self._check(lib().vpd_mlp_get_params(self.h, theta, norm))
        return theta, norm

    def loss_and_gradient(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = C.c_float()
        self._check(lib().vpd_mlp_loss_and_gradient(self.h, x, t, len(t), C.byref(loss), g))
        return loss.value, g

    def set_dataset(self, x, t, seed=0):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        self._check(lib().vpd_mlp_set_dataset(self.h, x, t, len(t), seed))

    def train(self, steps, batch=4096, lr=1e-3):
        loss = C.c_float()
        self._check(lib().vpd_mlp_train(self.h, steps, batch, lr, C.byref(loss)))
        return loss.value

    TRAIN_FP32, TRAIN_MFMA = 0, 1

    def set_training_form(self, form):
        """vpd_mlp_set_training_form: TRAIN_MFMA (bf16 matrix-core products, FP32
        accumulation; the default) or TRAIN_FP32."""
        self._check(lib().vpd_mlp_set_training_form(self.h, form))

    def predict(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros(len(x), np.float32)
        self._check(lib().vpd_mlp_predict(self.h, x, len(x), y))
        return y


def _return_map_on_device(mode, props, dt, nb, nst, ins, seed_ptr, rtol):
    D = DeviceBuffer
    outs = [D(np.zeros((6, nb))), D(np.zeros((nst, nb))), D(np.zeros(nb)), D(np.zeros(nb))]
    it = D(np.zeros(nb))  # 8 bytes per point: room for the int32 counts
    fb = D(np.zeros(nb))
    ms = C.c_double()
    rc = lib().vpd_return_map_device(mode, nb, nst, _f64(props), len(props), dt, rtol,
                                     *[b.ptr for b in ins], seed_ptr,
                                     *[b.ptr for b in outs], it.ptr, fb.ptr, C.byref(ms))
    if rc:
        raise VpdError(lib().vpd_last_error(None).decode())
    iters = it.read().view(np.int32)[:nb]
    fbs = fb.read().view(np.int32)[:nb]
    return dict(stress=outs[0].read(), state=outs[1].read(), eint=outs[2].read(),
                einel=outs[3].read(), iterations=iters, fallbacks=fbs, ms=ms.value)


def screen_accuracy():
    """vpd_screen_accuracy: (largest absolute error of the FP32 hardware log2 on [0.5, 1],
    largest relative error of the FP32 hardware exp2 on [0, 1]), measured on the device."""
    a, b = C.c_double(), C.c_double()
    if lib().vpd_screen_accuracy(C.byref(a), C.byref(b)):
        raise VpdError(lib().vpd_last_error(None).decode())
    return a.value, b.value


def return_map(mode, props, dt, strain_inc, stress_old, state_old, eint_old, einel_old,
               seed=None, rtol=1e-12):
    """vpd_return_map_device on device copies of column-major blocks (6, nb) / (nstatev, nb).
    Returns outputs, per-point iterations and fallbacks, and the kernel time in ms."""
    nb = strain_inc.shape[1]
    D = DeviceBuffer
    ins = [D(strain_inc), D(stress_old), D(state_old), D(eint_old), D(einel_old)]
    sd = D(seed) if seed is not None else None
    return _return_map_on_device(mode, props, dt, nb, state_old.shape[0], ins,
                                 sd.ptr if sd else None, rtol)


def return_map_learned(net, props, dt, strain_inc, stress_old, state_old, eint_old, einel_old,
                       rtol=1e-12, fused=True, want_seeds=True):
    """The learned return map on a device block. fused: vpd_return_map_learned_device, one
    timed kernel that seeds (state 5 where positive, the network elsewhere) and solves; else
    vpd_mlp_seed_device (timed as seed_ms) then vpd_return_map_device mode 3 (timed). Returns
    return_map's dict plus the seeds (and seed_ms for the two-call form). want_seeds False
    (fused only): no seed_out, so that only the network's seeds are written and read; the
    results are the same and no seeds are returned."""
    nb = strain_inc.shape[1]
    D = DeviceBuffer
    ins = [D(strain_inc), D(stress_old), D(state_old), D(eint_old), D(einel_old)]
    seed = D(np.zeros(nb))
    if fused:
        nst = state_old.shape[0]
        outs = [D(np.zeros((6, nb))), D(np.zeros((nst, nb))), D(np.zeros(nb)), D(np.zeros(nb))]
        it = D(np.zeros(nb))
        fb = D(np.zeros(nb))
        ms = C.c_double()
        if lib().vpd_return_map_learned_device(net.h, nb, nst, _f64(props), len(props), dt, rtol,
                                               *[b.ptr for b in ins],
                                               seed.ptr if want_seeds else None,
                                               *[b.ptr for b in outs], it.ptr, fb.ptr,
                                               C.byref(ms)):
            raise VpdError(lib().vpd_last_error(None).decode())
        out = dict(stress=outs[0].read(), state=outs[1].read(), eint=outs[2].read(),
                   einel=outs[3].read(), iterations=it.read().view(np.int32)[:nb],
                   fallbacks=fb.read().view(np.int32)[:nb], ms=ms.value)
        if want_seeds:
            out["seed"] = seed.read()
        return out
    seed_ms = C.c_double()
    if lib().vpd_mlp_seed_device(net.h, ins[0].ptr, nb, ins[2].ptr, nb, nb, seed.ptr,
                                 C.byref(seed_ms)):
        raise VpdError(lib().vpd_mlp_last_error(net.h).decode())
    out = _return_map_on_device(RETURN_MAP_LEARNED, props, dt, nb, state_old.shape[0], ins,
                                seed.ptr, rtol)
    out["seed_ms"] = seed_ms.value
    out["seed"] = seed.read()
    return out


def load_warm_start(path=None, max_batch=4096):
    """The trained warm-start network shipped with the package (warm_start_jc.npz)."""
    path = path or os.path.join(HERE, "warm_start_jc.npz")
    data = np.load(path)
    net = WarmStart(max_batch)
    net.set_params(data["theta"], data["norm"])
    return net


# ------------------------------------------------------------------ a bar case
class Bar:
    """The reference's regional-constraint bar (BoundaryValueCaseRunner.cpp:142-229) on the GPU:
    lattice, boundary bands, GPU neighbour CSR, K⁻¹, material and dynamics initialisation,
    then fused steps. `velocity` replaces the half-cosine ramp by a constant end velocity
    (u = v·t, v, a = 0) for the velocity-driven bars; with `velocity_rise` > 0 the band
    accelerates uniformly to v over that time first (u = v·t²/2t_r, v·t/t_r, v/t_r, then
    u = v·(t − t_r/2), v, 0)."""

    def __init__(self, nx, ny, nz, spacing, horizon, props, dt, density, amplitude=0.0,
                 rise_time=1.0, component=0, region_layers=3, plane_strain=False, nstatev=6,
                 tref=293.0, initial_state=None, damage=None, velocity=None,
                 material_density=7850.0, return_map=RETURN_MAP_BISECTION, rowptr=None,
                 col=None, rtol=1e-12, warm_start=None, points=None, velocity_rise=0.0):
        self.dt, self.amplitude, self.rise_time, self.velocity = dt, amplitude, rise_time, velocity
        self.velocity_rise = velocity_rise
        self.ctx = Context(horizon, density, props, nstatev, initial_state, tref, spacing,
                           material_density, damage)
        self.ctx.set_return_map(return_map, rtol, warm_start)
        # `points` = (x, volume) replaces the full lattice (e.g. voided_patch).
        self.x, self.vol = make_patch(nx, ny, nz, spacing) if points is None else points
        self.n = self.x.shape[0]
        self.ctx.set_particles(self.x, self.vol)
        if rowptr is None:
            self.ctx.build_neighbors()
        else:
            self.ctx.set_neighbors(rowptr, col)
        self.codes, self.anchor, self.pin = regional_constraint_codes(
            self.x, component, region_layers, spacing, plane_strain)
        self.ctx.set_constraints(self.codes)
        self.ctx.initialize_shape()
        self.ctx.initialize_material()
        self.ctx.initialize_dynamics()
        self.ctx.apply_constraints(self.ramp(0.0))
        self.step_index = 0

    def reset(self):
        """Back to the state before step 1 (zero displacement, fresh history, intact bonds)."""
        self.ctx.upload("u", np.zeros((self.n, 3)))
        self.ctx.initialize_material()
        self.ctx.initialize_dynamics()
        self.ctx.apply_constraints(self.ramp(0.0))
        self.step_index = 0

    def ramp(self, t):
        if self.velocity is not None:
            v, tr = self.velocity, self.velocity_rise
            if tr <= 0.0:
                return np.array([v * t, v, 0.0])
            if t < tr:
                return np.array([v * t * t / (2.0 * tr), v * t / tr, v / tr])
            return np.array([v * (t - 0.5 * tr), v, 0.0])
        return smooth_ramp(self.amplitude, self.rise_time, t)

    def step(self, k=1, fuse=True, want_fields=False):
        """k increments; the last one is left unfused so the state reads as the reference's
        end-of-step state."""
        for s in range(k):
            self.step_index += 1
            t = self.step_index * self.dt
            last = s == k - 1
            nxt = None if (last or not fuse) else self.ramp((self.step_index + 1) * self.dt)
            self.ctx.step(self.dt, t, (self.step_index - 1) * self.dt, self.ramp(t), nxt,
                          want_fields and last)

    def get(self, name):
        return self.ctx.download(name)

    STATE_FIELDS = ("u", "v", "a", "fN", "vN", "rN", "stressN", "statev", "temp", "eint",
                    "einel")

    def autotune_in_steps(self, reps=3, steps=40, warmup=10, rounds=2, top=3):
        """vpd_autotune, then its `top` fastest settings per bond pass timed in this bar's own
        steps: the force pass's with the F pass at the autotuner's pick, then the F pass's with
        the force pass so chosen. Each setting runs `warmup` + `steps` steps from the current
        state, `rounds` times in alternation with the others, and the fastest total is kept.
        The autotuner times each pass alone, and settings it ranks within a few percent can
        rank otherwise inside a step. The state is restored afterwards (bit for bit, every
        setting giving the same results) and the kept tuning is set and returned."""
        import time

        best = self.ctx.autotune(reps)
        snap = self.snapshot()
        for bond_pass in (1, 0):
            keys = [k for k in self.ctx.TUNE_KEYS
                    if k.startswith("force_" if bond_pass else "defgrad_")]
            settings = []
            for cand, _ in self.ctx.autotune_candidates(bond_pass, top):
                t = dict(best, **{k: cand[k] for k in keys})
                if t not in settings:
                    settings.append(t)
            if len(settings) < 2:
                continue
            total = [0.0] * len(settings)
            for _ in range(rounds):
                for i, t in enumerate(settings):
                    self.ctx.set_tuning(**t)
                    self.restore(snap)
                    self.step(warmup)
                    self.ctx.synchronize()
                    t0 = time.perf_counter()
                    self.step(steps)
                    self.ctx.synchronize()
                    total[i] += time.perf_counter() - t0
            best = settings[int(np.argmin(total))]
        self.ctx.set_tuning(**best)
        self.restore(snap)
        return best

    def snapshot(self):
        """The end-of-step state a following step reads (host copies), with the step index and
        the indices of the broken bonds in CSR order."""
        st = {k: self.get(k) for k in self.STATE_FIELDS}
        st["step"] = np.array(self.step_index)
        if self.ctx.damage:
            st["broken"] = np.flatnonzero(self.get("damage"))
        else:
            st["broken"] = np.zeros(0, np.int64)
        return st

    def restore(self, st):
        """Back to a snapshot: the next step is st['step'] + 1, bit for bit as if the run had
        gone on from there."""
        for k in self.STATE_FIELDS:
            self.ctx.upload(k, st[k])
        if self.ctx.damage:
            d = np.zeros(self.ctx.nnz)
            d[st["broken"]] = 1.0
            self.ctx.upload("damage", d)
        self.step_index = int(st["step"])


# ------------------------------------------------------------------ batched specimens
class Specimen:
    """One velocity-driven bar of a batch: lattice, horizon δ = horizon_factor·Δx, time step,
    end velocity (u = v·t on the loaded band) and step count."""

    def __init__(self, nx, ny, nz, spacing, horizon_factor, dt, velocity, steps,
                 component=0, region_layers=3):
        self.nx, self.ny, self.nz, self.spacing = nx, ny, nz, spacing
        self.horizon = horizon_factor * spacing
        self.dt, self.velocity, self.steps = dt, velocity, int(steps)
        self.component, self.region_layers = component, region_layers

    @property
    def n(self):
        return self.nx * self.ny * self.nz

    def bar_kwargs(self):
        return dict(nx=self.nx, ny=self.ny, nz=self.nz, spacing=self.spacing,
                    horizon=self.horizon, dt=self.dt, velocity=self.velocity,
                    component=self.component, region_layers=self.region_layers)


class SpecimenBatch:
    """Many specimens in one particle block, stepped together (vpd_set_specimens /
    vpd_step_specimens): the sensitivity sweep that benchmarks/tensile3d/run_cases.py:95-96
    runs as one process per case. Each specimen keeps its own coordinates, neighbour rows
    (built on the GPU with its own horizon; the block's CSR is their concatenation with
    column offsets), constraint codes, dt, velocity and step count; its results equal a
    `Bar` of the same specimen bit for bit. Material, density and damage are shared."""

    def __init__(self, specimens, props, density, nstatev=6, tref=293.0, damage=None,
                 material_density=7850.0, return_map=RETURN_MAP_BISECTION, rtol=1e-12,
                 warm_start=None, neighbor_cache=None):
        self.specimens = list(specimens)
        xs, vols, lens, cols, codes = [], [], [], [], []
        cache = {} if neighbor_cache is None else neighbor_cache
        offset = 0
        for sp in self.specimens:
            x, vol = make_patch(sp.nx, sp.ny, sp.nz, sp.spacing)
            key = (sp.nx, sp.ny, sp.nz, sp.spacing, sp.horizon)
            if key not in cache:  # the specimen's own neighbour rows, built on the GPU
                tmp = Context(sp.horizon, density, props, nstatev, None, tref, sp.spacing,
                              material_density, damage)
                tmp.set_particles(x, vol)
                tmp.build_neighbors()
                cache[key] = tmp.get_neighbors()
                tmp.close()
            rowptr, col = cache[key]
            xs.append(x)
            vols.append(vol)
            lens.append(np.diff(rowptr))
            cols.append((col, offset))
            codes.append(regional_constraint_codes(x, sp.component, sp.region_layers,
                                                   sp.spacing)[0])
            offset += sp.n
        if offset > np.iinfo(np.int32).max:
            raise VpdError("The batch holds more particles than int32 indices can address.")
        self.first = np.cumsum([0] + [sp.n for sp in self.specimens]).astype(np.int64)
        rowptr = np.zeros(offset + 1, np.int64)
        rowptr[1:] = np.cumsum(np.concatenate(lens))
        if rowptr[-1] > np.iinfo(np.int32).max:
            raise VpdError("The batch holds more bonds than the int32 CSR can index.")
        col = np.empty(max(int(rowptr[-1]), 1), np.int32)  # rows with column offsets
        pos = 0
        for c, off in cols:
            np.add(c, np.int32(off), out=col[pos:pos + len(c)])
            pos += len(c)
        self.x = np.concatenate(xs)
        self.vol = np.concatenate(vols)
        self.n = offset
        self.rowptr = rowptr.astype(np.int32)
        self.codes = np.concatenate(codes)
        self.ctx = Context(max(sp.horizon for sp in self.specimens), density, props, nstatev,
                           None, tref, self.specimens[0].spacing, material_density, damage)
        self.ctx.set_return_map(return_map, rtol, warm_start)
        self.ctx.set_particles(self.x, self.vol)
        self.ctx.set_neighbors(self.rowptr, col[:pos])
        del col, cols
        self.ctx.set_constraints(self.codes)
        self.ctx.initialize_shape()
        self.ctx.initialize_material()
        self.ctx.initialize_dynamics()
        self._table()
        self.step_index = 0

    def _table(self):
        sp = self.specimens
        self.ctx.set_specimens(self.first, [s.dt for s in sp], [s.velocity for s in sp],
                               [s.steps for s in sp])

    @property
    def max_steps(self):
        return max(s.steps for s in self.specimens)

    def reset(self):
        """Back to the state before step 1 for every specimen."""
        self.ctx.upload("u", np.zeros((self.n, 3)))
        self.ctx.initialize_material()
        self.ctx.initialize_dynamics()
        self._table()
        self.step_index = 0

    def step(self, k=None, fuse=True, want_fields=False):
        """k batch steps (default: until every specimen has finished). Specimens whose count
        is reached stop and keep their end-of-step state; the batch's last step is left
        unfused."""
        k = self.max_steps - self.step_index if k is None else k
        for s in range(k):
            self.step_index += 1
            last = s == k - 1
            self.ctx.step_specimens(self.step_index, fuse and not last, want_fields and last)

    def _part(self, name, out, s):
        a, b = self.first[s], self.first[s + 1]
        if name == "damage":
            return out[self.rowptr[a]:self.rowptr[b]]
        return out[a:b]

    def run(self, k=None):
        """k batch steps (default: until every specimen has finished) as replays of one
        captured HIP graph of a fused step. Specimens still running afterwards hold their
        prediction for the next step; `step` continues from here."""
        k = self.max_steps - self.step_index if k is None else k
        if k > 0:
            self.ctx.run_specimens(self.step_index + 1, k)
            self.step_index += k

    def get(self, name, specimen=None):
        """A field of the whole block, or of one specimen (bonds: its CSR range)."""
        out = self.ctx.download(name)
        return out if specimen is None else self._part(name, out, specimen)

    def split(self, name):
        """One download of a field, as a list of per-specimen views."""
        out = self.ctx.download(name)
        return [self._part(name, out, s) for s in range(len(self.specimens))]

