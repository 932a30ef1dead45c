"""ctypes binding of libfvmic2d_hip.so and device-resident objects.

The library is the product: every call below goes through its C ABI
(include/fvm_abi.h).  There is no CPU fallback -- if the shared library is
missing or fails to load, ``load()`` raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_LIB = None
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libfvmic2d_hip.so")
COMPAT_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libfvmic2d_compat.so")

ERRORS = {0: "FVM_OK", 1: "FVM_ERR_PERIODIC", 2: "FVM_ERR_CELL_OUT_OF_BOUNDS", 3: "FVM_ERR_BAD_STAG",
          4: "FVM_ERR_BAD_METHOD", 5: "FVM_ERR_NAN_RESIDUAL", 6: "FVM_ERR_VISCOSITY_LAW",
          7: "FVM_ERR_NOMEM", 8: "FVM_ERR_HIP", 9: "FVM_ERR_ARG", 10: "FVM_ERR_TOO_MANY_FIELDS",
          11: "FVM_ERR_NOT_CONVERGED", 12: "FVM_ERR_BC", 13: "FVM_ERR_IO", 14: "FVM_ERR_CAPACITY",
          15: "FVM_ERR_NEGATIVE_ADD", 16: "FVM_ERR_STALLED"}


class FvmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code


class fvm_markers(C.Structure):
    _fields_ = ([("nMark", C.c_int32), ("maxMark", C.c_int32)] +
                [(n, C.c_void_p) for n in abi.DCOLS] + [(n, C.c_void_p) for n in abi.ICOLS])


class fvm_grid(C.Structure):
    _fields_ = [("NX", C.c_int32), ("NY", C.c_int32), ("LX", C.c_double), ("LY", C.c_double),
                ("xperiodic", C.c_int32), ("x", C.c_void_p), ("y", C.c_void_p), ("xc", C.c_void_p),
                ("yc", C.c_void_p)]


class fvm_projection(C.Structure):
    _fields_ = [("markerField", C.c_void_p), ("materialProperty", C.c_void_p),
                ("nodalField", C.c_void_p), ("stagx", C.c_int32), ("stagy", C.c_int32),
                ("method", C.c_int32)]


class fvm_instance(C.Structure):
    _fields_ = [("grid", fvm_grid), ("markers", fvm_markers), ("nodal", abi.NodalArrays),
                ("nproj", C.c_int32), ("proj", C.c_void_p), ("materials", abi.Materials),
                ("bv", abi.BoundaryValues), ("dt", C.c_double), ("dsubgrid", C.c_double),
                ("jaumann", C.c_int32)]


class fvm_tuning(C.Structure):
    _fields_ = [("block", C.c_int32), ("lanes", C.c_int32), ("persistent", C.c_int32)]


P = C.c_void_p
I32 = C.c_int32
D = C.c_double
_SIGS = {
    "fvm_ctx_create": [C.POINTER(P), P],
    "fvm_ctx_destroy": [P],
    "fvm_last_error": [P],
    "fvm_ctx_stream": [P],
    "fvm_ctx_sync": [P],
    "fvm_ctx_set_jaumann": [P, I32],
    "fvm_check_marker_density": [P, P, P, P, P, C.c_uint64, C.POINTER(I32), C.POINTER(I32)],
    "fvm_limit_displacement_timestep_noalloc": [P, P, P, P, P],
    "fvm_ctx_workspace_bytes": [P],
    "fvm_ctx_set_lanes": [P, I32],
    "fvm_device_count": [C.POINTER(C.c_int)],
    "fvm_malloc": [C.POINTER(P), C.c_size_t],
    "fvm_free": [P],
    "fvm_memcpy_htod": [P, P, P, C.c_size_t],
    "fvm_memcpy_dtoh": [P, P, P, C.c_size_t],
    "fvm_memcpy_dtod": [P, P, P, C.c_size_t],
    "fvm_memset": [P, P, C.c_int, C.c_size_t],
    "fvm_markers_alloc": [P, I32, P],
    "fvm_markers_free": [P],
    "fvm_markers_upload_aos": [P, P, P, I32],
    "fvm_markers_download_aos": [P, P, P],
    "fvm_find_marker_cells": [P, P, P, I32, C.POINTER(I32)],
    "fvm_markers_reorder": [P, P, P, P],
    "fvm_project_marker_fields": [P, P, P, I32, P],
    "fvm_project_markers_nodes_all2": [P, P, P, P, P],
    "fvm_project_markers_nodes_from_scalar": [P, P, P, P, P],
    "fvm_project_nodal_field_to_markers": [P, P, P, P, P],
    "fvm_advect_markers_rk": [P, P, P, P, P, D, I32],
    "fvm_subgrid_temperature_changes": [P, P, P, P, P, P, D, D],
    "fvm_subgrid_stress_changes": [P, P, P, P, P, D, D],
    "fvm_update_marker_stress": [P, P, P, P, P],
    "fvm_update_marker_strain_pressure": [P, P, P, P, P, P, D],
    "fvm_check_plastic_yielding": [P, P, P, P, D, C.POINTER(I32), P, D],
    "fvm_limit_viscosity": [P, P, P],
    "fvm_update_global_strain_rate_residual": [P, P, P, C.POINTER(D)],
    "fvm_update_marker_density": [P, P, P],
    "fvm_max_marker_T": [P, P, C.POINTER(D)],
    "fvm_max_abs_sxx": [P, P, C.POINTER(D)],
    "fvm_retrieve_solutions": [P, P, P, P, D, C.POINTER(I32)],
    "fvm_nodal_stress_strain": [P, P, P, P, D, P],
    "fvm_limit_displacement_timestep": [P, P, P, C.POINTER(D), P],
    "fvm_save_markers_binary": [P, P, C.c_char_p, D, C.POINTER(I32)],
    "fvm_form_thermal_system": [P, P, P, P, P, D, I32, P, P],
    "fvm_solve_thermal_system": [P, P, P, P, P, D, I32, C.POINTER(I32), C.POINTER(D)],
    "fvm_restart_from_markers": [P, P, C.c_char_p, C.POINTER(D)],
    "fvm_save_nodal_fields": [P, P, P, P, C.c_char_p],
    "fvm_host_alloc": [C.POINTER(P), C.c_size_t],
    "fvm_host_free": [P],
    "fvm_nodal_upload": [P, C.c_int64, I32, P, P],
    "fvm_nodal_download": [P, C.c_int64, I32, P, P],
    "fvm_step_thermal": [P, P, P, P, P, P, P, D, I32, P],
    "fvm_step_vep": [P, P, P, P, P, P, P, P, D, C.POINTER(I32), C.POINTER(D)],
    "fvm_profile_enable": [P, I32],
    "fvm_profile_read": [P, C.POINTER(C.c_double), C.POINTER(I32)],
    "fvm_timer_start": [P],
    "fvm_timer_stop": [P, C.POINTER(C.c_float)],
    "fvm_initial_pressure_guess": [P, P, P, D, D, D, I32, P, P, C.POINTER(I32), C.POINTER(D)],
    "fvm_pressure_guess_apply": [P, P, P, D, D, P, P, P],
    "fvm_ensemble_create": [P, I32, P, C.POINTER(P)],
    "fvm_ensemble_destroy": [P],
    "fvm_ensemble_step": [P, P],
    "fvm_ensemble_set_tuning": [P, P, P],
    "fvm_ensemble_get_tuning": [P, P],
    "fvm_ensemble_markers": [P],
    "fvm_ensemble_capture": [P, P],
    "fvm_ensemble_launch": [P, P],
    "fvm_ensemble_autotune": [P, P, I32, P, C.POINTER(C.c_float)],
    "fvm_ensemble_counters": [P, P, C.POINTER(I32), C.POINTER(I32)],
}
# libfvmic2d_compat.so: the reference-named wrappers (fvm_compat.h); they keep
# the reference's names and are bound by callers with their own prototypes
_COMPAT_SIGS = {
    "fvm_compat_finalize": [],
    "fvm_compat_last_status": [],
}
_RESTYPES = {"fvm_last_error": C.c_char_p, "fvm_ctx_stream": P, "fvm_ctx_workspace_bytes": C.c_size_t,
             "fvm_ensemble_markers": C.c_int64}


def load(path=None):
    """Load the HIP library (raises if it is missing: there is no fallback)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError("%s not built; run python -m fvmic2d_amd.build" % p)
    lib = C.CDLL(p)
    for name, args in _SIGS.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = _RESTYPES.get(name, C.c_int)
    _LIB = lib
    return lib


def exported_symbols():
    return list(_SIGS) + list(_COMPAT_SIGS)


_COMPAT = None


def load_compat(path=None):
    """The reference-named wrappers (libfvmic2d_compat.so, over libfvmic2d_hip.so)."""
    global _COMPAT
    if _COMPAT is not None:
        return _COMPAT
    load()
    p = path or COMPAT_LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError("%s not built; run python -m fvmic2d_amd.build" % p)
    lib = C.CDLL(p)
    for name, args in _COMPAT_SIGS.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int
    _COMPAT = lib
    return lib


class Context:
    """An fvm_ctx (HIP stream + workspace) on the current device."""

    def __init__(self, stream=None):
        self.lib = load()
        self.h = P()
        self._check(self.lib.fvm_ctx_create(C.byref(self.h), stream), None)

    def _check(self, rc, h="self"):
        if rc != 0:
            msg = self.lib.fvm_last_error(self.h if h == "self" else None)
            raise FvmError(rc, (msg or b"").decode())

    def call(self, name, *args):
        self._check(getattr(self.lib, name)(self.h, *args))

    def sync(self):
        self.call("fvm_ctx_sync")

    def close(self):
        if self.h:
            self.lib.fvm_ctx_destroy(self.h)
            self.h = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # device memory -------------------------------------------------------
    def alloc(self, nbytes):
        p = P()
        self._check(self.lib.fvm_malloc(C.byref(p), nbytes))
        return p.value

    def free(self, p):
        if p:
            self.lib.fvm_free(P(p))

    def pinned(self, shape):
        """A numpy array over pinned host memory (fvm_host_alloc), freed with the array's owner."""
        nbytes = int(np.prod(shape)) * 8
        p = P()
        self._check(self.lib.fvm_host_alloc(C.byref(p), nbytes))
        buf = (C.c_double * int(np.prod(shape))).from_address(p.value)
        arr = np.frombuffer(buf, dtype=np.float64).reshape(shape)
        self._pinned = getattr(self, "_pinned", []) + [p.value]
        return arr

    def htod(self, dst, arr):
        a = np.ascontiguousarray(arr)
        self.call("fvm_memcpy_htod", P(dst), a.ctypes.data_as(P), a.nbytes)

    def dtoh(self, src, shape, dtype=np.float64):
        out = np.empty(shape, dtype=dtype)
        self.call("fvm_memcpy_dtoh", out.ctypes.data_as(P), P(src), out.nbytes)
        return out

    def timer_start(self):
        self.call("fvm_timer_start")

    def timer_stop(self):
        ms = C.c_float()
        self.call("fvm_timer_stop", C.byref(ms))
        return ms.value


class DeviceArray:
    def __init__(self, ctx, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        self.ctx, self.shape, self.nbytes = ctx, a.shape, a.nbytes
        self.ptr = ctx.alloc(a.nbytes)
        ctx.htod(self.ptr, a)

    def get(self):
        return self.ctx.dtoh(self.ptr, self.shape)

    def set(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.nbytes == self.nbytes
        self.ctx.htod(self.ptr, a)

    def __del__(self):
        try:
            self.ctx.free(self.ptr)
        except Exception:
            pass


class DeviceGrid:
    """Device copies of the padded coordinate arrays (element 2 is index 0)."""

    def __init__(self, ctx, w):
        self.ctx = ctx
        self.NX, self.NY = w.NX, w.NY
        self._arrs = [DeviceArray(ctx, a) for a in (w.xp, w.yp, w.xcp, w.ycp)]
        self.desc = fvm_grid(w.NX, w.NY, w.LX, w.LY, int(getattr(w, "xperiodic", 0)),
                             *[a.ptr + 16 for a in self._arrs])

    @property
    def ref(self):
        return C.byref(self.desc)


class DeviceMarkers:
    """fvm_markers: SoA columns on the device."""

    def __init__(self, ctx, maxMark):
        self.ctx = ctx
        self.desc = fvm_markers()
        ctx.call("fvm_markers_alloc", I32(maxMark), C.byref(self.desc))

    @property
    def ref(self):
        return C.byref(self.desc)

    @property
    def n(self):
        return self.desc.nMark

    def upload(self, aos):
        a = np.ascontiguousarray(aos)
        assert a.dtype == abi.MARKER_DTYPE
        self.ctx.call("fvm_markers_upload_aos", self.ref, a.ctypes.data_as(P), I32(len(a)))

    def download(self):
        out = np.zeros(self.desc.nMark, dtype=abi.MARKER_DTYPE)
        self.ctx.call("fvm_markers_download_aos", self.ref, out.ctypes.data_as(P))
        return out

    def col(self, name):
        return getattr(self.desc, name)

    def get(self, name):
        dt = np.int32 if name in abi.ICOLS else np.float64
        return self.ctx.dtoh(self.col(name), (self.n,), dt)

    def __del__(self):
        try:
            self.ctx.lib.fvm_markers_free(C.byref(self.desc))
        except Exception:
            pass


class DeviceNodal:
    """Named NX*NY device arrays and a NodalArrays descriptor pointing at them."""

    def __init__(self, ctx, NX, NY, names=abi.NODAL_NAMES, init=None):
        self.ctx, self.NX, self.NY = ctx, NX, NY
        self.arrays = {}
        init = init or {}
        for n in names:
            a = init.get(n)
            self.arrays[n] = DeviceArray(ctx, a if a is not None else np.zeros((NY, NX)))
        self.desc = abi.NodalArrays()
        for n, a in self.arrays.items():
            if n in abi.NODAL_NAMES:
                setattr(self.desc, n, a.ptr)

    @property
    def ref(self):
        return C.byref(self.desc)

    def __getitem__(self, n):
        return self.arrays[n]

    def ptr(self, n):
        return self.arrays[n].ptr

    def get(self, n):
        return self.arrays[n].get()

    def view(self, **alias):
        """A NodalArrays descriptor whose named fields point at other arrays."""
        d = abi.NodalArrays()
        C.memmove(C.byref(d), C.byref(self.desc), C.sizeof(d))
        for k, v in alias.items():
            setattr(d, k, self.arrays[v].ptr)
        return d
