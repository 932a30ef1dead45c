"""ctypes binding of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by
bench.py's cpu_baseline leg, always as the checker or the timed CPU baseline,
never as part of the product path.
"""
import ctypes as C
import os

import numpy as np

from fvmic2d_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_LIB = None

P = C.c_void_p
D = C.c_double
I = C.c_int
SZ = C.c_size_t
_SIGS = {
    "orc_find_marker_cells": [P, P, I],
    "orc_project_marker_field": [P, P, P, SZ, P, I, I, I],
    "orc_project_markers_nodes_all2": [P, P, P, P],
    "orc_project_markers_nodes_from_scalar": [P, P, P, P],
    "orc_project_nodal_field_to_markers": [P, P, SZ, P],
    "orc_advect_markers_rk": [P, P, P, P, D, I],
    "orc_subgrid_temperature_changes": [P, P, P, P, P, D, D],
    "orc_subgrid_stress_changes": [P, P, P, P, D, D],
    "orc_update_marker_stress": [P, P, P],
    "orc_update_marker_strain_pressure": [P, P, P, D],
    "orc_compute_rheology": [P, P, P, C.POINTER(D), C.POINTER(D)],
    "orc_check_plastic_yielding": [P, P, P, D, C.POINTER(I), P],
    "orc_limit_viscosity": [P, P],
    "orc_update_global_strain_rate_residual": [P, P, C.POINTER(D)],
    "orc_update_marker_density": [P, P],
    "orc_max_marker_T": [P],
    "orc_max_abs_sxx": [P],
    "orc_nodal_stress_strain": [P, P, P, D, P],
    "orc_retrieve_solutions": [P, P, P, D, C.POINTER(I)],
    "orc_limit_displacement_timestep": [P, P, C.POINTER(D), D],
    "orc_save_markers_binary": [P, C.c_char_p, D],
    "orc_restart_from_markers": [P, C.c_char_p, C.POINTER(D)],
    "orc_check_marker_density": [P, P, P, P, C.c_uint64, C.POINTER(I), C.POINTER(I)],
    "orc_pc_u01": [C.c_uint64, C.c_uint64, C.c_uint64],
}
_RES = {"orc_max_marker_T": D, "orc_max_abs_sxx": D, "orc_limit_viscosity": None, "orc_pc_u01": D,
        "orc_update_marker_density": None}


def load():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError("%s not built (make -C oracle)" % LIB_PATH)
        lib = C.CDLL(LIB_PATH)
        for n, a in _SIGS.items():
            f = getattr(lib, n)
            f.argtypes = a
            f.restype = _RES.get(n, I)
        _LIB = lib
    return _LIB


class OracleError(RuntimeError):
    pass


class State:
    """Host copy of a workload: AoS markers, padded grid, nodal arrays."""

    def __init__(self, w, nodal_names=abi.NODAL_NAMES, extra=(), capacity=None):
        self.lib = load()
        self.w = w
        n = len(w.markers)
        cap = max(capacity or n, n)
        self.markers = np.zeros(cap, dtype=abi.MARKER_DTYPE)
        self.markers[:n] = w.markers
        self.markers["cellX"][n:] = -1  # unused slots: zeros, no cell
        self.markers["cellY"][n:] = -1
        self.xp, self.yp = w.xp.copy(), w.yp.copy()
        self.xcp, self.ycp = w.xcp.copy(), w.ycp.copy()
        s = 8
        self.grid = abi.GridData()
        self.grid.LX, self.grid.LY, self.grid.NX, self.grid.NY = w.LX, w.LY, w.NX, w.NY
        self.grid.xperiodic = int(getattr(w, "xperiodic", 0))
        self.grid.x = self.xp.ctypes.data + 2 * s
        self.grid.y = self.yp.ctypes.data + 2 * s
        self.grid.xc = self.xcp.ctypes.data + 2 * s
        self.grid.yc = self.ycp.ctypes.data + 2 * s
        self.ms = abi.MarkerSet(n, cap, 0, self.markers.ctypes.data)
        self.nodal = {}
        for n in list(nodal_names) + list(extra):
            a = w.nodal.get(n)
            self.nodal[n] = np.ascontiguousarray(a.copy() if a is not None else np.zeros((w.NY, w.NX)))
        self.nf = self.descriptor()

    def clone(self):
        """An independent copy (markers and nodal arrays copied, grid shared): the
        scale computations of the tests run oracle functions on modified copies."""
        c = State.__new__(State)
        c.lib, c.w = self.lib, self.w
        c.markers = self.markers.copy()
        c.xp, c.yp, c.xcp, c.ycp, c.grid = self.xp, self.yp, self.xcp, self.ycp, self.grid
        c.ms = abi.MarkerSet(self.ms.nMark, self.ms.maxMark, self.ms.nOut, c.markers.ctypes.data)
        c.nodal = {k: v.copy() for k, v in self.nodal.items()}
        c.nf = c.descriptor()
        return c

    def descriptor(self, **alias):
        d = abi.NodalArrays()
        for n in abi.NODAL_NAMES:
            if n in self.nodal:
                setattr(d, n, self.nodal[alias.get(n, n)].ctypes.data)
        for k, v in alias.items():
            setattr(d, k, self.nodal[v].ctypes.data)
        return d

    def _c(self, name, *args):
        rc = getattr(self.lib, name)(*args)
        if rc:
            raise OracleError("%s returned %d" % (name, rc))

    # reference-named calls ------------------------------------------------
    def find_marker_cells(self):
        self._c("orc_find_marker_cells", C.byref(self.ms), C.byref(self.grid), 0)
        return self.ms.nOut

    def project_marker_field(self, src, out, stagx, stagy, method):
        if src.startswith("mat:"):
            tab = (C.c_double * abi.MAXMAT)(*getattr(self.w.materials, "material" + src[4:]))
            self._c("orc_project_marker_field", C.byref(self.grid), C.byref(self.ms), tab, 0,
                    self.nodal[out].ctypes.data, stagx, stagy, method)
        else:
            self._c("orc_project_marker_field", C.byref(self.grid), C.byref(self.ms), None,
                    abi.FIELD_OFFSET[src], self.nodal[out].ctypes.data, stagx, stagy, method)

    def project_markers_nodes_all2(self):
        self._c("orc_project_markers_nodes_all2", C.byref(self.ms), C.byref(self.grid),
                C.byref(self.nf), C.byref(self.w.materials))

    def project_from_scalar(self, field, out):
        f = np.ascontiguousarray(field, dtype=np.float64)
        self._c("orc_project_markers_nodes_from_scalar", C.byref(self.ms), C.byref(self.grid),
                f.ctypes.data, self.nodal[out].ctypes.data)

    def project_nodal_to_markers(self, name, field):
        self._c("orc_project_nodal_field_to_markers", self.nodal[name].ctypes.data,
                C.byref(self.ms), abi.FIELD_OFFSET[field], C.byref(self.grid))

    def advect(self, dt, jaumann=0, nf=None):
        self._c("orc_advect_markers_rk", C.byref(self.ms), C.byref(nf or self.nf),
                C.byref(self.grid), C.byref(self.w.bv), dt, jaumann)

    def subgrid_temperature(self, dt, dsubgrid, thisT="thisT", nf=None):
        self._c("orc_subgrid_temperature_changes", self.nodal[thisT].ctypes.data,
                C.byref(self.grid), C.byref(nf or self.nf), C.byref(self.ms),
                C.byref(self.w.materials), dt, dsubgrid)

    def subgrid_stress(self, dt, dsubgrid):
        self._c("orc_subgrid_stress_changes", C.byref(self.grid), C.byref(self.nf),
                C.byref(self.ms), C.byref(self.w.materials), dt, dsubgrid)

    def update_marker_stress(self):
        self._c("orc_update_marker_stress", C.byref(self.grid), C.byref(self.nf), C.byref(self.ms))

    def update_marker_strain_pressure(self, dt):
        self._c("orc_update_marker_strain_pressure", C.byref(self.grid), C.byref(self.nf),
                C.byref(self.ms), dt)

    def check_plastic_yielding(self, dt):
        y = C.c_int(0)
        self._c("orc_check_plastic_yielding", C.byref(self.grid), C.byref(self.ms),
                C.byref(self.w.materials), dt, C.byref(y), C.byref(self.w.options))
        return y.value

    def nodal_stress_strain(self, dt, heating):
        self._c("orc_nodal_stress_strain", C.byref(self.grid), C.byref(self.nf), C.byref(self.w.bv), dt,
                heating.ctypes.data)

    def retrieve_solutions(self, S, Kcont):
        fn = I(0)
        self._c("orc_retrieve_solutions", C.byref(self.grid), C.byref(self.nf), S.ctypes.data, Kcont,
                C.byref(fn))
        return fn.value

    def limit_displacement_timestep(self, dt):
        d = D(dt)
        self._c("orc_limit_displacement_timestep", C.byref(self.grid), C.byref(self.nf), C.byref(d),
                self.w.options.displacementStepLimit)
        return d.value

    def save_markers_binary(self, path, elapsed_time):
        self._c("orc_save_markers_binary", C.byref(self.ms), str(path).encode(), elapsed_time)

    def restart_from_markers(self, path, capacity):
        """Replaces the markers by the file's (capacity slots); returns the elapsed time."""
        self.markers = np.zeros(capacity, dtype=abi.MARKER_DTYPE)
        self.ms = abi.MarkerSet(0, capacity, 0, self.markers.ctypes.data)
        t = D()
        self._c("orc_restart_from_markers", C.byref(self.ms), str(path).encode(), C.byref(t))
        self.markers = self.markers[:capacity]
        return t.value

    def check_marker_density(self, seed):
        """checkMarkerDensity (markers.c:455-781) with the counter stream orc_pc_u01;
        returns (added, removed).  The active markers are markers[:ms.nMark]."""
        a, r = I(0), I(0)
        self._c("orc_check_marker_density", C.byref(self.ms), C.byref(self.grid), C.byref(self.w.options),
                C.byref(self.w.materials), C.c_uint64(seed), C.byref(a), C.byref(r))
        return a.value, r.value

    def limit_viscosity(self):
        self.lib.orc_limit_viscosity(C.byref(self.ms), C.byref(self.w.options))

    def strain_rate_residual(self):
        r = C.c_double()
        self._c("orc_update_global_strain_rate_residual", C.byref(self.grid), C.byref(self.ms),
                C.byref(r))
        return r.value

    def update_marker_density(self):
        self.lib.orc_update_marker_density(C.byref(self.ms), C.byref(self.w.materials))

    def max_marker_T(self):
        return self.lib.orc_max_marker_T(C.byref(self.ms))

    def max_abs_sxx(self):
        return self.lib.orc_max_abs_sxx(C.byref(self.ms))

    # whole steps, reference call order (main_subduction.c) --------------
    # ``hook(stage, state)``, when given, is called between the sub-steps (stage
    # names: "pre_projection", "projected", "strain_pressure", "subgrid_stress",
    # "update_stress", "subgrid_t"); the tests use it to derive error scales from
    # the intermediate states.
    def step_thermal(self, hook=None):
        w = self.w
        hook = hook or (lambda stage, st: None)
        self.find_marker_cells()
        self.update_marker_density()
        hook("pre_projection", self)
        if w.projections == "all2":
            self.project_markers_nodes_all2()
            self.thermal_inputs = {"lastT": "lastT", "rho": "rho", "Cp": "Cp", "kThermal": "kThermal"}
        else:
            for src, sx, sy, meth, out in w.projections:
                self.project_marker_field(src, out, sx, sy, meth)
            self.thermal_inputs = {"lastT": "T_b", "rho": "rho_b", "Cp": "Cp_b", "kThermal": "k_b"}
        nf = self.descriptor(**self.thermal_inputs)
        hook("projected", self)
        self.subgrid_temperature(w.dt, w.options.subgridTemperatureDiffusivity, nf=nf)
        hook("subgrid_t", self)
        self.advect(w.dt)

    def step_vep(self, hook=None):
        """Marker part of one VEP step; host-solve outputs are the workload's nodal inputs
        (kept in nodal[...] under their names, projection outputs under proj_*)."""
        w = self.w
        hook = hook or (lambda stage, st: None)
        self.find_marker_cells()
        self.update_marker_density()
        y = self.check_plastic_yielding(w.dt)
        self.limit_viscosity()
        hook("pre_projection", self)
        proj = self.descriptor(rho="proj_rho", etaN="proj_etaN", etaS="proj_etaS",
                               kThermal="proj_kThermal", lastT="proj_lastT", Cp="proj_Cp",
                               soxx="proj_soxx", soyy="proj_soyy", soxy="proj_soxy",
                               rhodot="proj_rhodot", muS="proj_muS", muN="proj_muN")
        self._c("orc_project_markers_nodes_all2", C.byref(self.ms), C.byref(self.grid),
                C.byref(proj), C.byref(w.materials))
        hook("projected", self)
        self.update_marker_strain_pressure(w.dt)
        r = self.strain_rate_residual()
        hook("strain_pressure", self)
        if w.options.subgridStressDiffusivity > 0.0:
            self.subgrid_stress(w.dt, w.options.subgridStressDiffusivity)
        hook("subgrid_stress", self)
        self.update_marker_stress()
        hook("update_stress", self)
        self.thermal_inputs = {"lastT": "proj_lastT", "rho": "proj_rho", "Cp": "proj_Cp",
                               "kThermal": "proj_kThermal"}
        th = self.descriptor(**self.thermal_inputs)
        self.subgrid_temperature(w.dt, w.options.subgridTemperatureDiffusivity, nf=th)
        hook("subgrid_t", self)
        self.advect(w.dt)
        return y, r


VEP_PROJ = ["proj_rho", "proj_etaN", "proj_etaS", "proj_kThermal", "proj_lastT", "proj_Cp",
            "proj_soxx", "proj_soyy", "proj_soxy", "proj_rhodot", "proj_muS", "proj_muN"]
