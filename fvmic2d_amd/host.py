"""Host-side mirror of the reference's marker-path interface over the HIP library.

``MarkerPath`` holds one problem instance on the GPU (grid, SoA markers, nodal
fields) and exposes the reference's functions under their own names
(findMarkerCells -> find_marker_cells, ...).  Every method is one call into
libfvmic2d_hip.so; nothing here computes on the CPU.
"""
import ctypes as C

import numpy as np

from . import abi
from .lib import (Context, DeviceGrid, DeviceMarkers, DeviceNodal, fvm_instance, fvm_projection,
                  fvm_tuning, I32, P)

VEP_PROJ = ["proj_rho", "proj_etaN", "proj_etaS", "proj_kThermal", "proj_lastT", "proj_Cp",
            "proj_soxx", "proj_soyy", "proj_soxy", "proj_rhodot", "proj_muS", "proj_muN"]


def _mat_table(materials, name):
    return (C.c_double * abi.MAXMAT)(*getattr(materials, "material" + name))


class MarkerPath:
    # stress rotation used when this path is an ensemble member (0 exact, 1
    # -DJAUMANN); a single-instance step takes the context's setting
    # (fvm_ctx_set_jaumann)
    jaumann = 0

    def __init__(self, w, ctx=None, extra=(), capacity=None):
        self.w = w
        self.ctx = ctx or Context()
        self.grid = DeviceGrid(self.ctx, w)
        self.markers = DeviceMarkers(self.ctx, capacity or max(1, w.n_markers))
        self.markers.upload(w.markers)
        names = list(abi.NODAL_NAMES) + list(extra)
        if isinstance(w.projections, list):
            names += [p[4] for p in w.projections]
        if w.kind == "vep":
            names += VEP_PROJ
        seen = []
        for n in names:
            if n not in seen:
                seen.append(n)
        self.nodal = DeviceNodal(self.ctx, w.NX, w.NY, seen, init=w.nodal)
        self._tables = []
        self.proj_list = None
        if isinstance(w.projections, list):
            self.proj_list = self.projection_list(w.projections)

    # helpers --------------------------------------------------------------
    def projection_list(self, entries):
        arr = (fvm_projection * len(entries))()
        for i, (src, sx, sy, meth, out) in enumerate(entries):
            e = arr[i]
            if src.startswith("mat:"):
                t = _mat_table(self.w.materials, src[4:])
                self._tables.append(t)
                e.materialProperty = C.cast(t, P)
            else:
                e.markerField = self.markers.col(src)
            e.nodalField = self.nodal.ptr(out)
            e.stagx, e.stagy, e.method = sx, sy, meth
        return arr

    def get_markers(self):
        return self.markers.download()

    def get_nodal(self, name):
        return self.nodal.get(name)

    def set_nodal(self, name, arr):
        self.nodal[name].set(arr)

    def call(self, name, *args):
        self.ctx.call(name, *args)

    # reference-named calls -----------------------------------------------
    def find_marker_cells(self, rank=0):
        nout = I32()
        self.call("fvm_find_marker_cells", self.grid.ref, self.markers.ref, I32(rank), C.byref(nout))
        return nout.value

    def reorder_markers(self):
        """fvm_markers_reorder: move the markers into cell order (the cells must
        be current); returns the old index of the marker now in each slot."""
        n = self.markers.n
        if n == 0:
            return np.zeros(0, np.int32)
        d = self.ctx.alloc(4 * n)
        try:
            self.call("fvm_markers_reorder", self.grid.ref, self.markers.ref, C.c_void_p(d))
            return self.ctx.dtoh(d, (n,), np.int32)
        finally:
            self.ctx.free(d)

    def project_marker_fields(self, entries):
        arr = self.projection_list(entries)
        self.call("fvm_project_marker_fields", self.grid.ref, self.markers.ref, I32(len(entries)),
                  C.cast(arr, P))

    def project_marker_field(self, src, out, stagx, stagy, method):
        self.project_marker_fields([(src, stagx, stagy, method, out)])

    def project_markers_nodes_all2(self, nf=None):
        self.call("fvm_project_markers_nodes_all2", self.grid.ref, self.markers.ref,
                  C.byref(nf) if nf is not None else self.nodal.ref, C.byref(self.w.materials))

    def project_from_scalar(self, field_ptr, out):
        self.call("fvm_project_markers_nodes_from_scalar", self.grid.ref, self.markers.ref,
                  P(field_ptr), P(self.nodal.ptr(out)))

    def project_nodal_to_markers(self, name, field):
        self.call("fvm_project_nodal_field_to_markers", self.grid.ref, P(self.nodal.ptr(name)),
                  self.markers.ref, P(self.markers.col(field)))

    def advect(self, dt, jaumann=0, nf=None):
        self.call("fvm_advect_markers_rk", self.markers.ref,
                  C.byref(nf) if nf is not None else self.nodal.ref, self.grid.ref,
                  C.byref(self.w.bv), dt, I32(jaumann))

    def subgrid_temperature(self, dt, dsubgrid, thisT="thisT", nf=None):
        self.call("fvm_subgrid_temperature_changes", P(self.nodal.ptr(thisT)), self.grid.ref,
                  C.byref(nf) if nf is not None else self.nodal.ref, self.markers.ref,
                  C.byref(self.w.materials), dt, dsubgrid)

    def subgrid_stress(self, dt, dsubgrid):
        self.call("fvm_subgrid_stress_changes", self.grid.ref, self.nodal.ref, self.markers.ref,
                  C.byref(self.w.materials), dt, dsubgrid)

    def update_marker_stress(self):
        self.call("fvm_update_marker_stress", self.grid.ref, self.nodal.ref, self.markers.ref,
                  C.byref(self.w.materials))

    # nodal step between the Stokes solve and the marker updates ------------
    def retrieve_solutions(self, S, Kcont):
        """retrieveSolutions: S a device array of NX*NY*3 (vx, vy, p interleaved),
        its pressure entries scaled in place."""
        fn = I32(0)
        self.call("fvm_retrieve_solutions", self.grid.ref, self.nodal.ref, C.c_void_p(S.ptr), Kcont,
                  C.byref(fn))
        return fn.value

    def nodal_stress_strain(self, dt, heating):
        """nodalStressStrain: heating a device array of NX*NY (shear heating out)."""
        self.call("fvm_nodal_stress_strain", self.grid.ref, self.nodal.ref, C.byref(self.w.bv), dt,
                  C.c_void_p(heating.ptr))

    def limit_displacement_timestep(self, dt):
        d = C.c_double(dt)
        self.call("fvm_limit_displacement_timestep", self.grid.ref, self.nodal.ref, C.byref(d),
                  C.byref(self.w.options))
        return d.value

    # energy equation -------------------------------------------------------
    def form_thermal_system(self, dt, do_steady, coef, rhs):
        """formThermalSystem: coef a device array of 5*NX*NY (diagonal, east, west,
        south, north), rhs a device array of NX*NY (heating in, system RHS out)."""
        self.call("fvm_form_thermal_system", self.grid.ref, self.nodal.ref, C.byref(self.w.materials),
                  C.byref(self.w.options), dt, I32(1 if do_steady else 0), C.c_void_p(coef.ptr),
                  C.c_void_p(rhs.ptr))

    def solve_thermal_system(self, coef, rhs, T, rtol=1e-12, maxit=20000):
        its, rel = I32(0), C.c_double()
        self.call("fvm_solve_thermal_system", self.grid.ref, C.c_void_p(coef.ptr), C.c_void_p(rhs.ptr),
                  C.c_void_p(T.ptr), rtol, I32(maxit), C.byref(its), C.byref(rel))
        return its.value, rel.value

    # marker files ----------------------------------------------------------
    def save_markers_binary(self, path, elapsed_time):
        """saveMarkersBinary: returns the number of markers written."""
        n = I32(0)
        self.call("fvm_save_markers_binary", self.markers.ref, str(path).encode(), elapsed_time, C.byref(n))
        return n.value

    def save_nodal_fields(self, path, ha=None):
        """saveNodalFields in PETSc's binary Vec format; ha a device array or None."""
        self.call("fvm_save_nodal_fields", self.grid.ref, self.nodal.ref,
                  C.c_void_p(ha.ptr) if ha is not None else None, str(path).encode())

    def restart_from_markers(self, path):
        """restartFromMarkers: returns the elapsed time stored in the file."""
        t = C.c_double()
        self.call("fvm_restart_from_markers", self.markers.ref, str(path).encode(), C.byref(t))
        return t.value

    def update_marker_strain_pressure(self, dt):
        self.call("fvm_update_marker_strain_pressure", self.grid.ref, self.nodal.ref,
                  self.markers.ref, C.byref(self.w.materials), C.byref(self.w.options), dt)

    def check_plastic_yielding(self, dt):
        y = I32(0)
        self.call("fvm_check_plastic_yielding", self.grid.ref, self.markers.ref,
                  C.byref(self.w.materials), dt, C.byref(y), C.byref(self.w.options), 0.0)
        return y.value

    def check_marker_density(self, seed):
        """checkMarkerDensity (markers.c:455-781); returns (added, removed)."""
        a, r = I32(0), I32(0)
        self.call("fvm_check_marker_density", self.grid.ref, self.markers.ref, C.byref(self.w.materials),
                  C.byref(self.w.options), C.c_uint64(seed), C.byref(a), C.byref(r))
        return a.value, r.value

    def limit_viscosity(self):
        self.call("fvm_limit_viscosity", self.markers.ref, C.byref(self.w.options))

    def strain_rate_residual(self):
        r = C.c_double()
        self.call("fvm_update_global_strain_rate_residual", self.grid.ref, self.markers.ref,
                  C.byref(r))
        return r.value

    def update_marker_density(self):
        self.call("fvm_update_marker_density", self.markers.ref, C.byref(self.w.materials))

    def max_marker_T(self):
        r = C.c_double()
        self.call("fvm_max_marker_T", self.markers.ref, C.byref(r))
        return r.value

    def max_abs_sxx(self):
        r = C.c_double()
        self.call("fvm_max_abs_sxx", self.markers.ref, C.byref(r))
        return r.value

    # whole steps ------------------------------------------------------------
    def thermal_view(self):
        if self.proj_list is None:
            return self.nodal.desc
        return self.nodal.view(lastT="T_b", rho="rho_b", Cp="Cp_b", kThermal="k_b")

    def step_thermal(self):
        """One marker step (locate, density, projections, subgrid T, RK4), async."""
        w = self.w
        nf = self.thermal_view()
        if self.proj_list is None:
            self.call("fvm_step_thermal", self.grid.ref, self.markers.ref, C.byref(nf),
                      C.byref(w.materials), C.byref(w.options), C.byref(w.bv), w.dt, I32(0), None)
        else:
            self.call("fvm_step_thermal", self.grid.ref, self.markers.ref, C.byref(nf),
                      C.byref(w.materials), C.byref(w.options), C.byref(w.bv), w.dt,
                      I32(len(self.proj_list)), C.cast(self.proj_list, P))

    def vep_views(self):
        proj = self.nodal.view(rho="proj_rho", etaN="proj_etaN", etaS="proj_etaS",
                               kThermal="proj_kThermal", lastT="proj_lastT", Cp="proj_Cp",
                               soxx="proj_soxx", soyy="proj_soyy", soxy="proj_soxy",
                               rhodot="proj_rhodot", muS="proj_muS", muN="proj_muN")
        return proj, self.nodal.desc

    def step_vep(self):
        w = self.w
        proj, host = self.vep_views()
        y, r = I32(0), C.c_double()
        self.call("fvm_step_vep", self.grid.ref, self.markers.ref, C.byref(proj), C.byref(host),
                  C.byref(w.materials), C.byref(w.options), C.byref(w.bv), w.dt, C.byref(y),
                  C.byref(r))
        return y.value, r.value

    def step(self):
        return self.step_vep() if self.w.kind == "vep" else self.step_thermal()


class Ensemble:
    """Many independent thermal-marker problems stepped together.

    ``paths`` are MarkerPath objects on one Context whose workloads carry a
    projection list of the same pattern; their device arrays are shared with
    the ensemble (stepping the ensemble advances each path's state).
    """

    def __init__(self, paths):
        assert paths and all(p.ctx is paths[0].ctx for p in paths)
        self.paths, self.ctx = list(paths), paths[0].ctx
        arr = (fvm_instance * len(paths))()
        self._views = []
        for i, p in enumerate(paths):
            assert p.proj_list is not None, "ensemble members need a projection list"
            w, e = p.w, arr[i]
            e.grid = p.grid.desc
            e.markers = p.markers.desc
            v = p.thermal_view()
            self._views.append(v)
            e.nodal = v
            e.nproj = len(p.proj_list)
            e.proj = C.cast(p.proj_list, P)
            e.materials = w.materials
            e.bv = w.bv
            e.dt = w.dt
            e.dsubgrid = w.options.subgridTemperatureDiffusivity
            e.jaumann = p.jaumann
        self._arr = arr
        h = P()
        self.ctx.call("fvm_ensemble_create", I32(len(paths)), C.cast(arr, P), C.byref(h))
        self.h = h

    @property
    def n_markers(self):
        return int(self.ctx.lib.fvm_ensemble_markers(self.h))

    def step(self):
        self.ctx.call("fvm_ensemble_step", self.h)

    def capture(self):
        self.ctx.call("fvm_ensemble_capture", self.h)

    def launch(self):
        self.ctx.call("fvm_ensemble_launch", self.h)

    def set_tuning(self, block, lanes, persistent=0):
        t = fvm_tuning(block, lanes, persistent)
        self.ctx.call("fvm_ensemble_set_tuning", self.h, C.byref(t))

    def get_tuning(self):
        t = fvm_tuning()
        self.ctx._check(self.ctx.lib.fvm_ensemble_get_tuning(self.h, C.byref(t)))
        return t.block, t.lanes, t.persistent

    def autotune(self, reps=5):
        """((block, lanes, persistent), ms per step) of the fastest launch shape."""
        t, ms = fvm_tuning(), C.c_float()
        self.ctx.call("fvm_ensemble_autotune", self.h, I32(reps), C.byref(t), C.byref(ms))
        return (t.block, t.lanes, t.persistent), ms.value

    def counters(self):
        a, b = I32(), I32()
        self.ctx.call("fvm_ensemble_counters", self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def close(self):
        if self.h:
            self.ctx.lib.fvm_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def initial_pressure_guess(ctx, grid, rho, gy, gx=0.0, rtol=1e-12, maxit=20000, S=None, p_out=None):
    """initialPressureGuess (initialPressureGuess.c:6-184) on the GPU.

    grid: DeviceGrid; rho, S (3*NX*NY interleaved [vx, vy, p]) and p_out
    (NX*NY) are DeviceArray objects (or None for S / p_out).  Returns
    (iterations, relative residual)."""
    it, rel = I32(), C.c_double()
    ctx.call("fvm_initial_pressure_guess", grid.ref, P(rho.ptr), gy, gx, rtol, I32(maxit),
             P(S.ptr) if S is not None else None, P(p_out.ptr) if p_out is not None else None,
             C.byref(it), C.byref(rel))
    return it.value, rel.value


def pressure_guess_apply(ctx, grid, rho, gy, x, y, b=None, gx=0.0):
    """y = A x for the pressure-guess operator (and its right-hand side into b)."""
    ctx.call("fvm_pressure_guess_apply", grid.ref, P(rho.ptr), gy, gx, P(x.ptr), P(y.ptr),
             P(b.ptr) if b is not None else None)
