"""Per-entry error scales for the parity tests.

A result that the GPU computes with a different association of its additions
(projections: bucket order and lane-group trees; everything downstream of a
projection) can differ from the oracle by a few units in the last place of the
*terms that were summed*, not of the result.  Each comparison therefore divides
|gpu - oracle| entry by entry by the magnitude of those terms, obtained from the
oracle itself:

* single-signed values (eta, mu, rho, weights, k, Cp, second invariants): the
  value itself, i.e. pointwise relative error (``assert_close`` with no scale);
* projected sums: the same projection of |f| on the same cells (the per-node
  sum of |w f| over the sum of w);
* marker values interpolated from mixed-sign nodal fields (p, strain rates,
  stress increments): the same oracle function run on sign-aligned inputs, so
  that every term of its sums has one sign and the result is the sum of |terms|
  (each function is linear in the aligned inputs with non-negative weights);
* temperature after subgrid diffusion: |T| plus the interpolated |thisT| and
  |lastT| (the temperatures that it differences); Tdot the same over dt;
* stress tensors that a rotation mixes: |sxx| + |syy| + 2|sxy| of the stress
  before the step plus the aligned magnitudes of the step's increments, one
  value for all three components.

Everything here runs the oracle only (the checker); nothing reads the GPU.
"""
import ctypes as C

import numpy as np

import oracle

# marker FP64 fields that may carry either sign (positions stay as they are:
# their sign decides the cell)
SIGNED = ("T", "Tdot", "sxx", "syy", "sxy", "p", "exx", "eyy", "exy", "Exx", "Eyy", "Exy",
          "wxy", "rho", "rhodot", "eta", "mu", "Eii", "strainRateResidual")
STRESS = ("sxx", "syy", "sxy")
DS = ("dsxx", "dsyy", "dsxy")


def abs_markers(c):
    """|f| for every signed FP64 marker field of an oracle State (in place)."""
    for f in SIGNED:
        c.markers[f] = np.abs(c.markers[f])
    return c


def interp_basic(st, nodal):
    """Bilinear interpolation of a nodal array at the markers' basic cells (the
    oracle's projectNodalFieldToMarkersS), one value per marker."""
    c = st.clone()
    c.nodal["_scale_tmp"] = np.ascontiguousarray(nodal, dtype=np.float64)
    c.project_nodal_to_markers("_scale_tmp", "Tdot")
    return c.markers["Tdot"].copy()


def stress_l1(m):
    return np.abs(m["sxx"]) + np.abs(m["syy"]) + 2.0 * np.abs(m["sxy"])


def abs_projections(st, vep=False, projections=None):
    """Every projection of a step (or the given list) on |marker fields|, same cells."""
    c = abs_markers(st.clone())
    w = st.w
    if vep:
        proj = c.descriptor(rho="proj_rho", etaN="proj_etaN", etaS="proj_etaS",
                            kThermal="proj_kThermal", lastT="proj_lastT", Cp="proj_Cp",
                            soxx="proj_soxx", soyy="proj_soyy", soxy="proj_soxy",
                            rhodot="proj_rhodot", muS="proj_muS", muN="proj_muN")
        c._c("orc_project_markers_nodes_all2", C.byref(c.ms), C.byref(c.grid), C.byref(proj),
             C.byref(w.materials))
        names = [n for n in oracle.VEP_PROJ if n not in ("proj_rhodot", "proj_muS", "proj_muN")]
    elif (projections or w.projections) == "all2":
        c.project_markers_nodes_all2()
        names = ["rho", "etaN", "etaS", "kThermal", "lastT", "Cp", "soxx", "soyy", "soxy"]
    else:
        for src, sx, sy, meth, o in (projections or w.projections):
            c.project_marker_field(src, o, sx, sy, meth)
        names = [p[4] for p in (projections or w.projections)]
    return {n: np.abs(c.nodal[n]) for n in names}


def strain_pressure_scales(st, dt):
    """updateMarkerStrainPressure on aligned inputs: p, wxy and the stress terms as
    |.|, the nodal strain rates as -|.| (they enter e - mexx with a minus)."""
    c = st.clone()
    for n in ("p", "wxy", "soxx", "soyy", "soxy", "dsxx", "dsyy", "dsxy"):
        c.nodal[n][:] = np.abs(c.nodal[n])
    for n in ("edotxx", "edotyy", "edotxy"):
        c.nodal[n][:] = -np.abs(c.nodal[n])
    c.update_marker_strain_pressure(dt)
    return {f: np.abs(c.markers[f]) for f in ("p", "wxy", "exx", "eyy", "exy", "strainRateResidual")}


def subgrid_stress_scales(st, dt, dsubgrid):
    """subgridStressChanges on aligned inputs: marker s -> -|s|, nodal so -> |so|,
    incoming ds -> -|ds|.  Every increment (so_interp - s) * sdif is then >= 0 and
    the outgoing ds = ds_in - proj(increment) is minus the sum of |terms|.
    Returns (per-marker |increment| summed like stress_l1, nodal ds scales)."""
    c = st.clone()
    for f in STRESS:
        c.markers[f] = -np.abs(c.markers[f])
    for n in ("soxx", "soyy", "soxy"):
        c.nodal[n][:] = np.abs(c.nodal[n])
    for n in DS:
        c.nodal[n][:] = -np.abs(c.nodal[n])
    before = {f: c.markers[f].copy() for f in STRESS}
    c.subgrid_stress(dt, dsubgrid)
    inc = sum((2.0 if f == "sxy" else 1.0) * np.abs(c.markers[f] - before[f]) for f in STRESS)
    return inc, {n: np.abs(c.nodal[n]) for n in DS}


def update_stress_scales(st, ds_mag):
    """updateMarkerStress on aligned inputs (s -> 0, nodal ds -> its magnitude):
    the per-marker sum of |interpolation terms|, summed like stress_l1."""
    c = st.clone()
    for f in STRESS:
        c.markers[f] = 0.0
    for n in DS:
        c.nodal[n][:] = ds_mag[n]
    c.update_marker_stress()
    return stress_l1(c.markers)


def subgrid_t_scales(st, T0, lastT, dt):
    """Temperature magnitude that subgrid diffusion differences at each marker."""
    mag = np.abs(T0) + interp_basic(st, np.abs(st.nodal["thisT"])) + interp_basic(st, np.abs(lastT))
    return {"T": mag, "Tdot": mag / dt}


def eii_scale(m, dt, e_mag=None):
    """Eii += dt * sqrt(0.5 (exx^2 + eyy^2) + exy^2) (advection): old |Eii| plus
    the increment formed from the strain-rate magnitudes."""
    ex, ey, exy = e_mag if e_mag is not None else (np.abs(m["exx"]), np.abs(m["eyy"]),
                                                   np.abs(m["exy"]))
    return np.abs(m["Eii"]) + dt * np.sqrt(0.5 * (ex * ex + ey * ey) + exy * exy)


class StepScales:
    """Runs one oracle step (thermal or VEP) and collects the per-entry scale of
    every output: ``marker[f]`` over all markers, ``nodal[n]`` as (NY, NX)."""

    def __init__(self, st, vep):
        self.vep = vep
        self.marker, self.nodal = {}, {}
        self.dt = st.w.dt
        self._e_mag = None
        self.result = (st.step_vep if vep else st.step_thermal)(hook=self._hook)
        for f in ("eta", "mu", "rho", "rhodot", "p", "exx", "eyy", "exy", "strainRateResidual"):
            if f not in self.marker:  # single-signed, or unchanged by the step
                self.marker[f] = np.abs(st.markers[f])
        for f in STRESS:
            self.marker[f] = self._s_mag

    def _hook(self, stage, st):
        w = st.w
        if stage == "pre_projection":
            self.nodal.update(abs_projections(st, vep=self.vep))
            self._s_mag = stress_l1(st.markers)
            self._T0 = st.markers["T"].copy()
        elif stage == "projected" and self.vep:
            sp = strain_pressure_scales(st, self.dt)
            for f in ("p", "exx", "eyy", "exy", "strainRateResidual"):
                self.marker[f] = sp[f]
            self._e_mag = (sp["exx"], sp["eyy"], sp["exy"])
        elif stage == "strain_pressure" and self.vep:
            d = w.options.subgridStressDiffusivity
            if d > 0.0:
                inc, self._ds_mag = subgrid_stress_scales(st, self.dt, d)
                self._s_mag = self._s_mag + inc
            else:
                self._ds_mag = {n: np.abs(st.nodal[n]) for n in DS}
            self.nodal.update(self._ds_mag)
        elif stage == "subgrid_stress" and self.vep:
            self._s_mag = self._s_mag + update_stress_scales(st, self._ds_mag)
        elif stage == "subgrid_t":
            # markers have not moved since the projection: same cells
            name = st.thermal_inputs["lastT"]
            lastT = self.nodal.get(name, np.abs(st.nodal[name]))  # projection of |T|
            self.marker.update(subgrid_t_scales(st, self._T0, lastT, self.dt))
            self.marker["Eii"] = eii_scale(st.markers, self.dt, self._e_mag)
