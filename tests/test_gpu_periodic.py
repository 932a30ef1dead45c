"""Periodic x grids (grid->xperiodic, DM_BOUNDARY_PERIODIC) on the GPU against
the oracle: cell location (markers.c:49,73), every projection family and method
(ghost columns add into the nodes they wrap to), the interpolations back to the
markers, RK4 advection across x = 0 and x = LX (markers.c:1909-1915, 1965,
2014-2023, 2225-2231), the stress and strain-rate updates
(updateMarkerStress.c:66,76,110, updateMarkerStrainPressure.c:137,146,211) and
whole thermal and VEP steps.  Same bars as the non-periodic parity tests: cells,
positions, velocities and spins bitwise, everything else 1e-12 relative entry
by entry against the magnitude of the summed terms."""
import ctypes as C

import numpy as np
import pytest

import oracle
from fvmic2d_amd import abi, workloads
from fvmic2d_amd.host import MarkerPath
from fvmic2d_amd.lib import FvmError
import scales as S
from helpers import RTOL, assert_bitwise, assert_close
from test_gpu_parity import FAMILIES, _abs_markers, _advect_scales, _cmp_markers

pytestmark = pytest.mark.gpu


def _pair(w, extra=()):
    print("CASE", w.name, w.NX, w.NY, flush=True)
    st = oracle.State(w, extra=list(extra) + (oracle.VEP_PROJ if w.kind == "vep" else []))
    mp = MarkerPath(w, extra=extra)
    return st, mp


@pytest.fixture(scope="module")
def plain():
    return workloads.periodic_box("plain", NX=64, NY=41, NM=4)


@pytest.fixture(scope="module")
def vep():
    return workloads.periodic_box("vep", NX=96, NY=49, NM=4)


def test_find_marker_cells_periodic(plain):
    st, mp = _pair(plain)
    assert mp.find_marker_cells() == st.find_marker_cells() == 2
    g = mp.get_markers()
    for f in ("cellX", "cellY", "cpu"):
        assert_bitwise(g[f], st.markers[f], f)
    # the last cell [x[NX-1], LX] exists and holds the markers at X = LX
    assert g["cellX"].max() == plain.NX - 1
    assert (g["cellX"][g["X"] == plain.LX] == plain.NX - 1).all()


@pytest.mark.parametrize("method", [abi.ARITHMETIC, abi.ARITHMETIC_LOCAL, abi.HARMONIC])
def test_projection_families_periodic(plain, method):
    st, mp = _pair(plain, extra=["out"])
    st.find_marker_cells()
    mp.find_marker_cells()
    for src in ("T", "X", "mat:kThermal"):
        for sx, sy in FAMILIES:
            print("CASE project %s %d %d %d" % (src, sx, sy, method), flush=True)
            st.project_marker_field(src, "out", sx, sy, method)
            mp.project_marker_field(src, "out", sx, sy, method)
            ref = st.nodal["out"].copy()
            got = mp.get_nodal("out")
            what = "%s (%d,%d) m%d" % (src, sx, sy, method)
            if src.startswith("mat:"):
                assert_close(got, ref, what)
                continue
            sa = oracle.State(_abs_markers(plain, src), extra=["out"])
            sa.find_marker_cells()
            sa.project_marker_field(src, "out", sx, sy, method)
            if method == abi.HARMONIC:
                with np.errstate(divide="ignore"):
                    assert_close(1.0 / got, 1.0 / ref, what, scale=np.abs(1.0 / sa.nodal["out"]))
            else:
                assert_close(got, ref, what, scale=np.abs(sa.nodal["out"]))


def test_project_all2_and_back_periodic(plain):
    st, mp = _pair(plain)
    for s in (st, mp):
        s.find_marker_cells()
    sc = S.abs_projections(st, projections="all2")
    st.project_markers_nodes_all2()
    mp.project_markers_nodes_all2()
    for n in ("rho", "etaN", "etaS", "kThermal", "lastT", "Cp", "soxx", "soyy", "soxy"):
        assert_close(mp.get_nodal(n), st.nodal[n], n, scale=sc[n])
    # projectNodalFieldToMarkersS reads the wrapped column for cell NX-1
    t_scale = S.interp_basic(st, np.abs(st.nodal["lastT"]))
    mp.set_nodal("lastT", st.nodal["lastT"])
    st.project_nodal_to_markers("lastT", "T")
    mp.project_nodal_to_markers("lastT", "T")
    _cmp_markers(st, mp, fields=["T"], scales={"T": t_scale})


@pytest.mark.parametrize("jaumann", [0, 1])
def test_advect_periodic(plain, jaumann):
    st, mp = _pair(plain)
    for s in (st, mp):
        s.find_marker_cells()
    X0 = st.markers["X"].copy()
    sc = _advect_scales(st, plain.dt * 1.7)
    st.advect(plain.dt * 1.7, jaumann)
    mp.advect(plain.dt * 1.7, jaumann)
    _cmp_markers(st, mp, exact=("X", "Y", "VX", "VY", "wxy"), scales=sc)
    X1 = mp.get_markers()["X"]
    act = st.markers["cellX"] != -1
    assert ((X0 > 0.9 * plain.LX) & (X1 < 0.1 * plain.LX) & act).sum() > 10  # wrapped at x = LX
    assert (X1[act] >= 0.0).all() and (X1[act] <= plain.LX).all()


def test_subgrid_temperature_periodic(plain):
    st, mp = _pair(plain)
    for s in (st, mp):
        s.find_marker_cells()
        s.update_marker_density()
        s.project_markers_nodes_all2()
    lastT = S.abs_projections(st, projections="all2")["lastT"]
    sc = S.subgrid_t_scales(st, st.markers["T"].copy(), lastT, plain.dt)
    st.subgrid_temperature(plain.dt, 1.0)
    mp.subgrid_temperature(plain.dt, 1.0)
    _cmp_markers(st, mp, fields=["T", "Tdot", "rho", "rhodot"], scales=sc)


def test_stress_updates_periodic(vep):
    st, mp = _pair(vep)
    for s in (st, mp):
        s.find_marker_cells()
    s_mag = S.stress_l1(st.markers)
    inc, ds_mag = S.subgrid_stress_scales(st, vep.dt, 1.0)
    s_mag = s_mag + inc
    st.subgrid_stress(vep.dt, 1.0)
    mp.subgrid_stress(vep.dt, 1.0)
    _cmp_markers(st, mp, fields=["sxx", "syy", "sxy"], what="subgrid_stress ",
                 scales={f: s_mag for f in S.STRESS})
    for n in ("dsxx", "dsyy", "dsxy"):
        assert_close(mp.get_nodal(n), st.nodal[n], n, scale=ds_mag[n])
    s_mag = s_mag + S.update_stress_scales(st, ds_mag)
    st.update_marker_stress()
    mp.update_marker_stress()
    _cmp_markers(st, mp, fields=["sxx", "syy", "sxy"], what="update_stress ",
                 scales={f: s_mag for f in S.STRESS})


def test_strain_pressure_periodic(vep):
    st, mp = _pair(vep)
    for s in (st, mp):
        s.find_marker_cells()
    sc = S.strain_pressure_scales(st, vep.dt)
    st.update_marker_strain_pressure(vep.dt)
    mp.update_marker_strain_pressure(vep.dt)
    _cmp_markers(st, mp, fields=["p", "wxy", "exx", "eyy", "exy", "strainRateResidual"], scales=sc)
    r_ref = st.strain_rate_residual()
    r_gpu = mp.strain_rate_residual()
    assert abs(r_gpu - r_ref) <= RTOL * abs(r_ref), (r_gpu, r_ref)


@pytest.mark.parametrize("lanes", [8, 4])
def test_step_thermal_periodic(plain, lanes):
    st, mp = _pair(plain)
    mp.ctx.call("fvm_ctx_set_lanes", C.c_int32(lanes))
    sc = S.StepScales(st, vep=False)
    mp.step_thermal()
    _cmp_markers(st, mp, exact=("X", "Y", "VX", "VY", "wxy"), scales=sc.marker)
    for n in ("rho", "etaN", "etaS", "kThermal", "lastT", "Cp", "soxx", "soyy", "soxy"):
        assert_close(mp.get_nodal(n), st.nodal[n], n, scale=sc.nodal[n])


def test_step_vep_periodic(vep):
    st, mp = _pair(vep)
    sc = S.StepScales(st, vep=True)
    y_r, r_r = sc.result
    y_g, r_g = mp.step_vep()
    assert y_r == y_g
    assert abs(r_g - r_r) <= RTOL * abs(r_r)
    _cmp_markers(st, mp, exact=("X", "Y", "VX", "VY", "wxy"), scales=sc.marker)
    for n in oracle.VEP_PROJ + ["dsxx", "dsyy", "dsxy"]:
        assert_close(mp.get_nodal(n), st.nodal[n], n, scale=sc.nodal.get(n))


@pytest.mark.timeout(300)
def test_fifty_steps_periodic():
    """50 thermal steps; the drift moves the markers about 13 of the 32 cells,
    so many cross x = LX: cells and positions stay bitwise the oracle's."""
    w = workloads.periodic_box("plain", NX=32, NY=21, NM=3, graded_y=False)
    st, mp = _pair(w)
    for _ in range(50):
        st.step_thermal()
        mp.step_thermal()
    g, r = mp.get_markers(), st.markers
    for f in ("cellX", "cellY"):
        assert_bitwise(g[f], r[f], f)
    act = r["cellX"] != -1
    for f in ("X", "Y", "VX", "VY"):
        assert_bitwise(g[f][act], r[f][act], f)
    assert_close(g["T"][act], r["T"][act], "T after 50 steps", rtol=1e-10)


def test_periodic_limits(plain):
    """A periodic grid needs NX >= 4; the nodal solves and population control
    refuse periodic grids (FVM_ERR_PERIODIC) before doing any work."""
    w = workloads.periodic_box("plain", NX=3, NY=5, NM=2)
    mp = MarkerPath(w)
    with pytest.raises(FvmError) as e:
        mp.find_marker_cells()
    assert e.value.code == 1
    mp = MarkerPath(plain, capacity=2 * plain.n_markers)
    mp.find_marker_cells()
    with pytest.raises(FvmError) as e:
        mp.check_marker_density(seed=1)
    assert e.value.code == 1
