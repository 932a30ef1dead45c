"""The reference-signature wrappers (include/fvm_compat.h) against the oracle:
a driver written for the reference calls findMarkerCells, projectMarkersNodesAll2,
subgridTemperatureChanges, advectMarkersRK, ... on its host AoS MarkerSet and
NodalFields, and the GPU path must leave them as the oracle does."""
import ctypes as C

import numpy as np
import pytest

import oracle
from fvmic2d_amd import abi, lib, workloads
import scales as S
from helpers import assert_bitwise, assert_close

pytestmark = pytest.mark.gpu

P = C.c_void_p


class HostVec(C.Structure):
    _fields_ = [("n", C.c_int32), ("a", C.c_void_p)]


NF_FIELDS = ["rho", "rhodot", "etaS", "etaN", "thisT", "lastT", "kThermal", "Cp", "boundaryT",
             "edotxx", "edotxy", "edotyy", "edotzz", "edotxz", "edotyz", "soxx", "soxy", "soyy",
             "dsxx", "dsyy", "dsxy", "muS", "muN", "vx", "vy", "p", "eii", "sii", "wxy", "ha",
             "nodalHeating", "VPTensorC", "VPTensorB", "strainRateResidual"]


class NodalFields(C.Structure):
    _fields_ = [(n, C.c_double if n == "boundaryT" else C.c_void_p) for n in NF_FIELDS]


class HostProblem:
    """The host-side structures a reference driver holds."""

    def __init__(self, w):
        self.w = w
        self.markers = np.ascontiguousarray(w.markers.copy())
        self.pads = [a.copy() for a in (w.xp, w.yp, w.xcp, w.ycp)]
        g = abi.GridData()
        g.LX, g.LY, g.NX, g.NY = w.LX, w.LY, w.NX, w.NY
        g.x, g.y, g.xc, g.yc = [a.ctypes.data + 16 for a in self.pads]
        self.grid = g
        self.ms = abi.MarkerSet(len(self.markers), len(self.markers), 0, self.markers.ctypes.data)
        self.arr, self.vecs = {}, {}
        self.nf = NodalFields()
        for n in NF_FIELDS:
            if n == "boundaryT":
                continue
            a = np.ascontiguousarray(w.nodal.get(n, np.zeros((w.NY, w.NX))).copy())
            v = HostVec(a.size, a.ctypes.data)
            self.arr[n], self.vecs[n] = a, v
            setattr(self.nf, n, C.addressof(v))

    def vec(self, n):
        return C.byref(self.vecs[n])

    def field_ptr(self, name):
        return P(self.markers.ctypes.data + abi.FIELD_OFFSET[name])


@pytest.fixture(scope="module")
def L():
    h = lib.load_compat()
    return h


def _rc(x):
    assert x == 0, x


def test_compat_thermal_step(L):
    w = workloads.convection_small()
    hp = HostProblem(w)
    st = oracle.State(w)
    # reference call order of one step (main_subduction.c:243-435), via the wrappers
    L.findMarkerCells(C.byref(hp.ms), C.byref(hp.grid))
    assert L.fvm_compat_last_status() == 0
    st.find_marker_cells()
    assert hp.ms.nOut == st.ms.nOut
    assert_bitwise(hp.markers["cellX"], st.markers["cellX"], "cellX")
    _rc(L.projectMarkersNodesAll2(C.byref(hp.ms), C.byref(hp.grid), C.byref(hp.nf),
                                  C.byref(w.materials), C.byref(w.options)))
    sc = S.abs_projections(st, projections="all2")
    st.project_markers_nodes_all2()
    for n in ("rho", "etaN", "etaS", "kThermal", "lastT", "Cp", "soxx", "soyy", "soxy", "muS"):
        assert_close(hp.arr[n], st.nodal[n], n, scale=sc.get(n))
    t_sc = S.subgrid_t_scales(st, st.markers["T"].copy(), sc["lastT"], w.dt)
    _rc(L.subgridTemperatureChanges(hp.vec("thisT"), C.byref(hp.grid), C.byref(hp.nf), C.byref(hp.ms),
                                    C.byref(w.materials), C.c_double(w.dt), C.byref(w.options)))
    st.subgrid_temperature(w.dt, w.options.subgridTemperatureDiffusivity)
    act = st.markers["cellX"] != -1
    assert_close(hp.markers["T"][act], st.markers["T"][act], "T", scale=t_sc["T"][act])
    _rc(L.advectMarkersRK(C.byref(hp.ms), C.byref(hp.nf), C.byref(hp.grid), C.byref(w.options),
                          C.byref(w.bv), C.c_double(w.dt)))
    st.advect(w.dt)
    for f in ("X", "Y", "VX", "VY", "wxy"):
        assert_bitwise(hp.markers[f], st.markers[f], f)
    # single-field projection named by a pointer into the AoS
    _rc(L.projectMarkerFieldToNodes(C.byref(hp.grid), C.byref(hp.ms), None, hp.field_ptr("T"),
                                    hp.vec("ha"), 1, 0, abi.ARITHMETIC))
    st.nodal["ha"] = np.zeros((w.NY, w.NX))
    st.project_marker_field("T", "ha", 1, 0, abi.ARITHMETIC)
    assert_close(hp.arr["ha"], st.nodal["ha"], "T->vy nodes")


def test_compat_vep_functions(L):
    w = workloads.vep_subduction(NX=61, NY=31, NM=3)
    hp = HostProblem(w)
    st = oracle.State(w)
    L.findMarkerCells(C.byref(hp.ms), C.byref(hp.grid))
    st.find_marker_cells()
    y = C.c_int32(0)
    _rc(L.checkPlasticYielding(C.byref(hp.grid), C.byref(hp.ms), C.byref(w.materials), C.c_double(w.dt),
                               C.byref(y), C.byref(w.options), C.c_double(0.0)))
    assert y.value == st.check_plastic_yielding(w.dt)
    L.limitViscosity(C.byref(hp.ms), C.byref(w.options), C.byref(w.materials))
    st.limit_viscosity()
    sc = S.strain_pressure_scales(st, w.dt)
    _rc(L.updateMarkerStrainPressure(C.byref(hp.grid), C.byref(hp.nf), C.byref(hp.ms),
                                     C.byref(w.materials), C.byref(w.options), C.c_double(w.dt)))
    st.update_marker_strain_pressure(w.dt)
    _rc(L.updateGlobalStrainRateResidual(C.byref(hp.grid), C.byref(hp.ms), 0))
    L.getGlobalStrainRateResidual.restype = C.c_double
    r = L.getGlobalStrainRateResidual(0)
    assert abs(r - st.strain_rate_residual()) <= 1e-12 * abs(r)
    s_mag = S.stress_l1(st.markers)
    inc, ds_mag = S.subgrid_stress_scales(st, w.dt, 1.0)
    _rc(L.subgridStressChanges(C.byref(hp.grid), C.byref(hp.nf), C.byref(hp.ms), C.byref(w.materials),
                               C.c_double(w.dt), C.c_double(1.0)))
    st.subgrid_stress(w.dt, 1.0)
    s_mag = s_mag + inc + S.update_stress_scales(st, ds_mag)
    _rc(L.updateMarkerStress(C.byref(hp.grid), C.byref(hp.nf), C.byref(hp.ms), C.byref(w.materials)))
    st.update_marker_stress()
    act = st.markers["cellX"] != -1
    sc.update({f: s_mag for f in S.STRESS})
    for f in ("eta", "mu", "sxx", "syy", "sxy", "exx", "eyy", "exy", "p", "wxy", "strainRateResidual"):
        sf = sc[f][act] if f in sc else None  # eta, mu: pointwise
        assert_close(hp.markers[f][act], st.markers[f][act], f, scale=sf)
    for n in ("dsxx", "dsyy", "dsxy"):
        assert_close(hp.arr[n], st.nodal[n], n, scale=ds_mag[n])
    L.fvm_compat_finalize()


def test_compat_nodal_step(L):
    """retrieveSolutions, nodalStressStrain, limitDisplacementTimestep through the
    reference's prototypes on host Vecs."""
    w = workloads.nodal_box(61, 31)
    hp = HostProblem(w)
    st = oracle.State(w)
    nn = w.NX * w.NY
    rng = np.random.default_rng(5)
    S = rng.normal(size=3 * nn)
    S_ref = S.copy()
    Sv = HostVec(3 * nn, S.ctypes.data)
    fn = C.c_int32(7)
    _rc(L.retrieveSolutions(C.byref(hp.grid), C.byref(hp.nf), C.byref(Sv), C.c_double(3.5), C.byref(fn)))
    st.retrieve_solutions(S_ref, 3.5)
    assert fn.value == 0 and (S == S_ref).all()
    for n in ("vx", "vy", "p"):
        assert (hp.arr[n] == st.nodal[n]).all(), n
    heat = np.zeros((w.NY, w.NX))
    hv = HostVec(nn, heat.ctypes.data)
    _rc(L.nodalStressStrain(C.byref(hp.grid), C.byref(hp.nf), C.byref(w.options), C.byref(w.bv),
                            C.c_double(w.dt), C.byref(hv), C.c_double(10.0)))
    h_ref = np.zeros((w.NY, w.NX))
    st.nodal_stress_strain(w.dt, h_ref)
    for n in ("edotxx", "edotyy", "edotxy", "wxy", "soxx", "soyy", "soxy", "dsxx", "dsyy", "dsxy", "eii", "sii"):
        assert_close(hp.arr[n], st.nodal[n], n)  # same expression order: pointwise
    assert_close(heat, h_ref, "nodalHeating")
    d = C.c_double(w.dt)
    _rc(L.limitDisplacementTimestep(C.byref(hp.grid), C.byref(hp.nf), C.byref(d), C.byref(w.options)))
    assert d.value == st.limit_displacement_timestep(w.dt)
    L.fvm_compat_finalize()


def test_compat_marker_files(L, tmp_path, monkeypatch):
    """saveMarkersBinary / restartFromMarkers write and read output/Markers.<iMonte>.<iTime>."""
    from test_oracle_markerfile import _workload
    monkeypatch.chdir(tmp_path)
    (tmp_path / "output").mkdir()
    w = _workload()
    hp = HostProblem(w)
    _rc(L.saveMarkersBinary(C.byref(hp.ms), 2, 17, C.c_double(9.5)))
    st = oracle.State(w)
    st.save_markers_binary(tmp_path / "ref", 9.5)
    assert open(tmp_path / "output" / "Markers.2.17", "rb").read() == open(tmp_path / "ref", "rb").read()
    ms = abi.MarkerSet()
    t = C.c_double()
    opt = abi.default_options()
    opt.maxMarkFactor = 1.5
    _rc(L.restartFromMarkers(C.byref(ms), C.byref(hp.grid), C.byref(w.materials), C.byref(opt), 2, 17,
                             C.byref(t)))
    assert t.value == 9.5
    n_act = int((w.markers["cellX"] != -1).sum())
    assert ms.nMark == n_act and ms.maxMark == int(n_act * 1.5)
    got = np.ctypeslib.as_array(C.cast(ms.markers, C.POINTER(C.c_uint8)), shape=(ms.maxMark * 216,))
    got = got.view(abi.MARKER_DTYPE)
    t_ref = st.restart_from_markers(tmp_path / "ref", ms.maxMark)
    assert t_ref == 9.5
    for f in abi.MARKER_DTYPE.names:
        assert_bitwise(got[f], st.markers[f], f)
    libc = C.CDLL(None)
    libc.free.argtypes = [C.c_void_p]
    libc.free(ms.markers)
    L.fvm_compat_finalize()


def test_compat_save_nodal_fields(L, tmp_path, monkeypatch):
    from test_gpu_markerfile import NODAL_FILE_ORDER, read_petsc_vecs
    monkeypatch.chdir(tmp_path)
    (tmp_path / "output").mkdir()
    w = workloads.nodal_box(19, 11)
    rng = np.random.default_rng(4)
    for n in NODAL_FILE_ORDER:
        w.nodal[n] = rng.normal(size=(w.NY, w.NX))
    hp = HostProblem(w)
    _rc(L.saveNodalFields(C.byref(hp.nf), C.byref(hp.grid), 1, 8, C.c_double(1.0), C.c_double(2.0), 0))
    vecs = read_petsc_vecs(tmp_path / "output" / "loadNodalFields_1_8.petscbin")
    for name, v in zip(NODAL_FILE_ORDER, vecs):
        assert_bitwise(v, hp.arr[name].ravel(), name)
    assert L.saveNodalFields(C.byref(hp.nf), C.byref(hp.grid), 1, 8, C.c_double(1.0), C.c_double(2.0), 1) != 0
    L.fvm_compat_finalize()


def test_compat_initial_pressure_guess(L):
    """initialPressureGuess through the reference's prototype: the pressure entries
    of S match the direct solve of the oracle's system, the velocity entries stay."""
    import scipy.sparse.linalg as sla
    from oracle import stokes as OS
    w = workloads.nodal_box(41, 29)
    rng = np.random.default_rng(6)
    w.nodal["rho"] = 3300.0 + 50.0 * rng.normal(size=(w.NY, w.NX))
    hp = HostProblem(w)
    n = w.NX * w.NY
    S = np.full(3 * n, 7.0)
    Sv = HostVec(3 * n, S.ctypes.data)
    opt = abi.default_options()
    opt.gy = 9.81
    _rc(L.initialPressureGuess(C.byref(hp.grid), C.byref(hp.nf), C.byref(opt), C.byref(Sv)))
    A, b = OS.pressure_guess_system(w, w.nodal["rho"], 9.81)
    ref = sla.spsolve(A.tocsc(), b)
    p = S.reshape(-1, 3)[:, 2]
    assert np.abs(p - ref).max() <= 1e-8 * np.abs(ref).max()
    assert (S.reshape(-1, 3)[:, :2] == 7.0).all()
    L.fvm_compat_finalize()
