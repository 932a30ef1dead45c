"""The nodal step between the Stokes solve and the marker updates on the GPU
(fvm_retrieve_solutions, fvm_nodal_stress_strain, fvm_limit_displacement_timestep)
against the oracle: strain rates, stresses, heating and invariants to 1e-12
relative entry by entry (the kernels evaluate each node's expressions in the
reference's order, so no sum is reassociated), the retrieved velocities,
pressure and the time step bitwise."""
import numpy as np
import pytest

import oracle
from fvmic2d_amd import workloads
from fvmic2d_amd.host import MarkerPath
from fvmic2d_amd.lib import DeviceArray, FvmError
from helpers import assert_bitwise, assert_close

pytestmark = pytest.mark.gpu

OUT = ["edotxx", "edotyy", "edotxy", "wxy", "soxx", "soyy", "soxy", "dsxx", "dsyy", "dsxy", "eii", "sii"]
CASES = [((37, 23), True, 1), ((64, 33), False, 1), ((801, 401), True, 0)]


def _case(shape, graded, bc):
    print("  case: nodal_box %dx%d graded=%s top/left type %d" % (shape[0], shape[1], graded, bc), flush=True)
    w = workloads.nodal_box(*shape, graded=graded)
    w.bv.mechBCTop.type[0] = bc
    w.bv.mechBCLeft.type[1] = bc
    w.bv.mechBCTop.value[0] = 2e-10 if bc == 0 else 0.0
    w.bv.mechBCLeft.value[1] = -1e-10 if bc == 0 else 0.0
    return w


@pytest.mark.parametrize("shape,graded,bc", CASES)
def test_nodal_stress_strain_matches_oracle(shape, graded, bc):
    w = _case(shape, graded, bc)
    st = oracle.State(w)
    mp = MarkerPath(w)
    h_ref = np.zeros((w.NY, w.NX))
    st.nodal_stress_strain(w.dt, h_ref)
    h = DeviceArray(mp.ctx, np.full((w.NY, w.NX), 5.0))
    mp.nodal_stress_strain(w.dt, h)
    for n in OUT:
        assert_close(mp.get_nodal(n), st.nodal[n], n, scale=None)
    assert_close(h.get().reshape(w.NY, w.NX), h_ref, "nodalHeating")


def test_nodal_stress_strain_rejects_undefined_ghost():
    w = workloads.nodal_box(9, 7)
    w.bv.mechBCLeft.type[1] = 2
    mp = MarkerPath(w)
    with pytest.raises(FvmError):
        mp.nodal_stress_strain(w.dt, DeviceArray(mp.ctx, np.zeros((w.NY, w.NX))))


@pytest.mark.parametrize("shape,graded,bc", CASES)
def test_retrieve_solutions_bitwise(shape, graded, bc):
    w = _case(shape, graded, bc)
    st = oracle.State(w)
    mp = MarkerPath(w)
    rng = np.random.default_rng(shape[0])
    S = rng.normal(size=(w.NY, w.NX, 3)).ravel()
    K = 1e21 / 0.0731
    S_ref = S.copy()
    st.retrieve_solutions(S_ref, K)
    dS = DeviceArray(mp.ctx, S)
    assert mp.retrieve_solutions(dS, K) == 0
    assert_bitwise(dS.get().ravel(), S_ref, "S (pressure scaled in place)")
    for n in ("vx", "vy", "p"):
        assert_bitwise(mp.get_nodal(n), st.nodal[n], n)


@pytest.mark.parametrize("shape,graded,bc", CASES)
def test_limit_displacement_timestep_bitwise(shape, graded, bc):
    w = _case(shape, graded, bc)
    st = oracle.State(w)
    mp = MarkerPath(w)
    for d0 in (1e30, w.dt, 1.0):
        assert mp.limit_displacement_timestep(d0) == st.limit_displacement_timestep(d0)
    # NaN velocities offer no candidate; a NaN first rhodot disables the cap
    w.nodal["vx"][2, 3] = np.nan
    w.nodal["rhodot"][0, 0] = np.nan
    st, mp = oracle.State(w), MarkerPath(w)
    assert mp.limit_displacement_timestep(1e30) == st.limit_displacement_timestep(1e30)


def test_nodal_step_in_sequence_at_full_size():
    """retrieveSolutions then nodalStressStrain then limitDisplacementTimestep on
    the 801 x 401 graded grid, in the order main_subduction.c:320-349 calls them."""
    w = _case((801, 401), True, 1)
    st = oracle.State(w)
    mp = MarkerPath(w)
    rng = np.random.default_rng(11)
    S = np.stack([w.nodal["vx"], w.nodal["vy"], rng.normal(size=(w.NY, w.NX))], axis=-1).ravel()
    S_ref = S.copy()
    st.retrieve_solutions(S_ref, 2.5)
    h_ref = np.zeros((w.NY, w.NX))
    st.nodal_stress_strain(w.dt, h_ref)
    d_ref = st.limit_displacement_timestep(w.dt)
    dS = DeviceArray(mp.ctx, S)
    mp.retrieve_solutions(dS, 2.5)
    h = DeviceArray(mp.ctx, np.zeros((w.NY, w.NX)))
    mp.nodal_stress_strain(w.dt, h)
    assert mp.limit_displacement_timestep(w.dt) == d_ref
    for n in OUT + ["p"]:
        assert_close(mp.get_nodal(n), st.nodal[n], n, scale=None)


@pytest.mark.parametrize("shape", [(2, 2), (3, 2), (2, 5), (3, 3)])
def test_nodal_step_minimal_grids(shape):
    """Grids with no interior node (heating, eii, sii all zero) or a single one."""
    w = _case(shape, False, 1)
    st, mp = oracle.State(w), MarkerPath(w)
    h_ref = np.zeros((w.NY, w.NX))
    st.nodal_stress_strain(w.dt, h_ref)
    h = DeviceArray(mp.ctx, np.full((w.NY, w.NX), 3.0))
    mp.nodal_stress_strain(w.dt, h)
    for n in OUT:
        assert_close(mp.get_nodal(n), st.nodal[n], n, scale=None)
    assert_close(h.get().reshape(w.NY, w.NX), h_ref, "nodalHeating")
    S = np.random.default_rng(1).normal(size=3 * w.NX * w.NY)
    S_ref = S.copy()
    st.retrieve_solutions(S_ref, 2.0)
    mp.retrieve_solutions(DeviceArray(mp.ctx, S), 2.0)
    assert_bitwise(mp.get_nodal("p"), st.nodal["p"], "p")
    assert mp.limit_displacement_timestep(1e30) == st.limit_displacement_timestep(1e30)


def test_packed_nodal_exchange():
    """fvm_nodal_upload / fvm_nodal_download: packed pinned host block <-> separate
    device arrays, bitwise, and the 32-array limit."""
    import ctypes as C
    w = workloads.nodal_box(37, 23)
    mp = MarkerPath(w)
    names = ["vx", "vy", "etaN", "muN", "soxx"]
    n = w.NX * w.NY
    h = mp.ctx.pinned((len(names), n))
    h[:] = np.random.default_rng(9).normal(size=h.shape)
    d = (C.c_void_p * len(names))(*[mp.nodal.ptr(x) for x in names])
    mp.ctx.call("fvm_nodal_upload", C.c_int64(n), len(names), C.cast(d, C.c_void_p), h.ctypes.data_as(C.c_void_p))
    for k, x in enumerate(names):
        assert_bitwise(mp.get_nodal(x).ravel(), h[k], x)
    out = mp.ctx.pinned((len(names), n))
    mp.ctx.call("fvm_nodal_download", C.c_int64(n), len(names), C.cast(d, C.c_void_p),
                out.ctypes.data_as(C.c_void_p))
    assert_bitwise(out.ravel(), h.ravel(), "round trip")
    with pytest.raises(FvmError):
        mp.ctx.call("fvm_nodal_upload", C.c_int64(n), 33, C.cast(d, C.c_void_p), h.ctypes.data_as(C.c_void_p))
