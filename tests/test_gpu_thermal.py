"""The energy equation on the GPU (fvm_form_thermal_system,
fvm_solve_thermal_system) against the oracle's formThermalSystem: every matrix
coefficient and right-hand-side entry to 1e-14 relative (the boundary
temperatures go through the device's erf/tan/cos), and the BiCGStab solution
against scipy's direct solve of the oracle's system to 1e-8."""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

from fvmic2d_amd import workloads
from fvmic2d_amd.host import MarkerPath
from fvmic2d_amd.lib import DeviceArray, FvmError
from oracle import thermal as TH

pytestmark = pytest.mark.gpu

CASES = [((41, 29), 0, False), ((64, 33), 1, False), ((41, 29), 0, True), ((801, 401), 0, False)]


def _run(shape, top, steady):
    print("  case: thermal_box %dx%d top type %d steady=%s" % (shape[0], shape[1], top, steady), flush=True)
    w = workloads.thermal_box(*shape, top_type=top)
    mp = MarkerPath(w)
    nn = w.NX * w.NY
    coef = DeviceArray(mp.ctx, np.zeros(5 * nn))
    rhs = DeviceArray(mp.ctx, w.nodal["heating"])
    mp.form_thermal_system(w.dt, steady, coef, rhs)
    A, b = TH.thermal_system(w.x, w.y, w.nodal, w.materials.materialRho[0], w.options, w.dt, steady,
                             w.nodal["heating"])
    return w, mp, coef, rhs, A, b


@pytest.mark.parametrize("shape,top,steady", CASES)
def test_form_thermal_system_matches_oracle(shape, top, steady):
    w, mp, coef, rhs, A, b = _run(shape, top, steady)
    bands = TH.five_bands(A, w.NX, w.NY)
    got = coef.get().reshape(5, -1)
    for r, name in enumerate(("diagonal", "east", "west", "south", "north")):
        scale = np.abs(bands[r]).max() or 1.0
        err = np.abs(got[r] - bands[r]).max() / scale
        assert err <= 1e-14, (name, err)
    gb = rhs.get().ravel()
    assert np.abs(gb - b).max() <= 1e-14 * np.abs(b).max()


@pytest.mark.parametrize("shape,top,steady", CASES[:3])
def test_solve_thermal_system_matches_direct(shape, top, steady):
    w, mp, coef, rhs, A, b = _run(shape, top, steady)
    T = DeviceArray(mp.ctx, np.zeros(w.NX * w.NY))
    its, rel = mp.solve_thermal_system(coef, rhs, T, rtol=1e-12, maxit=50000)
    ref = sla.spsolve(A.tocsc(), b)
    err = np.abs(T.get().ravel() - ref).max() / np.abs(ref).max()
    print("  %d iterations, relres %.2e, error vs direct %.2e" % (its, rel, err), flush=True)
    assert rel <= 1e-12 and err <= 1e-8


def test_top_type_without_row_is_an_error():
    w = workloads.thermal_box(9, 7)
    w.options.thermalBCTop.type[0] = 2
    mp = MarkerPath(w)
    nn = w.NX * w.NY
    with pytest.raises(FvmError):
        mp.form_thermal_system(w.dt, False, DeviceArray(mp.ctx, np.zeros(5 * nn)), DeviceArray(mp.ctx, np.zeros(nn)))


@pytest.mark.parametrize("shape,top", [((3, 3), 0), ((3, 4), 1), ((5, 3), 0)])
def test_thermal_minimal_grids(shape, top):
    w, mp, coef, rhs, A, b = _run(shape, top, False)
    bands = TH.five_bands(A, w.NX, w.NY)
    got = coef.get().reshape(5, -1)
    for r in range(5):
        scale = np.abs(bands[r]).max() or 1.0
        assert np.abs(got[r] - bands[r]).max() <= 1e-14 * scale
    assert np.abs(rhs.get().ravel() - b).max() <= 1e-14 * np.abs(b).max()
    T = DeviceArray(mp.ctx, np.zeros(w.NX * w.NY))
    mp.solve_thermal_system(coef, rhs, T, rtol=1e-12, maxit=1000)
    ref = sla.spsolve(A.tocsc(), b)
    assert np.abs(T.get().ravel() - ref).max() <= 1e-8 * np.abs(ref).max()
