"""Pins the oracle's formThermalSystem (oracle/thermal.py) with closed forms on
graded grids: the discrete operator of a linear temperature field is the
transient term alone, of x^2 (or y^2) with constant conductivity it is the
transient term minus 2k exactly on any spacing; the boundary temperatures
(slab inflow, plate geotherm, plate depth with its root) at points where they
are known; the right-wall branch by plate depth and the sign of vx; the top
gradient row; the steady variant."""
import math

import numpy as np
import pytest

from fvmic2d_amd import workloads
from oracle import thermal as TH


def _sys(w, steady=False):
    return TH.thermal_system(w.x, w.y, w.nodal, w.materials.materialRho[0], w.options, w.dt, steady,
                             w.nodal["heating"])


def _const_k(w):
    w.nodal["kThermal"][:] = 3.0
    return w


def _interior(w):
    m = np.zeros((w.NY, w.NX), bool)
    m[1:-1, 1:-1] = True
    return m.ravel()


@pytest.mark.parametrize("which", ["linear", "x2", "y2"])
def test_operator_on_polynomials(which):
    w = _const_k(workloads.thermal_box(23, 17))
    A, _ = _sys(w)
    X, Y = np.meshgrid(w.x, w.y)
    T = {"linear": 300.0 + 2e-3 * X + 5e-3 * Y, "x2": (X / 1e3) ** 2, "y2": (Y / 1e3) ** 2}[which].ravel()
    trans = ((w.materials.materialRho[0] + w.nodal["rho"]) * w.nodal["Cp"] / w.dt).ravel()
    expect = trans * T - {"linear": 0.0, "x2": 2 * 3.0 / 1e6, "y2": 2 * 3.0 / 1e6}[which]
    got = A @ T
    inner = _interior(w)
    scale = 2 * 3.0 / 1e6 + np.abs(trans * T).max()
    assert np.abs(got[inner] - expect[inner]).max() <= 1e-9 * scale


def test_boundary_temperatures_closed_forms():
    o = workloads.thermal_box(5, 5).options
    x = 100e3
    assert TH.slab_inflow_temperature(x, x * math.tan(o.slabAngle), o.slabAngle, o.slabAge) == 273.0
    assert abs(TH.slab_inflow_temperature(0.0, 1e9, o.slabAngle, o.slabAge) - 1623.0) < 1e-9
    assert TH.plate_depth(o.rootCenter, o) == o.overridingPlateThickness + o.rootThickness
    assert TH.plate_depth(o.rootCenter + 2 * o.rootWidth, o) == o.overridingPlateThickness
    half = TH.plate_depth(o.rootCenter + o.rootWidth / 2, o)
    assert abs(half - (o.overridingPlateThickness + o.rootThickness * 0.5)) < 1e-6
    assert TH.plate_geotherm(0.0, 0.0, o) == 273.0
    assert abs(TH.plate_geotherm(0.0, o.overridingPlateThickness, o) - 1573.0) < 1e-9


def test_rows_by_branch():
    w = workloads.thermal_box(15, 13)
    A, b = _sys(w)
    A = A.tocsr()
    NX, NY = w.NX, w.NY
    o = w.options
    for jy in range(NY):
        i = jy * NX
        assert A[i, i] == 1.0 and b[i] == TH.slab_inflow_temperature(w.x[0], w.y[jy], o.slabAngle, o.slabAge)
        i = jy * NX + NX - 1
        assert A[i, i] == 1.0 and A[i].nnz == 1
        if w.y[jy] <= TH.plate_depth(w.x[-1], o):
            assert b[i] == TH.plate_geotherm(w.x[-1], w.y[jy], o)
        elif (w.nodal["vx"][jy, -1] + w.nodal["vx"][jy - 1, -1]) / 2.0 < 0.0:
            assert b[i] == 1573.0
        else:
            assert b[i] == w.nodal["lastT"][jy, -1]
    for ix in range(1, NX - 1):
        assert b[ix] == 273.0 and A[ix, ix] == 1.0                       # top, type 0
        i = (NY - 1) * NX + ix
        assert b[i] == w.nodal["lastT"][-1, ix] and A[i, i] == 1.0       # bottom
    i = 5 * NX + 6                                                         # interior RHS
    tr = (w.materials.materialRho[0] + w.nodal["rho"][5, 6]) * w.nodal["Cp"][5, 6] / w.dt
    assert b[i] == w.nodal["heating"][5, 6] + tr * w.nodal["lastT"][5, 6]
    # zero-gradient top (type 1)
    w1 = workloads.thermal_box(15, 13, top_type=1)
    A1, b1 = _sys(w1)
    dy = w1.y[1] - w1.y[0]
    assert A1[3, 3] == -1.0 / dy and A1[3, 3 + NX] == 1.0 / dy and b1[3] == 0.0
    # steady: no transient term, RHS = heating on the interior
    As, bs = _sys(w, steady=True)
    assert bs[i] == w.nodal["heating"][5, 6]
    assert abs(As[i, i] - (A[i, i] - tr)) <= 1e-12 * A[i, i]
    with pytest.raises(ValueError):
        w2 = workloads.thermal_box(7, 7, top_type=2)
        _sys(w2)
