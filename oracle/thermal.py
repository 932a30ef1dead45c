"""CPU restatement of formThermalSystem (formThermalSystem.c:6-249) and the
subduction boundary temperatures it uses (benchmarkInitialConditions.c:75-135).

TEST INFRASTRUCTURE ONLY (the checker): imported by tests/ and by bench.py's
cpu_baseline leg, never by the product path.  One rank: every node owned.  Rows
are built node by node in the reference's branch order (left wall, right wall,
outflow, top, bottom, interior) with Python floats (IEEE doubles) and the C
library's erf/tan/cos/sqrt, in the reference's expression order.  Returns the
matrix as scipy CSR and the right-hand side; the reference's thermalRHS enters
holding the heating terms (``rhs_in``).  pow(c, 2.0) is written c*c, which is
what the reference's gcc -O3 build computes.
"""
import math

import numpy as np
import scipy.sparse as sp


def plate_depth(x, o):
    """benchmarkInitialConditions.c:80-93"""
    if abs(x - o.rootCenter) < o.rootWidth:
        c = math.cos(math.pi * (x - o.rootCenter) / o.rootWidth / 2.0)
        return o.overridingPlateThickness + o.rootThickness * (c * c)
    return o.overridingPlateThickness


def plate_geotherm(x, y, o):
    """:94-97 with mantle_temperature() = 1573 (:75-78)"""
    return y / plate_depth(x, o) * (1573.0 - 273.0) + 273.0


def slab_inflow_temperature(x, y, angle, age):
    """:128-135"""
    Ts, T0, kappa, s_in_year = 273.0, 1623.0, 0.7272e-6, 3.15576e7
    slab_age_s = age * 1.0e6 * s_in_year
    return Ts + (T0 - Ts) * math.erf((y - x * math.tan(angle)) / (2.0 * math.sqrt(kappa * slab_age_s)))


def thermal_system(x, y, nodal, rho0, opt, dt, do_steady=False, rhs_in=None):
    """formThermalSystem for one rank.  x, y: node coordinates; nodal: dict of
    (NY, NX) arrays kThermal, rho, Cp, lastT, vx.  Returns (A csr, b)."""
    NX, NY = len(x), len(y)
    k, rho, Cp, lastT, vx = (nodal[n] for n in ("kThermal", "rho", "Cp", "lastT", "vx"))
    b = np.zeros(NX * NY) if rhs_in is None else np.array(rhs_in, dtype=np.float64).ravel().copy()
    rows, cols, vals = [], [], []

    def put(i, j, v):
        rows.append(i)
        cols.append(j)
        vals.append(v)

    top = opt.thermalBCTop.type[0]
    if top not in (0, 1):
        raise ValueError("top thermal boundary type %d sets no row" % top)
    for jy in range(NY):
        for ix in range(NX):
            i = jy * NX + ix
            if ix == 0:
                put(i, i, 1.0)
                b[i] = slab_inflow_temperature(x[ix], y[jy], opt.slabAngle, opt.slabAge)
            elif ix == NX - 1 and (y[jy] <= plate_depth(x[NX - 1], opt) or
                                   (jy > 0 and (vx[jy, ix] + vx[jy - 1, ix]) / 2.0 < 0.0)):
                put(i, i, 1.0)
                b[i] = plate_geotherm(x[ix], y[jy], opt) if y[jy] <= plate_depth(x[ix], opt) else 1573.0
            elif ix == NX - 1:
                put(i, i, 1.0)
                b[i] = lastT[jy, ix]
            elif jy == 0:
                if top == 0:
                    put(i, i, 1.0)
                    b[i] = opt.thermalBCTop.value[0]
                else:
                    dyt1 = y[jy + 1] - y[jy]
                    put(i, i, -1.0 / dyt1)
                    put(i, i + NX, 1.0 / dyt1)
                    b[i] = 0.0
            elif jy == NY - 1:
                put(i, i, 1.0)
                b[i] = lastT[jy, ix]
            else:
                ka = (k[jy, ix] + k[jy, ix - 1]) / 2.0
                kb = (k[jy, ix] + k[jy, ix + 1]) / 2.0
                kc = (k[jy, ix] + k[jy - 1, ix]) / 2.0
                kd = (k[jy, ix] + k[jy + 1, ix]) / 2.0
                dx2 = x[ix + 1] - x[ix]
                dx1 = x[ix] - x[ix - 1]
                dy1 = y[jy] - y[jy - 1]
                dy2 = y[jy + 1] - y[jy]
                lap = 2.0 / (dx2 + dx1) * (kb / dx2 + ka / dx1) + 2.0 / (dy1 + dy2) * (kd / dy2 + kc / dy1)
                if not do_steady:
                    b[i] = b[i] + (rho0 + rho[jy, ix]) * Cp[jy, ix] / dt * lastT[jy, ix]
                    put(i, i, (rho0 + rho[jy, ix]) * Cp[jy, ix] / dt + lap)
                else:
                    put(i, i, lap)
                put(i, i + 1, -2.0 / (dx1 + dx2) * kb / dx2)
                put(i, i - 1, -2.0 / (dx1 + dx2) * ka / dx1)
                put(i, i - NX, -2.0 / (dy1 + dy2) * kc / dy1)
                put(i, i + NX, -2.0 / (dy1 + dy2) * kd / dy2)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(NX * NY, NX * NY))
    return A, b


def five_bands(A, NX, NY):
    """The matrix as the five node-ordered arrays (diagonal, east, west, south,
    north) that fvm_form_thermal_system writes; entries a row lacks are 0."""
    n = NX * NY
    A = A.tocsr()
    out = np.zeros((5, n))
    for off, r in ((0, 0), (1, 1), (-1, 2), (NX, 3), (-NX, 4)):
        d = A.diagonal(off)
        if off >= 0:
            out[r, :n - off] = d
        else:
            out[r, -off:] = d
    return out
