"""CPU restatement of the nodal solves around the marker path: the plain
initial pressure guess and the staggered visco-elasto-plastic Stokes system.

TEST INFRASTRUCTURE ONLY (the checker): imported by tests/ and by bench.py's
baseline legs, never by the product path.

- pressure_guess_system: initialPressureGuess.c:59-135, row for row (natural
  node order jy*NX + ix): dummy rows for ix == 0 or jy == 0 (:65-71), the pinned
  cell (3, 3) (:72-77), the 5-point variable-spacing Laplacian with the
  density-gradient right-hand side (:79-106), zero-flux sides (:110-118) and
  dp/dy = rho*g on top and bottom (:119-132).
- stokes_system: the generic free-slip box built from vep_system.c's stencils:
  the viscoelastic viscosity factor Z (:58-92), Kbond/Kcont (:115-123), the
  continuity rows and pressure ghost rows (:240-275), the interior x-Stokes
  (:325-411) and y-Stokes (:445-533) stencils.  Boundaries are this code's own
  free-slip box (the reference's subduction-specific kinematic constraints,
  open right/bottom boundaries and plate-depth pressure pin are not part of
  it): vx = 0 on the left/right walls, vy = 0 on the top/bottom walls, mirrored
  tangential velocity (zero shear) on the ghost rows, pressure pinned in cell
  `pin`.  Unknown order: [vx (NX*NY), vy (NX*NY), p (NX*NY)].
"""
import numpy as np
import scipy.sparse as sp


def _grid(w):
    """Unpadded views: x[0..NX-1], y[0..NY-1], xc[0..NX], yc[0..NY] plus one ghost on each side."""
    NX, NY = w.NX, w.NY
    x = w.xp[2:NX + 2]
    y = w.yp[2:NY + 2]
    xc = w.xcp[2:NX + 3]
    yc = w.ycp[2:NY + 3]
    return NX, NY, x, y, xc, yc, w.xcp, w.ycp


def pressure_guess_system(w, rho, gy, gx=0.0):
    """LHS (CSR) and RHS of initialPressureGuess.c for nodal density rho[NY, NX]."""
    NX, NY, x, y, xc, yc, _, _ = _grid(w)
    n = NX * NY
    rows, cols, vals = [], [], []
    b = np.zeros(n)

    def add(r, c, v):
        rows.append(r)
        cols.append(c)
        vals.append(v)

    for ix in range(NX):
        for jy in range(NY):
            r = jy * NX + ix
            if ix == 0 or jy == 0:
                dx = x[1] - x[0]
                dy = y[1] - y[0]
                add(r, r, 4.0 / (dx * dx + dy * dy))
            elif ix == 3 and jy == 3:
                dx = x[ix] - x[ix - 1]
                dy = y[jy] - y[jy - 1]
                add(r, r, 4.0 / (dx * dx + dy * dy))
            else:
                hx = x[ix] - x[ix - 1]
                hy = y[jy] - y[jy - 1]
                if ix != NX - 1:
                    add(r, r + 1, 1.0 / hx / (xc[ix + 1] - xc[ix]))
                add(r, r, 1.0 / hx * (-1.0 / (xc[ix + 1] - xc[ix]) - 1.0 / (xc[ix] - xc[ix - 1]))
                    + 1.0 / hy * (-1.0 / (yc[jy + 1] - yc[jy]) - 1.0 / (yc[jy] - yc[jy - 1])))
                if ix != 1:
                    add(r, r - 1, 1.0 / hx / (xc[ix] - xc[ix - 1]))
                if jy != NY - 1:
                    add(r, r + NX, 1.0 / hy / (yc[jy + 1] - yc[jy]))
                if jy != 1:
                    add(r, r - NX, 1.0 / hy / (yc[jy] - yc[jy - 1]))
                b[r] += (gy * ((rho[jy, ix] + rho[jy, ix - 1]) / 2.0 - (rho[jy - 1, ix] + rho[jy - 1, ix - 1]) / 2.0)
                         / hy
                         + gx * ((rho[jy, ix] + rho[jy - 1, ix]) / 2.0 - (rho[jy, ix - 1] + rho[jy - 1, ix - 1]) / 2.0)
                         / hx)
            if ix == 1:
                add(r, r, 1.0 / (x[ix] - x[ix - 1]) / (xc[ix] - xc[ix - 1]))
            elif ix == NX - 1:
                add(r, r, 1.0 / (x[ix] - x[ix - 1]) / (xc[ix + 1] - xc[ix]))
            # :119-132 apply to every row of lines 1 and NY-1 (ix == 0 included);
            # density outside the grid is a zero ghost (parity unpinned: it
            # rests on a fresh PETSc local vector being zero)
            R = lambda j, i: rho[j, i] if i >= 0 else 0.0
            if jy == 1:
                c = 1.0 / (y[jy] - y[jy - 1]) / (yc[jy] - yc[jy - 1])
                rhoc = (R(jy, ix) + R(jy, ix - 1) + R(jy - 1, ix) + R(jy - 1, ix - 1)) / 4.0
                dy = yc[jy] - yc[jy - 1]
                add(r, r, c)
                b[r] += c * rhoc * gy * dy
            elif jy == NY - 1:
                c = 1.0 / (y[jy] - y[jy - 1]) / (yc[jy + 1] - yc[jy])
                rhoc = (R(jy, ix) + R(jy, ix - 1) + R(jy - 1, ix) + R(jy - 1, ix - 1)) / 4.0
                dy = yc[jy] - yc[jy - 1]
                add(r, r, c)
                b[r] += -c * rhoc * gy * dy
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))  # duplicates are summed (ADD_VALUES)
    return A, b


def viscosity_z(etaS, etaN, muS, muN, dt):
    """vep_system.c:58-92: X = eta/(eta + dt*mu); X = -X + 1; etaZ = X*eta."""
    XN = etaN / (etaN + dt * muN)
    XN = XN * -1.0 + 1.0
    XS = etaS / (etaS + dt * muS)
    XS = XS * -1.0 + 1.0
    return XS * etaS, XN * etaN


def stokes_scales(w, etaSZ):
    """Kbond, Kcont (vep_system.c:115-123)."""
    NX, NY, x, y, xc, yc, _, _ = _grid(w)
    etamin = float(np.min(etaSZ))
    dx = x[1] - x[0]
    dy = y[1] - y[0]
    return 4 * etamin / ((dx + dy) * (dx + dy)), 2.0 * etamin / (dx + dy)


def stokes_system(w, etaS, etaN, muS, muN, rho, dt, gy, gx=0.0, pin=(1, 1)):
    """LHS (CSR), RHS, Kbond, Kcont of the free-slip box Stokes problem."""
    NX, NY, x, y, xc, yc, xcp, ycp = _grid(w)
    etaSZ, etaNZ = viscosity_z(etaS, etaN, muS, muN, dt)
    Kbond, Kcont = stokes_scales(w, etaSZ)
    n = NX * NY
    U, V, Pn = 0, n, 2 * n
    rows, cols, vals = [], [], []
    b = np.zeros(3 * n)

    def add(r, c, v):
        rows.append(r)
        cols.append(c)
        vals.append(v)

    def XC(i):  # xc with one ghost below index 0
        return xcp[i + 2]

    def YC(j):
        return ycp[j + 2]

    def vx(jy, ix):
        return U + jy * NX + ix

    def vy(jy, ix):
        return V + jy * NX + ix

    def p(jy, ix):
        return Pn + jy * NX + ix

    for jy in range(NY):
        for ix in range(NX):
            # continuity
            r = p(jy, ix)
            if jy == 0:
                add(r, r, Kbond)
                add(r, p(1, ix), -Kbond)
            elif ix == 0:
                add(r, r, Kbond)
                add(r, p(jy, 1), -Kbond)
            elif (ix, jy) == tuple(pin):
                add(r, r, Kbond)
            else:
                dx = x[ix] - x[ix - 1]
                dy = y[jy] - y[jy - 1]
                add(r, vx(jy - 1, ix), Kcont / dx)
                add(r, vx(jy - 1, ix - 1), -Kcont / dx)
                add(r, vy(jy, ix - 1), Kcont / dy)
                add(r, vy(jy - 1, ix - 1), -Kcont / dy)
            # x-momentum
            r = vx(jy, ix)
            if ix == 0 or ix == NX - 1:
                add(r, r, Kbond)
            elif jy == NY - 1:
                dyc = yc[NY] - yc[NY - 1]
                add(r, r, 1.0 / dyc * Kcont)
                add(r, vx(jy - 1, ix), -1.0 / dyc * Kcont)
            else:
                v0000, v0001, v0002, v0003 = XC(ix), Kcont, x[ix], y[jy]
                v0006 = etaSZ[jy, ix]
                v0007, v0008, v0009 = XC(ix + 1), y[jy + 1], YC(jy + 1)
                v0010 = etaSZ[jy + 1, ix]
                v0011 = etaNZ[jy + 1, ix]
                v0012 = etaNZ[jy + 1, ix + 1]
                v0013 = 1.0 / (v0000 - v0007)
                v0014 = 1.0 / (v0002 - x[ix + 1])
                v0015 = 1 / (v0003 - v0008)
                v0016 = 1 / (-v0009 + YC(jy))
                v0017 = 1 / (-v0002 + x[ix - 1])
                v0018 = 1 / (v0009 - YC(jy + 2))
                diag = ((-4 * v0012 * v0013 * v0014) / 3. - (4 * v0011 * v0013 * v0017) / 3.
                        + (v0006 * v0016 + v0010 * v0018) / (-v0003 + v0008))
                up = v0006 * v0015 * v0016
                if jy == 0:
                    diag = diag + up  # mirrored ghost row: vx[-1] = vx[0]
                else:
                    add(r, vx(jy - 1, ix), up)
                add(r, vx(jy, ix - 1), (4 * v0011 * v0013 * v0017) / 3.)
                add(r, vy(jy, ix - 1), ((3 * v0006 - 2 * v0011) * v0013 * v0015) / 3.)
                add(r, vy(jy + 1, ix - 1), ((-3 * v0010 + 2 * v0011) * v0013 * v0015) / 3.)
                add(r, r, diag)
                add(r, vy(jy, ix), ((-3 * v0006 + 2 * v0012) * v0013 * v0015) / 3.)
                add(r, vx(jy + 1, ix), v0010 * v0015 * v0018)
                add(r, vy(jy + 1, ix), ((3 * v0010 - 2 * v0012) * v0013 * v0015) / 3.)
                add(r, p(jy + 1, ix), v0001 / (-v0000 + v0007))
                add(r, vx(jy, ix + 1), (4 * v0012 * v0013 * v0014) / 3.)
                add(r, p(jy + 1, ix + 1), v0001 * v0013)
                b[r] = -gx * (rho[jy, ix] + rho[jy + 1, ix]) / 2.0
            # y-momentum
            r = vy(jy, ix)
            if jy == 0 or jy == NY - 1:
                add(r, r, Kbond)
            elif ix == NX - 1:
                dxc = xc[NX] - xc[NX - 1]
                add(r, r, 1.0 / dxc * Kcont)
                add(r, vy(jy, ix - 1), -1.0 / dxc * Kcont)
            else:
                v0000, v0001, v0002 = YC(jy), Kcont, y[jy]
                v0005 = etaSZ[jy, ix]
                v0006, v0007 = YC(jy + 1), XC(ix + 1)
                v0008 = etaSZ[jy, ix + 1]
                v0009 = etaNZ[jy, ix + 1]
                v0010 = etaNZ[jy + 1, ix + 1]
                v0011 = 1 / (v0000 - v0006)
                v0012 = 1 / (-v0000 + v0006)
                v0013 = 1 / (x[ix] - x[ix + 1])
                v0014 = 1 / (v0002 - y[jy + 1])
                v0015 = 1 / (-v0002 + y[jy - 1])
                v0016 = 1 / (-v0007 + XC(ix))
                v0017 = 1 / (v0007 - XC(ix + 2))
                diag = ((4 * v0012 * (v0010 * v0014 + v0009 * v0015)) / 3. - v0005 * v0013 * v0016
                        - v0008 * v0013 * v0017)
                left = v0005 * v0013 * v0016
                if ix == 0:
                    diag = diag + left  # mirrored ghost column: vy[., -1] = vy[., 0]
                else:
                    add(r, vy(jy, ix - 1), left)
                add(r, vx(jy - 1, ix), ((3 * v0005 - 2 * v0009) * v0011 * v0013) / 3.)
                add(r, vy(jy - 1, ix), (4 * v0009 * v0011 * v0015) / 3.)
                add(r, vx(jy, ix), ((-3 * v0005 + 2 * v0010) * v0011 * v0013) / 3.)
                add(r, r, diag)
                add(r, vy(jy + 1, ix), (4 * v0010 * v0011 * v0014) / 3.)
                add(r, vx(jy - 1, ix + 1), ((-3 * v0008 + 2 * v0009) * v0011 * v0013) / 3.)
                add(r, vx(jy, ix + 1), ((3 * v0008 - 2 * v0010) * v0011 * v0013) / 3.)
                add(r, vy(jy, ix + 1), v0008 * v0013 * v0017)
                add(r, p(jy, ix + 1), v0001 * v0012)
                add(r, p(jy + 1, ix + 1), v0001 * v0011)
                b[r] = -gy * (rho[jy, ix] + rho[jy, ix + 1]) / 2.0
    A = sp.csr_matrix((vals, (rows, cols)), shape=(3 * n, 3 * n))
    return A, b, Kbond, Kcont
