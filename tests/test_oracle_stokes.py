"""Pins of the CPU restatement of the nodal solves (oracle/stokes.py) by closed
forms: the plain pressure guess of initialPressureGuess.c gives the hydrostatic
pressure for uniform density; the Stokes box gives v = 0 and dp/dy = rho g for
density depending on depth only, and converges at second order to a
manufactured free-slip solution with variable viscosity."""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

from oracle import stokes as S
from fvmic2d_amd import workloads


class Box:
    def __init__(self, N, L=1.0, tanh=False):
        x = workloads.spacing_tanh_band(N, L) if tanh else workloads.spacing_uniform(N, L)
        self.xp, self.xcp = workloads.pad_grid(x, L)
        self.yp, self.ycp = self.xp, self.xcp
        self.NX = self.NY = N
        self.LX = self.LY = L
        self.x = self.y = x


@pytest.mark.parametrize("tanh", [False, True])
def test_pressure_guess_hydrostatic_closed_form(tanh):
    N, g, rho0 = 41, 9.81, 3300.0
    w = Box(N, tanh=tanh)
    A, b = S.pressure_guess_system(w, np.full((N, N), rho0), g)
    p = sla.spsolve(A.tocsc(), b).reshape(N, N)
    yc = w.ycp[2:N + 3]
    exact = rho0 * g * (yc[1:N] - yc[3])  # p pinned to 0 in cell (3, 3)
    # exact on a uniform grid; on the stretched grid the bottom condition uses
    # the inner spacing yc[NY-1]-yc[NY-2] for the ghost distance (:127), an
    # O(dh) departure where the spacing still varies at the wall
    tol = 1e-5 if tanh else 1e-12
    for ix in range(1, N):
        np.testing.assert_allclose(p[1:, ix], exact, rtol=0, atol=tol * np.abs(exact).max())
    # dummy rows (ix == 0 or jy == 0) decouple and stay 0
    assert np.all(p[0, :] == 0.0)


def test_pressure_guess_interior_rows_density_gradient():
    """rho = a + c*y: interior rows of A applied to p = g (a y + c y^2 / 2) at
    the cell centres equal the right-hand side g*d(rho)/dy (uniform grid: the
    5-point operator and the two-row density average are exact for it)."""
    N, g = 25, 10.0
    w = Box(N)
    rho = np.repeat((2.0 + 0.5 * w.y)[:, None], N, axis=1)
    A, b = S.pressure_guess_system(w, rho, g)
    yc = w.ycp[2:N + 3][:N]
    p = np.repeat((g * (2.0 * yc + 0.25 * yc * yc))[:, None], N, axis=1)
    r = (A @ p.ravel() - b).reshape(N, N)
    inner = r[3:N - 2, 2:N - 2].copy()
    inner[0, 1] = 0.0  # row (ix=3, jy=3) is the pinned cell
    assert np.abs(inner).max() < 1e-9 * g * 0.5
    assert np.all(b.reshape(N, N)[4:N - 2, 4:N - 2] > 0)  # g * d(rho)/dy with y downward


def test_stokes_hydrostatic():
    N = 17
    w = Box(N)
    eta = np.ones((N, N))
    mu = np.full((N, N), 1e99)
    Y = np.repeat(w.y[:, None], N, axis=1)
    rho = 1.0 + Y
    A, b, Kb, Kc = S.stokes_system(w, eta, eta, mu, mu, rho, 1.0, gy=10.0)
    sol = sla.spsolve(A.tocsc(), b)
    n = N * N
    assert np.abs(sol[:2 * n]).max() < 1e-12
    p = sol[2 * n:].reshape(N, N) * Kc
    yc = w.ycp[2:N + 3]
    dp = (p[2:, 5] - p[1:-1, 5]) / (yc[2:N] - yc[1:N - 1])
    np.testing.assert_allclose(dp, 10.0 * (rho[1:N - 1, 4] + rho[1:N - 1, 5]) / 2, rtol=1e-10)


def _manufactured(N, tanh=False):
    """Free-slip solution psi = sin(pi x) sin(2 pi y) with eta = 1 + 0.5 sin(pi x) cos(pi y)."""
    w = Box(N, tanh=tanh)
    pi = np.pi
    vx = lambda x, y: 2 * pi * np.sin(pi * x) * np.cos(2 * pi * y)
    vy = lambda x, y: -pi * np.cos(pi * x) * np.sin(2 * pi * y)
    pe = lambda x, y: np.cos(pi * x) * np.cos(pi * y)
    eta = lambda x, y: 1.0 + 0.5 * np.sin(pi * x) * np.cos(pi * y)
    h = 1e-4
    dx_ = lambda f: (lambda x, y: (f(x + h, y) - f(x - h, y)) / (2 * h))
    dy_ = lambda f: (lambda x, y: (f(x, y + h) - f(x, y - h)) / (2 * h))
    txx = lambda x, y: 2 * eta(x, y) * dx_(vx)(x, y)
    tyy = lambda x, y: 2 * eta(x, y) * dy_(vy)(x, y)
    txy = lambda x, y: eta(x, y) * (dy_(vx)(x, y) + dx_(vy)(x, y))
    Fx = lambda x, y: dx_(txx)(x, y) + dy_(txy)(x, y) - dx_(pe)(x, y)
    Fy = lambda x, y: dx_(txy)(x, y) + dy_(tyy)(x, y) - dy_(pe)(x, y)
    x, y = w.x, w.y
    xc, yc = w.xcp[2:N + 3], w.ycp[2:N + 3]
    etaS = eta(x[None, :], y[:, None])
    etaN = np.ones((N, N))
    etaN[1:, 1:] = eta(xc[None, 1:N], yc[1:N, None])
    mu = np.full((N, N), 1e99)
    A, b, Kb, Kc = S.stokes_system(w, etaS, etaN, mu, mu, np.zeros((N, N)), 1.0, gy=0.0)
    n = N * N
    b = b.copy()
    for jy in range(N - 1):
        for ix in range(1, N - 1):
            b[jy * N + ix] = Fx(x[ix], yc[jy + 1])
    for jy in range(1, N - 1):
        for ix in range(N - 1):
            b[n + jy * N + ix] = Fy(xc[ix + 1], y[jy])
    sol = sla.spsolve(A.tocsc(), b)
    ux = sol[:n].reshape(N, N)[:N - 1, 1:N - 1]
    ex = vx(x[None, 1:N - 1], yc[1:N, None])
    uy = sol[n:2 * n].reshape(N, N)[1:N - 1, :N - 1]
    ey = vy(xc[None, 1:N], y[1:N - 1, None])
    p = sol[2 * n:].reshape(N, N)[1:, 1:] * Kc
    pex = pe(xc[None, 1:N], yc[1:N, None])
    pex = pex - pex[0, 0]  # pinned cell (1, 1)
    return np.abs(ux - ex).max(), np.abs(uy - ey).max(), np.abs(p - pex).max()


def test_stokes_manufactured_second_order():
    e1 = _manufactured(17)
    e2 = _manufactured(33)
    for a, b_ in zip(e1, e2):
        assert b_ < a / 3.0, (e1, e2)  # second order: ratio ~4
    assert max(e2[:2]) < 0.02 and e2[2] < 0.1
