"""GPU nodal solves against the CPU restatement (oracle/stokes.py): the plain
initial pressure guess (initialPressureGuess.c) -- operator and right-hand side
to 1e-13, solution to the direct solve of the same system."""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

from oracle import stokes as S
from fvmic2d_amd import workloads
from fvmic2d_amd.host import initial_pressure_guess, pressure_guess_apply
from fvmic2d_amd.lib import Context, DeviceArray, DeviceGrid, FvmError

pytestmark = pytest.mark.gpu


class Box:
    def __init__(self, NX, NY, LX=1.0, LY=1.0, tanh=False):
        sx = workloads.spacing_tanh_band if tanh else workloads.spacing_uniform
        self.xp, self.xcp = workloads.pad_grid(sx(NX, LX), LX)
        self.yp, self.ycp = workloads.pad_grid(sx(NY, LY), LY)
        self.NX, self.NY, self.LX, self.LY = NX, NY, LX, LY


def _rho(w, seed):
    r = np.random.default_rng(seed)
    X, Y = np.meshgrid(w.xp[2:w.NX + 2], w.yp[2:w.NY + 2])
    f = 3300.0 + 100.0 * Y
    for _ in range(4):
        cx, cy, s = r.uniform(0.2, 0.8) * w.LX, r.uniform(0.2, 0.8) * w.LY, r.uniform(0.05, 0.2) * w.LX
        f += r.uniform(-50, 50) * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / s ** 2)
    return f


@pytest.mark.parametrize("NX,NY,tanh", [(17, 13, False), (41, 33, True)])
def test_pressure_guess_operator(NX, NY, tanh):
    print("CASE pressure guess operator", NX, NY, tanh, flush=True)
    ctx = Context()
    w = Box(NX, NY, tanh=tanh)
    g = DeviceGrid(ctx, w)
    rho = _rho(w, 1)
    A, b = S.pressure_guess_system(w, rho, 9.81, 0.3)
    x = np.random.default_rng(2).standard_normal(NX * NY)
    d_rho, d_x = DeviceArray(ctx, rho), DeviceArray(ctx, x)
    d_y, d_b = DeviceArray(ctx, np.zeros(NX * NY)), DeviceArray(ctx, np.zeros(NX * NY))
    pressure_guess_apply(ctx, g, d_rho, 9.81, d_x, d_y, d_b, gx=0.3)
    ref = A @ x
    np.testing.assert_allclose(d_y.get(), ref, rtol=0, atol=1e-13 * np.abs(ref).max())
    np.testing.assert_allclose(d_b.get(), b, rtol=0, atol=1e-13 * np.abs(b).max())


def test_pressure_guess_solve():
    print("CASE pressure guess solve 65x49", flush=True)
    ctx = Context()
    w = Box(65, 49, LX=2.0, LY=1.0, tanh=True)
    n = w.NX * w.NY
    g = DeviceGrid(ctx, w)
    rho = _rho(w, 3)
    A, b = S.pressure_guess_system(w, rho, 9.81)
    ref = sla.spsolve(A.tocsc(), b)
    Sv = np.full(3 * n, 7.0)  # velocity slots must stay untouched
    d_rho, d_S, d_p = DeviceArray(ctx, rho), DeviceArray(ctx, Sv), DeviceArray(ctx, np.zeros(n))
    it, rel = initial_pressure_guess(ctx, g, d_rho, 9.81, rtol=1e-12, S=d_S, p_out=d_p)
    assert rel <= 1e-12 and 0 < it < 20000
    p = d_p.get()
    np.testing.assert_allclose(p, ref, rtol=0, atol=1e-8 * np.abs(ref).max())
    s = d_S.get().reshape(n, 3)
    assert np.array_equal(s[:, 2], p)
    assert np.all(s[:, :2] == 7.0)
    # deterministic: a second solve is bitwise identical
    d_p2 = DeviceArray(ctx, np.zeros(n))
    it2, rel2 = initial_pressure_guess(ctx, g, d_rho, 9.81, rtol=1e-12, p_out=d_p2)
    assert it2 == it and rel2 == rel and np.array_equal(d_p2.get(), p)


def test_pressure_guess_not_converged_is_an_error():
    ctx = Context()
    w = Box(33, 33)
    g = DeviceGrid(ctx, w)
    d_rho, d_p = DeviceArray(ctx, _rho(w, 4)), DeviceArray(ctx, np.zeros(33 * 33))
    with pytest.raises(FvmError) as e:
        initial_pressure_guess(ctx, g, d_rho, 9.81, rtol=1e-14, maxit=2, p_out=d_p)
    assert e.value.code == 11
