"""The host Stokes solve (fvmic2d_amd/stokes.py) against the CPU restatement
(oracle/stokes.py): identical assembled systems, and FGMRES + field split
converging to the direct solution."""
import numpy as np
import scipy.sparse.linalg as sla

from oracle import stokes as S
from fvmic2d_amd import stokes, workloads


class Box:
    def __init__(self, NX, NY, tanh=False):
        sx = workloads.spacing_tanh_band if tanh else workloads.spacing_uniform
        self.xp, self.xcp = workloads.pad_grid(sx(NX, 1.0), 1.0)
        self.yp, self.ycp = workloads.pad_grid(sx(NY, 1.0), 1.0)
        self.NX, self.NY = NX, NY


def _fields(NX, NY, seed, contrast=1e4):
    r = np.random.default_rng(seed)
    X, Y = np.meshgrid(np.linspace(0, 1, NX), np.linspace(0, 1, NY))
    B = np.zeros((NY, NX))
    for _ in range(5):
        cx, cy, s = r.uniform(0.1, 0.9), r.uniform(0.1, 0.9), r.uniform(0.05, 0.2)
        B += np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / s ** 2)
    B /= B.max()
    etaS = 10.0 ** (np.log10(contrast) * B)
    etaN = 10.0 ** (np.log10(contrast) * np.roll(B, 1, axis=0))
    mu = np.full((NY, NX), 1e10)
    rho = 3300.0 - 100.0 * B
    return etaS, etaN, mu, rho


def test_assembly_matches_oracle():
    for NX, NY, tanh in [(13, 9, False), (21, 17, True)]:
        w = Box(NX, NY, tanh)
        etaS, etaN, mu, rho = _fields(NX, NY, 1)
        A0, b0, Kb0, Kc0 = S.stokes_system(w, etaS, etaN, mu, mu, rho, 1e3, gy=9.81, gx=0.5)
        A1, b1, Kb1, Kc1 = stokes.assemble(stokes.BoxGrid.of(w), etaS, etaN, mu, mu, rho, 1e3, 9.81, gx=0.5)
        assert (Kb0, Kc0) == (Kb1, Kc1)
        D = (A0 - A1).tocoo()
        scale = np.abs(A0).max()
        assert D.nnz == 0 or np.abs(D.data).max() <= 1e-14 * scale
        assert A0.nnz == A1.nnz
        np.testing.assert_allclose(b1, b0, rtol=0, atol=1e-14 * np.abs(b0).max())


def test_fgmres_fieldsplit_matches_direct():
    NX = NY = 33
    w = Box(NX, NY)
    etaS, etaN, mu, rho = _fields(NX, NY, 2)
    A, b, Kb, Kc = stokes.assemble(stokes.BoxGrid.of(w), etaS, etaN, mu, mu, rho, 1e3, 9.81)
    ref = sla.spsolve(A.tocsc(), b)
    x, its, rel, M = stokes.solve(A, b, Kc, stokes.viscosity_z(etaS, etaN, mu, mu, 1e3)[1], rtol=1e-10)
    assert rel <= 1e-10 and 0 < its < 500
    assert np.abs(x - ref).max() <= 1e-6 * np.abs(ref).max()
    # a guess closer to the solution needs fewer iterations to 1e-8
    x1, its1, rel1, _ = stokes.solve(A, b, Kc, None, x0=ref * (1 + 1e-4), rtol=1e-8, M=M)
    x0, its0, rel0, _ = stokes.solve(A, b, Kc, None, x0=None, rtol=1e-8, M=M)
    assert its1 < its0
