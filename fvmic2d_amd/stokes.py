"""Host-side Stokes solve: the measure of the learned initial guess.

The reference solves its visco-elasto-plastic Stokes system on the host with a
PETSc KSP (options prefix "stokes_", kspLinearSolve.c:13-97).  This module is
that host step for a variable-viscosity free-slip box: the system is assembled
from vep_system.c's stencils -- viscoelastic factor Z (:58-92), Kbond/Kcont
scaling (:115-123), continuity and pressure ghost rows (:240-275), interior
x-Stokes (:325-411) and y-Stokes (:445-533) stencils -- with this code's own
box boundaries (vx = 0 on the side walls, vy = 0 on top and bottom, mirrored
tangential velocity on the ghost rows, pressure pinned in one cell) in place
of the subduction-specific kinematic constraints.  Unknowns are ordered
[vx (NX*NY), vy (NX*NY), p (NX*NY)], nodal arrays [jy][ix]; p is the scaled
pressure (physical pressure / Kcont).

The solver is right-preconditioned restarted FGMRES (the residual it monitors
is the unpreconditioned ||b - A x|| / ||b||) with a u/p field split: the
velocity block solved exactly (sparse LU) and the Schur complement
approximated by the diagonal -Kcont^2 / etaN -- the analogue of
-stokes_ksp_type fgmres -stokes_pc_type fieldsplit with a nonzero initial
guess.  Iteration counts to rtol = 1e-8 are the figure the learned guess is
measured by.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla


class BoxGrid:
    """Node and centre coordinates of a grid (padded arrays as in workloads)."""

    def __init__(self, xp, yp, xcp, ycp, NX, NY):
        self.NX, self.NY = NX, NY
        self.x = xp[2:NX + 2]
        self.y = yp[2:NY + 2]
        self.xcp, self.ycp = xcp, ycp  # xc[i] = xcp[i + 2], valid for i in -2..NX+1

    @classmethod
    def of(cls, w):
        return cls(w.xp, w.yp, w.xcp, w.ycp, w.NX, w.NY)


def viscosity_z(etaS, etaN, muS, muN, dt):
    """vep_system.c:58-92: Z = 1 - eta/(eta + dt mu); etaZ = Z * eta."""
    zS = (etaS / (etaS + dt * muS)) * -1.0 + 1.0
    zN = (etaN / (etaN + dt * muN)) * -1.0 + 1.0
    return zS * etaS, zN * etaN


def assemble(g, etaS, etaN, muS, muN, rho, dt, gy, gx=0.0, pin=(1, 1)):
    """The box Stokes system: (A (CSR), b, Kbond, Kcont)."""
    NX, NY = g.NX, g.NY
    x, y = g.x, g.y
    XC = lambda i: g.xcp[np.asarray(i) + 2]
    YC = lambda j: g.ycp[np.asarray(j) + 2]
    eSZ, eNZ = viscosity_z(np.asarray(etaS, float), np.asarray(etaN, float), muS, muN, dt)
    etamin = float(eSZ.min())
    dx0, dy0 = x[1] - x[0], y[1] - y[0]
    Kbond = 4 * etamin / ((dx0 + dy0) * (dx0 + dy0))
    Kcont = 2.0 * etamin / (dx0 + dy0)
    n = NX * NY
    J, I = np.meshgrid(np.arange(NY), np.arange(NX), indexing="ij")
    J, I = J.ravel(), I.ravel()
    node = J * NX + I
    U, V, Pp = 0, n, 2 * n
    rows, cols, vals = [], [], []
    b = np.zeros(3 * n)

    def put(r, c, v, mask):
        rows.append(np.asarray(r)[mask])
        cols.append(np.asarray(c)[mask])
        vals.append(np.broadcast_to(v, mask.shape)[mask])

    # ---- continuity (p rows)
    r = Pp + node
    m_top = J == 0
    m_left = (J > 0) & (I == 0)
    m_pin = (J > 0) & (I > 0) & (I == pin[0]) & (J == pin[1])
    m_div = (J > 0) & (I > 0) & ~m_pin
    put(r, r, Kbond, m_top | m_left | m_pin)
    put(r, Pp + np.minimum(J + 1, NY - 1) * NX + I, -Kbond, m_top)
    put(r, Pp + J * NX + np.minimum(I + 1, NX - 1), -Kbond, m_left)
    Ic, Jc = np.maximum(I, 1), np.maximum(J, 1)
    hx = x[Ic] - x[Ic - 1]
    hy = y[Jc] - y[Jc - 1]
    put(r, U + (Jc - 1) * NX + Ic, Kcont / hx, m_div)
    put(r, U + (Jc - 1) * NX + Ic - 1, -Kcont / hx, m_div)
    put(r, V + Jc * NX + Ic - 1, Kcont / hy, m_div)
    put(r, V + (Jc - 1) * NX + Ic - 1, -Kcont / hy, m_div)

    # ---- x-momentum (vx rows)
    r = U + node
    m_wall = (I == 0) | (I == NX - 1)
    m_bot = ~m_wall & (J == NY - 1)
    m_in = ~m_wall & ~m_bot
    put(r, r, Kbond, m_wall)
    dyc = YC(NY) - YC(NY - 1)
    put(r, r, 1.0 / dyc * Kcont, m_bot)
    put(r, U + np.maximum(J - 1, 0) * NX + I, -1.0 / dyc * Kcont, m_bot)
    i = np.clip(I, 1, NX - 2)
    j = np.clip(J, 0, NY - 2)
    v0000, v0002, v0003 = XC(i), x[i], y[j]
    v0006 = eSZ[j, i]
    v0007, v0008, v0009 = XC(i + 1), y[j + 1], YC(j + 1)
    v0010 = eSZ[j + 1, i]
    v0011 = eNZ[j + 1, i]
    v0012 = eNZ[j + 1, i + 1]
    v0013 = 1.0 / (v0000 - v0007)
    v0014 = 1.0 / (v0002 - x[i + 1])
    v0015 = 1 / (v0003 - v0008)
    v0016 = 1 / (-v0009 + YC(j))
    v0017 = 1 / (-v0002 + x[i - 1])
    v0018 = 1 / (v0009 - YC(j + 2))
    diag = ((-4 * v0012 * v0013 * v0014) / 3. - (4 * v0011 * v0013 * v0017) / 3.
            + (v0006 * v0016 + v0010 * v0018) / (-v0003 + v0008))
    up = v0006 * v0015 * v0016
    top = J == 0
    diag = np.where(top, diag + up, diag)  # mirrored ghost row vx[-1] = vx[0]
    put(r, U + (j - 1) * NX + i, up, m_in & ~top)
    put(r, U + j * NX + i - 1, (4 * v0011 * v0013 * v0017) / 3., m_in)
    put(r, V + j * NX + i - 1, ((3 * v0006 - 2 * v0011) * v0013 * v0015) / 3., m_in)
    put(r, V + (j + 1) * NX + i - 1, ((-3 * v0010 + 2 * v0011) * v0013 * v0015) / 3., m_in)
    put(r, r, diag, m_in)
    put(r, V + j * NX + i, ((-3 * v0006 + 2 * v0012) * v0013 * v0015) / 3., m_in)
    put(r, U + (j + 1) * NX + i, v0010 * v0015 * v0018, m_in)
    put(r, V + (j + 1) * NX + i, ((3 * v0010 - 2 * v0012) * v0013 * v0015) / 3., m_in)
    put(r, Pp + (j + 1) * NX + i, Kcont / (-v0000 + v0007), m_in)
    put(r, U + j * NX + i + 1, (4 * v0012 * v0013 * v0014) / 3., m_in)
    put(r, Pp + (j + 1) * NX + i + 1, Kcont * v0013, m_in)
    b[U + node[m_in]] = (-gx * (rho[j, i] + rho[j + 1, i]) / 2.0)[m_in]

    # ---- y-momentum (vy rows)
    r = V + node
    m_wall = (J == 0) | (J == NY - 1)
    m_right = ~m_wall & (I == NX - 1)
    m_in = ~m_wall & ~m_right
    put(r, r, Kbond, m_wall)
    dxc = XC(NX) - XC(NX - 1)
    put(r, r, 1.0 / dxc * Kcont, m_right)
    put(r, V + J * NX + np.maximum(I - 1, 0), -1.0 / dxc * Kcont, m_right)
    i = np.clip(I, 0, NX - 2)
    j = np.clip(J, 1, NY - 2)
    v0000, v0002 = YC(j), y[j]
    v0005 = eSZ[j, i]
    v0006, v0007 = YC(j + 1), XC(i + 1)
    v0008 = eSZ[j, i + 1]
    v0009 = eNZ[j, i + 1]
    v0010 = eNZ[j + 1, i + 1]
    v0011 = 1 / (v0000 - v0006)
    v0012 = 1 / (-v0000 + v0006)
    v0013 = 1 / (x[i] - x[i + 1])
    v0014 = 1 / (v0002 - y[j + 1])
    v0015 = 1 / (-v0002 + y[j - 1])
    v0016 = 1 / (-v0007 + XC(i))
    v0017 = 1 / (v0007 - XC(i + 2))
    diag = ((4 * v0012 * (v0010 * v0014 + v0009 * v0015)) / 3. - v0005 * v0013 * v0016
            - v0008 * v0013 * v0017)
    left = v0005 * v0013 * v0016
    lft = I == 0
    diag = np.where(lft, diag + left, diag)  # mirrored ghost column vy[., -1] = vy[., 0]
    put(r, V + j * NX + i - 1, left, m_in & ~lft)
    put(r, U + (j - 1) * NX + i, ((3 * v0005 - 2 * v0009) * v0011 * v0013) / 3., m_in)
    put(r, V + (j - 1) * NX + i, (4 * v0009 * v0011 * v0015) / 3., m_in)
    put(r, U + j * NX + i, ((-3 * v0005 + 2 * v0010) * v0011 * v0013) / 3., m_in)
    put(r, r, diag, m_in)
    put(r, V + (j + 1) * NX + i, (4 * v0010 * v0011 * v0014) / 3., m_in)
    put(r, U + (j - 1) * NX + i + 1, ((-3 * v0008 + 2 * v0009) * v0011 * v0013) / 3., m_in)
    put(r, U + j * NX + i + 1, ((3 * v0008 - 2 * v0010) * v0011 * v0013) / 3., m_in)
    put(r, V + j * NX + i + 1, v0008 * v0013 * v0017, m_in)
    put(r, Pp + j * NX + i + 1, Kcont * v0012, m_in)
    put(r, Pp + (j + 1) * NX + i + 1, Kcont * v0011, m_in)
    b[V + node[m_in]] = (-gy * (rho[j, i] + rho[j, i + 1]) / 2.0)[m_in]

    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(3 * n, 3 * n))
    return A, b, Kbond, Kcont


class FieldSplit:
    """Block upper-triangular u/p preconditioner: z_p = S^-1 r_p with the
    diagonal Schur approximation -Kcont^2 / etaN (rows with their own
    pressure diagonal -- ghost and pinned rows -- use it), then
    z_u = A_uu^-1 (r_u - A_up z_p) by sparse LU."""

    def __init__(self, A, Kcont, etaN):
        n3 = A.shape[0]
        n = n3 // 3
        self.n = n
        A = A.tocsr()
        self.lu = sla.splu(A[:2 * n, :2 * n].tocsc())
        self.Aup = A[:2 * n, 2 * n:].tocsr()
        dpp = A[2 * n:, 2 * n:].diagonal()
        self.sdiag = np.where(dpp != 0.0, dpp, -Kcont * Kcont / np.asarray(etaN, float).ravel())

    def __call__(self, r):
        n = self.n
        zp = r[2 * n:] / self.sdiag
        zu = self.lu.solve(r[:2 * n] - self.Aup @ zp)
        return np.concatenate([zu, zp])


def fgmres(A, b, x0, M, rtol=1e-8, restart=100, maxit=5000):
    """Right-preconditioned restarted FGMRES.  Returns (x, iterations,
    relative residual); iterations counts inner (Arnoldi) steps, stopping at
    ||b - A x|| <= rtol ||b||."""
    x = np.array(x0, dtype=float, copy=True)
    bnorm = np.linalg.norm(b)
    if bnorm == 0.0:
        return np.zeros_like(x), 0, 0.0
    r = b - A @ x
    beta = np.linalg.norm(r)
    its = 0
    if beta <= rtol * bnorm:
        return x, 0, beta / bnorm
    while its < maxit:
        m = min(restart, maxit - its)
        Vb = np.zeros((m + 1, b.size))
        Z = np.zeros((m, b.size))
        H = np.zeros((m + 1, m))
        cs, sn = np.zeros(m), np.zeros(m)
        e = np.zeros(m + 1)
        e[0] = beta
        Vb[0] = r / beta
        k_done = 0
        for k in range(m):
            Z[k] = M(Vb[k])
            wv = A @ Z[k]
            for i in range(k + 1):  # modified Gram-Schmidt
                H[i, k] = wv @ Vb[i]
                wv = wv - H[i, k] * Vb[i]
            H[k + 1, k] = np.linalg.norm(wv)
            if H[k + 1, k] != 0.0:
                Vb[k + 1] = wv / H[k + 1, k]
            for i in range(k):  # apply the previous Givens rotations
                t = cs[i] * H[i, k] + sn[i] * H[i + 1, k]
                H[i + 1, k] = -sn[i] * H[i, k] + cs[i] * H[i + 1, k]
                H[i, k] = t
            den = np.hypot(H[k, k], H[k + 1, k])
            cs[k], sn[k] = H[k, k] / den, H[k + 1, k] / den
            H[k, k] = den
            H[k + 1, k] = 0.0
            e[k + 1] = -sn[k] * e[k]
            e[k] = cs[k] * e[k]
            its += 1
            k_done = k + 1
            if abs(e[k + 1]) <= rtol * bnorm:
                break
        y = np.linalg.solve(np.triu(H[:k_done, :k_done]), e[:k_done])
        x = x + Z[:k_done].T @ y
        r = b - A @ x
        beta = np.linalg.norm(r)
        if beta <= rtol * bnorm:
            break
    return x, its, beta / bnorm


def solve(A, b, Kcont, etaN, x0=None, rtol=1e-8, restart=100, maxit=5000, M=None):
    """FGMRES + field split from x0 (zeros if None): (x, iterations, relres, M)."""
    M = M or FieldSplit(A, Kcont, etaN)
    x0 = np.zeros(A.shape[0]) if x0 is None else x0
    x, its, rel = fgmres(A, b, x0, M, rtol=rtol, restart=restart, maxit=maxit)
    return x, its, rel, M
