"""Seeded synthetic inputs for the marker hot path.

Five workloads, each named by what it is:

* ``convection_small``   51x51 uniform grid, 4x4 markers per cell (~40k markers),
  isoviscous cellular flow; step = locate, density, projectMarkersNodesAll2,
  subgrid temperature, RK4 advection.
* ``convection_tdep``    201x201 uniform grid, 6x6 markers per cell (1.44M),
  eta = exp(-ln(1e5) T); T, eta, rho, k projected onto all four staggered node
  families (+ Cp on basic nodes); same step.
* ``vep_subduction``     801x401 tanh-stretched grid (4x refinement in a central
  band), 5x5 markers per cell (8M), 5 materials, Drucker-Prager yield; full
  visco-elasto-plastic step.
* ``ensemble``           N instances of the convection_tdep generator with random
  sizes 65..257 and 4..6 markers per cell per axis.

Random numbers come from a counter-based SplitMix64 stream, so every array is a
pure function of (seed, stream id, index) and is identical on every machine.
Marker order follows distributeMarkersUniformInDomain (markers.c:382-395): the
sub-lattice column index k is outer, the row index l inner.
"""
from dataclasses import dataclass, field
import math

import numpy as np

from . import abi

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _splitmix64(x):
    with np.errstate(over="ignore"):
        z = x + _GOLD
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


class Rng:
    """Counter-based stream: draw(n) returns the next n values of stream ``sid``."""

    def __init__(self, seed, sid):
        with np.errstate(over="ignore"):
            self.key = _splitmix64(np.uint64(seed) * np.uint64(1_000_003) + np.uint64(sid))
        self.pos = 0

    def u64(self, n):
        c = np.arange(self.pos, self.pos + n, dtype=np.uint64)
        self.pos += n
        with np.errstate(over="ignore"):
            return _splitmix64(c ^ self.key)

    def uniform(self, n, lo=0.0, hi=1.0):
        u = (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self, n):
        u1 = self.uniform(n)
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


# ---------------------------------------------------------------- grids ----
def spacing_uniform(N, L):
    """gridSpacing.c:8-17: x[i] = x[i-1] + dx, x[N-1] = L."""
    dx = L / (float(N) - 1.0)
    x = np.empty(N)
    x[0] = 0.0
    acc = 0.0
    for i in range(1, N - 1):
        acc = acc + dx
        x[i] = acc
    x[N - 1] = L
    return x


def spacing_tanh_band(N, L, refine=4.0, lo=0.35, hi=0.65, width=0.04):
    """Monotone spacing refined `refine`x inside [lo, hi]*L with tanh transitions."""
    s = (np.arange(N - 1) + 0.5) / (N - 1)
    band = 0.5 * (np.tanh((s - lo) / width) - np.tanh((s - hi) / width))
    h = 1.0 / (1.0 + (refine - 1.0) * band)
    # widths in a cell-count parametrisation: more cells where h is small
    cum = np.concatenate([[0.0], np.cumsum(h)])
    x = L * cum / cum[-1]
    x[0] = 0.0
    x[-1] = L
    return x


def pad_grid(x, L):
    """Padded node and centre arrays (length N+4, element 2 is index 0).

    Centres follow getCellCenters (gridSpacing.c:150-157); the two ghost slots on
    each side are filled by linear extrapolation (the reference leaves them unset
    and never reads them on a non-periodic grid)."""
    N = len(x)
    xp = np.empty(N + 4)
    xp[2:N + 2] = x
    h0, h1 = x[1] - x[0], x[N - 1] - x[N - 2]
    xp[1], xp[0] = x[0] - h0, x[0] - 2.0 * h0
    xp[N + 2], xp[N + 3] = x[N - 1] + h1, x[N - 1] + 2.0 * h1
    xc = np.empty(N + 1)
    xc[1:N] = (x[0:N - 1] + x[1:N]) / 2.0
    xc[0] = -xc[1]
    xc[N] = 2.0 * L - xc[N - 1]
    xcp = np.empty(N + 4)
    xcp[2:N + 3] = xc
    xcp[1], xcp[0] = xc[0] - (xc[1] - xc[0]), xc[0] - 2.0 * (xc[1] - xc[0])
    xcp[N + 3] = xc[N] + (xc[N] - xc[N - 1])
    return xp, xcp


def pad_periodic(x, L):
    """Padded node and centre arrays of a periodic x axis (length NX+4, element 2
    is index 0).  x holds the NX+1 nodes 0 = x[0] < ... < x[NX] = L of the NX
    cells; centres follow getCellCenters over NX nodes (gridSpacing.c:150-157),
    then the ghosts of makePeriodic (gridSpacing.c:159-168): x[-2], x[-1],
    x[NX], x[NX+1], xc[-1], xc[0], xc[NX], xc[NX+1].  xc[-2], which the
    reference leaves unset, is extrapolated."""
    NX = len(x) - 1
    xp = np.empty(NX + 4)
    xp[2:NX + 2] = x[:NX]
    xp[NX + 2] = L
    xp[NX + 3] = L + x[1]
    xp[1] = x[NX - 1] - L
    xp[0] = x[NX - 2] - L
    xc = np.empty(NX + 2)
    xc[1:NX] = (x[0:NX - 1] + x[1:NX]) / 2.0
    xc[NX] = (L + x[NX - 1]) / 2.0
    xc[NX + 1] = xc[1] + L
    xc[0] = 0.0 - (L - xc[NX])
    xcp = np.empty(NX + 4)
    xcp[2:NX + 4] = xc
    xcp[1] = 0.0 - (L - xc[NX - 1])
    xcp[0] = xcp[1] - (xcp[2] - xcp[1])
    return xp, xcp


# ------------------------------------------------------------ workload ----
@dataclass
class Workload:
    name: str
    NX: int
    NY: int
    LX: float
    LY: float
    xp: np.ndarray            # padded node coordinates (NX+4)
    yp: np.ndarray
    xcp: np.ndarray           # padded centre coordinates (NX+4)
    ycp: np.ndarray
    markers: np.ndarray       # MARKER_DTYPE, AoS 216 B
    nodal: dict               # name -> (NY, NX) float64 inputs
    materials: object
    options: object
    bv: object
    dt: float
    kind: str                 # "plain" or "vep"
    projections: list = field(default_factory=list)  # (src, stagx, stagy, method, out)
    seed: int = 0
    xperiodic: int = 0        # 1: x periodic (grid from pad_periodic)

    @property
    def x(self):
        return self.xp[2:2 + self.NX]

    @property
    def y(self):
        return self.yp[2:2 + self.NY]

    @property
    def xc(self):
        return self.xcp[2:2 + self.NX + 1]

    @property
    def yc(self):
        return self.ycp[2:2 + self.NY + 1]

    @property
    def n_markers(self):
        return len(self.markers)


def _bc_free_slip():
    bv = abi.BoundaryValues()
    for side in ("mechBCLeft", "mechBCRight", "mechBCTop", "mechBCBottom"):
        b = getattr(bv, side)
        for k in range(3):
            b.type[k] = 1
            b.value[k] = 0.0
    return bv


def _markers_in_domain(rng, x, y, NX, NY, NMX, NMY, LX, LY):
    """distributeMarkersUniformInDomain (markers.c:342-400) with a counter RNG."""
    mdx = LX / (NMX * (NX - 1))
    mdy = LY / (NMY * (NY - 1))
    nk = NMX * (NX - 1)
    nl = NMY * (NY - 1)
    k = np.repeat(np.arange(nk), nl)
    l = np.tile(np.arange(nl), nk)
    jx = rng.uniform(nk * nl, -mdx / 4.0, mdx / 4.0)
    jy = rng.uniform(nk * nl, -mdy / 4.0, mdy / 4.0)
    X = x[0] + mdx * (0.5 + k) + jx
    Y = y[0] + mdy * (0.5 + l) + jy
    return X, Y


def _markers_in_cells(rng, x, y, NMX, NMY):
    """NMX x NMY jittered sub-lattice in every cell of a nonuniform grid
    (the placement of distributeMarkersUniformInCells, markers.c:404-452)."""
    NX, NY = len(x), len(y)
    cx = np.repeat(np.arange(NX - 1), NMX)   # sub-column -> cell
    ix = np.tile(np.arange(NMX), NX - 1)
    cy = np.repeat(np.arange(NY - 1), NMY)
    iy = np.tile(np.arange(NMY), NY - 1)
    dxs = (x[cx + 1] - x[cx]) / NMX
    dys = (y[cy + 1] - y[cy]) / NMY
    colX = x[cx] + dxs * (0.5 + ix)
    rowY = y[cy] + dys * (0.5 + iy)
    nk, nl = len(colX), len(rowY)
    X = np.repeat(colX, nl) + rng.uniform(nk * nl, -0.25, 0.25) * np.repeat(dxs, nl)
    Y = np.tile(rowY, nk) + rng.uniform(nk * nl, -0.25, 0.25) * np.tile(dys, nk)
    return X, Y


def _new_markers(n):
    mk = np.zeros(n, dtype=abi.MARKER_DTYPE)
    mk["cellX"] = -1
    mk["cellY"] = -1
    return mk


def discrete_vorticity(vx, vy, xc, yc, bv):
    """wxy and exy at basic nodes (nodalStressStrain.c:157-236), ghost values by BC."""
    NY, NX = vx.shape
    vxm = np.empty_like(vx)
    vym = np.empty_like(vy)
    vym[:, 1:] = vy[:, :-1]
    if bv.mechBCLeft.type[1] == 0:
        vym[:, 0] = 2.0 * bv.mechBCLeft.value[1] - vy[:, 0]
    else:
        vym[:, 0] = vy[:, 0]
    vxm[1:, :] = vx[:-1, :]
    if bv.mechBCTop.type[0] == 0:
        vxm[0, :] = 2.0 * bv.mechBCTop.value[0] - vx[0, :]
    else:
        vxm[0, :] = vx[0, :]
    dx = (xc[1:NX + 1] - xc[0:NX])[None, :]
    dy = (yc[1:NY + 1] - yc[0:NY])[:, None]
    wxy = 0.5 * ((vx - vxm) / dy - (vy - vym) / dx)
    exy = 0.5 * ((vx - vxm) / dy + (vy - vym) / dx)
    return wxy, exy


def cell_strain_rates(vx, vy, x, y):
    """exx, eyy at cell centres (nodalStressStrain.c:55-75); row/col 0 are ghosts = 0."""
    NY, NX = vx.shape
    exx = np.zeros_like(vx)
    eyy = np.zeros_like(vx)
    dx = (x[1:] - x[:-1])[None, :]
    dy = (y[1:] - y[:-1])[:, None]
    dvx = (vx[:-1, 1:] - vx[:-1, :-1]) / dx
    dvy = (vy[1:, :-1] - vy[:-1, :-1]) / dy
    divv = dvx + dvy
    exx[1:, 1:] = dvx - divv / 3.0
    eyy[1:, 1:] = dvy - divv / 3.0
    return exx, eyy


def _cellular(w, A):
    """Velocity of the cellular flow vx = A sin(pi x) cos(pi y), vy = -A cos(pi x) sin(pi y)
    sampled at the staggered vx nodes (x[ix], yc[jy+1]) and vy nodes (xc[ix+1], y[jy])."""
    X = w.x[None, :] / w.LX
    Yc = w.yc[1:w.NY + 1][:, None] / w.LY
    vx = A * np.sin(np.pi * X) * np.cos(np.pi * Yc)
    Xc = w.xc[1:w.NX + 1][None, :] / w.LX
    Y = w.y[:, None] / w.LY
    vy = -A * np.cos(np.pi * Xc) * np.sin(np.pi * Y)
    return vx, vy


def convection_small(seed=1):
    """51x51, 4x4 markers/cell, T = 1-y+0.01 cos(pi x) sin(pi y), isoviscous."""
    NX = NY = 51
    LX = LY = 1.0
    x = spacing_uniform(NX, LX)
    y = spacing_uniform(NY, LY)
    xp, xcp = pad_grid(x, LX)
    yp, ycp = pad_grid(y, LY)
    rng = Rng(seed, 1)
    X, Y = _markers_in_domain(rng, x, y, NX, NY, 4, 4, LX, LY)
    mk = _new_markers(len(X))
    mk["X"], mk["Y"] = X, Y
    mk["Mat"] = 0
    mk["T"] = 1.0 - Y + 0.01 * np.cos(np.pi * X) * np.sin(np.pi * Y)
    mk["eta"] = 1.0
    mk["mu"] = 1e99
    mat = abi.default_materials()
    mat.materialEta[0] = 1.0
    mat.materialRho[0] = 1.0
    mat.materialAlpha[0] = 1e-3
    mat.materialkThermal[0] = 1.0
    mat.materialCp[0] = 1.0
    mat.materialMu[0] = 1e99
    opt = abi.default_options()
    opt.NX, opt.NY, opt.LX, opt.LY, opt.NMX, opt.NMY = NX, NY, LX, LY, 4, 4
    bv = _bc_free_slip()
    w = Workload("convection_small", NX, NY, LX, LY, xp, yp, xcp, ycp, mk, {}, mat, opt, bv,
                 0.0, "plain", seed=seed)
    vx, vy = _cellular(w, 1.0)
    wxy, _ = discrete_vorticity(vx, vy, w.xc, w.yc, bv)
    Xn, Yn = np.meshgrid(x, y)
    Tn = 1.0 - Yn + 0.01 * np.cos(np.pi * Xn) * np.sin(np.pi * Yn)
    w.nodal = {"vx": vx, "vy": vy, "wxy": wxy,
               "thisT": Tn + 1e-3 * np.sin(np.pi * Xn) * np.sin(np.pi * Yn)}
    w.dt = 0.5 * (LX / (NX - 1)) / 1.0
    w.projections = "all2"
    return w


def _erf(v):
    try:
        from scipy.special import erf
        return erf(v)
    except ImportError:  # pragma: no cover
        return np.vectorize(math.erf)(v)


def _bl_profile(y, delta=0.05):
    erf = _erf
    return 0.5 * erf(y / delta) + 0.5 * (1.0 + erf((y - 1.0) / delta))


def _tdep_projections():
    out = []
    fam = {(0, 0): "b", (-1, -1): "c", (0, 1): "vx", (1, 0): "vy"}
    for (sx, sy), tag in fam.items():
        out.append(("T", sx, sy, abi.ARITHMETIC, "T_" + tag))
        out.append(("eta", sx, sy, abi.ARITHMETIC, "eta_" + tag))
        out.append(("rho", sx, sy, abi.ARITHMETIC, "rho_" + tag))
        out.append(("mat:kThermal", sx, sy, abi.ARITHMETIC, "k_" + tag))
    out.append(("mat:Cp", 0, 0, abi.ARITHMETIC, "Cp_b"))
    return out


def convection_tdep(seed=2, N=201, NM=6, contrast=1e5, name="convection_tdep"):
    """N x N uniform grid, NM x NM markers/cell, eta = exp(-ln(contrast) T)."""
    NX = NY = N
    LX = LY = 1.0
    x = spacing_uniform(NX, LX)
    y = spacing_uniform(NY, LY)
    xp, xcp = pad_grid(x, LX)
    yp, ycp = pad_grid(y, LY)
    rng = Rng(seed, 1)
    X, Y = _markers_in_domain(rng, x, y, NX, NY, NM, NM, LX, LY)
    mk = _new_markers(len(X))
    mk["X"], mk["Y"] = X, Y
    mk["Mat"] = 0
    T = _bl_profile(Y) * (1.0 + 0.02 * Rng(seed, 2).normal(len(X)))
    mk["T"] = T
    qonr = math.log(contrast)
    mk["eta"] = np.exp(-qonr * T)
    mk["mu"] = 1e99
    mat = abi.default_materials()
    mat.materialEta[0] = 1.0
    mat.materialRho[0] = 1.0
    mat.materialAlpha[0] = 1e-3
    mat.materialkThermal[0] = 1.0
    mat.materialCp[0] = 1.0
    mat.materialMu[0] = 1e99
    mat.hasEtaT[0] = 5
    mat.QonR[0] = qonr
    opt = abi.default_options()
    opt.NX, opt.NY, opt.LX, opt.LY, opt.NMX, opt.NMY = NX, NY, LX, LY, NM, NM
    bv = _bc_free_slip()
    A = float(Rng(seed, 3).uniform(1, 0.5, 2.0)[0])
    w = Workload(name, NX, NY, LX, LY, xp, yp, xcp, ycp, mk, {}, mat, opt, bv, 0.0, "plain",
                 seed=seed)
    vx, vy = _cellular(w, A)
    wxy, _ = discrete_vorticity(vx, vy, w.xc, w.yc, bv)
    Xn, Yn = np.meshgrid(x, y)
    w.nodal = {"vx": vx, "vy": vy, "wxy": wxy,
               "thisT": _bl_profile(Yn) + 1e-3 * np.sin(np.pi * Xn) * np.sin(np.pi * Yn)}
    w.dt = 0.5 * (LX / (NX - 1)) / A
    w.projections = _tdep_projections()
    return w


def ensemble(n_instances=64, seed=4, nmin=65, nmax=257, first=0):
    """Independent convection_tdep instances with random sizes (nmin..nmax nodes a
    side), 16/25/36 markers per cell and viscosity contrast 1e2..1e6; instance i
    depends only on (seed, i), so ``first`` generates a slice of a larger set."""
    r = Rng(seed, 7)
    r.pos = 3 * first
    out = []
    for i in range(first, first + n_instances):
        u = r.uniform(3)
        N = int(nmin + min(int(u[0] * (nmax - nmin + 1)), nmax - nmin))
        NM = 4 + min(int(u[1] * 3), 2)
        contrast = 10.0 ** (2.0 + 4.0 * u[2])
        out.append(convection_tdep(seed=1000 * seed + i, N=N, NM=NM, contrast=contrast,
                                   name="ensemble_%d" % i))
    return out


def _streamfunction_velocity(w, rng, amp, nmodes=4):
    """Divergence-free staggered velocity from a smooth random streamfunction psi
    sampled at basic nodes: vx = dpsi/dy on vx nodes, vy = -dpsi/dx on vy nodes
    (discrete divergence zero on every cell, nodalStressStrain.c:67)."""
    x, y = w.x, w.y
    a = rng.normal(nmodes * nmodes).reshape(nmodes, nmodes)
    Xn, Yn = np.meshgrid(x / w.LX, y / w.LY)
    psi = np.zeros_like(Xn)
    for i in range(nmodes):
        for j in range(nmodes):
            psi += a[i, j] / (1.0 + i * i + j * j) * np.sin(np.pi * (i + 1) * Xn) * np.sin(np.pi * (j + 1) * Yn)
    psi *= amp
    NY, NX = psi.shape
    vx = np.zeros((NY, NX))
    vy = np.zeros((NY, NX))
    dy = (y[1:] - y[:-1])[:, None]
    dx = (x[1:] - x[:-1])[None, :]
    vx[:-1, :] = (psi[1:, :] - psi[:-1, :]) / dy
    vy[:, :-1] = -(psi[:, 1:] - psi[:, :-1]) / dx
    return vx, vy


def vep_subduction(seed=3, NX=801, NY=401, NM=5):
    """Layered 5-material slab model on a tanh-stretched grid with Drucker-Prager yield."""
    LX, LY = 2.0e5, 1.0e5
    x = spacing_tanh_band(NX, LX)
    y = spacing_tanh_band(NY, LY)
    xp, xcp = pad_grid(x, LX)
    yp, ycp = pad_grid(y, LY)
    rng = Rng(seed, 1)
    X, Y = _markers_in_cells(rng, x, y, NM, NM)
    n = len(X)
    mk = _new_markers(n)
    mk["X"], mk["Y"] = X, Y
    # layered materials with a dipping slab; interfaces perturbed by a random Fourier series
    fr = Rng(seed, 2).normal(12)
    pert = np.zeros(n)
    for k in range(6):
        pert += 1.5e3 * fr[k] / (1 + k) * np.sin((k + 1) * np.pi * X / LX + fr[6 + k])
    d = Y + pert
    matid = np.where(d < 1.0e4, 0, np.where(d < 3.0e4, 1, np.where(d < 6.0e4, 2, 3)))
    slab = (np.abs((Y - (X - 0.6e5) * 0.8)) < 1.2e4) & (X > 0.6e5)
    matid = np.where(slab, 4, matid)
    mk["Mat"] = matid
    T = 273.0 + 1300.0 * np.clip(Y / 1.0e5, 0, 1) * (1.0 + 0.01 * Rng(seed, 3).normal(n))
    T = np.where(slab, T - 400.0, T)
    mk["T"] = T
    mat = abi.default_materials()
    mat.nMaterials = 5
    etas = [1e23, 1e22, 1e21, 1e20, 1e24]
    for i in range(5):
        mat.materialEta[i] = etas[i]
        mat.materialRho[i] = 3300.0 + 20.0 * i
        mat.materialkThermal[i] = 3.0
        mat.materialCp[i] = 1000.0 + 50.0 * i
        mat.materialAlpha[i] = 3e-5
        mat.materialMu[i] = 1e10
        mat.materialCohesion[i] = 1e7 * (1 + i)
        mat.materialFriction[i] = 0.3 + 0.05 * i
        mat.hasPlasticity[i] = 1
        mat.hasEtaT[i] = 1 if i < 4 else 0
        mat.QonR[i] = 2.0e4
        mat.Tref[i] = 1573.0
    # strain-weakening law on material 2
    mat.hasPlasticity[2] = 2
    mat.C[2][0], mat.C[2][1] = 2e7, 5e6
    mat.F[2][0], mat.F[2][1] = 30.0, 10.0
    mat.gamma[2][0], mat.gamma[2][1] = 0.1, 1.0
    opt = abi.default_options()
    opt.NX, opt.NY, opt.LX, opt.LY, opt.NMX, opt.NMY = NX, NY, LX, LY, NM, NM
    opt.etamin, opt.etamax, opt.fractionalEtamin = 1e18, 1e25, 1e-4
    opt.subgridStressDiffusivity = 1.0
    opt.subgridTemperatureDiffusivity = 1.0
    r4 = Rng(seed, 4)
    eta0 = np.array(etas)[matid] * np.exp(0.1 * r4.normal(n))
    mk["eta"] = eta0
    mk["mu"] = 1e10
    mk["sxx"] = 1e7 * r4.normal(n)
    mk["syy"] = -mk["sxx"] + 1e6 * r4.normal(n)
    mk["sxy"] = 5e6 * r4.normal(n)
    mk["exx"] = 1e-15 * r4.normal(n)
    mk["eyy"] = -mk["exx"]
    mk["exy"] = 1e-15 * r4.normal(n)
    mk["p"] = 3300.0 * 9.8 * Y * (1.0 + 0.05 * r4.normal(n))
    mk["Eii"] = np.abs(r4.uniform(n, 0.0, 1.5))
    bv = _bc_free_slip()
    w = Workload("vep_subduction", NX, NY, LX, LY, xp, yp, xcp, ycp, mk, {}, mat, opt, bv, 0.0,
                 "vep", seed=seed)
    vel = 1e-9  # m/s
    vx, vy = _streamfunction_velocity(w, Rng(seed, 5), vel * LY)
    vmax = max(np.abs(vx).max(), np.abs(vy).max())
    hmin = min(np.diff(x).min(), np.diff(y).min())
    dt = 0.5 * hmin / vmax
    wxy, exy = discrete_vorticity(vx, vy, w.xc, w.yc, bv)
    exx, eyy = cell_strain_rates(vx, vy, x, y)
    # smooth viscosities on both node families and the nodal stress update of
    # nodalStressStrain.c:90-131, 256-271 with muN = muS = 1e99 (markerProjection.c:36-37)
    Xn, Yn = np.meshgrid(x, y)
    Xc, Yc = np.meshgrid(w.xc[0:NX], w.yc[0:NY])
    etaS = 1e21 * np.exp(2.0 * np.sin(np.pi * Xn / LX) * np.cos(np.pi * Yn / LY))
    etaN = 1e21 * np.exp(2.0 * np.sin(np.pi * Xc / LX) * np.cos(np.pi * Yc / LY))
    mu = 1e99
    r6 = Rng(seed, 6)
    so_old = {k: 1e7 * r6.normal(NX * NY).reshape(NY, NX) for k in ("soxx", "soyy", "soxy")}
    nod = {"vx": vx, "vy": vy, "wxy": wxy, "edotxy": exy, "edotxx": exx, "edotyy": eyy}
    XN = etaN / (etaN + dt * mu)
    X1N = etaN * (-2.0 * XN + 2.0)
    XS = etaS / (etaS + dt * mu)
    X1S = etaS * (-2.0 * XS + 2.0)
    for comp, X_, X1_, e_ in (("xx", XN, X1N, exx), ("yy", XN, X1N, eyy), ("xy", XS, X1S, exy)):
        snew = X_ * so_old["so" + comp] + X1_ * e_
        nod["ds" + comp] = snew - so_old["so" + comp]
        nod["so" + comp] = snew
    nod["p"] = 3300.0 * 9.8 * Yc * (1.0 + 0.01 * r6.normal(NX * NY).reshape(NY, NX))
    Tn = 273.0 + 1300.0 * np.clip(Yn / 1.0e5, 0, 1)
    nod["thisT"] = Tn + 5.0 * np.sin(np.pi * Xn / LX) * np.sin(np.pi * Yn / LY)
    w.nodal = nod
    w.dt = dt
    w.projections = "all2"
    return w


def _periodic_fourier(rng, X, Y, LX, LY, nmodes=3):
    """A smooth field, periodic in x over LX: sum of a_ij cos(2 pi i x / LX + p_ij)
    sin/cos(pi j y / LY) with seeded coefficients."""
    a = rng.normal(2 * nmodes * nmodes)
    v = np.zeros(np.broadcast(X, Y).shape)
    k = 0
    for i in range(nmodes):
        for j in range(nmodes):
            v = v + a[k] / (1.0 + i * i + j * j) * np.cos(2.0 * np.pi * i * X / LX + a[k + 1]) * \
                np.cos(np.pi * j * Y / LY + 0.3 * j)
            k += 2
    return v


def periodic_box(kind="plain", NX=64, NY=41, NM=4, seed=8, graded_y=True):
    """A box periodic in x (the reference's grid for mechBCLeft.type[0] == 2,
    gridGenerator.c:66-85): NX uniform cells over LX (x[NX] = LX closes the
    last one), a tanh-stretched or uniform y axis, NM x NM jittered markers per
    cell plus markers exactly on x = 0 and x = LX and a few outside the box.
    The velocity has a mean drift along x of about half a cell per step, so
    markers cross x = LX and x = 0 and are wrapped.

    kind "plain": isoviscous thermal convection inputs (the convection_small
    fields); kind "vep": five materials with yielding and every nodal input of
    the marker stress and strain updates (vep_subduction's fields, periodic)."""
    LX, LY = (2.0, 1.0) if kind == "plain" else (2.0e5, 1.0e5)
    xf = spacing_uniform(NX + 1, LX)
    xp, xcp = pad_periodic(xf, LX)
    y = spacing_tanh_band(NY, LY) if graded_y else spacing_uniform(NY, LY)
    yp, ycp = pad_grid(y, LY)
    rng = Rng(seed, 1)
    X, Y = _markers_in_cells(rng, xf, y, NM, NM)
    n = len(X)
    edges = [(0.0, 0.3), (LX, 0.5), (LX, 0.0), (0.0, 1.0), (LX * (1 - 1e-16), 0.7),
             (xf[NX - 1], 0.2), (xcp[NX + 2], 0.6), (xcp[3], 0.4), (-1e-9 * LX, 0.5), (LX * (1 + 1e-12), 0.5)]
    for k, (ex, ey) in enumerate(edges):
        X[(37 * k + 5) % n], Y[(37 * k + 5) % n] = ex, ey * LY
    mk = _new_markers(n)
    mk["X"], mk["Y"] = X, Y
    r2 = Rng(seed, 2)
    bv = _bc_free_slip()
    for side in ("mechBCLeft", "mechBCRight"):
        getattr(bv, side).type[0] = 2  # periodic (gridGenerator.c:67)
    opt = abi.default_options()
    opt.NX, opt.NY, opt.LX, opt.LY, opt.NMX, opt.NMY = NX, NY, LX, LY, NM, NM
    opt.mechBCLeft.type[0] = 2
    opt.mechBCRight.type[0] = 2
    Xn, Yn = np.meshgrid(xf[:NX], y)
    hx = LX / NX
    if kind == "plain":
        mk["Mat"] = 0
        mk["T"] = 1.0 - Y / LY + 0.05 * np.cos(2 * np.pi * X / LX) * np.sin(np.pi * Y / LY)
        mk["eta"] = 1.0
        mk["mu"] = 1e99
        mat = abi.default_materials()
        mat.materialEta[0] = mat.materialRho[0] = mat.materialkThermal[0] = mat.materialCp[0] = 1.0
        mat.materialAlpha[0] = 1e-3
        mat.materialMu[0] = 1e99
        U = 1.0
    else:
        matid = np.clip((Y / LY * 4).astype(int), 0, 3)
        matid = np.where(np.abs(Y - 0.5 * LY - 0.1 * LY * np.sin(2 * np.pi * X / LX)) < 0.08 * LY, 4, matid)
        mk["Mat"] = matid
        mk["T"] = 273.0 + 1300.0 * Y / LY * (1.0 + 0.01 * r2.normal(n))
        mat = abi.default_materials()
        mat.nMaterials = 5
        etas = [1e23, 1e22, 1e21, 1e20, 1e24]
        for i in range(5):
            mat.materialEta[i] = etas[i]
            mat.materialRho[i] = 3300.0 + 20.0 * i
            mat.materialkThermal[i] = 3.0
            mat.materialCp[i] = 1000.0 + 50.0 * i
            mat.materialAlpha[i] = 3e-5
            mat.materialMu[i] = 1e10
            mat.materialCohesion[i] = 1e7 * (1 + i)
            mat.materialFriction[i] = 0.3 + 0.05 * i
            mat.hasPlasticity[i] = 1
            mat.hasEtaT[i] = 1 if i < 4 else 0
            mat.QonR[i] = 2.0e4
            mat.Tref[i] = 1573.0
        opt.etamin, opt.etamax, opt.fractionalEtamin = 1e18, 1e25, 1e-4
        opt.subgridStressDiffusivity = 1.0
        opt.subgridTemperatureDiffusivity = 1.0
        r4 = Rng(seed, 4)
        mk["eta"] = np.array(etas)[matid] * np.exp(0.1 * r4.normal(n))
        mk["mu"] = 1e10
        mk["sxx"] = 1e7 * r4.normal(n)
        mk["syy"] = -mk["sxx"] + 1e6 * r4.normal(n)
        mk["sxy"] = 5e6 * r4.normal(n)
        mk["exx"] = 1e-15 * r4.normal(n)
        mk["eyy"] = -mk["exx"]
        mk["exy"] = 1e-15 * r4.normal(n)
        mk["p"] = 3300.0 * 9.8 * Y * (1.0 + 0.05 * r4.normal(n))
        mk["Eii"] = np.abs(r4.uniform(n, 0.0, 1.5))
        U = 1e-9
    w = Workload("periodic_" + kind, NX, NY, LX, LY, xp, yp, xcp, ycp, mk, {}, mat, opt, bv, 0.0,
                 "plain" if kind == "plain" else "vep", seed=seed, xperiodic=1)
    # staggered velocity: mean drift U along x plus a periodic perturbation
    r5 = Rng(seed, 5)
    Xv, Yv = np.meshgrid(xf[:NX], w.yc[1:NY + 1])          # vx nodes (x[ix], yc[jy+1])
    Xw, Yw = np.meshgrid(w.xc[1:NX + 1], y)                # vy nodes (xc[ix+1], y[jy])
    vx = U * (1.0 + 0.5 * _periodic_fourier(r5, Xv, Yv, LX, LY))
    vy = U * 0.5 * _periodic_fourier(r5, Xw, Yw, LX, LY)
    vx[NY - 1, :] = 0.0  # ghost row of the vx nodes
    vy[:, :] = np.where(np.arange(NY)[:, None] == NY - 1, 0.0, vy)
    w.dt = 0.5 * hx / (U * 2.0)
    r6 = Rng(seed, 6)
    nod = {"vx": vx, "vy": vy, "wxy": U / LY * _periodic_fourier(r6, Xn, Yn, LX, LY)}
    if kind == "plain":
        nod["thisT"] = 1.0 - Yn / LY + 0.05 * np.cos(2 * np.pi * Xn / LX) * np.sin(np.pi * Yn / LY) + \
            1e-3 * _periodic_fourier(r6, Xn, Yn, LX, LY)
    else:
        Xc, Yc = np.meshgrid(w.xc[0:NX], w.yc[0:NY])
        for name, scale, (XX, YY) in (("edotxx", 1e-15, (Xc, Yc)), ("edotyy", 1e-15, (Xc, Yc)),
                                      ("edotxy", 1e-15, (Xn, Yn)), ("soxx", 1e7, (Xc, Yc)),
                                      ("soyy", 1e7, (Xc, Yc)), ("soxy", 1e7, (Xn, Yn)),
                                      ("dsxx", 1e6, (Xc, Yc)), ("dsyy", 1e6, (Xc, Yc)),
                                      ("dsxy", 1e6, (Xn, Yn))):
            nod[name] = scale * _periodic_fourier(r6, XX, YY, LX, LY)
        nod["p"] = 3300.0 * 9.8 * Yc * (1.0 + 0.01 * _periodic_fourier(r6, Xc, Yc, LX, LY))
        nod["thisT"] = 273.0 + 1300.0 * Yn / LY + 5.0 * _periodic_fourier(r6, Xn, Yn, LX, LY)
    w.nodal = nod
    w.projections = "all2"
    return w


WORKLOADS = {
    "convection_small": convection_small,
    "convection_tdep": convection_tdep,
    "vep_subduction": vep_subduction,
    "periodic_box": periodic_box,
}


def nodal_box(NX=801, NY=401, LX=2.0, LY=1.0, seed=5, graded=True):
    """Nodal fields for the step between the Stokes solve and the marker updates
    (retrieveSolutions, nodalStressStrain, limitDisplacementTimestep): a
    tanh-stretched (or uniform) grid, smooth seeded velocities of either sign,
    viscosities 1e19..1e23, shear moduli 1e10..3e10 (mu = 0 on a few nodes),
    stresses of either sign, densities and density rates.  No markers."""
    rng = Rng(seed, 7)
    x = spacing_tanh_band(NX, LX) if graded else spacing_uniform(NX, LX)
    y = spacing_tanh_band(NY, LY) if graded else spacing_uniform(NY, LY)
    xp, xcp = pad_grid(x, LX)
    yp, ycp = pad_grid(y, LY)
    X, Y = np.meshgrid(x / LX, y / LY)

    def u(lo, hi):
        return lo + (hi - lo) * rng.uniform(NY * NX).reshape(NY, NX)

    a, b = u(0.5, 2.0)[0, 0], u(0.5, 2.0)[0, 0]
    nodal = {
        "vx": 1e-9 * (a * np.sin(np.pi * X) * np.cos(2 * np.pi * Y) + 0.1 * u(-1.0, 1.0)),
        "vy": -1e-9 * (b * np.cos(np.pi * X) * np.sin(2 * np.pi * Y) + 0.1 * u(-1.0, 1.0)),
        "etaN": 10.0 ** u(19.0, 23.0), "etaS": 10.0 ** u(19.0, 23.0),
        "muN": u(1e10, 3e10), "muS": u(1e10, 3e10),
        "soxx": u(-1e7, 1e7), "soyy": u(-1e7, 1e7), "soxy": u(-1e7, 1e7),
        "rho": u(3000.0, 3300.0), "rhodot": u(-1e-12, 1e-12),
    }
    nodal["muN"][::7, ::5] = 0.0
    nodal["muS"][::5, ::7] = 0.0
    if NY > 3 and NX > 4:
        nodal["vx"][3, 4] = 0.0
    if NY > 2 and NX > 2:
        nodal["rhodot"][2, 2] = 0.0
    mk = _new_markers(0)
    opt = abi.default_options()
    opt.displacementStepLimit = 0.1
    return Workload(name="nodal_box", NX=NX, NY=NY, LX=LX, LY=LY, xp=xp, yp=yp, xcp=xcp, ycp=ycp, markers=mk,
                    nodal=nodal, materials=abi.Materials(), options=opt, bv=_bc_free_slip(), dt=3.15e10,
                    kind="plain", projections=[], seed=seed)


def thermal_box(NX=801, NY=401, LX=660e3, LY=600e3, seed=6, graded=True, top_type=0):
    """Nodal fields for the energy equation of the subduction set-up
    (formThermalSystem): conductivity 2..4, density anomaly, heat capacity,
    a lastT between 273 and 1573 K, vx of either sign near the right wall,
    a subduction parameter set (30 degree slab of 50 Myr, 50 km plate with a
    10 km root at 400 km), and heating terms for the right-hand side."""
    rng = Rng(seed, 9)
    x = spacing_tanh_band(NX, LX) if graded else spacing_uniform(NX, LX)
    y = spacing_tanh_band(NY, LY) if graded else spacing_uniform(NY, LY)
    xp, xcp = pad_grid(x, LX)
    yp, ycp = pad_grid(y, LY)
    X, Y = np.meshgrid(x / LX, y / LY)

    def u(lo, hi):
        return lo + (hi - lo) * rng.uniform(NY * NX).reshape(NY, NX)

    nodal = {
        "kThermal": u(2.0, 4.0), "rho": u(-50.0, 50.0), "Cp": u(1000.0, 1250.0),
        "lastT": 273.0 + 1300.0 * np.clip(Y * 3.0, 0.0, 1.0) + u(-5.0, 5.0),
        "vx": 1e-9 * np.sin(2 * np.pi * Y * 3.0) + 1e-10 * u(-1.0, 1.0),
        "heating": u(0.0, 1e-6),
    }
    opt = abi.default_options()
    opt.thermalBCTop.type[0] = top_type
    opt.thermalBCTop.value[0] = 273.0
    opt.slabAngle = math.pi / 6.0
    opt.slabAge = 50.0
    opt.overridingPlateThickness = 50e3
    opt.rootThickness = 10e3
    opt.rootCenter = 400e3
    opt.rootWidth = 40e3
    mat = abi.Materials()
    mat.materialRho[0] = 3300.0
    return Workload(name="thermal_box", NX=NX, NY=NY, LX=LX, LY=LY, xp=xp, yp=yp, xcp=xcp, ycp=ycp,
                    markers=_new_markers(0), nodal=nodal, materials=mat, options=opt, bv=_bc_free_slip(),
                    dt=3.15e11, kind="plain", projections=[], seed=seed)
