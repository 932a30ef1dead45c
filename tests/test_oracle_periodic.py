"""Pins of the oracle's periodic-x restatement (grid->xperiodic; the PETSc DMDA
is DM_BOUNDARY_PERIODIC along x, nodalFields.c:11-20) by facts the periodic
geometry fixes, independent of the wrap code under test:

* findMarkerCell (markers.c:41-83): a brute-force search over the NX cells
  [x[i], x[i+1]], the last one closed by x[NX] = LX;
* projections, the interpolations back to the markers, subgrid temperature
  diffusion and the marker stress / strain-rate updates equal the oracle's
  NON-periodic functions (pinned by their own tests) on an extended grid with
  two extra cells on each side, the markers near x = 0 and x = LX replicated
  one period over and the nodal fields extended periodically -- node i of the
  periodic grid is node i + 2 of the extended one.  Coordinates and marker
  positions are dyadic, so both sides compute identical weights and differ only
  in how the wrapped nodes' sums associate.  Where the periodic code clamps
  (a cell-centre interpolation left of xc[1] or right of xc[NX]) it
  extrapolates from the two centre columns it clamps to, checked against that
  line;
* RK4 advection (markers.c:1865-2250) in a velocity field that alternates in
  sign from node column to node column: linear interpolation in cell i at
  fraction t gives U + A (-1)^i (1 - 2t) in closed form (also beyond the ends,
  where the walk stops at cell -1 or NX-1 and extrapolates), the new position
  is wrapped into [0, LX] (markers.c:2225-2231), and on a periodic grid the
  left-wall ghost rule for vy is off.
"""
import numpy as np
import pytest

import oracle
from fvmic2d_amd import abi, workloads
from helpers import assert_bitwise, assert_close

pytestmark = pytest.mark.usefixtures("oracle_lib")

Q = 2.0 ** 40  # marker positions on a 2^-40 lattice: shifts by multiples of h are exact


def _quantize(w):
    m = w.markers
    for f in ("X", "Y"):
        m[f] = np.round(m[f] * Q) / Q
    return w


def _periodic(kind="plain", NX=32, NY=17, NM=3, seed=8):
    return _quantize(workloads.periodic_box(kind, NX=NX, NY=NY, NM=NM, seed=seed, graded_y=False))


def _extended(w):
    """The non-periodic twin: NX + 5 uniform nodes over [0, LX + 4h] (x' = x + 2h),
    the markers shifted by 2h (originals first, in order) plus copies one period
    over of those within 2h of x = 0 or x = LX, and every nodal field extended
    periodically (column j' <- column (j' - 2) mod NX)."""
    NX, h = w.NX, w.LX / w.NX
    LXe = w.LX + 4 * h
    xe = workloads.spacing_uniform(NX + 5, LXe)
    xp, xcp = workloads.pad_grid(xe, LXe)
    m = w.markers
    act = (m["X"] >= 0) & (m["X"] <= w.LX)
    lo = np.flatnonzero(act & (m["X"] < 2 * h))
    hi = np.flatnonzero(act & (m["X"] > w.LX - 2 * h))
    mk = np.concatenate([m, m[lo], m[hi]])
    mk["X"][len(m):len(m) + len(lo)] += w.LX
    mk["X"][len(m) + len(lo):] -= w.LX
    mk["X"] += 2 * h
    mk["X"][:len(m)][~act] = -1.0  # outside stays outside
    cols = (np.arange(NX + 5) - 2) % NX
    nodal = {k: np.ascontiguousarray(v[:, cols]) for k, v in w.nodal.items()}
    e = workloads.Workload(w.name + "_ext", NX + 5, w.NY, LXe, w.LY, xp, w.yp, xcp, w.ycp, mk, nodal,
                           w.materials, w.options, w.bv, w.dt, w.kind, w.projections, w.seed, 0)
    return e, len(m)


def _pair(w, extra=()):
    e, n = _extended(w)
    vep = oracle.VEP_PROJ if w.kind == "vep" else []
    sp = oracle.State(w, extra=list(extra) + vep)
    se = oracle.State(e, extra=list(extra) + vep)
    sp.find_marker_cells()
    se.find_marker_cells()
    return sp, se, n


def _inner(a, NX):
    """Nodes of the extended grid that stand for the periodic nodes 0..NX-1."""
    return a[:, 2:NX + 2]


def test_find_marker_cell_periodic_brute_force():
    w = _periodic(NX=16, NY=9, NM=3)
    m = w.markers
    r = workloads.Rng(3, 1)
    k = np.arange(0, w.n_markers, 5)
    m["X"][k] = np.round(r.uniform(len(k), -0.05, 1.05) * w.LX * Q) / Q
    m["X"][k[:16]] = w.xp[2:18]  # on every node
    m["X"][k[16]] = w.LX
    st = oracle.State(w)
    nout = st.find_marker_cells()
    xn = w.xp[2:2 + w.NX + 1]  # x[0..NX], x[NX] = LX
    X = m["X"]
    inside = (X >= 0) & (X <= w.LX) & (m["Y"] >= 0) & (m["Y"] <= w.LY)
    want = np.array([max(i for i in range(w.NX) if xn[i] <= v) if ok else -1 for v, ok in zip(X, inside)])
    assert_bitwise(st.markers["cellX"], want, "cellX")
    assert nout == int((~inside).sum())
    assert st.markers["cellX"][k[16]] == w.NX - 1


@pytest.mark.parametrize("method", [abi.ARITHMETIC, abi.ARITHMETIC_LOCAL, abi.HARMONIC])
def test_projection_equals_extended_grid(method):
    w = _periodic()
    sp, se, _ = _pair(w, extra=["out"])
    for src in ("T", "sxy", "mat:kThermal"):
        for sx, sy in [(0, 0), (-1, -1), (0, 1), (1, 0), (1, 1), (-1, 1), (1, -1), (-1, 0), (0, -1)]:
            sp.project_marker_field(src, "out", sx, sy, method)
            se.project_marker_field(src, "out", sx, sy, method)
            assert_close(sp.nodal["out"], _inner(se.nodal["out"], w.NX),
                         "%s (%d,%d) m%d" % (src, sx, sy, method), rtol=1e-14)


def test_scatter_wraps_both_sides():
    """Each ghost column adds into the node it wraps to: one marker a quarter
    cell left of x = LX (value 3) and one a quarter cell right of x = 0
    (value 5), projected onto the basic nodes and the two x-staggered families
    (weights by hand, h = LX / 8)."""
    w = _periodic(NX=8, NY=5, NM=2)
    m = w.markers
    h = w.LX / w.NX
    m["X"][:] = -1.0  # all outside but two
    m["X"][0], m["Y"][0] = w.LX - 0.25 * h, 0.5 * w.LY
    m["X"][1], m["Y"][1] = 0.25 * h, 0.5 * w.LY
    m["T"][0], m["T"][1] = 3.0, 5.0
    st = oracle.State(w, extra=["out"])
    st.find_marker_cells()
    assert list(st.markers["cellX"][:2]) == [w.NX - 1, 0]
    j = int(np.searchsorted(w.y, 0.5 * w.LY, side="right") - 1)  # Y on node row j
    # stagx 0: marker 0 in cell NX-1 at mdx 3/4 (nodes NX-1: 1/4, NX = 0: 3/4),
    # marker 1 in cell 0 at mdx 1/4 (nodes 0: 3/4, 1: 1/4)
    st.project_marker_field("T", "out", 0, 0, abi.ARITHMETIC)
    row = st.nodal["out"][j]
    assert row[0] == (3.0 * 0.75 + 5.0 * 0.75) / 1.5
    assert row[w.NX - 1] == 3.0 and row[1] == 5.0
    # stagx -1: marker 0 (X >= xc[NX]) in lattice cell NX at mdx 1/4 (nodes
    # NX = 0: 3/4, NX+1 = 1: 1/4); marker 1 (X < xc[1]) in cell 0 at mdx 3/4
    st.project_marker_field("T", "out", -1, 0, abi.ARITHMETIC)
    row = st.nodal["out"][j]
    assert row[0] == 3.0 * 0.75 + 5.0 * 0.25 and row[1] == 3.0 * 0.25 + 5.0 * 0.75
    # stagx +1: marker 1 (X < xc[1]) in lattice cell -1 at mdx 3/4 (nodes -1 =
    # NX-1: 1/4, 0: 3/4); marker 0 in cell NX-1 at mdx 1/4 (NX-1: 3/4, NX = 0: 1/4)
    st.project_marker_field("T", "out", 1, 0, abi.ARITHMETIC)
    row = st.nodal["out"][j]
    assert row[0] == 5.0 * 0.75 + 3.0 * 0.25 and row[w.NX - 1] == 5.0 * 0.25 + 3.0 * 0.75
    assert np.isnan(row[2:w.NX - 1]).all()  # no weight: 0/0


def test_interpolations_equal_extended_grid():
    w = _periodic()
    sp, se, n = _pair(w)
    for s in (sp, se):
        s.nodal["lastT"][:] = s.nodal["thisT"]
        s.project_nodal_to_markers("lastT", "T")
    assert_bitwise(sp.markers["T"], se.markers["T"][:n], "projectNodalFieldToMarkersS")


def test_subgrid_temperature_equals_extended_grid():
    w = _periodic()
    sp, se, n = _pair(w)
    for s in (sp, se):
        s.update_marker_density()
        s.project_markers_nodes_all2()
        s.subgrid_temperature(w.dt, 1.0)
    act = sp.markers["cellX"] != -1
    for f in ("T", "Tdot"):
        assert_close(sp.markers[f][act], se.markers[f][:n][act], f, rtol=1e-13,
                     scale=np.abs(se.markers["T"][:n][act]).max())


def test_stress_updates_equal_extended_grid():
    """updateMarkerStress and updateMarkerStrainPressure: equal to the extended
    grid where no clamp acts; where the cell-centre cell is clamped (left of
    xc[1]: cell 1, right of xc[NX]: cell NX - 1) the value lies on the line
    through the two centre columns clamped to."""
    w = _periodic("vep", NX=24, NY=13, NM=3)
    sp, se, n = _pair(w)
    h = w.LX / w.NX
    X = sp.markers["X"]
    act = sp.markers["cellX"] != -1
    free = act & (X >= 0.5 * h) & (X < w.LX - 0.5 * h)
    clamped = act & ~free
    assert free.sum() > 100 and clamped.sum() > 20
    for s in (sp, se):
        s.update_marker_strain_pressure(w.dt)
        s.update_marker_stress()
    for f in ("p", "exx", "eyy", "exy", "wxy", "sxx", "syy", "sxy", "strainRateResidual"):
        sc = np.abs(se.markers[f][:n][free]).max()
        assert_close(sp.markers[f][free], se.markers[f][:n][free], f, rtol=1e-13, scale=sc)
    # clamped markers: pressure from the line through centre columns (1, 2) or
    # (NX-1, NX = 0), bilinear in y
    p = w.nodal["p"]
    yc = w.yc
    for k in np.flatnonzero(clamped)[:40]:
        Xk, Yk = X[k], sp.markers["Y"][k]
        c = 1 if Xk < 0.5 * h else w.NX - 1
        c1 = (c + 1) % w.NX
        xcl, xcr = w.xcp[2 + c], w.xcp[2 + c + 1]
        t = (Xk - xcl) / (xcr - xcl)
        cy = int(sp.markers["cellY"][k])
        if Yk >= yc[cy + 1] and cy < w.NY - 1:
            cy += 1
        cy = min(max(cy, 1), w.NY - 2)
        u = (Yk - yc[cy]) / (yc[cy + 1] - yc[cy])
        want = (1 - t) * (1 - u) * p[cy, c] + (1 - t) * u * p[cy + 1, c] + t * (1 - u) * p[cy, c1] + \
            t * u * p[cy + 1, c1]
        assert sp.markers["p"][k] == pytest.approx(want, rel=1e-13, abs=1e-13 * np.abs(p).max())


@pytest.mark.parametrize("jaumann", [0, 1])
def test_advect_alternating_field(jaumann):
    w = _periodic(NX=16, NY=9, NM=3)
    NX, NY, h = w.NX, w.NY, w.LX / w.NX
    U, A, B, Wv = 0.7, 0.3, 0.2, 0.5
    sgn = (-1.0) ** np.arange(NX)
    w.nodal["vx"] = np.ascontiguousarray(np.tile(U + A * sgn, (NY, 1)))
    w.nodal["vy"] = np.ascontiguousarray(np.tile(B * sgn, (NY, 1)))
    w.nodal["wxy"] = np.ascontiguousarray(np.tile(Wv * sgn, (NY, 1)))
    # vy(x) stays small, and Y is kept away from the walls so that no stage leaves
    # the rows; the left wall is a kinematic one (type 0, value 9): a periodic
    # grid must not use it
    w.bv.mechBCLeft.type[1], w.bv.mechBCLeft.value[1] = 0, 9.0
    m = w.markers
    m["Y"] = np.clip(m["Y"], 0.2 * w.LY, 0.8 * w.LY)
    dt = 1.3 * h / (U + A)
    st = oracle.State(w)
    st.find_marker_cells()
    X0, Y0 = m["X"].copy(), m["Y"].copy()
    st.advect(dt, jaumann)
    act = st.markers["cellX"] != -1

    def alt(x, base, amp, off):
        # node columns at off + i h; the walk of markers.c:1930-1936 from cell0
        i = np.floor((x - off) / h).astype(int)
        i = np.clip(i, -1, NX - 1)
        t = (x - off - i * h) / h
        return base + amp * np.where(i % 2 == 0, 1.0, -1.0) * (1.0 - 2.0 * t)

    def stage(x):
        return alt(x, U, A, 0.0)

    k1 = stage(X0)
    k2 = stage(X0 + dt / 2 * k1)
    k3 = stage(X0 + dt / 2 * k2)
    k4 = stage(X0 + dt * k3)
    VX = (k1 + 2 * k2 + 2 * k3 + k4) / 6
    Xn = X0 + dt * VX
    Xn = np.where(Xn > w.LX, Xn - w.LX, np.where(Xn < 0, Xn + w.LX, Xn))
    r = st.markers
    assert_close(r["VX"][act], VX[act], "VX", rtol=1e-13, scale=U + A)
    assert_close(r["X"][act], Xn[act], "X", rtol=1e-13, scale=w.LX)
    assert ((X0 > w.LX - h) & (r["X"] < h) & act).sum() > 3  # wrapped at x = LX
    assert (r["X"][act] >= 0).all() and (r["X"][act] <= w.LX).all()
    # vy nodes sit at the cell centres xc[i + 1] = (i + 1/2) h: the stage vy from
    # the same alternating line, with no left-wall ghost (the kinematic value 9
    # would pull it far off)
    def vy_stage(x):
        return alt(x, 0.0, B, 0.5 * h)
    xs = [X0, X0 + dt / 2 * k1, X0 + dt / 2 * k2, X0 + dt * k3]
    VY = (vy_stage(xs[0]) + 2 * vy_stage(xs[1]) + 2 * vy_stage(xs[2]) + vy_stage(xs[3])) / 6
    assert_close(r["VY"][act], VY[act], "VY", rtol=1e-13, scale=B)
    wx = (alt(xs[0], 0.0, Wv, 0.0) + 2 * alt(xs[1], 0.0, Wv, 0.0) +
          2 * alt(xs[2], 0.0, Wv, 0.0) + alt(xs[3], 0.0, Wv, 0.0)) / 6
    assert_close(r["wxy"][act], wx[act], "wxy", rtol=1e-13, scale=Wv)
