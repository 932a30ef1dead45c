"""Pinning the CPU oracle (oracle/oracle.c) without the reference build.

The reference cannot be compiled here (PETSc 3.5 is absent) and ships no tests or
golden vectors for this path, so each oracle function is pinned by what the
mathematics fixes: brute force on tiny inputs, closed forms, invariants and
hand-derived known answers.  Each check is chosen so that a plausible slip --
a dropped term, a swapped index, a wrong sign, a transposed operand -- fails it.
"""
import copy

import numpy as np
import pytest

import oracle
from fvmic2d_amd import abi, workloads

pytestmark = pytest.mark.usefixtures("oracle_lib")


def tiny(NX=4, NY=4, LX=3.0, LY=3.0, x=None, y=None, markers=()):
    """A small hand-made workload: grid, given markers (X, Y, **fields)."""
    x = np.asarray(x if x is not None else np.linspace(0, LX, NX))
    y = np.asarray(y if y is not None else np.linspace(0, LY, NY))
    NX, NY, LX, LY = len(x), len(y), x[-1], y[-1]
    xp, xcp = workloads.pad_grid(x, LX)
    yp, ycp = workloads.pad_grid(y, LY)
    mk = np.zeros(len(markers), dtype=abi.MARKER_DTYPE)
    mk["cellX"] = mk["cellY"] = -1
    mk["mu"] = 1e99
    mk["eta"] = 1.0
    for i, m in enumerate(markers):
        for k, v in m.items():
            mk[k][i] = v
    mat = abi.default_materials()
    w = workloads.Workload("tiny", NX, NY, LX, LY, xp, yp, xcp, ycp, mk, {}, mat,
                           abi.default_options(), workloads._bc_free_slip(), 0.1, "plain")
    return w


def state(w, extra=("out",)):
    st = oracle.State(w, extra=list(extra))
    st.find_marker_cells()
    return st


# ---- findMarkerCells (markers.c:41-106) -----------------------------------
def test_find_cells_matches_brute_force():
    rng = np.random.default_rng(0)
    for trial in range(20):
        NX, NY = rng.integers(2, 40, size=2)
        x = np.sort(np.concatenate([[0.0, 7.0], rng.uniform(0, 7.0, NX - 2)]))
        y = np.sort(np.concatenate([[0.0, 3.0], rng.uniform(0, 3.0, NY - 2)]))
        pts = [dict(X=v, Y=u) for v, u in zip(rng.uniform(-0.5, 7.5, 300), rng.uniform(-0.2, 3.2, 300))]
        pts += [dict(X=xi, Y=yj) for xi in x for yj in y[:3]]  # exactly on grid lines
        st = state(tiny(x=x, y=y, markers=pts))
        for m in st.markers:
            X, Y = m["X"], m["Y"]
            if 0 <= X <= 7.0 and 0 <= Y <= 3.0:
                # brute force: largest i <= N-2 with x[i] <= X
                bx = max(i for i in range(len(x) - 1) if x[i] <= X)
                by = max(j for j in range(len(y) - 1) if y[j] <= Y)
                assert (m["cellX"], m["cellY"]) == (bx, by), (X, Y)
            else:
                assert (m["cellX"], m["cellY"]) == (-1, -1)


def test_find_cells_special_values():
    w = tiny(markers=[dict(X=3.0, Y=3.0), dict(X=0.0, Y=0.0), dict(X=np.nan, Y=1.0),
                      dict(X=1.0, Y=np.inf), dict(X=-0.0, Y=2.0), dict(X=1.0, Y=1.0)])
    st = state(w)
    assert st.markers["cellX"].tolist() == [2, 0, -1, -1, 0, 1]
    assert st.markers["cellY"].tolist() == [2, 0, -1, -1, 2, 1]
    assert st.ms.nOut == 2


# ---- projectMarkerFieldToNodes (markerProjection.c:71-269) -----------------
FAMS = [(0, 0), (-1, -1), (0, 1), (1, 0)]


def test_projection_hand_computed_arithmetic():
    # one marker at 1/4, 1/2 of basic cell (1,1) on a unit-spaced 4x4 grid
    st = state(tiny(markers=[dict(X=1.25, Y=1.5, T=2.0)]))
    st.project_marker_field("T", "out", 0, 0, abi.ARITHMETIC)
    out = st.nodal["out"]
    touched = [(1, 1), (1, 2), (2, 1), (2, 2)]
    for (j, i) in touched:
        assert out[j, i] == 2.0
    mask = np.ones_like(out, bool)
    for (j, i) in touched:
        mask[j, i] = False
    assert np.isnan(out[mask]).all()
    # two markers: node (1,1) gets weights 0.75*0.5 and 0.25*0.25
    st = state(tiny(markers=[dict(X=1.25, Y=1.5, T=2.0), dict(X=1.75, Y=1.75, T=6.0)]))
    st.project_marker_field("T", "out", 0, 0, abi.ARITHMETIC)
    w1, w2 = 0.75 * 0.5, 0.25 * 0.25
    assert st.nodal["out"][1, 1] == pytest.approx((w1 * 2 + w2 * 6) / (w1 + w2), rel=1e-15)
    w1, w2 = 0.25 * 0.5, 0.75 * 0.75  # node (jy=2, ix=2)
    assert st.nodal["out"][2, 2] == pytest.approx((w1 * 2 + w2 * 6) / (w1 + w2), rel=1e-15)


def test_projection_staggered_cells_and_ghost_drop():
    # unit grid: centres at -0.5, 0.5, 1.5, 2.5, 3.5
    st = state(tiny(markers=[dict(X=0.25, Y=0.75, T=5.0)]))
    # (-1,-1): X >= xc[1]=0.5? no -> lattice x 0; Y >= yc[1]? yes -> lattice y 1
    st.project_marker_field("T", "out", -1, -1, abi.ARITHMETIC)
    fin = np.argwhere(np.isfinite(st.nodal["out"])).tolist()
    assert fin == [[1, 0], [1, 1], [2, 0], [2, 1]]
    # (1,0): X < xc[1] -> lattice x -1 -> nodes x -1 (dropped ghost) and 0
    st.project_marker_field("T", "out", 1, 0, abi.ARITHMETIC)
    fin = np.argwhere(np.isfinite(st.nodal["out"])).tolist()
    assert fin == [[0, 0], [1, 0]]
    # (0,1): Y >= yc[1] -> lattice y 0 -> rows 0, 1
    st.project_marker_field("T", "out", 0, 1, abi.ARITHMETIC)
    fin = np.argwhere(np.isfinite(st.nodal["out"])).tolist()
    assert fin == [[0, 0], [0, 1], [1, 0], [1, 1]]


def test_projection_local_and_harmonic():
    # LOCAL: marker at mdx=0.75, mdy=0.25 of cell (0,0) -> node (0,1), weight .75*.75
    st = state(tiny(markers=[dict(X=0.75, Y=0.25, T=3.0), dict(X=0.8, Y=0.1, T=7.0)]))
    st.project_marker_field("T", "out", 0, 0, abi.ARITHMETIC_LOCAL)
    out = st.nodal["out"]
    w1, w2 = 0.75 * 0.75, 0.8 * 0.9
    assert out[0, 1] == pytest.approx((w1 * 3 + w2 * 7) / (w1 + w2), rel=1e-15)
    assert np.isfinite(out).sum() == 1
    # HARMONIC: both to node (0,1): 2 / (1/3 + 1/7)
    st.project_marker_field("T", "out", 0, 0, abi.HARMONIC)
    assert st.nodal["out"][0, 1] == pytest.approx(2.0 / (1 / 3.0 + 1 / 7.0), rel=1e-15)
    # ties: mdx == 0.5 goes to the right node (>= 0.5)
    st = state(tiny(markers=[dict(X=0.5, Y=0.5, T=1.0)]))
    st.project_marker_field("T", "out", 0, 0, abi.ARITHMETIC_LOCAL)
    assert np.argwhere(np.isfinite(st.nodal["out"])).tolist() == [[1, 1]]


@pytest.mark.parametrize("fam", FAMS)
def test_projection_constant_field_and_permutation(fam):
    w = workloads.convection_small()
    w.markers["T"] = 0.3
    st = state(w)
    st.project_marker_field("T", "out", fam[0], fam[1], abi.ARITHMETIC)
    out = st.nodal["out"]
    fin = np.isfinite(out)
    assert np.allclose(out[fin], 0.3, rtol=1e-14, atol=0)
    # permutation of the marker order changes nothing but rounding
    w2 = workloads.convection_small()
    st1 = state(w2)
    st1.project_marker_field("T", "out", fam[0], fam[1], abi.ARITHMETIC)
    perm = np.random.default_rng(1).permutation(w2.n_markers)
    w3 = copy.copy(w2)
    w3.markers = w2.markers[perm].copy()
    st2 = state(w3)
    st2.project_marker_field("T", "out", fam[0], fam[1], abi.ARITHMETIC)
    a, b = st1.nodal["out"], st2.nodal["out"]
    assert np.array_equal(np.isnan(a), np.isnan(b))
    f = np.isfinite(a)
    assert np.max(np.abs(a[f] - b[f]) / np.abs(a[f])) < 1e-13


def test_projection_conserves_weighted_sum():
    """sum over nodes of (value x weight) equals the sum of the marker values:
    each marker's four bilinear weights add up to 1 and, on basic nodes, no
    contribution falls on a ghost.  The node weights come from an independent
    brute-force scatter written here."""
    w = workloads.convection_small()
    rng = np.random.default_rng(3)
    w.markers["T"] = rng.uniform(1, 2, w.n_markers)
    st = state(w)
    st.project_marker_field("T", "out", 0, 0, abi.ARITHMETIC)
    val = st.nodal["out"].copy()
    m = st.markers
    x, y = w.x, w.y
    wt = np.zeros((w.NY, w.NX))
    for X, Y, cx, cy in zip(m["X"], m["Y"], m["cellX"], m["cellY"]):
        if cx < 0:
            continue
        fx = (X - x[cx]) / (x[cx + 1] - x[cx])
        fy = (Y - y[cy]) / (y[cy + 1] - y[cy])
        wt[cy, cx] += (1 - fx) * (1 - fy)
        wt[cy, cx + 1] += fx * (1 - fy)
        wt[cy + 1, cx] += (1 - fx) * fy
        wt[cy + 1, cx + 1] += fx * fy
    assert np.nansum(val * wt) == pytest.approx(m["T"][m["cellX"] >= 0].sum(), rel=1e-12)


def test_from_scalar_equals_field_projection():
    w = workloads.convection_small()
    st = state(w, extra=("a", "b"))
    st.project_marker_field("T", "a", 0, 0, abi.ARITHMETIC)
    st.project_from_scalar(st.markers["T"], "b")
    assert np.array_equal(st.nodal["a"], st.nodal["b"])


def test_nodal_to_markers_is_exact_on_bilinear_fields():
    w = workloads.convection_small()
    st = state(w, extra=("f",))
    Xn, Yn = np.meshgrid(w.x, w.y)
    st.nodal["f"][:] = 1.0 + 2.0 * Xn - 3.0 * Yn + 0.5 * Xn * Yn
    st.project_nodal_to_markers("f", "T")
    m = st.markers[st.markers["cellX"] >= 0]
    exact = 1.0 + 2.0 * m["X"] - 3.0 * m["Y"] + 0.5 * m["X"] * m["Y"]
    assert np.max(np.abs(m["T"] - exact)) < 1e-13


# ---- advectMarkersRK (markers.c:1865-2250) ---------------------------------
def _flow_state(vx_fn, vy_fn, wxy_val=0.0, N=21):
    w = tiny(x=np.linspace(0, 1, N), y=np.linspace(0, 1, N))
    rng = np.random.default_rng(5)
    n = 400
    w.markers = np.zeros(n, dtype=abi.MARKER_DTYPE)
    w.markers["X"] = rng.uniform(0.2, 0.8, n)
    w.markers["Y"] = rng.uniform(0.2, 0.8, n)
    w.markers["sxx"], w.markers["syy"], w.markers["sxy"] = rng.normal(size=(3, n))
    w.markers["exx"], w.markers["eyy"], w.markers["exy"] = rng.normal(size=(3, n))
    st = state(w, extra=())
    st.nodal["vx"][:] = vx_fn(w.x[None, :], w.yc[1:N + 1][:, None])
    st.nodal["vy"][:] = vy_fn(w.xc[1:N + 1][None, :], w.y[:, None])
    st.nodal["wxy"][:] = wxy_val
    return w, st


def test_rk4_uniform_velocity():
    w, st = _flow_state(lambda x, y: 0.3 + 0 * x * y, lambda x, y: -0.2 + 0 * x * y)
    X0, Y0 = st.markers["X"].copy(), st.markers["Y"].copy()
    st.advect(0.05)
    assert np.max(np.abs(st.markers["X"] - (X0 + 0.05 * 0.3))) < 1e-15
    assert np.max(np.abs(st.markers["Y"] - (Y0 - 0.05 * 0.2))) < 1e-15
    assert np.allclose(st.markers["VX"], 0.3, rtol=1e-15, atol=0)


def test_rk4_linear_flow_closed_form():
    a, dt = 0.7, 0.1
    w, st = _flow_state(lambda x, y: a * x + 0 * y, lambda x, y: -a * y + 0 * x)
    X0, Y0 = st.markers["X"].copy(), st.markers["Y"].copy()
    st.advect(dt)
    h = a * dt
    gx = 1 + h + h ** 2 / 2 + h ** 3 / 6 + h ** 4 / 24
    gy = 1 - h + h ** 2 / 2 - h ** 3 / 6 + h ** 4 / 24
    assert np.max(np.abs(st.markers["X"] - X0 * gx)) < 1e-14
    assert np.max(np.abs(st.markers["Y"] - Y0 * gy)) < 1e-14


def test_rk4_stress_rotation_and_strain():
    wxy, dt = 0.4, 0.3
    w, st = _flow_state(lambda x, y: 0 * x * y, lambda x, y: 0 * x * y, wxy_val=wxy)
    s0 = st.markers[["sxx", "syy", "sxy"]].copy()
    e = st.markers[["exx", "eyy", "exy"]].copy()
    E0 = st.markers["Eii"].copy()
    st.advect(dt)
    th = -wxy * dt
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    for i in range(0, len(s0), 37):
        S = np.array([[s0["sxx"][i], s0["sxy"][i]], [s0["sxy"][i], s0["syy"][i]]])
        Sr = R @ S @ R.T
        got = st.markers[i]
        assert got["sxx"] == pytest.approx(Sr[0, 0], abs=1e-14)
        assert got["syy"] == pytest.approx(Sr[1, 1], abs=1e-14)
        assert got["sxy"] == pytest.approx(Sr[0, 1], abs=1e-14)
    tr0 = s0["sxx"] + s0["syy"]
    assert np.allclose(st.markers["sxx"] + st.markers["syy"], tr0, atol=1e-14)
    dE = dt * np.sqrt(0.5 * (e["exx"] ** 2 + e["eyy"] ** 2) + e["exy"] ** 2)
    assert np.allclose(st.markers["Eii"] - E0, dE, rtol=1e-14)
    assert np.allclose(st.markers["wxy"], wxy, rtol=1e-15)


def test_rk4_top_ghost_kinematic_bc():
    # vx = 1 everywhere, kinematic top value 3: markers in the top half-cell see
    # vx interpolated between the ghost (2*3 - 1 = 5) and 1
    w, st = _flow_state(lambda x, y: 1 + 0 * x * y, lambda x, y: 0 * x * y, N=11)
    w.bv.mechBCTop.type[0] = 0
    w.bv.mechBCTop.value[0] = 3.0
    st.markers["Y"][:5] = [0.01, 0.02, 0.03, 0.04, 0.049]
    st.find_marker_cells()
    st.advect(1e-9)
    yc1 = w.yc[1]
    for i in range(5):
        Y = st.markers["Y"][i]
        frac = (Y - w.yc[0]) / (yc1 - w.yc[0])
        assert st.markers["VX"][i] == pytest.approx(5 * (1 - frac) + 1 * frac, rel=1e-9)


# ---- subgridTemperatureChanges ---------------------------------------------
def _thermal_state(seed=7):
    w = workloads.convection_small(seed)
    st = state(w)
    st.update_marker_density()
    st.project_markers_nodes_all2()
    return w, st


def _interp(w, f, X, Y, cx, cy):
    x, y = w.x, w.y
    fx = (X - x[cx]) / (x[cx + 1] - x[cx])
    fy = (Y - y[cy]) / (y[cy + 1] - y[cy])
    return (f[cy, cx] * (1 - fx) * (1 - fy) + f[cy, cx + 1] * fx * (1 - fy)
            + f[cy + 1, cx] * (1 - fx) * fy + f[cy + 1, cx + 1] * fx * fy)


def test_subgrid_t_zero_diffusivity_is_plain_interpolation():
    w, st = _thermal_state()
    T0 = st.markers["T"].copy()
    d = st.nodal["thisT"] - st.nodal["lastT"]
    st.subgrid_temperature(w.dt, 0.0)
    m = st.markers
    exp = T0 + _interp(w, d, m["X"], m["Y"], m["cellX"], m["cellY"])
    assert np.max(np.abs(m["T"] - exp)) < 1e-14
    # (T0 + dtm) - T0 differs from dtm by one rounding of T: 1e-16 / dt absolute
    assert np.allclose(m["Tdot"], (m["T"] - T0) / w.dt, rtol=0, atol=1e-13)


def test_subgrid_t_infinite_dt_relaxes_to_nodes():
    w, st = _thermal_state()
    m = st.markers
    Tn = _interp(w, st.nodal["lastT"], m["X"], m["Y"], m["cellX"], m["cellY"])
    dts = Tn - m["T"]  # dTsub for dt -> infinity
    proj = oracle.State(w, extra=["p"])
    proj.find_marker_cells()
    proj.project_from_scalar(dts, "p")
    corr = st.nodal["thisT"] - st.nodal["lastT"] - proj.nodal["p"]
    T0 = m["T"].copy()
    st.subgrid_temperature(1e300, 1.0)
    exp = T0 + _interp(w, corr, m["X"], m["Y"], m["cellX"], m["cellY"]) + dts
    assert np.max(np.abs(st.markers["T"] - exp)) < 1e-13


# ---- subgrid stress, marker stress, strain/pressure ------------------------
def _vep_tiny():
    w = workloads.vep_subduction(NX=21, NY=11, NM=3)
    st = state(w, extra=())
    return w, st


def test_subgrid_stress_mu_infinite_relaxes_to_nodes():
    w, st = _vep_tiny()
    st.markers["mu"] = 1e99  # sdif = 1 - exp(-30)
    m = st.markers
    sxy_n = _interp(w, st.nodal["soxy"], m["X"], m["Y"], m["cellX"], m["cellY"])
    st.subgrid_stress(w.dt, 1.0)
    assert np.allclose(st.markers["sxy"], sxy_n, rtol=1e-12, atol=1e-13 * np.abs(sxy_n).max())


def test_subgrid_stress_zero_rate_leaves_markers():
    w, st = _vep_tiny()
    st.markers["mu"] = 1e-300  # eta/mu huge -> sdif = 1 - exp(-0) = 0
    s0 = st.markers[["sxx", "syy", "sxy"]].copy()
    ds0 = st.nodal["dsxy"].copy()
    st.subgrid_stress(w.dt, 1.0)
    for f in ("sxx", "syy", "sxy"):
        assert np.array_equal(st.markers[f], s0[f])
    fin = np.isfinite(st.nodal["dsxy"])
    assert np.array_equal(st.nodal["dsxy"][fin], ds0[fin])


def test_update_stress_constant_fields():
    w, st = _vep_tiny()
    for n, c in (("dsxx", 2.0), ("dsyy", -3.0), ("dsxy", 0.5)):
        st.nodal[n][:] = c
    s0 = st.markers[["sxx", "syy", "sxy"]].copy()
    st.update_marker_stress()
    assert np.allclose(st.markers["sxx"] - s0["sxx"], 2.0, rtol=1e-9)
    assert np.allclose(st.markers["syy"] - s0["syy"], -3.0, rtol=1e-9)
    assert np.allclose(st.markers["sxy"] - s0["sxy"], 0.5, rtol=1e-9)


def test_strain_pressure_constant_fields():
    w, st = _vep_tiny()
    vals = dict(p=5.0, edotxx=1.0, edotyy=-1.0, edotxy=0.25, soxx=4.0, soyy=-4.0, soxy=2.0,
                dsxx=1.0, dsyy=-1.0, dsxy=0.5, wxy=0.125)
    for n, c in vals.items():
        st.nodal[n][:] = c
    dt = 2.0
    st.update_marker_strain_pressure(dt)
    m = st.markers
    eta, mu = m["eta"], m["mu"]
    e11 = 4.0 / (2 * eta) + 1.0 / (2 * dt * mu)
    e12 = 2.0 / (2 * eta) + 0.5 / (2 * dt * mu)
    assert np.allclose(m["p"], 5.0, rtol=1e-14) and np.allclose(m["wxy"], 0.125, rtol=1e-14)
    assert np.allclose(m["exx"], e11, rtol=1e-13)
    assert np.allclose(m["eyy"], -4.0 / (2 * eta) - 1.0 / (2 * dt * mu), rtol=1e-13)
    assert np.allclose(m["exy"], e12, rtol=1e-13)
    assert np.allclose(m["strainRateResidual"], np.sqrt((e11 - 1.0) ** 2 + (e12 - 0.25) ** 2),
                       rtol=1e-12)


def test_strain_pressure_nan_residual_is_an_error():
    w, st = _vep_tiny()
    st.markers["eta"][3] = np.nan
    with pytest.raises(oracle.OracleError):
        st.update_marker_strain_pressure(w.dt)


# ---- checkPlasticYielding / viscosity --------------------------------------
def _yield_case(s, e, p, C=1.0, F=0.0, eta=1.0, mu=1e99, plast=1, law=0):
    w = tiny(markers=[dict(X=1.5, Y=1.5, sxx=s[0], syy=s[1], sxy=s[2], exx=e[0], eyy=e[1],
                           exy=e[2], p=p, eta=eta, mu=mu)])
    w.materials.hasPlasticity[0] = plast
    w.materials.materialCohesion[0] = C
    w.materials.materialFriction[0] = F
    w.materials.materialEta[0] = 7.0
    w.materials.materialMu[0] = mu
    w.materials.hasEtaT[0] = law
    w.options.etamin, w.options.etamax, w.options.fractionalEtamin = 1e-30, 1e30, 1e-30
    return w, state(w)


def test_yield_below_and_above():
    # mu*dt huge -> X = 0: forecast stress = 2 eta e
    w, st = _yield_case(s=(0.1, -0.1, 0.0), e=(0.1, -0.1, 0.0), p=0.0, C=10.0, eta=2.0)
    assert st.check_plastic_yielding(1.0) == 0
    assert st.markers["eta"][0] == 2.0  # not yielding: eta reverts to the pre-call value
    assert st.markers["mu"][0] == 1e99
    # above: sii_new = 2*2*0.1*sqrt(0.5*2) = 0.4 > sigma_y = 0.1 + 0.5*0.2 = 0.2
    w, st = _yield_case(s=(3.0, -3.0, 4.0), e=(0.1, -0.1, 0.0), p=0.2, C=0.1, F=0.5, eta=2.0)
    assert st.check_plastic_yielding(1.0) == 1
    sy = 0.2
    eii = np.sqrt(0.5 * (0.01 + 0.01))
    assert st.markers["eta"][0] == pytest.approx(sy / 2 / eii, rel=1e-15)
    m = st.markers[0]
    assert np.sqrt(0.5 * (m["sxx"] ** 2 + m["syy"] ** 2) + m["sxy"] ** 2) == pytest.approx(sy, rel=1e-15)
    # non-plastic material takes the law's viscosity (law 0: materialEta)
    w, st = _yield_case(s=(0, 0, 0), e=(0, 0, 0), p=0.0, plast=0)
    st.check_plastic_yielding(1.0)
    assert st.markers["eta"][0] == 7.0


def test_viscosity_laws():
    import ctypes as C
    mk = abi.Marker()
    mk.T, mk.Mat, mk.X, mk.Y = 0.4, 0, 0.0, 0.0
    mat = abi.default_materials()
    mat.materialEta[0], mat.QonR[0], mat.Tref[0], mat.materialMu[0] = 2.0, 3.0, 0.5, 9.0
    eta, mu = C.c_double(), C.c_double()
    lib = oracle.load()
    cases = {0: 2.0, 1: 2.0 * np.exp(3.0 * (1 / 0.4 - 1 / 0.5)),
             2: 2.0 * np.exp(3.0 / 0.5 * (1 - (0.4 - 0.5) / 0.5)), 3: 2.0 * 0.4,
             5: 2.0 * np.exp(-3.0 * 0.4)}
    for law, exp in cases.items():
        mat.hasEtaT[0] = law
        assert lib.orc_compute_rheology(C.byref(mk), C.byref(abi.default_options()), C.byref(mat),
                                        C.byref(eta), C.byref(mu)) == 0
        assert eta.value == pytest.approx(exp, rel=1e-15) and mu.value == 9.0
    mat.hasEtaT[0] = 4
    assert lib.orc_compute_rheology(C.byref(mk), C.byref(abi.default_options()), C.byref(mat),
                                    C.byref(eta), C.byref(mu)) != 0


def test_limit_viscosity_clamps_all_markers():
    w = tiny(markers=[dict(X=1, Y=1, eta=1e30), dict(X=-5, Y=1, eta=1e-5), dict(X=1, Y=1, eta=1e20)])
    st = state(w)
    st.limit_viscosity()
    assert st.markers["eta"].tolist() == [1e25, 1e14, 1e20]  # the outside marker too


# ---- residual and glue loops ------------------------------------------------
def test_strain_rate_residual_constant():
    w = workloads.convection_small()
    w.markers["strainRateResidual"] = 0.37
    st = state(w)
    assert st.strain_rate_residual() == pytest.approx(0.37, rel=1e-13)


def test_density_and_maxima():
    w = tiny(markers=[dict(X=1, Y=1, T=2.0, sxx=-5.0), dict(X=2, Y=2, T=3.0, sxx=4.0),
                      dict(X=-1, Y=2, T=9.0, sxx=-99.0)])
    w.materials.materialRho[0], w.materials.materialAlpha[0] = 3.0, 0.5
    st = state(w)
    st.update_marker_density()
    assert st.markers["rho"][:2].tolist() == [-3.0, -4.5] and st.markers["rho"][2] == 0.0
    assert st.max_marker_T() == 3.0 and st.max_abs_sxx() == 5.0
