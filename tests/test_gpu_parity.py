"""GPU parity: every entry point of libfvmic2d_hip.so against the CPU oracle on
the same seeded inputs.

Bars: cell indices and the advected positions / velocities / spins are
bit-identical (pure +,-,*,/ chains compiled without contraction on both sides);
every other FP64 value agrees to 1e-12 relative (``helpers.RTOL``) entry by
entry, against the value itself for single-signed fields and against the
magnitude of the summed terms (``scales``) where a result is a mixed-sign sum;
nodes with zero weight are compared as a NaN mask.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from fvmic2d_amd import abi, workloads
from fvmic2d_amd.host import MarkerPath
from fvmic2d_amd.lib import DeviceArray, FvmError
import scales as S
from helpers import RTOL, assert_bitwise, assert_close

pytestmark = pytest.mark.gpu

FAMILIES = [(0, 0), (-1, -1), (0, 1), (1, 0), (-1, 0), (0, -1), (1, 1), (-1, 1), (1, -1)]
MARKER_F = ["X", "Y", "VX", "VY", "T", "Tdot", "sxx", "syy", "sxy", "eta", "mu", "rho", "rhodot",
            "p", "exx", "eyy", "exy", "Eii", "wxy", "strainRateResidual"]


def _print_case(name):
    print("CASE", name, flush=True)


def _with_stress(w, seed=11):
    """Give the markers non-trivial stress / strain-rate / pressure state."""
    r = workloads.Rng(seed, 9)
    n = w.n_markers
    m = w.markers
    m["sxx"] = r.normal(n)
    m["syy"] = r.normal(n)
    m["sxy"] = r.normal(n)
    m["exx"] = 0.1 * r.normal(n)
    m["eyy"] = 0.1 * r.normal(n)
    m["exy"] = 0.1 * r.normal(n)
    m["Eii"] = r.uniform(n)
    m["p"] = r.normal(n)
    return w


def _edge_markers(w):
    """Markers exactly on grid lines, on the domain edges and outside it."""
    m = w.markers
    x, y = w.x, w.y
    k = 0
    for X, Y in [(x[5], y[7]), (x[0], y[0]), (w.LX, w.LY), (w.LX, y[3]), (x[9], w.LY),
                 (-1e-12, y[2]), (x[2], -1e-300), (w.LX * (1 + 1e-15), y[4]), (np.nan, 0.5),
                 (w.xc[6], w.yc[8]), (w.xc[1], w.yc[1]), (x[w.NX - 2], y[3])]:
        m["X"][k], m["Y"][k] = X, Y
        k += 17
    return w


@pytest.fixture(scope="module")
def small():
    return _edge_markers(_with_stress(workloads.convection_small()))


def _pair(w, extra=()):
    _print_case(w.name)
    st = oracle.State(w, extra=list(extra) + (
        [p[4] for p in w.projections] if isinstance(w.projections, list) else []) + (
        oracle.VEP_PROJ if w.kind == "vep" else []))
    mp = MarkerPath(w, extra=extra)
    return st, mp


def _cmp_markers(st, mp, fields=MARKER_F, exact=(), what="", scales=None):
    """Cells bitwise; ``exact`` fields bitwise; the others to RTOL entry by entry
    against max(|ref|, scales[f]) (pointwise relative where no scale is given)."""
    g = mp.get_markers()
    r = st.markers
    for f in ("cellX", "cellY", "cpu", "Mat"):
        assert_bitwise(g[f], r[f], what + f)
    act = r["cellX"] != -1
    scales = scales or {}
    for f in fields:
        if f in exact:
            assert_bitwise(g[f], r[f], what + f)
        else:
            sc = scales[f][act] if f in scales else None
            assert_close(g[f][act], r[f][act], what + f, scale=sc)
        # inactive markers are left untouched by every function
        assert_bitwise(g[f][~act], r[f][~act], what + f + "[inactive]")


def test_find_marker_cells_bitexact(small):
    st, mp = _pair(small)
    nout_r = st.find_marker_cells()
    nout_g = mp.find_marker_cells()
    assert nout_g == nout_r and nout_r >= 3
    g = mp.get_markers()
    for f in ("cellX", "cellY", "cpu"):
        assert_bitwise(g[f], st.markers[f], f)


def test_find_marker_cells_tanh_grid():
    w = workloads.vep_subduction(NX=81, NY=41, NM=3)
    st, mp = _pair(w)
    assert mp.find_marker_cells() == st.find_marker_cells()
    g = mp.get_markers()
    assert_bitwise(g["cellX"], st.markers["cellX"], "cellX")
    assert_bitwise(g["cellY"], st.markers["cellY"], "cellY")


def _abs_markers(w, field):
    w2 = workloads.Workload(**{k: getattr(w, k) for k in w.__dataclass_fields__})
    w2.markers = w.markers.copy()
    w2.markers[field] = np.abs(w2.markers[field])
    return w2


@pytest.mark.parametrize("method", [abi.ARITHMETIC, abi.ARITHMETIC_LOCAL, abi.HARMONIC])
def test_projection_families(small, method):
    st, mp = _pair(small, extra=["out"])
    st.find_marker_cells()
    mp.find_marker_cells()
    for src in ("T", "sxy", "mat:kThermal"):
        for sx, sy in FAMILIES:
            _print_case("project %s %d %d %d" % (src, sx, sy, method))
            st.project_marker_field(src, "out", sx, sy, method)
            mp.project_marker_field(src, "out", sx, sy, method)
            ref = st.nodal["out"].copy()
            got = mp.get_nodal("out")
            what = "%s (%d,%d) m%d" % (src, sx, sy, method)
            if src.startswith("mat:"):
                assert_close(got, ref, what)
                continue
            # magnitude of the summed terms: the same projection of |f|
            sa = oracle.State(_abs_markers(small, src), extra=["out"])
            sa.find_marker_cells()
            sa.project_marker_field(src, "out", sx, sy, method)
            if method == abi.HARMONIC:
                # wt / sum(1/f): the rounding sits in the sum of reciprocals,
                # judged against sum(|1/f|) / wt = 1 / harmonic(|f|)
                with np.errstate(divide="ignore"):
                    assert_close(1.0 / got, 1.0 / ref, what, scale=np.abs(1.0 / sa.nodal["out"]))
            else:
                assert_close(got, ref, what, scale=np.abs(sa.nodal["out"]))


@pytest.mark.parametrize("sx,sy", [(0, 0), (-1, -1)])
def test_projection_material_index_outside_table(sx, sy):
    """A marker whose material index lies outside the table (here 9 and -3)
    makes every node it touches NaN, as an out-of-bounds table read has no
    value; every other node is bitwise the projection with those markers given
    a valid index (the tables are staged per block in LDS: a wrong slot or a
    stale table shows up here too)."""
    import copy
    w = workloads.convection_small()
    bad = np.array([5, 1000, 20000])
    w.markers["Mat"][bad] = [9, -3, 9]
    print("CASE mat index outside the table, family (%d,%d)" % (sx, sy), flush=True)
    mp = MarkerPath(w, extra=["out"])
    mp.find_marker_cells()
    mp.project_marker_field("mat:kThermal", "out", sx, sy, abi.ARITHMETIC)
    got = mp.get_nodal("out")
    wv = copy.deepcopy(w)
    wv.markers["Mat"][bad] = 0
    mv = MarkerPath(wv, extra=["out"])
    mv.find_marker_cells()
    mv.project_marker_field("mat:kThermal", "out", sx, sy, abi.ARITHMETIC)
    ref = mv.get_nodal("out")
    m = mp.get_markers()
    touched = np.zeros(got.shape, bool)
    for i in bad:
        x, y = m["X"][i], m["Y"][i]
        # the lattice cell of family (sx, sy): basic nodes (0) or centres (-1)
        xs = w.x if sx == 0 else w.xc
        ys = w.y if sy == 0 else w.yc
        ix = int(np.searchsorted(xs, x, side="right")) - 1
        iy = int(np.searchsorted(ys, y, side="right")) - 1
        for dy in (0, 1):
            for dx in (0, 1):
                jy, jx = iy + dy, ix + dx
                if 0 <= jy < got.shape[0] and 0 <= jx < got.shape[1]:
                    touched[jy, jx] = True
    assert touched.any()
    assert np.isnan(got[touched]).all(), "touched nodes"
    assert_bitwise(got[~touched], ref[~touched], "untouched nodes")


def test_projection_is_deterministic(small):
    _, mp = _pair(small, extra=["a", "b"])
    mp.find_marker_cells()
    mp.project_marker_field("sxy", "a", -1, -1, abi.ARITHMETIC)
    mp.find_marker_cells()  # invalidates and rebuilds the bucket order
    mp.project_marker_field("sxy", "b", -1, -1, abi.ARITHMETIC)
    assert_bitwise(mp.get_nodal("a"), mp.get_nodal("b"), "repeat")


def test_project_markers_nodes_all2(small):
    st, mp = _pair(small)
    st.find_marker_cells()
    mp.find_marker_cells()
    sc = S.abs_projections(st, projections="all2")
    st.project_markers_nodes_all2()
    mp.project_markers_nodes_all2()
    for n in ("rho", "etaN", "etaS", "kThermal", "lastT", "Cp", "soxx", "soyy", "soxy"):
        assert_close(mp.get_nodal(n), st.nodal[n], n, scale=sc[n])
    assert np.all(mp.get_nodal("rhodot") == 0.0)
    assert np.all(mp.get_nodal("muS") == 1e99) and np.all(mp.get_nodal("muN") == 1e99)


def test_project_from_scalar_and_back(small):
    st, mp = _pair(small, extra=["out"])
    st.find_marker_cells()
    mp.find_marker_cells()
    f = small.markers["T"] * 3.0 - np.cos(small.markers["X"])
    fd = DeviceArray(mp.ctx, f)
    sa = st.clone()
    sa.project_from_scalar(np.abs(f), "out")  # projection of |f|: the summed magnitude
    st.project_from_scalar(f, "out")
    mp.project_from_scalar(fd.ptr, "out")
    assert_close(mp.get_nodal("out"), st.nodal["out"], "from_scalar", scale=sa.nodal["out"])
    # projectNodalFieldToMarkersS: nodal field -> marker T (the GPU interpolates
    # its own projection; its rounding is relative to the interpolated |f|)
    t_scale = S.interp_basic(st, sa.nodal["out"])
    st.nodal["lastT"][:] = st.nodal["out"]
    mp.set_nodal("lastT", mp.get_nodal("out"))
    st.project_nodal_to_markers("lastT", "T")
    mp.project_nodal_to_markers("lastT", "T")
    _cmp_markers(st, mp, fields=["T"], scales={"T": t_scale})


def _advect_scales(st, dt):
    """The rotation mixes the stress components: one magnitude for all three."""
    s = S.stress_l1(st.markers)
    return {"sxx": s, "syy": s, "sxy": s, "Eii": S.eii_scale(st.markers, dt)}


@pytest.mark.parametrize("jaumann", [0, 1])
@pytest.mark.parametrize("bc", [0, 1, 2])
def test_advect_rk4(small, jaumann, bc):
    w = small
    for side in ("mechBCTop", "mechBCLeft"):
        b = getattr(w.bv, side)
        for k in range(3):
            b.type[k] = bc
            b.value[k] = 0.3 if bc == 0 else 0.0
    try:
        st, mp = _pair(w)
        st.find_marker_cells()
        mp.find_marker_cells()
        sc = _advect_scales(st, w.dt * 1.7)
        st.advect(w.dt * 1.7, jaumann)
        mp.advect(w.dt * 1.7, jaumann)
        _cmp_markers(st, mp, exact=("X", "Y", "VX", "VY", "wxy"), what="bc%d j%d " % (bc, jaumann),
                     scales=sc)
    finally:
        for side in ("mechBCTop", "mechBCLeft"):
            b = getattr(w.bv, side)
            for k in range(3):
                b.type[k] = 1
                b.value[k] = 0.0


@pytest.mark.parametrize("case", ["tiny_grid", "tiny_positions", "extreme_velocities"])
def test_advect_rk4_extreme_magnitudes(case):
    """RK4 advection stays bit-identical far from the usual magnitudes:
    tiny_grid scales the whole grid, the markers and dt by 2^-110; tiny_positions
    puts every seventh marker at |X| < 1e-33 in cell 0; extreme_velocities scales
    vx, vy by 1e-130 in some columns and by 1e120 in some rows (interpolation
    numerators near the bottom of the exponent range, then positions thrown far
    outside the grid)."""
    w = _with_stress(workloads.convection_small())
    m = w.markers
    if case == "tiny_grid":
        s = 2.0 ** -110
        for a in ("xp", "yp", "xcp", "ycp"):
            setattr(w, a, getattr(w, a) * s)
        w.LX *= s
        w.LY *= s
        m["X"] *= s
        m["Y"] *= s
        w.dt *= s
    elif case == "tiny_positions":
        r = np.random.default_rng(5)
        sel = np.arange(0, w.n_markers, 7)
        m["X"][sel] = r.uniform(0.0, 1e-33, len(sel))
    else:
        w.nodal["vx"] = w.nodal["vx"].copy()
        w.nodal["vy"] = w.nodal["vy"].copy()
        for n in ("vx", "vy"):
            w.nodal[n][:, 3:9] *= 1e-130
            w.nodal[n][20:24, :] *= 1e120
    st, mp = _pair(w)
    st.find_marker_cells()
    mp.find_marker_cells()
    sc = _advect_scales(st, w.dt * 1.3)
    st.advect(w.dt * 1.3, 0)
    mp.advect(w.dt * 1.3, 0)
    _cmp_markers(st, mp, exact=("X", "Y", "VX", "VY", "wxy"), what=case + " ", scales=sc)


def test_subgrid_temperature(small):
    st, mp = _pair(small)
    for s in (st, mp):
        s.find_marker_cells()
        s.update_marker_density()
        s.project_markers_nodes_all2()
    lastT = S.abs_projections(st, projections="all2")["lastT"]
    sc = S.subgrid_t_scales(st, st.markers["T"].copy(), lastT, small.dt)
    st.subgrid_temperature(small.dt, 1.0)
    mp.subgrid_temperature(small.dt, 1.0)
    _cmp_markers(st, mp, fields=["T", "Tdot", "rho", "rhodot"], scales=sc)


def _vep_small():
    return workloads.vep_subduction(NX=121, NY=61, NM=4)


@pytest.fixture(scope="module")
def vep():
    return _vep_small()


def test_subgrid_stress_and_update_stress(vep):
    st, mp = _pair(vep)
    for s in (st, mp):
        s.find_marker_cells()
    s_mag = S.stress_l1(st.markers)
    inc, ds_mag = S.subgrid_stress_scales(st, vep.dt, 1.0)
    s_mag = s_mag + inc
    st.subgrid_stress(vep.dt, 1.0)
    mp.subgrid_stress(vep.dt, 1.0)
    _cmp_markers(st, mp, fields=["sxx", "syy", "sxy"], what="subgrid_stress ",
                 scales={f: s_mag for f in S.STRESS})
    for n in ("dsxx", "dsyy", "dsxy"):
        assert_close(mp.get_nodal(n), st.nodal[n], n, scale=ds_mag[n])
    s_mag = s_mag + S.update_stress_scales(st, ds_mag)
    st.update_marker_stress()
    mp.update_marker_stress()
    _cmp_markers(st, mp, fields=["sxx", "syy", "sxy"], what="update_stress ",
                 scales={f: s_mag for f in S.STRESS})


def test_update_marker_strain_pressure(vep):
    st, mp = _pair(vep)
    for s in (st, mp):
        s.find_marker_cells()
    sc = S.strain_pressure_scales(st, vep.dt)
    st.update_marker_strain_pressure(vep.dt)
    mp.update_marker_strain_pressure(vep.dt)
    _cmp_markers(st, mp, fields=["p", "wxy", "exx", "eyy", "exy", "strainRateResidual"], scales=sc)
    r_ref = st.strain_rate_residual()
    r_gpu = mp.strain_rate_residual()
    assert abs(r_gpu - r_ref) <= RTOL * abs(r_ref), (r_gpu, r_ref)


def test_strain_pressure_nan_is_an_error(vep):
    w = _vep_small()
    w.markers["eta"][5] = np.nan
    _, mp = _pair(w)
    mp.find_marker_cells()
    with pytest.raises(FvmError) as e:
        mp.update_marker_strain_pressure(w.dt)
    assert e.value.code == 5


def test_plastic_yielding(vep):
    st, mp = _pair(vep)
    for s in (st, mp):
        s.find_marker_cells()
    y_ref = st.check_plastic_yielding(vep.dt)
    y_gpu = mp.check_plastic_yielding(vep.dt)
    assert y_ref == y_gpu == 1
    _cmp_markers(st, mp, fields=["eta", "mu", "sxx", "syy", "sxy"])
    st.limit_viscosity()
    mp.limit_viscosity()
    _cmp_markers(st, mp, fields=["eta"])


def test_plastic_yielding_viscosity_laws():
    """checkPlasticYielding with every viscosity law of viscosity.c:12-67 on the
    GPU: materials 0-4 use laws 2 (linearised Arrhenius about Tref), 3 (eta T),
    5 (exp(-QonR T)), 6 (the slab decoupling zone, entered by part of the
    markers) and 0, with Drucker-Prager yielding (plasticity 1) and the
    strain-weakening law 2 on material 2."""
    w = _vep_small()
    m, o = w.materials, w.options
    for i, law in enumerate([2, 3, 5, 6, 0]):
        m.hasEtaT[i] = law
    m.QonR[2] = 1e-3
    o.slabAngle, o.decoupleThickness = 0.5, 3.0e4
    o.minDecoupleDepth, o.maxDecoupleDepth, o.decoupleEtaScaler = 1.0e4, 7.0e4, 0.01
    st, mp = _pair(w)
    for s in (st, mp):
        s.find_marker_cells()
    mk = w.markers
    X, Y = mk["X"], mk["Y"]
    zone = (mk["Mat"] == 3) & (X > Y / np.tan(0.5)) & (X < Y / np.tan(0.5) + 3.0e4) & (Y > 1e4) & (Y < 7e4)
    assert zone.sum() > 100 and all((mk["Mat"] == i).sum() > 100 for i in range(5))
    y_ref = st.check_plastic_yielding(w.dt)
    y_gpu = mp.check_plastic_yielding(w.dt)
    assert y_ref == y_gpu
    _cmp_markers(st, mp, fields=["eta", "mu", "sxx", "syy", "sxy"])


def test_glue_loops(vep):
    st, mp = _pair(vep)
    for s in (st, mp):
        s.find_marker_cells()
        s.update_marker_density()
    _cmp_markers(st, mp, fields=["rho", "rhodot"])
    assert mp.max_marker_T() == st.max_marker_T()
    assert mp.max_abs_sxx() == st.max_abs_sxx()


def test_step_convection_small():
    w = _with_stress(workloads.convection_small())
    st, mp = _pair(w)
    sc = S.StepScales(st, vep=False)
    mp.step_thermal()
    _cmp_markers(st, mp, exact=("X", "Y", "VX", "VY", "wxy"), scales=sc.marker)
    for n in ("rho", "etaN", "etaS", "kThermal", "lastT", "Cp", "soxx", "soyy", "soxy"):
        assert_close(mp.get_nodal(n), st.nodal[n], n, scale=sc.nodal[n])


@pytest.mark.parametrize("lanes", [4, 8, 16, 32])
def test_step_convection_tdep_full(lanes):
    w = workloads.convection_tdep()
    st, mp = _pair(w)
    mp.ctx.call("fvm_ctx_set_lanes", C.c_int32(lanes))
    sc = S.StepScales(st, vep=False)
    mp.step_thermal()
    _cmp_markers(st, mp, exact=("X", "Y", "VX", "VY", "wxy"), fields=["X", "Y", "T", "Tdot", "rho"],
                 scales=sc.marker)
    for _, _, _, _, out in w.projections:
        assert_close(mp.get_nodal(out), st.nodal[out], out, scale=sc.nodal[out])


def test_step_crowded_cells():
    """Cells with more than 64 markers take the per-bucket ordering path (and
    gather their payload there): buckets of up to 256 ranked in registers, a
    bucket of 300 by the insertion-sort fallback; results still match the oracle."""
    w = workloads.convection_tdep(N=41, NM=4)
    r = workloads.Rng(5, 3)
    m, x, y = w.markers, w.x, w.y
    a = np.arange(0, 400 * 7, 7)  # 400 markers spread over basic cell (10, 12)
    m["X"][a] = x[10] + (x[11] - x[10]) * r.uniform(len(a))
    m["Y"][a] = y[12] + (y[13] - y[12]) * r.uniform(len(a))
    b = np.arange(2801, 2951)  # 150 markers inside one quadrant of cell (20, 5)
    m["X"][b] = x[20] + 0.49 * (x[21] - x[20]) * r.uniform(len(b))
    m["Y"][b] = y[6] - 0.49 * (y[6] - y[5]) * r.uniform(len(b))
    c = np.arange(3001, 3301 * 3, 3)[:300]  # 300 markers inside one quadrant of cell (30, 30)
    m["X"][c] = x[31] - 0.49 * (x[31] - x[30]) * r.uniform(len(c))
    m["Y"][c] = y[30] + 0.49 * (y[31] - y[30]) * r.uniform(len(c))
    st, mp = _pair(w)
    sc = S.StepScales(st, vep=False)
    mp.step_thermal()
    _cmp_markers(st, mp, exact=("X", "Y", "VX", "VY", "wxy"), fields=["X", "Y", "T", "Tdot", "rho"],
                 scales=sc.marker)
    for _, _, _, _, out in w.projections:
        assert_close(mp.get_nodal(out), st.nodal[out], out, scale=sc.nodal[out])


def _cmp_vep(st, mp, sc, y_r, r_r, y_g, r_g):
    assert y_r == y_g
    assert abs(r_g - r_r) <= RTOL * abs(r_r)
    _cmp_markers(st, mp, exact=("X", "Y", "VX", "VY", "wxy"), scales=sc.marker)
    for n in oracle.VEP_PROJ + ["dsxx", "dsyy", "dsxy"]:
        assert_close(mp.get_nodal(n), st.nodal[n], n, scale=sc.nodal.get(n))


@pytest.mark.parametrize("lanes", [8, 4])
def test_step_vep(vep, lanes):
    st, mp = _pair(vep)
    mp.ctx.call("fvm_ctx_set_lanes", C.c_int32(lanes))
    sc = S.StepScales(st, vep=True)
    y_r, r_r = sc.result
    y_g, r_g = mp.step_vep()
    _cmp_vep(st, mp, sc, y_r, r_r, y_g, r_g)


def test_step_vep_wide_grid():
    """A 3100 x 12 graded grid: its coordinates do not fit the LDS staging of the
    location (48 KB) or advection (32 KB) kernels, which then read them from
    global memory; strongly graded cells take the bisection fallback."""
    w = workloads.vep_subduction(NX=3100, NY=12, NM=2)
    st, mp = _pair(w)
    sc = S.StepScales(st, vep=True)
    y_r, r_r = sc.result
    y_g, r_g = mp.step_vep()
    _cmp_vep(st, mp, sc, y_r, r_r, y_g, r_g)


@pytest.mark.timeout(600)
def test_step_vep_full_size():
    """The visco-elasto-plastic step at full size: 801 x 401 tanh-stretched grid,
    25 markers per cell (8M markers), five materials with Drucker-Prager
    yielding, against the oracle."""
    import time
    t0 = time.time()
    w = workloads.vep_subduction()
    assert (w.NX, w.NY, w.n_markers) == (801, 401, 8_000_000)
    st, mp = _pair(w)
    print("  workload and both states in %.1f s" % (time.time() - t0), flush=True)
    sc = S.StepScales(st, vep=True)
    y_r, r_r = sc.result
    y_g, r_g = mp.step_vep()
    print("  oracle and GPU steps done at %.1f s" % (time.time() - t0), flush=True)
    _cmp_vep(st, mp, sc, y_r, r_r, y_g, r_g)
    print("  compared at %.1f s" % (time.time() - t0), flush=True)


def test_step_is_deterministic():
    w = workloads.convection_tdep(N=101, NM=5)
    a = MarkerPath(w)
    b = MarkerPath(w)
    for _ in range(3):
        a.step_thermal()
        b.step_thermal()
    ma, mb = a.get_markers(), b.get_markers()
    for f in MARKER_F:
        assert_bitwise(ma[f], mb[f], f)
    for _, _, _, _, out in w.projections:
        assert_bitwise(a.get_nodal(out), b.get_nodal(out), out)


def test_empty_marker_set():
    w = workloads.convection_small()
    w.markers = w.markers[:0]
    mp = MarkerPath(w)
    assert mp.find_marker_cells() == 0
    mp.project_markers_nodes_all2()
    assert np.isnan(mp.get_nodal("lastT")).all()  # zero weight everywhere -> NaN
    mp.step_thermal()


def test_hundred_steps_stay_bitwise():
    """100 consecutive thermal steps of the 51x51 convection workload: cells and
    the advected positions and velocities stay bit-identical to the oracle's;
    temperature, which each step updates from projected and interpolated nodal
    fields, drifts by at most 1e-10 of its scale over the run."""
    w = workloads.convection_small()
    st, mp = _pair(w)
    for _ in range(100):
        st.step_thermal()
        mp.step_thermal()
    g, r = mp.get_markers(), st.markers
    for f in ("cellX", "cellY"):
        assert_bitwise(g[f], r[f], f)
    act = r["cellX"] != -1
    for f in ("X", "Y", "VX", "VY"):
        assert_bitwise(g[f][act], r[f][act], f)
    assert_close(g["T"][act], r["T"][act], "T after 100 steps", rtol=1e-10)
