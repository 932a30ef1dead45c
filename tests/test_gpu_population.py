"""Marker population control on the GPU (fvm_check_marker_density) against the
oracle's checkMarkerDensity (markers.c:455-781).

Both draw from the same counter stream, so every decision (which quarters are
refilled or thinned, which markers leave, where new markers go, which slots
they recycle, whose fields they copy) is bitwise the oracle's.  The fields of
slab-inflow markers come from erf and the viscosity law (GPU and host libm):
1e-12 relative, entry by entry.  Over 100 thermal steps with population control
every quarter cell stays within [minMarkers, maxMarkers], and positions and
cells stay bitwise the oracle's.
"""
import numpy as np
import pytest

import oracle
from fvmic2d_amd import abi, workloads
from fvmic2d_amd.host import MarkerPath
from fvmic2d_amd.lib import FvmError
from helpers import assert_bitwise, assert_close

pytestmark = pytest.mark.gpu

FP = [n for n in abi.MARKER_DTYPE.names if abi.MARKER_DTYPE[n].kind == "f"]
IP = [n for n in abi.MARKER_DTYPE.names if abi.MARKER_DTYPE[n].kind == "i"]


def _options(w):
    o = w.options
    # 4 markers per quarter; target ceil(6*50/2.0*0.01)^2 = ceil(1.5)^2 = 4
    o.NMX = o.NMY = 6
    o.minMarkers, o.maxMarkers = 2, 6
    o.overridingPlateThickness, o.rootThickness, o.rootCenter, o.rootWidth = 0.5, 0.1, 0.2, 0.1
    o.slabAngle, o.slabAge = 0.5, 50.0
    m = w.materials
    m.hasEtaT[0], m.materialEta[0], m.QonR[0], m.Tref[0] = 1, 1e21, 5.0, 1.0


def _keep_one_per_quarter(m, sel, h=0.01):
    q = np.floor(m["X"][sel] / h).astype(int) * 100000 + np.floor(m["Y"][sel] / h).astype(int)
    first = np.unique(q, return_index=True)[1]
    return np.setdiff1d(sel, sel[first])


def _population_workload():
    """convection_small with three thinned bands (interior: nearest-neighbour
    copies; left column: slab-inflow markers; right column below the plate:
    mantle-inflow markers), the thinned markers crowded into one column of
    quarters, and a few markers outside the domain (recycled first)."""
    w = workloads.convection_small()
    _options(w)
    m = w.markers
    X, Y = m["X"], m["Y"]
    bands = [(0.3, 0.36, 0.5, 0.6), (0.0, 0.02, 0.2, 0.3), (0.98, 1.0, 0.7, 0.8)]
    move = []
    for x0, x1, y0, y1 in bands:
        sel = np.flatnonzero((X > x0) & (X < x1) & (Y > y0) & (Y < y1))
        move.append(_keep_one_per_quarter(m, sel))
    move = np.concatenate(move)
    X[move] = 0.7105 + 0.0035 * np.arange(len(move)) / len(move)
    out = np.arange(5, w.n_markers, 997)
    X[out] = 1.5
    return w


def _pair(w, cap):
    st = oracle.State(w, extra=(), capacity=cap)
    mp = MarkerPath(w, capacity=cap)
    return st, mp


def _compare(st, mp, what, rtol_special=1e-12):
    g = mp.get_markers()
    n = st.ms.nMark
    assert len(g) == n, (what, len(g), n)
    r = st.markers[:n]
    for f in IP:
        assert_bitwise(g[f], r[f], what + f)
    for f in FP:
        if f in ("T", "eta", "mu"):
            assert_close(g[f], r[f], what + f, rtol=rtol_special)
        else:
            assert_bitwise(g[f], r[f], what + f)


def test_population_control_matches_oracle():
    print("CASE population control convection_small bands", flush=True)
    w = _population_workload()
    cap = 2 * w.n_markers
    st, mp = _pair(w, cap)
    st.find_marker_cells()
    mp.find_marker_cells()
    ra = st.check_marker_density(seed=7)
    ga = mp.check_marker_density(seed=7)
    assert ra == ga and ra[0] > 0 and ra[1] > 0, (ra, ga)
    _compare(st, mp, "")
    # the three kinds of added marker occur
    r = st.markers[:st.ms.nMark]
    assert ((r["T"] > 273.0) & (r["T"] < 1623.0) & (r["X"] < w.x[1]) & (r["X"] >= 0)).any()  # slab
    assert ((r["T"] == 1573.0) & (r["X"] >= w.x[-2])).any()  # mantle inflow
    # a second call on the refilled set changes nothing but the crowded column
    st.find_marker_cells()
    mp.find_marker_cells()
    assert st.check_marker_density(seed=8) == mp.check_marker_density(seed=8)
    _compare(st, mp, "second call ")


def test_population_capacity_error():
    w = _population_workload()
    st, mp = _pair(w, w.n_markers + 10)
    mp.find_marker_cells()
    with pytest.raises(FvmError) as e:
        mp.check_marker_density(seed=7)
    assert e.value.code == 14
    st.find_marker_cells()
    with pytest.raises(oracle.OracleError):
        st.check_marker_density(seed=7)


def _quarter_counts(m, w):
    xq = np.sort(np.concatenate([w.x, w.xc[1:w.NX]]))
    yq = np.sort(np.concatenate([w.y, w.yc[1:w.NY]]))
    ok = m["cellX"] != -1
    i = np.clip(np.searchsorted(xq, m["X"][ok], side="right") - 1, 0, 2 * (w.NX - 1) - 1)
    j = np.clip(np.searchsorted(yq, m["Y"][ok], side="right") - 1, 0, 2 * (w.NY - 1) - 1)
    c = np.zeros((2 * (w.NY - 1), 2 * (w.NX - 1)), int)
    np.add.at(c, (j, i), 1)
    return c


@pytest.mark.timeout(300)
def test_hundred_steps_with_population_control():
    print("CASE population control, 100 thermal steps", flush=True)
    w = workloads.convection_small()
    _options(w)
    cap = 2 * w.n_markers
    st, mp = _pair(w, cap)
    added = removed = 0
    for step in range(100):
        for s in (st, mp):
            s.find_marker_cells()
        ra = st.check_marker_density(seed=1000 + step)
        ga = mp.check_marker_density(seed=1000 + step)
        assert ra == ga, (step, ra, ga)
        added, removed = added + ga[0], removed + ga[1]
        mp.find_marker_cells()
        c = _quarter_counts(mp.get_markers(), w)
        assert c.min() >= 2 and c.max() <= 6, (step, c.min(), c.max())
        st.step_thermal()
        mp.step_thermal()
    print("  added %d, removed %d over 100 steps" % (added, removed), flush=True)
    assert added > 0 and removed > 0
    g, r = mp.get_markers(), st.markers[:st.ms.nMark]
    assert len(g) == len(r)
    for f in ("cellX", "cellY", "Mat"):
        assert_bitwise(g[f], r[f], f)
    for f in ("X", "Y", "VX", "VY"):
        assert_bitwise(g[f], r[f], f)
    act = r["cellX"] != -1
    assert_close(g["T"][act], r["T"][act], "T after 100 steps", rtol=1e-10)


@pytest.mark.timeout(300)
def test_population_control_with_reorder():
    """60 thermal steps with population control every step and the store
    reordered every 10 steps (fvm_markers_reorder); the oracle's markers are
    permuted by the same order, so every later choice by index (removal, hole
    recycling, inheritance) must still be the oracle's."""
    print("CASE population control + reorder, 60 thermal steps", flush=True)
    w = workloads.convection_small()
    _options(w)
    st, mp = _pair(w, 2 * w.n_markers)
    for step in range(60):
        for s in (st, mp):
            s.find_marker_cells()
        if step % 10 == 5:
            order = mp.reorder_markers()
            n = st.ms.nMark
            assert len(order) == n
            st.markers[:n] = st.markers[:n][order]
        ra = st.check_marker_density(seed=500 + step)
        ga = mp.check_marker_density(seed=500 + step)
        assert ra == ga, (step, ra, ga)
        st.step_thermal()
        mp.step_thermal()
    g, r = mp.get_markers(), st.markers[:st.ms.nMark]
    assert len(g) == len(r)
    for f in ("cellX", "cellY", "Mat", "X", "Y", "VX", "VY"):
        assert_bitwise(g[f], r[f], f)
    act = r["cellX"] != -1
    assert_close(g["T"][act], r["T"][act], "T after 60 steps", rtol=1e-10)
