"""Pinning the oracle's checkMarkerDensity (markers.c:455-781).

The reference draws from drand48 / PetscRandom; the oracle (and the GPU path)
use the counter stream orc_pc_u01 instead, so these checks fix everything that
does not depend on which random numbers come out, and compute the rest from
the stream itself:

* the stream: SplitMix64 as published (Steele, Lea, Flood 2014), checked here
  against an independent Python implementation and for uniformity;
* quarter-cell counts after the call, recounted from positions with numpy:
  every under-full quarter holds exactly its target ceil(NMX*(NX-1)/2*dx/LX) *
  ceil(NMY*(NY-1)/2*dy/LY), every over-full one exactly maxMarkers, the others
  are unchanged;
* which markers leave an over-full quarter (the lowest keys), the out list and
  its recycling from the end, new positions inside their quarter, and the
  nearest-neighbour copy by brute force on small inputs;
* the slab and inflow markers of benchmarkInitialConditions.c:52-92 in closed
  form; the capacity and negative-target errors.
"""
import math

import numpy as np
import pytest

import oracle
from fvmic2d_amd import abi, workloads
from test_oracle import tiny

pytestmark = pytest.mark.usefixtures("oracle_lib")

M64 = (1 << 64) - 1


def _mix(z):
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def u01(seed, stream, ctr):
    key = _mix((seed + 0x9E3779B97F4A7C15 * (stream + 1)) & M64)
    return (_mix(key ^ ctr) >> 11) * (1.0 / 9007199254740992.0)


def test_counter_stream_is_splitmix64():
    lib = oracle.load()
    # SplitMix64's first output from state 0 (the published reference value)
    assert _mix(0) == 0xE220A8397B1DCDAF
    for seed, stream, ctr in [(0, 1, 0), (7, 2, 5), (2**40 + 3, 3, (17 << 32) | 9), (M64, 1, M64)]:
        assert lib.orc_pc_u01(seed, stream, ctr) == u01(seed, stream, ctr)
    v = np.array([lib.orc_pc_u01(11, 2, i) for i in range(20000)])
    assert 0.0 <= v.min() and v.max() < 1.0
    assert abs(v.mean() - 0.5) < 0.01 and abs(v.var() - 1.0 / 12.0) < 0.003


# ---- a hand-made 3 x 3-node grid: 2 x 2 cells, 4 x 4 quarter cells -----------
def _grid_case(markers, NMX=8, NMY=8, minM=2, maxM=6, capacity=None, law=0, plate=0.5):
    w = tiny(x=[0.0, 1.0, 2.0], y=[0.0, 1.0, 2.0], markers=markers)
    o = w.options
    o.NMX, o.NMY, o.minMarkers, o.maxMarkers = NMX, NMY, minM, maxM
    o.slabAngle, o.slabAge = 30.0 / 180.0 * math.pi, 50.0
    o.overridingPlateThickness, o.rootThickness, o.rootCenter, o.rootWidth = plate, 0.25, 1.9, 0.2
    w.materials.hasEtaT[0], w.materials.materialEta[0], w.materials.materialMu[0] = law, 3.0, 5.0
    st = oracle.State(w, extra=(), capacity=capacity)
    st.find_marker_cells()
    return w, st


def _quarter_counts(st, NX=3, NY=3):
    """Recount markers per quarter cell from positions (numpy), reference order."""
    m = st.markers[:st.ms.nMark]
    x, y = st.w.x, st.w.y
    xq = np.sort(np.concatenate([x, st.w.xc[1:NX]]))  # quarter boundaries
    yq = np.sort(np.concatenate([y, st.w.yc[1:NY]]))
    inside = (m["X"] >= 0) & (m["X"] <= x[-1]) & (m["Y"] >= 0) & (m["Y"] <= y[-1])
    i = np.clip(np.searchsorted(xq, m["X"][inside], side="right") - 1, 0, 2 * (NX - 1) - 1)
    j = np.clip(np.searchsorted(yq, m["Y"][inside], side="right") - 1, 0, 2 * (NY - 1) - 1)
    c = np.zeros((2 * (NY - 1), 2 * (NX - 1)), int)
    np.add.at(c, (j, i), 1)
    return c


def _fill(rng, qx, qy, k, **f):
    """k markers inside quarter (qx, qy) of the unit-cell grid."""
    return [dict(X=0.5 * qx + 0.5 * rng.uniform(0.05, 0.95), Y=0.5 * qy + 0.5 * rng.uniform(0.05, 0.95),
                 T=float(rng.uniform(1, 2)), eta=float(rng.uniform(1, 9)), Mat=0, **f) for _ in range(k)]


def _base(rng, counts):
    mk = []
    for (qx, qy), k in counts.items():
        mk += _fill(rng, qx, qy, k)
    return mk


def test_targets_removal_and_counts():
    rng = np.random.default_rng(1)
    counts = {(qx, qy): 4 for qx in range(4) for qy in range(4)}
    counts[(2, 1)] = 0   # empty: refilled to the target
    counts[(1, 2)] = 1   # under minMarkers 2
    counts[(3, 3)] = 10  # over maxMarkers 6: 4 leave
    counts[(0, 3)] = 7   # over by one
    w, st = _grid_case(_base(rng, counts), capacity=200)
    before = _quarter_counts(st)
    n0 = st.ms.nMark
    old = st.markers[:n0].copy()
    added, removed = st.check_marker_density(seed=5)
    # target per quarter: ceil(8*2/2.0*0.5/2)^2 = 2*2 = 4
    assert (added, removed) == ((4 - 0) + (4 - 1), 4 + 1)
    after = _quarter_counts(st)
    exp = before.copy()
    exp[1, 2], exp[2, 1] = 4, 4
    exp[3, 3], exp[3, 0] = 6, 6
    assert np.array_equal(after, exp)
    # removed ones: the lowest keys of their quarter's members in index order
    q = 3 + 8 * 3 // 2  # quarter (i=3, j=3) in reference order i + 2m*j, 2m = 4
    assert q == 15
    members = [k for k in range(n0) if 1.5 <= old["X"][k] < 2.0 and 1.5 <= old["Y"][k] < 2.0]
    keys = [u01(5, 1, (15 << 32) | k) for k in range(len(members))]
    gone = {members[k] for k in np.argsort(keys, kind="stable")[:4]}
    out = {k for k in range(n0) if st.markers["cellX"][k] == -1 and old["cellX"][k] != -1}
    # recycled slots were refilled: a removed slot is either still out or holds a new marker
    for k in gone:
        assert st.markers["X"][k] == -1.0 or k in (set(range(n0)) - out)
    assert len(gone) == 4


def test_recycling_order_positions_and_neighbours():
    rng = np.random.default_rng(2)
    counts = {(qx, qy): 3 for qx in range(4) for qy in range(4)}
    counts[(2, 2)] = 0
    counts[(1, 1)] = 9
    mk = _base(rng, counts)
    mk.insert(3, dict(X=5.0, Y=0.5))     # outside: first entry of the out list
    mk.insert(10, dict(X=0.5, Y=-2.0))   # outside: second
    # plate deeper than the box: no inflow markers; quarter (2, 2) is right of x[1]: no slab
    w, st = _grid_case(mk, minM=1, maxM=6, capacity=200, plate=5.0)
    n0 = st.ms.nMark
    old = st.markers[:n0].copy()
    members = [k for k in range(n0) if 0.5 <= old["X"][k] < 1.0 and 0.5 <= old["Y"][k] < 1.0]
    keys = [u01(9, 1, ((1 + 4 * 1) << 32) | k) for k in range(len(members))]
    removed = [members[k] for k in np.argsort(keys, kind="stable")[:3]]  # in removal order
    isout = [3, 10] + removed
    added, nrem = st.check_marker_density(seed=9)
    assert (added, nrem) == (4, 3)
    assert st.ms.nMark == n0  # 4 added, 5 slots to recycle: no growth
    # new marker iMark goes to isout[-1 - iMark]; positions from streams 2 / 3
    for i in range(4):
        slot = isout[-1 - i]
        X = (1.5 - 1.0) * u01(9, 2, i) + 1.0
        Y = (1.5 - 1.0) * u01(9, 3, i) + 1.0
        assert (st.markers["X"][slot], st.markers["Y"][slot]) == (X, Y)
        # brute-force nearest neighbour among the original markers still in the
        # domain (and recycled slots filled before), within one cell width
        filled = set(isout[len(isout) - i:])  # slots the new markers before this one took
        best, bd, src = -1, 4.0, None
        for m1 in range(n0):
            if m1 == slot:
                continue
            if m1 in filled:
                c = st.markers[m1]   # refilled earlier in this call: its new content
            elif m1 in removed:
                continue             # removed and not (yet) refilled
            else:
                c = old[m1]
            if c["cellX"] == -1:
                continue
            if abs(c["X"] - X) <= 1.0 and abs(c["Y"] - Y) <= 1.0:
                d = math.sqrt((c["X"] - X) ** 2 + (c["Y"] - Y) ** 2)
                if d < bd:
                    best, bd, src = m1, d, c
        assert best >= 0
        for f in ("T", "eta", "Mat", "cellX", "cellY"):
            assert st.markers[f][slot] == src[f], (i, f)
    # the fifth slot of the out list (the first outside marker) stays out
    assert st.markers["cellX"][3] == -1 and st.markers["X"][3] == 5.0


def test_slab_and_inflow_markers():
    rng = np.random.default_rng(3)
    counts = {(qx, qy): 4 for qx in range(4) for qy in range(4)}
    counts[(0, 1)] = 0  # left column: x < x[1] -> slab markers
    counts[(3, 3)] = 0  # right column, deep: x >= x[NX-2] and y >= plate depth -> inflow
    w, st = _grid_case(_base(rng, counts), capacity=200, law=1)
    w.materials.QonR[0], w.materials.Tref[0] = 100.0, 1000.0
    n0 = st.ms.nMark
    added, _ = st.check_marker_density(seed=13)
    assert added == 8 and st.ms.nMark == n0 + 8
    o = w.options
    kappa, age = 0.7272e-6, o.slabAge * 1.0e6 * 3.15576e7
    for k in range(n0, n0 + 8):
        m = st.markers[k]
        if m["X"] < 1.0:
            T = 273 + (1623 - 273) * math.erf((m["Y"] - m["X"] * math.tan(o.slabAngle)) / (2.0 * math.sqrt(kappa * age)))
        else:
            assert m["Y"] >= 0.5  # plate depth at LX = 2.0: outside the root, 0.5
            T = 1573.0
        assert m["T"] == pytest.approx(T, rel=1e-15)
        assert m["eta"] == pytest.approx(3.0 * math.exp(100.0 * (1.0 / T - 1.0 / 1000.0)), rel=1e-14)
        assert m["mu"] == 5.0 and m["Mat"] == 0
        assert (m["sxx"], m["Tdot"], m["p"], m["Eii"], m["isYielding"]) == (0.0, 0.0, 0.0, 0.0, 0)
        assert m["cellX"] == int(m["X"] >= 1.0) and m["cellY"] == int(m["Y"] >= 1.0)


def test_capacity_and_negative_target():
    rng = np.random.default_rng(4)
    counts = {(qx, qy): 4 for qx in range(4) for qy in range(4)}
    counts[(1, 1)] = 0
    n = sum(counts.values())
    # exactly full after adding: the reference neither adds nor aborts
    w, st = _grid_case(_base(rng, counts), capacity=n + 4)
    assert st.check_marker_density(seed=1) == (0, 0) and st.ms.nMark == n
    w, st = _grid_case(_base(rng, counts), capacity=n + 3)
    with pytest.raises(oracle.OracleError):
        st.check_marker_density(seed=1)
    # minMarkers above the target: the reference aborts
    counts[(1, 1)] = 3
    w, st = _grid_case(_base(rng, counts), minM=5, capacity=200)
    with pytest.raises(oracle.OracleError):
        st.check_marker_density(seed=1)


def test_counts_on_a_larger_workload():
    """convection_small (4 markers per quarter) with a band of quarters thinned
    to one marker each and a column crowded: after the call every quarter is
    within [minMarkers, maxMarkers] (the target, 4, lies inside)."""
    w = workloads.convection_small()
    m = w.markers
    sel = np.flatnonzero((m["X"] > 0.3) & (m["X"] < 0.36) & (m["Y"] > 0.5) & (m["Y"] < 0.6))
    q = (np.floor(m["X"][sel] / 0.01).astype(int), np.floor(m["Y"][sel] / 0.01).astype(int))
    first = np.unique(q[0] * 1000 + q[1], return_index=True)[1]
    move = np.setdiff1d(sel, sel[first])  # keep one marker per quarter
    m["X"][move] = 0.71 + 0.003 * np.arange(len(move)) / len(move)  # crowd a column of quarters
    # target ceil(6*50/2.0*0.01/1)^2 = ceil(1.5)^2 = 4 (an argument away from an integer,
    # so the spacing's rounding cannot tip the ceiling)
    w.options.NMX = w.options.NMY = 6
    w.options.minMarkers, w.options.maxMarkers = 2, 6
    st = oracle.State(w, extra=(), capacity=2 * w.n_markers)
    st.find_marker_cells()
    c0 = _quarter_counts(st, w.NX, w.NY)
    added, removed = st.check_marker_density(seed=3)
    assert added == 3 * int((c0 < 2).sum()) and removed == int((c0[c0 > 6] - 6).sum()) > 0
    st.find_marker_cells()
    c = _quarter_counts(st, w.NX, w.NY)
    assert c.min() >= 2 and c.max() <= 6
    assert np.array_equal(c[(c0 >= 2) & (c0 <= 6)], c0[(c0 >= 2) & (c0 <= 6)])
