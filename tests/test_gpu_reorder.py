"""fvm_markers_reorder: the store moved into cell order.

The expected order is formed here by brute force from the downloaded cells
and positions (a lexicographic sort by (bucket key, index), then the unlocated
markers by index), so a wrong key, quadrant test or tie rule fails it.  Every
column must move with its marker bitwise; projections right after a reorder
are bitwise the projections before it (the in-bucket order is kept); and a run
that reorders between steps stays with the oracle's run on the original order
(positions and cells bitwise, T within the multi-step bar)."""
import numpy as np
import pytest

import oracle
from fvmic2d_amd import abi, workloads
from fvmic2d_amd.host import MarkerPath
from helpers import assert_bitwise, assert_close

pytestmark = pytest.mark.gpu


def _scrambled(w, seed=3, n_out=0):
    """The workload's markers in a random index order, n_out of them moved
    outside the domain."""
    rng = np.random.default_rng(seed)
    w.markers = w.markers[rng.permutation(w.n_markers)].copy()
    if n_out:
        idx = rng.choice(w.n_markers, n_out, replace=False)
        w.markers["X"][idx] = w.LX * 1.5
    return w


def _expected_order(m, w):
    ncx = w.NX - 1 + int(getattr(w, "xperiodic", 0))
    cx, cy = m["cellX"].astype(np.int64), m["cellY"].astype(np.int64)
    act = cx != -1
    # padded centre arrays: element 2 is index 0
    key = np.full(len(m), -1, np.int64)
    a = np.flatnonzero(act)
    bx = (m["X"][a] >= w.xcp[cx[a] + 3]).astype(np.int64)
    by = (m["Y"][a] >= w.ycp[cy[a] + 3]).astype(np.int64)
    key[a] = ((cy[a] * ncx + cx[a]) << 2) | bx | (by << 1)
    ordered = a[np.lexsort((a, key[a]))]
    return np.concatenate([ordered, np.flatnonzero(~act)]), key


def _check_reorder(w):
    mp = MarkerPath(w)
    mp.find_marker_cells()
    before = mp.get_markers()
    want, key = _expected_order(before, w)
    order = mp.reorder_markers()
    assert_bitwise(order, want, "order")
    after = mp.get_markers()
    for f in abi.MARKER_DTYPE.names:
        assert_bitwise(after[f], before[f][order], f)
    k = key[order]
    nact = int((before["cellX"] != -1).sum())
    assert (np.diff(k[:nact]) >= 0).all() and (k[nact:] == -1).all()
    return mp


def test_reorder_order_and_columns():
    w = _scrambled(workloads.convection_small(), n_out=37)
    print("CASE reorder convection_small scrambled", w.n_markers, flush=True)
    _check_reorder(w)


def test_reorder_periodic():
    w = _scrambled(workloads.periodic_box("plain", NX=64, NY=41, NM=4), seed=5)
    print("CASE reorder periodic_box scrambled", w.n_markers, flush=True)
    mp = _check_reorder(w)
    # the last periodic cell column (cellX = NX - 1) is ordered like the others
    assert (mp.get_markers()["cellX"] == w.NX - 1).any()


def test_reorder_twice_is_identity_and_projections_unchanged():
    w = _scrambled(workloads.convection_small(), n_out=11)
    mp = MarkerPath(w)
    mp.find_marker_cells()
    mp.project_markers_nodes_all2()
    names = ("rho", "etaN", "etaS", "kThermal", "lastT", "Cp", "soxx", "soyy", "soxy")
    first = {n: mp.get_nodal(n) for n in names}
    mp.reorder_markers()
    mp.project_markers_nodes_all2()
    for n in names:
        assert_bitwise(mp.get_nodal(n), first[n], n + " after reorder")
    # a second reorder of an ordered store moves nothing
    mp.find_marker_cells()
    order = mp.reorder_markers()
    assert_bitwise(order, np.arange(w.n_markers, dtype=np.int32), "second order")


@pytest.mark.timeout(300)
def test_steps_with_reorder_follow_oracle():
    """20 thermal steps from a scrambled store, reordered before steps 0, 5, 10
    and 15, against the oracle's 20 steps on the original order."""
    w = _scrambled(workloads.convection_small(), seed=9, n_out=5)
    print("CASE 20 thermal steps with reorder", w.n_markers, flush=True)
    st = oracle.State(w, extra=[])
    mp = MarkerPath(w)
    ref = np.arange(w.n_markers)
    for step in range(20):
        if step % 5 == 0:
            mp.find_marker_cells()
            ref = ref[mp.reorder_markers()]
        st.step_thermal()
        mp.step_thermal()
    g, r = mp.get_markers(), st.markers[ref]
    for f in ("cellX", "cellY", "Mat"):
        assert_bitwise(g[f], r[f], f)
    act = r["cellX"] != -1
    for f in ("X", "Y", "VX", "VY"):
        assert_bitwise(g[f][act], r[f][act], f)
    assert_close(g["T"][act], r["T"][act], "T after 20 steps", rtol=1e-10)
    for n in ("rho", "etaN", "kThermal", "lastT", "Cp"):
        assert_close(mp.get_nodal(n), st.nodal[n], n, rtol=1e-10)


@pytest.mark.timeout(300)
def test_ensemble_members_reordered_between_steps():
    """Ensemble members are the callers' stores: reordered in place between
    batched steps (cells located first; every 10 of 40 graph-replayed steps),
    each member stays bitwise equal to the same problem stepped alone and
    reordered the same way."""
    from fvmic2d_amd.host import Ensemble
    from fvmic2d_amd.lib import Context
    print("CASE ensemble with reorder, 4 instances", flush=True)
    ctx = Context()
    ws = [_scrambled(w, seed=20 + i) for i, w in enumerate(workloads.ensemble(4, seed=9, nmin=17, nmax=61))]
    alone, batch = [MarkerPath(w, ctx=ctx) for w in ws], [MarkerPath(w, ctx=ctx) for w in ws]
    ens = Ensemble(batch)
    ens.capture()
    for s in range(40):
        for a, b in zip(alone, batch):
            if s % 10 == 0:
                for p in (a, b):
                    p.find_marker_cells()
                assert_bitwise(b.reorder_markers(), a.reorder_markers(), "order %d" % s)
            a.step_thermal()
        ens.launch()
        if s % 10 != 9:
            continue
        ctx.sync()
        for i, (a, b) in enumerate(zip(alone, batch)):
            ma, mb = a.get_markers(), b.get_markers()
            for f in ("X", "Y", "VX", "VY", "T", "Tdot", "rho", "wxy", "cellX", "cellY"):
                assert_bitwise(mb[f], ma[f], "step %d instance %d %s" % (s, i, f))
            for o in [p[4] for p in a.w.projections]:
                assert_bitwise(b.get_nodal(o), a.get_nodal(o), "step %d instance %d %s" % (s, i, o))
    ens.close()


def test_reorder_edge_cases():
    """No markers: nothing to do; every marker outside the domain: the order is
    the identity (unlocated markers keep their index order); one cell holding a
    crowded bucket of 300 markers (the one-wavefront ordering path)."""
    w = workloads.convection_small()
    w.markers = w.markers[:0]
    mp = MarkerPath(w)
    mp.find_marker_cells()
    assert len(mp.reorder_markers()) == 0
    w = workloads.convection_small()
    w.markers = w.markers[:500].copy()
    w.markers["X"] = -1.0
    mp = MarkerPath(w)
    assert mp.find_marker_cells() == 500
    assert_bitwise(mp.reorder_markers(), np.arange(500, dtype=np.int32), "all outside")
    w = _scrambled(workloads.convection_small(), seed=4)
    m = w.markers
    a = np.arange(0, 300 * 11, 11)
    m["X"][a] = w.x[7] + 0.2 * (w.x[8] - w.x[7]) * np.linspace(0.1, 0.9, len(a))
    m["Y"][a] = w.y[3] + 0.2 * (w.y[4] - w.y[3]) * np.linspace(0.9, 0.1, len(a))
    _check_reorder(w)
