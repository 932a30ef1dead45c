"""Marker files written and read on the GPU (fvm_save_markers_binary,
fvm_restart_from_markers) against the oracle: byte-identical files, bitwise
equal restarted markers, a save -> restart -> save round trip at 1.44M markers,
and the refusals (wrong layout, capacity)."""
import numpy as np
import pytest

import oracle
from fvmic2d_amd import workloads
from fvmic2d_amd.host import MarkerPath
from fvmic2d_amd.lib import FvmError
from helpers import assert_bitwise
from test_oracle_markerfile import _workload

pytestmark = pytest.mark.gpu


def test_file_is_byte_identical_to_the_oracle(tmp_path):
    print("  case: convection_small with random fields, every 7th marker inactive", flush=True)
    w = _workload()
    st, mp = oracle.State(w), MarkerPath(w)
    st.save_markers_binary(tmp_path / "ref", 42.0)
    n = mp.save_markers_binary(tmp_path / "gpu", 42.0)
    ref = open(tmp_path / "ref", "rb").read()
    assert n == int(np.frombuffer(ref, "<i4", 1)[0])
    assert open(tmp_path / "gpu", "rb").read() == ref


def test_restart_matches_the_oracle(tmp_path):
    w = _workload()
    st = oracle.State(w)
    st.save_markers_binary(tmp_path / "f", 7.25)
    cap = w.n_markers + 5
    t_ref = st.restart_from_markers(tmp_path / "f", cap)
    mp = MarkerPath(w, capacity=cap)
    assert mp.restart_from_markers(tmp_path / "f") == t_ref
    got = mp.markers.download()
    ref = st.markers[:st.ms.nMark]
    assert len(got) == len(ref)
    for f in ref.dtype.names:
        assert_bitwise(got[f], ref[f], f)


def test_round_trip_at_full_size(tmp_path):
    print("  case: convection_tdep 201x201, 1.44M markers", flush=True)
    w = workloads.convection_tdep()
    mp = MarkerPath(w)
    mp.step_thermal()
    before = mp.markers.download()
    n = mp.save_markers_binary(tmp_path / "a", 1.5)
    assert n == int((before["cellX"] != -1).sum())
    mp2 = MarkerPath(w, ctx=mp.ctx)
    assert mp2.restart_from_markers(tmp_path / "a") == 1.5
    after = mp2.markers.download()
    act = before[before["cellX"] != -1]
    for f in ("X", "Y", "VX", "VY", "T", "Tdot", "eta", "mu", "exx", "exy", "Eii", "sxx", "syy", "sxy", "p",
              "rho", "rhodot", "Mat", "cellX", "cellY", "cpu"):
        assert_bitwise(after[f], act[f], f)
    mp2.save_markers_binary(tmp_path / "b", 1.5)
    assert open(tmp_path / "a", "rb").read() == open(tmp_path / "b", "rb").read()


def test_refusals(tmp_path):
    w = _workload()
    mp = MarkerPath(w)
    mp.save_markers_binary(tmp_path / "f", 0.0)
    raw = open(tmp_path / "f", "rb").read()
    open(tmp_path / "short", "wb").write(raw[:-4])
    with pytest.raises(FvmError):
        mp.restart_from_markers(tmp_path / "short")
    with pytest.raises(FvmError):
        mp.restart_from_markers(tmp_path / "missing")
    w_small = _workload()
    w_small.markers = w_small.markers[:2]
    small = MarkerPath(w_small, capacity=3)
    with pytest.raises(FvmError):
        small.restart_from_markers(tmp_path / "f")


def test_empty_and_all_inactive_sets(tmp_path):
    w = _workload()
    w.markers["cellX"][:] = -1
    mp = MarkerPath(w, capacity=w.n_markers + 3)
    assert mp.save_markers_binary(tmp_path / "none", 4.0) == 0
    assert open(tmp_path / "none", "rb").read() == np.int32(0).tobytes() + np.float64(4.0).tobytes()
    assert mp.restart_from_markers(tmp_path / "none") == 4.0
    assert mp.markers.n == 0


NODAL_FILE_ORDER = ["vx", "vy", "p", "rho", "rhodot", "etaN", "etaS", "kThermal", "lastT", "Cp", "muN", "muS",
                    "soxx", "soyy", "soxy", "edotxx", "edotyy", "edotxy", "wxy", "ha"]


def read_petsc_vecs(path):
    """PETSc's binary Vec records as PetscBinaryRead reads them: big-endian int32
    class id (1211214 for a Vec), big-endian int32 length, big-endian doubles."""
    raw = open(path, "rb").read()
    off, out = 0, []
    while off < len(raw):
        cid, n = np.frombuffer(raw, ">i4", 2, off)
        assert cid == 1211214
        out.append(np.frombuffer(raw, ">f8", int(n), off + 8).astype(np.float64))
        off += 8 + 8 * int(n)
    assert off == len(raw)
    return out


@pytest.mark.parametrize("with_ha", [True, False])
def test_nodal_fields_petsc_binary(tmp_path, with_ha):
    w = workloads.nodal_box(37, 23)
    rng = np.random.default_rng(2)
    for n in NODAL_FILE_ORDER[:-1]:
        if n not in w.nodal:
            w.nodal[n] = rng.normal(size=(w.NY, w.NX))
    mp = MarkerPath(w)
    ha = rng.normal(size=(w.NY, w.NX))
    from fvmic2d_amd.lib import DeviceArray
    mp.save_nodal_fields(tmp_path / "n.petscbin", DeviceArray(mp.ctx, ha) if with_ha else None)
    vecs = read_petsc_vecs(tmp_path / "n.petscbin")
    assert len(vecs) == 20
    for name, v in zip(NODAL_FILE_ORDER, vecs):
        want = (ha if with_ha else np.zeros_like(ha)) if name == "ha" else mp.get_nodal(name)
        assert_bitwise(v, np.asarray(want).ravel(), name)
