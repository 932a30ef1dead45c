"""fvm_limit_displacement_timestep_noalloc (no context, no allocation, partials
in a static device array, the step read from and written to device memory)
against fvm_limit_displacement_timestep on the same nodal fields: bitwise,
on two calls in a row, the second on a smaller grid."""
import ctypes as C

import numpy as np
import pytest

from fvmic2d_amd import lib, workloads
from fvmic2d_amd.host import MarkerPath
from fvmic2d_amd.lib import DeviceArray

pytestmark = pytest.mark.gpu


def _noalloc():
    return lib.load().fvm_limit_displacement_timestep_noalloc


def _case(shape, graded, seed, rhodot_scale):
    w = workloads.nodal_box(*shape, graded=graded)
    rng = np.random.default_rng(seed)
    # rhodot nonzero on part of the grid, so the density candidate and the
    # 0.01/max|rhodot| cap both take part
    rd = np.zeros((w.NY, w.NX))
    rd[1::3, ::2] = rhodot_scale * rng.normal(size=rd[1::3, ::2].shape)
    w.nodal["rhodot"] = rd
    return w


@pytest.mark.parametrize("dt0", [1e30, 1e3])
def test_noalloc_equals_limit_displacement_timestep(dt0):
    f = _noalloc()
    results = []
    for shape, graded, seed, rds in (((301, 157), True, 1, 1e-3), ((37, 23), False, 2, 3e-2)):
        print("CASE noalloc dt limit %dx%d graded=%s dt0=%g" % (shape[0], shape[1], graded, dt0), flush=True)
        w = _case(shape, graded, seed, rds)
        mp = MarkerPath(w)
        ref = mp.limit_displacement_timestep(dt0)
        d = DeviceArray(mp.ctx, np.array([dt0]))
        stream = mp.ctx.lib.fvm_ctx_stream(mp.ctx.h)
        rc = f(stream, mp.grid.ref, mp.nodal.ref, C.byref(w.options), C.c_void_p(d.ptr))
        assert rc == 0
        mp.ctx.sync()
        got = float(d.get()[0])
        assert got == ref, (shape, got, ref)
        assert np.isfinite(got) and got <= dt0
        results.append(got)
    assert results[0] != results[1]  # two different problems, two different steps


def test_noalloc_rejects_bad_arguments():
    f = _noalloc()
    w = _case((9, 7), False, 3, 0.0)
    mp = MarkerPath(w)
    d = DeviceArray(mp.ctx, np.array([1.0]))
    stream = mp.ctx.lib.fvm_ctx_stream(mp.ctx.h)
    opt = C.byref(w.options)
    assert f(stream, mp.grid.ref, mp.nodal.ref, opt, None) != 0
    assert f(stream, None, mp.nodal.ref, opt, C.c_void_p(d.ptr)) != 0
    mp.ctx.sync()
    assert float(d.get()[0]) == 1.0  # nothing was launched
