"""rmetis_hip_kway_device_noalloc: the k-way device entry point with every
device buffer carved from the library's fixed per-GPU workspace.  Its
results equal rmetis_hip_kway_device's bit for bit (part and objval), over
two calls in a row on the same workspace, the second on a smaller graph;
a graph beyond the workspace is rejected with METIS_ERROR_MEMORY before any
kernel runs, and the entry point works again afterwards."""
import ctypes as C

import numpy as np
import pytest

import rmetis_amd as R
import rmetis_amd.synth as S
from rmetis_amd import _lib

pytestmark = pytest.mark.gpu

i32 = _lib.i32p


def _call(name, g, k, seed, ncuts=1, ubvec=None):
    DA = _lib.DeviceArray
    dx, da = DA(g.xadj), DA(g.adjncy)
    dvw = DA(g.vwgt) if g.vwgt is not None else None
    daw = DA(g.adjwgt) if g.adjwgt is not None else None
    dp = DA(np.full(g.n, -7, np.int32))
    opts = R.Options(rtype=R.RType.Greedy, seed=seed, ncuts=ncuts).to_c_array()
    ub = None if ubvec is None else np.asarray(ubvec, np.float32)
    obj = C.c_int32(-1)
    args = [g.n, g.ncon, dx.ptr, da.ptr, dvw.ptr if dvw else None, daw.ptr if daw else None, k, None,
            ub.ctypes.data_as(_lib.f32p) if ub is not None else None, opts.ctypes.data_as(i32), C.byref(obj), dp.ptr]
    if name == "rmetis_hip_kway_device":
        args.append(None)
    rc = getattr(_lib.lib(), name)(*args)
    return rc, obj.value, dp.numpy()


def _same(g, k, seed, **kw):
    rc1, o1, p1 = _call("rmetis_hip_kway_device", g, k, seed, **kw)
    rc2, o2, p2 = _call("rmetis_hip_kway_device_noalloc", g, k, seed, **kw)
    assert rc1 == 1 and rc2 == 1, _lib.last_error()
    assert o1 == o2 and np.array_equal(p1, p2)
    return o2


def test_noalloc_equals_kway_device_on_two_calls_the_second_smaller():
    # weighted RGG (explicit vwgt and adjwgt), two attempts
    g1 = S.rgg(60000, 3, vwgt_max=3)
    print("first", _same(g1, 16, 1, ncuts=2))
    # smaller, unweighted grid: the workspace still holds the first call's data
    g2 = S.grid(64, 48)
    print("second", _same(g2, 4, 2))
    # two constraints with ubvec, smaller again
    g3, k3 = S._batch_member((5, 9000, 8, 17, 2, 8))
    print("third", _same(g3, k3, 3, ubvec=[1.05, 1.05]))


def test_noalloc_rejects_a_graph_beyond_its_workspace():
    big = S.rgg(1 << 20, 1)
    rc, obj, part = _call("rmetis_hip_kway_device_noalloc", big, 64, 1)
    assert rc == R.METIS_ERROR_MEMORY and "MiB of device workspace" in _lib.last_error()
    assert obj == -1 and (part == -7).all()
    small = S.grid(32, 32)
    _same(small, 4, 1)
