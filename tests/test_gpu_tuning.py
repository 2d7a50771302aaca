"""Every entry of the tuning table is exact: forcing each suitor matching
shape on every level bucket, or extreme values of the two thresholds, gives
the same partitions (single-graph and batched calls, both matchings)."""
import ctypes as C

import numpy as np
import pytest

import rmetis_amd as R
import rmetis_amd.synth as S
from rmetis_amd import _lib

pytestmark = pytest.mark.gpu

NT = 26


def _get(L):
    v = (C.c_int32 * NT)()
    assert L.rmetis_hip_get_tuning(v, NT) == NT
    return list(v)


def _set(L, t):
    assert L.rmetis_hip_set_tuning((C.c_int32 * NT)(*t), NT) == 1, _lib.last_error()


def _graphs():
    out = []
    for s in (S.grid(200, 150), S.rgg(1 << 15, 4, vwgt_max=3), S.rmat(13, 16, 2)):
        out.append(R.Graph(s.n, s.ncon, s.xadj, s.adjncy, s.vwgt, s.adjwgt))
    return out


@pytest.mark.parametrize("ctype", [R.CType.Shem, R.CType.Rm])
def test_suitor_shapes_and_thresholds_do_not_change_partitions(ctype):
    L = _lib.lib()
    base = _get(L)
    graphs = _graphs()
    members = S.uneven_batch(12, 5, nmin_log2=10, nmax_log2=14)
    bg = [R.Graph(g.n, g.ncon, g.xadj, g.adjncy, g.vwgt, g.adjwgt) for g, _ in members]
    bk = [k for _, k in members]
    settings = [[s] * 24 + base[24:] for s in range(4)]
    settings.append(base[:24] + [1, 1])                     # incremental passes and speculation everywhere
    settings.append(base[:24] + [1 << 30, 1 << 30])         # neither
    try:
        ref = None
        for t in settings:
            _set(L, t)
            opts = R.Options(seed=3, rtype=R.RType.Greedy, ctype=ctype)
            single = [R.part_graph_kway(g, 16, options=opts) for g in graphs]
            batch = R.part_graph_kway_batch(bg, bk, options=R.Options(seed=3, rtype=R.RType.Greedy, ctype=ctype,
                                                                      ncuts=2))
            got = [(r.objval, r.part.tobytes()) for r in single + list(batch)]
            if ref is None:
                ref = got
            else:
                assert got == ref, t
    finally:
        _set(L, base)
