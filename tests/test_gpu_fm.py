"""k-way FM refinement on the GPU (refine/fm_kway.rs:12-158) against the
oracle's restatement, bit for bit: the part after every pass sequence and the
final cut, over graphs whose moves recompute neighbours in registers, in the
dense part table (many distinct neighbour parts, k above the LDS table) and
with hubs whose moves rebuild the whole queue tree.  Equal gains pop the
smallest vertex first in both (the reference's order among equal gains lives
in its priority-queue crate).  Then the drop-in: RMETIS_PartGraphKway with
NULL options (the reference default RTYPE Fm, types.rs:82) returns METIS_OK,
objval = edge cut, and its cuts are judged against the oracle's FM pipeline
like the greedy ones."""
import ctypes as C

import numpy as np
import pytest

import pyoracle as O
import rmetis_amd as R
import rmetis_amd.synth as S
from graphs import build_undirected, grid_sorted, random_graph
from test_gpu_quality import _geomean, _over_limit

pytestmark = pytest.mark.gpu


def _both(x, a, vw, aw, ncon, part, k, tpw, ub, niter=10):
    g = R.Graph(len(x) - 1, ncon, x, a, vw, aw)
    p_g, c_g = R.fm_kway_refine(g, part, k, tpw, ub, niter)
    p_o, c_o = O.fm_kway_refine(ncon, x, a, vw, aw, part, k, tpw, ub, niter)
    assert c_g == c_o and np.array_equal(p_g, p_o)
    return c_o


def test_fm_reference_case():
    # fm_kway.rs:209-219: grid3x3, k = 3, row stripes, tpwgts 3, ub 1.1
    x, a = grid_sorted(3, 3)
    cut = _both(x, a, None, None, 1, np.arange(9) // 3, 3, [3.0] * 3, [1.1], 5)
    assert cut <= O.edge_cut(x, a, None, np.arange(9) // 3)


@pytest.mark.parametrize("k,ncon,wmax", [(2, 1, 1), (5, 1, 7), (16, 2, 3), (64, 1, 10)])
def test_fm_random_graphs(k, ncon, wmax):
    rng = np.random.default_rng(k * 10 + ncon)
    for trial in range(6):
        n = int(rng.integers(200, 3000))
        x, a, aw = random_graph(n, float(rng.choice([0.004, 0.01, 0.03])), 700 + trial, wmax=wmax)
        vw = rng.integers(1, 5, n * ncon).astype(np.int32) if trial % 2 else None
        part = rng.integers(0, k, n).astype(np.int32)
        tot = (vw.reshape(n, ncon).sum(0) if vw is not None else np.full(ncon, n)).astype(np.float64)
        _both(x, a, vw, aw, ncon, part, k, np.tile(tot / k, k), np.full(ncon, 1.05))


def test_fm_grids_and_rgg_banded():
    g = S.rgg(40000, 3)
    k = 32
    part = ((np.arange(g.n) * k) // g.n).astype(np.int32)        # Morton bands: a good start, few moves
    _both(g.xadj, g.adjncy, g.vwgt, g.adjwgt, 1, part, k, np.full(k, g.n / k), [1.03])
    x, a = grid_sorted(120, 100)
    part = (np.arange(12000) % 100 // 25 + 4 * (np.arange(12000) // 100 // 30)).astype(np.int32)
    _both(x, a, None, None, 1, part, 16, np.full(16, 750.0), [1.03])


def test_fm_many_distinct_parts_and_large_k():
    # dense rows with neighbours in many parts (the dense-table recompute),
    # and k above the LDS table (its global copy)
    x, a, aw = random_graph(1500, 0.05, 9, wmax=4)
    rng = np.random.default_rng(1)
    part = rng.integers(0, 40, 1500).astype(np.int32)
    _both(x, a, None, aw, 1, part, 40, np.full(40, 1500 / 40), [1.2])
    x, a, aw = random_graph(6000, 0.002, 10, wmax=3)
    k = 5000
    part = rng.integers(0, k, 6000).astype(np.int32)
    _both(x, a, None, aw, 1, part, k, np.full(k, 6000 / k), [2.0], niter=3)


def test_fm_hubs():
    # a hub with thousands of neighbours moves: the whole queue tree is rebuilt
    rng = np.random.default_rng(2)
    n = 9000
    edges = [(0, i) for i in range(1, 5000)] + [(i, i + 1) for i in range(1, n - 1)]
    x, a = build_undirected(n, edges)
    part = (np.arange(n) * 4 // n).astype(np.int32)
    part[0] = 3
    _both(x, a, None, None, 1, part, 4, np.full(4, n / 4), [1.5])
    g = S.rmat(13, 16, 4)
    part = rng.integers(0, 8, g.n).astype(np.int32)
    _both(g.xadj, g.adjncy, None, None, 1, part, 8, np.full(8, g.n / 8), [1.05])


def _fm_quality(g, k, seeds=range(1, 9)):
    rg = R.Graph(g.n, g.ncon, g.xadj, g.adjncy, g.vwgt, g.adjwgt)
    gc, oc, gov, oov = [], [], 0, 0
    for s in seeds:
        r = R.part_graph_kway(rg, k, options=R.Options(seed=s))            # rtype Fm: the default
        assert r.objval == rg.edge_cut(r.part)
        rc, obj, opart = O.part_graph_kway(g.xadj, g.adjncy, k, vwgt=g.vwgt, adjwgt=g.adjwgt, seed=s,
                                           rtype_greedy=False)
        assert rc == 1
        gc.append(r.objval)
        oc.append(obj)
        gov += _over_limit(g, r.part, k, None)
        oov += _over_limit(g, opart, k, None)
    return gc, oc, gov, oov


def test_kway_default_options_run_fm():
    # NULL options parse to Options::default(), RTYPE Fm (c_api.rs:211-222)
    from rmetis_amd import _lib
    s = S.grid(40, 40)
    n, nc, k = C.c_int32(s.n), C.c_int32(1), C.c_int32(4)
    part = np.zeros(s.n, np.int32)
    obj = C.c_int32()
    rc = _lib.lib().RMETIS_PartGraphKway(C.byref(n), C.byref(nc), s.xadj.ctypes.data_as(_lib.i32p),
                                         s.adjncy.ctypes.data_as(_lib.i32p), None, None, None, C.byref(k), None,
                                         None, None, C.byref(obj), part.ctypes.data_as(_lib.i32p))
    assert rc == 1 and 0 <= part.min() and part.max() < 4
    assert obj.value == O.edge_cut(s.xadj, s.adjncy, None, part)


@pytest.mark.parametrize("name", ["grid", "rgg", "batch_member"])
def test_fm_pipeline_quality(name):
    if name == "grid":
        g, k = S.grid(128, 128), 8
    elif name == "rgg":
        g, k = S.rgg(1 << 15, 2), 32
    else:
        g, k = S._batch_member((1, 20000, 16, 55))
    # FM (fm_kway.rs) only makes moves that keep their target part within
    # its limit and has no repair pass: parts over the limit are left from
    # the initial partition of the coarsest level, which differs from the
    # oracle's (the GPU matching builds different coarse levels).  Cuts are
    # held to the 3 % bar; over-limit counts only to the oracle's order.
    gc, oc, gov, oov = _fm_quality(g, k)
    print(f"FM {name}: GPU geomean cut {_geomean(gc):.1f}, oracle {_geomean(oc):.1f}, over-limit {gov} vs {oov}")
    assert _geomean(gc) <= 1.03 * _geomean(oc)
    assert gov <= 2 * oov + 4
