"""The served form: requests submitted to a Server while others run are
batched with whatever is pending, and each result must equal what
part_graph_kway returns for that request alone (tolerance 0: every
per-graph computation is deterministic and independent of its batch)."""
import math
import time

import numpy as np
import pytest

import rmetis_amd as R
from rmetis_amd import synth

pytestmark = pytest.mark.gpu


def _rgg(n, seed, ncon=1, deg=12.0, vmax=4):
    s = synth.rgg(n, seed, r=math.sqrt(deg / (math.pi * n)), wmax=10, vwgt_max=vmax, ncon=ncon)
    return R.Graph(s.n, ncon, s.xadj, s.adjncy, s.vwgt, s.adjwgt)


def _grid(rows, cols):
    g = synth.grid(rows, cols)
    return R.Graph(g.n, 1, g.xadj, g.adjncy)


def _opts(seed, **kw):
    return R.Options.for_kway(seed=seed, rtype=R.RType.Greedy, **kw)


def _check(served, alone):
    assert served.objval == alone.objval
    assert np.array_equal(served.part, alone.part)


def test_served_equal_alone_with_overlapping_arrivals():
    reqs = [(_rgg(60000, 1), 64, 3), (_grid(40, 50), 8, 4), (_rgg(3000, 2), 16, 5), (_rgg(20000, 3, deg=9), 32, 6),
            (_grid(8, 9), 2, 7), (_rgg(700, 4, deg=15), 4, 8), (_rgg(120000, 5), 128, 9), (_grid(100, 30), 64, 10)]
    with R.Server(slots=2) as srv:
        ids = []
        for i, (g, k, seed) in enumerate(reqs):
            ids.append(srv.submit(g, k, options=_opts(1), seed=seed))
            if i % 3 == 0:
                time.sleep(0.003)        # later requests arrive while earlier ones run
        res = [srv.result(rid) for rid in ids]
        st = srv.stats()
    assert st["served"] == len(reqs) and st["batches"] >= 2       # the unweighted grids batch apart
    for (g, k, seed), r in zip(reqs, res):
        _check(r, R.part_graph_kway(g, k, options=_opts(seed)))
        assert r.total_ms >= r.running_ms >= 0.0 and r.batch_graphs >= 1


def test_served_setups_batch_apart():
    # different ncon, ubvec and niter never share a batch; each equals alone
    a = [(_rgg(5000, 11, ncon=2, vmax=8), 32, np.array([1.05, 1.05], np.float32), _opts(21)),
         (_rgg(9000, 12, ncon=2, vmax=8), 512, np.array([1.05, 1.05], np.float32), _opts(22)),
         (_rgg(4000, 13), 8, None, _opts(23)),
         (_rgg(4000, 14), 8, None, _opts(24, niter=3)),
         (_rgg(6000, 15, ncon=2, vmax=8), 256, np.array([1.05, 1.05], np.float32), _opts(25))]
    with R.Server(slots=3) as srv:
        ids = [srv.submit(g, k, ubvec=ub, options=o) for g, k, ub, o in a]
        res = [srv.result(rid) for rid in ids]
    for (g, k, ub, o), r in zip(a, res):
        _check(r, R.part_graph_kway(g, k, ubvec=ub, options=o))


def test_served_batches_overlap_on_two_slots():
    # two incompatible requests submitted together run as two batches at once
    g1, g2 = _rgg(150000, 31), _rgg(150000, 32, ncon=2, vmax=8)
    with R.Server(slots=2) as srv:
        i1 = srv.submit(g1, 64, options=_opts(3))
        i2 = srv.submit(g2, 64, ubvec=[1.05, 1.05], options=_opts(3))
        r1, r2 = srv.result(i1), srv.result(i2)
        st = srv.stats()
    assert st["batches"] == 2 and st["max_inflight_batches"] == 2
    _check(r1, R.part_graph_kway(g1, 64, options=_opts(3)))
    _check(r2, R.part_graph_kway(g2, 64, ubvec=[1.05, 1.05], options=_opts(3)))


def test_server_errors():
    g = _grid(4, 4)
    with R.Server(slots=1) as srv:
        with pytest.raises(R.InputError):
            srv.submit(g, 17, options=_opts(1))                   # nparts > nvtxs
        with pytest.raises(R.InputError):
            srv.submit(g, 1, options=_opts(1))                    # nparts < 2
        with pytest.raises(R.InputError):
            srv.submit(g, 2, options=R.Options.for_kway(seed=1))  # Fm refinement
        rid = srv.submit(g, 2, options=_opts(1))
        r = srv.result(rid)
        assert r.part.shape == (16,)
        with pytest.raises(R.InputError):
            srv.result(rid)                                        # already collected


def test_served_device_requests_equal_alone():
    # graphs already in HBM: submitted by device pointer, part written in place
    from rmetis_amd import _lib
    DA = _lib.DeviceArray
    graphs = [(_rgg(30000, 41, ncon=2, vmax=8), 128, 7), (_rgg(2000, 42, ncon=2, vmax=8), 8, 8),
              (_rgg(70000, 43, ncon=2, vmax=8), 512, 9)]
    ub = np.array([1.05, 1.05], np.float32)
    dev = [(DA(g.xadj), DA(g.adjncy), DA(g.vwgt), DA(g.adjwgt), DA(count=g.nvtxs, dtype=np.int32))
           for g, _, _ in graphs]
    with R.Server(slots=2) as srv:
        ids = [srv.submit_device(g.nvtxs, g.ncon, *d[:4], k, d[4], ubvec=ub, options=_opts(1), seed=s)
               for (g, k, s), d in zip(graphs, dev)]
        host = srv.submit(graphs[1][0], 16, ubvec=ub, options=_opts(2))       # a host request in the mix
        res = [srv.result(rid) for rid in ids]
        hres = srv.result(host)
    for (g, k, s), d, r in zip(graphs, dev, res):
        alone = R.part_graph_kway(g, k, ubvec=ub, options=_opts(s))
        assert r.part is None and r.objval == alone.objval
        assert np.array_equal(d[4].numpy(), alone.part)
    _check(hres, R.part_graph_kway(graphs[1][0], 16, ubvec=ub, options=_opts(2)))


def test_served_stream_on_four_slots():
    # four slots (batched calls in flight at once, one stream each) under a
    # stream of mixed arrivals; results are waited for with a time limit so
    # that a stall fails the test with every slot's phase instead of hanging
    rng = np.random.default_rng(5)
    pool = [(_rgg(int(n), 60 + i, ncon=2, vmax=8), 2 ** int(rng.integers(3, 9)))
            for i, n in enumerate(rng.integers(2000, 90000, 12))]
    ub = np.array([1.05, 1.05], np.float32)
    reqs = [(pool[int(rng.integers(len(pool)))], 100 + j) for j in range(48)]
    with R.Server(slots=4) as srv:
        ids = []
        for j, ((g, k), seed) in enumerate(reqs):
            ids.append(srv.submit(g, min(k, g.nvtxs), ubvec=ub, options=_opts(1), seed=seed))
            if j % 4 == 3:
                time.sleep(0.002)
        res = [srv.result(rid, timeout_ms=120000) for rid in ids]
        st = srv.stats()
    assert st["served"] == len(reqs) and st["max_inflight_batches"] >= 2
    for ((g, k), seed), r in zip(reqs, res):
        _check(r, R.part_graph_kway(g, min(k, g.nvtxs), ubvec=ub, options=_opts(seed)))
