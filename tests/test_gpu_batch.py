"""The batched k-way form: every graph of a batch must get exactly the
partition and objval that a separate part_graph_kway call with the same
options and seed returns (tolerance 0; every per-graph computation is
deterministic and independent of the other graphs of the batch)."""
import numpy as np
import pytest

import rmetis_amd as R
from rmetis_amd import synth

pytestmark = pytest.mark.gpu


def _weighted_grid(rows, cols, seed):
    g = synth.grid(rows, cols)
    rng = np.random.default_rng(seed)
    src = np.repeat(np.arange(g.n), np.diff(g.xadj))
    lo = np.minimum(src, g.adjncy).astype(np.int64)
    hi = np.maximum(src, g.adjncy).astype(np.int64)
    _, inv = np.unique(lo * g.n + hi, return_inverse=True)
    w = rng.integers(1, 11, size=inv.max() + 1).astype(np.int32)
    return R.Graph(g.n, 1, g.xadj, g.adjncy, rng.integers(1, 5, g.n).astype(np.int32), w[inv])


def _rgg(n, seed, ncon=1, deg=12.0):
    import math
    s = synth.rgg(n, seed, r=math.sqrt(deg / (math.pi * n)), wmax=10, vwgt_max=4, ncon=ncon)
    return R.Graph(s.n, ncon, s.xadj, s.adjncy, s.vwgt, s.adjwgt)


def _isolated_mix(seed):
    """Grid rows plus isolated vertices interleaved (rows without entries)."""
    g = synth.grid(12, 20)
    n0 = g.n
    iso = 40
    n = n0 + iso
    # vertex ids: the grid keeps 0..n0-1, isolated vertices are n0..n-1
    xadj = np.concatenate([g.xadj, np.full(iso, g.xadj[-1])]).astype(np.int32)
    rng = np.random.default_rng(seed)
    return R.Graph(n, 1, xadj, g.adjncy, rng.integers(1, 5, n).astype(np.int32), None)


def _batch():
    graphs = [_weighted_grid(20, 30, 1), _rgg(3000, 2), _weighted_grid(64, 64, 3), _rgg(20000, 4, deg=9.0),
              _weighted_grid(10, 10, 5), _isolated_mix(6), _rgg(700, 7, deg=15.0), _weighted_grid(3, 50, 8)]
    ks = [4, 16, 8, 64, 2, 8, 32, 2]
    return graphs, ks


def _alone(graphs, ks, opts_for):
    return [R.part_graph_kway(g, k, options=opts_for(i)) for i, (g, k) in enumerate(zip(graphs, ks))]


def _check(batch, alone):
    assert len(batch) == len(alone)
    for i, (b, a) in enumerate(zip(batch, alone)):
        assert b.objval == a.objval, f"graph {i}: objval {b.objval} vs {a.objval}"
        assert np.array_equal(b.part, a.part), f"graph {i}: part differs at {np.nonzero(b.part != a.part)[0][:5]}"


@pytest.mark.parametrize("ncuts", [1, 3])
def test_batch_equals_alone(ncuts):
    graphs, ks = _batch()
    opts = R.Options.for_kway(seed=11, ncuts=ncuts, rtype=R.RType.Greedy)
    _check(R.part_graph_kway_batch(graphs, ks, options=opts), _alone(graphs, ks, lambda i: opts))


def test_batch_per_graph_seeds_and_rm():
    graphs, ks = _batch()
    seeds = [3 + 7 * i for i in range(len(graphs))]
    base = dict(ncuts=2, rtype=R.RType.Greedy, ctype=R.CType.Rm, niter=4)
    batch = R.part_graph_kway_batch(graphs, ks, options=R.Options.for_kway(seed=1, **base), seeds=seeds)
    _check(batch, _alone(graphs, ks, lambda i: R.Options.for_kway(seed=seeds[i], **base)))


def test_batch_two_constraints_ubvec():
    graphs = [_rgg(5000, 21, ncon=2), _rgg(1200, 22, ncon=2), _rgg(9000, 23, ncon=2, deg=8.0)]
    ks = [16, 8, 32]
    ub = np.array([1.05, 1.05], np.float32)
    opts = R.Options.for_kway(seed=5, ncuts=2, rtype=R.RType.Greedy)
    batch = R.part_graph_kway_batch(graphs, ks, ubvec=ub, options=opts)
    alone = [R.part_graph_kway(g, k, ubvec=ub, options=opts) for g, k in zip(graphs, ks)]
    _check(batch, alone)


def test_batch_random_initial_partition():
    graphs, ks = _batch()
    opts = R.Options.for_kway(seed=9, ncuts=1, rtype=R.RType.Greedy, iptype=R.IpType.Random)
    _check(R.part_graph_kway_batch(graphs, ks, options=opts), _alone(graphs, ks, lambda i: opts))


def test_batch_single_graph_and_order_independence():
    graphs, ks = _batch()
    opts = R.Options.for_kway(seed=2, ncuts=1, rtype=R.RType.Greedy)
    one = R.part_graph_kway_batch([graphs[3]], [ks[3]], options=opts)
    rev = R.part_graph_kway_batch(graphs[::-1], ks[::-1], options=opts)[::-1]
    alone = _alone(graphs, ks, lambda i: opts)
    _check(one, [alone[3]])
    _check(rev, alone)


def test_batch_unweighted_graphs():
    # without vertex weights coarse vertices keep weight 1 (graph/mod.rs:300-314)
    graphs = [_grid_plain(20, 30), _grid_plain(64, 64), _grid_plain(7, 90), _grid_plain(128, 40)]
    ks = [4, 16, 2, 64]
    opts = R.Options.for_kway(seed=4, ncuts=2, rtype=R.RType.Greedy)
    _check(R.part_graph_kway_batch(graphs, ks, options=opts), _alone(graphs, ks, lambda i: opts))
    with pytest.raises(R.InputError):                                   # weighted and unweighted mixed
        R.part_graph_kway_batch([graphs[0], _weighted_grid(10, 10, 1)], [2, 2], options=opts)


def _grid_plain(rows, cols):
    g = synth.grid(rows, cols)
    return R.Graph(g.n, 1, g.xadj, g.adjncy)


def test_batch_errors():
    graphs, ks = _batch()
    opts = R.Options.for_kway(seed=1, rtype=R.RType.Greedy)
    with pytest.raises(R.InputError):
        R.part_graph_kway_batch(graphs, ks[:-1], options=opts)
    bad = list(ks)
    bad[4] = 101                      # graph 4 has 100 vertices
    with pytest.raises(R.InputError):
        R.part_graph_kway_batch(graphs, bad, options=opts)
    with pytest.raises(R.InputError):
        R.part_graph_kway_batch(graphs, ks, options=R.Options.for_kway(seed=1))   # Fm: single-graph calls only
