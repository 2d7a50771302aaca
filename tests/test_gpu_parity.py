"""GPU parity: every hot-path function of librmetis_hip.so against the oracle
(a CPU restatement of javagg/rmetis) on the same seeded inputs.  Integer
results must be bit-exact."""
import numpy as np
import pytest

import pyoracle as O
import rmetis_amd as R
from graphs import cycle4, grid_parallel_order, grid_sorted, path, random_graph

pytestmark = pytest.mark.gpu


def _graph(xadj, adjncy, vwgt=None, adjwgt=None, ncon=1, vsize=None):
    return R.Graph(len(xadj) - 1, ncon, xadj, adjncy, vwgt, adjwgt, vsize)


def _cases():
    """(name, xadj, adjncy, adjwgt) small/medium graphs."""
    out = [("cycle4",) + cycle4() + (None,), ("path9",) + path(9) + (None,),
           ("grid16x16",) + grid_sorted(16, 16) + (None,),
           ("gridpar12x7",) + grid_parallel_order(12, 7) + (None,)]
    for i, (n, p, w) in enumerate([(60, 0.1, 1), (300, 0.03, 10), (2000, 0.004, 5), (1500, 0.02, 3)]):
        x, a, aw = random_graph(n, p, 100 + i, wmax=w)
        out.append((f"er{n}", x, a, aw))
    return out


CASES = _cases()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("k", [2, 5, 64])
def test_edge_cut(case, k):
    _, x, a, aw = case
    n = len(x) - 1
    rng = np.random.default_rng(k)
    for _ in range(3):
        part = rng.integers(0, k, n).astype(np.int32)
        g = _graph(x, a, adjwgt=aw)
        assert g.edge_cut(part) == O.edge_cut(x, a, aw, part)


def test_edge_cut_hubs_and_batches():
    # hubs of more than 1,024 entries (summed by a whole wave), tiles whose
    # entries span several staged batches, isolated vertices, large weights
    # (the i32 wrap of the reference's sum)
    rng = np.random.default_rng(3)
    n = 9000
    edges = set()
    for hub, cnt in ((0, 3000), (4500, 1500), (8999, 1100)):
        for u in rng.choice(n, cnt, replace=False):
            if u != hub:
                edges.add((min(hub, u), max(hub, u)))
    for v in range(100, 4000):                       # dense rows: several batches per tile
        for d in (1, 2, 3, 5, 8, 13, 21, 34, 55):
            if v + d < 4400:
                edges.add((v, v + d))
    src = np.array([e[0] for e in edges] + [e[1] for e in edges], np.int64)
    dst = np.array([e[1] for e in edges] + [e[0] for e in edges], np.int64)
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    x = np.zeros(n + 1, np.int32)
    np.add.at(x, src + 1, 1)
    x = np.cumsum(x).astype(np.int32)
    a = dst.astype(np.int32)
    lo, hi = np.minimum(src, dst), np.maximum(src, dst)
    _, inv = np.unique(lo * n + hi, return_inverse=True)
    aw = rng.integers(1, 1 << 28, inv.max() + 1).astype(np.int32)[inv]
    for k in (2, 37):
        part = rng.integers(0, k, n).astype(np.int32)
        assert _graph(x, a, adjwgt=aw).edge_cut(part) == O.edge_cut(x, a, aw, part)
        assert _graph(x, a).edge_cut(part) == O.edge_cut(x, a, None, part)


def test_edge_cut_reference_values():
    x, a = cycle4()
    g = _graph(x, a)
    assert g.edge_cut([0, 0, 1, 1]) == 2          # graph/mod.rs:455-461
    # 256x256 grid split into k horizontal bands: cut = (k-1)*cols
    from rmetis_amd import synth
    s = synth.grid(256, 256)
    g = R.Graph(s.n, 1, s.xadj, s.adjncy)
    for k in (2, 8, 64):
        part = (np.arange(s.n) // 256 * k // 256).astype(np.int32)
        assert g.edge_cut(part) == (k - 1) * 256


@pytest.mark.parametrize("ncon", [1, 2, 3])
@pytest.mark.parametrize("k", [2, 7, 300, 9000])
def test_partition_weights(ncon, k):
    rng = np.random.default_rng(ncon * 1000 + k)
    n = 5000
    x, a = path(n)
    part = rng.integers(0, k, n).astype(np.int32)
    vw = rng.integers(1, 9, n * ncon).astype(np.int32)
    g = _graph(x, a, vwgt=vw, ncon=ncon)
    assert np.array_equal(g.partition_weights(part, k), O.partition_weights(ncon, vw, part, k))
    g1 = _graph(x, a, ncon=ncon)
    assert np.array_equal(g1.partition_weights(part, k), O.partition_weights(ncon, None, part, k))


def _check_gain_table(x, a, aw, part, k):
    g = _graph(x, a, adjwgt=aw)
    rec, g0, cp, cg = g.gain_table(part, k)
    n = len(x) - 1
    bnd = O.boundary_vertices(x, a, part)
    assert np.array_equal(np.nonzero(rec[:, 0] >= 0)[0], bnd)
    for v in bnd:
        pv = part[v]
        nc = rec[v, 2]
        nb = rec[v, 3] & 0x7fffffff
        lst = [(int(cp[x[v] + i]), int(cg[x[v] + i])) for i in range(nc)]
        expl = {p for p, _ in lst}
        for p in range(k):
            if p == pv:
                continue
            og = O.move_gain(x, a, aw, v, part, pv, p)
            got = dict(lst).get(p, g0[v])
            assert got == og, (v, p)
            if p not in expl:
                assert og == g0[v]
        # admission order: gain descending, ties to the lowest part
        assert lst == sorted(lst, key=lambda t: (-t[1], t[0]))
        assert nb == sum(1 for _, gg in lst if gg > g0[v])
    return rec


@pytest.mark.parametrize("case", CASES[:6], ids=[c[0] for c in CASES[:6]])
@pytest.mark.parametrize("k", [2, 4, 9])
def test_gain_table(case, k):
    _, x, a, aw = case
    n = len(x) - 1
    rng = np.random.default_rng(7 * k)
    part = rng.integers(0, k, n).astype(np.int32)
    _check_gain_table(x, a, aw, part, k)


def test_gain_table_heavy_vertices():
    # a star hub (degree 900) plus a dense block: exercises the block-per-vertex path
    n = 1000
    edges = [(0, i) for i in range(1, 901)] + [(i, i + 1) for i in range(901, 999)]
    from graphs import build_undirected
    x, a = build_undirected(n, edges)
    rng = np.random.default_rng(3)
    aw_u = rng.integers(1, 6, len(a) // 2)
    # symmetric weights
    src = np.repeat(np.arange(n), np.diff(x))
    key = np.minimum(src, a).astype(np.int64) * n + np.maximum(src, a)
    _, inv = np.unique(key, return_inverse=True)
    aw = aw_u[inv].astype(np.int32)
    for k in (3, 40, 700):
        part = rng.integers(0, k, n).astype(np.int32)
        _check_gain_table(x, a, aw, part, k)


def _balanced_targets(n_total, k, ncon, ub):
    tpw = np.full(k * ncon, n_total / k, np.float64)
    return tpw, np.full(ncon, ub, np.float64)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("k,ub,skew", [(2, 1.03, 0.0), (4, 1.001, 0.6), (8, 1.05, 0.3), (16, 1.0, 0.9)])
def test_greedy_pass(case, k, ub, skew):
    _, x, a, aw = case
    n = len(x) - 1
    rng = np.random.default_rng(k * 31 + n)
    # skewed starting partition so that the repair pass has work
    part = rng.integers(0, k, n).astype(np.int32)
    part[rng.random(n) < skew] = 0
    vw = rng.integers(1, 4, n).astype(np.int32)
    g = _graph(x, a, vwgt=vw, adjwgt=aw)
    tpw, ubv = _balanced_targets(vw.sum(), k, 1, ub)
    pw = O.partition_weights(1, vw, part, k)
    for mode in (0, 1, 1, 1):
        p_o, pw_o, mv_o = O.greedy_pass(1, x, a, vw, aw, part, k, tpw, ubv, mode, pw)
        p_g, pw_g, mv_g = R.greedy_pass(g, part, k, tpw, ubv, mode, pw)
        assert np.array_equal(mv_g, mv_o), mode
        assert np.array_equal(p_g, p_o)
        assert np.array_equal(pw_g, pw_o)
        part, pw = p_o, pw_o


@pytest.mark.parametrize("k,ncon,wmax", [
    (64, 1, 3),      # register slack, one block; speculation with 2 weight planes
    (64, 2, 200),    # two constraints, 8 weight planes
    (100, 1, 255),   # two register blocks
    (200, 2, 3),     # four register blocks, two constraints
    (64, 1, 1000),   # weights past the speculation limit: exact walk only
    (300, 1, 5),     # slack table in LDS
    (100, 3, 4),     # three constraints: table walk
    (1500, 3, 2),    # k*ncon > 4096: tables in global memory
])
def test_greedy_pass_admission_variants(k, ncon, wmax):
    # every admission variant (register / LDS / global slack, speculation on
    # and off) against the oracle, on a grid whose vertices are nearly all
    # boundary movers
    x, a = grid_sorted(60, 60)
    n = len(x) - 1
    rng = np.random.default_rng(k * 7 + ncon)
    aw = rng.integers(1, 6, len(a)).astype(np.int32)
    # symmetric edge weights: weight of (u, v) from the smaller endpoint's entry
    src = np.repeat(np.arange(n), np.diff(x))
    key = {}
    for e in range(len(a)):
        u, v = int(src[e]), int(a[e])
        key.setdefault((min(u, v), max(u, v)), int(aw[e]))
        aw[e] = key[(min(u, v), max(u, v))]
    vw = rng.integers(1, wmax + 1, n * ncon).astype(np.int32)
    part = rng.integers(0, k, n).astype(np.int32)
    part[rng.random(n) < 0.3] = 0
    tot = vw.reshape(n, ncon).sum(0).astype(np.float64)
    tpw = np.tile(tot / k, k)
    ubv = np.full(ncon, 1.05)
    g = _graph(x, a, vwgt=vw, adjwgt=aw, ncon=ncon)
    pw = O.partition_weights(ncon, vw, part, k)
    for mode in (0, 1, 1, 1):
        p_o, pw_o, mv_o = O.greedy_pass(ncon, x, a, vw, aw, part, k, tpw, ubv, mode, pw)
        p_g, pw_g, mv_g = R.greedy_pass(g, part, k, tpw, ubv, mode, pw)
        assert np.array_equal(mv_g, mv_o), (mode, len(mv_g), len(mv_o))
        assert np.array_equal(p_g, p_o)
        assert np.array_equal(pw_g, pw_o)
        part, pw = p_o, pw_o


def test_greedy_pass_with_hubs():
    # tiles mixing heavy (block-per-vertex) and ordinary vertices
    from graphs import build_undirected
    rng = np.random.default_rng(8)
    n = 3000
    edges = set()
    for hub in (5, 700, 2900):
        for u in rng.choice(n, 400, replace=False):
            if u != hub:
                edges.add((min(hub, u), max(hub, u)))
    for _ in range(6000):
        u, v = rng.integers(0, n, 2)
        if u != v:
            edges.add((min(u, v), max(u, v)))
    x, a = build_undirected(n, sorted(edges))
    src = np.repeat(np.arange(n), np.diff(x))
    key = np.minimum(src, a).astype(np.int64) * n + np.maximum(src, a)
    _, inv = np.unique(key, return_inverse=True)
    aw = rng.integers(1, 7, inv.max() + 1)[inv].astype(np.int32)
    vw = rng.integers(1, 4, n).astype(np.int32)
    g = _graph(x, a, vwgt=vw, adjwgt=aw)
    for k in (4, 37):
        part = rng.integers(0, k, n).astype(np.int32)
        part[: n // 4] = 1
        tpw = np.full(k, vw.sum() / k)
        ub = np.array([1.02])
        pw = O.partition_weights(1, vw, part, k)
        for mode in (0, 1, 1):
            p_o, pw_o, mv_o = O.greedy_pass(1, x, a, vw, aw, part, k, tpw, ub, mode, pw)
            p_g, pw_g, mv_g = R.greedy_pass(g, part, k, tpw, ub, mode, pw)
            assert np.array_equal(mv_g, mv_o) and np.array_equal(p_g, p_o) and np.array_equal(pw_g, pw_o)
            part, pw = p_o, pw_o


@pytest.mark.parametrize("ncon", [1, 2])
@pytest.mark.parametrize("k", [3, 8, 33])
def test_greedy_refine(ncon, k):
    x, a, aw = random_graph(3000, 0.003, 77 + k, wmax=7)
    n = len(x) - 1
    rng = np.random.default_rng(k)
    vw = rng.integers(1, 5, n * ncon).astype(np.int32)
    part = rng.integers(0, k, n).astype(np.int32)
    part[: n // 3] = 0
    tot = vw.reshape(n, ncon).sum(0)
    tpw = np.tile(tot / k, k).astype(np.float64)
    ub = np.full(ncon, 1.03)
    g = _graph(x, a, vwgt=vw, adjwgt=aw, ncon=ncon)
    p_o, cut_o = O.greedy_refine(ncon, x, a, vw, aw, part, k, tpw, ub, 10)
    p_g, cut_g = R.greedy_refine(g, part, k, tpw, ub, 10)
    assert cut_g == cut_o
    assert np.array_equal(p_g, p_o)


def test_greedy_refine_large_level_incremental():
    # a level of at least 2^18 vertices runs its later cut passes
    # incrementally (only the tiles around the previous pass's moves): the
    # result must still equal the oracle's full passes
    from rmetis_amd import synth
    s = synth.grid(640, 440)
    n = s.n
    rng = np.random.default_rng(5)
    src = np.repeat(np.arange(n), np.diff(s.xadj))
    lo = np.minimum(src, s.adjncy).astype(np.int64)
    hi = np.maximum(src, s.adjncy).astype(np.int64)
    _, inv = np.unique(lo * n + hi, return_inverse=True)
    aw = rng.integers(1, 9, size=inv.max() + 1).astype(np.int32)[inv]
    vw = rng.integers(1, 4, n).astype(np.int32)
    k = 16
    part = (np.arange(n) * k // n).astype(np.int32)
    flip = rng.random(n) < 0.2
    part[flip] = rng.integers(0, k, int(flip.sum()))
    tpw = np.full(k, vw.sum() / k)
    ub = np.array([1.03])
    g = _graph(s.xadj, s.adjncy, vwgt=vw, adjwgt=aw)
    p_o, cut_o = O.greedy_refine(1, s.xadj, s.adjncy, vw, aw, part, k, tpw, ub, 10)
    p_g, cut_g = R.greedy_refine(g, part, k, tpw, ub, 10)
    assert n >= 1 << 18
    assert cut_g == cut_o
    assert np.array_equal(p_g, p_o)


def test_greedy_reference_cases():
    # tests/parallel.rs:91-116: cycle4 with tpwgts [2,2], ub 1.03
    x, a = cycle4()
    g = _graph(x, a)
    for init, cut in (([0, 0, 1, 1], 2), ([0, 1, 0, 1], 4), ([0, 1, 1, 0], 2)):
        p, c = R.greedy_refine(g, init, 2, [2.0, 2.0], [1.03], 10)
        assert c == cut and list(p) == init
    # tests/parallel.rs:118-132: grid4x4 row stripes, ub 1.05 -> no cut increase
    x, a = grid_parallel_order(4, 4)
    g = _graph(x, a)
    part = np.arange(16) // 4
    before = g.edge_cut(part)
    p, c = R.greedy_refine(g, part, 4, [4.0] * 4, [1.05], 10)
    assert c <= before
    p_o, c_o = O.greedy_refine(1, x, a, None, None, part, 4, [4.0] * 4, [1.05], 10)
    assert c == c_o and np.array_equal(p, p_o)


def _coarse_equal(x, a, vw, aw, vs, ncon, match):
    g = _graph(x, a, vwgt=vw, adjwgt=aw, ncon=ncon, vsize=vs)
    cg, cmap = g.build_coarse_graph(match)
    o = O.build_coarse_graph(ncon, x, a, vw, aw, vs, match)
    assert cg.nvtxs == o["n"]
    assert np.array_equal(cmap, o["cmap"])
    assert np.array_equal(cg.xadj, o["cxadj"])
    assert np.array_equal(cg.adjncy, o["cadjncy"])
    cw = cg.adjwgt if cg.adjwgt is not None else np.ones(len(cg.adjncy), np.int32)
    assert np.array_equal(cw, o["cadjwgt"])
    assert (cg.adjwgt is not None) == o["adjwgt_present"]
    assert (cg.vwgt is None) == (o["cvwgt"] is None)
    if cg.vwgt is not None:
        assert np.array_equal(cg.vwgt, o["cvwgt"])
    assert (cg.vsize is None) == (o["cvsize"] is None)
    if cg.vsize is not None:
        assert np.array_equal(cg.vsize, o["cvsize"])


def test_coarse_reference_path():
    # graph/mod.rs:499-514: path 0-1-2-3 matched (0,1),(2,3)
    x, a = path(4)
    g = _graph(x, a)
    cg, cmap = g.build_coarse_graph([1, 0, 3, 2])
    assert cg.nvtxs == 2 and cg.nedges == 1
    assert list(cmap) == [0, 0, 1, 1]
    assert list(cg.xadj) == [0, 1, 2] and list(cg.adjncy) == [1, 0]
    assert cg.adjwgt is None and cg.vwgt is None


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("ncon,weighted", [(1, False), (1, True), (2, True)])
def test_build_coarse_graph(case, ncon, weighted):
    _, x, a, aw = case
    n = len(x) - 1
    rng = np.random.default_rng(n + ncon)
    vw = rng.integers(1, 6, n * ncon).astype(np.int32) if weighted else None
    vs = rng.integers(1, 4, n).astype(np.int32) if weighted else None
    match = O.shem_matching(x, a, aw, O.rng(n))
    _coarse_equal(x, a, vw, aw, vs, ncon, match)
    match = O.random_matching(x, a, O.rng(n + 1))
    _coarse_equal(x, a, vw, aw, vs, ncon, match)


@pytest.mark.parametrize("rows,cols", [(64, 63), (128, 128), (128, 129), (257, 256)])
def test_build_coarse_graph_scan_sizes(rows, cols):
    # the contraction's prefix sums: one workgroup in one launch up to 16,384
    # elements (4,032 and 16,384 vertices here), three launches above (16,512
    # and 65,792)
    x, a = grid_sorted(rows, cols)
    n = len(x) - 1
    rng = np.random.default_rng(n)
    vw = rng.integers(1, 6, n).astype(np.int32)
    aw = rng.integers(1, 9, len(a)).astype(np.int32)
    aw = _symmetric_weights(x, a, aw)
    match = O.shem_matching(x, a, aw, O.rng(n))
    _coarse_equal(x, a, vw, aw, None, 1, match)


@pytest.mark.parametrize("seed", [1, 2])
def test_build_coarse_graph_degree_mix(seed):
    # degree sums of coarse vertices spread over 1..~500 within the same
    # 64-vertex tiles: LDS batches of one to dozens of vertices, a vertex
    # filling a batch alone, and the global-hash path above 256; neighbours
    # drawn mostly from a window around the vertex, so that many entries merge
    from graphs import build_undirected
    rng = np.random.default_rng(seed)
    n = 3000
    deg = np.where(rng.random(n) < 0.05, rng.integers(100, 420, n), rng.integers(1, 12, n))
    edges = set()
    for v in range(n):
        for _ in range(int(deg[v])):
            u = int(v + rng.integers(-150, 151)) if rng.random() < 0.8 else int(rng.integers(0, n))
            if 0 <= u < n and u != v:
                edges.add((min(u, v), max(u, v)))
    x, a = build_undirected(n, sorted(edges))
    aw = _symmetric_weights(x, a, rng.integers(1, 9, len(a)).astype(np.int32))
    assert np.diff(x).max() > 256
    for m in (O.shem_matching(x, a, aw, O.rng(seed)), O.random_matching(x, a, O.rng(seed + 7))):
        _coarse_equal(x, a, None, aw, None, 1, m)
        _coarse_equal(x, a, None, None, None, 1, m)


def _symmetric_weights(x, a, aw):
    """Edge weights with w(u,v) == w(v,u): the weight of the smaller-id end's entry."""
    n = len(x) - 1
    src = np.repeat(np.arange(n), np.diff(x))
    key = {}
    for i in range(len(a)):
        u, v = int(src[i]), int(a[i])
        if u < v:
            key[(u, v)] = int(aw[i])
    return np.array([key[(min(int(src[i]), int(a[i])), max(int(src[i]), int(a[i])))] for i in range(len(a))],
                    np.int32)


def test_build_coarse_graph_hub():
    # degree sums far above the per-group LDS capacity -> global-hash path
    n = 3000
    from graphs import build_undirected
    edges = [(0, i) for i in range(1, 2500)] + [(1, i) for i in range(2, 1800)] + \
            [(i, i + 1) for i in range(2500, 2999)]
    x, a = build_undirected(n, edges)
    for seed in (1, 2):
        match = O.shem_matching(x, a, None, O.rng(seed))
        _coarse_equal(x, a, None, None, None, 1, match)


def test_gpu_matching_properties():
    from rmetis_amd import synth
    s = synth.rgg(1 << 14, 5)
    g = R.Graph(s.n, 1, s.xadj, s.adjncy, s.vwgt, s.adjwgt)
    m1 = R.gpu_matching(g, 99)
    m2 = R.gpu_matching(g, 99)
    assert np.array_equal(m1, m2)                      # deterministic
    assert np.array_equal(m1[m1], np.arange(s.n))      # symmetric (shem.rs:96-99)
    src = np.repeat(np.arange(s.n), np.diff(s.xadj))
    pairs = m1 != np.arange(s.n)
    # matched partners are neighbours
    nbr = set(zip(src.tolist(), s.adjncy.tolist()))
    assert all((v, int(m1[v])) in nbr for v in np.nonzero(pairs)[0])
    assert pairs.mean() > 0.7
    m3 = R.gpu_matching(g, 100)
    assert not np.array_equal(m1, m3)                  # seeded


_M32 = 0xFFFFFFFF


def _mix32(x):
    x &= _M32
    x ^= x >> 16
    x = (x * 0x7FEB352D) & _M32
    x ^= x >> 15
    x = (x * 0x846CA68B) & _M32
    x ^= x >> 16
    return x


def _degree_rank(d):
    d = max(d, 1)
    return min(d, 4) + max(0, d.bit_length() - 1 - 2)


def _greedy_matching(x, a, aw, seed, shem):
    """The matching the GPU computes, by its definition: the sequential greedy
    matching over the edges in descending order of (key, then the smaller and
    larger endpoint ascending), key = (2^32-1 minus the rank of the smaller
    endpoint degree d -- min(d, 4) + max(0, floor(log2 d) - 2) --, edge weight,
    seeded lowbias32 hash) for heavy-edge matching and the hash alone for
    random matching."""
    n = len(x) - 1
    deg = np.diff(x)
    edges = []
    for v in range(n):
        for j in range(x[v], x[v + 1]):
            u = int(a[j])
            if v < u:
                w = int(aw[j]) if aw is not None else 1
                h = _mix32(v ^ _mix32(u ^ (seed & _M32)) ^ ((seed >> 32) & _M32))
                if shem:
                    hi = ((_M32 - _degree_rank(min(int(deg[v]), int(deg[u])))) << 32) | ((w & _M32) ^ 0x80000000)
                else:
                    hi = 0
                edges.append((-hi, -h, v, u))
    edges.sort()
    m = np.arange(n, dtype=np.int32)
    free = np.ones(n, bool)
    for _, _, v, u in edges:
        if free[v] and free[u]:
            m[v], m[u] = u, v
            free[v] = free[u] = False
    return m


@pytest.mark.parametrize("shem", [1, 0])
def test_gpu_matching_is_the_greedy_matching(shem):
    from rmetis_amd import synth
    cases = [synth.rgg(3000, 11), synth.grid(40, 37)]
    x2, a2, aw2 = random_graph(1500, 0.01, 4, wmax=3)           # many equal weights
    for s in cases:
        g = R.Graph(s.n, 1, s.xadj, s.adjncy, s.vwgt, s.adjwgt)
        for seed in (1, 0x1234567890ABCDEF):
            ctype = R.CType.Shem if shem else R.CType.Rm
            assert np.array_equal(R.gpu_matching(g, seed, ctype), _greedy_matching(s.xadj, s.adjncy, s.adjwgt, seed, shem))
    g2 = _graph(x2, a2, adjwgt=aw2)
    ctype = R.CType.Shem if shem else R.CType.Rm
    assert np.array_equal(R.gpu_matching(g2, 7, ctype), _greedy_matching(x2, a2, aw2, 7, shem))


def _greedy_matching_np(x, a, aw, seed, shem, vwgt=None, cap=None):
    """_greedy_matching with the keys and the edge order computed by numpy
    (for large levels); with cap, only the edges whose endpoints' summed
    weights fit cap[c] for every constraint c take part (the keys still use
    the full graph's degrees)."""
    n = len(x) - 1
    deg = np.diff(x).astype(np.int64)
    src = np.repeat(np.arange(n, dtype=np.int64), deg)
    keep = src < a
    if cap is not None:
        vw = np.asarray(vwgt, np.int64).reshape(n, len(cap))
        keep &= ((vw[src] + vw[a.astype(np.int64)]) <= np.asarray(cap, np.int64)).all(axis=1)
    v, u = src[keep], a[keep].astype(np.int64)
    w = (aw[keep] if aw is not None else np.ones(len(v), np.int32)).astype(np.int64)

    def mix(z):
        z = z.astype(np.uint64) & np.uint64(_M32)
        z ^= z >> np.uint64(16)
        z = (z * np.uint64(0x7FEB352D)) & np.uint64(_M32)
        z ^= z >> np.uint64(15)
        z = (z * np.uint64(0x846CA68B)) & np.uint64(_M32)
        z ^= z >> np.uint64(16)
        return z

    lo, hi_s = np.uint64(seed & _M32), np.uint64((seed >> 32) & _M32)
    h = mix(v.astype(np.uint64) ^ mix(u.astype(np.uint64) ^ lo) ^ hi_s)
    if shem:
        md = np.maximum(np.minimum(deg[v], deg[u]), 1)
        lg = np.zeros(len(md), np.int64)                    # floor(log2(md))
        for b in range(1, 32):
            lg[md >= (1 << b)] = b
        rank = (np.minimum(md, 4) + np.maximum(lg - 2, 0)).astype(np.uint64)
        hi = ((np.uint64(_M32) - rank) << np.uint64(32)) | ((w.astype(np.uint64) & np.uint64(_M32)) ^ np.uint64(0x80000000))
    else:
        hi = np.zeros(len(v), np.uint64)
    order = np.lexsort((u, v, ~h, ~hi))
    m = np.arange(n, dtype=np.int32)
    free = np.ones(n, bool)
    for e in order.tolist():
        a_, b_ = int(v[e]), int(u[e])
        if free[a_] and free[b_]:
            m[a_], m[b_] = b_, a_
            free[a_] = free[b_] = False
    return m


@pytest.mark.parametrize("shem", [1, 0])
def test_gpu_matching_large_levels(shem):
    # levels of 2^16 vertices and more (wide grids of suitor chains); the
    # result is still the greedy matching in edge-key order
    from rmetis_amd import synth
    ctype = R.CType.Shem if shem else R.CType.Rm
    cases = [synth.rgg(70000, 12), synth.grid(300, 240)]
    xr, ar, awr = random_graph(66000, 0.00012, 8, wmax=2)       # many equal keys, isolated vertices
    for s in cases:
        g = R.Graph(s.n, 1, s.xadj, s.adjncy, s.vwgt, s.adjwgt)
        for seed in (5, 0xABCDEF0123456789):
            assert np.array_equal(R.gpu_matching(g, seed, ctype), _greedy_matching_np(s.xadj, s.adjncy, s.adjwgt,
                                                                                    seed, shem))
    g = _graph(xr, ar, adjwgt=awr)
    assert np.array_equal(R.gpu_matching(g, 9, ctype), _greedy_matching_np(xr, ar, awr, 9, shem))


@pytest.mark.parametrize("shem", [1, 0])
def test_gpu_matching_weight_cap(shem):
    # the capped matching is the greedy matching over the pairs that fit;
    # caps that every pair fits give the uncapped matching
    from rmetis_amd import synth
    ctype = R.CType.Shem if shem else R.CType.Rm
    cases = [(synth.rgg(3000, 21, vwgt_max=4), 1, [5]), (synth.rgg(70000, 22, vwgt_max=4), 1, [6]),
             (synth.rgg(5000, 23, vwgt_max=8, ncon=2), 2, [9, 12]), (synth.rgg(4000, 24, vwgt_max=3), 1, [2])]
    for s, ncon, cap in cases:
        g = R.Graph(s.n, ncon, s.xadj, s.adjncy, s.vwgt, s.adjwgt)
        for seed in (4, 0x0F1E2D3C4B5A6978):
            got = R.gpu_matching(g, seed, ctype, cap=cap)
            assert np.array_equal(got, _greedy_matching_np(s.xadj, s.adjncy, s.adjwgt, seed, shem, s.vwgt, cap))
            vw = s.vwgt.reshape(s.n, ncon).astype(np.int64)
            pair = got != np.arange(s.n)
            assert ((vw[pair] + vw[got[pair]]) <= np.array(cap)).all()
        free = [int(2 * s.vwgt.reshape(s.n, ncon)[:, c].max()) for c in range(ncon)]
        assert np.array_equal(R.gpu_matching(g, 4, ctype, cap=free), R.gpu_matching(g, 4, ctype))
    with pytest.raises(R.InputError):
        s = cases[0][0]
        R.gpu_matching(R.Graph(s.n, 1, s.xadj, s.adjncy, s.vwgt, s.adjwgt), 1, ctype, cap=[-1])


@pytest.mark.parametrize("shem", [1, 0])
def test_gpu_matching_wide_suitor_words(shem):
    # suitor words hold (smaller degree, weight) only for weights below 2^16;
    # heavier edges take the other encoding, which reloads the suitor's degree
    ctype = R.CType.Shem if shem else R.CType.Rm
    x, a, aw = random_graph(2000, 0.006, 31, wmax=3)
    remap = np.array([0, 65535, 65536, 1], np.int32)            # 65535 compact, 65536 not
    aw = remap[aw]
    x2, a2, aw2 = random_graph(1500, 0.008, 32, wmax=200000)     # mostly above 2^16
    for xx, aa, ww in ((x, a, aw), (x2, a2, aw2)):
        g = _graph(xx, aa, adjwgt=ww)
        for seed in (3, 0xFEDCBA9876543210):
            assert np.array_equal(R.gpu_matching(g, seed, ctype), _greedy_matching(xx, aa, ww, seed, shem))


def test_kway_reference_cases():
    # tests/correctness.rs:119-128 (greedy, ncuts 5, balanced)
    x, a = grid_sorted(4, 4)
    g = _graph(x, a)
    opts = R.Options.for_kway(seed=1, ncuts=5, rtype=R.RType.Greedy)
    r = R.part_graph_kway(g, 4, options=opts)
    assert r.part.min() >= 0 and r.part.max() < 4
    pw = g.partition_weights(r.part, 4)
    assert (pw[:, 0] <= 16 * 1.03 / 4 + 1).all()
    assert r.objval == g.edge_cut(r.part)
    # tests/correctness.rs:100-108 with greedy refinement
    x, a = cycle4()
    g = _graph(x, a)
    r = R.part_graph_kway(g, 2, options=R.Options.for_kway(seed=1, rtype=R.RType.Greedy))
    assert r.objval <= 4
    # determinism (tests/correctness.rs:179-187)
    x, a = grid_sorted(4, 4)
    g = _graph(x, a)
    o = R.Options.for_kway(seed=123, rtype=R.RType.Greedy)
    r1, r2 = R.part_graph_kway(g, 4, options=o), R.part_graph_kway(g, 4, options=o)
    assert r1.objval == r2.objval and np.array_equal(r1.part, r2.part)
    # tests/parallel.rs:163-187: objval == edge_cut, every part non-empty
    x, a = grid_parallel_order(4, 4)
    g = _graph(x, a)
    r = R.part_graph_kway(g, 4, options=R.Options.for_kway(seed=1, ncuts=3, rtype=R.RType.Greedy))
    assert r.objval == g.edge_cut(r.part)
    assert (np.bincount(r.part, minlength=4) > 0).all()
    # tests/parallel.rs:202-215: more attempts never worse
    x, a = grid_parallel_order(6, 6)
    g = _graph(x, a)
    r1 = R.part_graph_kway(g, 4, options=R.Options.for_kway(seed=42, ncuts=1, rtype=R.RType.Greedy))
    r4 = R.part_graph_kway(g, 4, options=R.Options.for_kway(seed=42, ncuts=4, rtype=R.RType.Greedy))
    assert r4.objval <= r1.objval


def test_kway_errors():
    x, a = cycle4()
    g = _graph(x, a)
    with pytest.raises(R.InputError):
        R.part_graph_kway(g, 10, options=R.Options(rtype=R.RType.Greedy))
    with pytest.raises(R.InputError):
        R.part_graph_kway(g, 1, options=R.Options(rtype=R.RType.Greedy))
    r = R.part_graph_kway(g, 2, options=R.Options(seed=1))   # the default RTYPE Fm runs k-way FM
    assert r.objval == g.edge_cut(r.part)


def test_kway_weighted():
    # tests/correctness.rs:221-235 with greedy refinement
    g = R.Graph(4, 1, [0, 2, 4, 6, 8], [1, 3, 0, 2, 1, 3, 0, 2], [2, 1, 2, 1], [3, 1, 3, 1, 3, 1, 3, 1])
    r = R.part_graph_kway(g, 2, options=R.Options.for_kway(seed=5, rtype=R.RType.Greedy))
    assert r.part.min() >= 0 and r.part.max() < 2
    assert r.objval == g.edge_cut(r.part)
