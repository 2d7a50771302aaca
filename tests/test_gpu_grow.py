"""The device BFS growing of the coarsest level (initial/grow.rs:12-103)
through rmetis_hip_grow_partition: bit-exact against the pinned oracle for
the same attempt seed.  The kernel resolves each round's parts in parallel
(sub-rounds with claim stamps); the cases cover contested frontiers, many
parts, vertices left over after the rounds (isolated vertices and unreached
components, which go to the lightest part on constraint 0), two
constraints, k = n and k above the 4,096 parts whose leftover weights the
kernel keeps in LDS."""
import numpy as np
import pytest

from rmetis_amd import _lib
from rmetis_amd import synth

import sys, os
sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import pyoracle as O  # noqa: E402

from graphs import random_graph  # noqa: E402

pytestmark = pytest.mark.gpu


def _gpu_grow(n, ncon, x, a, vw, k, tpw, seed):
    part = np.zeros(n, np.int32)
    rc = _lib.lib().rmetis_hip_grow_partition(n, ncon, x.ctypes.data_as(_lib.i32p), a.ctypes.data_as(_lib.i32p),
                                              vw.ctypes.data_as(_lib.i32p) if vw is not None else None, k,
                                              tpw.ctypes.data_as(_lib.f64p), seed, part.ctypes.data_as(_lib.i32p))
    assert rc == 1, _lib.last_error()
    return part


def _targets(n, ncon, vw, k):
    tot = (vw.reshape(n, ncon).sum(0) if vw is not None else np.full(ncon, n)).astype(np.float64)
    return np.tile(tot / k, k)


def _check(n, ncon, x, a, vw, k, seed, tpw=None):
    tpw = _targets(n, ncon, vw, k) if tpw is None else tpw
    got = _gpu_grow(n, ncon, x, a, vw, k, tpw, seed)
    want = O.grow_partition(ncon, x, a, vw, k, tpw, O.rng(seed))
    assert np.array_equal(got, want), (n, k, seed)
    return got


def test_grow_random_graphs():
    rng = np.random.default_rng(3)
    for trial in range(40):
        n = int(rng.integers(4, 300))
        x, a, aw = random_graph(n, float(rng.choice([0.005, 0.02, 0.1])), trial, wmax=3)
        ncon = int(rng.integers(1, 3))
        vw = rng.integers(1, 6, n * ncon).astype(np.int32) if trial % 3 else None
        k = int(min(n, rng.choice([2, 3, 8, 17, n])))
        _check(n, ncon, x, a, vw, k, int(rng.integers(1, 2**40)))


def test_grow_grids_and_rggs_many_parts():
    # frontiers of many parts meet (contested claims in most rounds)
    for rows, cols, k, seed in ((60, 70, 64, 5), (200, 150, 256, 9), (33, 31, 700, 11)):
        g = synth.grid(rows, cols)
        _check(g.n, 1, g.xadj, g.adjncy, None, k, seed)
    g = synth.rgg(20000, 4, vwgt_max=4)
    _check(g.n, 1, g.xadj, g.adjncy, g.vwgt, 128, 13)
    g = synth.rgg(6000, 5, vwgt_max=8, ncon=2)
    _check(g.n, 2, g.xadj, g.adjncy, g.vwgt, 40, 17)


def test_grow_leftovers_and_uneven_targets():
    # mostly isolated vertices and small components: most vertices are left
    # over and placed one by one on the lightest part
    rng = np.random.default_rng(8)
    x, a, aw = random_graph(5000, 0.00015, 21, wmax=2)
    vw = rng.integers(1, 9, 5000).astype(np.int32)
    got = _check(5000, 1, x, a, vw, 37, 23)
    assert len(np.unique(got)) == 37
    # targets as the k-way path passes user tpwgts (not uniform)
    n = 3000
    x, a, _ = random_graph(n, 0.004, 22, wmax=2)
    frac = rng.random(16)
    frac /= frac.sum()
    _check(n, 1, x, a, None, 16, 29, tpw=frac * n)


def test_grow_k_equals_n_and_many_parts_beyond_lds():
    x, a, _ = random_graph(600, 0.01, 31, wmax=2)
    got = _check(600, 1, x, a, None, 600, 37)
    assert np.array_equal(np.sort(got), np.arange(600))          # every vertex its own part
    # k above 4,096: the leftover walk keeps the weights in global memory
    rng = np.random.default_rng(9)
    x, a, _ = random_graph(12000, 0.00005, 32, wmax=2)
    vw = rng.integers(1, 4, 12000).astype(np.int32)
    _check(12000, 1, x, a, vw, 5000, 41)


def test_grow_equal_leftover_weights_take_the_periodic_rule():
    # leftovers of one weight: after the first i0 steps of the walk every part
    # lies in one band and the rest are dealt by period; unit weights (every
    # vertex 1) and a weighted graph whose isolated vertices all weigh 3
    x, a, _ = random_graph(20000, 0.00008, 51, wmax=2)
    for k, seed in ((37, 53), (256, 57), (3, 59)):
        _check(20000, 1, x, a, None, k, seed)
    deg = np.diff(x)
    rng = np.random.default_rng(10)
    vw = rng.integers(1, 9, 20000).astype(np.int32)
    vw[deg == 0] = 3
    _check(20000, 1, x, a, vw, 64, 61)
    # two constraints: the rule reads constraint 0 only
    vw2 = rng.integers(1, 5, 40000).astype(np.int32)
    vw2[0::2][deg == 0] = 2
    _check(20000, 2, x, a, vw2, 48, 63)
    # leftovers fewer than the walk's i0: the walk alone
    x2, a2, _ = random_graph(3000, 0.002, 52, wmax=2)
    _check(3000, 1, x2, a2, None, 200, 67)


def _csr(n, edges, rng=None):
    """CSR of the undirected edge list (each pair listed both ways, repeats
    kept), rows shuffled when rng is given."""
    rows = [[] for _ in range(n)]
    for u, v in edges:
        rows[u].append(v)
        rows[v].append(u)
    if rng is not None:
        for r in rows:
            rng.shuffle(r)
    x = np.zeros(n + 1, np.int32)
    x[1:] = np.cumsum([len(r) for r in rows])
    a = np.array([u for r in rows for u in r], np.int32)
    return x, a


def test_grow_long_rows_are_taken_by_the_whole_workgroup():
    # rows longer than 96 entries are scanned by the whole workgroup in every
    # phase (hubs of R-MAT graphs, stars); the order of the taken neighbours
    # in the queue must stay the row order
    g = synth.rmat(14, 16, 3)
    for k, seed in ((32, 71), (256, 73)):
        _check(g.n, 1, g.xadj, g.adjncy, None, k, seed)
    rng = np.random.default_rng(11)
    edges = [(0, i) for i in range(1, 501)]                      # a star of 500 ...
    nxt = 501
    for i in range(1, 501):                                      # ... each with 3 leaves
        for _ in range(3):
            edges.append((i, nxt))
            nxt += 1
    x, a = _csr(nxt, edges, rng)
    for k, seed in ((2, 77), (8, 79), (64, 83)):
        _check(nxt, 1, x, a, None, k, seed)
    # repeated entries (every edge of the hubs listed twice, in shuffled rows:
    # some repeats share a chunk of the row, some do not)
    edges2 = [(0, i) for i in range(1, 700)] + [(1, i) for i in range(700, 1400)]
    edges2 += [(i, i + 1) for i in range(2, 1399)]
    edges2 = edges2 + edges2[:900]
    x2, a2 = _csr(1400, edges2, rng)
    vw = rng.integers(1, 5, 1400).astype(np.int32)
    for k, seed in ((3, 89), (40, 97)):
        _check(1400, 1, x2, a2, vw, k, seed)


def test_grow_rows_without_pairwise_symmetry_are_scanned():
    # the validation (validate.rs:84-108) checks only that each vertex
    # appears as often as its degree; a graph whose pairs appear unequally
    # often in the two rows (twisted triangles a:[b,b,c], b:[a,c,c],
    # c:[a,a,b] on a ring) is valid, and the kernel then scans rows instead of
    # keeping unassigned-neighbour counts
    rows = {v: [] for v in range(900)}
    for t in range(300):
        a_, b_, c_ = 3 * t, 3 * t + 1, 3 * t + 2
        rows[a_] += [b_, b_, c_]
        rows[b_] += [a_, c_, c_]
        rows[c_] += [a_, a_, b_]
    for t in range(299):                                   # ring between triangles (symmetric)
        rows[3 * t + 2].append(3 * t + 3)
        rows[3 * t + 3].append(3 * t + 2)
    x = np.zeros(901, np.int32)
    x[1:] = np.cumsum([len(rows[v]) for v in range(900)])
    a = np.array([u for v in range(900) for u in rows[v]], np.int32)
    for k, seed in ((4, 101), (30, 103)):
        _check(900, 1, x, a, None, k, seed)
