"""Pins the oracle (CPU restatement of javagg/rmetis) before it is trusted as
the checker: the reference's own unit values, closed forms, published RNG
vectors, invariants, and an independent brute-force transcription of the
greedy pass on tiny graphs."""
import itertools
import math

import numpy as np
import pytest

import pyoracle as O
from graphs import build_undirected, cycle4, grid_parallel_order, grid_sorted, path, random_graph

M64 = (1 << 64) - 1


# ---------------------------------------------------------------- RNG
def test_splitmix64_published_vector():
    # SplitMix64 (Steele, Lea, Flood 2014) seeded with 0
    assert O.splitmix64(0) == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                               0x06C45D188009454F, 0xF88BB8A8724C81EC]


def test_xoshiro256pp_reference_vector():
    # xoshiro256++ from state (1, 2, 3, 4): first output is closed form
    # rotl(s0 + s3, 23) + s0 = 5 << 23 | 1; the next ones are the published
    # values of the reference implementation (xoshiro.di.unimi.it).
    r = O.rng_from_words([1, 2, 3, 4])
    out = [O.next_u64(r) for _ in range(6)]
    assert out[0] == (5 << 23) + 1 == 41943041
    assert out[:6] == [41943041, 58720359, 3588806011781223, 3591011842654386,
                       9228616714210784205, 9973669472204895162]


def test_next_u32_is_upper_half():
    r1, r2 = O.rng(7), O.rng(7)
    for _ in range(10):
        assert O.next_u32(r1) == O.next_u64(r2) >> 32


def test_range_is_unbiased_widening_multiply():
    # gen_range(0..n): hi 64 bits of v*n with rejection zone (n << lz) - 1
    r1, r2 = O.rng(11), O.rng(11)
    for n in (1, 2, 3, 10, 1000, 2**33 + 5):
        got = O.range_incl(r1, 0, n - 1)
        zone = ((n << (64 - n.bit_length())) - 1) & M64
        while True:
            v = O.next_u64(r2)
            p = v * n
            if (p & M64) <= zone:
                assert got == p >> 64
                break
        assert 0 <= got < n


# ---------------------------------------------------------------- reductions
def test_reference_unit_values():
    # graph/mod.rs:455-485
    x, a = cycle4()
    assert O.edge_cut(x, a, None, [0, 0, 1, 1]) == 2
    assert O.partition_weights(1, None, [0, 0, 1, 1], 2).tolist() == [[2], [2]]
    assert O.total_weight(4, 1, None).tolist() == [4]
    assert O.boundary_vertices(x, a, [0, 0, 1, 1]).tolist() == [0, 1, 2, 3]


@pytest.mark.parametrize("rows,cols,k", [(8, 8, 2), (16, 12, 4), (256, 256, 8), (32, 64, 32)])
def test_banded_grid_closed_forms(rows, cols, k):
    x, a = grid_sorted(rows, cols)
    n = rows * cols
    band = (np.arange(n) // cols) * k // rows
    assert O.edge_cut(x, a, None, band) == (k - 1) * cols
    pw = O.partition_weights(1, None, band, k)
    assert (pw[:, 0] == rows * cols // k).all()
    # rows next to a band edge: 2(k-1) of them, or all rows when bands are 1 row high
    assert len(O.boundary_vertices(x, a, band)) == min(rows, 2 * (k - 1)) * cols


def test_reductions_brute_force_tiny():
    rng = np.random.default_rng(5)
    for trial in range(30):
        n = int(rng.integers(2, 9))
        x, a, aw = random_graph(n, 0.5, trial, wmax=5)
        k = int(rng.integers(2, 4))
        vw = rng.integers(1, 5, n * 2).astype(np.int32)
        for part in itertools.islice(itertools.product(range(k), repeat=n), 200):
            part = np.array(part, np.int32)
            cut = 0
            for v in range(n):
                for j in range(x[v], x[v + 1]):
                    if part[a[j]] != part[v]:
                        cut += aw[j] if aw is not None else 1
            assert O.edge_cut(x, a, aw, part) == cut // 2
            pw = np.zeros((k, 2), np.int64)
            for v in range(n):
                pw[part[v]] += vw[2 * v:2 * v + 2]
            assert np.array_equal(O.partition_weights(2, vw, part, k), pw)
            b = [v for v in range(n) if any(part[a[j]] != part[v] for j in range(x[v], x[v + 1]))]
            assert O.boundary_vertices(x, a, part).tolist() == b


def test_edge_cut_wraps_like_i32():
    x, a = path(2)
    w = np.array([2**31 - 1, 2**31 - 1], np.int32)
    # i32 sum 2*(2^31-1) wraps to -2, halved (truncating) -> -1
    assert O.edge_cut(x, a, w, [0, 1]) == -1


def test_move_gain_closed_forms():
    # star centre 0 with leaves 1..5 (weights 1..5): gain(0 -> p) = ed[p] - ed[pv]
    x, a = build_undirected(6, [(0, i) for i in range(1, 6)])
    src = np.repeat(np.arange(6), np.diff(x))
    w = np.maximum(src, a).astype(np.int32)   # weight of edge (0,i) = i, symmetric
    part = np.array([0, 1, 1, 2, 0, 2], np.int32)
    assert O.move_gain(x, a, w, 0, part, 0, 1) == (1 + 2) - 4
    assert O.move_gain(x, a, w, 0, part, 0, 2) == (3 + 5) - 4
    assert O.move_gain(x, a, w, 0, part, 0, 3) == -4
    assert O.move_gain(x, a, w, 3, part, 2, 0) == 3     # leaf 3's only neighbour is in part 0


def test_target_weights():
    # graph/mod.rs:392-410: (tpwgts or 1/k) * total
    vw = np.array([1, 2, 3, 4, 5, 6], np.int32)
    t = O.target_weights(3, 2, vw, 2)
    assert np.allclose(t, [4.5, 6.0, 4.5, 6.0])
    t = O.target_weights(3, 2, vw, 2, tpwgts=[0.25, 0.5, 0.75, 0.5])
    assert t.tolist() == [float(np.float32(0.25)) * 9, 6.0, float(np.float32(0.75)) * 9, 6.0]


# ---------------------------------------------------------------- validation
def test_validate_reference_cases():
    # graph/validate.rs:151-188
    x, a = [0, 2, 4, 6, 8], [1, 3, 0, 2, 1, 3, 0, 2]
    assert O.validate_csr(4, 1, x, a) == 1
    assert O.validate_csr(4, 1, [0, 2, 4], a) == -2
    assert O.validate_csr(4, 1, [0, 3, 5, 7, 9], [0, 1, 3, 0, 2, 1, 3, 0, 2]) == -2
    assert O.validate_csr(4, 1, [0, 1, 1, 1, 1], [1]) == -2
    assert O.validate_csr(4, 1, x, a, vwgt=[1, 1, 1]) == -2
    assert O.validate_csr(4, 1, x, a, vwgt=[1, 0, 1, 1]) == -2


# ---------------------------------------------------------------- greedy
def brute_greedy_pass(x, a, aw, vw, ncon, part, k, tpw, ub, mode, pw):
    """Second transcription of refine/greedy.rs:60-123 / 127-182, written
    independently (whole-vector form, Python ints) for tiny graphs."""
    n = len(x) - 1
    part = list(part)
    pw = [list(r) for r in np.asarray(pw).reshape(k, ncon)]
    limit = [[tpw[p * ncon + c] * ub[c] for c in range(ncon)] for p in range(k)]
    nbrs = [[(int(a[j]), int(aw[j]) if aw is not None else 1) for j in range(x[v], x[v + 1])] for v in range(n)]
    boundary = [v for v in range(n) if any(part[u] != part[v] for u, _ in nbrs[v])]
    moves = []
    for v in boundary:
        pv = part[v]
        w = [int(vw[v * ncon + c]) if vw is not None else 1 for c in range(ncon)]
        if mode == 0 and not any(pw[pv][c] > limit[pv][c] for c in range(ncon)):
            continue
        ed = [0] * k
        for u, ww in nbrs[v]:
            ed[part[u]] += ww
        feas = [p for p in range(k) if p != pv and all(pw[p][c] + w[c] <= limit[p][c] for c in range(ncon))]
        gains = [(ed[p] - ed[pv], -p) for p in feas]
        if mode == 1:
            gains = [g for g in gains if g[0] > 0]
        if not gains:
            continue
        best = -max(gains)[1]
        moves.append((v, best))
        for c in range(ncon):
            pw[pv][c] -= w[c]
            pw[best][c] += w[c]
    for v, p in moves:
        part[v] = p
    return np.array(part, np.int32), np.array(pw, np.int32), np.array(moves, np.int32).reshape(-1, 2)


def test_greedy_pass_vs_independent_transcription():
    rng = np.random.default_rng(17)
    for trial in range(150):
        n = int(rng.integers(3, 14))
        x, a, aw = random_graph(n, float(rng.uniform(0.2, 0.7)), 1000 + trial, wmax=int(rng.integers(1, 6)))
        k = int(rng.integers(2, 5))
        ncon = int(rng.integers(1, 3))
        vw = rng.integers(1, 4, n * ncon).astype(np.int32)
        part = rng.integers(0, k, n).astype(np.int32)
        tot = vw.reshape(n, ncon).sum(0)
        tpw = np.tile(tot / k, k).astype(np.float64)
        ub = np.full(ncon, float(rng.choice([1.0, 1.03, 1.2])))
        pw = O.partition_weights(ncon, vw, part, k)
        for mode in (0, 1):
            got = O.greedy_pass(ncon, x, a, vw, aw, part, k, tpw, ub, mode, pw)
            exp = brute_greedy_pass(x, a, aw, vw, ncon, part, k, tpw, ub, mode, pw)
            assert np.array_equal(got[0], exp[0])
            assert np.array_equal(got[1], exp[1])
            assert np.array_equal(got[2], exp[2])


def test_greedy_reference_expectations():
    # greedy.rs:229-251 and tests/parallel.rs:91-132
    x, a = cycle4()
    for init, cut in (([0, 0, 1, 1], 2), ([0, 1, 0, 1], 4), ([0, 1, 1, 0], 2)):
        # maxw = floor(2 * 1.03) = 2: no move is ever feasible
        p, c = O.greedy_refine(1, x, a, None, None, init, 2, [2.0, 2.0], [1.03], 10)
        assert c == cut and p.tolist() == init
    x, a = grid_parallel_order(4, 4)
    part = np.arange(16) // 4
    before = O.edge_cut(x, a, None, part)
    _, c = O.greedy_refine(1, x, a, None, None, part, 4, [4.0] * 4, [1.05], 10)
    assert c <= before


def test_greedy_invariants():
    rng = np.random.default_rng(3)
    x, a, aw = random_graph(400, 0.02, 9, wmax=6)
    n, k = 400, 6
    vw = rng.integers(1, 5, n).astype(np.int32)
    part = rng.integers(0, k, n).astype(np.int32)
    part[:150] = 0
    tpw = np.full(k, vw.sum() / k)
    ub = np.array([1.03])
    pw = O.partition_weights(1, vw, part, k)
    for mode in (0, 1, 1):
        newp, newpw, moves = O.greedy_pass(1, x, a, vw, aw, part, k, tpw, ub, mode, pw)
        # weight conservation and pw == partition_weights(new part)
        assert newpw.sum() == pw.sum()
        assert np.array_equal(newpw, O.partition_weights(1, vw, newp, k))
        # moves ascend in vertex id (boundary order) and target another part
        assert (np.diff(moves[:, 0]) > 0).all()
        assert all(part[v] != p for v, p in moves)
        part, pw = newp, newpw


# ---------------------------------------------------------------- contraction
def test_coarse_reference_path():
    # graph/mod.rs:499-514
    x, a = path(4)
    c = O.build_coarse_graph(1, x, a, None, None, None, [1, 0, 3, 2])
    assert c["n"] == 2 and c["cxadj"].tolist() == [0, 1, 2] and c["cadjncy"].tolist() == [1, 0]
    assert c["cmap"].tolist() == [0, 0, 1, 1]
    assert not c["adjwgt_present"] and c["cvwgt"] is None


def test_coarse_grid_pairs():
    # horizontal pairing of a 4x4 grid -> 4x2 grid, vertical edges weight 2,
    # horizontal weight 1; vwgt 2 only when the input carries vwgt
    x, a = grid_sorted(4, 4)
    match = np.array([v + 1 if v % 2 == 0 else v - 1 for v in range(16)], np.int32)
    c = O.build_coarse_graph(1, x, a, None, None, None, match)
    assert c["n"] == 8 and c["cvwgt"] is None and c["adjwgt_present"]
    cx, ca, cw = c["cxadj"], c["cadjncy"], c["cadjwgt"]
    for cv in range(8):
        r, col = divmod(cv, 2)
        for j in range(cx[cv], cx[cv + 1]):
            r2, c2 = divmod(int(ca[j]), 2)
            assert cw[j] == (2 if c2 == col else 1)
            assert abs(r2 - r) + abs(c2 - col) == 1
    c = O.build_coarse_graph(1, x, a, np.ones(16, np.int32), None, None, match)
    assert c["cvwgt"].tolist() == [2] * 8


def test_coarse_invariants():
    rng = np.random.default_rng(2)
    for trial in range(20):
        x, a, aw = random_graph(200, 0.05, 50 + trial, wmax=9)
        vw = rng.integers(1, 5, 200).astype(np.int32)
        m = O.shem_matching(x, a, aw, O.rng(trial))
        c = O.build_coarse_graph(1, x, a, vw, aw, None, m)
        assert c["cvwgt"].sum() == vw.sum()
        src = np.repeat(np.arange(200), np.diff(x))
        internal = aw[c["cmap"][src] == c["cmap"][a]].sum()
        assert c["cadjwgt"].sum() == aw.sum() - internal
        # symmetric coarse CSR with matching weights
        cs = np.repeat(np.arange(c["n"]), np.diff(c["cxadj"]))
        d1 = dict(zip(zip(cs.tolist(), c["cadjncy"].tolist()), c["cadjwgt"].tolist()))
        assert all(d1[(v, u)] == w for (u, v), w in d1.items())
        assert O.validate_csr(c["n"], 1, c["cxadj"], c["cadjncy"], c["cvwgt"], c["cadjwgt"]) == 1


# ---------------------------------------------------------------- matching
@pytest.mark.parametrize("fn", ["shem", "rm"])
def test_matchings_symmetric_maximal(fn):
    for seed in range(10):
        x, a, aw = random_graph(120, 0.05, seed, wmax=4)
        r = O.rng(seed)
        m = O.shem_matching(x, a, aw, r) if fn == "shem" else O.random_matching(x, a, r)
        assert np.array_equal(m[m], np.arange(120))           # shem.rs:96-99, rm.rs:56-64
        src = np.repeat(np.arange(120), np.diff(x))
        un = m == np.arange(120)
        assert not (un[src] & un[a]).any()                     # maximal
        for v in np.nonzero(~un)[0]:
            assert m[v] in a[x[v]:x[v + 1]]


def test_shem_star_prefers_heavy_edge():
    # shem.rs:80-106: star centre 0, leaf weights 1, 3, 1
    x = np.array([0, 3, 4, 5, 6], np.int32)
    a = np.array([1, 2, 3, 0, 0, 0], np.int32)
    w = np.array([1, 3, 1, 1, 3, 1], np.int32)
    for s in range(20):
        m = O.shem_matching(x, a, w, O.rng(s))
        assert np.array_equal(m[m], np.arange(4))
        assert (m != np.arange(4)).sum() >= 2


# ---------------------------------------------------------------- pipeline
def _balanced(xadj, part, k, ufactor, vwgt=None):
    n = len(xadj) - 1
    pw = np.bincount(part, weights=vwgt, minlength=k)
    tot = vwgt.sum() if vwgt is not None else n
    return (pw <= tot * (1 + ufactor / 1000) / k + 1).all()


def test_pipeline_reference_cases():
    x, a = grid_sorted(4, 4)
    rc, obj, part = O.part_graph_kway(x, a, 4, seed=1, ncuts=5)
    assert rc == 1 and part.min() >= 0 and part.max() < 4 and _balanced(x, part, 4, 30)
    x, a = cycle4()
    rc, obj, part = O.part_graph_kway(x, a, 2, seed=1)
    assert rc == 1 and obj <= 4
    x, a = grid_sorted(4, 4)
    r1 = O.part_graph_kway(x, a, 4, seed=123)
    r2 = O.part_graph_kway(x, a, 4, seed=123)
    assert r1[1] == r2[1] and np.array_equal(r1[2], r2[2])
    x, a = grid_parallel_order(4, 4)
    rc, obj, part = O.part_graph_kway(x, a, 4, seed=1, ncuts=3)
    assert obj == O.edge_cut(x, a, None, part)
    assert (np.bincount(part, minlength=4) > 0).all()
    x, a = grid_parallel_order(6, 6)
    o1 = O.part_graph_kway(x, a, 4, seed=42, ncuts=1)[1]
    o4 = O.part_graph_kway(x, a, 4, seed=42, ncuts=4)[1]
    assert o4 <= o1
    x, a = cycle4()
    assert O.part_graph_kway(x, a, 10)[0] == -2
    assert O.part_graph_kway(x, a, 1)[0] == -2
