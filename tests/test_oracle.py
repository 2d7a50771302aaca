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


def test_range_power_of_two_closed_form():
    # rand 0.8.5 UniformInt<usize>::sample_single_inclusive keeps a draw v
    # iff lo(v * range) <= (range << lz) - 1.  For range = 2^j that zone is
    # 2^63 - 1 whatever j is (the "conservative but fast approximation", not
    # the exact 2^64 - 1 that a power of two would allow), and lo = v << j,
    # so v is kept iff its bit 63 - j is clear, and the result is then v's top
    # j bits.  j = 0 (gen_range(0..1)) keeps only draws below 2^63 and returns
    # 0.  Both the result and the number of draws consumed are checked.
    rejected = 0
    for j in (0, 1, 3, 8, 17, 33, 62):
        n = 1 << j
        r1, r2 = O.rng(100 + j), O.rng(100 + j)
        for _ in range(48):
            got = O.range_incl(r1, 0, n - 1)
            while True:
                v = O.next_u64(r2)
                if (v >> (63 - j)) & 1 == 0:
                    break
                rejected += 1
            assert got == v >> (64 - j)
        assert O.next_u64(r1) == O.next_u64(r2)      # same number of draws consumed
    assert rejected > 100                            # about half of all draws


def test_range_small_odd_bounds():
    # range 3: zone (3 << 62) - 1; the result is floor(3 v / 2^64), i.e. 0, 1
    # or 2 as v falls below 2^64/3, 2*2^64/3 or above, and a draw is rejected
    # iff 3 v mod 2^64 >= 3 * 2^62, which for v in [0, 2^64) is the top quarter
    # of each third: v mod ceil(2^64/3) lands in [2^62, ...).  Checked against
    # exact rational arithmetic on the raw stream.
    from fractions import Fraction
    r1, r2 = O.rng(5), O.rng(5)
    third = Fraction(1 << 64, 3)
    for _ in range(200):
        got = O.range_incl(r1, 0, 2)
        while True:
            v = O.next_u64(r2)
            q = int(Fraction(v) / third)                 # floor(3v / 2^64)
            frac = Fraction(v) - q * third               # position inside its third
            if frac * 3 < (3 << 62):                    # lo = 3 * frac < 3 * 2^62
                break
        assert got == q
    # gen_range(a..b) shifts by a: low + hi
    r1, r2 = O.rng(9), O.rng(9)
    assert O.range_incl(r1, 1000, 1000 + 7) == 1000 + O.range_incl(r2, 0, 7)


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


# ---------------------------------------------------------------- grow
def _distinct_draws(seed, n, k):
    """The seed vertices grow.rs:24-32 picks: distinct gen_range(0..n) draws."""
    r = O.rng(seed)
    out = []
    while len(out) < k:
        v = O.range_incl(r, 0, n - 1)
        if v not in out:
            out.append(v)
    return out


def _seed_for(n, want, limit=200000):
    """An attempt seed whose first distinct draws over 0..n are `want`."""
    for s in range(1, limit):
        if _distinct_draws(s, n, len(want)) == list(want):
            return s
    raise AssertionError(f"no seed draws {want}")


def test_grow_k_equals_n_is_the_draw_order():
    # every vertex is a seed and at its target (1/k * n = 1): part[seed_p] = p
    x, a = grid_sorted(3, 3)
    for seed in (1, 2, 77):
        order = _distinct_draws(seed, 9, 9)
        part = O.grow_partition(1, x, a, None, 9, np.ones(9), O.rng(seed))
        assert [int(part[v]) for v in order] == list(range(9))


@pytest.mark.parametrize("n", [7, 8])
def test_grow_path_from_both_ends(n):
    # seeds 0 (part 0) and n-1 (part 1); each round part 0 claims the next
    # vertex to the right, then part 1 the next to the left, until a part
    # reaches its target n/2: part 0 ends with [0, ceil(n/2)), because it
    # moves first in every round, and part 1 with the rest
    x, a = path(n)
    seed = _seed_for(n, [0, n - 1])
    part = O.grow_partition(1, x, a, None, 2, np.full(2, n / 2.0), O.rng(seed))
    h = (n + 1) // 2
    assert part.tolist() == [0] * h + [1] * (n - h)


def test_grow_claims_every_free_neighbour_of_a_popped_vertex():
    # star centre 0, leaves 1..5; seeds 0 and 3: part 0's first pop claims
    # all free leaves at once (past its target 3), part 1 never grows
    x, a = build_undirected(6, [(0, i) for i in range(1, 6)])
    seed = _seed_for(6, [0, 3])
    part = O.grow_partition(1, x, a, None, 2, np.full(2, 3.0), O.rng(seed))
    assert part.tolist() == [0, 0, 0, 1, 0, 0]


def test_grow_leftovers_go_to_the_first_lightest_part_on_constraint_0():
    # edge (0, 1) plus isolated vertices; seeds 0 and 1 block each other, so
    # every other vertex is a leftover, assigned in ascending order to the
    # part lightest on constraint 0, the lower index on ties
    x, a = build_undirected(8, [(0, 1)])
    seed = _seed_for(8, [0, 1])
    part = O.grow_partition(1, x, a, None, 2, np.full(2, 4.0), O.rng(seed))
    assert part.tolist() == [0, 1, 0, 1, 0, 1, 0, 1]
    # two constraints: constraint 1 would prefer the other part every time
    x, a = build_undirected(6, [(0, 1)])
    seed = _seed_for(6, [0, 1])
    vw = np.array([1, 9, 2, 9, 5, 1, 1, 1, 1, 1, 1, 1], np.int32)
    part = O.grow_partition(2, x, a, vw, 2, np.full(4, 100.0), O.rng(seed))
    # pw0 = 1, pw1 = 2 on constraint 0: v2 -> 0 (pw0 6), v3, v4, v5 -> 1 (3, 4, 5)
    assert part.tolist() == [0, 1, 0, 1, 1, 1]


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


def _kway_composed(x, a, k, seed, vwgt=None, adjwgt=None, ncon=1, ncuts=1, niter=10, ufactor=30,
                   tpwgts=None, shem=True):
    """partition/kway.rs:20-141 composed in Python from the pinned oracle
    pieces (matching, contraction, targets, grow, greedy refine, cut): one
    attempt RNG per attempt drawn in the driver's order (coarsen/mod.rs:49-69
    matchings, then grow.rs seeds), the f32 stop rule, projection, strictly
    better cut kept."""
    ub = np.full(ncon, 1.0 + ufactor / 1000.0)
    best = None
    for att in range(ncuts):
        r = O.rng(seed + att)
        cur = (x, a, vwgt, adjwgt)
        levels = []
        limit = max(20 * k, 20)
        while True:
            cx, ca, cvw, caw = cur
            m = O.shem_matching(cx, ca, caw, r) if shem else O.random_matching(cx, ca, r)
            cg = O.build_coarse_graph(ncon, cx, ca, cvw, caw, None, m)
            nxt = (cg["cxadj"], cg["cadjncy"], cg["cvwgt"], cg["cadjwgt"] if cg["adjwgt_present"] else None)
            levels.append((cur, cg["cmap"]))
            ratio = np.float32(cg["n"]) / np.float32(len(cx) - 1)
            cur = nxt
            if cg["n"] <= limit or ratio > np.float32(0.95):
                break
        cx, ca, cvw, caw = cur
        tpw = O.target_weights(len(cx) - 1, ncon, cvw, k, tpwgts)
        part = O.grow_partition(ncon, cx, ca, cvw, k, tpw, r)
        part, _ = O.greedy_refine(ncon, cx, ca, cvw, caw, part, k, tpw, ub, niter)
        for (fx, fa, fvw, faw), cmap in reversed(levels):
            part = part[cmap]
            tpw = O.target_weights(len(fx) - 1, ncon, fvw, k, tpwgts)
            part, _ = O.greedy_refine(ncon, fx, fa, fvw, faw, part, k, tpw, ub, niter)
        obj = O.edge_cut(x, a, adjwgt, part)
        if best is None or obj < best[0]:
            best = (obj, part)
    return best


@pytest.mark.parametrize("case", ["grid", "weighted_rgg", "rm", "ncuts", "tpwgts"])
def test_part_graph_kway_is_the_composition_of_its_pinned_pieces(case):
    # the driver's composition and RNG consumption order, checked against an
    # independent transcription over the functions pinned above
    import rmetis_amd.synth as S
    kw = {}
    if case == "grid":
        x, a = grid_sorted(24, 20)
        k, vw, aw = 6, None, None
    else:
        g = S.rgg(1500, 3, vwgt_max=3)
        x, a, vw, aw, k = g.xadj, g.adjncy, g.vwgt, g.adjwgt, 8
        if case == "rm":
            kw["shem"] = False
        if case == "ncuts":
            kw["ncuts"] = 3
        if case == "tpwgts":
            kw["tpwgts"] = np.array([0.05, 0.1, 0.15, 0.2, 0.1, 0.1, 0.2, 0.1], np.float32)
    for seed in (1, 5):
        obj, part = _kway_composed(x, a, k, seed, vwgt=vw, adjwgt=aw, **kw)
        rc, o_obj, o_part = O.part_graph_kway(x, a, k, vwgt=vw, adjwgt=aw, seed=seed,
                                              ctype_shem=kw.get("shem", True), ncuts=kw.get("ncuts", 1),
                                              tpwgts=kw.get("tpwgts"))
        assert rc == 1 and o_obj == obj and np.array_equal(o_part, part)


# ---------------------------------------------------------------- k-way FM
def _fm_pass_transcribed(x, a, aw, vw, ncon, part, k, tpw, ub, pw):
    """refine/fm_kway.rs:33-158 transcribed independently: a binary heap of
    (-gain, v) with lazy deletion (a popped entry counts only if it is the
    vertex's current key), equal gains smallest vertex first."""
    import heapq
    n = len(x) - 1
    part = list(map(int, part))
    pw = [list(map(int, row)) for row in np.asarray(pw).reshape(k, ncon)]
    w = (lambda j: int(aw[j])) if aw is not None else (lambda j: 1)
    vwt = (lambda v, c: int(vw[v * ncon + c])) if vw is not None else (lambda v, c: 1)

    def best(v):
        et = {}
        for j in range(x[v], x[v + 1]):
            et[part[a[j]]] = et.get(part[a[j]], 0) + w(j)
        bp, bg = None, 0
        for p in range(k):
            if p != part[v] and et.get(p, 0) != 0:
                g = et[p] - et.get(part[v], 0)
                if g > bg:
                    bp, bg = p, g
        return (bp, bg) if bp is not None else None

    cur = {}
    heap = []
    for v in range(n):
        if any(part[a[j]] != part[v] for j in range(x[v], x[v + 1])):
            b = best(v)
            if b:
                cur[v] = b
                heapq.heappush(heap, (-b[1], v))
    locked = [False] * n
    moves = 0
    while heap:
        g, v = heapq.heappop(heap)
        if v not in cur or cur[v][1] != -g:
            continue                                   # stale entry
        to = cur.pop(v)[0]
        if not all(pw[to][c] + vwt(v, c) <= tpw[to * ncon + c] * ub[c] for c in range(ncon)):
            continue
        frm = part[v]
        part[v] = to
        for c in range(ncon):
            pw[frm][c] -= vwt(v, c)
            pw[to][c] += vwt(v, c)
        locked[v] = True
        moves += 1
        for j in range(x[v], x[v + 1]):
            u = int(a[j])
            if locked[u]:
                continue
            b = best(u)
            if b:
                if u not in cur or cur[u] != b:
                    cur[u] = b
                    heapq.heappush(heap, (-b[1], u))
            else:
                cur.pop(u, None)
    return moves > 0, np.array(part, np.int32), np.array(pw, np.int32), moves


def test_fm_reference_case():
    # fm_kway.rs:209-219: grid3x3, k=3, row stripes, tpwgts 3, ub 1.1
    x, a = grid_sorted(3, 3)
    part = np.arange(9) // 3
    before = O.edge_cut(x, a, None, part)
    after_part, after = O.fm_kway_refine(1, x, a, None, None, part, 3, [3.0] * 3, [1.1], 5)
    assert after <= before and after == O.edge_cut(x, a, None, after_part)


def test_fm_pass_vs_independent_transcription():
    rng = np.random.default_rng(8)
    checked = 0
    for trial in range(120):
        n = int(rng.integers(4, 60))
        x, a, aw = random_graph(n, float(rng.choice([0.05, 0.15, 0.4])), 100 + trial,
                                wmax=int(rng.choice([1, 3, 9])))
        ncon = int(rng.integers(1, 3))
        k = int(rng.integers(2, 6))
        vw = rng.integers(1, 4, n * ncon).astype(np.int32) if trial % 2 else None
        part = rng.integers(0, k, n).astype(np.int32)
        tot = (vw.reshape(n, ncon).sum(0) if vw is not None else np.full(ncon, n)).astype(np.float64)
        tpw = np.tile(tot / k, k)
        ub = np.full(ncon, float(rng.choice([1.0, 1.03, 1.2, 1.6])))
        pw = O.partition_weights(ncon, vw, part, k)
        ok_o, p_o, pw_o, nm_o = O.fm_kway_pass(ncon, x, a, vw, aw, part, k, tpw, ub, pw)
        ok_t, p_t, pw_t, nm_t = _fm_pass_transcribed(x, a, aw, vw, ncon, part, k, tpw, ub, pw)
        assert (ok_o, nm_o) == (ok_t, nm_t) and np.array_equal(p_o, p_t) and np.array_equal(pw_o, pw_t), trial
        checked += nm_o > 0
    assert checked > 60


def test_fm_invariants():
    # every kept move has positive gain, so the cut falls by exactly their
    # sum and the rollback keeps every move; after a pass that moves
    # nothing, no vertex's best move (fm_kway.rs:47-72) is feasible
    rng = np.random.default_rng(9)
    for trial in range(30):
        n = int(rng.integers(30, 200))
        x, a, aw = random_graph(n, 0.05, 300 + trial, wmax=5)
        k = int(rng.integers(2, 7))
        part = rng.integers(0, k, n).astype(np.int32)
        tpw = np.full(k, n / k)
        ub = np.array([1.05])
        pw = O.partition_weights(1, None, part, k)
        before = O.edge_cut(x, a, aw, part)
        ok, p1, pw1, nm = O.fm_kway_pass(1, x, a, None, aw, part, k, tpw, ub, pw)
        after = O.edge_cut(x, a, aw, p1)
        assert ok == (nm > 0) and (after < before if nm else after == before)
        assert np.array_equal(pw1, O.partition_weights(1, None, p1, k))
        assert (pw1[:, 0] <= np.floor(tpw * ub[0]) + 0).all() or (pw[:, 0] > np.floor(tpw * ub[0])).any()
        p_f, cut_f = O.fm_kway_refine(1, x, a, None, aw, part, k, tpw, ub, 50)
        assert cut_f == O.edge_cut(x, a, aw, p_f) and cut_f <= before
        pwf = O.partition_weights(1, None, p_f, k)[:, 0]
        for v in range(n):
            et = np.zeros(k, np.int64)
            for j in range(x[v], x[v + 1]):
                et[p_f[a[j]]] += aw[j] if aw is not None else 1
            gains = [(et[p] - et[p_f[v]], -p) for p in range(k) if p != p_f[v] and et[p] != 0]
            if gains and max(gains)[0] > 0:
                bp = -max(gains)[1]
                assert pwf[bp] + 1 > tpw[bp] * ub[0], (trial, v)


def test_part_graph_kway_fm_is_the_composition():
    # rtype Fm (kway.rs:158-160): the same driver with fm_kway_refine per level
    import rmetis_amd.synth as S
    g = S.rgg(1200, 4, vwgt_max=2)
    for seed in (1, 3):
        obj, part = _kway_composed_fm(g.xadj, g.adjncy, 8, seed, g.vwgt, g.adjwgt)
        rc, o_obj, o_part = O.part_graph_kway(g.xadj, g.adjncy, 8, vwgt=g.vwgt, adjwgt=g.adjwgt, seed=seed,
                                              rtype_greedy=False)
        assert rc == 1 and o_obj == obj and np.array_equal(o_part, part)


def _kway_composed_fm(x, a, k, seed, vwgt, adjwgt, niter=10, ufactor=30):
    ub = np.full(1, 1.0 + ufactor / 1000.0)
    r = O.rng(seed)
    cur = (x, a, vwgt, adjwgt)
    levels = []
    while True:
        cx, ca, cvw, caw = cur
        m = O.shem_matching(cx, ca, caw, r)
        cg = O.build_coarse_graph(1, cx, ca, cvw, caw, None, m)
        levels.append((cur, cg["cmap"]))
        ratio = np.float32(cg["n"]) / np.float32(len(cx) - 1)
        cur = (cg["cxadj"], cg["cadjncy"], cg["cvwgt"], cg["cadjwgt"] if cg["adjwgt_present"] else None)
        if cg["n"] <= max(20 * k, 20) or ratio > np.float32(0.95):
            break
    cx, ca, cvw, caw = cur
    tpw = O.target_weights(len(cx) - 1, 1, cvw, k)
    part = O.grow_partition(1, cx, ca, cvw, k, tpw, r)
    part, _ = O.fm_kway_refine(1, cx, ca, cvw, caw, part, k, tpw, ub, niter)
    for (fx, fa, fvw, faw), cmap in reversed(levels):
        part = part[cmap]
        tpw = O.target_weights(len(fx) - 1, 1, fvw, k)
        part, _ = O.fm_kway_refine(1, fx, fa, fvw, faw, part, k, tpw, ub, niter)
    return O.edge_cut(x, a, adjwgt, part), part
