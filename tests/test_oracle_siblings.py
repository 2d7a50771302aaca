"""The oracle's restatements of the C-ABI siblings of the k-way path: the
2-way FM of refine/fm.rs (recursive bisection's refinement), multilevel
recursive bisection (partition/recursive.rs) and nested dissection
(partition/nd.rs with separator/mod.rs).  Each is pinned by an independent
Python transcription over pieces pinned in test_oracle.py, by the
reference's own test cases and by invariants."""
import heapq

import numpy as np
import pytest

import pyoracle as O
from graphs import cycle4, grid_sorted, path, random_graph


def _fm2_pass_transcribed(x, a, aw, vw, ncon, part, tpw, ub):
    """refine/fm.rs:31-145 transcribed independently: a binary heap of
    (-gain, v) with lazy deletion (an entry counts only while it is the
    vertex's current key), equal gains smallest vertex first."""
    n = len(x) - 1
    part = list(map(int, part))
    w = (lambda j: int(aw[j])) if aw is not None else (lambda j: 1)
    vwt = (lambda v, c: int(vw[v * ncon + c])) if vw is not None else (lambda v, c: 1)
    pw = [[0] * ncon for _ in range(2)]
    for v in range(n):
        for c in range(ncon):
            pw[part[v]][c] += vwt(v, c)

    def gain(v):
        ext = sum(w(j) for j in range(x[v], x[v + 1]) if part[a[j]] != part[v])
        inn = sum(w(j) for j in range(x[v], x[v + 1]) if part[a[j]] == part[v])
        return ext - inn

    def boundary(v):
        return any(part[a[j]] != part[v] for j in range(x[v], x[v + 1]))

    def cut():
        return sum(w(j) for v in range(n) for j in range(x[v], x[v + 1]) if part[a[j]] != part[v]) // 2

    key, heap = {}, []
    for v in range(n):
        if boundary(v):
            key[v] = gain(v)
            heapq.heappush(heap, (-key[v], v))
    locked = [False] * n
    moves = []
    cur = best = cut()
    best_len = 0
    while heap:
        g, v = heapq.heappop(heap)
        if key.get(v) != -g:
            continue                                   # stale entry
        del key[v]
        if locked[v]:
            continue
        frm, to = part[v], 1 - part[v]
        if not all(pw[to][c] + vwt(v, c) <= tpw[to * ncon + c] * ub[c] for c in range(ncon)):
            continue
        part[v] = to
        for c in range(ncon):
            pw[frm][c] -= vwt(v, c)
            pw[to][c] += vwt(v, c)
        locked[v] = True
        moves.append((v, frm))
        cur -= -g
        if cur < best:
            best, best_len = cur, len(moves)
        for j in range(x[v], x[v + 1]):
            u = int(a[j])
            if locked[u]:
                continue
            if u in key:
                key[u] = gain(u)
                heapq.heappush(heap, (-key[u], u))
            elif boundary(u):
                key[u] = gain(u)
                heapq.heappush(heap, (-key[u], u))
    improved = best_len > 0 and best < cut()
    for v, frm in reversed(moves[best_len:]):
        part[v] = frm
    return improved, np.array(part, np.int32)


def test_fm2_pass_vs_independent_transcription():
    rng = np.random.default_rng(4)
    cases = 0
    for trial in range(60):
        n = int(rng.integers(6, 120))
        x, a, aw = random_graph(n, float(rng.choice([0.05, 0.1, 0.25])), trial, wmax=int(rng.choice([1, 4])))
        ncon = int(rng.integers(1, 3))
        vw = rng.integers(1, 5, n * ncon).astype(np.int32) if trial % 2 else None
        part = rng.integers(0, 2, n).astype(np.int32)
        tot = (vw.reshape(n, ncon).sum(0) if vw is not None else np.full(ncon, n)).astype(np.float64)
        frac = float(rng.choice([0.5, 0.3]))
        tpw = np.concatenate([frac * tot, (1 - frac) * tot])
        ub = np.full(ncon, float(rng.choice([1.001, 1.05, 1.3])))
        want = _fm2_pass_transcribed(x, a, aw if aw is not None and aw.max() > 1 else None, vw, ncon, part, tpw, ub)
        got = O.fm2_pass(ncon, x, a, vw, aw if aw is not None and aw.max() > 1 else None, part, tpw, ub)
        assert got[0] == want[0] and np.array_equal(got[1], want[1]), trial
        cases += 1
    assert cases == 60


def test_fm2_reference_case_and_invariants():
    # refine/fm.rs:152-159: the cut of cycle6 {0,0,0,1,1,1} does not increase
    x = np.array([0, 2, 4, 6, 8, 10, 12], np.int32)
    a = np.array([1, 5, 0, 2, 1, 3, 2, 4, 3, 5, 4, 0], np.int32)
    part0 = np.array([0, 0, 0, 1, 1, 1], np.int32)
    before = O.edge_cut(x, a, None, part0)
    _, after = O.fm2_refine(1, x, a, None, None, part0, np.array([3.0, 3.0]), np.array([1.05]), 10)
    assert after <= before
    # each pass keeps its best prefix: the cut never rises over passes
    rng = np.random.default_rng(7)
    for trial in range(20):
        n = int(rng.integers(20, 200))
        x, a, aw = random_graph(n, 0.05, 100 + trial, wmax=3)
        part = rng.integers(0, 2, n).astype(np.int32)
        c0 = O.edge_cut(x, a, aw, part)
        p2, c2 = O.fm2_refine(1, x, a, None, aw, part, np.array([n / 2, n / 2]), np.array([1.1]), 10)
        assert c2 <= c0 and c2 == O.edge_cut(x, a, aw, p2)


def _extract(x, a, aw, vw, ncon, part, t):
    """graph/mod.rs:209-270"""
    keep = np.flatnonzero(np.asarray(part) == t)
    o2n = np.full(len(x) - 1, -1, np.int64)
    o2n[keep] = np.arange(len(keep))
    nx, na, nw = [0], [], []
    for v in keep:
        for j in range(x[v], x[v + 1]):
            if part[a[j]] == t:
                na.append(int(o2n[a[j]]))
                if aw is not None:
                    nw.append(int(aw[j]))
        nx.append(len(na))
    svw = None if vw is None else np.asarray(vw).reshape(-1, ncon)[keep].reshape(-1).astype(np.int32)
    return (np.array(nx, np.int32), np.array(na, np.int32), None if aw is None else np.array(nw, np.int32), svw,
            keep)


def _bisect_composed(x, a, aw, vw, ncon, tp2, ub, r, greedy, niter=10, iptype=4):
    """partition/recursive.rs:140-187 over the pinned pieces (nparts 2:
    coarsening limit 40)."""
    n = len(x) - 1
    if n <= 2:
        return np.arange(n, dtype=np.int32)
    cur, levels = (x, a, vw, aw), []
    while True:
        cx, ca, cvw, caw = cur
        m = O.shem_matching(cx, ca, caw, r)
        cg = O.build_coarse_graph(ncon, cx, ca, cvw, caw, None, m)
        nxt = (cg["cxadj"], cg["cadjncy"], cg["cvwgt"], cg["cadjwgt"] if cg["adjwgt_present"] else None)
        levels.append((cur, cg["cmap"]))
        ratio = np.float32(cg["n"]) / np.float32(len(cx) - 1)
        cur = nxt
        if cg["n"] <= 40 or ratio > np.float32(0.95):
            break

    def refine(gx, ga, gvw, gaw, part, tpw):
        if greedy:
            return O.greedy_refine(ncon, gx, ga, gvw, gaw, part, 2, tpw, ub, niter)[0]
        return O.fm2_refine(ncon, gx, ga, gvw, gaw, part, tpw, ub, niter)[0]

    cx, ca, cvw, caw = cur
    tpw = O.target_weights(len(cx) - 1, ncon, cvw, 2, tp2)
    if iptype == 1:
        part = np.array([O.range_incl(r, 0, 1) for _ in range(len(cx) - 1)], np.int32)
    else:
        part = O.grow_partition(ncon, cx, ca, cvw, 2, tpw, r)
    part = refine(cx, ca, cvw, caw, part, tpw)
    for (fx, fa, fvw, faw), cmap in reversed(levels):
        part = part[cmap]
        part = refine(fx, fa, fvw, faw, part, O.target_weights(len(fx) - 1, ncon, fvw, 2, tp2))
    return part


def _rb_composed(x, a, aw, vw, ncon, k, tp, ub, r, greedy):
    """partition/recursive.rs:60-137 with f32 sums of the targets."""
    n = len(x) - 1
    if k == 1:
        return np.zeros(n, np.int32)
    if n <= k:
        return np.minimum(np.arange(n), k - 1).astype(np.int32)
    k1, k2 = k // 2, k - k // 2
    tp = np.asarray(tp, np.float32).reshape(k, ncon)
    lw = [np.float32(0)] * ncon
    rw = [np.float32(0)] * ncon
    for c in range(ncon):
        for p in range(k1):
            lw[c] = np.float32(lw[c] + tp[p, c])
        for p in range(k1, k):
            rw[c] = np.float32(rw[c] + tp[p, c])
    total = np.float32(lw[0] + rw[0])
    tp2 = np.array([np.float32(w / total) for w in lw] + [np.float32(w / total) for w in rw], np.float32)
    bp = _bisect_composed(x, a, aw, vw, ncon, tp2, ub, r, greedy)
    x0, a0, w0, v0, n2o0 = _extract(x, a, aw, vw, ncon, bp, 0)
    x1, a1, w1, v1, n2o1 = _extract(x, a, aw, vw, ncon, bp, 1)
    t0 = (tp[:k1] / np.float32(lw[0])).astype(np.float32)
    t1 = (tp[k1:] / np.float32(rw[0])).astype(np.float32)
    p0 = _rb_composed(x0, a0, w0, v0, ncon, k1, t0, ub, r, greedy)
    p1 = _rb_composed(x1, a1, w1, v1, ncon, k2, t1, ub, r, greedy)
    part = np.zeros(n, np.int32)
    part[n2o0] = p0
    part[n2o1] = p1 + k1
    return part


@pytest.mark.parametrize("greedy", [False, True])
def test_part_graph_recursive_is_the_composition_of_its_pinned_pieces(greedy):
    import rmetis_amd.synth as S
    cases = [(grid_sorted(20, 18) + (None, None, 1), 4), (grid_sorted(9, 31) + (None, None, 1), 7)]
    g = S.rgg(900, 3, vwgt_max=3)
    cases.append(((g.xadj, g.adjncy, g.adjwgt, g.vwgt, 1), 5))
    g2 = S.rgg(700, 4, vwgt_max=4, ncon=2)
    cases.append(((g2.xadj, g2.adjncy, g2.adjwgt, g2.vwgt, 2), 3))
    for (x, a, aw, vw, ncon), k in cases:
        for seed, tp in ((1, None), (6, np.linspace(1, 2, k * ncon).astype(np.float32))):
            if tp is not None:
                tp = (tp.reshape(k, ncon) / tp.reshape(k, ncon).sum(0)).astype(np.float32).reshape(-1)
            ub = np.full(ncon, 1.0 + 1 / 1000.0)
            want = _rb_composed(x, a, aw, vw, ncon, k, tp if tp is not None else np.full(k * ncon, np.float32(1.0 / k)),
                                ub, O.rng(seed), greedy)
            rc, obj, got = O.part_graph_recursive(x, a, k, vwgt=vw, adjwgt=aw, ncon=ncon, tpwgts=tp, seed=seed,
                                                 rtype_greedy=greedy, ufactor=1)
            assert rc == 1 and np.array_equal(got, want), (k, seed)
            assert obj == O.edge_cut(x, a, aw, got)


def test_recursive_reference_cases():
    # tests/correctness.rs:130-152, 189-197 (Options::for_recursive)
    x, a = cycle4()
    rc, obj, part = O.part_graph_recursive(x, a, 2, seed=7)
    assert rc == 1 and sorted(np.bincount(part, minlength=2)) == [2, 2]
    x, a = grid_sorted(4, 4)
    for k in (4, 8):
        rc, obj, part = O.part_graph_recursive(x, a, k, seed=42)
        assert rc == 1 and part.min() >= 0 and part.max() < k and obj == O.edge_cut(x, a, None, part)
    r1 = O.part_graph_recursive(x, a, 4, seed=99)
    r2 = O.part_graph_recursive(x, a, 4, seed=99)
    assert r1[1] == r2[1] and np.array_equal(r1[2], r2[2])
    assert O.part_graph_recursive(x, a, 1, seed=1)[0] == -2          # recursive.rs:18-20
    assert O.part_graph_recursive(x, a, 17, seed=1)[0] == -2         # recursive.rs:21-26


def _nd_composed(x, a, vw, seed):
    """partition/nd.rs:48-120 with separator/mod.rs:10-61, over the pinned
    recursive bisection."""
    n = len(x) - 1
    perm = np.zeros(n, np.int32)
    counter = [0]

    def rec(gx, ga, gvw, orig):
        gn = len(gx) - 1
        if gn <= 100:
            for o in orig:
                perm[o] = counter[0]
                counter[0] += 1
            return
        rc, _, part = O.part_graph_recursive(gx, ga, 2, vwgt=gvw, tpwgts=np.array([0.5, 0.5], np.float32),
                                             ubvec=np.array([1.05], np.float32), iptype=0, ufactor=30,
                                             seed=seed)
        if rc != 1:
            part = (np.arange(gn) >= gn // 2).astype(np.int32)
        pw = O.partition_weights(1, gvw, part, 2)
        small = 0 if pw[0] <= pw[1] else 1
        sep = [v for v in range(gn) if part[v] == small and any(part[ga[j]] != part[v] for j in range(gx[v], gx[v + 1]))]
        ins = set(sep)
        left = [v for v in range(gn) if v not in ins and part[v] == small]
        right = [v for v in range(gn) if v not in ins and part[v] != small]
        if len(sep) >= gn // 2:
            for o in orig:
                perm[o] = counter[0]
                counter[0] += 1
            return
        side = np.full(gn, 2, np.int32)
        side[left] = 0
        side[right] = 1
        for t, lst in ((0, left), (1, right)):
            if lst:
                sx, sa, _, svw, n2o = _extract(gx, ga, None, gvw, 1, side, t)
                rec(sx, sa, svw, [orig[i] for i in n2o])
        for v in sep:
            perm[orig[v]] = counter[0]
            counter[0] += 1

    rec(x, a, vw, list(range(n)))
    return perm


def test_node_nd_is_the_composition_and_a_permutation():
    import rmetis_amd.synth as S
    g = S.grid(30, 24)
    cases = [(g.xadj, g.adjncy, None, 3), (path(250)[0], path(250)[1], None, 42)]
    r = S.rgg(1200, 5, vwgt_max=4)
    cases.append((r.xadj, r.adjncy, r.vwgt, 8))
    for x, a, vw, seed in cases:
        perm, iperm = O.node_nd(x, a, vwgt=vw, seed=seed)
        assert np.array_equal(np.sort(perm), np.arange(len(x) - 1))
        assert np.array_equal(iperm[perm], np.arange(len(x) - 1))
        assert np.array_equal(perm, _nd_composed(x, a, vw, seed)), seed
    # nd.rs:125-150 (path of 10, seed 42) and tests/correctness.rs:158-173
    for x, a, seed in (path(10) + (42,), cycle4() + (1,), grid_sorted(4, 4) + (42,)):
        perm, iperm = O.node_nd(x, a, seed=seed)
        assert np.array_equal(np.sort(perm), np.arange(len(x) - 1))
        assert np.array_equal(perm[iperm], np.arange(len(x) - 1))
