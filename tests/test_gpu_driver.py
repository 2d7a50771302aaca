"""The k-way driver's own rules on the GPU path (partition/kway.rs:20-141):
user target weights, the coarsening stop rule, random matching end to end,
and the part-weight table carried across projections.

* tpwgts: target[p][c] = tpwgts[p*ncon+c] as f64 * total[c]
  (graph/mod.rs:392-410) at every level.  Judged like the uniform case: GPU
  geomean cut <= 1.03 x the oracle's over seeds 1..8 and no more parts over
  target * ub than the oracle's partitions of the same seeds.
* Stop rule (coarsen/mod.rs:36-71): at least one level; the loop continues
  only while n' > max(20k, 20) and the f32 ratio n'/n <= 0.95; the last level
  satisfies the stop condition.  Read from the per-level statistics of
  rmetis_hip_kway_device.
* RM (ctype Rm, coarsen/rm.rs:13-43): judged end to end against the
  oracle's RM like SHEM.
* With explicit vertex weights the part-weight table of the coarser level is
  carried across a projection instead of recomputed; a debug toggle turns
  the carry off and the partitions must be identical.
"""
import ctypes as C

import numpy as np
import pytest

import rmetis_amd as R
import rmetis_amd.synth as S
from rmetis_amd import _lib
from graphs import build_undirected, grid_sorted
from test_gpu_quality import _check, _run

pytestmark = pytest.mark.gpu


def test_tpwgts_one_constraint():
    g = S.rgg(1 << 15, 3, vwgt_max=4)
    k = 8
    tp = np.array([0.05, 0.1, 0.15, 0.2, 0.1, 0.1, 0.2, 0.1], np.float32)
    _check("rgg2^15 k=8 tpwgts", [_run(g, k, tpwgts=tp)])


def test_tpwgts_two_constraints_and_ubvec():
    # each part's target fraction is the same for both constraints (with
    # independent random fractions per constraint the targets are rarely
    # reachable together: the oracle's own partitions of that case leave
    # dozens of parts over)
    g, _ = S._batch_member((3, 20000, 16, 99, 2, 8))
    k = 16
    rng = np.random.default_rng(2)
    w = rng.uniform(0.5, 1.5, k)
    tp = np.repeat(w / w.sum(), 2).astype(np.float32)
    _check("rgg 20k ncon=2 k=16 tpwgts", [_run(g, k, tpwgts=tp, ubvec=[1.05, 1.08])])


def _levels(g, k, seed=1):
    """(fine n of each coarsening pass, coarsest n) of one partition."""
    DA = _lib.DeviceArray
    dx, da = DA(g[0]), DA(g[1])
    n = len(g[0]) - 1
    dp = DA(count=n, dtype=np.int32)
    st = _lib.HipStats()
    obj = C.c_int32()
    opts = R.Options(rtype=R.RType.Greedy, seed=seed).to_c_array()
    rc = _lib.lib().rmetis_hip_kway_device(n, 1, dx.ptr, da.ptr, None, None, k, None, None,
                                           opts.ctypes.data_as(_lib.i32p), C.byref(obj), dp.ptr, C.byref(st))
    assert rc == 1, _lib.last_error()
    # one entry per coarsening pass (its fine level), then the coarsest level
    fine = [st.level_n[i] for i in range(st.nlevels - 1)]
    return fine, st.coarsest_n, st.level_n[st.nlevels - 1]


def _check_stop_rule(fine, coarsest, k):
    limit = max(20 * k, 20)
    assert len(fine) >= 1                                   # always at least one level
    sizes = fine + [coarsest]
    for i in range(len(fine)):
        cn, n = sizes[i + 1], sizes[i]
        stop = cn <= limit or np.float32(cn) / np.float32(n) > np.float32(0.95)
        assert stop == (i == len(fine) - 1), (i, n, cn)


def test_stop_rule_star_stalls_after_one_level():
    # hub 0 with 199 leaves: one pair matches, n' = 199, ratio 0.995 > 0.95
    x, a = build_undirected(200, [(0, i) for i in range(1, 200)])
    fine, cn, last = _levels((x, a), 2)
    assert fine == [200] and cn == 199 == last


def test_stop_rule_mostly_isolated():
    # a 10x10 grid (pairs up to 50 coarse vertices) beside 1,000 isolated
    # vertices: n' >= 1,050, ratio >= 0.954 > 0.95 after the first level
    xg, ag = grid_sorted(10, 10)
    x = np.concatenate([xg, np.full(1000, xg[-1], np.int32)])
    fine, cn, _ = _levels((x, ag), 4)
    assert fine == [1100] and 1050 <= cn <= 1100
    _check_stop_rule(fine, cn, 4)


def test_stop_rule_small_graph_still_coarsens_once():
    x, a = grid_sorted(8, 8)                    # n 64 below max(20k, 20) = 80
    fine, cn, _ = _levels((x, a), 4)
    assert fine == [64] and cn <= 64
    _check_stop_rule(fine, cn, 4)


@pytest.mark.parametrize("rows,cols,k", [(256, 256, 8), (64, 300, 2), (100, 100, 64)])
def test_stop_rule_grids(rows, cols, k):
    x, a = grid_sorted(rows, cols)
    for seed in (1, 2):
        fine, cn, _ = _levels((x, a), k, seed)
        _check_stop_rule(fine, cn, k)


def test_random_matching_end_to_end():
    res = [_run(S.rgg(1 << 16, 5), 16, ctype=R.CType.Rm), _run(S.grid(128, 96), 8, ctype=R.CType.Rm)]
    _check("RM: rgg2^16 k=16, grid128x96 k=8", res)


@pytest.mark.parametrize("ncon", [1, 2])
def test_part_weight_carry_equals_recompute(ncon):
    g, _ = S._batch_member((1, 60000, 32, 1234, ncon, 4))
    rg = R.Graph(g.n, ncon, g.xadj, g.adjncy, g.vwgt, g.adjwgt)
    ub = None if ncon == 1 else [1.05] * ncon
    try:
        for seed in (1, 2, 3):
            opts = R.Options(rtype=R.RType.Greedy, seed=seed)
            _lib.lib().rmetis_hip_debug_flags(0)
            r1 = R.part_graph_kway(rg, 32, ubvec=ub, options=opts)
            _lib.lib().rmetis_hip_debug_flags(1)
            r2 = R.part_graph_kway(rg, 32, ubvec=ub, options=opts)
            assert r1.objval == r2.objval and np.array_equal(r1.part, r2.part), seed
    finally:
        _lib.lib().rmetis_hip_debug_flags(0)
