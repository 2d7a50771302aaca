"""ctypes binding of the oracle (TEST INFRASTRUCTURE ONLY).

The oracle is a CPU restatement of the javagg/rmetis k-way greedy path
(oracle.cpp).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may load this module; the product never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")
_lib = None

i32p = C.POINTER(C.c_int32)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)


class Rng(C.Structure):
    _fields_ = [("s", C.c_uint64 * 4)]


class Stats(C.Structure):
    _fields_ = [("t_coarsen_s", C.c_double), ("t_initial_s", C.c_double),
                ("t_refine_s", C.c_double), ("coarsen_edges", C.c_int64),
                ("refine_edges", C.c_int64), ("nlevels", C.c_int32),
                ("coarsest_n", C.c_int32)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        L.orc_rng_next_u64.restype = C.c_uint64
        L.orc_rng_next_u32.restype = C.c_uint32
        L.orc_rng_range_incl.restype = C.c_uint64
        L.orc_rng_range_incl.argtypes = [C.POINTER(Rng), C.c_uint64, C.c_uint64]
        L.orc_rng_seed_from_u64.argtypes = [C.POINTER(Rng), C.c_uint64]
        L.orc_splitmix64_next.restype = C.c_uint64
        L.orc_splitmix64_next.argtypes = [C.POINTER(C.c_uint64)]
        L.orc_part_graph_kway.argtypes = [C.c_int32, C.c_int32, i32p, i32p, i32p, i32p, i32p,
                                          C.c_int32, f32p, f32p, C.c_int32, C.c_int32, C.c_int32,
                                          C.c_int32, C.c_int32, C.c_uint64, C.c_int32, i32p, i32p,
                                          C.POINTER(Stats)]
        L.orc_part_graph_recursive.argtypes = [C.c_int32, C.c_int32, i32p, i32p, i32p, i32p, C.c_int32, f32p,
                                               f32p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint64,
                                               C.c_int32, i32p, i32p]
        L.orc_node_nd.argtypes = [C.c_int32, i32p, i32p, i32p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                  C.c_uint64, C.c_int32, i32p, i32p]
        _lib = L
    return _lib


def _p(a, t=i32p):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def rng(seed):
    r = Rng()
    lib().orc_rng_seed_from_u64(C.byref(r), C.c_uint64(seed))
    return r


def rng_from_words(words):
    r = Rng()
    arr = (C.c_uint64 * 4)(*words)
    lib().orc_rng_from_words(C.byref(r), arr)
    return r


def next_u64(r):
    return lib().orc_rng_next_u64(C.byref(r))


def next_u32(r):
    return lib().orc_rng_next_u32(C.byref(r))


def range_incl(r, lo, hi):
    return lib().orc_rng_range_incl(C.byref(r), lo, hi)


def splitmix64(state):
    s = C.c_uint64(state)
    out = []
    for _ in range(4):
        out.append(lib().orc_splitmix64_next(C.byref(s)))
    return out


def edge_cut(xadj, adjncy, adjwgt, part):
    xadj, adjncy, adjwgt, part = map(_i32, (xadj, adjncy, adjwgt, part))
    return lib().orc_edge_cut(len(xadj) - 1, _p(xadj), _p(adjncy), _p(adjwgt), _p(part))


def partition_weights(ncon, vwgt, part, nparts):
    part, vwgt = _i32(part), _i32(vwgt)
    pw = np.zeros(nparts * ncon, np.int32)
    lib().orc_partition_weights(len(part), ncon, _p(vwgt), _p(part), nparts, _p(pw))
    return pw.reshape(nparts, ncon)


def total_weight(n, ncon, vwgt):
    vwgt = _i32(vwgt)
    tw = np.zeros(ncon, np.int32)
    lib().orc_total_weight(n, ncon, _p(vwgt), _p(tw))
    return tw


def boundary_vertices(xadj, adjncy, part):
    xadj, adjncy, part = map(_i32, (xadj, adjncy, part))
    out = np.zeros(len(part), np.int32)
    k = lib().orc_boundary_vertices(len(xadj) - 1, _p(xadj), _p(adjncy), _p(part), _p(out))
    return out[:k]


def target_weights(n, ncon, vwgt, nparts, tpwgts=None):
    vwgt = _i32(vwgt)
    tp = None if tpwgts is None else np.ascontiguousarray(tpwgts, np.float32)
    out = np.zeros(nparts * ncon, np.float64)
    lib().orc_compute_target_weights(n, ncon, _p(vwgt), nparts, _p(tp, f32p), _p(out, f64p))
    return out


def validate_csr(n, ncon, xadj, adjncy, vwgt=None, adjwgt=None):
    xadj, adjncy, vwgt, adjwgt = map(_i32, (xadj, adjncy, vwgt, adjwgt))
    L = lib()
    L.orc_validate_csr.argtypes = [C.c_int32, C.c_int32, i32p, C.c_int64, i32p, C.c_int64,
                                   i32p, C.c_int64, i32p, C.c_int64]
    return L.orc_validate_csr(n, ncon, _p(xadj), len(xadj), _p(adjncy), len(adjncy),
                              _p(vwgt), 0 if vwgt is None else len(vwgt),
                              _p(adjwgt), 0 if adjwgt is None else len(adjwgt))


def move_gain(xadj, adjncy, adjwgt, v, part, frm, to):
    xadj, adjncy, adjwgt, part = map(_i32, (xadj, adjncy, adjwgt, part))
    return lib().orc_compute_move_gain(_p(xadj), _p(adjncy), _p(adjwgt), int(v), _p(part), int(frm),
                                       int(to))


def greedy_pass(ncon, xadj, adjncy, vwgt, adjwgt, part, nparts, tpw, ub, mode, pw):
    """One pass; returns (new_part, new_pw, moves[(v,p)...])."""
    xadj, adjncy, vwgt, adjwgt = map(_i32, (xadj, adjncy, vwgt, adjwgt))
    part = np.array(part, dtype=np.int32)
    pw = np.array(pw, dtype=np.int32).reshape(-1)
    tpw = np.ascontiguousarray(tpw, np.float64)
    ub = np.ascontiguousarray(ub, np.float64)
    moves = np.zeros(2 * len(part) + 2, np.int32)
    nm = lib().orc_greedy_pass(len(part), ncon, _p(xadj), _p(adjncy), _p(vwgt), _p(adjwgt),
                               _p(part), nparts, _p(tpw, f64p), _p(ub, f64p), mode, _p(pw),
                               _p(moves))
    return part, pw.reshape(nparts, ncon), moves[:2 * nm].reshape(-1, 2)


def greedy_refine(ncon, xadj, adjncy, vwgt, adjwgt, part, nparts, tpw, ub, niter):
    xadj, adjncy, vwgt, adjwgt = map(_i32, (xadj, adjncy, vwgt, adjwgt))
    part = np.array(part, dtype=np.int32)
    tpw = np.ascontiguousarray(tpw, np.float64)
    ub = np.ascontiguousarray(ub, np.float64)
    cut = lib().orc_greedy_refine(len(part), ncon, _p(xadj), _p(adjncy), _p(vwgt), _p(adjwgt),
                                  _p(part), nparts, _p(tpw, f64p), _p(ub, f64p), niter)
    return part, cut


def fm_kway_refine(ncon, xadj, adjncy, vwgt, adjwgt, part, nparts, tpw, ub, niter):
    xadj, adjncy, vwgt, adjwgt = map(_i32, (xadj, adjncy, vwgt, adjwgt))
    part = np.array(part, dtype=np.int32)
    tpw = np.ascontiguousarray(tpw, np.float64)
    ub = np.ascontiguousarray(ub, np.float64)
    cut = lib().orc_fm_kway_refine(len(part), ncon, _p(xadj), _p(adjncy), _p(vwgt), _p(adjwgt),
                                   _p(part), nparts, _p(tpw, f64p), _p(ub, f64p), niter)
    return part, cut


def fm_kway_pass(ncon, xadj, adjncy, vwgt, adjwgt, part, nparts, tpw, ub, pw):
    """One FM pass; returns (improved, new_part, new_pw, kept moves)."""
    xadj, adjncy, vwgt, adjwgt = map(_i32, (xadj, adjncy, vwgt, adjwgt))
    part = np.array(part, dtype=np.int32)
    pw = np.array(pw, dtype=np.int32).reshape(-1)
    tpw = np.ascontiguousarray(tpw, np.float64)
    ub = np.ascontiguousarray(ub, np.float64)
    nm = C.c_int32()
    ok = lib().orc_fm_kway_pass(len(part), ncon, _p(xadj), _p(adjncy), _p(vwgt), _p(adjwgt), _p(part),
                                nparts, _p(tpw, f64p), _p(ub, f64p), _p(pw), C.byref(nm))
    return bool(ok), part, pw.reshape(nparts, ncon), nm.value


def fm2_refine(ncon, xadj, adjncy, vwgt, adjwgt, part, tpw, ub, niter):
    """refine/fm.rs:12-28 (2-way FM); returns (part, cut)."""
    xadj, adjncy, vwgt, adjwgt = map(_i32, (xadj, adjncy, vwgt, adjwgt))
    part = np.array(part, dtype=np.int32)
    tpw = np.ascontiguousarray(tpw, np.float64)
    ub = np.ascontiguousarray(ub, np.float64)
    cut = lib().orc_fm2_refine(len(part), ncon, _p(xadj), _p(adjncy), _p(vwgt), _p(adjwgt), _p(part),
                               _p(tpw, f64p), _p(ub, f64p), niter)
    return part, cut


def fm2_pass(ncon, xadj, adjncy, vwgt, adjwgt, part, tpw, ub):
    """One 2-way FM pass (fm.rs:31-145); returns (improved, new part)."""
    xadj, adjncy, vwgt, adjwgt = map(_i32, (xadj, adjncy, vwgt, adjwgt))
    part = np.array(part, dtype=np.int32)
    tpw = np.ascontiguousarray(tpw, np.float64)
    ub = np.ascontiguousarray(ub, np.float64)
    ok = lib().orc_fm2_pass(len(part), ncon, _p(xadj), _p(adjncy), _p(vwgt), _p(adjwgt), _p(part),
                            _p(tpw, f64p), _p(ub, f64p))
    return bool(ok), part


def part_graph_recursive(xadj, adjncy, nparts, vwgt=None, adjwgt=None, ncon=1, tpwgts=None, ubvec=None,
                         ctype_shem=True, iptype=4, rtype_greedy=False, niter=10, seed=1, ufactor=1):
    """partition/recursive.rs:11-187 (defaults: Options::for_recursive())."""
    xadj, adjncy, vwgt, adjwgt = map(_i32, (xadj, adjncy, vwgt, adjwgt))
    n = len(xadj) - 1
    part = np.zeros(n, np.int32)
    obj = C.c_int32()
    tp = None if tpwgts is None else np.ascontiguousarray(tpwgts, np.float32)
    ub = None if ubvec is None else np.ascontiguousarray(ubvec, np.float32)
    rc = lib().orc_part_graph_recursive(n, ncon, _p(xadj), _p(adjncy), _p(vwgt), _p(adjwgt), nparts, _p(tp, f32p),
                                        _p(ub, f32p), int(ctype_shem), iptype, int(rtype_greedy), niter, seed,
                                        ufactor, C.byref(obj), _p(part))
    return rc, obj.value, part


def node_nd(xadj, adjncy, vwgt=None, ctype_shem=True, iptype=0, rtype_greedy=False, niter=10, seed=1,
            ufactor=30):
    """partition/nd.rs:17-120 (defaults: Options::default()); returns (perm, iperm)."""
    xadj, adjncy, vwgt = map(_i32, (xadj, adjncy, vwgt))
    n = len(xadj) - 1
    perm = np.zeros(n, np.int32)
    iperm = np.zeros(n, np.int32)
    lib().orc_node_nd(n, _p(xadj), _p(adjncy), _p(vwgt), int(ctype_shem), iptype, int(rtype_greedy), niter, seed,
                      ufactor, _p(perm), _p(iperm))
    return perm, iperm


def shem_matching(xadj, adjncy, adjwgt, r):
    xadj, adjncy, adjwgt = map(_i32, (xadj, adjncy, adjwgt))
    m = np.zeros(len(xadj) - 1, np.int32)
    lib().orc_shem_matching(len(xadj) - 1, _p(xadj), _p(adjncy), _p(adjwgt), C.byref(r), _p(m))
    return m


def random_matching(xadj, adjncy, r):
    xadj, adjncy = map(_i32, (xadj, adjncy))
    m = np.zeros(len(xadj) - 1, np.int32)
    lib().orc_random_matching(len(xadj) - 1, _p(xadj), _p(adjncy), C.byref(r), _p(m))
    return m


def build_coarse_graph(ncon, xadj, adjncy, vwgt, adjwgt, vsize, matching):
    """Returns dict(cxadj, cadjncy, cadjwgt, cvwgt|None, cvsize|None, cmap, adjwgt_present)."""
    xadj, adjncy, vwgt, adjwgt, vsize, matching = map(
        _i32, (xadj, adjncy, vwgt, adjwgt, vsize, matching))
    n = len(xadj) - 1
    m = int(xadj[-1])
    cxadj = np.zeros(n + 1, np.int32)
    cadj = np.zeros(max(m, 1), np.int32)
    cw = np.zeros(max(m, 1), np.int32)
    cvw = np.zeros(max(n * ncon, 1), np.int32)
    cvs = np.zeros(max(n, 1), np.int32)
    cmap = np.zeros(n, np.int32)
    cm = C.c_int32()
    flags = C.c_int32()
    cn = lib().orc_build_coarse_graph(n, ncon, _p(xadj), _p(adjncy), _p(vwgt), _p(adjwgt),
                                      _p(vsize), _p(matching), _p(cxadj), _p(cadj), _p(cw),
                                      _p(cvw), _p(cvs), _p(cmap), C.byref(cm), C.byref(flags))
    f = flags.value
    return dict(n=cn, cxadj=cxadj[:cn + 1].copy(), cadjncy=cadj[:cm.value].copy(),
                cadjwgt=cw[:cm.value].copy(),
                cvwgt=cvw[:cn * ncon].copy() if f & 1 else None,
                cvsize=cvs[:cn].copy() if f & 2 else None,
                cmap=cmap, adjwgt_present=bool(f & 4))


def grow_partition(ncon, xadj, adjncy, vwgt, nparts, tpw, r):
    xadj, adjncy, vwgt = map(_i32, (xadj, adjncy, vwgt))
    n = len(xadj) - 1
    part = np.zeros(n, np.int32)
    tpw = np.ascontiguousarray(tpw, np.float64)
    rc = lib().orc_grow_partition(n, ncon, _p(xadj), _p(adjncy), _p(vwgt), nparts,
                                  _p(tpw, f64p), C.byref(r), _p(part))
    if rc != 0:
        raise ValueError("graph has fewer vertices than parts")
    return part


def part_graph_kway(xadj, adjncy, nparts, vwgt=None, adjwgt=None, vsize=None, ncon=1,
                    tpwgts=None, ubvec=None, ctype_shem=True, iptype=0, rtype_greedy=True,
                    ncuts=1, niter=10, seed=1, ufactor=30, want_stats=False):
    xadj, adjncy, vwgt, adjwgt, vsize = map(_i32, (xadj, adjncy, vwgt, adjwgt, vsize))
    n = len(xadj) - 1
    part = np.zeros(n, np.int32)
    obj = C.c_int32()
    tp = None if tpwgts is None else np.ascontiguousarray(tpwgts, np.float32)
    ub = None if ubvec is None else np.ascontiguousarray(ubvec, np.float32)
    st = Stats()
    rc = lib().orc_part_graph_kway(n, ncon, _p(xadj), _p(adjncy), _p(vwgt), _p(vsize),
                                   _p(adjwgt), nparts, _p(tp, f32p), _p(ub, f32p),
                                   int(ctype_shem), iptype, int(rtype_greedy), ncuts, niter,
                                   seed, ufactor, C.byref(obj), _p(part), C.byref(st))
    if want_stats:
        return rc, obj.value, part, st
    return rc, obj.value, part
