"""CPU-side checks of the product library: it loads, exports every symbol that
include/rmetis.h declares, and its host-side boundary logic (CSR validation,
option parsing, argument checks, error codes) follows src/ffi/c_api.rs and
src/graph/validate.rs.  No compute call reaches a GPU here."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

import rmetis_amd as R
from rmetis_amd import _lib, synth
from graphs import cycle4
import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "rmetis.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:RMETIS_|rmetis_)\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    syms = _declared_symbols()
    assert "RMETIS_PartGraphKway" in syms and "RMETIS_SetDefaultOptions" in syms
    lib = C.CDLL(_lib.LIB_PATH)
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(_lib.SIGNATURES), "ctypes table and header disagree"


def test_set_default_options():
    # c_api.rs:14-23 + types.rs:227-245: all -1, then the defaults; index 0 stays -1
    o = np.zeros(40, np.int32)
    assert _lib.lib().RMETIS_SetDefaultOptions(o.ctypes.data_as(_lib.i32p)) == 1
    exp = np.full(40, -1, np.int32)
    exp[[2, 3, 4, 1, 9, 7, 6, 8, 10, 11, 16, 17, 5]] = [1, 0, 0, 0, 0, 1, 10, 0, 0, 0, 30, 0, 0]
    assert np.array_equal(o, exp)
    assert _lib.lib().RMETIS_SetDefaultOptions(None) == -2
    assert np.array_equal(R.Options().to_raw(), exp)


def _parse(arr):
    out = (C.c_int64 * 13)()
    a = None if arr is None else np.ascontiguousarray(arr, np.int32)
    assert _lib.lib().rmetis_hip_parse_options(None if a is None else a.ctypes.data_as(_lib.i32p), out) == 1
    return list(out)


DEFAULTS = [1, 0, 0, 0, 0, 1, 10, 0, 0, 0, 30, 0, 0]


def test_option_parsing_quirks():
    assert _parse(None) == DEFAULTS
    a = np.zeros(40, np.int32)
    _lib.lib().RMETIS_SetDefaultOptions(a.ctypes.data_as(_lib.i32p))
    a[4] = 1                     # edited but options[0] == -1: still all defaults
    assert _parse(a) == DEFAULTS
    a[0] = 0
    p = _parse(a)
    assert p[2] == 1             # RTYPE 1 -> Greedy
    b = np.zeros(40, np.int32)   # all zero: CTYPE 0 -> Rm, RTYPE 0 -> Fm, counts keep defaults
    assert _parse(b) == [0, 0, 0, 0, 0, 1, 10, 0, 0, 0, 30, 0, 0]
    b[3] = 1; b[8] = -5; b[16] = 7; b[7] = 3; b[6] = 2; b[9] = 2
    p = _parse(b)
    assert p[1] == 1 and p[7] == 0 and p[10] == 7 and p[5] == 3 and p[6] == 2 and p[4] == 1
    b[3] = 4; b[8] = 99
    p = _parse(b)
    assert p[1] == 4 and p[7] == 99


def test_validate_csr_reference_cases():
    # graph/validate.rs:151-188 through Graph.__init__
    x, a = cycle4()
    R.Graph(4, 1, x, a)
    for bad in ([0, 2, 4], [0, 3, 5, 7, 9]):
        with pytest.raises(R.InputError):
            R.Graph(4, 1, bad, a)
    with pytest.raises(R.InputError):   # self loop
        R.Graph(4, 1, [0, 3, 5, 7, 9], [0, 1, 3, 0, 2, 1, 3, 0, 2])
    with pytest.raises(R.InputError):   # asymmetric degree counts
        R.Graph(4, 1, [0, 1, 1, 1, 1], [1])
    with pytest.raises(R.InputError):   # vwgt length
        R.Graph(4, 1, x, a, vwgt=[1, 1, 1])
    with pytest.raises(R.InputError):   # non-positive vertex weight
        R.Graph(4, 1, x, a, vwgt=[1, 0, 1, 1])
    with pytest.raises(R.InputError):   # adjwgt length
        R.Graph(4, 1, x, a, adjwgt=[1, 1])
    with pytest.raises(R.InputError):
        R.Graph(4, 0, x, a)


def _kway_raw(n, ncon, xadj, adjncy, nparts, part=True):
    L = _lib.lib()
    nv, nc, npt = C.c_int32(n), C.c_int32(ncon), C.c_int32(nparts)
    xa = None if xadj is None else np.ascontiguousarray(xadj, np.int32)
    aa = None if adjncy is None else np.ascontiguousarray(adjncy, np.int32)
    out = np.zeros(max(n, 1), np.int32) if part else None
    obj = C.c_int32(-99)
    rc = L.RMETIS_PartGraphKway(C.byref(nv), C.byref(nc), None if xa is None else xa.ctypes.data_as(_lib.i32p),
                                None if aa is None else aa.ctypes.data_as(_lib.i32p), None, None, None,
                                C.byref(npt), None, None, None, C.byref(obj),
                                None if out is None else out.ctypes.data_as(_lib.i32p))
    return rc, obj.value, out


def test_c_abi_input_errors():
    # c_api.rs:157-167 and kway.rs:31-38: errors come back before any GPU work
    x, a = cycle4()
    assert _kway_raw(4, 1, None, a, 2)[0] == -2
    assert _kway_raw(4, 1, x, None, 2)[0] == -2
    assert _kway_raw(4, 1, x, a, 2, part=False)[0] == -2
    assert _kway_raw(0, 1, x, a, 2)[0] == -2
    assert _kway_raw(4, 0, x, a, 2)[0] == -2
    assert _kway_raw(4, 1, x, a, 1)[0] == -2
    assert _kway_raw(4, 1, x, a, 10)[0] == -2
    rc, obj, part = _kway_raw(4, 1, [0, 1, 1, 1, 1], [1], 2)
    assert rc == -2 and obj == -99 and not part.any()      # outputs untouched on error


def test_compute_without_gpu_fails_loudly():
    if R.device_count() > 0:
        pytest.skip("a GPU is present")
    x, a = cycle4()
    g = R.Graph(4, 1, x, a)
    with pytest.raises(R.InternalError, match="no HIP device"):
        g.edge_cut([0, 0, 1, 1])
    with pytest.raises(R.InternalError):
        R.part_graph_kway(g, 2, options=R.Options(rtype=R.RType.Greedy))
    rc, obj, part = _kway_raw(4, 1, x, a, 2)
    assert rc == -4 and obj == -99


def test_generators_shape():
    g = synth.grid(256, 256)
    assert g.n == 65536 and g.m == 261120 and g.vwgt is None and g.adjwgt is None
    R.Graph(g.n, 1, g.xadj, g.adjncy)
    r = synth.rgg(1 << 12, 3)
    R.Graph(r.n, 1, r.xadj, r.adjncy, r.vwgt, r.adjwgt)
    assert 1 <= r.adjwgt.min() and r.adjwgt.max() <= 10 and (r.vwgt == 1).all()
    src = np.repeat(np.arange(r.n), np.diff(r.xadj))
    o1 = np.lexsort((r.adjncy, src))
    o2 = np.lexsort((src, r.adjncy))
    assert np.array_equal(r.adjwgt[o1], r.adjwgt[o2])     # symmetric weights
    m = synth.rmat(12, 16, 1)
    R.Graph(m.n, 1, m.xadj, m.adjncy)
    b = synth.uneven_batch(6, 1, nmin_log2=8, nmax_log2=10)
    for gg, k in b:
        R.Graph(gg.n, 1, gg.xadj, gg.adjncy, gg.vwgt, gg.adjwgt)
        assert k <= gg.n


def test_pack_batch_layout():
    from graphs import path
    g1 = R.Graph(4, 1, *cycle4(), vwgt=np.ones(4, np.int32))
    x, a = path(5)
    g2 = R.Graph(5, 1, x, a, vwgt=np.arange(1, 6), adjwgt=np.full(len(a), 3))
    gvtx, xadj, adjncy, vwgt, adjwgt, ncon = R.pack_batch([g1, g2])
    assert list(gvtx) == [0, 4, 9] and ncon == 1
    assert xadj[-1] == len(g1.adjncy) + len(g2.adjncy)
    assert np.array_equal(xadj[:5], g1.xadj)
    assert np.array_equal(xadj[4:] - xadj[4], g2.xadj)
    assert np.array_equal(adjncy[:8], g1.adjncy) and np.array_equal(adjncy[8:] - 4, g2.adjncy)
    assert np.array_equal(vwgt, [1, 1, 1, 1, 1, 2, 3, 4, 5])
    assert np.array_equal(adjwgt, [1] * 8 + [3] * len(a))
    # the packed union is itself a valid CSR with no edge between graphs
    R.Graph(9, 1, xadj, adjncy, vwgt, adjwgt)
    # vertex weights: all or none (absent weights are not ones for a batch)
    assert R.pack_batch([R.Graph(4, 1, *cycle4()), R.Graph(5, 1, x, a)])[3] is None
    with pytest.raises(R.InputError):
        R.pack_batch([R.Graph(4, 1, *cycle4()), g2])


def _batch_raw(gvtx, xadj, adjncy, vwgt, nparts):
    arr = [None if v is None else np.ascontiguousarray(v, np.int32) for v in (gvtx, xadj, adjncy, vwgt, nparts)]
    ptr = [None if v is None else v.ctypes.data_as(_lib.i32p) for v in arr]
    G = len(nparts)
    obj = np.full(G, -99, np.int32)
    part = np.zeros(int(gvtx[-1]), np.int32)
    rc = _lib.lib().rmetis_hip_kway_batch(G, ptr[0], 1, ptr[1], ptr[2], ptr[3], None, ptr[4], None, None, None,
                                          obj.ctypes.data_as(_lib.i32p), part.ctypes.data_as(_lib.i32p))
    return rc, obj


def test_batch_abi_input_errors():
    # per-graph graph/validate.rs:16-142 and kway.rs:31-38 checks, before any GPU work
    x, a = cycle4()
    xx = np.concatenate([x, x[1:] + x[-1]])
    aa = np.concatenate([a, a + 4])
    ones = np.ones(8, np.int32)
    assert _batch_raw([0, 4, 8], xx, aa, None, [2, 2])[0] in (1, -4)       # valid without vwgt (-4: no GPU)
    assert _batch_raw([0, 4, 8], xx, aa, ones, [2, 5])[0] == -2            # nparts > nvtxs
    assert _batch_raw([0, 4, 8], xx, aa, ones, [1, 2])[0] == -2            # nparts < 2
    assert _batch_raw([0, 4, 4, 8], np.concatenate([x, x[1:] + x[-1]]), aa, ones, [2, 2, 2])[0] == -2  # empty graph
    cross = aa.copy()
    cross[0] = 5                                                          # an edge into the other graph
    assert _batch_raw([0, 4, 8], xx, cross, ones, [2, 2])[0] == -2
    rc, obj = _batch_raw([0, 4, 8], xx, aa, np.zeros(8, np.int32), [2, 2])
    assert rc == -2 and (obj == -99).all()                                # outputs untouched on error


def test_server_without_gpu_fails_loudly():
    if R.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(R.InternalError, match="no HIP device"):
        R.Server(slots=2)
    assert _lib.lib().rmetis_hip_server_create(0, 0, 0) is None      # slots < 1
    assert _lib.lib().rmetis_hip_server_create(9, 0, 0) is None      # slots > 8


def test_product_grow_input_errors():
    # nparts > nvtxs is an input error (the reference's seed loop never ends),
    # refused before any device work; the growing itself is compared with the
    # oracle on the GPU (test_gpu_grow.py)
    L = _lib.lib()
    x, a = cycle4_arrays()
    part = np.zeros(4, np.int32)
    tpw = np.ones(5)
    assert L.rmetis_hip_grow_partition(4, 1, x.ctypes.data_as(_lib.i32p), a.ctypes.data_as(_lib.i32p), None, 5,
                                       tpw.ctypes.data_as(_lib.f64p), 1, part.ctypes.data_as(_lib.i32p)) == -2


def cycle4_arrays():
    return cycle4()


def test_level_thresholds():
    # floor(target * ub) with target = (tpwgts as f64 or 1/k) * total
    # (graph/mod.rs:392-410) and ub = ubvec as f64 or 1 + ufactor/1000
    # (kway.rs:44-49): w <= fmax <=> w as f64 <= target*ub (greedy.rs:87,144)
    L = _lib.lib()
    cases = [
        (8, 1, None, [65536], None, 30),
        (64, 1, None, [1048576], None, 30),
        (3, 2, None, [1000, 999], [1.05, 1.05], 30),
        (4, 1, [0.1, 0.2, 0.3, 0.4], [12345], None, 1),
        (5, 1, None, [1000], [1.0], 30),          # exact integer targets
        (2, 1, [0.3, 0.7], [10], None, 0),        # 0.7f * 10 is just below 7
        (7, 2, None, [2**30 - 1, 3], [1.001, 1.5], 30),
    ]
    for k, ncon, tp, total, ubv, uf in cases:
        tpa = None if tp is None else np.array(tp, np.float32)
        tot = np.array(total, np.int32)
        uba = None if ubv is None else np.array(ubv, np.float32)
        fm = np.zeros(k * ncon, np.int64)
        om = np.zeros(k * ncon, np.int64)
        rc = L.rmetis_hip_level_thresholds(k, ncon, tpa.ctypes.data_as(_lib.f32p) if tpa is not None else None,
                                           tot.ctypes.data_as(_lib.i32p),
                                           uba.ctypes.data_as(_lib.f32p) if uba is not None else None, uf,
                                           fm.ctypes.data_as(C.POINTER(C.c_int64)),
                                           om.ctypes.data_as(C.POINTER(C.c_int64)))
        assert rc == 1
        tgt = [((float(tpa[p * ncon + c]) if tpa is not None else 1.0 / k) * float(tot[c]))
               for p in range(k) for c in range(ncon)]
        ub = [float(uba[c]) if uba is not None else 1.0 + uf / 1000.0 for c in range(ncon)]
        for p in range(k):
            for c in range(ncon):
                t = tgt[p * ncon + c] * ub[c]
                f = int(math.floor(t))
                assert fm[p * ncon + c] == f and om[p * ncon + c] == f
                # the defining property on the integers either side
                assert float(f) <= t < float(f + 1)
