"""rmetis_amd -- MI355X (gfx950) multilevel k-way partitioning, a drop-in for
javagg/rmetis's ``part_graph_kway`` / ``Graph`` / ``Options`` API (src/lib.rs:48-56,
src/graph/mod.rs:13-50, src/types.rs:12-245) and ``RMETIS_PartGraphKway`` C ABI
(src/ffi/c_api.rs:55-81), with its siblings ``part_graph_recursive`` /
``node_nd`` (``RMETIS_PartGraphRecursive`` / ``RMETIS_NodeND``, c_api.rs:25-134)
and METIS ``.graph`` files (``read_graph`` / ``write_graph``).

The host-side Python mirror calls the C ABI of ``librmetis_hip.so``; every
computation runs in the HIP kernels on the current GPU.  Names, argument
meanings and errors follow the reference: invalid input raises
``InputError`` (METIS_ERROR_INPUT), device memory exhaustion ``MemoryError_``
(METIS_ERROR_MEMORY), everything else ``InternalError`` (METIS_ERROR).
"""
import ctypes as C
import enum
import os
from dataclasses import dataclass, field

import numpy as np

from . import _lib

METIS_OK = 1
METIS_ERROR_INPUT = -2
METIS_ERROR_MEMORY = -3
METIS_ERROR = -4
METIS_NOPTIONS = 40

# option-array indices, src/types.rs:173-186
METIS_OPTION_PTYPE = 0
METIS_OPTION_OBJTYPE = 1
METIS_OPTION_CTYPE = 2
METIS_OPTION_IPTYPE = 3
METIS_OPTION_RTYPE = 4
METIS_OPTION_DBGLVL = 5
METIS_OPTION_NITER = 6
METIS_OPTION_NCUTS = 7
METIS_OPTION_SEED = 8
METIS_OPTION_NO2HOP = 9
METIS_OPTION_MINCONN = 10
METIS_OPTION_CONTIG = 11
METIS_OPTION_UFACTOR = 16
METIS_OPTION_NUMBERING = 17


class MetisError(Exception):
    """src/types.rs:145-161."""
    code = METIS_ERROR


class InputError(MetisError):
    code = METIS_ERROR_INPUT


class MemoryError_(MetisError):
    code = METIS_ERROR_MEMORY


class InternalError(MetisError):
    code = METIS_ERROR


def _raise(code):
    if code == METIS_OK:
        return
    msg = _lib.last_error()
    if code == METIS_ERROR_INPUT:
        raise InputError(msg or "input error")
    if code == METIS_ERROR_MEMORY:
        raise MemoryError_(msg or "out of memory")
    raise InternalError(msg or "internal error")


class CType(enum.IntEnum):
    Rm = 0
    Shem = 1


class IpType(enum.IntEnum):
    Grow = 0
    Random = 1
    MetisRb = 4


class RType(enum.IntEnum):
    Fm = 0
    Greedy = 1


class ObjType(enum.IntEnum):
    Cut = 0
    Vol = 1


@dataclass
class Options:
    """src/types.rs:52-95 (defaults), presets 99-116, raw array 190-245."""
    ctype: CType = CType.Shem
    iptype: IpType = IpType.Grow
    rtype: RType = RType.Fm
    objtype: ObjType = ObjType.Cut
    no2hop: bool = False
    ncuts: int = 1
    niter: int = 10
    seed: int = 0
    minconn: bool = False
    contig: bool = False
    ufactor: int = 30
    numbering: int = 0
    dbglvl: int = 0

    @classmethod
    def for_recursive(cls, **kw):
        return cls(**{**dict(iptype=IpType.MetisRb, rtype=RType.Fm, ufactor=1), **kw})

    @classmethod
    def for_kway(cls, **kw):
        return cls(**{**dict(iptype=IpType.Grow, rtype=RType.Fm, ufactor=30), **kw})

    def to_raw(self):
        """types.rs:227-245 (index 0 stays -1)."""
        a = np.full(METIS_NOPTIONS, -1, np.int32)
        a[METIS_OPTION_CTYPE] = int(self.ctype)
        a[METIS_OPTION_IPTYPE] = int(self.iptype)
        a[METIS_OPTION_RTYPE] = int(self.rtype)
        a[METIS_OPTION_OBJTYPE] = int(self.objtype)
        a[METIS_OPTION_NO2HOP] = int(self.no2hop)
        a[METIS_OPTION_NCUTS] = self.ncuts
        a[METIS_OPTION_NITER] = self.niter
        a[METIS_OPTION_SEED] = np.int64(self.seed).astype(np.int32)
        a[METIS_OPTION_MINCONN] = int(self.minconn)
        a[METIS_OPTION_CONTIG] = int(self.contig)
        a[METIS_OPTION_UFACTOR] = self.ufactor
        a[METIS_OPTION_NUMBERING] = self.numbering
        a[METIS_OPTION_DBGLVL] = self.dbglvl
        return a

    def to_c_array(self):
        """Array handed to the C ABI: to_raw() with index 0 set so that the
        C side parses it (c_api.rs:217 treats options[0] == -1 as defaults)."""
        a = self.to_raw()
        a[METIS_OPTION_PTYPE] = 0
        return a


@dataclass
class PartitionResult:
    """src/types.rs:125-130."""
    part: np.ndarray
    objval: int


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def _p(a, t=_lib.i32p):
    return None if a is None else a.ctypes.data_as(t)


class Graph:
    """CSR graph, src/graph/mod.rs:13-50.  Arrays are int32; None weights mean
    unit weights.  Construction validates with graph/validate.rs:16-142 and
    raises InputError like Graph::new."""

    def __init__(self, nvtxs, ncon, xadj, adjncy, vwgt=None, adjwgt=None, vsize=None):
        self.nvtxs = int(nvtxs)
        self.ncon = int(ncon)
        self.xadj = _i32(xadj)
        self.adjncy = _i32(adjncy)
        self.vwgt = _i32(vwgt)
        self.adjwgt = _i32(adjwgt)
        self.vsize = _i32(vsize)
        rc = _lib.lib().rmetis_hip_validate_csr(
            self.nvtxs, self.ncon, _p(self.xadj), len(self.xadj), _p(self.adjncy), len(self.adjncy),
            _p(self.vwgt), 0 if self.vwgt is None else len(self.vwgt),
            _p(self.adjwgt), 0 if self.adjwgt is None else len(self.adjwgt))
        _raise(rc)
        self.nedges = len(self.adjncy) // 2

    @classmethod
    def new_unweighted(cls, nvtxs, xadj, adjncy):
        return cls(nvtxs, 1, xadj, adjncy)

    def __repr__(self):
        return f"Graph(nvtxs={self.nvtxs}, nedges={self.nedges}, ncon={self.ncon})"

    def neighbors(self, v):
        return self.adjncy[self.xadj[v]:self.xadj[v + 1]]

    def degree(self, v):
        return int(self.xadj[v + 1] - self.xadj[v])

    def total_weight(self):
        """graph/mod.rs:152-160 (i32 sums)."""
        if self.vwgt is None:
            return np.full(self.ncon, self.nvtxs, np.int32)
        s = self.vwgt.reshape(self.nvtxs, self.ncon).astype(np.int64).sum(axis=0)
        return s.astype(np.uint32).astype(np.int32)

    def compute_target_weights(self, nparts, tpwgts=None):
        """graph/mod.rs:392-410 -> float64 [nparts][ncon]."""
        total = self.total_weight().astype(np.float64)
        if tpwgts is None:
            frac = np.full((nparts, self.ncon), 1.0 / float(nparts))
        else:
            frac = np.asarray(tpwgts, np.float32).astype(np.float64).reshape(nparts, self.ncon)
        return frac * total[None, :]

    # ---- reductions and gain tables on the GPU ----
    def edge_cut(self, part):
        """graph/mod.rs:123-137."""
        part = _i32(part)
        out = C.c_int32()
        _raise(_lib.lib().rmetis_hip_edge_cut(self.nvtxs, _p(self.xadj), _p(self.adjncy), _p(self.adjwgt),
                                              _p(part), C.byref(out)))
        return out.value

    def partition_weights(self, part, nparts):
        """graph/mod.rs:140-149 -> int32 [nparts][ncon]."""
        part = _i32(part)
        pw = np.zeros(nparts * self.ncon, np.int32)
        _raise(_lib.lib().rmetis_hip_partition_weights(self.nvtxs, self.ncon, _p(self.vwgt), _p(part), nparts,
                                                       _p(pw)))
        return pw.reshape(nparts, self.ncon)

    def gain_table(self, part, nparts):
        """Boundary records and gain lists (greedy.rs:197-218); see
        rmetis_hip_gain_table in include/rmetis.h."""
        part = _i32(part)
        m = len(self.adjncy)
        rec = np.zeros(4 * self.nvtxs, np.int32)
        g0 = np.zeros(self.nvtxs, np.int32)
        cp = np.zeros(max(m, 1), np.int32)
        cg = np.zeros(max(m, 1), np.int32)
        _raise(_lib.lib().rmetis_hip_gain_table(self.nvtxs, _p(self.xadj), _p(self.adjncy), _p(self.adjwgt),
                                                _p(part), nparts, _p(rec), _p(g0), _p(cp), _p(cg)))
        return rec.reshape(self.nvtxs, 4), g0, cp, cg

    def boundary_vertices(self, part, nparts=None):
        """graph/mod.rs:164-171 (ascending)."""
        part = _i32(part)
        k = int(part.max()) + 1 if nparts is None else nparts
        rec, _, _, _ = self.gain_table(part, k)
        return np.nonzero(rec[:, 0] >= 0)[0].astype(np.int64)

    def build_coarse_graph(self, matching):
        """graph/mod.rs:284-384 -> (Graph, cmap)."""
        matching = _i32(matching)
        n, m = self.nvtxs, len(self.adjncy)
        cx = np.zeros(n + 1, np.int32)
        ca = np.zeros(max(m, 1), np.int32)
        cw = np.zeros(max(m, 1), np.int32)
        cvw = np.zeros(max(n * self.ncon, 1), np.int32)
        cvs = np.zeros(max(n, 1), np.int32)
        cmap = np.zeros(n, np.int32)
        cn, cm, fl = C.c_int32(), C.c_int32(), C.c_int32()
        _raise(_lib.lib().rmetis_hip_build_coarse_graph(
            n, self.ncon, _p(self.xadj), _p(self.adjncy), _p(self.vwgt), _p(self.adjwgt), _p(self.vsize),
            _p(matching), _p(cx), _p(ca), _p(cw), _p(cvw), _p(cvs), _p(cmap), C.byref(cn), C.byref(cm),
            C.byref(fl)))
        cn, cm, f = cn.value, cm.value, fl.value
        g = Graph(cn, self.ncon, cx[:cn + 1], ca[:cm],
                  cvw[:cn * self.ncon] if f & 1 else None,
                  cw[:cm] if f & 4 else None,
                  cvs[:cn] if f & 2 else None)
        return g, cmap


def greedy_refine(graph, part, nparts, tpwgts, ubvec, niter):
    """refine/greedy.rs:12-22 -> (new part, cut).  tpwgts: f64 targets
    [nparts*ncon]; ubvec: f64 [ncon]."""
    part = np.array(part, dtype=np.int32)
    tp = np.ascontiguousarray(tpwgts, np.float64)
    ub = np.ascontiguousarray(ubvec, np.float64)
    cut = C.c_int32()
    _raise(_lib.lib().rmetis_hip_greedy_refine(
        graph.nvtxs, graph.ncon, _p(graph.xadj), _p(graph.adjncy), _p(graph.vwgt), _p(graph.adjwgt), _p(part),
        nparts, _p(tp, _lib.f64p), _p(ub, _lib.f64p), niter, C.byref(cut)))
    return part, cut.value


def fm_kway_refine(graph, part, nparts, tpwgts, ubvec, niter):
    """refine/fm_kway.rs:12-30 -> (new part, cut).  tpwgts: f64 targets
    [nparts*ncon]; ubvec: f64 [ncon]."""
    part = np.array(part, dtype=np.int32)
    tp = np.ascontiguousarray(tpwgts, np.float64)
    ub = np.ascontiguousarray(ubvec, np.float64)
    cut = C.c_int32()
    _raise(_lib.lib().rmetis_hip_fm_refine(
        graph.nvtxs, graph.ncon, _p(graph.xadj), _p(graph.adjncy), _p(graph.vwgt), _p(graph.adjwgt), _p(part),
        nparts, _p(tp, _lib.f64p), _p(ub, _lib.f64p), niter, C.byref(cut)))
    return part, cut.value


def fm2_refine(graph, part, tpwgts, ubvec, niter):
    """refine/fm.rs:12-29 (2-way FM) -> (new part, cut).  part: 0/1 per
    vertex; tpwgts: f64 targets [2*ncon]; ubvec: f64 [ncon]."""
    part = np.array(part, dtype=np.int32)
    tp = np.ascontiguousarray(tpwgts, np.float64)
    ub = np.ascontiguousarray(ubvec, np.float64)
    cut = C.c_int32()
    _raise(_lib.lib().rmetis_hip_fm2_refine(
        graph.nvtxs, graph.ncon, _p(graph.xadj), _p(graph.adjncy), _p(graph.vwgt), _p(graph.adjwgt), _p(part),
        _p(tp, _lib.f64p), _p(ub, _lib.f64p), niter, C.byref(cut)))
    return part, cut.value


def greedy_pass(graph, part, nparts, tpwgts, ubvec, mode, pw):
    """One synchronous greedy pass (mode 0 repair, 1 cut) -> (part, pw, moves)."""
    part = np.array(part, dtype=np.int32)
    pw = np.array(pw, dtype=np.int32).reshape(-1)
    tp = np.ascontiguousarray(tpwgts, np.float64)
    ub = np.ascontiguousarray(ubvec, np.float64)
    moves = np.zeros(2 * graph.nvtxs + 2, np.int32)
    nm = C.c_int32()
    _raise(_lib.lib().rmetis_hip_greedy_pass(
        graph.nvtxs, graph.ncon, _p(graph.xadj), _p(graph.adjncy), _p(graph.vwgt), _p(graph.adjwgt), _p(part),
        nparts, _p(tp, _lib.f64p), _p(ub, _lib.f64p), mode, _p(pw), _p(moves), C.byref(nm)))
    return part, pw.reshape(nparts, graph.ncon), moves[:2 * nm.value].reshape(-1, 2)


def gpu_matching(graph, seed, ctype=CType.Shem, cap=None):
    """The seeded GPU matching that replaces shem/rm matching; cap[ncon]
    (int64) limits the matched pairs' summed vertex weights."""
    m = np.zeros(graph.nvtxs, np.int32)
    if cap is None:
        _raise(_lib.lib().rmetis_hip_match(graph.nvtxs, _p(graph.xadj), _p(graph.adjncy), _p(graph.adjwgt),
                                           int(ctype), C.c_uint64(seed), _p(m)))
    else:
        cp = np.ascontiguousarray(cap, np.int64)
        if cp.shape != (graph.ncon,) or graph.vwgt is None:
            raise InputError("cap needs one entry per constraint and a weighted graph")
        _raise(_lib.lib().rmetis_hip_match_capped(graph.nvtxs, graph.ncon, _p(graph.xadj), _p(graph.adjncy),
                                                  _p(graph.vwgt), _p(graph.adjwgt), int(ctype), C.c_uint64(seed),
                                                  _p(cp, _lib.i64p), _p(m)))
    return m


def part_graph_kway(graph, nparts, tpwgts=None, ubvec=None, options=None):
    """src/lib.rs:48-56 through RMETIS_PartGraphKway (c_api.rs:55-81)."""
    options = options if options is not None else Options()
    n = C.c_int32(graph.nvtxs)
    nc = C.c_int32(graph.ncon)
    np_ = C.c_int32(int(nparts))
    tp = None if tpwgts is None else np.ascontiguousarray(tpwgts, np.float32)
    ub = None if ubvec is None else np.ascontiguousarray(ubvec, np.float32)
    opts = options.to_c_array()
    part = np.zeros(graph.nvtxs, np.int32)
    obj = C.c_int32()
    rc = _lib.lib().RMETIS_PartGraphKway(C.byref(n), C.byref(nc), _p(graph.xadj), _p(graph.adjncy),
                                         _p(graph.vwgt), _p(graph.vsize), _p(graph.adjwgt), C.byref(np_),
                                         _p(tp, _lib.f32p), _p(ub, _lib.f32p), _p(opts), C.byref(obj), _p(part))
    _raise(rc)
    return PartitionResult(part=part, objval=obj.value)


def part_graph_recursive(graph, nparts, tpwgts=None, ubvec=None, options=None):
    """src/lib.rs:44 (partition/recursive.rs:11-57) through
    RMETIS_PartGraphRecursive (c_api.rs:25-48)."""
    options = options if options is not None else Options()
    n = C.c_int32(graph.nvtxs)
    nc = C.c_int32(graph.ncon)
    np_ = C.c_int32(int(nparts))
    tp = None if tpwgts is None else np.ascontiguousarray(tpwgts, np.float32)
    ub = None if ubvec is None else np.ascontiguousarray(ubvec, np.float32)
    opts = options.to_c_array()
    part = np.zeros(graph.nvtxs, np.int32)
    obj = C.c_int32()
    rc = _lib.lib().RMETIS_PartGraphRecursive(C.byref(n), C.byref(nc), _p(graph.xadj), _p(graph.adjncy),
                                              _p(graph.vwgt), _p(graph.vsize), _p(graph.adjwgt), C.byref(np_),
                                              _p(tp, _lib.f32p), _p(ub, _lib.f32p), _p(opts), C.byref(obj),
                                              _p(part))
    _raise(rc)
    return PartitionResult(part=part, objval=obj.value)


@dataclass
class NDResult:
    """src/types.rs:132-139: perm[v] = new position of vertex v, iperm its
    inverse."""
    perm: np.ndarray
    iperm: np.ndarray


def node_nd(graph, options=None):
    """src/lib.rs:45 (partition/nd.rs:16-44) through RMETIS_NodeND
    (c_api.rs:83-134).  Uses the graph's CSR and vertex weights (ncon 1)."""
    options = options if options is not None else Options()
    n = C.c_int32(graph.nvtxs)
    opts = options.to_c_array()
    perm = np.zeros(graph.nvtxs, np.int32)
    iperm = np.zeros(graph.nvtxs, np.int32)
    vw = graph.vwgt if graph.ncon == 1 else None
    rc = _lib.lib().RMETIS_NodeND(C.byref(n), _p(graph.xadj), _p(graph.adjncy), _p(vw), _p(opts), _p(perm),
                                  _p(iperm))
    _raise(rc)
    return NDResult(perm=perm, iperm=iperm)


def read_graph(path):
    """A METIS .graph file (README.md:208-218) -> Graph (0-based CSR, the
    file's vertex weights, sizes and edge weights when present)."""
    n, nc = C.c_int32(), C.c_int32()
    px, pa, pw, pz, pe = (_lib.i32p() for _ in range(5))
    L = _lib.lib()
    _raise(L.rmetis_read_graph(os.fsencode(path), C.byref(n), C.byref(nc), C.byref(px), C.byref(pa), C.byref(pw),
                               C.byref(pz), C.byref(pe)))
    try:
        def arr(ptr, count):
            if not ptr:
                return None
            return np.ctypeslib.as_array(ptr, shape=(count,)).copy() if count > 0 else np.zeros(0, np.int32)
        xadj = arr(px, n.value + 1)
        m = int(xadj[-1])
        adjncy = arr(pa, m)
        vwgt = arr(pw, n.value * nc.value)
        vsize = arr(pz, n.value)
        adjwgt = arr(pe, m)
    finally:
        for ptr in (px, pa, pw, pz, pe):
            if ptr:
                L.rmetis_free(ptr)
    return Graph(n.value, nc.value, xadj, adjncy, vwgt=vwgt, adjwgt=adjwgt, vsize=vsize)


def write_graph(path, graph):
    """Writes graph as a METIS .graph file (1-based, weights when present)."""
    _raise(_lib.lib().rmetis_write_graph(os.fsencode(path), graph.nvtxs, graph.ncon, _p(graph.xadj),
                                         _p(graph.adjncy), _p(graph.vwgt), _p(graph.vsize), _p(graph.adjwgt)))


def pack_batch(graphs):
    """Lays graphs out one after another in one CSR with global vertex ids.
    Returns (gvtx, xadj, adjncy, vwgt, adjwgt, ncon).  Vertex weights must be
    given for all graphs or for none (absent weights are not the same as
    ones: coarse vertices of an unweighted graph keep weight 1); missing edge
    weights become ones when any graph has edge weights."""
    if not graphs:
        raise InputError("a batch needs at least one graph")
    ncon = graphs[0].ncon
    if any(g.ncon != ncon for g in graphs):
        raise InputError("every graph of a batch needs the same ncon")
    nv = [len(g.xadj) - 1 for g in graphs]
    gvtx = np.zeros(len(graphs) + 1, np.int64)
    np.cumsum(nv, out=gvtx[1:])
    eoff = np.zeros(len(graphs) + 1, np.int64)
    np.cumsum([len(g.adjncy) for g in graphs], out=eoff[1:])
    if gvtx[-1] >= 2**31 or eoff[-1] >= 2**31:
        raise InputError("batch too large for int32 CSR")
    xadj = np.concatenate([g.xadj[:-1].astype(np.int64) + eoff[i] for i, g in enumerate(graphs)]
                          + [np.array([eoff[-1]])]).astype(np.int32)
    adjncy = np.concatenate([g.adjncy.astype(np.int64) + gvtx[i] for i, g in enumerate(graphs)]).astype(np.int32)
    has_vw = [g.vwgt is not None for g in graphs]
    if any(has_vw) and not all(has_vw):
        raise InputError("vertex weights must be given for every graph of a batch or for none")
    vwgt = np.concatenate([g.vwgt for g in graphs]).astype(np.int32) if all(has_vw) else None
    adjwgt = None
    if any(g.adjwgt is not None for g in graphs):
        adjwgt = np.concatenate([g.adjwgt if g.adjwgt is not None else np.ones(len(g.adjncy), np.int32)
                                 for g in graphs]).astype(np.int32)
    return gvtx.astype(np.int32), xadj, adjncy, vwgt, adjwgt, ncon


def part_graph_kway_batch(graphs, nparts, ubvec=None, options=None, seeds=None):
    """Partitions many graphs in one call (rmetis_hip_kway_batch): graph i with
    nparts[i] parts gets the PartitionResult part_graph_kway(graphs[i],
    nparts[i], None, ubvec, options) gives it when options' seed is seeds[i]
    (the options' seed for every graph when seeds is None)."""
    gvtx, xadj, adjncy, vwgt, adjwgt, ncon = pack_batch(graphs)
    G = len(graphs)
    kk = np.ascontiguousarray(nparts, np.int32)
    if kk.shape != (G,):
        raise InputError("nparts needs one entry per graph")
    options = options if options is not None else Options()
    opts = options.to_c_array()
    ub = None if ubvec is None else np.ascontiguousarray(ubvec, np.float32)
    sd = None if seeds is None else np.ascontiguousarray(seeds, np.uint64)
    obj = np.zeros(G, np.int32)
    part = np.zeros(int(gvtx[-1]), np.int32)
    rc = _lib.lib().rmetis_hip_kway_batch(G, _p(gvtx), ncon, _p(xadj), _p(adjncy), _p(vwgt), _p(adjwgt), _p(kk),
                                          _p(ub, _lib.f32p), _p(opts), _p(sd, C.POINTER(C.c_uint64)), _p(obj),
                                          _p(part))
    _raise(rc)
    return [PartitionResult(part=part[gvtx[i]:gvtx[i + 1]].copy(), objval=int(obj[i])) for i in range(G)]


@dataclass
class ServedResult:
    """A served request's result: part and objval as PartitionResult, the
    queued / running / total milliseconds, and the graphs of its batch."""
    part: np.ndarray
    objval: int
    queued_ms: float
    running_ms: float
    total_ms: float
    batch_graphs: int


class Server:
    """k-way requests with continuous batching (rmetis_hip_server_*): up to
    `slots` batched calls run at once; every request's result equals that of
    part_graph_kway with its seed, whatever it is batched with."""

    def __init__(self, slots=2, memory_budget=0, max_batch_vertices=0):
        self._h = _lib.lib().rmetis_hip_server_create(int(slots), int(memory_budget), int(max_batch_vertices))
        if not self._h:
            raise InternalError(_lib.last_error())
        self._n = {}

    def submit(self, graph, nparts, ubvec=None, options=None, seed=0):
        options = options if options is not None else Options()
        ub = None if ubvec is None else np.ascontiguousarray(ubvec, np.float32)
        rid = _lib.lib().rmetis_hip_server_submit(
            self._h, graph.nvtxs, graph.ncon, _p(graph.xadj), _p(graph.adjncy), _p(graph.vwgt), _p(graph.adjwgt),
            int(nparts), _p(ub, _lib.f32p), _p(options.to_c_array()), C.c_uint64(int(seed)))
        if rid < 0:
            _raise(int(rid))
        self._n[rid] = graph.nvtxs
        return rid

    def submit_device(self, nvtxs, ncon, d_xadj, d_adjncy, d_vwgt, d_adjwgt, nparts, d_part, ubvec=None,
                      options=None, seed=0):
        """A request whose CSR already lives in device memory (_lib.DeviceArray
        or raw pointers); the part is written to d_part, and the arrays must
        stay alive until result() has returned."""
        ptr = lambda a: None if a is None else (a.ptr if hasattr(a, "ptr") else a)
        options = options if options is not None else Options()
        ub = None if ubvec is None else np.ascontiguousarray(ubvec, np.float32)
        rid = _lib.lib().rmetis_hip_server_submit_device(
            self._h, int(nvtxs), int(ncon), ptr(d_xadj), ptr(d_adjncy), ptr(d_vwgt), ptr(d_adjwgt), int(nparts),
            _p(ub, _lib.f32p), _p(options.to_c_array()), C.c_uint64(int(seed)), ptr(d_part))
        if rid < 0:
            _raise(int(rid))
        self._n[rid] = None
        return rid

    def result(self, rid, wait=True, timeout_ms=None):
        """ServedResult (part None for device requests), or None while
        pending (wait=False).  With timeout_ms, waits at most that long and
        then raises InternalError naming every slot's phase."""
        if rid not in self._n:
            raise InputError(f"unknown or already collected request id {rid}")
        part = np.zeros(self._n[rid], np.int32) if self._n[rid] is not None else None
        obj = C.c_int32()
        t = np.zeros(3, np.float64)
        bg = C.c_int32()
        w = 0 if not wait else (max(2, int(timeout_ms)) if timeout_ms else 1)
        rc = _lib.lib().rmetis_hip_server_result(self._h, int(rid), w, _p(part), C.byref(obj),
                                                 _p(t, _lib.f64p), C.byref(bg))
        if rc == 0:
            return None
        del self._n[rid]
        _raise(rc)
        return ServedResult(part, obj.value, float(t[0]), float(t[1]), float(t[2]), bg.value)

    def stats(self):
        out = np.zeros(5, np.int64)
        _raise(_lib.lib().rmetis_hip_server_stats(self._h, out.ctypes.data_as(C.POINTER(C.c_int64))))
        return dict(zip(["batches", "served", "inflight_batches", "max_inflight_batches", "max_batch_graphs"],
                        map(int, out)))

    def close(self):
        if self._h:
            _lib.lib().rmetis_hip_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_count():
    return _lib.lib().rmetis_hip_device_count()
