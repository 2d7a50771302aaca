"""Deterministic synthetic graph generators for the k-way benchmarks and tests.

All generators draw from numpy's PCG64 stream (``np.random.default_rng(seed)``)
and return 0-based CSR arrays with ascending neighbour lists, each undirected
edge stored twice with equal weights, no self-loops and no duplicates -- the
shape graph/validate.rs:16-142 accepts and tests/correctness.rs:36-50 builds.

Workloads:
  * ``grid(rows, cols)``           5-point grid, row-major ids, unit weights.
  * ``rgg(n, seed, r=None)``       random geometric graph in the unit square,
                                   edge iff distance < r, r = sqrt(12/(pi n)) by
                                   default (mean degree ~12), Morton-ordered ids,
                                   int32 vertex weight 1, edge weights U{1..10}.
  * ``rmat(scale, ef, seed)``      R-MAT (a,b,c,d) = (.57,.19,.19,.05),
                                   symmetrised, self-loops and duplicates
                                   removed, ids randomly permuted, unit weights.
  * ``uneven_batch(count, seed)``  the 1,024-graph mixed batch (grids with a
                                   random aspect ratio and RGGs of mean degree
                                   8-16, n log-uniform in 2^10..2^18).
"""
import math

import numpy as np


class CSR:
    """Plain CSR container (int32 arrays; None weights mean unit weights)."""

    def __init__(self, xadj, adjncy, vwgt=None, adjwgt=None, ncon=1, name=""):
        self.xadj = np.ascontiguousarray(xadj, dtype=np.int32)
        self.adjncy = np.ascontiguousarray(adjncy, dtype=np.int32)
        self.vwgt = None if vwgt is None else np.ascontiguousarray(vwgt, dtype=np.int32)
        self.adjwgt = None if adjwgt is None else np.ascontiguousarray(adjwgt, dtype=np.int32)
        self.ncon = ncon
        self.name = name

    @property
    def n(self):
        return len(self.xadj) - 1

    @property
    def m(self):
        return int(self.xadj[-1])

    def __repr__(self):
        return f"CSR({self.name}, n={self.n}, m={self.m}, ncon={self.ncon})"


def _csr_from_undirected(n, u, v, w=None):
    """u < v undirected edge list (no duplicates) -> symmetric CSR, sorted rows."""
    src = np.concatenate([u, v]).astype(np.int64)
    dst = np.concatenate([v, u]).astype(np.int64)
    # one int64 key per entry (row-major): a plain sort when there are no
    # weights to carry, else a stable argsort of the keys
    key = src * np.int64(n) + dst
    ww = None
    if w is None:
        key.sort()
    else:
        order = np.argsort(key, kind="stable")
        key = key[order]
        ww = np.concatenate([w, w])[order].astype(np.int32)
    src = key // np.int64(n)
    dst = key - src * np.int64(n)
    counts = np.bincount(src, minlength=n)
    xadj = np.zeros(n + 1, np.int64)
    np.cumsum(counts, out=xadj[1:])
    if xadj[-1] >= 2**31:
        raise ValueError("graph too large for int32 CSR")
    return xadj.astype(np.int32), dst.astype(np.int32), ww


def grid(rows, cols, name=None):
    """5-point grid, id = r*cols + c, unit weights (vwgt/adjwgt None)."""
    ids = np.arange(rows * cols, dtype=np.int64).reshape(rows, cols)
    hu, hv = ids[:, :-1].ravel(), ids[:, 1:].ravel()
    vu, vv = ids[:-1, :].ravel(), ids[1:, :].ravel()
    xadj, adj, _ = _csr_from_undirected(rows * cols, np.concatenate([hu, vu]),
                                        np.concatenate([hv, vv]))
    return CSR(xadj, adj, name=name or f"grid{rows}x{cols}")


def _morton_key(x, y):
    def spread(a):
        a = a.astype(np.uint64) & np.uint64(0xFFFF)
        a = (a | (a << np.uint64(8))) & np.uint64(0x00FF00FF)
        a = (a | (a << np.uint64(4))) & np.uint64(0x0F0F0F0F)
        a = (a | (a << np.uint64(2))) & np.uint64(0x33333333)
        a = (a | (a << np.uint64(1))) & np.uint64(0x55555555)
        return a
    qx = np.minimum((x * 65536.0).astype(np.int64), 65535)
    qy = np.minimum((y * 65536.0).astype(np.int64), 65535)
    return spread(qx) | (spread(qy) << np.uint64(1))


def _rgg_edges(pts, r):
    """All pairs (i<j) with |p_i - p_j| < r via a uniform cell grid."""
    n = len(pts)
    nc = max(1, int(1.0 / r))
    cx = np.minimum((pts[:, 0] * nc).astype(np.int64), nc - 1)
    cy = np.minimum((pts[:, 1] * nc).astype(np.int64), nc - 1)
    cell = cy * nc + cx
    order = np.argsort(cell, kind="stable")
    cell_s = cell[order]
    start = np.searchsorted(cell_s, np.arange(nc * nc), side="left")
    end = np.searchsorted(cell_s, np.arange(nc * nc), side="right")
    P = pts[order]
    us, vs = [], []
    r2 = r * r
    # forward neighbour cells: (0,0) with i<j, then (1,0), (-1,1), (0,1), (1,1)
    for dx, dy in ((0, 0), (1, 0), (-1, 1), (0, 1), (1, 1)):
        ox, oy = cx[order] + dx, cy[order] + dy
        ok = (ox >= 0) & (ox < nc) & (oy < nc)
        i_idx = np.nonzero(ok)[0]
        oc = oy[i_idx] * nc + ox[i_idx]
        cnt = end[oc] - start[oc]
        tot = int(cnt.sum())
        if tot == 0:
            continue
        ii = np.repeat(i_idx, cnt)
        offs = np.arange(tot) - np.repeat(np.cumsum(cnt) - cnt, cnt)
        jj = np.repeat(start[oc], cnt) + offs
        if dx == 0 and dy == 0:
            keep = jj > ii
            ii, jj = ii[keep], jj[keep]
        d = P[ii] - P[jj]
        keep = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) < r2
        us.append(order[ii[keep]])
        vs.append(order[jj[keep]])
    u = np.concatenate(us) if us else np.zeros(0, np.int64)
    v = np.concatenate(vs) if vs else np.zeros(0, np.int64)
    lo, hi = np.minimum(u, v), np.maximum(u, v)
    return lo, hi


def rgg(n, seed, r=None, wmax=10, vwgt_max=1, ncon=1, name=None, rng=None):
    """Random geometric graph, Morton-ordered ids, adjwgt U{1..wmax},
    vwgt U{1..vwgt_max} per constraint (explicit int32 array)."""
    rng = rng if rng is not None else np.random.default_rng(seed)
    if r is None:
        r = math.sqrt(12.0 / (math.pi * n))
    pts = rng.random((n, 2))
    key = _morton_key(pts[:, 0], pts[:, 1])
    order = np.argsort(key, kind="stable")          # ties by original index
    pts = pts[order]
    u, v = _rgg_edges(pts, r)
    w = rng.integers(1, wmax + 1, size=len(u), dtype=np.int64) if wmax > 1 else np.ones(len(u), np.int64)
    xadj, adj, ww = _csr_from_undirected(n, u, v, w)
    if vwgt_max > 1:
        vw = rng.integers(1, vwgt_max + 1, size=n * ncon, dtype=np.int64).astype(np.int32)
    else:
        vw = np.ones(n * ncon, np.int32)
    return CSR(xadj, adj, vw, ww, ncon=ncon, name=name or f"rgg{n}")


def _rmat_chunk(args):
    """Unique (lo * n + hi) keys of one chunk of R-MAT pairs.  The chunk's
    draws are the ones a single sequential stream would give it: scale
    uniform doubles per pair and level, level by level, after all earlier
    chunks (PCG64 advanced past them)."""
    seed, scale, off, cnt, a, b, c = args
    bg = np.random.PCG64(seed)
    bg.advance(scale * off)
    rng = np.random.Generator(bg)
    src = np.zeros(cnt, np.int64)
    dst = np.zeros(cnt, np.int64)
    ab, abc = a + b, a + b + c
    for lvl in range(scale):
        rr = rng.random(cnt)
        # quadrant 0 (a), 1 (b: right), 2 (c: down), 3 (d: down-right)
        q = (rr >= a).view(np.uint8) + (rr >= ab).view(np.uint8) + (rr >= abc).view(np.uint8)
        dst <<= 1
        dst |= q & 1
        src <<= 1
        src |= q >> 1
    keep = src != dst
    src, dst = src[keep], dst[keep]
    lo, hi = np.minimum(src, dst), np.maximum(src, dst)
    return np.unique(lo * (np.int64(1) << scale) + hi)


def rmat(scale, ef, seed, a=0.57, b=0.19, c=0.19, name=None, chunk=1 << 23, workers=None):
    """R-MAT (scale, edge factor) graph: symmetrised, deduplicated, no
    self-loops, vertex ids randomly permuted, unit weights.  Chunks of pairs
    are drawn in parallel processes from positions of one PCG64 stream, so
    the graph does not depend on the number of workers."""
    import os
    n = 1 << scale
    total = ef * n
    jobs = [(seed, scale, off, min(chunk, total - off), a, b, c) for off in range(0, total, chunk)]
    if workers is None:
        workers = min(16, os.cpu_count() or 1, len(jobs))
    if workers <= 1 or len(jobs) == 1:
        keys = [_rmat_chunk(j) for j in jobs]
    else:
        import multiprocessing as mp
        with mp.get_context("fork").Pool(workers) as pool:
            keys = pool.map(_rmat_chunk, jobs, chunksize=1)
    key = np.unique(np.concatenate(keys))
    del keys
    bg = np.random.PCG64(seed)
    bg.advance(scale * total)
    perm = np.random.Generator(bg).permutation(n)
    u, v = perm[key >> scale], perm[key & (n - 1)]
    del key
    lo, hi = np.minimum(u, v), np.maximum(u, v)
    xadj, adj, _ = _csr_from_undirected(n, lo, hi)
    return CSR(xadj, adj, name=name or f"rmat{scale}")


def _batch_member(args):
    """Graph i of the uneven batch from its drawn (i, n, k, gseed[, ncon,
    vwgt_max])."""
    i, n, k, gseed = args[:4]
    ncon, vmax = (args[4], args[5]) if len(args) > 4 else (1, 4)
    grng = np.random.default_rng(gseed)
    if i % 2 == 0:
        aspect = 2.0 ** grng.uniform(-2, 2)
        rows = max(2, int(round(math.sqrt(n * aspect))))
        cols = max(2, int(round(n / rows)))
        g = grid(rows, cols, name=f"b{i}_grid{rows}x{cols}")
        m2 = g.m // 2
        # symmetric U{1..10} weights: draw per undirected edge, mirror.
        src = np.repeat(np.arange(g.n), np.diff(g.xadj))
        lo = np.minimum(src, g.adjncy).astype(np.int64)
        hi = np.maximum(src, g.adjncy).astype(np.int64)
        ukey, inv = np.unique(lo * g.n + hi, return_inverse=True)
        w = grng.integers(1, 11, size=len(ukey), dtype=np.int64).astype(np.int32)
        g.adjwgt = np.ascontiguousarray(w[inv])
        g.vwgt = grng.integers(1, vmax + 1, size=g.n * ncon, dtype=np.int64).astype(np.int32)
        g.ncon = ncon
        assert len(ukey) == m2
    else:
        deg = grng.uniform(8, 16)
        r = math.sqrt(deg / (math.pi * n))
        g = rgg(n, gseed, r=r, wmax=10, vwgt_max=vmax, ncon=ncon, name=f"b{i}_rgg{n}", rng=grng)
    return g, min(k, g.n)


def uneven_batch(count, seed, nmin_log2=10, nmax_log2=18, workers=1):
    """The mixed batch: n log-uniform in [2^nmin_log2, 2^nmax_log2]; even
    indices are grids with aspect 2^U(-2,2), odd ones RGGs of mean degree
    U(8,16); vwgt U{1..4}, adjwgt U{1..10}; k from {2,4,8,16,32,64}.
    Every graph's (n, k, seed) is drawn first from the batch seed, so the
    result does not depend on `workers` (processes building the graphs).
    Returns a list of (CSR, k)."""
    rng = np.random.default_rng(seed)
    params = []
    for i in range(count):
        n = int(round(2.0 ** rng.uniform(nmin_log2, nmax_log2)))
        k = int(rng.choice([2, 4, 8, 16, 32, 64]))
        gseed = int(rng.integers(0, 2**63 - 1))
        params.append((i, n, k, gseed))
    if workers <= 1:
        return [_batch_member(p) for p in params]
    import multiprocessing as mp
    # largest graphs first so the pool stays busy to the end
    order = sorted(range(count), key=lambda i: -params[i][1])
    with mp.get_context("fork").Pool(workers) as pool:
        built = pool.map(_batch_member, [params[i] for i in order], chunksize=1)
    out = [None] * count
    for j, i in enumerate(order):
        out[i] = built[j]
    return out


def serve_pool(count, seed, nmin_log2=12, nmax_log2=22, workers=1):
    """Graphs of the open-loop serving stream: built as in uneven_batch, with
    n log-uniform in [2^nmin_log2, 2^nmax_log2] and two constraints with
    vertex weights U{1..8} each.  Returns a list of CSR (ncon 2)."""
    rng = np.random.default_rng(seed)
    params = []
    for i in range(count):
        n = int(round(2.0 ** rng.uniform(nmin_log2, nmax_log2)))
        gseed = int(rng.integers(0, 2**63 - 1))
        params.append((i, n, 2, gseed, 2, 8))
    if workers <= 1:
        return [_batch_member(p)[0] for p in params]
    import multiprocessing as mp
    order = sorted(range(count), key=lambda i: -params[i][1])
    with mp.get_context("fork").Pool(workers) as pool:
        built = pool.map(_batch_member, [params[i] for i in order], chunksize=1)
    out = [None] * count
    for j, i in enumerate(order):
        out[i] = built[j][0]
    return out


def config_graph(cfg, seed=1):
    """Graphs of the headline configurations: 'grid256' (65,536 v, k=8),
    'rgg20' (2^20 v, k=64), 'rmat22' (2^22 v, k=256).  Returns (CSR, k)."""
    if cfg == "grid256":
        return grid(256, 256), 8
    if cfg == "rgg20":
        return rgg(1 << 20, seed), 64
    if cfg == "rmat22":
        return rmat(22, 16, seed), 256
    if cfg.startswith("rgg"):
        return rgg(1 << int(cfg[3:]), seed), 64
    if cfg.startswith("rmat"):
        return rmat(int(cfg[4:]), 16, seed), 256
    raise ValueError(cfg)
