"""Small test graphs, built as the reference's own tests build them
(tests/correctness.rs:10-50, tests/parallel.rs:20-38), plus seeded random
graphs for parity sweeps."""
import numpy as np


def cycle4():
    return np.array([0, 2, 4, 6, 8], np.int32), np.array([1, 3, 0, 2, 1, 3, 0, 2], np.int32)


def build_undirected(n, edges):
    """tests/correctness.rs:36-50: neighbour lists sorted ascending."""
    adj = [[] for _ in range(n)]
    for u, v in edges:
        adj[u].append(v)
        adj[v].append(u)
    xadj = [0]
    adjncy = []
    for v in range(n):
        adj[v].sort()
        adjncy.extend(adj[v])
        xadj.append(len(adjncy))
    return np.array(xadj, np.int32), np.array(adjncy, np.int32)


def grid2x3():
    return build_undirected(6, [(0, 1), (1, 2), (3, 4), (4, 5), (0, 3), (1, 4), (2, 5)])


def grid_sorted(rows, cols):
    edges = []
    for r in range(rows):
        for c in range(cols):
            v = r * cols + c
            if c + 1 < cols:
                edges.append((v, v + 1))
            if r + 1 < rows:
                edges.append((v, v + cols))
    return build_undirected(rows * cols, edges)


def grid_parallel_order(rows, cols):
    """tests/parallel.rs:20-38: neighbour order up, down, left, right."""
    xadj = [0]
    adjncy = []
    for r in range(rows):
        for c in range(cols):
            nb = []
            if r > 0:
                nb.append((r - 1) * cols + c)
            if r + 1 < rows:
                nb.append((r + 1) * cols + c)
            if c > 0:
                nb.append(r * cols + c - 1)
            if c + 1 < cols:
                nb.append(r * cols + c + 1)
            adjncy.extend(nb)
            xadj.append(len(adjncy))
    return np.array(xadj, np.int32), np.array(adjncy, np.int32)


def path(n):
    return build_undirected(n, [(i, i + 1) for i in range(n - 1)])


def random_graph(n, p, seed, wmax=1, shuffle_rows=True):
    """Erdos-Renyi G(n, p), symmetric weights U{1..wmax}; rows shuffled so
    neighbour order is not sorted (exercises CSR-order semantics)."""
    rng = np.random.default_rng(seed)
    iu = np.triu_indices(n, 1)
    keep = rng.random(len(iu[0])) < p
    u, v = iu[0][keep], iu[1][keep]
    w = rng.integers(1, wmax + 1, size=len(u)) if wmax > 1 else np.ones(len(u), np.int64)
    adj = [[] for _ in range(n)]
    for a, b, ww in zip(u, v, w):
        adj[a].append((b, ww))
        adj[b].append((a, ww))
    xadj = [0]
    adjncy, adjwgt = [], []
    for x in range(n):
        row = adj[x]
        if shuffle_rows:
            rng.shuffle(row)
        for b, ww in row:
            adjncy.append(b)
            adjwgt.append(ww)
        xadj.append(len(adjncy))
    return (np.array(xadj, np.int32), np.array(adjncy, np.int32),
            np.array(adjwgt, np.int32) if wmax > 1 else None)
