"""Small seeded graphs for the tests (test infrastructure, numpy only).

The CSR follows the reference's conventions (util/graph/graph.cpp:463-519):
undirected edges, self-loops and duplicates dropped, both directions stored,
neighbour lists sorted (graph.cpp:764-776).  Edge weights and labels are drawn
once per undirected edge and mirrored, so w(u,v) == w(v,u).
"""
import os

import numpy as np


class CSR:
    def __init__(self, degree, adj, weight=None, label=None):
        self.degree = np.ascontiguousarray(degree, dtype=np.int32)
        self.adj = np.ascontiguousarray(adj, dtype=np.int32)
        self.offset = np.zeros(len(self.degree) + 1, dtype=np.int64)
        np.cumsum(self.degree, out=self.offset[1:])
        self.weight = None if weight is None else np.ascontiguousarray(weight, dtype=np.float64)
        self.label = None if label is None else np.ascontiguousarray(label, dtype=np.int32)

    @property
    def nv(self):
        return int(len(self.degree))

    @property
    def ne(self):
        return int(len(self.adj))

    def neighbors(self, v):
        return self.adj[self.offset[v]:self.offset[v + 1]]


def csr_from_undirected(n, pairs, rng=None, weights=True, labels=True, n_labels=5):
    """pairs: (m, 2) int array of undirected edges (any order, may repeat)."""
    pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    lo = np.minimum(pairs[:, 0], pairs[:, 1])
    hi = np.maximum(pairs[:, 0], pairs[:, 1])
    keep = lo != hi
    key = np.unique(lo[keep] * n + hi[keep])
    lo, hi = key // n, key % n
    m = len(key)
    rng = rng or np.random.default_rng(0)
    w_und = 1.0 + 4.0 * rng.random(m) if weights else None
    l_und = rng.integers(0, n_labels, m).astype(np.int32) if labels else None
    src = np.concatenate([lo, hi])
    dst = np.concatenate([hi, lo])
    eid = np.concatenate([np.arange(m), np.arange(m)])
    order = np.lexsort((dst, src))
    src, dst, eid = src[order], dst[order], eid[order]
    degree = np.bincount(src, minlength=n).astype(np.int32)
    return CSR(degree, dst.astype(np.int32),
               None if w_und is None else w_und[eid],
               None if l_und is None else l_und[eid])


def chung_lu(n, undirected_edges, gamma, dmin, dmax, seed, weights=True, labels=True):
    """Seeded Chung-Lu power-law graph: expected degrees theta_i = dmin*(1-U)^(-1/(gamma-1)),
    truncated at dmax; endpoints drawn i.i.d. proportional to theta."""
    rng = np.random.default_rng(seed)
    theta = dmin * (1.0 - rng.random(n)) ** (-1.0 / (gamma - 1.0))
    theta = np.minimum(theta, dmax)
    p = theta / theta.sum()
    cdf = np.cumsum(p)
    got = np.zeros((0, 2), dtype=np.int64)
    while True:
        need = int((undirected_edges - len(got)) * 1.3) + 16
        u = np.searchsorted(cdf, rng.random(need) * cdf[-1], side="right")
        v = np.searchsorted(cdf, rng.random(need) * cdf[-1], side="right")
        u = np.minimum(u, n - 1)
        v = np.minimum(v, n - 1)
        cand = np.concatenate([got, np.stack([u, v], 1)])
        lo = np.minimum(cand[:, 0], cand[:, 1])
        hi = np.maximum(cand[:, 0], cand[:, 1])
        keep = lo != hi
        key = np.unique(lo[keep] * n + hi[keep])
        got = np.stack([key // n, key % n], 1)
        if len(got) >= undirected_edges:
            got = got[rng.permutation(len(got))[:undirected_edges]]
            break
    return csr_from_undirected(n, got, rng, weights, labels)


def hub_graph(n, hub_degree, extra_edges, seed):
    """A sparse random graph plus vertex 0 joined to hub_degree random vertices."""
    rng = np.random.default_rng(seed)
    nb = rng.choice(np.arange(1, n), size=hub_degree, replace=False)
    hub = np.stack([np.zeros(hub_degree, dtype=np.int64), nb], 1)
    rest = rng.integers(0, n, (extra_edges, 2))
    return csr_from_undirected(n, np.concatenate([hub, rest]), rng)


def write_reference_csr(g, directory):
    """Write the reference's binary CSR files (util/graph/graph.cpp:18-30, 123-212, 233-252)."""
    os.makedirs(directory, exist_ok=True)
    with open(os.path.join(directory, "b_degree.bin"), "wb") as f:
        f.write(np.uint32(4).tobytes())
        f.write(np.int32(g.nv).tobytes())
        f.write(np.int64(g.ne).tobytes())
        f.write(g.degree.tobytes())
    g.adj.tofile(os.path.join(directory, "b_adj.bin"))
    if g.weight is not None:
        g.weight.tofile(os.path.join(directory, "b_edge_weight.bin"))
    if g.label is not None:
        g.label.tofile(os.path.join(directory, "b_edge_label.bin"))
