"""Free-running mode (Philox stream) judged statistically, on graphs of 2^20 vertices, over every step of
the walk.  Replay mode pins the arithmetic bit for bit elsewhere; these tests pin the free-running
stream and its keying: that the walks the GPU produces follow the reference's laws.

Laws (the reference's Moves with the apps' weights):
  * static alias, ITS, rejection (alias_sampling.h:147-172, inverse_transformation_sampling.h:259-286,
    rejection_sampling.h:334-366) and uniform (uniform_sampling.h:47-63): P(u | v) = w(v,u) / sum w
    (1 / deg for uniform), the same at every step;
  * Node2Vec (node2vec.cpp:24-38 with rejection_sampling.h:368-402), second order:
    P(x | prev, cur) proportional to {1/p: x == prev, 1: x a neighbour of prev, 1/q: otherwise}, times
    w(cur, x) when weighted; the first move follows w(src, x) (uniform when unweighted);
  * MetaPath (metapath.cpp:22-30 with dynamic ITS): the next vertex is uniform over the current
    vertex's neighbours whose edge label is schema[(length - 1) % |schema|].

Tests and thresholds (all seeds fixed, so a run is reproducible):
  * every move follows an edge (and, for MetaPath, an edge of the step's label);
  * per-vertex (or per (prev, cur) pair, or per (vertex, label)) Pearson chi-square of the
    neighbour counts against the law, over every vertex with at least 5 expected moves per neighbour,
    pooled: sum of statistics against chi-square with the summed degrees of freedom, p > 1e-4, and no
    single vertex below 1e-4 / (number of vertices tested);
  * keying: moves out of the same vertex by the same walker at steps i and i + 2 (a walk that came
    back) agree as often as independent draws would (sum over the law of p^2), within 5 standard
    deviations -- a step counter that did not advance would make them agree every time; and two
    walkers from the same source agree on their first move as often as independent draws would."""
import numpy as np
import pytest
from scipy import stats

import thunderrw_amd as T

pytestmark = pytest.mark.gpu

P_POOLED = 1e-4


class Csr:
    def __init__(self, hg):
        deg, adj, w, lab = hg.arrays()
        self.deg = deg.copy()
        self.adj = adj.copy()
        self.w = None if w is None else w.copy()
        self.lab = None if lab is None else lab.copy()
        self.nv = len(self.deg)
        self.off = np.zeros(self.nv + 1, np.int64)
        np.cumsum(self.deg, out=self.off[1:])
        self.src = np.repeat(np.arange(self.nv, dtype=np.int64), self.deg)
        self.keys = self.src * self.nv + self.adj

    def edges_of(self, keys):
        """Edge ids of sorted unique move keys (a * nv + b); asserts each is an edge."""
        e = np.minimum(np.searchsorted(self.keys, keys), len(self.keys) - 1)
        assert np.all(self.keys[e] == keys), "a move that follows no edge"
        return e

    def move_counts(self, a, b):
        """Moves a -> b counted per edge; asserts every move follows an edge."""
        u, c = np.unique(a.astype(np.int64) * self.nv + b, return_counts=True)
        out = np.zeros(len(self.adj), np.float64)
        out[self.edges_of(u)] = c
        return out

    def reverse_edge(self):
        return self.edges_of_unsorted(self.adj.astype(np.int64) * self.nv + self.src)

    def edges_of_unsorted(self, keys):
        order = np.argsort(keys)
        e = np.empty(len(keys), np.int64)
        e[order] = self.edges_of(keys[order])
        return e

    def seg(self, v):
        return slice(int(self.off[v]), int(self.off[v + 1]))


@pytest.fixture(scope="module")
def sparse_graph(gpu_ctx):
    """2^20 vertices, Chung-Lu exponent 2.1, 2M undirected edges (average degree 4), dmin 2, dmax
    4096, FP64 U[1,5) weights and labels U{0..4}, mirrored: sparse enough that 2^20 walks of length
    80 move along every edge about 20 times."""
    hg = T.HostGraph.chung_lu(1 << 20, 2 << 20, 2.1, 2.0, 4096.0, 61, True, 5)
    g = Csr(hg)
    hg.close()
    dg = T.DeviceGraph(gpu_ctx, g.deg, g.adj, g.w, g.lab)
    yield g, dg
    dg.close()


def pooled(stats_dofs, label):
    s = np.array([x[0] for x in stats_dofs])
    d = np.array([x[1] for x in stats_dofs])
    p_each = stats.chi2.sf(s, d)
    p_all = stats.chi2.sf(s.sum(), d.sum())
    print("%s: %d groups, %d dof, pooled chi2 %.1f, p %.3g, min p %.3g"
          % (label, len(s), d.sum(), s.sum(), p_all, p_each.min()), flush=True)
    assert len(s) >= 200, label
    assert p_all > P_POOLED, label + " pooled chi-square"
    assert p_each.min() > P_POOLED / len(s), label + " single group"


def first_order_laws(g, sampler):
    """Per-edge P(u | v) of the first-order samplers."""
    if sampler == "uniform":
        return np.repeat(1.0 / np.maximum(g.deg, 1), g.deg)
    tot = np.add.reduceat(g.w, g.off[:-1][g.deg > 0])
    full = np.zeros(g.nv)
    full[g.deg > 0] = tot
    return g.w / np.repeat(full, g.deg)


def check_first_order(g, paths, law, label):
    assert np.all(paths >= 0)
    check_law(g, paths, law, label)
    check_keying_within_walkers(g, paths, law, label)
    check_keying_across_walkers(g, paths, law, label)


def check_law(g, paths, law, label):
    counts = g.move_counts(paths[:, :-1].ravel(), paths[:, 1:].ravel())
    live = g.deg > 0
    vis = np.zeros(g.nv)
    vis[live] = np.add.reduceat(counts, g.off[:-1][live])
    # vertices whose every neighbour expects at least 5 moves
    minp = np.zeros(g.nv)
    minp[live] = np.minimum.reduceat(law, g.off[:-1][live])
    cand = np.nonzero(live & (g.deg >= 2) & (vis * minp >= 5.0))[0]
    rng = np.random.default_rng(5)
    cand = rng.choice(cand, min(len(cand), 3000), replace=False)
    res = []
    for v in cand:
        s = g.seg(v)
        obs, exp = counts[s], vis[v] * law[s]
        res.append((float(((obs - exp) ** 2 / exp).sum()), len(obs) - 1))
    pooled(res, label)


def check_keying_within_walkers(g, paths, law, label):
    # Moves out of v at steps i and i + 2 of a walk that came back to v.
    # Given the return, the first move u has probability p(u|v) p(v|u) / sum_u' p(u'|v) p(v|u'), the
    # second p(u|v) independently, so they agree with probability
    # E_v = sum_u p(u|v)^2 p(v|u) / sum_u p(u|v) p(v|u).
    rev = law[g.reverse_edge()]
    num = np.bincount(g.src, weights=law * law * rev, minlength=g.nv)
    den = np.bincount(g.src, weights=law * rev, minlength=g.nv)
    ev = np.divide(num, den, out=np.zeros(g.nv), where=den > 0)
    rows, cols = np.nonzero(paths[:, :-3] == paths[:, 2:-1])
    v = paths[rows, cols]
    agree = paths[rows, cols + 1] == paths[rows, cols + 3]
    pr = ev[v]
    mean, sd = pr.sum(), np.sqrt((pr * (1 - pr)).sum())
    print("%s keying within walkers: %d returns, %d agree, %.1f expected (sd %.1f)"
          % (label, len(v), agree.sum(), mean, sd), flush=True)
    assert len(v) > 100000
    assert abs(agree.sum() - mean) < 5 * sd, label + " keying within walkers"


def check_keying_across_walkers(g, paths, law, label):
    # First moves of walkers that share a source agree with probability sum_u p(u|v)^2.
    law2 = np.bincount(g.src, weights=law * law, minlength=g.nv)
    src = paths[:, 0]
    order = np.argsort(src, kind="stable")
    s_sorted, f_sorted = src[order], paths[order, 1]
    pair = s_sorted[1:] == s_sorted[:-1]
    agree = (f_sorted[1:] == f_sorted[:-1])[pair]
    pr = law2[s_sorted[1:][pair]]
    mean, sd = pr.sum(), np.sqrt((pr * (1 - pr)).sum())
    print("%s keying across walkers: %d pairs, %d agree, %.1f expected (sd %.1f)"
          % (label, pair.sum(), agree.sum(), mean, sd), flush=True)
    assert pair.sum() > 100000
    assert abs(agree.sum() - mean) < 5 * sd, label + " keying across walkers"


FIRST_ORDER = [("alias", "fused", None), ("its", "fused", None), ("its", "fused", "search"),
               ("rejection", "fused", None), ("uniform", "fused", None), ("alias", "queue", None),
               ("its", "step", None)]


@pytest.mark.parametrize("sm,form,variant", FIRST_ORDER, ids=lambda x: str(x))
def test_first_order_transitions_chi_square_every_step(sparse_graph, sm, form, variant):
    g, dg = sparse_graph
    n, L = 1 << 20, 80
    src = T.sources_random(g.deg, n, seed=71)
    if sm == "uniform":
        prm = T.make_params("deepwalk", execution="uniform", sampling="uniform", length=L, rng="philox", seed=72,
                            form=form)
    else:
        dg.build_tables(sm, "edge")
        prm = T.make_params("deepwalk", execution="static", sampling=sm, length=L, rng="philox", seed=72, form=form)
    dg.set_tuning(its_guide_off=1 if variant == "search" else 0)
    try:
        r = T.walk(dg, prm, src)
    finally:
        dg.set_tuning(its_guide_off=0)
    assert r["steps"] == n * (L - 1)
    check_first_order(g, r["paths"], first_order_laws(g, sm), "%s/%s/%s" % (sm, form, variant))


def node2vec_law(g, prev, cur, p_, q_, weighted):
    s = g.seg(cur)
    nb = g.adj[s]
    pn = g.adj[g.seg(prev)]
    alpha = np.where(nb == prev, p_, np.where(np.isin(nb, pn), 1.0, q_))
    a = alpha * (g.w[s] if weighted else 1.0)
    return a / a.sum()


def chi2_of(obs, exp):
    return float(((obs - exp) ** 2 / exp).sum()), len(obs) - 1


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("pq", [(2.0, 0.5), (0.5, 2.0)])
def test_node2vec_second_order_transitions_chi_square(sparse_graph, weighted, pq):
    """Walkers from 256 sources of degree 4..16, 8192 each, length 20: every (prev, cur) pair with
    at least 5 expected moves per neighbour of cur is tested against the second-order law, over
    every step; first moves against w (or uniform).  The same counts fail the first-order law."""
    g, dg = sparse_graph
    p, q = pq
    p_, q_ = 1.0 / p, 1.0 / q
    rng = np.random.default_rng(81)
    pool = np.nonzero((g.deg >= 4) & (g.deg <= 16))[0]
    sources = rng.choice(pool, 256, replace=False).astype(np.int32)
    per = 8192
    src = np.repeat(sources, per)
    L = 20
    prm = T.make_params("node2vec", length=L, p=p, q=q, rng="philox", seed=82,
                        weights="edge" if weighted else "app")
    r = T.walk(dg, prm, src)
    paths = r["paths"]
    assert r["steps"] == len(src) * (L - 1)
    g.move_counts(paths[:, :-1].ravel(), paths[:, 1:].ravel())  # every move follows an edge
    tag = "node2vec weighted=%s p,q=%s" % (weighted, pq)
    # First moves.
    res = []
    first = paths[:, 1].reshape(len(sources), per)
    for i, s0 in enumerate(sources):
        s = g.seg(s0)
        law = g.w[s] / g.w[s].sum() if weighted else np.full(g.deg[s0], 1.0 / g.deg[s0])
        obs = np.bincount(np.searchsorted(g.adj[s], first[i]), minlength=g.deg[s0])
        res.append(chi2_of(obs, per * law))
    pooled(res, tag + " first move")
    # Second-order moves at every later step: counts per (prev, cur, next), grouped by (prev, cur).
    k3 = ((paths[:, :-2].astype(np.int64) * g.nv + paths[:, 1:-1]) << 20) | paths[:, 2:]
    u, c = np.unique(k3.ravel(), return_counts=True)
    del k3
    pair, nxt = u >> 20, (u & ((1 << 20) - 1)).astype(np.int32)
    pu, pstart, pcnt = np.unique(pair, return_index=True, return_counts=True)
    tot = np.add.reduceat(c, pstart)
    cur_u = (pu % g.nv).astype(np.int64)
    cand = np.nonzero((g.deg[cur_u] >= 2) & (g.deg[cur_u] <= 64) & (tot >= 20 * g.deg[cur_u]))[0]
    res, bad = [], []
    for j in cand[:4000]:
        pv, cv = int(pu[j] // g.nv), int(pu[j] % g.nv)
        law = node2vec_law(g, pv, cv, p_, q_, weighted)
        n = tot[j]
        if (n * law).min() < 5:
            continue
        s = g.seg(cv)
        obs = np.zeros(g.deg[cv])
        sl = slice(pstart[j], pstart[j] + pcnt[j])
        obs[np.searchsorted(g.adj[s], nxt[sl])] = c[sl]
        res.append(chi2_of(obs, n * law))
        first_law = g.w[s] / g.w[s].sum() if weighted else np.full(g.deg[cv], 1.0 / g.deg[cv])
        bad.append(chi2_of(obs, n * first_law))
    pooled(res, tag + " second order")
    sb = np.array(bad)
    assert stats.chi2.sf(sb[:, 0].sum(), sb[:, 1].sum()) < 1e-12, tag


def test_metapath_moves_are_uniform_over_matching_labels(sparse_graph):
    """2^21 walkers, 4096 from each of 512 sources of degree 8..64, schema 0,1,2,3,4, length 80:
    every move follows an edge of the step's label, and per (vertex, label) the next vertex is
    uniform over the matching neighbours; walks end only where no neighbour matches
    (gather-terminated)."""
    g, dg = sparse_graph
    schema = np.array([0, 1, 2, 3, 4])
    L = 80
    rng = np.random.default_rng(91)
    pool = np.nonzero((g.deg >= 8) & (g.deg <= 64))[0]
    src = np.repeat(rng.choice(pool, 512, replace=False), 4096).astype(np.int32)
    prm = T.make_params("metapath", length=L, schema=schema.tolist(), rng="philox", seed=92)
    r = T.walk(dg, prm, src)
    paths, lengths = r["paths"], r["lengths"]
    assert set(np.unique(r["status"]).tolist()) <= {0, 1}
    cols = np.arange(L - 1)[None, :]
    valid = cols < (lengths[:, None] - 1)
    a = paths[:, :-1][valid].astype(np.int64)
    b = paths[:, 1:][valid]
    lab_step = np.broadcast_to(schema[cols % len(schema)], valid.shape)[valid]
    u, c = np.unique(((a * g.nv + b) << 3) | lab_step, return_counts=True)
    e = g.edges_of(u >> 3)  # every move follows an edge ...
    assert np.array_equal(g.lab[e], (u & 7).astype(g.lab.dtype))  # ... of the step's label
    counts = np.zeros(len(g.adj))
    np.add.at(counts, e, c)
    # gather-terminated walkers: no neighbour of the end vertex carries the step's label
    gt = np.nonzero(r["status"] == 1)[0]
    assert len(gt) > 1000
    for i in gt[:2000]:
        v = paths[i, lengths[i] - 1]
        assert not np.any(g.lab[g.seg(v)] == schema[(lengths[i] - 1) % len(schema)])
    # uniformity per (vertex, label): the moves out of v along label-l edges
    ev = g.src[e]
    groups, gcnt = np.unique(ev * 8 + g.lab[e], return_counts=True)
    res = []
    for key in groups[gcnt >= 2]:
        v, lb = int(key // 8), int(key % 8)
        s = g.seg(v)
        m = g.lab[s] == lb
        obs = counts[s][m]
        k, tot = len(obs), obs.sum()
        if k < 2 or tot < 5 * k:
            continue
        res.append(chi2_of(obs, np.full(k, tot / k)))
        if len(res) >= 4000:
            break
    pooled(res, "metapath")
