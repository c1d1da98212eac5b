"""Pins the PPR checks bench.py applies to served queries (and tests/test_serving.py to its 2^20-vertex
graph) before they judge the GPU: the host power iteration equals the oracle's; what sampling alone
gives for L1 (de Moivre's mean absolute deviation) equals the mean over simulated multinomial draws;
on the CPU oracle's PPR walks (weighted alias, replay SFMT draws, the reference's Moves) the checks
pass, and they fail on a planted mistake: the law of unweighted moves, or a walk that tests
termination before it moves (the reference moves first, creeper.h:306 and ppr.cpp:27-29)."""
import importlib.util
import os

import numpy as np
import pytest
import scipy.sparse as sp

import graphs
import oracle_bind as ob

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_spec = importlib.util.spec_from_file_location("trw_bench", os.path.join(ROOT, "bench.py"))
B = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(B)

ALPHA = 0.2


@pytest.fixture(scope="module")
def weighted_graph():
    g = graphs.chung_lu(1 << 13, 1 << 15, 2.1, 2, 1024, 91)
    strength = np.zeros(g.nv)
    nz = g.degree > 0
    strength[nz] = np.add.reduceat(g.weight, g.offset[:-1][nz])
    A = sp.csr_matrix((g.weight, g.adj, g.offset), shape=(g.nv, g.nv))
    return g, A, strength


def oracle_ppr_walks(g, src, nw, seed):
    """End vertices of nw weighted-alias PPR walks from the oracle, walker i from src[i % |S|]."""
    alias, _ = ob.alias_build(g, g.weight)
    og = ob.Graph(g, alias)
    starts = src[np.arange(nw) % len(src)].astype(np.int32)
    K = 160  # draws per walker; a walk needs more with probability 0.8^160
    d = ob.sfmt_stream(seed, nw * K).reshape(nw, K)
    td = ob.sfmt_stream(seed + 1, nw * K).reshape(nw, K)
    r = ob.walk(og, "ppr", 1, 2, 1, starts, d, K, td, K, stop_prob=ALPHA)
    assert np.all(r["status"] == 0)
    v, c = np.unique(r["ends"], return_counts=True)
    return starts, v, c.astype(np.float64)


@pytest.fixture(scope="module")
def served_like(weighted_graph):
    """Three oracle-walked queries: one source, a few, many."""
    g, _, _ = weighted_graph
    rng = np.random.default_rng(92)
    live = np.nonzero(g.degree > 0)[0]
    out = []
    for k, nw, seed in ((1, 20000, 3), (6, 30000, 5), (300, 4000, 7)):
        src = rng.choice(live, k).astype(np.int32)
        starts, v, c = oracle_ppr_walks(g, src, nw, seed)
        out.append(((src, nw, v, c), starts))
    return out


def test_power_iteration_equals_the_oracle(weighted_graph, served_like):
    g, A, strength = weighted_graph
    starts = [s for _, s in served_like]
    pi, iters, left = B.ppr_power_host(A, strength, starts, ALPHA, tol=1e-13)
    assert left <= 1e-13
    for j, s in enumerate(starts):
        ref, _ = ob.ppr_power(g, ALPHA, s, weighted=True, tol=1e-13)
        assert np.abs(pi[:, j] - ref).max() < 1e-12
        assert abs(pi[:, j].sum() - 1.0) < 1e-9


def test_expected_l1_is_the_mean_over_multinomial_draws():
    rng = np.random.default_rng(93)
    for states, n in ((40, 3000), (5000, 800)):
        p = rng.dirichlet(np.full(states, 0.3))
        draws = rng.multinomial(n, p, size=4000)
        l1 = np.abs(draws / n - p).sum(axis=1)
        assert abs(l1.mean() - B.expected_l1(p, n)) < 4 * l1.std() / np.sqrt(len(l1)) + 1e-12


def test_checks_pass_on_the_oracle_walks(weighted_graph, served_like):
    _, A, strength = weighted_graph
    checks = [c for c, _ in served_like]
    pi, _, _ = B.ppr_power_host(A, strength, [s for _, s in served_like], ALPHA, tol=1e-12)
    res = B.judge_ppr(pi, checks)
    for r in res:
        assert r["ok"], r
        assert r["chi2_dof"] > 10, r


def test_a_wrong_law_fails(weighted_graph, served_like):
    g, A, strength = weighted_graph
    checks = [c for c, _ in served_like]
    starts = [s for _, s in served_like]
    unweighted = sp.csr_matrix((np.ones(g.ne), g.adj, g.offset), shape=(g.nv, g.nv))
    pi_u, _, _ = B.ppr_power_host(unweighted, g.degree.astype(np.float64), starts, ALPHA, tol=1e-12)
    res = B.judge_ppr(pi_u, checks)
    assert not res[0]["ok"] and not res[1]["ok"], res
    # Termination tested before the move: sum_{k>=0} a (1-a)^k x0 P^k = a x0 + (1-a) pi.
    pi, _, _ = B.ppr_power_host(A, strength, starts, ALPHA, tol=1e-12)
    x0 = np.zeros_like(pi)
    for j, s in enumerate(starts):
        np.add.at(x0[:, j], s, 1.0 / len(s))
    res = B.judge_ppr(ALPHA * x0 + (1 - ALPHA) * pi, checks)
    assert not res[0]["ok"] and not res[1]["ok"], res


def test_symmetry_sample_sees_an_unmirrored_weight(weighted_graph):
    g, _, _ = weighted_graph
    assert B.symmetric_sample(g.degree, g.adj, g.weight, g.offset)
    w = g.weight.copy()
    w[17] += 1.0
    assert not B.symmetric_sample(g.degree, g.adj, w, g.offset, n=g.ne)
