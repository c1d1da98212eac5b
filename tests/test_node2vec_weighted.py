"""Node2Vec over FP64 edge weights (the weighted extension of node2vec.cpp:24-38): a candidate's
weight is its edge weight times Node2Vec's factor, the rejection bound the largest edge weight
times max(1, p_, q_), sampled by max-weight rejection (rejection_sampling.h:368-402).

Pinned by the reference's own engine (tests/golden/functor_custom7_*.npz: oracle/_ref/harness_customwalk
running the weighted functor of tests/dropin_apps/customwalk.cpp) and by the oracle (itself checked
against those fixtures and against the exact second-order law in tests/test_oracle.py).  Replay
mode: bit-exact paths, lengths, statuses and draw counts in every form.  All calls through the C ABI."""
import json
import os

import numpy as np
import pytest

import graphs
import oracle_bind as ob
import thunderrw_amd as T
from conftest import GOLDEN, load_golden_graph

pytestmark = pytest.mark.gpu

FORMS = ["fused", "queue", "step", "graph"]


def params(p, q, L, form, rng="replay", seed=0, max_trials=0):
    return T.make_params("node2vec", execution="dynamic", sampling="max_weight_rejection", length=L, p=p, q=q,
                         rng=rng, seed=seed, form=form, max_trials=max_trials, weights="edge")


def oracle(g, src, L, p, q, K, seed, max_trials=0):
    d = ob.sfmt_stream(seed, len(src) * K).reshape(len(src), K)
    return ob.walk(ob.Graph(g), "node2vec", 2, 4, L, src, d, K, p=p, q=q, max_trials=max_trials, weighted=True)


def assert_same(r, ref, tag):
    for key in ("paths", "lengths", "ends", "status", "draws_used"):
        assert np.array_equal(r[key], ref[key]), (tag, key)
    assert r["steps"] == ref["steps"], tag


def custom7_cases():
    with open(os.path.join(GOLDEN, "index.json")) as f:
        return [c for c in json.load(f)["functor_walks"] if c["case"].startswith("custom7")]


@pytest.mark.parametrize("case", custom7_cases(), ids=lambda c: c["case"])
def test_reference_engine_fixtures_in_every_form(gpu_ctx, case):
    g, _ = load_golden_graph(case["graph"])
    z = np.load(os.path.join(GOLDEN, "functor_%s.npz" % case["case"]))
    src = T.sources_all(g.degree)
    K = case["K"]
    p, q = float(case["extra"]["-p"]), float(case["extra"]["-q"])
    dg = T.DeviceGraph(gpu_ctx, g.degree, g.adj, g.weight)
    try:
        draws = T.replay_fill(1234, len(src), K)
        for form in FORMS:
            r = T.walk(dg, params(p, q, case["length"], form), src, draws)
            assert np.array_equal(r["paths"], z["paths"]), form
            assert np.array_equal(r["lengths"], z["lengths"]), form
            assert np.array_equal(r["draws_used"], z["draws_used"]), form
            assert np.all(r["status"] == 0), form
            assert r["steps"] == int(z["steps"][0]), form
    finally:
        dg.close()


def skewed(g, seed):
    """Weights over three decades, a few exact zeros and ties, mirrored per undirected edge."""
    rng = np.random.default_rng(seed)
    src = np.repeat(np.arange(g.nv), g.degree)
    lo, hi = np.minimum(src, g.adj), np.maximum(src, g.adj)
    key = lo.astype(np.int64) * g.nv + hi
    uniq, inv = np.unique(key, return_inverse=True)
    w = 10.0 ** rng.uniform(-1.5, 1.5, len(uniq))
    w[rng.random(len(uniq)) < 0.02] = 0.0
    w[rng.random(len(uniq)) < 0.05] = 1.0
    return graphs.CSR(g.degree, g.adj, w[inv], g.label)


def hub_heavy(seed):
    rng = np.random.default_rng(seed)
    n = 40000
    parts = [rng.integers(0, n, (90000, 2))]
    for hub, deg in [(0, 30000)] + [(h, 2500) for h in range(1, 5)] + [(h, 120) for h in range(5, 30)]:
        nb = rng.choice(np.arange(30, n), size=deg, replace=False)
        parts.append(np.stack([np.full(deg, hub), nb], 1))
    parts.append(np.array([(a, b) for a in range(30) for b in range(a + 1, 30) if rng.random() < 0.5]))
    return graphs.csr_from_undirected(n, np.concatenate(parts), rng)


def directed(n, seed):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, 6 * n)
    dst = rng.integers(0, n, 6 * n)
    key = np.unique(src.astype(np.int64) * n + dst)
    key = key[key // n != key % n]
    s, d = key // n, key % n
    deg = np.bincount(s, minlength=n).astype(np.int32)
    return graphs.CSR(deg, d.astype(np.int32), 1.0 + 4.0 * rng.random(len(d)))


@pytest.mark.parametrize("gkind", ["chung_lu", "hubs", "skewed", "directed"])
def test_matches_oracle_in_every_form(gpu_ctx, gkind):
    if gkind == "chung_lu":
        g = graphs.chung_lu(8000, 40000, 2.0, 2, 3000, 301)
    elif gkind == "hubs":
        g = hub_heavy(302)
    elif gkind == "skewed":
        g = skewed(graphs.chung_lu(8000, 40000, 2.1, 2, 2000, 303), 304)
    else:
        g = directed(4000, 305)
    dg = T.DeviceGraph(gpu_ctx, g.degree, g.adj, g.weight)
    try:
        rng = np.random.default_rng(306)
        live = np.nonzero(g.degree > 0)[0].astype(np.int32)
        top = np.argsort(g.degree)[::-1][:6].astype(np.int32)
        src = np.concatenate([rng.choice(live, 5000), np.repeat(top, 48), [0, g.nv - 1]]).astype(np.int32)
        L = 20
        for p, q in ((2.0, 0.5), (0.5, 2.0), (1.0, 1.0)):
            K = 40 * L if gkind == "skewed" else 24 * L
            print("weighted node2vec", gkind, p, q, flush=True)
            ref = oracle(g, src, L, p, q, K, 307)
            draws = T.replay_fill(307, len(src), K)
            for form in FORMS:
                r = T.walk(dg, params(p, q, L, form), src, draws)
                assert_same(r, ref, (gkind, p, q, form))
    finally:
        dg.close()


def test_without_weighted_records_or_membership_tables(gpu_ctx):
    """trw_walk_noalloc builds nothing: on a graph without the weighted records (membership tables
    only) and on one with neither, the walk reads the weight array and searches adjacencies, and
    still equals the oracle."""
    g = graphs.chung_lu(6000, 30000, 2.0, 2, 2500, 311)
    src = np.nonzero(g.degree > 0)[0].astype(np.int32)
    L, K = 16, 24 * 16
    ref = oracle(g, src, L, 2.0, 0.5, K, 312)
    draws = T.replay_fill(312, len(src), K)
    for with_tables in (True, False):
        dg = T.DeviceGraph(gpu_ctx, g.degree, g.adj, g.weight)
        try:
            if with_tables:
                dg.build_membership()
            for form in ("fused", "queue", "step"):
                r = T.walk(dg, params(2.0, 0.5, L, form), src, draws, noalloc=True)
                assert_same(r, ref, (with_tables, form))
        finally:
            dg.close()


def test_trial_cap_matches_oracle(gpu_ctx):
    g = skewed(graphs.chung_lu(5000, 30000, 2.0, 2, 1500, 321), 322)
    dg = T.DeviceGraph(gpu_ctx, g.degree, g.adj, g.weight)
    try:
        src = np.nonzero(g.degree > 0)[0].astype(np.int32)
        L, K = 30, 4 * 30 + 64
        for cap in (1, 3):
            ref = oracle(g, src, L, 2.0, 0.5, K, 323, max_trials=cap)
            assert (ref["status"] == T._abi.WALKER_TRIAL_CAP).sum() > 0
            draws = T.replay_fill(323, len(src), K)
            for form in FORMS:
                r = T.walk(dg, params(2.0, 0.5, L, form, max_trials=cap), src, draws)
                assert_same(r, ref, (cap, form))
    finally:
        dg.close()


def test_forms_agree_in_free_running_mode_and_differ_from_unweighted(gpu_ctx):
    g = graphs.chung_lu(20000, 100000, 2.0, 2, 5000, 331)
    dg = T.DeviceGraph(gpu_ctx, g.degree, g.adj, g.weight)
    try:
        src = np.nonzero(g.degree > 0)[0].astype(np.int32)
        res = [T.walk(dg, params(2.0, 0.5, 40, form, rng="philox", seed=9), src) for form in FORMS]
        for r in res[1:]:
            assert np.array_equal(r["paths"], res[0]["paths"])
            assert r["steps"] == res[0]["steps"] == 39 * len(src)
        un = T.walk(dg, T.make_params("node2vec", length=40, p=2.0, q=0.5, rng="philox", seed=9), src)
        assert not np.array_equal(un["paths"], res[0]["paths"])
    finally:
        dg.close()
