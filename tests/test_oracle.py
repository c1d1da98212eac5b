"""Pins the CPU oracle (oracle/liboracle.so) before it is trusted as the checker.

Every oracle function is checked against something other than itself:
reference-held known answers (SFMT KAT), published vectors (Philox), outputs
of the reference's own functions (golden fixtures written by
oracle/_ref/harness_<app>), library routines (numpy/searchsorted/sets) and
closed forms.  A few tests plant a mistake in an oracle output and check that
the pin notices.
"""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

import graphs
import oracle_bind as ob
from conftest import GOLDEN, ROOT, golden_case_args, load_golden_graph

REF_HARNESS = os.path.join(ROOT, "oracle", "_ref")


def test_sfmt_known_answers():
    kat = np.load(os.path.join(GOLDEN, "sfmt_kat.npz"))
    assert np.array_equal(ob.sfmt_stream(1234, 1000), kat["init_gen_rand_1234"])
    assert np.array_equal(ob.sfmt_by_array([0x1234, 0x5678, 0x9ABC, 0xDEF0], 1000),
                          kat["init_by_array_1234_5678_9abc_def0"])


def test_sfmt_stream_skip_is_a_suffix():
    full = ob.sfmt_stream(99, 3000)
    assert np.array_equal(ob.sfmt_stream(99, 1000, skip=2000), full[2000:])


def test_real_conversions_closed_form():
    assert ob.real2(0) == 0.0
    assert ob.real1(0) == 0.0
    assert ob.real1(0xFFFFFFFF) == 1.0
    top = ob.real2(0xFFFFFFFF)
    assert top < 1.0 and top == (2.0 ** 32 - 1) / 2.0 ** 32
    rng = np.random.default_rng(5)
    for u in rng.integers(0, 2 ** 32, 200, dtype=np.uint64):
        assert ob.real2(int(u)) == float(u) / 2.0 ** 32          # exact: scaling by a power of two
        assert abs(ob.real1(int(u)) - float(u) / 4294967295.0) <= 2.3e-16 * max(1.0, float(u) / 4294967295.0)


def test_philox_published_vectors():
    # Known-answer vectors for Philox-4x32-10 (Random123 distribution, kat_vectors).
    assert [int(x) for x in ob.philox([0, 0, 0, 0], [0, 0])] == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert [int(x) for x in ob.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)] == \
        [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert [int(x) for x in ob.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                                      [0xA4093822, 0x299F31D0])] == [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def _weights(g, app):
    return g.weight if app == "deepwalk" else np.ones(g.ne)


@pytest.mark.parametrize("gname", ["small", "hub"])
@pytest.mark.parametrize("app", ["deepwalk", "ppr"])
def test_tables_match_reference(gname, app):
    g, _ = load_golden_graph(gname)
    z = np.load(os.path.join(GOLDEN, f"tables_{gname}_{app}.npz"))
    w = _weights(g, app)
    alias, zero = ob.alias_build(g, w)
    assert zero == 0
    assert np.array_equal(alias.view(np.uint8).reshape(-1, 16), z["alias"])
    assert np.array_equal(ob.its_build(g, w).view(np.uint64), z["prefix"].view(np.uint64))
    assert np.array_equal(ob.rs_build(g, w), z["rs_max"])
    assert np.array_equal(z["rs_weight"], w)


def alias_reconstruction_error(g, w, table):
    """max relative error of the per-vertex distribution the alias table encodes vs w / sum(w)."""
    worst = 0.0
    for v in range(g.nv):
        lo, hi = g.offset[v], g.offset[v + 1]
        d = hi - lo
        if d == 0:
            continue
        t = table[lo:hi]
        nb = g.adj[lo:hi]
        idx = {int(x): i for i, x in enumerate(nb)}
        p = np.zeros(d)
        for row in t:
            p[idx[int(row["first"])]] += row["prob"] / d
            p[idx[int(row["second"])]] += (1.0 - row["prob"]) / d
        target = w[lo:hi] / w[lo:hi].sum()
        worst = max(worst, float(np.max(np.abs(p - target) / target)))
        assert np.all((t["prob"] >= 0.0) & (t["prob"] <= 1.0))
    return worst


def test_alias_tables_encode_the_weights():
    g, _ = load_golden_graph("hub")
    table, _ = ob.alias_build(g, g.weight)
    assert alias_reconstruction_error(g, g.weight, table) < 1e-12


def test_alias_invariant_detects_a_planted_swap():
    g, _ = load_golden_graph("small")
    table, _ = ob.alias_build(g, g.weight)
    bad = table.copy()
    v = int(np.argmax(g.degree))
    row = bad[g.offset[v] + 1]
    row["first"], row["second"] = row["second"], row["first"]
    bad[g.offset[v] + 1] = row
    assert alias_reconstruction_error(g, g.weight, bad) > 1e-6


def test_its_build_is_sequential_prefix():
    g, _ = load_golden_graph("hub")
    P = ob.its_build(g, g.weight)
    for v in range(0, g.nv, 7):
        lo, hi = g.offset[v], g.offset[v + 1]
        if hi > lo:
            assert np.array_equal(P[lo:hi], np.add.accumulate(g.weight[lo:hi]))


def test_rs_build_is_segment_max():
    g, _ = load_golden_graph("hub")
    m = ob.rs_build(g, g.weight)
    live = g.degree > 0
    assert np.array_equal(m[live], np.maximum.reduceat(g.weight, g.offset[:-1][live]))
    assert np.all(m[~live] == 0.0)


def test_upper_bound_equals_searchsorted():
    rng = np.random.default_rng(3)
    for n in [1, 2, 3, 7, 64, 1000, 4097]:
        a = np.cumsum(rng.integers(0, 3, n).astype(np.float64))   # non-decreasing, with ties
        for t in np.concatenate([a[rng.integers(0, n, 20)], rng.random(20) * (a[-1] + 1.0), [-1.0]]):
            assert ob.lib().orc_upper_bound(a.ctypes.data, n, float(t)) == np.searchsorted(a, t, side="right")


def test_is_neighbor_equals_set_membership():
    g, _ = load_golden_graph("hub")
    og = ob.Graph(g)
    rng = np.random.default_rng(4)
    pairs = np.concatenate([rng.integers(0, g.nv, (3000, 2)),
                            np.stack([np.zeros(500, np.int64), g.neighbors(0)[:500]], 1)])
    for u, v in pairs:
        expect = int(v) in set(g.neighbors(int(u)).tolist())
        assert bool(ob.lib().orc_is_neighbor(ctypes.byref(og.s), int(u), int(v))) == expect


def run_golden_walk(case, graph, sources, seed, tseed):
    g = graph
    w = _weights(g, case["app"])
    alias, _ = ob.alias_build(g, w)
    og = ob.Graph(g, alias, ob.its_build(g, w), w, ob.rs_build(g, w))
    K, KT = case["K"], max(case["KT"], 1)
    n = len(sources)
    d = ob.sfmt_stream(seed, n * K).reshape(n, K)
    td = ob.sfmt_stream(tseed, n * KT).reshape(n, KT)
    return ob.walk(og, case["app"], case["execution"], case["sampling"], case["length"], sources, d, K, td, KT,
                   **{k: v for k, v in golden_case_args(case).items() if k != "stop_prob"},
                   stop_prob=golden_case_args(case)["stop_prob"])


def test_walks_match_reference(golden_index, golden_graphs):
    for case in golden_index["walks"]:
        g, src = golden_graphs[case["graph"]]
        z = np.load(os.path.join(GOLDEN, f"walk_{case['case']}.npz"))
        r = run_golden_walk(case, g, src, golden_index["seed"], golden_index["tseed"])
        for key in ("paths", "lengths", "ends", "draws_used"):
            assert np.array_equal(r[key], z[key]), (case["case"], key)
        assert np.array_equal((r["status"] == 1).astype(np.int32), z["status"]), case["case"]
        assert r["steps"] == int(z["steps"][0]), case["case"]


def test_walk_fixture_detects_planted_draw_shift(golden_index, golden_graphs):
    case = next(c for c in golden_index["walks"] if c["case"] == "deepwalk_static_alias_small")
    g, src = golden_graphs[case["graph"]]
    z = np.load(os.path.join(GOLDEN, f"walk_{case['case']}.npz"))
    r = run_golden_walk(case, g, src, golden_index["seed"] + 1, golden_index["tseed"])
    assert not np.array_equal(r["paths"], z["paths"])


def test_ppr_power_two_vertex_closed_form():
    g = graphs.csr_from_undirected(2, [[0, 1]], weights=False, labels=False)
    for a in (0.15, 0.2, 0.5):
        pi, _ = ob.ppr_power(g, a, [0])
        assert abs(pi[1] - 1.0 / (2.0 - a)) < 1e-12
        assert abs(pi[0] - (1.0 - a) / (2.0 - a)) < 1e-12


def test_ppr_power_star_closed_form():
    m = 9
    g = graphs.csr_from_undirected(m + 1, [[0, i] for i in range(1, m + 1)], weights=False, labels=False)
    a = 0.2
    pi, _ = ob.ppr_power(g, a, [0])
    assert abs(pi[0] - (1 - a) / (2 - a)) < 1e-12
    assert np.allclose(pi[1:], (1.0 / (2 - a)) / m, rtol=0, atol=1e-12)


def test_ppr_power_agrees_with_walk_oracle():
    g, _ = load_golden_graph("small")
    a = 0.2
    src = np.full(200000, int(np.argmax(g.degree)), dtype=np.int32)
    pi, _ = ob.ppr_power(g, a, src[:1])
    og = ob.Graph(g)
    K = 64
    d = ob.sfmt_stream(7, len(src) * K).reshape(-1, K)
    td = ob.sfmt_stream(8, len(src) * K).reshape(-1, K)
    r = ob.walk(og, "ppr", 0, 0, 1, src, d, K, td, K, stop_prob=a)
    assert np.all(r["status"] == 0)
    est = np.bincount(r["ends"], minlength=g.nv) / len(src)
    bound = np.sum(np.sqrt(pi * (1 - pi) / len(src)))
    assert np.abs(est - pi).sum() < 1.5 * bound


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_HARNESS, "harness_node2vec")),
                    reason="reference harness not built (make -C oracle ref; build container only)")
@pytest.mark.parametrize("seed", [21, 22])
def test_oracle_matches_reference_on_fresh_graphs(seed):
    """Fresh random graphs, all apps: the reference harness and the oracle agree."""
    g = graphs.chung_lu(300 + seed, 1500, 2.0, 2, 120, seed)
    src = np.nonzero(g.degree > 0)[0].astype(np.int32)
    cases = [("deepwalk", 1, 2, 12, {}), ("deepwalk", 1, 1, 12, {}), ("deepwalk", 1, 3, 12, {}),
             ("deepwalk", 0, 0, 12, {}), ("ppr", 0, 0, 1, {"-sp": "0.3"}),
             ("node2vec", 2, 4, 12, {"-p": "4", "-q": "0.25"}), ("metapath", 2, 1, 12, {"-s": "2,0,4"})]
    with tempfile.TemporaryDirectory() as tmp:
        gdir = os.path.join(tmp, "g")
        graphs.write_reference_csr(g, gdir)
        src.tofile(os.path.join(tmp, "src.bin"))
        for app, em, sm, L, extra in cases:
            out = os.path.join(tmp, f"o_{app}_{em}_{sm}")
            os.makedirs(out)
            K, KT = 512, 64
            cmd = [os.path.join(REF_HARNESS, f"harness_{app}"), "-f", gdir, "-ew", "-el", "-o", out, "-mode", "walk",
                   "-em", str(em), "-sm", str(sm), "-l", str(L), "-src", os.path.join(tmp, "src.bin"), "-K", str(K),
                   "-KT", str(KT), "-seed", str(seed), "-tseed", str(seed + 100)]
            for k, v in extra.items():
                cmd += [k, v]
            subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
            ref_paths = np.fromfile(os.path.join(out, "paths.bin"), np.int32).reshape(len(src), -1)
            ref_used = np.fromfile(os.path.join(out, "draws_used.bin"), np.int32)
            case = {"app": app, "execution": em, "sampling": sm, "length": L, "K": K, "KT": KT, "extra": extra}
            r = run_golden_walk(case, g, src, seed, seed + 100)
            assert np.array_equal(r["paths"], ref_paths), (app, em, sm)
            assert np.array_equal(r["draws_used"], ref_used), (app, em, sm)


def _custom7_cases():
    import json
    with open(os.path.join(GOLDEN, "index.json")) as f:
        return [c for c in json.load(f)["functor_walks"] if c["case"].startswith("custom7")]


@pytest.mark.parametrize("case", _custom7_cases(), ids=lambda c: c["case"])
def test_weighted_node2vec_matches_the_reference_engine(case):
    """Node2Vec over edge weights (factor times w_e, bound max_e w_e times max(1, p_, q_)): the
    oracle equals the reference's own dynamic_max_weight_rs_move (rejection_sampling.h:368-402)
    driving the weighted functor of tests/dropin_apps/customwalk.cpp (kind 7), bit for bit."""
    import thunderrw_amd as T

    g, _ = load_golden_graph(case["graph"])
    z = np.load(os.path.join(GOLDEN, "functor_%s.npz" % case["case"]))
    src = T.sources_all(g.degree)  # the fixture's walkers: generate_all_source order
    K = case["K"]
    d = ob.sfmt_stream(1234, len(src) * K).reshape(len(src), K)
    p, q = float(case["extra"]["-p"]), float(case["extra"]["-q"])
    r = ob.walk(ob.Graph(g), "node2vec", 2, 4, case["length"], src, d, K, p=p, q=q, weighted=True)
    assert np.array_equal(r["paths"], z["paths"])
    assert np.array_equal(r["lengths"], z["lengths"])
    assert r["steps"] == int(z["steps"][0])
    assert np.array_equal(r["draws_used"], z["draws_used"])
    # The unweighted walk from the same draws is a different one (the fixture pins the weights).
    u = ob.walk(ob.Graph(g), "node2vec", 2, 4, case["length"], src, d, K, p=p, q=q)
    assert not np.array_equal(u["paths"], z["paths"])


def _tiny_weighted_graph():
    pairs = [(0, 1), (0, 2), (0, 3), (1, 2), (2, 3), (3, 4), (1, 4), (4, 5), (2, 5)]
    w = [1.0, 4.0, 2.0, 3.0, 1.5, 4.5, 2.5, 1.0, 3.5]
    g = graphs.csr_from_undirected(6, pairs, weights=False, labels=False)
    wt = {}
    for (a, b), x in zip(pairs, w):
        wt[(a, b)] = wt[(b, a)] = x
    weight = np.array([wt[(int(v), int(u))] for v in range(6) for u in g.neighbors(v)])
    return graphs.CSR(g.degree, g.adj, weight), wt


def _weighted_n2v_law(g, wt, prev, cur, p_, q_):
    """P(next | prev, cur) of Node2Vec over edge weights: w(cur, x) * {p_: x == prev, 1: x a neighbour
    of prev, q_: otherwise}, normalised (the law max-weight rejection samples, node2vec.cpp:24-38)."""
    nb = [int(x) for x in g.neighbors(cur)]
    pn = set(int(x) for x in g.neighbors(prev))
    a = np.array([wt[(cur, x)] * (p_ if x == prev else 1.0 if x in pn else q_) for x in nb])
    return nb, a / a.sum()


@pytest.mark.parametrize("pq", [(2.0, 0.5), (0.5, 2.0)])
def test_weighted_node2vec_second_order_law_brute_force(pq):
    """Tiny graph, exact second-order law by enumeration: the oracle's move from (source, first)
    follows it (chi-square, every (prev, cur) pair with enough visits), and its first move follows
    w / sum w.  A planted mistake -- the unweighted functor -- fails the same test."""
    from scipy import stats

    g, wt = _tiny_weighted_graph()
    p, q = pq
    p_, q_ = 1.0 / p, 1.0 / q
    n, K = 240000, 96
    src = np.repeat(np.arange(6, dtype=np.int32), n // 6)
    d = ob.sfmt_stream(77, n * K).reshape(n, K)

    def pvalues(weighted):
        r = ob.walk(ob.Graph(g), "node2vec", 2, 4, 3, src, d, K, p=p, q=q, weighted=weighted)
        assert np.all(r["lengths"] == 3)
        path = r["paths"]
        out = []
        for s in range(6):
            first = path[path[:, 0] == s, 1]
            nb = [int(x) for x in g.neighbors(s)]
            ws = np.array([wt[(s, x)] for x in nb])
            obs = np.array([(first == x).sum() for x in nb])
            out.append(stats.chisquare(obs, ws / ws.sum() * obs.sum()).pvalue)
            for a in nb:
                sel = path[(path[:, 0] == s) & (path[:, 1] == a), 2]
                if len(sel) < 500:
                    continue
                nb2, law = _weighted_n2v_law(g, wt, s, a, p_, q_)
                obs = np.array([(sel == x).sum() for x in nb2])
                out.append(stats.chisquare(obs, law * obs.sum()).pvalue)
        return np.maximum(np.array(out), 1e-300)

    good = pvalues(True)
    assert len(good) > 20
    assert stats.combine_pvalues(good).pvalue > 1e-3
    assert good.min() > 1e-5
    bad = pvalues(False)
    assert stats.combine_pvalues(bad).pvalue < 1e-12
