"""Parity of the HIP engine with the reference (golden fixtures) and the oracle.

Bar: bit-exact tables, paths, lengths, end vertices, statuses, draw counts and
step counts in replay mode; statistical agreement in free-running (Philox) mode.
All calls go through the C ABI (libthunderrw.so)."""
import os

import numpy as np
import pytest

import graphs
import oracle_bind as ob
import thunderrw_amd as T
from conftest import GOLDEN, golden_case_args, load_golden_graph

pytestmark = pytest.mark.gpu

STATIC_SM = {1: "its", 2: "alias", 3: "rejection"}


def dgraph(ctx, g):
    return T.DeviceGraph(ctx, g.degree, g.adj, g.weight, g.label)


def case_params(case, form="auto", rng="replay", seed=0):
    a = golden_case_args(case)
    return T.make_params(case["app"], execution=case["execution"], sampling=case["sampling"], length=case["length"],
                         stop_probability=a["stop_prob"], p=a["p"], q=a["q"], schema=a["schema"], rng=rng,
                         seed=seed, form=form)


def forms_for(app):
    return ["queue", "step", "graph"] if app == "ppr" else ["fused", "queue", "step", "graph"]


@pytest.mark.parametrize("gname", ["small", "hub"])
@pytest.mark.parametrize("app", ["deepwalk", "ppr"])
def test_device_tables_are_byte_identical_to_reference(gpu_ctx, gname, app):
    g, _ = load_golden_graph(gname)
    z = np.load(os.path.join(GOLDEN, f"tables_{gname}_{app}.npz"))
    dg = dgraph(gpu_ctx, g)
    w = "edge" if app == "deepwalk" else "unit"
    try:
        for sm in ("alias", "its", "rejection"):
            dg.build_tables(sm, w)
        assert np.array_equal(dg.alias_table().view(np.uint8).reshape(-1, 16), z["alias"])
        assert np.array_equal(dg.prefix_sum().view(np.uint64), z["prefix"].view(np.uint64))
        assert np.array_equal(dg.rejection_max(), z["rs_max"])
    finally:
        dg.close()


def test_golden_walks_in_every_form(gpu_ctx, golden_index, golden_graphs):
    seed, tseed = golden_index["seed"], golden_index["tseed"]
    devs = {name: dgraph(gpu_ctx, g) for name, (g, _) in golden_graphs.items()}
    try:
        for case in golden_index["walks"]:
            if case["app"] == "metapath" and case["sampling"] != 1:
                continue  # dynamic alias over label weights is not provided on the device
            g, src = golden_graphs[case["graph"]]
            z = np.load(os.path.join(GOLDEN, f"walk_{case['case']}.npz"))
            n = len(src)
            draws = T.replay_fill(seed, n, case["K"])
            tdraws = T.replay_fill(tseed, n, max(case["KT"], 1))
            for form in forms_for(case["app"]):
                r = T.walk(devs[case["graph"]], case_params(case, form), src, draws, tdraws)
                tag = (case["case"], form)
                assert np.array_equal(r["paths"], z["paths"]), tag
                assert np.array_equal(r["lengths"], z["lengths"]), tag
                assert np.array_equal(r["ends"], z["ends"]), tag
                assert np.array_equal(r["draws_used"], z["draws_used"]), tag
                assert np.array_equal((r["status"] == 1).astype(np.int32), z["status"]), tag
                assert np.all((r["status"] == 0) | (r["status"] == 1)), tag
                assert r["steps"] == int(z["steps"][0]), tag
    finally:
        for d in devs.values():
            d.close()


def oracle_walk(g, prm_case, src, K, KT, seed, tseed):
    w = g.weight if prm_case["app"] != "ppr" else np.ones(g.ne)
    alias, _ = ob.alias_build(g, w)
    og = ob.Graph(g, alias, ob.its_build(g, w), w, ob.rs_build(g, w))
    n = len(src)
    d = ob.sfmt_stream(seed, n * K).reshape(n, K)
    td = ob.sfmt_stream(tseed, n * KT).reshape(n, KT)
    a = golden_case_args(prm_case)
    return ob.walk(og, prm_case["app"], prm_case["execution"], prm_case["sampling"], prm_case["length"], src, d, K,
                   td, KT, stop_prob=a["stop_prob"], p=a["p"], q=a["q"], schema=a["schema"])


RANDOM_CASES = [
    ("deepwalk", 1, 2, 30, {}), ("deepwalk", 1, 1, 30, {}), ("deepwalk", 1, 3, 30, {}), ("deepwalk", 0, 0, 30, {}),
    ("deepwalk", 2, 1, 30, {}), ("genericwalk", 1, 2, 17, {}), ("ppr", 0, 0, 1, {"-sp": "0.15"}),
    ("ppr", 1, 2, 3, {"-sp": "0.2"}), ("node2vec", 2, 4, 30, {"-p": "2", "-q": "0.5"}),
    ("node2vec", 2, 4, 30, {"-p": "0.25", "-q": "4"}), ("metapath", 2, 1, 30, {"-s": "0,1,2,3,4"}),
    ("metapath", 2, 1, 30, {"-s": "3"}),
]


@pytest.mark.parametrize("seed", [31, 32])
def test_random_graphs_match_oracle(gpu_ctx, seed):
    g = graphs.chung_lu(6000, 30000, 2.0, 2, 2500, seed)
    g2 = graphs.hub_graph(4000, 3500, 9000, seed + 1)
    for gg in (g, g2):
        dg = dgraph(gpu_ctx, gg)
        try:
            live = np.nonzero(gg.degree > 0)[0].astype(np.int32)
            rng = np.random.default_rng(seed)
            src = np.concatenate([live, rng.choice(live, 3000), np.full(64, int(np.argmax(gg.degree)), np.int32)])
            src = src.astype(np.int32)
            for app, em, sm, L, extra in RANDOM_CASES:
                case = {"app": app, "execution": em, "sampling": sm, "length": L, "extra": extra}
                K, KT = 4 * L + 256, 160
                ref = oracle_walk(gg, case, src, K, KT, seed, seed + 7)
                draws = T.replay_fill(seed, len(src), K)
                tdraws = T.replay_fill(seed + 7, len(src), KT)
                for form in forms_for(app):
                    r = T.walk(dg, case_params(case, form), src, draws, tdraws)
                    tag = (app, em, sm, form)
                    for key in ("paths", "lengths", "ends", "status", "draws_used"):
                        assert np.array_equal(r[key], ref[key]), (tag, key)
                    assert r["steps"] == ref["steps"], tag
        finally:
            dg.close()



def directed_graph(n, seed):
    """A directed CSR (the upload takes any CSR) in which about 5% of the vertices have no
    out-edges, so walks reach dead ends mid-walk; vertex 0 is isolated, vertex 2 a hub."""
    rng = np.random.default_rng(seed)
    deg = rng.integers(1, 40, n)
    deg[rng.random(n) < 0.05] = 0
    deg[:3] = [0, 1, 700]
    rows = []
    for v in range(n):
        cand = rng.choice(n - 1, int(deg[v]), replace=False)
        rows.append(np.sort(cand + (cand >= v)))  # no self-loops, sorted as the reference keeps lists
    adj = np.concatenate(rows).astype(np.int32)
    return graphs.CSR(deg, adj, rng.uniform(1.0, 5.0, len(adj)), rng.integers(0, 5, len(adj)))


def test_dead_ends_and_empty_batches_match_oracle(gpu_ctx):
    """A walker at a vertex without out-edges stops with status DEAD_END (the reference's
    u % 0 is undefined there), at its source or mid-walk, in every app, sampler and form;
    an empty batch is a no-op in every form."""
    g = directed_graph(3000, 41)
    dg = dgraph(gpu_ctx, g)
    try:
        rng = np.random.default_rng(41)
        src = np.concatenate([np.arange(g.nv), rng.integers(0, g.nv, 500)]).astype(np.int32)
        for app, em, sm, L, extra in RANDOM_CASES:
            case = {"app": app, "execution": em, "sampling": sm, "length": L, "extra": extra}
            K, KT = 4 * L + 256, 160
            ref = oracle_walk(g, case, src, K, KT, 43, 44)
            if app not in ("ppr", "metapath"):
                assert np.count_nonzero(ref["status"] == T._abi.WALKER_DEAD_END) > len(src) // 10, app
            draws = T.replay_fill(43, len(src), K)
            tdraws = T.replay_fill(44, len(src), KT)
            for form in forms_for(app):
                r = T.walk(dg, case_params(case, form), src, draws, tdraws)
                tag = (app, em, sm, form)
                for key in ("paths", "lengths", "ends", "status", "draws_used"):
                    assert np.array_equal(r[key], ref[key]), (tag, key)
                assert r["steps"] == ref["steps"], tag
                assert r["dead_ends"] == np.count_nonzero(ref["status"] == T._abi.WALKER_DEAD_END), tag
        empty = np.zeros(0, np.int32)
        for app, em, sm, L, extra in RANDOM_CASES[:1] + RANDOM_CASES[6:7]:
            case = {"app": app, "execution": em, "sampling": sm, "length": L, "extra": extra}
            for form in forms_for(app):
                for rngk in ("replay", "philox"):
                    draws = T.replay_fill(1, 0, 4 * L) if rngk == "replay" else None
                    tdraws = T.replay_fill(2, 0, 8) if rngk == "replay" else None
                    r = T.walk(dg, case_params(case, form, rng=rngk), empty, draws, tdraws)
                    assert r["steps"] == 0 and len(r["lengths"]) == 0 and r["paths"].shape == (0, max(L, 1))
    finally:
        dg.close()

@pytest.mark.parametrize("n_labels,schema", [(8, "7,0,3,9,7,-1,5"), (12, "11,0,3,7,9")])
def test_metapath_label_alphabets_match_oracle(gpu_ctx, n_labels, schema):
    """Labels in [0, 8) take the label-start records (schema labels outside the alphabet
    find no match); wider alphabets take the binary searches.  Both equal the oracle."""
    base = graphs.chung_lu(6000, 30000, 2.0, 2, 2500, 35)
    rng = np.random.default_rng(n_labels)
    g = graphs.CSR(base.degree, base.adj, base.weight, rng.integers(0, n_labels, base.ne))
    dg = dgraph(gpu_ctx, g)
    try:
        live = np.nonzero(g.degree > 0)[0].astype(np.int32)
        src = np.concatenate([live, np.full(64, int(np.argmax(g.degree)), np.int32)]).astype(np.int32)
        case = {"app": "metapath", "execution": 2, "sampling": 1, "length": 40, "extra": {"-s": schema}}
        K, KT = 200, 1
        ref = oracle_walk(g, case, src, K, KT, 36, 37)
        assert ref["steps"] > len(src)
        draws, tdraws = T.replay_fill(36, len(src), K), T.replay_fill(37, len(src), KT)
        for form in forms_for("metapath"):
            r = T.walk(dg, case_params(case, form), src, draws, tdraws)
            for key in ("paths", "lengths", "ends", "status", "draws_used"):
                assert np.array_equal(r[key], ref[key]), (form, key)
            assert r["steps"] == ref["steps"], form
    finally:
        dg.close()


def test_metapath_schema_records_on_a_directed_graph_match_oracle(gpu_ctx):
    """A schema of distinct labels reads the graph's 16-byte schema records.  On a directed graph
    whose sinks (vertices without out-edges) are reachable, a walk that moves onto one ends as a
    dead end there, and one whose next label is missing as gather-terminated, as in the oracle; a
    schema with a repeated label reads the general records; both equal the oracle in every form."""
    base = graphs.chung_lu(6000, 30000, 2.0, 2, 2500, 41)
    rng = np.random.default_rng(42)
    sink = rng.random(base.nv) < 0.1
    keep = ~np.repeat(sink, base.degree)
    deg = np.where(sink, 0, base.degree).astype(np.int32)
    g = graphs.CSR(deg, base.adj[keep], base.weight[keep], rng.integers(0, 5, int(keep.sum())))
    dg = dgraph(gpu_ctx, g)
    try:
        live = np.nonzero(g.degree > 0)[0].astype(np.int32)
        src = np.concatenate([live, np.full(64, int(np.argmax(g.degree)), np.int32)]).astype(np.int32)
        for schema in ("0,1,2,3,4", "2,4", "1,0,1,3"):
            case = {"app": "metapath", "execution": 2, "sampling": 1, "length": 30, "extra": {"-s": schema}}
            K, KT = 120, 1
            ref = oracle_walk(g, case, src, K, KT, 46, 47)
            assert np.any(ref["status"] == T._abi.WALKER_DEAD_END), schema
            assert np.any(ref["status"] == T._abi.WALKER_GATHER_TERMINATED), schema
            draws, tdraws = T.replay_fill(46, len(src), K), T.replay_fill(47, len(src), KT)
            for form in forms_for("metapath"):
                r = T.walk(dg, case_params(case, form), src, draws, tdraws)
                for key in ("paths", "lengths", "ends", "status", "draws_used"):
                    assert np.array_equal(r[key], ref[key]), (schema, form, key)
    finally:
        dg.close()


def test_mixed_metapath_kinds_with_two_schemas(gpu_ctx):
    """Two MetaPath kinds of different schemas in one batch: the graph keeps one schema's records,
    the other kind reads the general records; each half walks as it does alone."""
    g = graphs.chung_lu(20000, 120000, 2.1, 2, 3000, 43)
    dg = dgraph(gpu_ctx, g)
    try:
        live = np.nonzero(g.degree > 0)[0]
        n, split = 60000, 30000
        src = np.random.default_rng(9).choice(live, n).astype(np.int32)
        for form in ("queue", "step", "graph"):
            a = T.make_params("metapath", length=40, schema=[0, 1, 2, 3, 4], rng="philox", seed=8, form=form)
            b = T.make_params("metapath", length=40, schema=[3, 1, 4], rng="philox", seed=8, form=form)
            mixed = T.walk(dg, a, src, mix=(b, split))
            ra = T.walk(dg, a, src[:split])
            rb = T.walk(dg, b, src[split:], walker_id_base=split)
            for key in ("paths", "lengths", "ends", "status"):
                assert np.array_equal(mixed[key][:split], ra[key]), (form, key)
                assert np.array_equal(mixed[key][split:], rb[key]), (form, key)
            assert mixed["steps"] == ra["steps"] + rb["steps"]
    finally:
        dg.close()


def test_tuning_changes_no_walk(gpu_ctx):
    """ITS through guide records, and with them switched off through the search tree on every
    vertex, on hubs only, and off; the queue form with several chunk sizes; the fused form with
    1, 2 and 4 walkers per lane; the launch shapes (fused threads per block and path tile, queue
    blocks per CU, Node2Vec threads per block): identical replay walks, equal to the oracle's.
    Weights with a negative entry build no tree and no guide."""
    g = graphs.hub_graph(5000, 4500, 20000, 71)
    dg = dgraph(gpu_ctx, g)
    try:
        live = np.nonzero(g.degree > 0)[0].astype(np.int32)
        src = np.concatenate([live, np.full(256, int(np.argmax(g.degree)), np.int32)]).astype(np.int32)
        K = 4 * 40 + 64
        draws = T.replay_fill(72, len(src), K)
        cases = [({"app": "deepwalk", "execution": 1, "sampling": 1, "length": 40, "extra": {}},
                  [dict(its_guide_off=0)] + [dict(its_guide_off=1, its_tree_min_degree=v) for v in (0, 64, 1 << 30)],
                  ["fused", "queue", "step"]),
                 ({"app": "deepwalk", "execution": 1, "sampling": 2, "length": 40, "extra": {}},
                  [dict(walkers_per_thread=v) for v in (1, 2, 4)], ["fused"]),
                 ({"app": "metapath", "execution": 2, "sampling": 1, "length": 40, "extra": {"-s": "0,1,2,3,4"}},
                  [dict(queue_chunk=v) for v in (64, 256, 4096)], ["queue"]),
                 # Launch shapes: fused threads per block and path tile, queue blocks per CU,
                 # Node2Vec threads per block.
                 ({"app": "deepwalk", "execution": 1, "sampling": 2, "length": 40, "extra": {}},
                  [dict(walkers_per_thread=w, fused_block=b, fused_tile=t)
                   for w, b, t in ((1, 64, 0), (1, 256, 16), (2, 128, 32), (4, 256, 0), (4, 64, 16))], ["fused"]),
                 ({"app": "deepwalk", "execution": 1, "sampling": 3, "length": 40, "extra": {}},
                  [dict(walkers_per_thread=0, fused_block=b, fused_tile=t) for b, t in ((64, 0), (128, 16), (256, 32))],
                  ["fused"]),
                 ({"app": "node2vec", "execution": 2, "sampling": 4, "length": 40, "extra": {"-p": "2", "-q": "0.5"}},
                  [dict(fused_block=0, fused_tile=0, n2v_block=v) for v in (128, 256)], ["fused"]),
                 ({"app": "metapath", "execution": 2, "sampling": 1, "length": 40, "extra": {"-s": "0,1,2,3,4"}},
                  [dict(queue_chunk=0, queue_blocks_per_cu=v) for v in (1, 2)] +
                  [dict(queue_blocks_per_cu=0, fused_block=b, fused_tile=t) for b, t in ((64, 16), (256, 0))],
                  ["queue", "fused"]),
                 # GRAPH-form steps per launch (0: the built-in 16) over MetaPath's uneven walks.
                 ({"app": "metapath", "execution": 2, "sampling": 1, "length": 40, "extra": {"-s": "0,1,2,3,4"}},
                  [dict(fused_block=0, fused_tile=0, graph_burst=v) for v in (1, 3, 0, 64)], ["graph"])]
        for case, knobs, forms in cases:
            ref = oracle_walk(g, case, src, K, 1, 72, 73)
            for kn in knobs:
                dg.set_tuning(**kn)
                for form in forms:
                    r = T.walk(dg, case_params(case, form), src, draws, T.replay_fill(73, len(src), 1))
                    for key in ("paths", "lengths", "ends", "status", "draws_used"):
                        assert np.array_equal(r[key], ref[key]), (case["app"], kn, form, key)
            dg.set_tuning(walkers_per_thread=0, its_tree_min_degree=0, queue_chunk=0, its_guide_off=0, fused_block=0,
                          fused_tile=0, queue_blocks_per_cu=0, n2v_block=0, graph_burst=0)
    finally:
        dg.close()
    # Negative weights: no tree; the walk still equals the oracle's (binary search, clamped pick).
    w = np.array(g.weight)
    w[::7] *= -0.25
    gn = graphs.CSR(g.degree, g.adj, w, g.label)
    dg = dgraph(gpu_ctx, gn)
    try:
        case = {"app": "deepwalk", "execution": 1, "sampling": 1, "length": 20, "extra": {}}
        src = np.nonzero(gn.degree > 0)[0].astype(np.int32)
        ref = oracle_walk(gn, case, src, 100, 1, 74, 75)
        r = T.walk(dg, case_params(case, "fused"), src, T.replay_fill(74, len(src), 100))
        for key in ("paths", "lengths", "status"):
            assert np.array_equal(r[key], ref[key]), key
    finally:
        dg.close()


def test_autotune_picks_knobs_and_keeps_walks(gpu_ctx):
    g = graphs.chung_lu(30000, 200000, 2.1, 2, 4000, 77)
    dg = dgraph(gpu_ctx, g)
    try:
        src = np.nonzero(g.degree > 0)[0].astype(np.int32)
        for app, kw, knobs, move in [
                ("deepwalk", dict(execution="static", sampling="its", length=30),
                 ("its_tree_min_degree", "fused_block", "fused_tile"), "its"),
                ("deepwalk", dict(execution="static", sampling="alias", length=30),
                 ("walkers_per_thread", "fused_block", "fused_tile"), "alias"),
                ("node2vec", dict(length=30, p=2.0, q=0.5),
                 ("n2v_block", "fused_block", "fused_tile"), "node2vec"),
                ("ppr", dict(stop_probability=0.2), ("queue_chunk", "queue_blocks_per_cu"), "uniform")]:
            prm = T.make_params(app, rng="philox", seed=3, **kw)
            before = T.walk(dg, prm, src)
            rep = dg.autotune(prm, sample_walkers=1 << 16, seed=1)
            assert rep["move"] == move, (app, rep["move"])
            assert set(knobs) <= set(rep["chosen"]) and rep["tuned_steps_per_s"] >= rep["default_steps_per_s"] > 0
            tuned = dg.tuning(move)
            for knob in knobs:
                assert tuned[knob] == rep["chosen"][knob], (app, knob)
            after = T.walk(dg, prm, src)
            for key in ("paths", "lengths", "ends", "status"):
                assert np.array_equal(before[key], after[key]), (app, key)
    finally:
        dg.close()


def test_autotune_mix_picks_queue_knobs_and_keeps_walks(gpu_ctx):
    """trw_autotune_mix on a PPR + MetaPath sample: it tunes the queue chunk and blocks per CU
    (kept for PPR's Move, the first kind), and the mixed batch walks the same after as before."""
    g = graphs.chung_lu(30000, 200000, 2.1, 2, 4000, 78)
    dg = dgraph(gpu_ctx, g)
    try:
        live = np.nonzero(g.degree > 0)[0].astype(np.int32)
        src = np.random.default_rng(5).choice(live, 20000).astype(np.int32)
        a = T.make_params("ppr", length=1, stop_probability=0.15, rng="philox", seed=4, form="queue")
        b = T.make_params("metapath", length=30, schema=[0, 1, 2, 3, 4], rng="philox", seed=4, form="queue")
        before = T.walk(dg, a, src, mix=(b, 10000))
        rep = dg.autotune_mix(a, b, 1 << 15, sample_walkers=1 << 16, seed=2)
        assert rep["move"] == "uniform"
        assert set(rep["chosen"]) == {"queue_chunk", "queue_blocks_per_cu"}
        assert rep["tuned_steps_per_s"] >= rep["default_steps_per_s"] > 0
        assert dg.tuning("uniform")["queue_blocks_per_cu"] == rep["chosen"]["queue_blocks_per_cu"]
        after = T.walk(dg, a, src, mix=(b, 10000))
        for key in ("paths", "lengths", "ends", "status"):
            assert np.array_equal(before[key], after[key]), key
        with pytest.raises(T.TrwError):  # the sample split must lie inside the sample
            dg.autotune_mix(a, b, (1 << 16) + 1, sample_walkers=1 << 16)
    finally:
        dg.close()


def test_replay_exhaustion_and_bad_sources_are_reported(gpu_ctx):
    g, src = load_golden_graph("small")
    dg = dgraph(gpu_ctx, g)
    try:
        src = np.concatenate([src[:50], np.array([-1, g.nv, 10 ** 6], np.int32)]).astype(np.int32)
        K = 11  # fewer than the draws a 20-vertex walk needs with any of these Moves (19 or more)
        for app, em, sm, extra in (("deepwalk", 1, 2, {}), ("deepwalk", 1, 1, {}), ("deepwalk", 1, 3, {}),
                                   ("node2vec", 2, 4, {"-p": "2", "-q": "0.5"})):
            case = {"app": app, "execution": em, "sampling": sm, "length": 20, "extra": extra}
            ref = oracle_walk(g, case, src, K, 1, 5, 6)
            assert np.all(ref["status"][:50] == T._abi.WALKER_REPLAY_EXHAUSTED), (app, sm)
            assert np.all(ref["status"][50:] == T._abi.WALKER_BAD_SOURCE), (app, sm)
            draws = T.replay_fill(5, len(src), K)
            for form in forms_for(app):
                r = T.walk(dg, case_params(case, form), src, draws, T.replay_fill(6, len(src), 1))
                for key in ("paths", "lengths", "ends", "status", "draws_used"):
                    assert np.array_equal(r[key], ref[key]), (app, sm, form, key)
                assert r["errors"] == len(src), (app, sm, form)
    finally:
        dg.close()


def test_forms_agree_in_free_running_mode(gpu_ctx):
    g = graphs.chung_lu(20000, 120000, 2.1, 2, 3000, 41)
    dg = dgraph(gpu_ctx, g)
    try:
        src = np.nonzero(g.degree > 0)[0].astype(np.int32)
        for app, em, sm, L in [("deepwalk", 1, 2, 40), ("deepwalk", 1, 1, 40), ("node2vec", 2, 4, 40),
                               ("ppr", 0, 0, 1), ("metapath", 2, 1, 40)]:
            outs = []
            for form in forms_for(app):
                prm = T.make_params(app, execution=em, sampling=sm, length=L, rng="philox", seed=77, form=form,
                                    schema=[0, 1, 2, 3, 4], stop_probability=0.15)
                outs.append(T.walk(dg, prm, src))
            for o in outs[1:]:
                for key in ("paths", "lengths", "ends", "status"):
                    assert np.array_equal(o[key], outs[0][key]), (app, key)
                assert o["steps"] == outs[0]["steps"]
    finally:
        dg.close()


def _chi2_pvalue(counts, probs):
    from scipy.stats import chisquare
    n = counts.sum()
    return chisquare(counts, probs * n).pvalue


@pytest.mark.parametrize("sm", ["alias", "its", "rejection", "uniform"])
def test_free_running_first_step_transitions_chi_square(gpu_ctx, sm):
    g = graphs.hub_graph(3000, 400, 6000, 51)
    dg = dgraph(gpu_ctx, g)
    try:
        hub = int(np.argmax(g.degree))
        n = 400000
        src = np.full(n, hub, np.int32)
        em = "uniform" if sm == "uniform" else "static"
        prm = T.make_params("deepwalk", execution=em, sampling=sm, length=2, rng="philox", seed=123)
        r = T.walk(dg, prm, src)
        nb = g.neighbors(hub)
        lo, hi = g.offset[hub], g.offset[hub + 1]
        w = np.ones(len(nb)) if sm == "uniform" else g.weight[lo:hi]
        pos = np.searchsorted(nb, r["paths"][:, 1])
        counts = np.bincount(pos, minlength=len(nb)).astype(np.float64)
        assert _chi2_pvalue(counts, w / w.sum()) > 1e-4
    finally:
        dg.close()


def test_free_running_ppr_matches_power_iteration(gpu_ctx):
    g = graphs.chung_lu(5000, 20000, 2.1, 2, 500, 61)
    dg = dgraph(gpu_ctx, g)
    try:
        live = np.nonzero(g.degree > 0)[0]
        sources = live[:7].astype(np.int32)
        n = 1 << 20
        src = np.resize(sources, n).astype(np.int32)
        for prm, weighted in [(T.make_params("ppr", rng="philox", seed=9, stop_probability=0.2), False),
                              (T.make_params("ppr", execution="static", sampling="alias", rng="philox", seed=9,
                                             stop_probability=0.2), False)]:
            r = T.walk(dg, prm, src, want_paths=False)
            assert np.all(r["status"] == 0)
            pi, _ = ob.ppr_power(g, 0.2, sources, weighted=weighted)
            est = np.bincount(r["ends"], minlength=g.nv) / n
            bound = np.sum(np.sqrt(pi * (1 - pi) / n))
            assert np.abs(est - pi).sum() < 1.5 * bound
            mean_len = (r["lengths"] - 1).mean()
            assert abs(mean_len - 1 / 0.2) < 0.05
    finally:
        dg.close()


# Kind pairs of a mixed batch (the kind's mode is read per walker at run time in every form).
MIX_PAIRS = [
    (("ppr", {}), ("metapath", {"schema": [0, 1, 2, 3, 4]})),
    (("ppr", {"execution": "static", "sampling": "alias"}), ("deepwalk", {"execution": "static", "sampling": "rejection"})),
    (("deepwalk", {"execution": "static", "sampling": "its"}), ("metapath", {"schema": [1, 0, 2]})),
    (("ppr", {"execution": "static", "sampling": "alias"}), ("node2vec", {"p": 2.0, "q": 0.5})),
    # The first kind's Move after the second's: the pair kernel runs walkers [0, split) as its second Move.
    (("metapath", {"schema": [0, 1, 2, 3, 4]}), ("ppr", {})),
]


@pytest.mark.parametrize("pair", MIX_PAIRS,
                         ids=["ppr+metapath", "alias+rejection", "its+metapath", "alias+node2vec", "metapath+ppr"])
def test_mixed_batch_queue_equals_step_and_single_kinds(gpu_ctx, pair):
    (app_a, kw_a), (app_b, kw_b) = pair
    g = graphs.chung_lu(30000, 200000, 2.1, 2, 4000, 71)
    dg = dgraph(gpu_ctx, g)
    try:
        live = np.nonzero(g.degree > 0)[0]
        rng = np.random.default_rng(3)
        n = 200000
        src = rng.choice(live, n).astype(np.int32)
        split = n // 2
        res = {}
        for form in ("queue", "step", "graph"):
            a = T.make_params(app_a, length=80, stop_probability=0.15, rng="philox", seed=5, form=form, **kw_a)
            b = T.make_params(app_b, length=80, stop_probability=0.15, rng="philox", seed=5, form=form, **kw_b)
            res[form] = T.walk(dg, a, src, mix=(b, split))
        for key in ("paths", "lengths", "ends", "status"):
            assert np.array_equal(res["queue"][key], res["step"][key]), key
            assert np.array_equal(res["queue"][key], res["graph"][key]), key
        assert res["queue"]["steps"] == res["step"]["steps"] == res["graph"]["steps"]
        # Each kind alone (keyed by the same global walker ids) gives the same walks as in the mix.
        a = T.make_params(app_a, length=80, stop_probability=0.15, rng="philox", seed=5, **kw_a)
        b = T.make_params(app_b, length=80, stop_probability=0.15, rng="philox", seed=5, **kw_b)
        ra = T.walk(dg, a, src[:split])
        rb = T.walk(dg, b, src[split:], walker_id_base=split)
        for key in ("paths", "lengths", "ends", "status"):
            assert np.array_equal(res["queue"][key][:split], ra[key]), key
            assert np.array_equal(res["queue"][key][split:], rb[key]), key
        assert res["queue"]["steps"] == ra["steps"] + rb["steps"]
    finally:
        dg.close()


# Kind pairs with rows of their own lengths: PPR's length-1 rows beside MetaPath's 80 (the bench's
# mix), rows at most one 16-position chunk long beside an odd length, and lengths either side of a
# chunk boundary.  Odd splits leave the second kind's rows off 16-byte alignment.
KIND_ROWS = [
    (("ppr", {"length": 1}), ("metapath", {"length": 80, "schema": [0, 1, 2, 3, 4]})),
    (("deepwalk", {"execution": "static", "sampling": "its", "length": 12}),
     ("metapath", {"length": 37, "schema": [1, 0, 2]})),
    (("ppr", {"length": 16}), ("metapath", {"length": 17, "schema": [0, 1, 2, 3, 4]})),
    (("metapath", {"length": 80, "schema": [0, 1, 2, 3, 4]}), ("ppr", {"length": 1})),
]


@pytest.mark.parametrize("unpadded", [False, True], ids=["padded", "unpadded"])
@pytest.mark.parametrize("pair", KIND_ROWS,
                         ids=["ppr1+metapath80", "its12+metapath37", "ppr16+metapath17", "metapath80+ppr1"])
def test_mixed_batch_rows_per_kind(gpu_ctx, pair, unpadded):
    """io.path_stride_b: each kind's rows as long as its walks.  Every form writes the same rows
    as with one stride for both kinds, padded or not."""
    (app_a, kw_a), (app_b, kw_b) = pair
    g = graphs.chung_lu(20000, 120000, 2.1, 2, 3000, 73)
    dg = dgraph(gpu_ctx, g)
    try:
        live = np.nonzero(g.degree > 0)[0]
        n = 100003
        src = np.random.default_rng(8).choice(live, n).astype(np.int32)
        split = 50001
        La, Lb = kw_a["length"], kw_b["length"]
        first = None
        for form in ("queue", "step", "graph"):
            a = T.make_params(app_a, stop_probability=0.15, rng="philox", seed=6, form=form, **kw_a)
            b = T.make_params(app_b, stop_probability=0.15, rng="philox", seed=6, form=form, **kw_b)
            one = T.walk(dg, a, src, mix=(b, split), unpadded=unpadded)
            per = T.walk(dg, a, src, mix=(b, split), unpadded=unpadded, kind_rows=True)
            assert np.array_equal(per["paths_a"], one["paths"][:split, :La]), form
            assert np.array_equal(per["paths_b"], one["paths"][split:, :Lb]), form
            for key in ("lengths", "ends", "status"):
                assert np.array_equal(per[key], one[key]), (form, key)
            if not unpadded:
                # Padded rows: the walk, then -1 to the kind's length.
                for rows, lens in ((per["paths_a"], per["lengths"][:split]), (per["paths_b"], per["lengths"][split:])):
                    pos = np.arange(rows.shape[1])[None, :]
                    assert np.all((rows >= 0) == (pos < lens[:, None])), form
            if first is None:
                first = per
            else:
                for key in ("paths_a", "paths_b", "lengths", "ends", "status"):
                    assert np.array_equal(per[key], first[key]), (form, key)
    finally:
        dg.close()


def test_mixed_batch_row_strides_are_checked(gpu_ctx):
    """Each kind's rows must hold its walks: path_stride >= the first kind's length and
    path_stride_b >= the second's (or, with path_stride_b 0, path_stride >= both)."""
    g = graphs.chung_lu(2000, 8000, 2.1, 2, 300, 75)
    dg = dgraph(gpu_ctx, g)
    try:
        src = np.nonzero(g.degree > 0)[0][:100].astype(np.int32)
        n, split = len(src), 50
        a = T.make_params("ppr", length=1, stop_probability=0.2, rng="philox", seed=1, form="queue")
        b = T.make_params("metapath", length=20, schema=[0, 1, 2], rng="philox", seed=1, form="queue")
        d_src = T.DeviceBuffer.from_array(gpu_ctx, src)
        paths = T.DeviceBuffer(gpu_ctx, n * 20 * 4)
        try:
            for stride, stride_b in ((1, 19), (1, -1), (0, 20), (19, 0)):
                io = T.make_io(d_src, n, stride, paths=paths, path_stride_b=stride_b)
                with pytest.raises(T.TrwError) as e:
                    T.walk_mix_async(dg, a, b, split, io)
                assert e.value.code == T._abi.E_INVALID, (stride, stride_b)
            for stride, stride_b in ((1, 20), (20, 0), (3, 25)):
                T.walk_mix_async(dg, a, b, split, T.make_io(d_src, n, stride, paths=paths, path_stride_b=stride_b))
            T.synchronize(gpu_ctx)
        finally:
            d_src.free()
            paths.free()
    finally:
        dg.close()


def test_mixed_kinds_on_one_static_table_from_different_weights_are_refused(gpu_ctx):
    """A graph holds one table per sampler: PPR's unit-weight alias table and DeepWalk's
    edge-weight one cannot serve one batch (the second build would replace the first under the
    first kind's walkers).  The mix is refused; the same kinds on one weight source walk as alone."""
    g = graphs.chung_lu(3000, 12000, 2.1, 2, 400, 57)
    dg = dgraph(gpu_ctx, g)
    try:
        src = np.nonzero(g.degree > 0)[0].astype(np.int32)
        a = T.make_params("ppr", execution="static", sampling="alias", stop_probability=0.2, rng="philox", seed=3,
                          form="queue")
        b = T.make_params("deepwalk", execution="static", sampling="alias", length=12, rng="philox", seed=3,
                          form="queue")
        for form in ("queue", "step", "graph"):
            a.form = b.form = T.engine._FORM[form]
            with pytest.raises(T.TrwError) as e:
                T.walk(dg, a, src, mix=(b, len(src) // 2))
            assert e.value.code == T._abi.E_UNSUPPORTED
        a_edge = T.make_params("ppr", execution="static", sampling="alias", stop_probability=0.2, rng="philox",
                               seed=3, form="queue", weights="edge")
        split = len(src) // 2
        mixed = T.walk(dg, a_edge, src, mix=(b, split))
        alone_a = T.walk(dg, a_edge, src[:split], walker_id_base=0)
        alone_b = T.walk(dg, b, src[split:], walker_id_base=split)
        assert np.array_equal(mixed["ends"][:split], alone_a["ends"])
        assert np.array_equal(mixed["paths"][split:, :12], alone_b["paths"])
    finally:
        dg.close()


def test_unpadded_rows_hold_exactly_the_walk(gpu_ctx):
    """TRW_PATHS_UNPADDED: each row's first lengths[i] entries are the padded walk's, and nothing
    past them is written (the buffer's -2 stays), in every form and in a mixed batch."""
    g = graphs.chung_lu(6000, 30000, 2.1, 2, 1500, 58)
    lbl = np.random.default_rng(3).integers(0, 5, g.ne).astype(np.int32)
    dg = T.DeviceGraph(gpu_ctx, g.degree, g.adj, g.weight, lbl)
    try:
        src = np.nonzero(g.degree > 0)[0].astype(np.int32)
        cases = [("metapath", {"schema": [0, 1, 2, 3, 4], "length": 40}, ["fused", "queue", "step", "graph"]),
                 ("ppr", {"stop_probability": 0.15}, ["queue", "step", "graph"]),
                 ("node2vec", {"length": 30}, ["fused", "queue", "step"]),
                 ("deepwalk", {"length": 30}, ["fused", "queue"])]
        for app, kw, forms in cases:
            for form in forms:
                prm = T.make_params(app, rng="philox", seed=5, form=form, **kw)
                a = T.walk(dg, prm, src)
                b = T.walk(dg, prm, src, unpadded=True)
                L = a["paths"].shape[1]
                rec = np.minimum(a["lengths"], L)
                inside = np.arange(L)[None, :] < rec[:, None]
                assert np.array_equal(a["paths"][inside], b["paths"][inside]), (app, form)
                if form != "fused":
                    assert np.all(b["paths"][~inside] == -2), (app, form)
        a_p = T.make_params("ppr", stop_probability=0.15, rng="philox", seed=6, form="queue")
        b_p = T.make_params("metapath", length=40, schema=[0, 1, 2, 3, 4], rng="philox", seed=6, form="queue")
        split = len(src) // 2
        a = T.walk(dg, a_p, src, mix=(b_p, split))
        b = T.walk(dg, a_p, src, mix=(b_p, split), unpadded=True)
        rec = np.where(np.arange(len(src)) < split, 1, np.minimum(a["lengths"], 40))
        inside = np.arange(40)[None, :] < rec[:, None]
        assert np.array_equal(a["paths"][inside], b["paths"][inside])
        assert np.all(b["paths"][~inside] == -2)
    finally:
        dg.close()


def test_walker_ids_key_the_free_running_stream(gpu_ctx):
    """A walker's Philox stream depends only on its global id: a permuted batch with
    per-walker ids gives each walker the walk it has in the ordered batch."""
    g = graphs.chung_lu(20000, 120000, 2.1, 2, 3000, 43)
    dg = dgraph(gpu_ctx, g)
    try:
        live = np.nonzero(g.degree > 0)[0]
        rng = np.random.default_rng(8)
        n = 50000
        src = rng.choice(live, n).astype(np.int32)
        perm = rng.permutation(n)
        base = 1 << 40
        for app, sm, form in [("deepwalk", "alias", "fused"), ("ppr", "alias", "queue"), ("node2vec", "max_weight_rejection", "queue")]:
            prm = T.make_params(app, execution="static" if app != "node2vec" else "dynamic", sampling=sm, length=30,
                                rng="philox", seed=21, form=form, stop_probability=0.2)
            ordered = T.walk(dg, prm, src, walker_id_base=base)
            shuffled = T.walk(dg, prm, src[perm], walker_ids=base + perm.astype(np.uint64))
            for key in ("paths", "lengths", "ends", "status"):
                assert np.array_equal(shuffled[key], ordered[key][perm]), (app, key)
    finally:
        dg.close()


def test_weights_selector_overrides_the_app_weights(gpu_ctx):
    """PPR samples unit weights by default (its functor's weight() is 1.0); weights="edge"
    makes the same walk sample the edge weights.  Both equal the oracle's replay walks over
    alias tables built from those weights."""
    g = graphs.chung_lu(5000, 30000, 2.1, 2, 600, 44)
    dg = dgraph(gpu_ctx, g)
    try:
        src = np.nonzero(g.degree > 0)[0].astype(np.int32)
        K, KT = 256, 128
        draws, tdraws = T.replay_fill(31, len(src), K), T.replay_fill(32, len(src), KT)
        d, td = draws.T.copy(), tdraws.T.copy()
        out = {}
        for sel, w in [("app", np.ones(g.ne)), ("edge", np.asarray(g.weight, np.float64))]:
            prm = T.make_params("ppr", execution="static", sampling="alias", rng="replay", weights=sel,
                                stop_probability=0.2)
            r = T.walk(dg, prm, src, draws, tdraws)
            alias, _ = ob.alias_build(g, w)
            ref = ob.walk(ob.Graph(g, alias), "ppr", 1, 2, 1, src, d, K, td, KT, stop_prob=0.2)
            for key in ("lengths", "ends", "status", "draws_used"):
                assert np.array_equal(r[key], ref[key]), (sel, key)
            assert np.all(r["status"] == 0)
            out[sel] = r["ends"]
        assert not np.array_equal(out["app"], out["edge"])
    finally:
        dg.close()


def test_full_size_deepwalk_alias_graph(gpu_ctx):
    """2^22-vertex Chung-Lu graph: device tables equal the oracle's, replay paths of a
    walker subset equal the oracle's, and every free-running step follows an edge."""
    h = T.HostGraph.chung_lu(1 << 22, 8 << 22, 2.1, 2, 65536, 2, True, 0)
    deg, adj, w, _ = h.arrays()
    g = graphs.CSR(deg, adj, w)
    dg = T.DeviceGraph(gpu_ctx, deg, adj, w)
    try:
        dg.build_tables("alias", "edge")
        table = dg.alias_table()
        ref_table, _ = ob.alias_build(g, w)
        assert np.array_equal(table.view(np.uint8), ref_table.view(np.uint8))
        # The ITS and rejection tables too (hubs of degree >= 1024 take the wavefront builds).
        dg.build_tables("its", "edge")
        assert np.array_equal(dg.prefix_sum().view(np.uint64), ob.its_build(g, w).view(np.uint64))
        dg.build_tables("rejection", "edge")
        assert np.array_equal(dg.rejection_max().view(np.uint64), ob.rs_build(g, w).view(np.uint64))
        dg.build_tables("alias", "unit")
        ref_unit, _ = ob.alias_build(g, np.ones(len(adj)))
        assert np.array_equal(dg.alias_table().view(np.uint8), ref_unit.view(np.uint8))
        dg.build_tables("alias", "edge")

        src = T.sources_all(deg)
        sub = src[:: max(1, len(src) // 65536)][:65536].astype(np.int32)
        case = {"app": "deepwalk", "execution": 1, "sampling": 2, "length": 80, "extra": {}}
        K = 158
        og = ob.Graph(g, ref_table)
        d = ob.sfmt_stream(1234, len(sub) * K).reshape(len(sub), K)
        ref = ob.walk(og, "deepwalk", 1, 2, 80, sub, d, K)
        r = T.walk(dg, case_params(case, "fused"), sub, T.replay_fill(1234, len(sub), K))
        for key in ("paths", "lengths", "ends", "status", "draws_used"):
            assert np.array_equal(r[key], ref[key]), key

        prm = T.make_params("deepwalk", rng="philox", seed=3)
        r = T.walk(dg, prm, src)
        assert np.all(r["lengths"] == 80) and np.all(r["status"] == 0)
        assert r["steps"] == 79 * len(src)
        rows = r["paths"][::97]
        assert np.array_equal(rows[:, 0], src[::97])
        nv = np.int64(len(deg))
        edge_keys = np.repeat(np.arange(nv, dtype=np.int64), deg) * nv + adj  # sorted: rows are sorted
        q = rows[:, :-1].ravel().astype(np.int64) * nv + rows[:, 1:].ravel()
        pos = np.minimum(np.searchsorted(edge_keys, q), len(edge_keys) - 1)
        assert np.all(edge_keys[pos] == q)
    finally:
        dg.close()


# (Move, execution, sampling, every how many of the walkers, draws per walker)
FULL_SIZE_MOVES = [("alias", 1, 2, 1, 158), ("its", 1, 1, 1, 158), ("uniform", 0, 0, 1, 79), ("rejection", 1, 3, 4, 400)]


@pytest.mark.parametrize("sm,execution,sampling,every,K", FULL_SIZE_MOVES, ids=[m[0] for m in FULL_SIZE_MOVES])
def test_full_size_replay_of_every_walker_is_bit_exact(gpu_ctx, sm, execution, sampling, every, K):
    """The DeepWalk configuration at full size, as the headline runs it: the 2^22-vertex Chung-Lu graph
    (exponent 2.1, average degree 16, dmin 2, dmax 65536, FP64 U[1,5) weights), one walker per
    non-isolated vertex (4,077,405), length 80, replay draws from the SFMT stream; static alias, static
    ITS (guide records) and uniform on every walker, static rejection (a trial per two draws) on every
    4th: every walker's path, length, end, status and draws used equal the oracle's (run over walker
    ranges on the host's cores)."""
    from concurrent.futures import ThreadPoolExecutor

    h = T.HostGraph.chung_lu(1 << 22, 8 << 22, 2.1, 2, 65536, 2, True, 0)
    deg, adj, w, _ = h.arrays()
    g = graphs.CSR(deg.copy(), adj.copy(), w.copy())
    h.close()
    dg = T.DeviceGraph(gpu_ctx, g.degree, g.adj, g.weight)
    try:
        if sm != "uniform":
            dg.build_tables(sm, "edge")
        og = (ob.Graph(g, ob.alias_build(g, g.weight)[0]) if sm == "alias"
              else ob.Graph(g, prefix=ob.its_build(g, g.weight)) if sm == "its"
              else ob.Graph(g, rs_weight=g.weight, rs_max=ob.rs_build(g, g.weight)) if sm == "rejection"
              else ob.Graph(g))
        src = T.sources_all(g.degree)[::every].copy()
        n, L = len(src), 80
        case = {"app": "deepwalk", "execution": execution, "sampling": sampling, "length": L, "extra": {}}
        r = T.walk(dg, case_params(case, "fused"), src, T.replay_fill(1234, n, K))
        print("full-size replay (%s): %d walkers, %d steps" % (sm, n, r["steps"]), flush=True)
        assert r["steps"] == n * (L - 1)
        d = ob.sfmt_stream(1234, n * K).reshape(n, K)
        cuts = np.linspace(0, n, 33).astype(np.int64)

        def part(i):
            a, b = int(cuts[i]), int(cuts[i + 1])
            return ob.walk(og, "deepwalk", execution, sampling, L, src[a:b], d[a:b], K)

        with ThreadPoolExecutor(16) as pool:
            parts = list(pool.map(part, range(len(cuts) - 1)))
        for key in ("paths", "lengths", "ends", "status", "draws_used"):
            assert np.array_equal(r[key], np.concatenate([p[key] for p in parts])), key
    finally:
        dg.close()


@pytest.mark.parametrize("weighted", [False, True])
def test_full_size_node2vec_replay_is_bit_exact(gpu_ctx, weighted):
    """The Node2Vec configuration at full size: the 2^22-vertex Chung-Lu graph (exponent 2.0, average
    degree 32, dmin 2, dmax 65536, FP64 U[1,5) weights; hubs of degree up to 57,998 take the hashed
    membership tables), p = 2, q = 0.5, length 80, max-weight rejection as the reference's functor
    (weights ignored) and over the edge weights; every 8th of the one-walker-per-vertex sources
    (519,518 walkers), replay draws (1,600 per walker, 2,400 over the weights: a walker that runs out
    ends as exhausted in both): paths, lengths, ends, statuses and draws used equal the oracle's."""
    from concurrent.futures import ThreadPoolExecutor

    h = T.HostGraph.chung_lu(1 << 22, 16 << 22, 2.0, 2.0, 65536.0, 3, True, 0)
    deg, adj, w, _ = h.arrays()
    g = graphs.CSR(deg.copy(), adj.copy(), w.copy())
    h.close()
    dg = T.DeviceGraph(gpu_ctx, g.degree, g.adj, g.weight)
    try:
        src = T.sources_all(g.degree)[::8].copy()
        n, L, K = len(src), 80, 2400 if weighted else 1600
        prm = T.make_params("node2vec", length=L, p=2.0, q=0.5, rng="replay", weights="edge" if weighted else "app")
        r = T.walk(dg, prm, src, T.replay_fill(5, n, K))
        print("full-size node2vec replay (weighted %s): %d walkers, %d steps" % (weighted, n, r["steps"]), flush=True)
        og = ob.Graph(g)
        d = ob.sfmt_stream(5, n * K).reshape(n, K)
        cuts = np.linspace(0, n, 33).astype(np.int64)

        def part(i):
            a, b = int(cuts[i]), int(cuts[i + 1])
            return ob.walk(og, "node2vec", 2, 4, L, src[a:b], d[a:b], K, p=2.0, q=0.5, weighted=weighted)

        with ThreadPoolExecutor(16) as pool:
            parts = list(pool.map(part, range(len(cuts) - 1)))
        st = np.concatenate([p["status"] for p in parts])
        print("oracle statuses:", np.bincount(st), flush=True)
        assert np.mean(st == 0) > 0.9999  # the draws cover (nearly) every walk in full
        for key in ("paths", "lengths", "ends", "status", "draws_used"):
            assert np.array_equal(r[key], np.concatenate([p[key] for p in parts])), key
        assert r["steps"] == sum(int(p["steps"]) for p in parts)
    finally:
        dg.close()


@pytest.mark.parametrize("sources", ["all", "single"])
def test_full_size_c1_ppr_replay_is_bit_exact(gpu_ctx, sources):
    """The reference's own CPU-runnable PPR configuration at full size: the 2^18-vertex Chung-Lu graph
    (exponent 2.1, dmin 2, dmax 4096, 1.5 M undirected edges, unweighted), uniform Moves, termination
    0.2, one walker per non-isolated vertex (generate_all_source) or |V| walkers from the max-degree
    vertex (the app's default single source); replay draws: every walker's end, length, status, draws
    used and source entry equal the oracle's in the QUEUE, STEP and GRAPH forms."""
    from concurrent.futures import ThreadPoolExecutor

    h = T.HostGraph.chung_lu(1 << 18, 1500000, 2.1, 2.0, 4096.0, 1, False, 0)
    deg, adj, _, _ = h.arrays()
    g = graphs.CSR(deg.copy(), adj.copy())
    h.close()
    dg = T.DeviceGraph(gpu_ctx, g.degree, g.adj)
    try:
        src = T.sources_all(g.degree) if sources == "all" else T.sources_single(g.degree, g.nv)
        n, K = len(src), 128
        draws, tdraws = T.replay_fill(51, n, K), T.replay_fill(52, n, K)
        d = ob.sfmt_stream(51, n * K).reshape(n, K)
        td = ob.sfmt_stream(52, n * K).reshape(n, K)
        og = ob.Graph(g)
        cuts = np.linspace(0, n, 17).astype(np.int64)

        def part(i):
            a, b = int(cuts[i]), int(cuts[i + 1])
            return ob.walk(og, "ppr", 0, 0, 1, src[a:b], d[a:b], K, td[a:b], K, stop_prob=0.2)

        with ThreadPoolExecutor(16) as pool:
            parts = list(pool.map(part, range(len(cuts) - 1)))
        ref = {key: np.concatenate([p[key] for p in parts]) for key in ("paths", "lengths", "ends", "status",
                                                                       "draws_used")}
        assert np.all(ref["status"] == 0)
        for form in ("queue", "step", "graph"):
            prm = T.make_params("ppr", stop_probability=0.2, rng="replay", form=form)
            r = T.walk(dg, prm, src, draws, tdraws)
            for key in ref:
                assert np.array_equal(r[key], ref[key]), (form, key)
            assert r["steps"] == sum(int(p["steps"]) for p in parts), form
        print("C1 replay (%s sources): %d walkers, %d steps" % (sources, n, r["steps"]), flush=True)
    finally:
        dg.close()


def test_full_size_ppr_metapath_mix(gpu_ctx):
    """The uneven mix at full size: the 2^22-vertex headline graph with edge labels U{0..4}, 2^23
    walkers from uniform random sources, walkers [0, 2^22) PPR (termination 0.15, length 1), the rest
    MetaPath on schema 0,1,2,3,4 (length 80), each kind's rows at its own length.
      * Free-running: the QUEUE form (padded and unpadded rows), the hipGraph step form and the plain
        step form give the same walks, step for step, and each kind alone gives what it gives in the mix.
      * Replay: every 16th walker of the batch (2^19, the same split) in the QUEUE form equals the
        oracle's PPR and MetaPath walks."""
    from concurrent.futures import ThreadPoolExecutor

    h = T.HostGraph.chung_lu(1 << 22, 8 << 22, 2.1, 2.0, 65536.0, 2, True, 5)
    deg, adj, w, lab = h.arrays()
    g = graphs.CSR(deg.copy(), adj.copy(), w.copy(), lab.copy())
    h.close()
    dg = T.DeviceGraph(gpu_ctx, g.degree, g.adj, g.weight, g.label)
    schema = [0, 1, 2, 3, 4]
    try:
        n = 1 << 23
        split = n // 2
        src = T.sources_random(g.degree, n, 4)

        def kinds(form, rng="philox"):
            return (T.make_params("ppr", length=1, stop_probability=0.15, rng=rng, seed=4, form=form),
                    T.make_params("metapath", length=80, schema=schema, rng=rng, seed=4, form=form))

        res = {}
        for form, unpadded in (("queue", False), ("queue", True), ("graph", False), ("step", False)):
            a, b = kinds(form)
            res[(form, unpadded)] = T.walk(dg, a, src, mix=(b, split), kind_rows=True, unpadded=unpadded)
            print("full-size mix, %s%s: %d steps" % (form, " unpadded" if unpadded else "",
                                                    res[(form, unpadded)]["steps"]), flush=True)
        ref = res[("queue", False)]
        for k, r in res.items():
            for key in ("lengths", "ends", "status", "paths_a"):
                assert np.array_equal(r[key], ref[key]), (k, key)
            assert r["steps"] == ref["steps"], k
            pb = r["paths_b"]
            if k[1]:  # unpadded: the entries of each walk, the rest untouched
                inside = np.arange(80)[None, :] < ref["lengths"][split:, None]
                assert np.array_equal(np.where(inside, pb, -1), ref["paths_b"]), k
            else:
                assert np.array_equal(pb, ref["paths_b"]), k
        a, b = kinds("queue")
        ra = T.walk(dg, a, src[:split])
        rb = T.walk(dg, b, src[split:], walker_id_base=split)
        assert np.array_equal(ra["paths"][:, :1], ref["paths_a"]) and np.array_equal(rb["paths"], ref["paths_b"])
        for key in ("lengths", "ends", "status"):
            assert np.array_equal(np.concatenate([ra[key], rb[key]]), ref[key]), key
        del res, ra, rb

        sub = src[::16].copy()
        m, ms = len(sub), len(sub) // 2
        K = KT = 160
        a, b = kinds("queue", "replay")
        r = T.walk(dg, a, sub, T.replay_fill(41, m, K), T.replay_fill(42, m, KT), mix=(b, ms), kind_rows=True)
        d = ob.sfmt_stream(41, m * K).reshape(m, K)
        td = ob.sfmt_stream(42, m * KT).reshape(m, KT)
        og = ob.Graph(g)
        cuts = np.linspace(0, m, 33).astype(np.int64)

        def part(i):
            lo, hi = int(cuts[i]), int(cuts[i + 1])
            out = []
            if lo < ms:
                e = min(hi, ms)
                out.append(ob.walk(og, "ppr", 0, 0, 1, sub[lo:e], d[lo:e], K, td[lo:e], KT, stop_prob=0.15))
            if hi > ms:
                s0 = max(lo, ms)
                out.append(ob.walk(og, "metapath", 2, 1, 80, sub[s0:hi], d[s0:hi], K, td[s0:hi], KT, schema=schema))
            return out

        with ThreadPoolExecutor(16) as pool:
            parts = [x for p in pool.map(part, range(len(cuts) - 1)) for x in p]
        oracle = {key: np.concatenate([p[key] for p in parts]) for key in ("lengths", "ends", "status", "draws_used")}
        print("full-size mix replay: %d walkers, oracle statuses %s" % (m, np.bincount(oracle["status"])), flush=True)
        for key in oracle:
            assert np.array_equal(r[key], oracle[key]), key
        pa = np.concatenate([p["paths"] for p in parts if p["paths"].shape[1] == 1])
        pb = np.concatenate([p["paths"] for p in parts if p["paths"].shape[1] == 80])
        assert np.array_equal(r["paths_a"], pa) and np.array_equal(r["paths_b"], pb)
        assert r["steps"] == sum(int(p["steps"]) for p in parts)
    finally:
        dg.close()


def test_upload_rejects_vertex_ids_out_of_range(gpu_ctx):
    g = graphs.chung_lu(3000, 20000, 2.1, 2, 300, 17)
    for bad in (-1, g.nv):
        adj = g.adj.copy()
        adj[len(adj) * 2 // 3] = bad
        with pytest.raises(T.TrwError) as e:
            T.DeviceGraph(gpu_ctx, g.degree, adj, g.weight)
        assert e.value.code == T._abi.E_INVALID
    T.DeviceGraph(gpu_ctx, g.degree, g.adj, g.weight).close()


def test_fused_path_tiles_for_every_walk_length(gpu_ctx):
    """The fused kernels stage whole path rows when they fit their LDS budget and chunks of a
    multiple of 32 (or 4) positions otherwise, per Move and walkers per lane: every shape --
    rows shorter than, equal to and longer than a tile, a partial last chunk, a batch that is
    not a whole number of blocks, a row stride wider than the walk -- writes the oracle's paths."""
    g = graphs.chung_lu(3000, 15000, 2.1, 2, 900, 81)
    dg = dgraph(gpu_ctx, g)
    try:
        src = np.nonzero(g.degree > 0)[0].astype(np.int32)[:2777]
        n = len(src)
        cases = [("deepwalk", 1, 2, {}, [1, 2, 4]), ("deepwalk", 0, 0, {}, [1, 4]), ("deepwalk", 1, 1, {}, [0]),
                 ("deepwalk", 1, 3, {}, [0]), ("node2vec", 2, 4, {"-p": "2", "-q": "0.5"}, [0])]
        for L in (2, 3, 31, 33, 64, 127, 128, 129, 300):
            for app, em, sm, extra, wpts in cases:
                case = {"app": app, "execution": em, "sampling": sm, "length": L, "extra": extra}
                K = 32 * L if sm in (3, 4) else 2 * L
                ref = oracle_walk(g, case, src, K, 1, 82, 83)
                draws = T.replay_fill(82, n, K)
                for wpt in wpts:
                    a = golden_case_args(case)
                    prm = T.make_params(app, execution=em, sampling=sm, length=L, p=a["p"], q=a["q"], rng="replay",
                                        form="fused", walkers_per_thread=wpt)
                    r = T.walk(dg, prm, src, draws)
                    tag = (app, em, sm, L, wpt)
                    assert np.array_equal(r["paths"], ref["paths"]), tag
                    assert np.array_equal(r["lengths"], ref["lengths"]), tag
                    assert r["steps"] == ref["steps"], tag
        # A row stride wider than the walk: rows are no longer contiguous; the gap stays untouched.
        L, stride = 80, 83
        case = {"app": "deepwalk", "execution": 1, "sampling": 2, "length": L, "extra": {}}
        ref = oracle_walk(g, case, src, 2 * L, 1, 84, 85)
        sb = T.DeviceBuffer.from_array(gpu_ctx, src)
        db = T.DeviceBuffer.from_array(gpu_ctx, T.replay_fill(84, n, 2 * L))
        pb = T.DeviceBuffer.filled(gpu_ctx, n * stride, np.int32, -7)
        try:
            io = T.make_io(sb, n, stride, 0, db, 2 * L, None, 0, pb)
            T.walk_async(dg, T.make_params("deepwalk", length=L, rng="replay", form="fused"), io)
            T.synchronize(gpu_ctx)
            rows = pb.to_numpy(np.int32, (n, stride))
        finally:
            for b in (sb, db, pb):
                b.free()
        assert np.array_equal(rows[:, :L], ref["paths"])
        assert np.all(rows[:, L:] == -7)
    finally:
        dg.close()


def test_node2vec_membership_structures_match_oracle(gpu_ctx):
    """Node2Vec's is_neighbor tests run through hashed membership tables (128-byte buckets) for
    adjacencies longer than 32 (8, 256 and 4096 buckets here: hubs of degree 100, 3000 and 40000)
    and by binary search below; on an undirected and on a directed graph (where which endpoint's
    list is searched decides the answer) the walks equal the oracle's."""
    rng = np.random.default_rng(91)
    n = 60000
    parts = [rng.integers(0, n, (150000, 2))]
    for hub, deg in [(0, 40000)] + [(h, 3000) for h in range(1, 6)] + [(h, 100) for h in range(6, 40)]:
        nb = rng.choice(np.arange(40, n), size=deg, replace=False)
        parts.append(np.stack([np.full(deg, hub), nb], 1))
    parts.append(np.array([(a, b) for a in range(40) for b in range(a + 1, 40) if rng.random() < 0.5]))
    g = graphs.csr_from_undirected(n, np.concatenate(parts), rng)
    d = directed_graph(3000, 93)
    for gg in (g, d):
        dg = dgraph(gpu_ctx, gg)
        try:
            live = np.nonzero(gg.degree > 0)[0].astype(np.int32)
            top = np.argsort(gg.degree)[::-1][:8].astype(np.int32)
            src = np.concatenate([rng.choice(live, 6000), np.repeat(top, 64)]).astype(np.int32)
            for p, q in ((2.0, 0.5), (0.5, 2.0), (1.0, 4.0)):
                case = {"app": "node2vec", "execution": 2, "sampling": 4, "length": 24,
                        "extra": {"-p": str(p), "-q": str(q)}}
                K = 64 * 24
                ref = oracle_walk(gg, case, src, K, 1, 94, 95)
                draws = T.replay_fill(94, len(src), K)
                for form in ("fused", "queue"):
                    r = T.walk(dg, case_params(case, form), src, draws)
                    for key in ("paths", "lengths", "status", "draws_used"):
                        assert np.array_equal(r[key], ref[key]), (p, q, form, key)
        finally:
            dg.close()


def test_tables_from_caller_weights_equal_the_edge_weight_tables(gpu_ctx):
    """trw_build_tables_from builds the static tables from a device array of per-edge weights
    (the functor engine's static_gather output): from the edge weights themselves it gives the
    tables trw_build_tables(EDGE) gives, byte for byte, tagged CUSTOM in the device view."""
    import ctypes
    from thunderrw_amd import _abi as A

    g, _ = load_golden_graph("hub")
    dg = dgraph(gpu_ctx, g)
    wbuf = T.DeviceBuffer.from_array(gpu_ctx, np.ascontiguousarray(g.weight, dtype=np.float64))
    try:
        ref = {}
        for sm in ("alias", "its", "rejection"):
            dg.build_tables(sm, "edge")
        ref = (dg.alias_table().copy(), dg.prefix_sum().copy(), dg.rejection_max().copy())
        for code in (A.SM_ALIAS, A.SM_ITS, A.SM_REJECTION):
            A.check(T.lib().trw_build_tables_from(dg.h, code, ctypes.c_void_p(wbuf.data_ptr())))
        assert np.array_equal(dg.alias_table().view(np.uint8), ref[0].view(np.uint8))
        assert np.array_equal(dg.prefix_sum().view(np.uint64), ref[1].view(np.uint64))
        assert np.array_equal(dg.rejection_max(), ref[2])
        view = A.DeviceView()
        A.check(T.lib().trw_graph_device_view(dg.h, ctypes.byref(view)))
        assert (view.num_vertices, view.num_edges) == (g.nv, g.ne)
        assert view.max_degree == int(g.degree.max())
        assert (view.alias_weights, view.its_weights, view.rejection_weights) == (A.WEIGHTS_CUSTOM,) * 3
        for f in ("offset", "degree", "adj", "edge_weight", "edge_label", "alias_table", "prefix_sum",
                  "rejection_max", "rejection_weight"):
            assert getattr(view, f), f
        # The rejection weights are the library's own copy, not the caller's buffer.
        assert view.rejection_weight != wbuf.data_ptr()
        # A built-in walk that needs edge-weight tables rebuilds them: the walk equals the oracle's.
        case = {"app": "deepwalk", "execution": 1, "sampling": 2, "length": 12, "extra": {}}
        src = np.nonzero(g.degree > 0)[0].astype(np.int32)
        refw = oracle_walk(g, case, src, 24, 1, 96, 97)
        r = T.walk(dg, case_params(case, "fused"), src, T.replay_fill(96, len(src), 24))
        assert np.array_equal(r["paths"], refw["paths"])
    finally:
        wbuf.free()
        dg.close()


def _its_change_points(P, total):
    """For each i in [1, d): the least u in [0, 2^32] with P[i - 1] <= real2(u) * total, the draw
    at which the ITS pick (upper bound of real2(u) * total) moves past edge i - 1; bisection on
    the Move's own expression (2^32: never)."""
    lo = np.zeros(len(P) - 1, np.int64)  # predicate false below lo unless it holds at 0
    hi = np.full(len(P) - 1, 1 << 32, np.int64)
    key = P[:-1]
    t0 = (np.float64(0.0) * (1.0 / 4294967296.0)) * total
    at0 = key <= t0
    while True:
        open_ = (hi - lo > 1) & ~at0
        if not open_.any():
            break
        mid = (lo + hi) // 2
        ok = key <= (mid.astype(np.float64) * (1.0 / 4294967296.0)) * total
        hi = np.where(open_ & ok, mid, hi)
        lo = np.where(open_ & ~ok, mid, lo)
    return np.where(at0, 0, hi)


def _bucket_edges(m):
    """First and last draw of each of m buckets floor(u * m / 2^32)."""
    k = np.arange(m + 1, dtype=object)
    first = np.array([((int(x) << 32) + m - 1) // m for x in k], dtype=np.int64)
    return np.concatenate([first[:-1], first[1:] - 1])


def test_its_guide_records_decide_every_pick_boundary(gpu_ctx):
    """Static ITS in the fused form walks one gather per step through guide records (buckets
    of the draw space whose picks span at most two edges, with the draw threshold between
    them).  Replay draws placed at every pick change point (u - 1, u, u + 1) and at the edges
    of the buckets at one to four buckets per edge, on vertices whose weights span six decades
    (so some buckets span many edges and take the prefix-sum search), unit-range weights, a
    hub, a degree-one vertex and a vertex whose weights are all zero: every pick equals the
    oracle's upper bound, in the fused and queue forms, and with the guide switched off (the
    search structures) too."""
    rng = np.random.default_rng(77)
    g = graphs.hub_graph(3000, 2500, 6000, 78)
    w = rng.uniform(1.0, 5.0, g.ne)
    skew = np.repeat(rng.random(g.nv) < 0.5, g.degree)
    w[skew] = 10.0 ** rng.uniform(-4.0, 2.0, int(skew.sum()))
    live = np.nonzero(g.degree > 0)[0]
    zero_v = int(live[np.argmin(g.degree[live] + (g.degree[live] < 3) * 10**6)])  # a small vertex of degree >= 3
    w[g.offset[zero_v]:g.offset[zero_v + 1]] = 0.0
    gg = graphs.CSR(g.degree, g.adj, w, g.label)
    dg = dgraph(gpu_ctx, gg)
    try:
        dg.build_tables("its", "edge")
        P_all = dg.prefix_sum()
        hub = int(np.argmax(g.degree))
        one = live[g.degree[live] == 1]
        pick = list(rng.choice(live, 120, replace=False)) + [hub, zero_v] + list(one[:3])
        src, draws = [], []
        for v in dict.fromkeys(int(x) for x in pick):
            d = int(g.degree[v])
            P = P_all[g.offset[v]:g.offset[v + 1]]
            total = P[-1]
            u = [0, 1, (1 << 32) - 1, 1 << 31]
            if total > 0 and d > 1:
                c = _its_change_points(P, total)
                c = c[c < (1 << 32)]
                u += list(c - 1) + list(c) + list(c + 1)
            for c in (1, 2, 3, 4):
                u += list(_bucket_edges(c * d))
            u = np.unique(np.clip(np.array(u, np.int64), 0, (1 << 32) - 1)).astype(np.uint32)
            src.append(np.full(len(u), v, np.int32))
            draws.append(u)
        src = np.concatenate(src)
        draws = np.concatenate(draws)
        og = ob.Graph(gg, None, ob.its_build(gg, w), w, None)
        ref = ob.walk(og, "deepwalk", 1, 1, 2, src, draws.reshape(-1, 1), 1)
        assert (ref["status"] == 1).sum() > 0  # the zero-weight vertex ends its walkers without a draw
        for off in (0, 1):
            dg.set_tuning(its_guide_off=off)
            for form in ("fused", "queue"):
                r = T.walk(dg, T.make_params("deepwalk", execution="static", sampling="its", length=2,
                                             rng="replay", form=form), src, draws.reshape(1, -1))
                for key in ("paths", "lengths", "ends", "status", "draws_used"):
                    assert np.array_equal(r[key], ref[key]), (off, form, key)
                assert r["steps"] == ref["steps"], (off, form)
    finally:
        dg.close()


@pytest.mark.parametrize("seed", [41, 42])
def test_its_guide_walks_match_oracle_on_skewed_weights(gpu_ctx, seed):
    """Whole ITS walks (length 40) where a third of the vertices have weights spanning six
    decades: replay mode against the oracle in every form, and the Philox stream identical through
    the guide records and through the search structures."""
    rng = np.random.default_rng(seed)
    g = graphs.chung_lu(5000, 25000, 2.0, 2, 2000, seed)
    w = g.weight.copy()
    skew = np.repeat(rng.random(g.nv) < 0.33, g.degree)
    w[skew] = 10.0 ** rng.uniform(-4.0, 2.0, int(skew.sum()))
    gg = graphs.CSR(g.degree, g.adj, w, g.label)
    dg = dgraph(gpu_ctx, gg)
    try:
        live = np.nonzero(gg.degree > 0)[0].astype(np.int32)
        src = np.concatenate([live, rng.choice(live, 2000)]).astype(np.int32)
        case = {"app": "deepwalk", "execution": 1, "sampling": 1, "length": 40, "extra": {}}
        K = 60
        ref = oracle_walk(gg, case, src, K, 1, seed, seed + 7)
        draws = T.replay_fill(seed, len(src), K)
        for form in forms_for("deepwalk"):
            r = T.walk(dg, case_params(case, form), src, draws, T.replay_fill(seed + 7, len(src), 1))
            for key in ("paths", "lengths", "ends", "status", "draws_used"):
                assert np.array_equal(r[key], ref[key]), (form, key)
            assert r["steps"] == ref["steps"], form
        outs = []
        for off, form in ((0, "fused"), (1, "queue")):
            dg.set_tuning(its_guide_off=off)
            outs.append(T.walk(dg, T.make_params("deepwalk", execution="static", sampling="its", length=40,
                                                 rng="philox", seed=seed, form=form), src))
        for key in ("paths", "lengths", "status"):
            assert np.array_equal(outs[0][key], outs[1][key]), key
    finally:
        dg.close()


def test_rejection_records_walk_like_the_oracle(gpu_ctx, tmp_path):
    """Static rejection reads one 32-byte record per trial (the edge's weight beside the
    neighbour's range and maximum weight) and carries the maximum into the next step.  Against
    the oracle in replay mode, in every form: weights spanning four decades (many rejected
    trials), vertices whose weights are all zero (the walk ends without a draw), unit weights,
    a trial cap, and a table loaded from the reference's file instead of built."""
    rng = np.random.default_rng(91)
    g = graphs.chung_lu(4000, 20000, 2.0, 2, 1500, 92)
    w = 10.0 ** rng.uniform(-2.0, 2.0, g.ne)
    live = np.nonzero(g.degree > 0)[0]
    for v in rng.choice(live, 40, replace=False):
        w[g.offset[v]:g.offset[v + 1]] = 0.0
    gg = graphs.CSR(g.degree, g.adj, w, g.label)
    src = np.concatenate([live, rng.choice(live, 1500)]).astype(np.int32)
    L, K = 25, 2500
    draws_wm = ob.sfmt_stream(93, len(src) * K).reshape(len(src), K)
    draws = np.ascontiguousarray(draws_wm.T)
    dg = dgraph(gpu_ctx, gg)
    try:
        for weights, cap in (("edge", 0), ("unit", 0), ("edge", 3)):
            ww = w if weights == "edge" else np.ones(gg.ne)
            og = ob.Graph(gg, None, None, ww, ob.rs_build(gg, ww))
            ref = ob.walk(og, "deepwalk", 1, 3, L, src, draws_wm, K, max_trials=cap)
            if weights == "edge":
                assert (ref["status"] == T._abi.WALKER_GATHER_TERMINATED).sum() > 0
            if cap:
                assert (ref["status"] == T._abi.WALKER_TRIAL_CAP).sum() > 0
            for form in forms_for("deepwalk"):
                prm = T.make_params("deepwalk", execution="static", sampling="rejection", length=L, rng="replay",
                                    form=form, max_trials=cap, weights=weights)
                r = T.walk(dg, prm, src, draws)
                tag = (weights, cap, form)
                for key in ("paths", "lengths", "ends", "status", "draws_used"):
                    assert np.array_equal(r[key], ref[key]), (tag, key)
                assert r["steps"] == ref["steps"], tag
        dg.build_tables("rejection", "edge")
        dg.store_tables(str(tmp_path))
        og = ob.Graph(gg, None, None, w, ob.rs_build(gg, w))
        ref = ob.walk(og, "deepwalk", 1, 3, L, src, draws_wm, K)
    finally:
        dg.close()
    dg2 = dgraph(gpu_ctx, gg)
    try:
        dg2.load_tables(str(tmp_path), "rejection", "edge")
        r = T.walk(dg2, T.make_params("deepwalk", execution="static", sampling="rejection", length=L, rng="replay",
                                      form="fused"), src, draws)
        for key in ("paths", "lengths", "ends", "status", "draws_used"):
            assert np.array_equal(r[key], ref[key]), ("loaded", key)
    finally:
        dg2.close()


def _member_bucket(u, bits):
    return ((u * 0x9E3779B1) & 0xFFFFFFFF) >> (32 - bits)


def test_node2vec_membership_bucket_overflow_matches_oracle(gpu_ctx):
    """A vertex of degree 60 (four buckets of 32 slots) whose neighbours all hash to its first
    bucket: 32 of them fill it and 28 overflow into the next.  Walks from it and around a clique
    of hubs test hubs that are its neighbours (in either bucket) and hubs that are not against its
    table (it is the smaller endpoint), in replay mode against the oracle."""
    rng = np.random.default_rng(101)
    n = 20000
    same = [u for u in range(1, n) if _member_bucket(u, 2) == 0]
    other = [u for u in range(1, n) if _member_bucket(u, 2) != 0]
    nb0 = np.array(same[:60])
    hubs = np.concatenate([nb0[:6], np.array(other[:6])])
    parts = [np.stack([np.zeros(60, np.int64), nb0], 1)]
    for h in hubs:
        far = rng.choice(np.arange(1, n), 300, replace=False)
        parts.append(np.stack([np.full(300, h), far], 1))
    parts.append(np.array([(a, b) for i, a in enumerate(hubs) for b in hubs[i + 1:]]))
    bg = rng.integers(1, n, (40000, 2))
    parts.append(bg)
    g = graphs.csr_from_undirected(n, np.concatenate(parts), rng)
    assert g.degree[0] == 60
    dg = dgraph(gpu_ctx, g)
    try:
        src = np.concatenate([np.zeros(20000, np.int32), np.repeat(hubs.astype(np.int32), 500)])
        for p, q in ((2.0, 0.5), (1.0, 4.0)):
            case = {"app": "node2vec", "execution": 2, "sampling": 4, "length": 12,
                    "extra": {"-p": str(p), "-q": str(q)}}
            K = 64 * 12
            ref = oracle_walk(g, case, src, K, 1, 102, 103)
            draws = T.replay_fill(102, len(src), K)
            for form in ("fused", "queue"):
                r = T.walk(dg, case_params(case, form), src, draws)
                for key in ("paths", "lengths", "status", "draws_used"):
                    assert np.array_equal(r[key], ref[key]), (p, q, form, key)
    finally:
        dg.close()


def test_its_guide_from_a_loaded_prefix_sum_file(gpu_ctx, tmp_path):
    """The ITS table read from the reference's b_edge_weight_prefix_sum.bin (instead of built)
    gets guide records too, and the fused walk over them equals the oracle's in replay mode."""
    g = graphs.chung_lu(3000, 15000, 2.0, 2, 800, 111)
    src = np.nonzero(g.degree > 0)[0].astype(np.int32)
    L, K = 30, 40
    draws_wm = ob.sfmt_stream(112, len(src) * K).reshape(len(src), K)
    og = ob.Graph(g, None, ob.its_build(g, g.weight), g.weight, None)
    ref = ob.walk(og, "deepwalk", 1, 1, L, src, draws_wm, K)
    dg = dgraph(gpu_ctx, g)
    try:
        dg.build_tables("its", "edge")
        dg.store_tables(str(tmp_path))
    finally:
        dg.close()
    dg2 = dgraph(gpu_ctx, g)
    try:
        dg2.load_tables(str(tmp_path), "its", "edge")
        r = T.walk(dg2, T.make_params("deepwalk", execution="static", sampling="its", length=L, rng="replay",
                                      form="fused"), src, np.ascontiguousarray(draws_wm.T))
        for key in ("paths", "lengths", "ends", "status", "draws_used"):
            assert np.array_equal(r[key], ref[key]), key
    finally:
        dg2.close()


def test_node2vec_trial_cap_matches_oracle(gpu_ctx):
    """A rejection-trial cap (TRW_WALKER_TRIAL_CAP) ends Node2Vec walkers where the oracle's
    capped Move does, in every form (the fused form's per-lane state machine included), with
    the draws consumed up to the cap."""
    g = graphs.chung_lu(5000, 30000, 2.0, 2, 1500, 121)
    dg = dgraph(gpu_ctx, g)
    try:
        src = np.nonzero(g.degree > 0)[0].astype(np.int32)
        K = 4 * 30 + 64
        for cap in (1, 2):
            case = {"app": "node2vec", "execution": 2, "sampling": 4, "length": 30, "extra": {"-p": "2", "-q": "0.5"}}
            a = golden_case_args(case)
            og = ob.Graph(g)
            d_wm = ob.sfmt_stream(122, len(src) * K).reshape(len(src), K)
            ref = ob.walk(og, "node2vec", 2, 4, 30, src, d_wm, K, p=a["p"], q=a["q"], max_trials=cap)
            assert (ref["status"] == T._abi.WALKER_TRIAL_CAP).sum() > 0, cap
            for form in forms_for("node2vec"):
                prm = T.make_params("node2vec", execution="dynamic", sampling="max_weight_rejection", length=30,
                                    p=2.0, q=0.5, rng="replay", form=form, max_trials=cap)
                r = T.walk(dg, prm, src, np.ascontiguousarray(d_wm.T))
                for key in ("paths", "lengths", "ends", "status", "draws_used"):
                    assert np.array_equal(r[key], ref[key]), (cap, form, key)
    finally:
        dg.close()
