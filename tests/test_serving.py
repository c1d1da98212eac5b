"""Query server (trw_server_*): continuous batching over a resident graph.

Results are keyed by (seed, walker id), so each query's outcome is exactly the
same served alone, served in a mix of queries of other sizes and kinds, split
over several batches, or walked directly with trw_walk_async; PPR estimates agree
with the oracle's FP64 power iteration within the sampling bound."""

import numpy as np
import pytest

import graphs
import oracle_bind as ob
from test_bench_checks import B as BENCH
import thunderrw_amd as T
from thunderrw_amd import _abi as A

pytestmark = pytest.mark.gpu

PPR = dict(execution="static", sampling="alias", weights="edge", stop_probability=0.2, rng="philox", seed=5)


@pytest.fixture(scope="module")
def served(gpu_ctx):
    g = graphs.chung_lu(40000, 240000, 2.1, 2, 4000, 81)
    dg = T.DeviceGraph(gpu_ctx, g.degree, g.adj, g.weight)
    dg.build_tables("alias", "edge")
    yield g, dg
    dg.close()


def make_queries(g, n, seed):
    rng = np.random.default_rng(seed)
    live = np.nonzero(g.degree > 0)[0]
    out = []
    for i in range(n):
        k = int(np.exp(rng.uniform(0, np.log(64))))
        srcs = rng.choice(live, k).astype(np.int32)
        walkers = int(np.exp(rng.uniform(np.log(500), np.log(30000))))
        kind = "paths" if i % 5 == 3 else "ppr"
        out.append((kind, srcs, walkers, (i + 1) << 32))
    return out


def params_for(kind):
    if kind == "ppr":
        return T.make_params("ppr", **PPR)
    return T.make_params("deepwalk", execution="static", sampling="alias", length=24, rng="philox", seed=5)


def collect(srv, qid, kind):
    info = srv.wait(qid)
    assert info["state"] == A.QUERY_DONE and info["errors"] == 0
    if kind == "ppr":
        v, c = srv.ppr(qid)
        res = dict(v=v, c=c, steps=info["steps"])
    else:
        res = srv.walk_result(qid, length=24)
        res["steps"] = info["steps"]
    srv.release(qid)
    return res


def same(a, b):
    return all(np.array_equal(a[k], b[k]) if isinstance(a[k], np.ndarray) else a[k] == b[k] for k in a if a[k] is not None)


def test_each_result_is_the_same_alone_and_in_the_mix(served):
    g, dg = served
    qs = make_queries(g, 30, 1)
    alone = []
    with T.Server(dg, streams=1) as srv:
        for kind, srcs, n, base in qs:
            qid = srv.submit(params_for(kind), srcs, n, base, result=kind)
            alone.append(collect(srv, qid, kind))
    # Small batches split the large queries; four slots overlap batches on the device.
    with T.Server(dg, streams=4, max_batch_walkers=16384, path_bytes_per_batch=16384 * 24 * 4 // 2,
                  max_queries_per_batch=8) as srv:
        ids = [srv.submit(params_for(kind), srcs, n, base, result=kind) for kind, srcs, n, base in qs]
        mixed = [collect(srv, qid, q[0]) for qid, q in zip(ids, qs)]
        st = srv.stats()
    total = sum(n for _, _, n, _ in qs)
    assert max(n for _, _, n, _ in qs) > 16384  # at least one query was split
    assert st["batches"] >= total / 16384 and st["mean_batch_queries"] > 1.5
    for i, (a, b) in enumerate(zip(alone, mixed)):
        assert same(a, b), i
    # A direct walk of the same walkers (sources[i % |S|], ids base + i) gives the same result.
    for (kind, srcs, n, base), a in zip(qs[:6], alone[:6]):
        src = srcs[np.arange(n) % len(srcs)]
        r = T.walk(dg, params_for(kind), src, walker_ids=base + np.arange(n, dtype=np.uint64),
                   want_paths=kind != "ppr")
        if kind == "ppr":
            v, c = np.unique(r["ends"], return_counts=True)
            assert np.array_equal(v, a["v"]) and np.array_equal(c, a["c"])
        else:
            assert np.array_equal(r["paths"], a["paths"])
        assert r["steps"] == a["steps"]


def test_served_ppr_matches_power_iteration(served):
    g, dg = served
    live = np.nonzero(g.degree > 0)[0]
    sources = live[[3, 50, 700, 9000, 20000]].astype(np.int32)
    n = 1 << 21
    with T.Server(dg, streams=2, max_batch_walkers=1 << 19) as srv:
        qid = srv.submit(params_for("ppr"), sources, n, 1 << 40)
        info = srv.wait(qid)
        v, c = srv.ppr(qid)
    assert c.sum() == n
    est = np.zeros(g.nv)
    est[v] = c / n
    pi, _ = ob.ppr_power(g, 0.2, sources, weighted=True)
    bound = np.sum(np.sqrt(pi * (1 - pi) / n))
    assert np.abs(est - pi).sum() < 1.5 * bound
    assert abs(info["steps"] / n - 1 / 0.2) < 0.05


def test_admission_defers_and_rejects(served):
    g, dg = served
    live = np.nonzero(g.degree > 0)[0].astype(np.int32)
    with T.Server(dg, streams=1, max_batch_walkers=4096, max_pending_queries=4) as srv:
        ids, busy = [], 0
        for i in range(64):
            try:
                ids.append(srv.submit(params_for("ppr"), live[i:i + 3], 200000, (i + 1) << 32))
            except T.TrwError as e:
                assert e.code == A.E_BUSY
                busy += 1
        assert busy > 0 and len(ids) >= 4
        for qid in ids:
            assert srv.wait(qid)["state"] == A.QUERY_DONE
        assert srv.stats()["rejected"] == busy
    with T.Server(dg, streams=1, path_bytes_per_batch=64) as srv:
        with pytest.raises(T.TrwError):  # one 24-vertex path row does not fit 64 bytes
            srv.submit(params_for("paths"), live[:2], 10, result="paths")
        with pytest.raises(T.TrwError):  # replay streams cannot be served
            srv.submit(T.make_params("ppr", rng="replay"), live[:2], 10)
        with pytest.raises(T.TrwError):  # sources are checked at submit
            srv.submit(params_for("ppr"), np.array([g.nv], np.int32), 10)
        for form in ("step", "graph"):  # forms whose walker state is graph-wide scratch
            with pytest.raises(T.TrwError) as e:
                srv.submit(T.make_params("ppr", **dict(PPR, form=form)), live[:2], 10)
            assert e.value.code == A.E_UNSUPPORTED


def test_open_loop_reports_latency_percentiles(served):
    g, dg = served
    qs = make_queries(g, 300, 2)
    with T.Server(dg, streams=4) as srv:
        structs = [srv.query(params_for(kind), srcs, n, base, result=kind) for kind, srcs, n, base in qs]
        rep, ids, lat = srv.run_open_loop(structs, qps=3000.0, seed=7)
        assert rep["rejected"] == 0
        assert np.all(lat > 0)
        assert rep["p50_ms"] <= rep["p99_ms"] <= rep["max_ms"]
        assert abs(rep["p99_ms"] - np.quantile(lat, 0.99, method="inverted_cdf")) < 1e-9
        total = 0
        for qid in ids:
            total += srv.info(int(qid))["steps"]
            srv.release(int(qid))
        assert rep["steps"] == total > 0
        print("open loop: %.0f qps offered, %.0f achieved, p50 %.3f ms, p99 %.3f ms, %.3g steps/s, "
              "%.1f queries per batch" % (rep["offered_qps"], rep["achieved_qps"], rep["p50_ms"], rep["p99_ms"],
                                          rep["steps_per_s"], rep["mean_batch_queries"]))


def test_concurrent_clients_get_their_own_results(served):
    """Eight host threads submit and wait on one server at once (ctypes releases the GIL
    during each call); every result equals the one served alone."""
    import threading
    g, dg = served
    qs = make_queries(g, 48, 9)
    with T.Server(dg, streams=1) as srv:
        alone = [collect(srv, srv.submit(params_for(k), s, n, b, result=k), k) for k, s, n, b in qs]
    results = [None] * len(qs)
    errors = []
    with T.Server(dg, streams=4, max_batch_walkers=32768) as srv:
        def client(idx):
            try:
                for i in idx:
                    k, s, n, b = qs[i]
                    results[i] = collect(srv, srv.submit(params_for(k), s, n, b, result=k), k)
            except Exception as e:  # surfaced below
                errors.append(e)
        threads = [threading.Thread(target=client, args=(range(t, len(qs), 8),)) for t in range(8)]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
    assert not errors, errors
    for i, (a, b) in enumerate(zip(alone, results)):
        assert same(a, b), i


def test_served_ppr_l1_on_a_2p20_vertex_graph(gpu_ctx):
    """The serving query mix (|S| ~ logU[1, 4096] sources uniform over the non-isolated vertices,
    N ~ logU[1e3, 1e5] walkers, termination 0.2, weighted alias) on a 2^20-vertex Chung-Lu graph
    (exponent 2.1, average degree 16, dmin 2, dmax 65536, FP64 U[1,5) weights): each query's
    estimate count / N against the oracle's FP64 power iteration started from the same walker
    sources (ppr.cpp:27-29 semantics), L1 printed and below 1.5 x the summed per-vertex standard
    errors sqrt(pi (1 - pi) / N); and judged as bench.py judges the 2^24-vertex serving run (pinned in
    test_bench_checks.py): L1 within 4 / sqrt(N) of what multinomial sampling alone gives, and
    Pearson's chi-square over the vertices expected to end 5 or more walks above 1e-4."""
    hg = T.HostGraph.chung_lu(1 << 20, 8 << 20, 2.1, 2.0, 65536.0, 51, True, 0)
    deg, adj, w, _ = hg.arrays()
    g = graphs.CSR(deg.copy(), adj.copy(), w.copy())
    hg.close()
    dg = T.DeviceGraph(gpu_ctx, g.degree, g.adj, g.weight)
    try:
        dg.build_tables("alias", "edge")
        rng = np.random.default_rng(52)
        live = np.nonzero(g.degree > 0)[0].astype(np.int32)
        qs = []
        for i in range(4):
            k = int(np.exp(rng.uniform(0.0, np.log(4096.0))))
            nw = int(np.exp(rng.uniform(np.log(1e3), np.log(1e5))))
            qs.append((rng.choice(live, k).astype(np.int32), nw))
        with T.Server(dg) as srv:
            ids = [srv.submit(T.make_params("ppr", **PPR), s, nw, (i + 1) << 32) for i, (s, nw) in enumerate(qs)]
            results = []
            for qid in ids:
                assert srv.wait(qid)["state"] == A.QUERY_DONE
                results.append(srv.ppr(qid))
                srv.release(qid)
        for (s, nw), (v, c) in zip(qs, results):
            assert c.sum() == nw
            start = s[np.arange(nw) % len(s)]  # walker i starts at sources[i % |S|]
            pi, _ = ob.ppr_power(g, 0.2, start, weighted=True, tol=1e-10)
            est = np.zeros(g.nv)
            est[v] = c / nw
            l1 = np.abs(est - pi).sum()
            bound = np.sqrt(pi * (1 - pi) / nw).sum()
            judged = BENCH.judge_ppr(pi[:, None], [(s, nw, v, c.astype(np.float64))])[0]
            print("served PPR: |S| %d, N %d, L1 %.4f, summed standard errors %.4f, %s" % (len(s), nw, l1, bound, judged),
                  flush=True)
            assert l1 < 1.5 * bound
            assert judged["ok"], judged
    finally:
        dg.close()


@pytest.mark.parametrize("nv", [37, 4101, 2 * 131072 + 77, 5 * 131072])
def test_ppr_counts_at_every_bucket_shape(gpu_ctx, nv):
    """PPR batches count end vertices per (segment, chunk of 2^17 vertex ids) bucket in LDS:
    a bucket of at most 2048 walkers in a hash, a larger one by rank counters in passes.
    One chunk holding every walker (so large buckets), one ragged chunk, three and five
    chunks: many small batches on one slot and on four slots give, per query, exactly
    np.unique of a direct walk's ends."""
    g = graphs.hub_graph(nv, min(nv - 1, 30), 3 * nv, 90 + nv)
    dg = T.DeviceGraph(gpu_ctx, g.degree, g.adj, g.weight)
    try:
        dg.build_tables("alias", "edge")
        rng = np.random.default_rng(nv)
        live = np.nonzero(g.degree > 0)[0].astype(np.int32)
        qs = []
        for i in range(12):
            k = int(rng.integers(1, 9))
            qs.append((rng.choice(live, k).astype(np.int32), int(rng.integers(1, 9000)), (i + 3) << 32))
        want = []
        for s, n, base in qs:
            r = T.walk(dg, params_for("ppr"), s[np.arange(n) % len(s)], walker_ids=base + np.arange(n, dtype=np.uint64))
            v, c = np.unique(r["ends"], return_counts=True)
            want.append((v, c, r["steps"]))
        for streams, cap in ((1, 4096), (4, 2500)):
            with T.Server(dg, streams=streams, max_batch_walkers=cap, max_queries_per_batch=5) as srv:
                ids = [srv.submit(params_for("ppr"), s, n, base) for s, n, base in qs]
                for qid, (v, c, steps) in zip(ids, want):
                    info = srv.wait(qid)
                    got_v, got_c = srv.ppr(qid)
                    srv.release(qid)
                    assert info["errors"] == 0 and info["steps"] == steps and info["ppr_entries"] == len(v)
                    assert np.array_equal(got_v, v) and np.array_equal(got_c, c)
    finally:
        dg.close()
