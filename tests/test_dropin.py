"""The reference's own app sources, compiled unchanged against thunderrw_amd/dropin/creeper_main.h
(thunderrw_amd/dropin/bin/<app>, built from the reference checkout), run their walks on the GPU;
in replay mode their dumped paths equal the oracle's.  Functors the library has no Move kernel
for (the reference apps in other mode/sampler pairs, tests/dropin_apps/customwalk.cpp) run in
the functor engine and walk as the reference's own engine does (harness fixtures), and as the
library's kernels do where the functor equals a built-in app."""
import json
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import graphs
import oracle_bind as ob
import thunderrw_amd as T
from conftest import ROOT, load_golden_graph
from thunderrw_amd import _abi as A

BIN = os.path.join(ROOT, "thunderrw_amd", "dropin", "bin")
pytestmark = [pytest.mark.gpu]
needs_reference_apps = pytest.mark.skipif(not os.path.exists(os.path.join(BIN, "deepwalk")),
                                          reason="drop-in apps not built (needs the reference checkout at build time)")

# app, args, oracle (app, execution, sampling, length, extra), sources kind, K, KT
CASES = [
    ("deepwalk", ["-ew", "-l", "20"], ("deepwalk", 1, 2, 20, {}), "all", 38, 0),
    ("deepwalk", ["-ew", "-l", "20", "-sm", "1"], ("deepwalk", 1, 1, 20, {}), "all", 19, 0),
    ("deepwalk", ["-ew", "-l", "20", "-sm", "3"], ("deepwalk", 1, 3, 20, {}), "all", 608, 0),
    ("node2vec", ["-l", "20"], ("node2vec", 2, 4, 20, {"p": 2.0, "q": 0.5}), "all", 608, 0),
    ("node2vec", ["-l", "20", "-p", "0.25", "-q", "4"], ("node2vec", 2, 4, 20, {"p": 0.25, "q": 4.0}), "all", 608, 0),
    ("metapath", ["-el", "-l", "20", "-s", "0,1,2,3,4"], ("metapath", 2, 1, 20, {"schema": [0, 1, 2, 3, 4]}), "all",
     19, 0),
    ("ppr", ["-sp", "0.2", "-nw", "300"], ("ppr", 0, 0, 1, {"stop_prob": 0.2}), "single", 4096, 4096),
    ("genericwalk", ["-ew", "-l", "20", "-st", "2", "-wn", "500", "-sm", "2"], ("deepwalk", 1, 2, 20, {}), "random",
     38, 0),
]


@needs_reference_apps
@pytest.mark.parametrize("gname", ["small", "hub"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0] + "_" + "_".join(c[1][1:4]))
def test_reference_app_runs_unchanged_on_the_gpu(gname, case):
    app, args, (oapp, em, sm, L, extra), skind, K, KT = case
    g, _ = load_golden_graph(gname)
    with tempfile.TemporaryDirectory() as tmp:
        gdir = os.path.join(tmp, "g")
        graphs.write_reference_csr(g, gdir)
        out = os.path.join(tmp, "out")
        os.makedirs(out)
        env = dict(os.environ, TRW_RNG="replay", TRW_SEED="1234", TRW_TSEED="4321", TRW_REPLAY_K=str(K),
                   TRW_REPLAY_KT=str(max(KT, 1)), TRW_DUMP_DIR=out)
        r = subprocess.run([os.path.join(BIN, app), "-f", gdir] + args, env=env, capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr[-2000:]
        assert "Throughput (steps per second)" in r.stderr
        paths = np.fromfile(os.path.join(out, "paths.bin"), np.int32)
        lengths = np.fromfile(os.path.join(out, "lengths.bin"), np.int32)
    if skind == "all":
        src = T.sources_all(g.degree)
    elif skind == "single":
        src = T.sources_single(g.degree, 300)
    else:
        src = T.sources_random(g.degree, 500, 0)
    n = len(src)
    w = g.weight if oapp != "ppr" else np.ones(g.ne)
    alias, _ = ob.alias_build(g, w)
    og = ob.Graph(g, alias, ob.its_build(g, w), w, ob.rs_build(g, w))
    d = ob.sfmt_stream(1234, n * K).reshape(n, K)
    KT1 = max(KT, 1)
    td = ob.sfmt_stream(4321, n * KT1).reshape(n, KT1)
    ref = ob.walk(og, oapp, em, sm, L, src, d, K, td, KT1, **extra)
    assert np.array_equal(paths.reshape(n, -1), ref["paths"])
    assert np.array_equal(lengths, ref["lengths"])


# ---- the functor engine: functors the library has no Move kernel for ----

def _functor_cases():
    with open(os.path.join(ROOT, "tests", "golden", "index.json")) as f:
        return json.load(f)["functor_walks"]


def _run_app(app, g, args, env_extra):
    """Runs dropin/bin/<app> on g (written as reference CSR files) and returns its dumped walk,
    the steps it logged and its configuration line."""
    with tempfile.TemporaryDirectory() as tmp:
        gdir = os.path.join(tmp, "g")
        graphs.write_reference_csr(g, gdir)
        out = os.path.join(tmp, "out")
        os.makedirs(out)
        env = dict(os.environ, TRW_DUMP_DIR=out, **env_extra)
        r = subprocess.run([os.path.join(BIN, app), "-f", gdir, "-ew", "-el"] + args, env=env, capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stderr[-2000:]
        res = {k: np.fromfile(os.path.join(out, k + ".bin"), np.int32) for k in ("paths", "lengths", "status")}
    res["steps"] = int(re.search(r"Num of steps: (\d+)", r.stderr).group(1))
    res["config"] = re.search(r"Configuration: (.*)", r.stderr).group(1)
    return res


def _app_args(case):
    args = ["-l", str(case["length"]), "-em", str(case["execution"]), "-sm", str(case["sampling"])]
    for k, v in case["extra"].items():
        args += [k, v]
    return args


@pytest.mark.parametrize("case", _functor_cases(), ids=lambda c: c["case"])
def test_functor_engine_matches_the_reference_harness(case):
    """Replay mode: the reference's own Moves, gather and update driving the same functor
    (oracle/_ref/harness_<app>, fixtures tests/golden/functor_*.npz) and the functor engine
    running it on the GPU give the same walks, bit for bit."""
    if not os.path.exists(os.path.join(BIN, case["app"])):
        pytest.skip("drop-in app %s not built" % case["app"])
    g, _ = load_golden_graph(case["graph"])
    ref = np.load(os.path.join(ROOT, "tests", "golden", "functor_%s.npz" % case["case"]))
    r = _run_app(case["app"], g, _app_args(case),
                 {"TRW_RNG": "replay", "TRW_SEED": "1234", "TRW_REPLAY_K": str(case["K"])})
    assert r["config"].endswith("functor engine")
    n = len(ref["lengths"])
    assert np.array_equal(r["paths"].reshape(n, -1), ref["paths"])
    assert np.array_equal(r["lengths"], ref["lengths"])
    assert np.array_equal(r["status"], ref["status"])  # 0 done, 1 gather-terminated in both
    assert r["steps"] == int(ref["steps"][0])


# customwalk kinds 3 and 4 are DeepWalk and Node2Vec under other names, kind 7 Node2Vec over edge
# weights: in free-running mode the functor engine must walk exactly as the library's kernels for
# those apps (kind 7: the Node2Vec kernel with weights EDGE).
TWINS = [(3, 1, 2, "deepwalk"), (3, 1, 1, "deepwalk"), (3, 1, 3, "deepwalk"), (3, 2, 1, "deepwalk"),
         (3, 2, 2, "deepwalk"), (3, 2, 3, "deepwalk"), (3, 0, 0, "deepwalk"), (4, 2, 4, "node2vec"),
         (7, 2, 4, "node2vec")]


@pytest.mark.parametrize("gname", ["small", "hub"])
@pytest.mark.parametrize("twin", TWINS, ids=lambda t: "kind%d_em%d_sm%d" % t[:3])
def test_functor_engine_walks_as_the_library_in_free_running_mode(gpu_ctx, gname, twin):
    kind, em, sm, app = twin
    if not os.path.exists(os.path.join(BIN, "customwalk")):
        pytest.skip("test app not built")
    g, _ = load_golden_graph(gname)
    r = _run_app("customwalk", g, ["-l", "20", "-em", str(em), "-sm", str(sm), "-kind", str(kind)],
                 {"TRW_RNG": "philox", "TRW_SEED": "77"})
    assert r["config"].endswith("functor engine")
    src = T.sources_all(g.degree)
    dg = T.DeviceGraph(gpu_ctx, g.degree, g.adj, g.weight, g.label)
    try:
        lib = T.walk(dg, T.make_params(app, execution=em, sampling=sm, length=20, rng="philox", seed=77,
                                       weights="edge" if kind == 7 else "app"), src)
    finally:
        dg.close()
    n = len(src)
    assert np.array_equal(r["paths"].reshape(n, -1), lib["paths"])
    assert np.array_equal(r["lengths"], lib["lengths"])
    assert r["steps"] == lib["steps"]


def test_functor_engine_caps_walks_the_reference_would_never_end(gpu_ctx):
    """A functor whose update() never ends a walk stops at the engine's move cap
    (TRW_WALKER_STEP_CAP) with its first moves those of the walk it is; one whose weights are
    all zero stops at the rejection trial cap (TRW_WALKER_TRIAL_CAP).  Every kernel ends."""
    if not os.path.exists(os.path.join(BIN, "customwalk")):
        pytest.skip("test app not built")
    g, _ = load_golden_graph("small")
    src = T.sources_all(g.degree)
    n, L = len(src), 20
    r = _run_app("customwalk", g, ["-l", str(L), "-em", "1", "-sm", "2", "-kind", "5"],
                 {"TRW_RNG": "philox", "TRW_SEED": "5", "TRW_MAX_MOVES": "50"})
    assert np.all(r["status"] == A.WALKER_STEP_CAP)
    assert np.all(r["lengths"] == 51)
    assert r["steps"] == 50 * n
    dg = T.DeviceGraph(gpu_ctx, g.degree, g.adj, g.weight, g.label)
    try:
        lib = T.walk(dg, T.make_params("deepwalk", execution="static", sampling="alias", length=L, rng="philox",
                                       seed=5), src)
    finally:
        dg.close()
    assert np.array_equal(r["paths"].reshape(n, -1), lib["paths"])  # the rows hold the first L positions
    r = _run_app("customwalk", g, ["-l", str(L), "-em", "2", "-sm", "4", "-kind", "6"],
                 {"TRW_RNG": "philox", "TRW_SEED": "5", "TRW_MAX_TRIALS": "100"})
    assert np.all(r["status"] == A.WALKER_TRIAL_CAP)
    assert np.all(r["lengths"] == 1)
    assert r["steps"] == 0
