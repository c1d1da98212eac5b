"""The reference's own app sources, compiled unchanged against thunderrw_amd/dropin/creeper_main.h
(thunderrw_amd/dropin/bin/<app>, built from the reference checkout), run their walks on the GPU;
in replay mode their dumped paths equal the oracle's."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import graphs
import oracle_bind as ob
import thunderrw_amd as T
from conftest import ROOT, load_golden_graph

BIN = os.path.join(ROOT, "thunderrw_amd", "dropin", "bin")
pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(os.path.join(BIN, "deepwalk")),
                                 reason="drop-in apps not built (needs the reference checkout at build time)")]

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
