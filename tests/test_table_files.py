"""The reference's table files (b_edge_weight_alias_table.bin, b_edge_weight_prefix_sum.bin,
b_edge_weight_max.bin; util/graph/graph.cpp:18-30, 184-212): tables built on the device and
stored equal, byte for byte, the oracle's tables and the files the reference's own
alias_preprocess / its_preprocess / rejection_preprocess write (compiled into oracle/_ref);
tables loaded from such files walk exactly like built ones."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import graphs
import oracle_bind as ob
import thunderrw_amd as T
from conftest import ROOT

pytestmark = pytest.mark.gpu

REF = os.path.join(ROOT, "oracle", "_ref")
FILES = {"alias": "b_edge_weight_alias_table.bin", "its": "b_edge_weight_prefix_sum.bin",
         "rejection": "b_edge_weight_max.bin"}
TOOLS = {"alias": "alias_preprocess.out", "its": "its_preprocess.out", "rejection": "rejection_preprocess.out"}


def read_bytes(path):
    with open(path, "rb") as f:
        return f.read()


@pytest.mark.parametrize("gname", ["chung_lu", "hub"])
def test_stored_tables_match_oracle_and_reference_tools(gpu_ctx, gname):
    g = (graphs.chung_lu(4000, 30000, 2.1, 2, 1500, 91) if gname == "chung_lu"
         else graphs.hub_graph(3000, 2800, 8000, 92))
    w = np.asarray(g.weight, np.float64)
    oracle = {"alias": ob.alias_build(g, w)[0].view(np.uint8).tobytes(), "its": ob.its_build(g, w).tobytes(),
              "rejection": ob.rs_build(g, w).tobytes()}
    dg = T.DeviceGraph(gpu_ctx, g.degree, g.adj, g.weight)
    try:
        with tempfile.TemporaryDirectory() as tmp:
            with pytest.raises(T.TrwError):
                dg.store_tables(tmp)  # nothing built yet
            for sm in FILES:
                dg.build_tables(sm, "edge")
            dg.store_tables(tmp)
            for sm, name in FILES.items():
                assert read_bytes(os.path.join(tmp, name)) == oracle[sm], sm
            ref_dir = os.path.join(tmp, "ref")
            graphs.write_reference_csr(g, ref_dir)
            for sm, tool in TOOLS.items():
                exe = os.path.join(REF, tool)
                if not os.path.exists(exe):
                    pytest.skip("reference tool %s not built (make -C oracle ref): the files are checked "
                                "against the oracle only" % tool)
                r = subprocess.run([exe, "-f", ref_dir], capture_output=True, text=True, timeout=120)
                assert r.returncode == 0, r.stderr[-1000:]
                ref_b, ours_b = read_bytes(os.path.join(ref_dir, FILES[sm])), read_bytes(os.path.join(tmp, FILES[sm]))
                if sm == "rejection":
                    # rejection_preprocess.cpp:26 dereferences max_element of an empty range for an
                    # isolated vertex (it stores the next vertex's first weight); the build defines
                    # that maximum as 0.0.  Every other vertex must match bit for bit.
                    live = g.degree > 0
                    ref_m, ours_m = np.frombuffer(ref_b, np.uint64), np.frombuffer(ours_b, np.uint64)
                    assert np.array_equal(ref_m[live], ours_m[live])
                    assert np.all(np.frombuffer(ours_b, np.float64)[~live] == 0.0)
                else:
                    assert ref_b == ours_b, sm
            # Loading the files instead of building gives the same walks.
            dl = T.DeviceGraph(gpu_ctx, g.degree, g.adj, g.weight)
            try:
                src = np.nonzero(g.degree > 0)[0].astype(np.int32)
                K = 4 * 30 + 64
                draws = T.replay_fill(93, len(src), K)
                for sm in FILES:
                    dl.load_tables(tmp, sm, "edge")
                    prm = T.make_params("deepwalk", execution="static", sampling=sm, length=30, rng="replay")
                    a = T.walk(dg, prm, src, draws)
                    b = T.walk(dl, prm, src, draws)
                    for key in ("paths", "lengths", "status", "draws_used"):
                        assert np.array_equal(a[key], b[key]), (sm, key)
                os.makedirs(os.path.join(tmp, "bad"))
                with open(os.path.join(tmp, "bad", FILES["alias"]), "wb") as f:
                    f.write(bytes(16))
                with pytest.raises(T.TrwError) as e:
                    dl.load_tables(os.path.join(tmp, "bad"), "alias")
                assert e.value.code == T._abi.E_IO
            finally:
                dl.close()
    finally:
        dg.close()
