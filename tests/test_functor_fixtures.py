"""CPU checks of the functor-engine fixtures (tests/golden/functor_*.npz): walks of functors the
library has no Move kernel for, written by the reference's own engine (oracle/_ref/harness_<app>
driving Moves, dynamic_gather and update with one live ring slot).  The GPU side of the same
cases is tests/test_dropin.py.

What pins them beyond being the reference's output: every path is a walk on the graph, the
step count is the moves taken, a walk ends only as update() or the gather says (closed forms
of each functor's weight and update), and a fresh harness run reproduces them."""
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

import graphs
import thunderrw_amd as T
from conftest import GOLDEN, ROOT, load_golden_graph

REF_HARNESS = os.path.join(ROOT, "oracle", "_ref")


def cases():
    with open(os.path.join(GOLDEN, "index.json")) as f:
        return json.load(f)["functor_walks"]


def load_case(case):
    return np.load(os.path.join(GOLDEN, "functor_%s.npz" % case["case"]))


def neighbours(g, v):
    off = np.concatenate([[0], np.cumsum(g.degree, dtype=np.int64)])
    return g.adj[off[v]:off[v + 1]], np.arange(off[v], off[v + 1])


@pytest.mark.parametrize("case", cases(), ids=lambda c: c["case"])
def test_fixture_walks_follow_edges_and_end_as_the_functor_says(case):
    g, _ = load_golden_graph(case["graph"])
    z = load_case(case)
    src = T.sources_all(g.degree)
    L = case["length"]
    paths, lengths, status = z["paths"], z["lengths"], z["status"]
    assert paths.shape == (len(src), L)
    assert np.array_equal(paths[:, 0], src)
    assert int(z["steps"][0]) == int((lengths - 1).sum())  # one step per move (creeper.h:305)
    kind = int(case["extra"].get("-kind", -1))
    schema = [int(x) for x in case["extra"]["-s"].split(",")] if "-s" in case["extra"] else None
    for i in range(len(src)):
        n = int(lengths[i])
        assert 1 <= n <= L
        row = paths[i]
        assert np.all(row[n:] == -1)
        for a, b in zip(row[:n - 1], row[1:n]):
            assert b in neighbours(g, a)[0]
        end = row[n - 1]
        if status[i] == 1:
            # The gather found every weight zero at the end vertex (creeper.h:267-272, 282-292).
            nb, eids = neighbours(g, end)
            if case["app"] == "metapath":
                assert not np.any(g.label[eids] == schema[(n - 1) % len(schema)])
            else:
                assert kind == 2
                assert not np.any((nb + n) % 4 == 0)
        else:
            assert status[i] == 0
            # update(): length reached, or (kind 2) a vertex of degree 1.
            assert n == L or (kind == 2 and g.degree[end] == 1)


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_HARNESS, "harness_customwalk")),
                    reason="reference harness not built (make -C oracle ref; build container only)")
@pytest.mark.parametrize("case", cases()[::5], ids=lambda c: c["case"])
def test_fixture_matches_a_fresh_harness_run(case):
    g, _ = load_golden_graph(case["graph"])
    src = T.sources_all(g.degree)
    z = load_case(case)
    with tempfile.TemporaryDirectory() as tmp:
        gdir = os.path.join(tmp, "g")
        graphs.write_reference_csr(g, gdir)
        src.tofile(os.path.join(tmp, "src.bin"))
        cmd = [os.path.join(REF_HARNESS, "harness_" + case["app"]), "-f", gdir, "-ew", "-el", "-o", tmp, "-mode",
               "walk", "-em", str(case["execution"]), "-sm", str(case["sampling"]), "-l", str(case["length"]), "-src",
               os.path.join(tmp, "src.bin"), "-K", str(case["K"]), "-KT", "1", "-seed", "1234", "-tseed", "4321"]
        for k, v in case["extra"].items():
            cmd += [k, v]
        subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
        paths = np.fromfile(os.path.join(tmp, "paths.bin"), np.int32).reshape(len(src), -1)
    assert np.array_equal(paths, z["paths"])
