"""Edge-list loading (Graph::load_edge_list, util/graph/graph.cpp:428-519) against the
reference's own converter, toolset/edgelist_to_csr.cpp compiled into oracle/_ref."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import thunderrw_amd as T
from conftest import ROOT

TOOL = os.path.join(ROOT, "oracle", "_ref", "edgelist_to_csr")


def write_lines(d, lines):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "b_edge_list.bin"), "w") as f:
        f.write("".join(l + "\n" for l in lines))


def random_edge_lines(seed, n=300, m=2000):
    rng = np.random.default_rng(seed)
    lines = ["# comment line", "% other comment"]
    for _ in range(m):
        a, b = (int(x) for x in rng.integers(0, n, 2))
        r = rng.random()
        if r < 0.05:
            lines.append("%d %d" % (a, a))              # self-loop
        elif r < 0.1:
            lines.append("  %d\t%d  trailing" % (b, a))  # reversed, whitespace, trailing text
        else:
            lines.append("%d %d" % (a, b))
        if rng.random() < 0.1:
            lines.append(lines[-1])                      # duplicate
    lines.append("%d %d" % (n + 7, n + 7))               # the largest id only on a self-loop
    return lines


@pytest.mark.skipif(not os.path.exists(TOOL), reason="reference converter not built (needs the reference checkout)")
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_edge_list_csr_equals_reference_converter(seed):
    with tempfile.TemporaryDirectory() as tmp:
        src, ref_out, our_out = (os.path.join(tmp, x) for x in ("in", "ref", "ours"))
        write_lines(src, random_edge_lines(seed))
        os.makedirs(ref_out)
        # Only lines starting with the skip character are comments: the '%' line is an
        # error for the reference (exit -1) and for the loader (TRW_E_IO).
        r = subprocess.run([TOOL, src, ref_out, "#"], capture_output=True, text=True, timeout=60)
        assert r.returncode != 0
        with pytest.raises(T.TrwError):
            T.HostGraph.load_edge_list(src, "#")
        with open(os.path.join(src, "b_edge_list.bin")) as f:
            text = f.read().replace("% other comment\n", "")
        with open(os.path.join(src, "b_edge_list.bin"), "w") as f:
            f.write(text)
        r = subprocess.run([TOOL, src, ref_out, "#"], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stderr
        h = T.HostGraph.load_edge_list(src, "#")
        h.store(our_out)
        h.close()
        for name in ("b_degree.bin", "b_adj.bin"):
            a = open(os.path.join(ref_out, name), "rb").read()
            b = open(os.path.join(our_out, name), "rb").read()
            assert a == b, name


def test_edge_list_rejects_what_the_reference_cannot_parse():
    with tempfile.TemporaryDirectory() as tmp:
        # 2147483647 parses as an int32 but leaves no int32 vertex count (largest id + 1).
        for bad in (["1 2", "3"], ["1 2", "x 4"], ["1 2", ""], ["1 -2"], ["1 99999999999"], ["2147483647 3"],
                    ["0 2147483647"]):
            write_lines(tmp, bad)
            with pytest.raises(T.TrwError) as e:
                T.HostGraph.load_edge_list(tmp, "#")
            assert e.value.code == T._abi.E_IO
        write_lines(tmp, ["# only comments"])
        h = T.HostGraph.load_edge_list(tmp, "#")
        d, a, _, _ = h.arrays()
        assert len(d) == 1 and len(a) == 0
        h.close()
