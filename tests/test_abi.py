"""The C-ABI library loads, exports what include/thunderrw.h declares, keeps the
struct layouts the header defines, and its host-side parts (replay stream,
walker sources, generator, CSR files, argument checks) behave.  No GPU compute."""
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import graphs
import oracle_bind as ob
import thunderrw_amd as T
from conftest import ROOT
from thunderrw_amd import _abi as A

HEADER = os.path.join(ROOT, "include", "thunderrw.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(trw_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_function():
    lib = T.lib()
    names = declared_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(A.exported_names()) == names


STRUCTS = {"trw_walk_params": A.WalkParams, "trw_walk_io": A.WalkIO, "trw_alias_slot": A.AliasSlot,
           "trw_server_config": A.ServerConfig, "trw_query": A.Query, "trw_query_info": A.QueryInfo,
           "trw_server_stats": A.ServerStats, "trw_open_loop_report": A.OpenLoopReport,
           "trw_tune_report": A.TuneReport, "trw_device_view": A.DeviceView}


def test_struct_layouts_match_the_header():
    """Every field offset and struct size of the ctypes mirror equals the C compiler's."""
    lines = []
    for cname, cls in STRUCTS.items():
        lines.append('  printf("%s.sizeof %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in cls._fields_:
            lines.append('  printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    src = "#include <stddef.h>\n#include <stdio.h>\n#include \"thunderrw.h\"\nint main(void) {\n%s\n  return 0;\n}\n" % (
        "\n".join(lines))
    with tempfile.TemporaryDirectory() as tmp:
        c = os.path.join(tmp, "layout.c")
        open(c, "w").write(src)
        exe = os.path.join(tmp, "layout")
        subprocess.run(["gcc", "-I", os.path.dirname(HEADER), "-o", exe, c], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    import ctypes
    seen = 0
    for line in out:
        if not line:
            continue
        name, val = line.split()
        t, f = name.split(".")
        cls = STRUCTS[t]
        want = ctypes.sizeof(cls) if f == "sizeof" else getattr(cls, f).offset
        assert want == int(val), name
        seen += 1
    assert seen == sum(len(c._fields_) + 1 for c in STRUCTS.values())
    assert ctypes.sizeof(A.AliasSlot) == 16


def test_replay_fill_is_the_chunked_sfmt_stream():
    n, K = 53, 23
    d = T.replay_fill(1234, n, K)
    s = ob.sfmt_stream(1234, n * K).reshape(n, K)
    assert d.shape == (K, n)
    assert np.array_equal(d, s.T)
    assert np.array_equal(T.replay_fill(1234, 7, K, first_walker=40), s[40:47].T)


def test_sources_single_and_random_follow_walker_generator():
    g, _ = graphs.chung_lu(500, 1500, 2.1, 2, 60, 3), None
    deg = g.degree.copy()
    deg[:50] = 0
    assert np.all(T.sources_single(deg, 10) == int(np.argmax(deg)))
    got = T.sources_random(deg, 300, seed=9)
    # generate_random_source (walker_generator.cpp:27-40): SFMT(seed) u % n, skipping degree 0.
    stream = ob.sfmt_stream(9, 5000)
    expect = [int(u % len(deg)) for u in stream if deg[int(u % len(deg))] != 0][:300]
    assert got.tolist() == expect


def test_sources_all_is_a_permutation_of_live_vertices():
    deg = np.array([3, 0, 1, 0, 2, 5, 0, 1], dtype=np.int32)
    s = T.sources_all(deg)
    assert sorted(s.tolist()) == [0, 2, 4, 5, 7]
    assert np.array_equal(s, T.sources_all(deg))


@pytest.mark.parametrize("gname", ["small", "hub"])
def test_sources_equal_the_reference_walker_generator(gname, golden_index):
    """tests/golden/sources_<graph>.npz holds the walkers the reference's own walker_generator.cpp
    (generate_all_source's std::shuffle with std::mt19937(0), generate_random_source,
    generate_single_source) made on the golden graphs."""
    from conftest import load_golden_graph

    g, _ = load_golden_graph(gname)
    z = np.load(os.path.join(ROOT, "tests", "golden", "sources_%s.npz" % gname))
    args = golden_index["sources"]
    assert np.array_equal(T.sources_all(g.degree), z["all"])
    assert np.array_equal(T.sources_random(g.degree, args["wn"], seed=args["wseed"]), z["random"])
    assert np.array_equal(T.sources_single(g.degree, args["wn"]), z["single"])


def test_dropin_sfmt_header_known_answers():
    """thunderrw_amd/dropin/SFMT.h (the SFMT names the reference's functors call, e.g. PPR::sfmt_)
    against the reference-held KAT: init_gen_rand(1234), first 1000 words; and real1/real2."""
    kat = np.load(os.path.join(ROOT, "tests", "golden", "sfmt_kat.npz"))["init_gen_rand_1234"]
    src = ('#include <stdio.h>\n#include "SFMT.h"\nint main(void) {\n  sfmt_t s;\n  sfmt_init_gen_rand(&s, 1234);\n'
           '  for (int i = 0; i < 1000; ++i) printf("%u\\n", sfmt_genrand_uint32(&s));\n'
           '  printf("%.17g %.17g\\n", sfmt_to_real1(4294967295u), sfmt_to_real2(4294967295u));\n  return 0;\n}\n')
    with tempfile.TemporaryDirectory() as tmp:
        c = os.path.join(tmp, "kat.c")
        open(c, "w").write(src)
        exe = os.path.join(tmp, "kat")
        subprocess.run(["gcc", "-O2", "-I", os.path.join(ROOT, "thunderrw_amd", "dropin"), "-o", exe, c], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    assert np.array_equal(np.array([int(x) for x in out[:1000]], np.uint32), kat)
    assert float(out[1000]) == 1.0 and float(out[1001]) == (2.0 ** 32 - 1) / 2.0 ** 32


def test_generator_is_a_simple_symmetric_sorted_graph():
    h = T.HostGraph.chung_lu(20000, 80000, 2.1, 2, 500, 7, True, 5)
    deg, adj, w, lab = h.arrays()
    n = len(deg)
    assert abs(deg.sum() / 2 - 80000) <= 0.02 * 80000
    off = np.zeros(n + 1, np.int64)
    np.cumsum(deg, out=off[1:])
    u = np.repeat(np.arange(n), deg)
    assert np.all(u != adj)
    row_break = np.zeros(len(adj), bool)
    row_break[off[1:-1][deg[1:] > 0]] = True
    assert np.all((np.diff(adj) > 0) | row_break[1:])
    key = np.minimum(u, adj).astype(np.int64) * n + np.maximum(u, adj)
    o = np.argsort(key, kind="stable")
    assert np.array_equal(key[o][0::2], key[o][1::2])
    assert np.array_equal(w[o][0::2], w[o][1::2]) and np.array_equal(lab[o][0::2], lab[o][1::2])
    assert w.min() >= 1.0 and w.max() < 5.0 and set(np.unique(lab)) <= set(range(5))
    h2 = T.HostGraph.chung_lu(20000, 80000, 2.1, 2, 500, 7, True, 5)
    assert all(np.array_equal(a, b) for a, b in zip(h.arrays(), h2.arrays()))


def test_reference_csr_files_round_trip():
    g = graphs.chung_lu(300, 900, 2.1, 2, 50, 5)
    with tempfile.TemporaryDirectory() as tmp:
        graphs.write_reference_csr(g, tmp)
        h = T.HostGraph.load(tmp, True, True)
        deg, adj, w, lab = h.arrays()
        assert np.array_equal(deg, g.degree) and np.array_equal(adj, g.adj)
        assert np.array_equal(w, g.weight) and np.array_equal(lab, g.label)
        out = os.path.join(tmp, "copy")
        h.store(out)
        for f in ("b_degree.bin", "b_adj.bin", "b_edge_weight.bin", "b_edge_label.bin"):
            assert open(os.path.join(tmp, f), "rb").read() == open(os.path.join(out, f), "rb").read()


def test_reference_csr_load_rejects_inconsistent_files():
    g = graphs.chung_lu(100, 300, 2.1, 2, 30, 5)
    with tempfile.TemporaryDirectory() as tmp:
        graphs.write_reference_csr(g, tmp)
        with open(os.path.join(tmp, "b_adj.bin"), "r+b") as f:
            f.truncate(8)
        with pytest.raises(T.TrwError) as e:
            T.HostGraph.load(tmp, False, False)
        assert e.value.code == A.E_IO
        with pytest.raises(T.TrwError):
            T.HostGraph.load(os.path.join(tmp, "missing"), False, False)


def test_argument_errors_are_status_codes():
    assert T.lib().trw_replay_fill(1, 0, -1, 3, None) == A.E_INVALID
    assert b"bad argument" in T.lib().trw_last_error()
    with pytest.raises(T.TrwError):
        T.HostGraph.chung_lu(10, 1000, 2.1, 2, 5, 1)
    with pytest.raises(ValueError):
        T.make_params("metapath", schema=list(range(17)))


@pytest.mark.skipif(T.device_count() > 0, reason="checks the no-device path")
def test_context_without_device_reports_hip_error():
    with pytest.raises(T.TrwError) as e:
        T.Context(0)
    assert e.value.code == A.E_HIP


def test_walk_supported_names_the_library_moves():
    """trw_walk_supported: the (app, mode, sampler) pairs the library has Move kernels for;
    the rest of what the reference's apps accept is the drop-in functor engine's."""
    def rc(app, em, sm):
        return T.lib().trw_walk_supported(ctypes.byref(T.make_params(app, execution=em, sampling=sm)))

    import ctypes
    for app in ("deepwalk", "genericwalk", "ppr", "node2vec", "metapath"):
        assert rc(app, 0, 0) == A.OK  # uniform_move for every app
    for app in ("deepwalk", "genericwalk", "ppr"):
        for em in (1, 2):
            for sm in (1, 2, 3):
                assert rc(app, em, sm) == A.OK
            assert rc(app, em, 4) == A.E_UNSUPPORTED
            assert rc(app, em, 0) == A.E_UNSUPPORTED
    assert rc("node2vec", 2, 4) == A.OK
    for sm in (1, 2, 3):
        assert rc("node2vec", 2, sm) == A.E_UNSUPPORTED
        assert rc("node2vec", 1, sm) == A.E_UNSUPPORTED
    assert rc("metapath", 2, 1) == A.OK
    for sm in (2, 3, 4):
        assert rc("metapath", 2, sm) == A.E_UNSUPPORTED
    bad = T.make_params("deepwalk")
    bad.execution = 7
    assert T.lib().trw_walk_supported(ctypes.byref(bad)) == A.E_INVALID


def test_tuning_knob_ids_match_the_header():
    """Every TRW_TUNE_* knob the header defines has the same id in the Python mirror, the
    mirror's knob names (per-graph knobs and per-Move launch shapes) cover exactly those ids,
    and TRW_MOVE_* index the mirror's Move names in order."""
    text = open(HEADER).read()
    defs = dict(re.findall(r"^#define TRW_TUNE_(\w+)\s+(\d+)", text, re.M))
    assert defs, "no TRW_TUNE_* defines found"
    for name, val in defs.items():
        assert getattr(A, "TUNE_" + name) == int(val), name
    knobs = {int(v) for k, v in defs.items() if k != "KNOBS"}
    assert set(T.DeviceGraph.TUNING.values()) | set(T.DeviceGraph.SHAPES.values()) == knobs
    assert not set(T.DeviceGraph.TUNING.values()) & set(T.DeviceGraph.SHAPES.values())
    assert len(knobs) == int(defs["KNOBS"])
    moves = dict(re.findall(r"^#define TRW_MOVE_(\w+)\s+(\d+)", text, re.M))
    assert int(re.search(r"^#define TRW_MOVES\s+(\d+)", text, re.M).group(1)) == len(A.MOVES) == len(moves)
    assert {A.MOVES[int(v)]: int(v) for v in moves.values()} == {k.lower(): int(v) for k, v in moves.items()}
