"""Write the golden fixtures under tests/golden/ (run in the build container only).

Two sources, both reference-held or reference-computed:
  * sfmt_kat.npz: the known-answer vectors shipped with the vendored RNG,
    dependency/SFMT-src-1.5.1/SFMT.19937.out.txt (init_gen_rand(1234) and
    init_by_array({0x1234, 0x5678, 0x9abc, 0xdef0}), first 1000 words each;
    driver test.c:35-90).
  * graph_<name>.npz, tables_<name>_<app>.npz, walk_<case>.npz: outputs of the
    reference's own functions, run by oracle/_ref/harness_<app> (built from the
    reference sources by `make -C oracle ref`) on small seeded graphs.
  * functor_<case>.npz: the same harness driving functors the library has no Move
    kernel for (reference apps in other mode/sampler pairs, and
    tests/dropin_apps/customwalk.cpp), walkers in generate_all_source order.

Usage:  python tests/golden/make_golden.py [--ref /root/reference] [--only functor]
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
import graphs  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref")

# Replay contract: walker i's move draws are words [i*K, (i+1)*K) of the
# SFMT-19937 stream init_gen_rand(SEED); PPR termination draws are words
# [i*KT, (i+1)*KT) of init_gen_rand(TSEED).
SEED, TSEED = 1234, 4321

GRAPHS = {
    "small": dict(kind="chung_lu", n=200, m=600, gamma=2.1, dmin=2, dmax=60, seed=11),
    "hub": dict(kind="hub", n=3000, hub_degree=2500, extra=6000, seed=12),
}

# name: (graph, app, em, sm, length, extra-args dict, K, KT)
CASES = [
    ("deepwalk_static_alias_small", "small", "deepwalk", 1, 2, 20, {}, 256, 0),
    ("deepwalk_static_alias_hub", "hub", "deepwalk", 1, 2, 20, {}, 256, 0),
    ("deepwalk_static_its_small", "small", "deepwalk", 1, 1, 20, {}, 256, 0),
    ("deepwalk_static_its_hub", "hub", "deepwalk", 1, 1, 20, {}, 256, 0),
    ("deepwalk_static_rs_small", "small", "deepwalk", 1, 3, 20, {}, 256, 0),
    ("deepwalk_static_rs_hub", "hub", "deepwalk", 1, 3, 20, {}, 256, 0),
    ("deepwalk_uniform_small", "small", "deepwalk", 0, 0, 20, {}, 256, 0),
    ("deepwalk_uniform_hub", "hub", "deepwalk", 0, 0, 20, {}, 256, 0),
    ("deepwalk_dynamic_its_small", "small", "deepwalk", 2, 1, 20, {}, 256, 0),
    ("deepwalk_dynamic_alias_small", "small", "deepwalk", 2, 2, 20, {}, 256, 0),
    ("deepwalk_dynamic_rs_small", "small", "deepwalk", 2, 3, 20, {}, 256, 0),
    ("ppr_uniform_small", "small", "ppr", 0, 0, 1, {"-sp": "0.2"}, 64, 64),
    ("ppr_uniform_hub", "hub", "ppr", 0, 0, 1, {"-sp": "0.2"}, 64, 64),
    ("ppr_static_alias_small", "small", "ppr", 1, 2, 1, {"-sp": "0.15"}, 128, 64),
    ("node2vec_small", "small", "node2vec", 2, 4, 20, {"-p": "2", "-q": "0.5"}, 512, 0),
    ("node2vec_hub", "hub", "node2vec", 2, 4, 20, {"-p": "2", "-q": "0.5"}, 512, 0),
    ("node2vec_inv_small", "small", "node2vec", 2, 4, 20, {"-p": "0.5", "-q": "2"}, 512, 0),
    ("metapath_its_small", "small", "metapath", 2, 1, 20, {"-s": "0,1,2,3,4"}, 256, 0),
    ("metapath_its_hub", "hub", "metapath", 2, 1, 20, {"-s": "0,1,2,3,4"}, 256, 0),
    ("metapath_alias_small", "small", "metapath", 2, 2, 20, {"-s": "1,3"}, 256, 0),
]

# Walks of functors the library has no Move kernel for (the drop-in functor engine): the
# reference apps in (mode, sampler) pairs outside the library's kernels, and the test app
# tests/dropin_apps/customwalk.cpp.  Sources are generate_all_source's order, as the apps
# themselves walk them.  name: (graph, app, em, sm, length, extra-args dict, K)
FUNCTOR_CASES = [
    ("node2vec_dyn_its_small", "small", "node2vec", 2, 1, 20, {"-p": "2", "-q": "0.5"}, 64),
    ("node2vec_dyn_its_hub", "hub", "node2vec", 2, 1, 20, {"-p": "2", "-q": "0.5"}, 64),
    ("node2vec_dyn_alias_small", "small", "node2vec", 2, 2, 20, {"-p": "2", "-q": "0.5"}, 64),
    ("node2vec_dyn_alias_hub", "hub", "node2vec", 2, 2, 20, {"-p": "2", "-q": "0.5"}, 64),
    ("node2vec_dyn_rs_small", "small", "node2vec", 2, 3, 20, {"-p": "0.5", "-q": "2"}, 1024),
    ("node2vec_dyn_rs_hub", "hub", "node2vec", 2, 3, 20, {"-p": "2", "-q": "0.5"}, 1024),
    ("metapath_dyn_alias_small", "small", "metapath", 2, 2, 20, {"-s": "0,1,2,3,4"}, 64),
    ("metapath_dyn_alias_hub", "hub", "metapath", 2, 2, 20, {"-s": "0,1,2,3,4"}, 64),
    ("metapath_dyn_rs_small", "small", "metapath", 2, 3, 20, {"-s": "1,3"}, 1024),
    ("metapath_dyn_rs_hub", "hub", "metapath", 2, 3, 20, {"-s": "0,1,2,3,4"}, 1024),
    ("deepwalk_dyn_maxrs_small", "small", "deepwalk", 2, 4, 20, {}, 1024),
    ("deepwalk_dyn_maxrs_hub", "hub", "deepwalk", 2, 4, 20, {}, 1024),
    ("custom0_static_its_small", "small", "customwalk", 1, 1, 20, {"-kind": "0"}, 64),
    ("custom0_static_its_hub", "hub", "customwalk", 1, 1, 20, {"-kind": "0"}, 64),
    ("custom0_static_alias_small", "small", "customwalk", 1, 2, 20, {"-kind": "0"}, 64),
    ("custom0_static_alias_hub", "hub", "customwalk", 1, 2, 20, {"-kind": "0"}, 64),
    ("custom0_static_rs_small", "small", "customwalk", 1, 3, 20, {"-kind": "0"}, 4096),
    ("custom0_dyn_its_small", "small", "customwalk", 2, 1, 20, {"-kind": "0"}, 64),
    ("custom0_dyn_alias_small", "small", "customwalk", 2, 2, 20, {"-kind": "0"}, 64),
    ("custom0_dyn_rs_small", "small", "customwalk", 2, 3, 20, {"-kind": "0"}, 4096),
    ("custom1_dyn_its_small", "small", "customwalk", 2, 1, 20, {"-kind": "1"}, 64),
    ("custom1_dyn_its_hub", "hub", "customwalk", 2, 1, 20, {"-kind": "1"}, 64),
    ("custom1_dyn_alias_small", "small", "customwalk", 2, 2, 20, {"-kind": "1"}, 64),
    ("custom1_dyn_alias_hub", "hub", "customwalk", 2, 2, 20, {"-kind": "1"}, 64),
    ("custom1_dyn_rs_small", "small", "customwalk", 2, 3, 20, {"-kind": "1"}, 2048),
    ("custom1_dyn_rs_hub", "hub", "customwalk", 2, 3, 20, {"-kind": "1"}, 2048),
    ("custom1_dyn_maxrs_small", "small", "customwalk", 2, 4, 20, {"-kind": "1"}, 2048),
    ("custom1_dyn_maxrs_hub", "hub", "customwalk", 2, 4, 20, {"-kind": "1"}, 2048),
    ("custom2_dyn_its_small", "small", "customwalk", 2, 1, 20, {"-kind": "2"}, 64),
    ("custom2_dyn_its_hub", "hub", "customwalk", 2, 1, 20, {"-kind": "2"}, 64),
    ("custom2_dyn_alias_small", "small", "customwalk", 2, 2, 20, {"-kind": "2"}, 64),
    ("custom2_dyn_alias_hub", "hub", "customwalk", 2, 2, 20, {"-kind": "2"}, 64),
    ("custom2_dyn_rs_small", "small", "customwalk", 2, 3, 20, {"-kind": "2"}, 2048),
    ("custom2_dyn_rs_hub", "hub", "customwalk", 2, 3, 20, {"-kind": "2"}, 2048),
    ("custom2_uniform_small", "small", "customwalk", 0, 0, 20, {"-kind": "2"}, 64),
    ("custom2_uniform_hub", "hub", "customwalk", 0, 0, 20, {"-kind": "2"}, 64),
    ("custom4_dyn_maxrs_small", "small", "customwalk", 2, 4, 20, {"-kind": "4", "-p": "2", "-q": "0.5"}, 1024),
    ("custom7_dyn_maxrs_small", "small", "customwalk", 2, 4, 20, {"-kind": "7", "-p": "2", "-q": "0.5"}, 2048),
    ("custom7_dyn_maxrs_hub", "hub", "customwalk", 2, 4, 20, {"-kind": "7", "-p": "2", "-q": "0.5"}, 2048),
    ("custom7_inv_dyn_maxrs_small", "small", "customwalk", 2, 4, 20, {"-kind": "7", "-p": "0.5", "-q": "2"}, 4096),
]


def build_graph(spec):
    if spec["kind"] == "chung_lu":
        return graphs.chung_lu(spec["n"], spec["m"], spec["gamma"], spec["dmin"], spec["dmax"], spec["seed"])
    return graphs.hub_graph(spec["n"], spec["hub_degree"], spec["extra"], spec["seed"])


def sources_for(g, name):
    rng = np.random.default_rng(99)
    live = np.nonzero(g.degree > 0)[0].astype(np.int32)
    if name == "small":
        return live
    pick = rng.choice(live, size=300, replace=False).astype(np.int32)
    return np.concatenate([np.zeros(8, np.int32), pick])


def parse_kat(path):
    with open(path) as f:
        text = f.read()
    a, b = text.split("init_gen_rand__________")[1].split("init_by_array__________")
    gen = np.array([int(t) for t in a.split()[:1000]], dtype=np.uint32)
    arr = np.array([int(t) for t in b.split()[:1000]], dtype=np.uint32)
    assert len(gen) == 1000 and len(arr) == 1000
    return gen, arr


def run(cmd):
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)


def load(path, dtype):
    return np.fromfile(path, dtype=dtype)


def run_case(tmp, case, gdir, src, app, em, sm, length, extra, K, KT):
    out = os.path.join(tmp, case)
    os.makedirs(out)
    src_path = os.path.join(out, "src.bin")
    src.tofile(src_path)
    cmd = [os.path.join(REF_BIN, f"harness_{app}"), "-f", gdir, "-ew", "-el", "-o", out,
           "-mode", "walk", "-em", str(em), "-sm", str(sm), "-l", str(length), "-src", src_path,
           "-K", str(K), "-KT", str(max(KT, 1)), "-seed", str(SEED), "-tseed", str(TSEED)]
    for k, v in extra.items():
        cmd += [k, v]
    run(cmd)
    return dict(paths=load(os.path.join(out, "paths.bin"), np.int32).reshape(len(src), -1),
                lengths=load(os.path.join(out, "lengths.bin"), np.int32),
                ends=load(os.path.join(out, "ends.bin"), np.int32),
                status=load(os.path.join(out, "status.bin"), np.int32),
                draws_used=load(os.path.join(out, "draws_used.bin"), np.int32),
                steps=load(os.path.join(out, "steps.bin"), np.uint64))


def functor_fixtures(index):
    """functor_<case>.npz from the committed graphs (graph_<name>.npz), sources in
    generate_all_source order."""
    sys.path.insert(0, ROOT)
    import thunderrw_amd as T

    index["functor_walks"] = []
    with tempfile.TemporaryDirectory() as tmp:
        gdirs, gobjs = {}, {}
        for name in GRAPHS:
            z = np.load(os.path.join(HERE, f"graph_{name}.npz"))
            g = graphs.CSR(z["degree"], z["adj"], z["weight"], z["label"])
            d = os.path.join(tmp, "graph_" + name)
            graphs.write_reference_csr(g, d)
            gdirs[name], gobjs[name] = d, g
        for case, gname, app, em, sm, length, extra, K in FUNCTOR_CASES:
            src = T.sources_all(gobjs[gname].degree)
            r = run_case(tmp, case, gdirs[gname], src, app, em, sm, length, extra, K, 0)
            # Every walker's draws must fit its chunk (a walker that needs more reads its
            # neighbour's chunk in the harness, which is no replay contract).
            assert (r["draws_used"] >= 0).all(), case
            np.savez_compressed(os.path.join(HERE, f"functor_{case}.npz"), **r)
            index["functor_walks"].append({"case": case, "graph": gname, "app": app, "execution": em, "sampling": sm,
                                           "length": length, "extra": extra, "K": K})


# Walker sources (walker_generator.cpp:9-63) on the golden graphs: generate_all_source's shuffle
# (std::shuffle with std::mt19937(0), libstdc++'s algorithm), generate_random_source and
# generate_single_source, written by the reference's own functions.
SOURCE_ARGS = {"wn": 5000, "wseed": 7}


def source_fixtures(index):
    index["sources"] = dict(SOURCE_ARGS, graphs=list(GRAPHS))
    with tempfile.TemporaryDirectory() as tmp:
        for name in GRAPHS:
            z = np.load(os.path.join(HERE, f"graph_{name}.npz"))
            g = graphs.CSR(z["degree"], z["adj"], z["weight"], z["label"])
            d = os.path.join(tmp, "graph_" + name)
            graphs.write_reference_csr(g, d)
            out = os.path.join(tmp, "out_" + name)
            os.makedirs(out)
            run([os.path.join(REF_BIN, "harness_deepwalk"), "-f", d, "-ew", "-o", out, "-mode", "sources",
                 "-wn", str(SOURCE_ARGS["wn"]), "-wseed", str(SOURCE_ARGS["wseed"])])
            np.savez_compressed(os.path.join(HERE, f"sources_{name}.npz"),
                                **{k: load(os.path.join(out, f"sources_{k}.bin"), np.int32)
                                   for k in ("all", "random", "single")})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--only", choices=["functor", "sources"], help="write only these fixtures (index entries kept)")
    args = ap.parse_args()

    if args.only:
        with open(os.path.join(HERE, "index.json")) as f:
            index = json.load(f)
        if args.only == "functor":
            functor_fixtures(index)
        else:
            source_fixtures(index)
        with open(os.path.join(HERE, "index.json"), "w") as f:
            json.dump(index, f, indent=1, sort_keys=True)
        print("wrote", args.only, "fixtures")
        return

    gen, arr = parse_kat(os.path.join(args.ref, "dependency/SFMT-src-1.5.1/SFMT.19937.out.txt"))
    np.savez_compressed(os.path.join(HERE, "sfmt_kat.npz"), init_gen_rand_1234=gen,
                        init_by_array_1234_5678_9abc_def0=arr)

    index = {"seed": SEED, "tseed": TSEED, "graphs": {}, "tables": [], "walks": []}
    with tempfile.TemporaryDirectory() as tmp:
        gdirs, gobjs = {}, {}
        for name, spec in GRAPHS.items():
            g = build_graph(spec)
            d = os.path.join(tmp, name)
            graphs.write_reference_csr(g, d)
            gdirs[name], gobjs[name] = d, g
            np.savez_compressed(os.path.join(HERE, f"graph_{name}.npz"), degree=g.degree, adj=g.adj,
                                weight=g.weight, label=g.label, sources=sources_for(g, name))
            index["graphs"][name] = spec

            for app in ("deepwalk", "ppr"):
                out = os.path.join(tmp, f"tables_{name}_{app}")
                os.makedirs(out)
                run([os.path.join(REF_BIN, f"harness_{app}"), "-f", d, "-ew", "-el", "-o", out, "-mode", "tables",
                     "-l", "20"])
                g_alias = load(os.path.join(out, "alias.bin"), np.uint8).reshape(-1, 16)
                np.savez_compressed(os.path.join(HERE, f"tables_{name}_{app}.npz"), alias=g_alias,
                                    prefix=load(os.path.join(out, "prefix.bin"), np.float64),
                                    rs_weight=load(os.path.join(out, "rs_weight.bin"), np.float64),
                                    rs_max=load(os.path.join(out, "rs_max.bin"), np.float64))
                index["tables"].append({"graph": name, "app": app})

        for case, gname, app, em, sm, length, extra, K, KT in CASES:
            g = gobjs[gname]
            src = sources_for(g, gname)
            r = run_case(tmp, case, gdirs[gname], src, app, em, sm, length, extra, K, KT)
            np.savez_compressed(os.path.join(HERE, f"walk_{case}.npz"), **r)
            index["walks"].append({"case": case, "graph": gname, "app": app, "execution": em, "sampling": sm,
                                   "length": length, "extra": extra, "K": K, "KT": KT})

    functor_fixtures(index)
    source_fixtures(index)
    with open(os.path.join(HERE, "index.json"), "w") as f:
        json.dump(index, f, indent=1, sort_keys=True)
    print("wrote", len(index["walks"]), "walk fixtures,", len(index["functor_walks"]), "functor-engine walk fixtures and",
          len(index["tables"]), "table fixtures")


if __name__ == "__main__":
    main()
