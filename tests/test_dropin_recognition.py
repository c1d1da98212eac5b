"""The drop-in compute<F> recognises the reference's five functors (apps/{ppr,node2vec,
metapath,deepwalk,genericwalk}.cpp) by their members and their type name, so their walks run in the
library's Move kernels: the name may carry any namespace (named or anonymous) or enclosing class, and
an alias names the type it aliases; a functor with the same members under another name is not one of
them and goes to the functor engine.  Host compile and run only (g++; no GPU call)."""
import os
import subprocess
import tempfile

import graphs
from conftest import ROOT

SRC = r'''
#include <cstdio>
#include "creeper_main.h"

namespace app { struct PPR { double tp_ = 0.25; sfmt_t sfmt_; }; }
namespace { struct Node2Vec { int l_ = 80; double p_ = 0.5, q_ = 2.0; }; }
struct Outer { struct DeepWalk { int l_ = 80; }; };
namespace lib { namespace v2 { struct GenericWalk { int l_ = 80; }; } }
struct MetaPath { int l_ = 80; int schema_length_ = 3; int schema_[3] = {0, 1, 2}; };
using RenamedPPR = app::PPR;
struct PPRLike { double tp_ = 0.25; sfmt_t sfmt_; };    // PPR's members, another name
struct WalkNode2Vec { int l_ = 80; double p_ = 1, q_ = 1; };  // a name that only ends like one

template <class F>
void show(const char *label, const F &f) {
    trw_walk_params p;
    std::memset(&p, 0, sizeof(p));
    const bool builtin = trw_dropin::describe(f, p);
    std::printf("%s %d %d %.3f %.3f %.3f %d\n", label, builtin ? 1 : 0, builtin ? p.app : -1, p.stop_probability,
                p.p, p.q, p.schema_length);
}

// The drop-in's main() loads the graph (host side only) and calls the app's execute().
void execute(Graph &, InputParser &) {
    show("ppr", app::PPR{});
    show("alias", RenamedPPR{});
    show("node2vec", Node2Vec{});
    show("deepwalk", Outer::DeepWalk{});
    show("generic", lib::v2::GenericWalk{});
    show("metapath", MetaPath{});
    show("ppr_like", PPRLike{});
    show("walk_node2vec", WalkNode2Vec{});
}
'''


def test_reference_functors_are_recognised_whatever_their_namespace():
    import thunderrw_amd._abi as A
    with tempfile.TemporaryDirectory() as tmp:
        c = os.path.join(tmp, "recognise.cpp")
        open(c, "w").write(SRC)
        exe = os.path.join(tmp, "recognise")
        pkg = os.path.join(ROOT, "thunderrw_amd")
        subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(pkg, "dropin"), "-I",
                        os.path.join(ROOT, "include"), c, "-o", exe, "-L", pkg, "-lthunderrw",
                        "-Wl,-rpath," + pkg], check=True)
        gdir = os.path.join(tmp, "graph")
        os.makedirs(gdir)
        graphs.write_reference_csr(graphs.chung_lu(200, 600, 2.1, 2, 30, 5), gdir)
        out = subprocess.run([exe, "-f", gdir], check=True, capture_output=True, text=True).stdout.split("\n")
    labels = ("ppr", "alias", "node2vec", "deepwalk", "generic", "metapath", "ppr_like", "walk_node2vec")
    got = {f[0]: (int(f[1]), int(f[2]), float(f[3]), float(f[4]), float(f[5]), int(f[6]))
           for f in (ln.split() for ln in out) if f and f[0] in labels and len(f) == 7}
    assert got["ppr"] == (1, A.APP_PPR, 0.25, 0.0, 0.0, 0)
    assert got["alias"] == got["ppr"]
    assert got["node2vec"] == (1, A.APP_NODE2VEC, 0.0, 0.5, 2.0, 0)
    assert got["deepwalk"][:2] == (1, A.APP_DEEPWALK)
    assert got["generic"][:2] == (1, A.APP_DEEPWALK)
    assert got["metapath"] == (1, A.APP_METAPATH, 0.0, 0.0, 0.0, 3)
    assert got["ppr_like"][:2] == (0, -1)
    assert got["walk_node2vec"][:2] == (0, -1)
