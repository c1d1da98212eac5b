"""trw_walk_noalloc: the walk trw_walk_async runs, with no allocation (scratch in a fixed device
array).  Its results must equal trw_walk_async's bit for bit, on two calls in a row, the second
with fewer walkers (the array keeps the first call's contents).  All calls through the C ABI."""
import numpy as np
import pytest

import graphs
import thunderrw_amd as T
from thunderrw_amd import _abi as A

pytestmark = pytest.mark.gpu

# (app, execution, sampling, length, form, rng)
CASES = [
    ("deepwalk", "static", "alias", 24, "fused", "replay"),
    ("deepwalk", "static", "alias", 24, "queue", "philox"),
    ("deepwalk", "static", "its", 24, "step", "replay"),
    ("deepwalk", "static", "rejection", 24, "queue", "replay"),
    ("ppr", "uniform", "uniform", 1, "queue", "replay"),
    ("ppr", "uniform", "uniform", 1, "step", "philox"),
    ("node2vec", "dynamic", "max_weight_rejection", 24, "step", "philox"),
    ("node2vec", "dynamic", "max_weight_rejection", 24, "fused", "replay"),
    ("metapath", "dynamic", "its", 24, "queue", "philox"),
    ("metapath", "dynamic", "its", 24, "auto", "replay"),
]


@pytest.fixture(scope="module")
def setup(gpu_ctx):
    g = graphs.chung_lu(20000, 90000, 2.1, 2, 3000, 71)
    lbl = np.random.default_rng(5).integers(0, 5, g.ne).astype(np.int32)
    dg = T.DeviceGraph(gpu_ctx, g.degree, g.adj, g.weight, lbl)
    for sm in ("alias", "its", "rejection"):
        dg.build_tables(sm, "edge")
    dg.build_membership()
    live = np.nonzero(g.degree > 0)[0].astype(np.int32)
    rng = np.random.default_rng(17)
    src = rng.choice(live, 30000).astype(np.int32)
    yield dg, src
    dg.close()


def run_pair(dg, prm, src, K, KT, seed):
    n = len(src)
    draws = T.replay_fill(seed, n, K) if prm.rng == A.RNG_REPLAY else None
    tdraws = T.replay_fill(seed + 1, n, KT) if prm.rng == A.RNG_REPLAY else None
    a = T.walk(dg, prm, src, draws, tdraws, noalloc=True)
    b = T.walk(dg, prm, src, draws, tdraws)
    return a, b


@pytest.mark.parametrize("app,ex,sm,L,form,rng", CASES)
def test_noalloc_equals_walk_async_twice_second_smaller(setup, app, ex, sm, L, form, rng):
    dg, src = setup
    prm = T.make_params(app, execution=ex, sampling=sm, length=L, stop_probability=0.15, schema=[0, 1, 2, 3, 4],
                        rng=rng, seed=99, form=form)
    K, KT = 4 * L + 128, 96
    for n in (len(src), 7001):
        print("noalloc case", app, sm, form, rng, "walkers", n, flush=True)
        a, b = run_pair(dg, prm, src[:n], K, KT, 1234 + n)
        for key in ("paths", "lengths", "ends", "status", "draws_used"):
            assert np.array_equal(a[key], b[key]), (app, form, n, key)
        for key in ("steps", "gather_terminated", "dead_ends", "errors"):
            assert a[key] == b[key], (app, form, n, key)
        assert a["steps"] > 0


def test_noalloc_refuses_what_would_allocate(gpu_ctx, setup):
    dg, src = setup
    # STEP form beyond the fixed walker-state array: refused before any launch.
    prm = T.make_params("ppr", rng="philox", form="step")
    one = T.DeviceBuffer.from_array(gpu_ctx, src[:1])
    io = T.make_io(one, (1 << 18) + 1)
    with pytest.raises(T.TrwError) as e:
        T.walk_noalloc(dg, prm, io)
    assert e.value.code == A.E_INVALID
    one.free()
    # The GRAPH form instantiates a hipGraph.
    with pytest.raises(T.TrwError) as e:
        T.walk(dg, T.make_params("ppr", rng="philox", form="graph"), src[:100], noalloc=True)
    assert e.value.code == A.E_UNSUPPORTED
    # Tables are never built by this entry point.
    g = graphs.chung_lu(500, 1500, 2.1, 2, 60, 3)
    fresh = T.DeviceGraph(gpu_ctx, g.degree, g.adj, g.weight)
    try:
        with pytest.raises(T.TrwError) as e:
            T.walk(fresh, T.make_params("deepwalk", length=10, rng="philox"), np.zeros(4, np.int32), noalloc=True)
        assert e.value.code == A.E_STATE
    finally:
        fresh.close()
