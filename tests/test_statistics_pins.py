"""Pins the free-running checks of tests/test_statistics.py before they judge the GPU: run on the
CPU oracle's walks (replay SFMT draws, the reference's Moves), the checks pass on the correct
walks and fail on walks with a planted mistake -- a wrong law, draws that repeat every step of a
walker (a step counter that does not advance), draws shared by every walker (a walker id that
does not key the stream)."""
import numpy as np
import pytest

import oracle_bind as ob
import test_statistics as S
import thunderrw_amd as T


@pytest.fixture(scope="module")
def oracle_graph():
    hg = T.HostGraph.chung_lu(1 << 20, 2 << 20, 2.1, 2.0, 4096.0, 61, True, 5)
    g = S.Csr(hg)
    hg.close()

    class View:
        nv, ne = g.nv, len(g.adj)
        offset, adj, degree, weight, label = g.off, g.adj, g.deg, g.w, g.lab

    alias, _ = ob.alias_build(View, g.w)
    return g, ob.Graph(View, alias)


def walks(og, draws, n, L=80):
    src = T.sources_random(og.g.degree, n, seed=71)
    return ob.walk(og, "deepwalk", 1, 2, L, src, draws, draws.shape[1])["paths"]


N = 1 << 19
K = 2 * 79


def test_checks_pass_on_the_oracle_walks(oracle_graph):
    g, og = oracle_graph
    paths = walks(og, ob.sfmt_stream(7, N * K).reshape(N, K), N)
    S.check_first_order(g, paths, S.first_order_laws(g, "alias"), "oracle alias")


def test_a_wrong_law_fails(oracle_graph):
    g, og = oracle_graph
    paths = walks(og, ob.sfmt_stream(8, N * K).reshape(N, K), N)
    with pytest.raises(AssertionError, match="pooled chi-square"):
        S.check_first_order(g, paths, S.first_order_laws(g, "uniform"), "alias walks judged as uniform")


def test_draws_repeating_every_step_fail_the_keying_check(oracle_graph):
    g, og = oracle_graph
    d = ob.sfmt_stream(9, N * 2).reshape(N, 2)
    paths = walks(og, np.tile(d, (1, 79)), N)
    law = S.first_order_laws(g, "alias")
    # A walker back at a vertex repeats its move there: the return moves agree every time (and the
    # counts are overdispersed).
    with pytest.raises(AssertionError, match="keying within walkers"):
        S.check_keying_within_walkers(g, paths, law, "repeat")
    with pytest.raises(AssertionError, match="pooled chi-square"):
        S.check_law(g, paths, law, "repeat")


def test_draws_shared_by_every_walker_fail_the_keying_check(oracle_graph):
    g, og = oracle_graph
    d = ob.sfmt_stream(10, K).reshape(1, K)
    # Shared draws still give each walker a valid walk; within a walker steps differ, across
    # walkers from one source the first moves always agree.
    paths = walks(og, np.tile(d, (N, 1)), N)
    law = S.first_order_laws(g, "alias")
    with pytest.raises(AssertionError, match="keying across walkers"):
        S.check_keying_across_walkers(g, paths, law, "shared")
    with pytest.raises(AssertionError, match="pooled chi-square"):
        S.check_law(g, paths, law, "shared")
