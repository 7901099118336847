"""Light-edge mode of SSSP (c2g_tune.light) on the GPU: bit-exact against the
oracle's Dijkstra for every light bound, on full and pruned graphs, single
source, batched, budgeted (host tail) and precedence with delta / split.

A vertex lowered in the last round relaxes only its edges of weight < light
and its full row once a round passes without a new improvement.  The
relaxation rule itself is the reference's (src/sssp/base/Main.cu:217-221);
only the order in which edges are relaxed changes, so the distances are the
unique shortest-path distances whatever the bound.
"""
import numpy as np
import pytest

import c2g
import oracle
from conftest import random_graph, rmat_graph

pytestmark = pytest.mark.gpu


def _graph(G, **kw):
    return c2g.Graph.from_csr(G["vlist"], G["elist"], G["wlist"], device=0, **kw)


def _sources(G, n, seed):
    deg = np.diff(G["vlist"])
    rng = np.random.default_rng(seed)
    hubs = np.argsort(-deg, kind="stable")[: max(1, n // 4)]
    ones = np.flatnonzero(deg == 1)
    rest = rng.choice(np.flatnonzero(deg > 0), n - len(hubs) - min(len(ones), n // 4), replace=False)
    return np.concatenate([hubs, ones[: n // 4], rest]).astype(np.int32)


@pytest.mark.parametrize("light", [0, 1, 3, 16, 64, 1 << 20])
def test_light_bounds_equal_dijkstra(gpu, light):
    # every bound, from none (0) to every edge light (2^20): single-source for
    # three sources and a ragged two-group batch of 70
    G = rmat_graph(15, 3, wmax=255)
    srcs = _sources(G, 70, light)
    refs = oracle.dijkstra_many(G["vlist"], G["elist"], G["wlist"], srcs)
    with _graph(G) as g:
        g.set_tune(light=light)
        assert g.tune().light == light
        for q in (0, 20, 69):
            d, st = g.sssp(int(srcs[q]))
            assert np.array_equal(d, refs[q]), (light, q)
        db, _ = g.sssp_batch(srcs)
        for q in range(len(srcs)):
            assert np.array_equal(db[q], refs[q]), (light, q)


def test_light_mode_relaxes_fewer_edges(gpu):
    # the point of the mode: on R-MAT with uniform weights a bound of about
    # 1/16 of the mean weight cuts relaxations (single source) and (query,
    # edge) relaxations (batched) well below plain rounds, same distances
    G = rmat_graph(16, 4, wmax=1023)
    srcs = _sources(G, 64, 4)
    with _graph(G) as g:
        assert g.tune().light == 32  # mean weight 512 / 16
        out = {}
        for light in (0, 32):
            g.set_tune(light=light)
            d, st = g.sssp(int(srcs[0]))
            sums, reach, sb = g.sssp_batch_resident(srcs)
            out[light] = (d, st.edges_relaxed, sums, reach, sb.edges_relaxed, sb.csr_edges)
    d0, e0, s0, r0, b0, c0 = out[0]
    d1, e1, s1, r1, b1, c1 = out[32]
    assert np.array_equal(d0, d1) and np.array_equal(s0, s1) and np.array_equal(r0, r1)
    assert e1 < 0.75 * e0, (e1, e0)
    assert b1 < 0.6 * b0 and c1 < 0.6 * c0, (b1, b0, c1, c0)


def test_light_zero_weights_and_pruned_graph(gpu):
    # weight-0 edges are light; a purn-sssp graph keeps its hidden lists out
    # of the mode (they are relaxed with every improvement, as without it)
    G = random_graph(400, 2400, seed=8, wmax=12, zero_ok=True)
    assert (G["wlist"] == 0).any()
    ss = np.arange(0, 400, 9, dtype=np.int32)
    refs = oracle.dijkstra_many(G["vlist"], G["elist"], G["wlist"], ss)
    with _graph(G) as g:
        for light in (1, 4):
            g.set_tune(light=light)
            db, _ = g.sssp_batch(ss)
            for q, s in enumerate(ss):
                assert np.array_equal(g.sssp(int(s))[0], refs[q]) and np.array_equal(db[q], refs[q]), (light, s)
    H = random_graph(300, 800, seed=5, wmax=9)
    P = c2g.prune_sssp_mem(H["vlist"], H["elist"], H["wlist"])
    assert len(P["helist"]) > 0
    hs = np.arange(300, dtype=np.int32)
    hrefs = oracle.dijkstra_many(H["vlist"], H["elist"], H["wlist"], hs)
    with c2g.Graph.from_csr(P["vlist"], P["elist"], P["wlist"], hvlist=P["hvlist"], helist=P["helist"],
                            hwlist=P["hwlist"], variant=c2g.VARIANT_SSSP_PURN, device=0) as g:
        for light in (2, 5):
            g.set_tune(light=light)
            db, _ = g.sssp_batch(hs)
            for s in range(0, 300, 7):
                assert np.array_equal(g.sssp(s)[0], hrefs[s]) and np.array_equal(db[s], hrefs[s]), (light, s)
            assert all(np.array_equal(db[s], hrefs[s]) for s in range(300))


@pytest.mark.parametrize("light", [4, 40])
def test_light_budget_hands_over_pending_vertices(gpu, light):
    # a budgeted run stops with vertices whose full rows are still pending;
    # they are part of the handed-over frontier, so the host tail converges
    G = rmat_graph(14, 6, wmax=64)
    src = int(np.argmax(np.diff(G["vlist"])))
    ref = oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], src)
    with _graph(G) as g:
        g.set_tune(light=light)
        for budget in (1, 2, 3, 4, 6, 100):
            d, fr, st = g.sssp_budget(src, budget)
            assert st.iterations <= budget and np.all(d >= ref)
            d2, _ = g.host_bellman_ford(d, fr)
            assert np.array_equal(d2, ref), budget
            d3, _ = g.sssp_migrate(src, budget)
            assert np.array_equal(d3, ref), budget


def test_light_precedence_and_rebuild(gpu):
    # delta > 0 or the XCD split turn the mode off (same counters as light 0);
    # the light rows are rebuilt when the bound changes, and a graph of unit
    # weights starts with the mode off
    G = rmat_graph(14, 7, wmax=255)
    s = int(np.argmax(np.diff(G["vlist"])))
    ref = oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], s)
    with _graph(G) as g:
        for kw in ({"delta": 32}, {"xcd_split": 1}):
            g.set_tune(light=0, **kw)
            d0, st0 = g.sssp(s)
            g.set_tune(light=16, **kw)
            d1, st1 = g.sssp(s)
            assert np.array_equal(d0, ref) and np.array_equal(d1, ref)
            assert st0.edges_relaxed == st1.edges_relaxed and st0.iterations == st1.iterations, kw
            g.set_tune(delta=0, xcd_split=0)
        for light in (16, 3, 0, 16, 200):
            g.set_tune(light=light)
            assert np.array_equal(g.sssp(s)[0], ref), light
            assert np.array_equal(g.sssp_batch(np.array([s, 0, 5], np.int32))[0][0], ref), light
        with pytest.raises(c2g.C2GError):
            g.set_tune(light=-1)
    U = dict(G)
    U["wlist"] = np.ones_like(G["wlist"])
    with _graph(U) as g:
        assert g.tune().light == 0
        d, _ = g.sssp(s)
    assert np.array_equal(d, oracle.dijkstra(U["vlist"], U["elist"], U["wlist"], s))
