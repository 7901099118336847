"""Pruned-graph SSSP on the GPU (CalcuSSSP + hidden seeding, src/sssp/purn/Main.cu:
67-90, 136-193): distances on the purn-sssp graph equal Dijkstra on the FULL graph."""
import numpy as np
import pytest

import c2g
import oracle
from conftest import check_sssp_certificate, random_graph, rmat_graph

pytestmark = pytest.mark.gpu


def _pruned(G):
    P = c2g.prune_sssp_mem(G["vlist"], G["elist"], G["wlist"])
    g = c2g.Graph.from_csr(P["vlist"], P["elist"], P["wlist"], hvlist=P["hvlist"], helist=P["helist"],
                           hwlist=P["hwlist"], variant=c2g.VARIANT_SSSP_PURN, device=0)
    return P, g


@pytest.mark.parametrize("delta", [0, 4])
def test_all_sources_sparse(gpu, delta):
    for seed in range(8):
        G = random_graph(60, 130, seed=seed, wmax=9)
        P, g = _pruned(G)
        assert len(P["helist"]) > 0
        with g:
            g.set_tune(delta=delta)
            srcs = np.arange(G["V"], dtype=np.int32)
            db, _ = g.sssp_batch(srcs)
            for s in range(G["V"]):
                ref = oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], s)
                d, st = g.sssp(s)
                assert np.array_equal(d, ref), (seed, s)
                assert np.array_equal(db[s], ref), (seed, s, "batched")


def test_rmat16_pruned(gpu):
    G = rmat_graph(16, 2, wmax=255)
    P, g = _pruned(G)
    deg = np.diff(G["vlist"])
    rng = np.random.default_rng(3)
    srcs = rng.choice(np.flatnonzero(deg > 0), 32, replace=False).astype(np.int32)
    # include processed (pruned) vertices as sources: their hidden lists matter
    srcs[:8] = P["porder"][:8]
    with g:
        db, _ = g.sssp_batch(srcs)
        for q, s in enumerate(srcs):
            ref = oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], int(s))
            d, _ = g.sssp(int(s))
            assert np.array_equal(d, ref), s
            assert np.array_equal(db[q], ref), s
        # migration lane on the pruned graph
        d, st = g.sssp_migrate(int(srcs[0]), 2)
        assert np.array_equal(d, oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], int(srcs[0])))


@pytest.mark.slow
def test_scale20_pruned_certificate(gpu):
    # C2's recipe at scale 20 (weights U[1,255]); exactness via the full-graph certificate.
    G = rmat_graph(20, 2, wmax=255)
    P, g = _pruned(G)
    deg = np.diff(G["vlist"])
    rng = np.random.default_rng(2)
    srcs = rng.choice(np.flatnonzero(deg > 0), 32, replace=False).astype(np.int32)
    with g:
        sums, reached, _ = g.sssp_batch_resident(srcs)
        for q in (0, 1, 2):
            d, _ = g.sssp(int(srcs[q]))
            check_sssp_certificate(G["vlist"], G["elist"], G["wlist"], int(srcs[q]), d)
            fin = d != np.iinfo(np.int32).max
            assert sums[q] == d[fin].astype(np.int64).sum() and reached[q] == fin.sum()


def test_site_graph_where_pruning_bites(gpu):
    # the web-like site graph keeps about 0.74 of its edges under purn-sssp:
    # single-source and batched pruned distances equal full-graph Dijkstra
    import importlib.util
    import os
    from conftest import ROOT
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    G = bench.site_graph(16, seed=9)
    P, g = _pruned(G)
    assert len(P["elist"]) < 0.8 * G["E"]
    deg = np.diff(G["vlist"])
    rng = np.random.default_rng(4)
    srcs = np.concatenate([P["porder"][:32], rng.choice(np.flatnonzero(deg > 0), 32, replace=False)]).astype(np.int32)
    refs = oracle.dijkstra_many(G["vlist"], G["elist"], G["wlist"], srcs)
    with g:
        db, _ = g.sssp_batch(srcs)
        for q, s in enumerate(srcs):
            assert np.array_equal(db[q], refs[q]), (int(s), "batched")
            assert np.array_equal(g.sssp(int(s))[0], refs[q]), int(s)
