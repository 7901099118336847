"""Batched multi-source SSSP (lane = query) on the GPU: every query's distances
must be bit-identical to the oracle's Dijkstra and to the single-source path.
"""
import numpy as np
import pytest

import c2g
import oracle
from conftest import INT_MAX, check_sssp_certificate, random_graph, rmat_graph

pytestmark = pytest.mark.gpu


def _graph(G):
    return c2g.Graph.from_csr(G["vlist"], G["elist"], G["wlist"], device=0)


@pytest.mark.parametrize("delta", [0, 32])
def test_batch_c1_against_dijkstra(gpu, delta):
    G = rmat_graph(16, 1, wmax=64)
    deg = np.diff(G["vlist"])
    rng = np.random.default_rng(7)
    srcs = np.concatenate([[np.argmax(deg)], rng.choice(np.flatnonzero(deg > 0), 40),
                           rng.choice(np.flatnonzero(deg == 1), 20), rng.choice(np.flatnonzero(deg == 0), 3)])
    with _graph(G) as g:
        g.set_tune(delta=delta)
        d, st = g.sssp_batch(srcs)
    assert st.lane == c2g.LANE_GPU
    for q, s in enumerate(srcs):
        ref = oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], int(s))
        assert np.array_equal(d[q], ref), (q, s)


def test_batch_multi_group_matches_single(gpu):
    # 150 queries = 3 groups (64, 64, 22 lanes), duplicate sources included
    G = rmat_graph(16, 1, wmax=64)
    deg = np.diff(G["vlist"])
    rng = np.random.default_rng(8)
    srcs = rng.choice(np.flatnonzero(deg > 0), 150)
    srcs[5] = srcs[70] = srcs[140]
    with _graph(G) as g:
        d, _ = g.sssp_batch(srcs)
        sums, reached, _ = g.sssp_batch_resident(srcs)
        sums_e, reached_e, _ = g.sssp_each_resident(srcs)
        for q in range(0, 150, 7):
            ds, _ = g.sssp(int(srcs[q]))
            assert np.array_equal(d[q], ds), q
    fin = d != INT_MAX
    assert np.array_equal(sums, np.where(fin, d, 0).sum(axis=1))
    assert np.array_equal(reached, fin.sum(axis=1))
    assert np.array_equal(sums, sums_e) and np.array_equal(reached, reached_e)


@pytest.mark.parametrize("chunk,delta", [(1, 0), (5, 7), (32, 0), (64, 40), (200, 1)])
def test_batch_chunk_sizes(gpu, chunk, delta):
    G = random_graph(300, 3000, seed=chunk, wmax=20)
    srcs = np.arange(0, 300, 3)
    with _graph(G) as g:
        g.set_tune(batch_chunk=chunk, delta=delta)
        d, _ = g.sssp_batch(srcs)
    for q, s in enumerate(srcs):
        assert np.array_equal(d[q], oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], int(s))), (chunk, s)


def test_batch_graph_replay_off(gpu):
    G = random_graph(500, 4000, seed=2, wmax=9)
    srcs = np.arange(64)
    with _graph(G) as g:
        g.set_tune(use_graph=0, iters_per_graph=2)
        d0, _ = g.sssp_batch(srcs)
        g.set_tune(use_graph=1, iters_per_graph=16)
        d1, _ = g.sssp_batch(srcs)
        g.set_tune(profile=1)
        d2, st = g.sssp_batch(srcs)
    assert np.array_equal(d0, d1) and np.array_equal(d0, d2)
    assert st.kernel_ms > 0


def test_batch_errors(gpu):
    G = random_graph(20, 40, seed=1)
    with _graph(G) as g:
        with pytest.raises(c2g.C2GError) as e:
            g.sssp_batch(np.array([1, 20]))
        assert e.value.code == c2g.EINVAL


@pytest.mark.slow
@pytest.mark.parametrize("delta", [0, 256])
def test_batch_scale20_certificate(gpu, delta):
    # uneven reach: half hubs, half out-degree-1 sources (the C4 recipe at scale 20)
    G = rmat_graph(20, 4, wmax=1023)
    deg = np.diff(G["vlist"])
    rng = np.random.default_rng(4)
    top = np.argsort(-deg)[: len(deg) // 100]
    srcs = np.concatenate([rng.choice(top, 64, replace=False), rng.choice(np.flatnonzero(deg == 1), 64)])
    with _graph(G) as g:
        g.set_tune(delta=delta)
        d, st = g.sssp_batch(srcs)
        sums_e, reached_e, _ = g.sssp_each_resident(srcs)
    fin = d != INT_MAX
    assert np.array_equal(np.where(fin, d, 0).sum(axis=1), sums_e)
    assert np.array_equal(fin.sum(axis=1), reached_e)
    for q in (0, 1, 63, 64, 100, 127):
        check_sssp_certificate(G["vlist"], G["elist"], G["wlist"], int(srcs[q]), d[q])
    ref = oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], int(srcs[0]))
    assert np.array_equal(d[0], ref)


def test_retune_between_runs_takes_effect(gpu):
    # captured launch sequences must follow set_tune (they hold kernel arguments)
    G = rmat_graph(16, 1, wmax=64)
    srcs = np.flatnonzero(np.diff(G["vlist"]) > 0)[:64].astype(np.int32)
    with _graph(G) as g:
        _, _, s0 = g.sssp_batch_resident(srcs)
        g.set_tune(delta=8)
        sums, reached, s1 = g.sssp_batch_resident(srcs)
        d, st = g.sssp(int(srcs[0]))
    assert s1.iterations > s0.iterations  # narrow buckets need more rounds
    assert np.array_equal(d, oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], int(srcs[0])))


def test_batch_duplicate_edges(gpu):
    # repeated (u, v) edges with different weights (graphs given as CSR are
    # not deduplicated): several lanes and steps update the same distance
    rng = np.random.default_rng(3)
    V, E = 400, 6000
    u = rng.integers(0, V, E)
    v = rng.integers(0, V, E)
    keep = u != v
    u, v = np.concatenate([u[keep], u[keep]]), np.concatenate([v[keep], v[keep]])
    w = rng.integers(1, 50, len(u))
    order = np.lexsort((v, u))
    u, v, w = u[order], v[order], w[order]
    vl = np.zeros(V + 1, np.int64)
    np.add.at(vl, u + 1, 1)
    G = dict(V=V, E=len(v), vlist=np.cumsum(vl).astype(np.int32), elist=v.astype(np.int32),
             wlist=w.astype(np.int32))
    srcs = np.arange(0, V, 5).astype(np.int32)
    with _graph(G) as g:
        d, _ = g.sssp_batch(srcs)
        ds, _ = g.sssp(int(srcs[0]))
    assert np.array_equal(ds, d[0])
    for q, s in enumerate(srcs):
        assert np.array_equal(d[q], oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], int(s))), q


@pytest.mark.parametrize("light", [0, 16])
def test_lane_order_keeps_caller_order(gpu, light):
    # batch_order 1 assigns queries to lanes by their sources' two-hop reach;
    # distances, checksums and reached counts come back in the caller's order,
    # identical to the caller's lane order (ragged 150 sources, 3 groups)
    G = rmat_graph(15, 9, wmax=255)
    deg = np.diff(G["vlist"])
    rng = np.random.default_rng(9)
    srcs = np.concatenate([np.argsort(-deg)[:50], rng.choice(np.flatnonzero(deg > 0), 100, replace=False)])
    srcs = rng.permutation(srcs).astype(np.int32)
    with c2g.Graph.from_csr(G["vlist"], G["elist"], G["wlist"], device=0) as g:
        assert g.tune().batch_order == 1
        g.set_tune(light=light)
        d1, s1 = g.sssp_batch(srcs)
        sums1, reach1, _ = g.sssp_batch_resident(srcs)
        g.set_tune(batch_order=0)
        d0, s0 = g.sssp_batch(srcs)
        sums0, reach0, _ = g.sssp_batch_resident(srcs)
        with pytest.raises(c2g.C2GError):
            g.set_tune(batch_order=2)
    assert np.array_equal(d1, d0)
    assert np.array_equal(sums1, sums0) and np.array_equal(reach1, reach0)
    assert s1.reached == s0.reached == int(reach0.sum())
    refs = oracle.dijkstra_many(G["vlist"], G["elist"], G["wlist"], srcs[[0, 75, 149]])
    for k, q in enumerate((0, 75, 149)):
        assert np.array_equal(d1[q], refs[k])


def test_consecutive_batches_reset_only_what_they_wrote(gpu):
    # after a converged batch only the rows of vertices with an in-edge and
    # the last sources' entries are reset: consecutive batches with fewer and
    # more groups, sources without in-edges (reachable only as sources), and a
    # batch after a served query, all against Dijkstra
    G = rmat_graph(14, 12, wmax=200)
    V = G["V"]
    indeg = np.bincount(G["elist"], minlength=V)
    deg = np.diff(G["vlist"])
    no_in = np.flatnonzero((indeg == 0) & (deg > 0))
    assert len(no_in) > 100
    rng = np.random.default_rng(12)
    runs = [np.concatenate([no_in[:40], rng.choice(V, 90, replace=False)]),
            rng.choice(V, 30, replace=False),
            np.concatenate([no_in[40:140], rng.choice(np.flatnonzero(deg > 0), 100, replace=False)])]
    with c2g.Graph.from_csr(G["vlist"], G["elist"], G["wlist"], device=0) as g:
        for srcs in runs:
            srcs = srcs.astype(np.int32)
            d, st = g.sssp_batch(srcs)
            sums, reach, _ = g.sssp_batch_resident(srcs)
            refs = oracle.dijkstra_many(G["vlist"], G["elist"], G["wlist"], srcs)
            for q in range(len(srcs)):
                assert np.array_equal(d[q], refs[q]), (len(srcs), q)
                fin = refs[q] != INT_MAX
                assert sums[q] == int(refs[q][fin].astype(np.int64).sum()) and reach[q] == int(fin.sum()), q
            m = sum(int(deg[refs[q] != INT_MAX].sum()) for q in range(len(srcs)))
            assert st.reached_edges == m
        # the server's lanes use group 0 of the same workspace
        with c2g.Server(g, sssp_slots=-1) as srv:  # lane group only
            qid = srv.submit(c2g.Q_SSSP, int(no_in[5]))
            _, out = srv.poll(qid, -1)
        srcs = runs[0].astype(np.int32)
        d, _ = g.sssp_batch(srcs)
    refs = oracle.dijkstra_many(G["vlist"], G["elist"], G["wlist"], srcs)
    assert all(np.array_equal(d[q], refs[q]) for q in range(len(srcs)))
    assert np.array_equal(out, oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], int(no_in[5])))


def test_item_size_and_light_changes_between_batches(gpu):
    # the expand counts a block's items from per-vertex bitmaps built for one
    # item size and light bound (rows of more than one item, light rows of
    # more than one item); every change between batches must rebuild them
    G = rmat_graph(15, 16, wmax=1023)
    deg = np.diff(G["vlist"])
    assert (deg > 200).sum() > 10  # hub rows span many items at every size below
    rng = np.random.default_rng(21)
    srcs = np.concatenate([np.argsort(deg)[-40:], rng.choice(np.flatnonzero(deg > 0), 60)]).astype(np.int32)
    refs = oracle.dijkstra_many(G["vlist"], G["elist"], G["wlist"], srcs)
    with _graph(G) as g:
        for chunk, light in [(64, 32), (7, 32), (7, 300), (200, 0), (1, 5), (64, 32)]:
            g.set_tune(batch_chunk=chunk, light=light)
            d, st = g.sssp_batch(srcs)
            sums, reach, _ = g.sssp_batch_resident(srcs)
            for q in range(len(srcs)):
                assert np.array_equal(d[q], refs[q]), (chunk, light, q)
                fin = refs[q] != INT_MAX
                assert sums[q] == int(refs[q][fin].astype(np.int64).sum()) and reach[q] == int(fin.sum())


def test_relabel_toggles_between_batches(gpu):
    # the batched kernels run on a copy of the graph with the vertices that
    # have an in-edge first (batch_relabel = 1) or on the graph itself (0):
    # toggling between batches, with light-bound and item-size changes in
    # between, must give Dijkstra's distances, checksums and edge counts in
    # the caller's vertex ids, and so must the server's lanes afterwards
    G = rmat_graph(15, 12, wmax=500)
    V = G["V"]
    deg = np.diff(G["vlist"])
    indeg = np.bincount(G["elist"], minlength=V)
    no_in = np.flatnonzero((indeg == 0) & (deg > 0))
    rng = np.random.default_rng(33)
    srcs = np.concatenate([no_in[:30], np.argsort(deg)[-20:], rng.choice(V, 50, replace=False)]).astype(np.int32)
    refs = oracle.dijkstra_many(G["vlist"], G["elist"], G["wlist"], srcs)
    m = sum(int(deg[refs[q] != INT_MAX].sum()) for q in range(len(srcs)))
    with _graph(G) as g:
        assert g.tune().batch_relabel == 1
        for relabel, light, chunk in [(1, 30, 64), (0, 30, 64), (1, 30, 64), (1, 0, 64), (1, 90, 9),
                                      (0, 90, 9), (1, 30, 64)]:
            g.set_tune(batch_relabel=relabel, light=light, batch_chunk=chunk)
            d, st = g.sssp_batch(srcs)
            sums, reach, _ = g.sssp_batch_resident(srcs)
            assert st.reached_edges == m
            for q in range(len(srcs)):
                assert np.array_equal(d[q], refs[q]), (relabel, light, chunk, q)
                fin = refs[q] != INT_MAX
                assert sums[q] == int(refs[q][fin].astype(np.int64).sum()) and reach[q] == int(fin.sum())
        with pytest.raises(c2g.C2GError):
            g.set_tune(batch_relabel=2)
        for relabel in (1, 0):
            g.set_tune(batch_relabel=relabel)
            with c2g.Server(g, sssp_slots=-1, max_batch=16) as srv:  # lane group only
                ids = [srv.submit(c2g.Q_SSSP, int(s)) for s in srcs[:40]]
                for k, i in enumerate(ids):
                    _, out = srv.poll(i, -1)
                    assert np.array_equal(out, refs[k]), (relabel, k)


@pytest.mark.parametrize("relabel", [0, 1])
def test_relabel_pruned_graph(gpu, relabel):
    # hidden lists are relabelled with the live rows; processed (pruned)
    # vertices among the sources
    G = rmat_graph(14, 3, wmax=255)
    P = c2g.prune_sssp_mem(G["vlist"], G["elist"], G["wlist"])
    assert len(P["helist"]) > 0
    deg = np.diff(G["vlist"])
    rng = np.random.default_rng(5)
    srcs = rng.choice(np.flatnonzero(deg > 0), 80, replace=False).astype(np.int32)
    srcs[:10] = P["porder"][:10]
    refs = oracle.dijkstra_many(G["vlist"], G["elist"], G["wlist"], srcs)
    with c2g.Graph.from_csr(P["vlist"], P["elist"], P["wlist"], hvlist=P["hvlist"], helist=P["helist"],
                            hwlist=P["hwlist"], variant=c2g.VARIANT_SSSP_PURN, device=0) as g:
        g.set_tune(batch_relabel=relabel)
        for _ in range(2):
            d, _ = g.sssp_batch(srcs)
            for q in range(len(srcs)):
                assert np.array_equal(d[q], refs[q]), (relabel, q)
