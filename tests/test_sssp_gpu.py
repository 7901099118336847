"""SSSP on the GPU through the C-ABI: bit-exact against the oracle's Dijkstra.

Reference semantics: CalcuSSSP (src/sssp/base/Main.cu:109-356) converges to the
unique shortest-path distances; unreached vertices are INT32_MAX here (the
reference writes 1999999999, Main.cu:121).
"""
import numpy as np
import pytest

import c2g
import oracle
from conftest import INT_MAX, check_sssp_certificate, random_graph, rmat_graph

pytestmark = pytest.mark.gpu


def _graph(G):
    return c2g.Graph.from_csr(G["vlist"], G["elist"], G["wlist"], device=0)


@pytest.mark.parametrize("delta", [0, 16, 64])
def test_c1_max_degree_source(gpu, delta):
    # Workload C1: R-MAT scale 16, ef 16, seed 1, weights U[1,64], max out-degree source.
    G = rmat_graph(16, 1, wmax=64)
    src = int(np.argmax(np.diff(G["vlist"])))
    with _graph(G) as g:
        g.set_tune(delta=delta)
        d, st = g.sssp(src)
    ref = oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], src)
    assert np.array_equal(d, ref)
    assert st.lane == c2g.LANE_GPU and st.iterations > 0
    assert st.reached == int((ref != INT_MAX).sum())


@pytest.mark.parametrize("seed", [11, 12, 13])
def test_c1_random_sources(gpu, seed):
    G = rmat_graph(16, 1, wmax=64)
    rng = np.random.default_rng(seed)
    srcs = rng.integers(0, G["V"], 4)
    with _graph(G) as g:
        for s in srcs:
            d, _ = g.sssp(int(s))
            assert np.array_equal(d, oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], int(s))), s


def test_tiny_and_edge_cases(gpu):
    # isolated source, zero-out-degree source, two-vertex graph, zero weights
    G = random_graph(50, 120, seed=3, wmax=5, zero_ok=True)
    with _graph(G) as g:
        for s in range(G["V"]):
            d, _ = g.sssp(s)
            assert np.array_equal(d, oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], s)), s
    vl = np.array([0, 0, 0], np.int32)
    with c2g.Graph.from_csr(vl, np.zeros(0, np.int32), np.zeros(0, np.int32), device=0) as g:
        d, st = g.sssp(1)
        assert d.tolist() == [INT_MAX, 0] and st.reached == 1


def test_bad_source_is_error(gpu):
    G = random_graph(20, 40, seed=1)
    with _graph(G) as g:
        with pytest.raises(c2g.C2GError) as e:
            g.sssp(20)
        assert e.value.code == c2g.EINVAL


def test_hub_rows_split_across_chunks(gpu):
    # a star with 10k leaves + chain: exercises items split at `chunk` edges
    n = 10001
    vl = np.zeros(n + 1, np.int32)
    vl[1:] = n - 1
    el = np.arange(1, n, dtype=np.int32)
    w = (np.arange(1, n) % 7 + 1).astype(np.int32)
    with c2g.Graph.from_csr(vl, el, w, device=0) as g:
        for chunk in (1, 7, 64, 4096):
            for delta in (0, 3):
                g.set_tune(chunk=chunk, delta=delta)
                d, _ = g.sssp(0)
                assert d[0] == 0 and np.array_equal(d[1:], w), (chunk, delta)


@pytest.mark.slow
@pytest.mark.parametrize("delta", [0, 300])
def test_scale20_against_dijkstra(gpu, delta):
    G = rmat_graph(20, 4, wmax=1023)
    deg = np.diff(G["vlist"])
    rng = np.random.default_rng(5)
    srcs = [int(np.argmax(deg))] + [int(x) for x in rng.choice(np.flatnonzero(deg > 0), 3)]
    with _graph(G) as g:
        g.set_tune(delta=delta)
        for s in srcs:
            d, _ = g.sssp(s)
            assert np.array_equal(d, oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], s)), s


@pytest.mark.slow
def test_scale22_certificate(gpu):
    # Workload C2's graph size (scale 22, weights U[1,255]); exactness by certificate.
    G = rmat_graph(22, 2, wmax=255)
    deg = np.diff(G["vlist"])
    rng = np.random.default_rng(2)
    srcs = rng.choice(np.flatnonzero(deg > 0), 3)
    with _graph(G) as g:
        for s in srcs:
            d, _ = g.sssp(int(s))
            check_sssp_certificate(G["vlist"], G["elist"], G["wlist"], int(s), d)


@pytest.mark.parametrize("delta", [0, 20])
def test_budget_and_migration(gpu, delta):
    # sssp-purn-M semantics: GPU head for `budget` rounds, host Bellman-Ford tail;
    # with delta-stepping the deferred vertices are part of the handed-over frontier
    G = rmat_graph(16, 1, wmax=64)
    src = int(np.argmax(np.diff(G["vlist"])))
    ref = oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], src)
    with _graph(G) as g:
        g.set_tune(delta=delta)
        for budget in (1, 2, 3, 5, 100):
            d, fr, st = g.sssp_budget(src, budget)
            assert st.iterations <= budget
            assert np.all(d >= ref)
            d2, _ = g.host_bellman_ford(d, fr)
            assert np.array_equal(d2, ref), budget
            d3, st3 = g.sssp_migrate(src, budget)
            assert np.array_equal(d3, ref)


@pytest.mark.slow
def test_multi_chunk_distance_copy(gpu):
    # binary tree on V = 17,000,000 vertices (i -> 2i+1, 2i+2, weight 1 + c % 7):
    # 68 MB of distances leave the device in two pinned 32 MiB chunks and a
    # partial third that reuses the first buffer; every distance has a closed form
    V = 17_000_000
    nchild = np.clip(V - 1 - 2 * np.arange(V, dtype=np.int64), 0, 2)
    vl = np.concatenate([[0], np.cumsum(nchild)]).astype(np.int32)
    el = np.arange(1, V, dtype=np.int32)
    wl = (1 + el % 7).astype(np.int32)
    want = np.zeros(V, np.int64)
    lo = 1
    while lo < V:  # level by level: parents come before children
        hi = min(V, 2 * lo + 1)
        c = np.arange(lo, hi)
        want[c] = want[(c - 1) // 2] + 1 + c % 7
        lo = hi
    with c2g.Graph.from_csr(vl, el, wl, device=0) as g:
        d, st = g.sssp(0)
    assert 2 * (32 << 20) < d.nbytes < 3 * (32 << 20)
    assert np.array_equal(d, want.astype(np.int32))


def test_distances_near_int32_max(gpu):
    # chain 0 -> 1 -> 2 -> 3 with weights 2^30 - 1, 2^30 - 1, 5 and a shortcut
    # 0 -> 3 of weight 2^31 - 10: every sum stays below INT32_MAX except
    # 1 -> 2 -> 3, which must not wrap (a wrapped sum would be negative and win)
    vl = np.array([0, 2, 3, 4, 4], np.int32)
    el = np.array([1, 3, 2, 3], np.int32)
    big = 2**30 - 1
    wl = np.array([big, 2**31 - 10, big, 5], np.int32)
    ref = oracle.dijkstra(vl, el, wl, 0)
    assert ref.tolist() == [0, big, 2 * big, 2**31 - 10]  # 2*big + 5 exceeds INT32_MAX
    with c2g.Graph.from_csr(vl, el, wl, device=0) as g:
        d, _ = g.sssp(0)
        db, _ = g.sssp_batch(np.array([0, 1, 2], np.int32))
    assert np.array_equal(d, ref)
    assert np.array_equal(db[0], ref)
    assert db[1].tolist() == [INT_MAX, 0, big, big + 5]
    assert db[2].tolist() == [INT_MAX, INT_MAX, 0, 5]


# ---- split mode (c2g_tune.xcd_split): frontier rows cut by the target's
# partition into eight queues, one per XCD's workgroups

@pytest.mark.parametrize("delta,block,chunk", [(0, 256, 32), (0, 512, 8), (16, 128, 64), (64, 256, 1)])
def test_split_mode_equals_dijkstra(gpu, delta, block, chunk):
    G = rmat_graph(16, 1, wmax=64)
    deg = np.diff(G["vlist"])
    rng = np.random.default_rng(delta + block)
    srcs = [int(np.argmax(deg))] + [int(x) for x in rng.choice(np.flatnonzero(deg > 0), 5)]
    refs = oracle.dijkstra_many(G["vlist"], G["elist"], G["wlist"], srcs)
    with c2g.Graph.from_csr(G["vlist"], G["elist"], G["wlist"], device=0) as g:
        g.set_tune(xcd_split=1, delta=delta, block_single=block, chunk=chunk)
        for s, ref in zip(srcs, refs):
            d, st = g.sssp(s)
            assert np.array_equal(d, ref), s
        sums, reach, _ = g.sssp_each_resident(np.array(srcs, np.int32))
        for k, ref in enumerate(refs):
            fin = ref != INT_MAX
            assert sums[k] == ref[fin].astype(np.int64).sum() and reach[k] == fin.sum()
        # budgeted runs read their queue back and run unsplit
        d, fr, _ = g.sssp_budget(srcs[0], 2)
        d2, _ = g.sssp_migrate(srcs[0], 3)
        assert np.array_equal(d2, refs[0])


def test_split_mode_pruned_and_unsorted_rows(gpu):
    from conftest import random_graph
    for seed in range(4):
        G = random_graph(500, 1600, seed=seed, wmax=9)
        P = c2g.prune_sssp_mem(G["vlist"], G["elist"], G["wlist"])
        with c2g.Graph.from_csr(P["vlist"], P["elist"], P["wlist"], hvlist=P["hvlist"], helist=P["helist"],
                                hwlist=P["hwlist"], variant=c2g.VARIANT_SSSP_PURN, device=0) as g:
            g.set_tune(xcd_split=1)
            for s in list(range(0, 500, 23)) + [int(x) for x in P["porder"][:10]]:
                assert np.array_equal(g.sssp(s)[0], oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], s)), (seed, s)
    # rows whose targets are not sorted go whole into one partition
    G = random_graph(3000, 20000, seed=7, wmax=20)
    el, wl = G["elist"].copy(), G["wlist"].copy()
    rng = np.random.default_rng(1)
    for u in range(0, 3000, 3):
        b, e = G["vlist"][u], G["vlist"][u + 1]
        perm = rng.permutation(e - b)
        el[b:e], wl[b:e] = el[b:e][perm], wl[b:e][perm]
    with c2g.Graph.from_csr(G["vlist"], el, wl, device=0) as g:
        g.set_tune(xcd_split=1)
        for s in (0, 5, 777, 2999):
            assert np.array_equal(g.sssp(s)[0], oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], s)), s


@pytest.mark.slow
def test_split_mode_scale20(gpu):
    G = rmat_graph(20, 2, wmax=255)
    deg = np.diff(G["vlist"])
    rng = np.random.default_rng(2)
    srcs = [int(np.argmax(deg))] + [int(x) for x in rng.choice(np.flatnonzero(deg > 0), 3)]
    refs = oracle.dijkstra_many(G["vlist"], G["elist"], G["wlist"], srcs)
    with c2g.Graph.from_csr(G["vlist"], G["elist"], G["wlist"], device=0) as g:
        g.set_tune(xcd_split=1)
        for s, ref in zip(srcs, refs):
            assert np.array_equal(g.sssp(s)[0], ref), s


def test_round_trace_single_and_batched(gpu):
    # per-round counters (the reference's iteration_act_num): the rounds add up
    # to the query's totals, the first round relaxes the source alone, and
    # the distances equal the untraced query's
    G = rmat_graph(16, 1, wmax=64)
    s = int(np.argmax(np.diff(G["vlist"])))
    with c2g.Graph.from_csr(G["vlist"], G["elist"], G["wlist"], device=0) as g:
        d0, st0 = g.sssp(s)
        d, st, rs = g.sssp_rounds(s)
        # a round whose queue is empty (light-edge mode: only vertices kept
        # for their full rows) is traced but does not count as an iteration
        assert np.array_equal(d, d0) and len(rs) >= st.iterations
        assert rs[0]["frontier"] == 1 and rs[0]["edges"] == G["vlist"][s + 1] - G["vlist"][s]
        assert sum(r["frontier"] for r in rs) == st.frontier_total
        assert sum(r["edges"] for r in rs) == st.edges_relaxed
        assert sum(r["improvements"] for r in rs) == st.improvements
        assert sum(r["items"] for r in rs) == st.items and all(r["ms"] > 0 for r in rs)
        srcs = np.flatnonzero(np.diff(G["vlist"]) > 0)[:100].astype(np.int32)
        stb, rb = g.sssp_batch_rounds(srcs)
        sums, reach, st2 = g.sssp_batch_resident(srcs)
        assert len(rb) >= stb.iterations
        assert sum(r["frontier"] for r in rb) == stb.frontier_total
        assert sum(r["lane_edges"] for r in rb) == stb.edges_relaxed
        assert sum(r["edges"] for r in rb) == stb.csr_edges
        assert sum(r["improvements"] for r in rb) == stb.improvements
        _, _, rs2 = g.sssp_rounds(s, cap=2)
        assert len(rs2) == 2  # the trace is cut at cap
