"""Query server on the GPU: a mixed SSSP / PPR / adsorption stream with
continuous batching.  Each result must equal the same query run alone:
SSSP bit-exact (and equal to Dijkstra); PPR within L1 2e-5 of the standalone
GPU push (FP32 atomics in arbitrary order); adsorption within max-abs 2e-6 of
the host lane's FP32 Jacobi."""
import time

import numpy as np
import pytest

import c2g
import oracle
from conftest import rmat_graph

pytestmark = pytest.mark.gpu


def _G():
    return rmat_graph(14, 9, wmax=255, want_f=True)


def _graph(G):
    return c2g.Graph.from_csr(G["vlist"], G["elist"], G["wlist"], G["flist"], device=0)


def test_mixed_stream_matches_alone(gpu):
    G = _G()
    rng = np.random.default_rng(5)
    deg = np.diff(G["vlist"])
    kinds = rng.choice([c2g.Q_SSSP] * 6 + [c2g.Q_PPR] * 3 + [c2g.Q_ADS], 80)
    srcs = [int(rng.choice(np.flatnonzero(deg > 0))) if k != c2g.Q_ADS else int(rng.integers(0, 4)) for k in kinds]
    with _graph(G) as g:
        with c2g.Server(g, ads_labels=16, ads_iters=20) as s:
            ids = []
            for k, v in zip(kinds, srcs):
                ids.append(s.submit(k, v))
                time.sleep(rng.exponential(0.002))  # Poisson arrivals while others run
            res = [s.poll(i) for i in ids]
        ads_alone = {}
        with c2g.Server(g, ads_labels=16, ads_iters=20, device_failed=True, host_lane=True) as h:
            for v in sorted(set(v for k, v in zip(kinds, srcs) if k == c2g.Q_ADS)):
                ads_alone[v] = h.poll(h.submit(c2g.Q_ADS, v))[1]
        for (info, r), k, v in zip(res, kinds, srcs):
            assert info["status"] == 0 and info["lane"] == c2g.LANE_GPU
            if k == c2g.Q_SSSP:
                d, _ = g.sssp(v)
                assert np.array_equal(r, d)
                assert np.array_equal(r, oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], v))
            elif k == c2g.Q_PPR:
                p, _, _ = g.ppr(v, alpha=0.15, eps=1e-7)
                assert np.abs(r.astype(np.float64) - p).sum() <= 2e-5
            else:
                assert np.abs(r - ads_alone[v]).max() <= 2e-6
    lat = [i["t_done_ms"] - i["t_submit_ms"] for i, _ in res]
    assert min(lat) >= 0


def test_memory_admission_and_fallback(gpu):
    # lanes are admitted by their device workspace: a budget of one slot's
    # bytes starts the SSSP slot pool with one slot and no other lane
    G = _G()
    with _graph(G) as g:
        with c2g.Server(g, sssp_slots=1) as s:
            one_slot = s.stats()["lane_bytes"][0]
        with c2g.Server(g, fake_free_hbm=one_slot) as s:
            st = s.stats()
            assert st["lanes_enabled"] == c2g.SLANE_SLOTS and st["sssp_slots"] == 1
            ids = [s.submit(c2g.Q_SSSP, v) for v in range(0, 400, 40)]  # queued behind the one slot
            for i, v in zip(ids, range(0, 400, 40)):
                info, d = s.poll(i)
                assert info["lane"] == c2g.LANE_GPU
                assert np.array_equal(d, oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], v))
            with pytest.raises(c2g.C2GError) as e:  # no PPR lane fits, no host lane
                s.submit(c2g.Q_PPR, 0)
            assert e.value.code == c2g.EBUSY
        # no room at all: refused without the host lane, host lane with it
        with c2g.Server(g, fake_free_hbm=1024) as s:
            with pytest.raises(c2g.C2GError) as e:
                s.submit(c2g.Q_SSSP, 0)
            assert e.value.code == c2g.EBUSY
        with c2g.Server(g, fake_free_hbm=1024, host_lane=True) as s:
            info, d = s.poll(s.submit(c2g.Q_SSSP, 7))
            assert info["lane"] == c2g.LANE_HOST
            assert np.array_equal(d, oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], 7))
        # device marked failed: the host lane takes everything
        with c2g.Server(g, device_failed=True, host_lane=True) as s:
            info, p = s.poll(s.submit(c2g.Q_PPR, 3))
            assert info["lane"] == c2g.LANE_HOST


def test_two_servers_on_one_handle(gpu):
    # the slot pool is sized per server (16 slots, then 4) and starts empty
    G = _G()
    srcs = [int(v) for v in np.flatnonzero(np.diff(G["vlist"]) > 0)[::53][:40]]
    with _graph(G) as g:
        for slots in (16, 4):
            with c2g.Server(g, sssp_slots=slots) as s:
                assert s.stats()["sssp_slots"] == slots
                ids = [s.submit(c2g.Q_SSSP, v) for v in srcs]
                for i, v in zip(ids, srcs):
                    info, d = s.poll(i)
                    assert info["status"] == 0 and info["lane"] == c2g.LANE_GPU
                    assert np.array_equal(d, oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], v)), (slots, v)
        d, _ = g.sssp(srcs[0])  # the handle is the caller's again
        assert np.array_equal(d, oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], srcs[0]))


def test_handle_is_owned_while_serving(gpu):
    G = _G()
    with _graph(G) as g:
        with c2g.Server(g) as s:
            for call in (lambda: g.sssp(0), lambda: g.sssp_batch([0, 1]), lambda: g.ppr(0),
                         lambda: g.adsorption([0], [0], L=1)):
                with pytest.raises(c2g.C2GError) as e:
                    call()
                assert e.value.code == c2g.EBUSY
            info, d = s.poll(s.submit(c2g.Q_SSSP, 5))
            assert np.array_equal(d, oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], 5))


@pytest.mark.parametrize("lane,slots,kind", [(c2g.SLANE_SLOTS, 0, c2g.Q_SSSP), (c2g.SLANE_GROUP, -1, c2g.Q_SSSP),
                                             (c2g.SLANE_PPR, 0, c2g.Q_PPR)])
def test_injected_lane_failure_recovers(gpu, lane, slots, kind):
    # a device lane's step reports an error (forced from the host; every kernel
    # still runs normally): its running requests finish on the host lane, the
    # lane is reset and serves the next requests on the GPU
    G = _G()
    deg = np.diff(G["vlist"])
    srcs = [int(v) for v in np.flatnonzero(deg > 0)[::101][:12]]
    with _graph(G) as g:
        with c2g.Server(g, host_lane=True, fail_inject=lane, fail_after=2, sssp_slots=slots) as s:
            res = []
            for v in srcs:  # one at a time: a known request runs into the failing step
                res.append((v,) + s.poll(s.submit(kind, v)))
            st = s.stats()
        lanes = [info["lane"] for _, info, _ in res]
        assert st["lane_failures"] == 1 and st["failed"] == 0 and st["to_host"] == 1
        assert lanes.count(c2g.LANE_HOST) == 1 and lanes[-1] == c2g.LANE_GPU
        for v, info, r in res:
            assert info["status"] == 0
            if kind == c2g.Q_SSSP:
                assert np.array_equal(r, oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], v)), v
            else:
                p, _, _ = g.ppr(v, alpha=0.15, eps=1e-7)
                assert np.abs(r.astype(np.float64) - p).sum() <= 2e-5, v


@pytest.mark.parametrize("delta", [0, 100])
def test_lane_reuse_many_queries(gpu, delta):
    # more queries than lanes: lanes are freed and reused, columns and bucket bounds reset
    G = _G()
    deg = np.diff(G["vlist"])
    srcs = np.flatnonzero(deg > 0)[:150]
    with _graph(G) as g:
        g.set_tune(delta=delta)
        with c2g.Server(g, max_batch=16) as s:
            ids = [s.submit(c2g.Q_SSSP, int(v)) for v in srcs]
            for i, v in zip(ids, srcs):
                info, d = s.poll(i)
                assert np.array_equal(d, oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], int(v))), v


def test_ppr_burst_is_batched(gpu):
    # a burst of PPR requests queues behind the first and runs as one batched
    # pass; every result must match the query run alone
    G = _G()
    deg = np.diff(G["vlist"])
    srcs = np.flatnonzero(deg > 0)[::97][:40]
    with _graph(G) as g:
        with c2g.Server(g) as s:
            ids = [s.submit(c2g.Q_PPR, int(v)) for v in srcs]
            res = [s.poll(i) for i in ids]
        for (info, p), v in zip(res, srcs):
            assert info["status"] == 0 and info["lane"] == c2g.LANE_GPU
            ref, _, _ = g.ppr(int(v), alpha=0.15, eps=1e-7)
            assert np.abs(p.astype(np.float64) - ref).sum() <= 2e-5, v
    admits = [i["t_admit_ms"] for i, _ in res]
    assert len(set(admits)) < len(admits)  # a batch shares one admission time


@pytest.mark.parametrize("slots,lanes", [(64, 64), (-1, 64), (4, 8), (1, 1)])
def test_sssp_pools(gpu, slots, lanes):
    # slot pool only, lane group only, and a burst that overflows a small slot
    # pool into the lane group; every pool's results equal Dijkstra
    G = _G()
    srcs = np.flatnonzero(np.diff(G["vlist"]) > 0)[::37][:120]
    with _graph(G) as g:
        with c2g.Server(g, max_batch=lanes, sssp_slots=slots) as s:
            ids = [s.submit(c2g.Q_SSSP, int(v)) for v in srcs]
            for i, v in zip(ids, srcs):
                info, d = s.poll(i)
                assert info["status"] == 0
                assert np.array_equal(d, oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], int(v))), v


def test_pruned_graph_server(gpu):
    # pruned graph (hidden lists relaxed per slot / per lane): results equal
    # Dijkstra on the full graph, through both SSSP pools
    from conftest import random_graph
    for seed in range(3):
        G = random_graph(300, 700, seed=seed, wmax=9)
        P = c2g.prune_sssp_mem(G["vlist"], G["elist"], G["wlist"])
        assert len(P["helist"]) > 0
        with c2g.Graph.from_csr(P["vlist"], P["elist"], P["wlist"], hvlist=P["hvlist"], helist=P["helist"],
                                hwlist=P["hwlist"], variant=c2g.VARIANT_SSSP_PURN, device=0) as g:
            with c2g.Server(g, max_batch=16, sssp_slots=8) as s:
                ids = [s.submit(c2g.Q_SSSP, v) for v in range(300)]
                for v, i in enumerate(ids):
                    info, d = s.poll(i)
                    assert np.array_equal(d, oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], v)), (seed, v)


@pytest.mark.parametrize("relabel", [0, 1])
def test_adsorption_burst_is_batched(gpu, relabel):
    # a burst of adsorption requests runs as passes of up to 4 seed sets side by
    # side (4 x 16 labels); each result equals the same seed set run alone, also
    # when the passes run on the in-degree-relabelled copy of the graph
    G = _G()
    sets = list(range(9))
    with _graph(G) as g:
        g.set_tune(ads_relabel=relabel)
        with c2g.Server(g, ads_labels=16, ads_iters=20) as s:
            ids = [s.submit(c2g.Q_ADS, k) for k in sets]
            res = [s.poll(i) for i in ids]
        with c2g.Server(g, ads_labels=16, ads_iters=20, device_failed=True, host_lane=True) as h:
            alone = [h.poll(h.submit(c2g.Q_ADS, k))[1] for k in sets]
    for (info, y), ya in zip(res, alone):
        assert info["status"] == 0 and info["lane"] == c2g.LANE_GPU
        assert np.abs(y - ya).max() <= 2e-6
    admits = [i["t_admit_ms"] for i, _ in res]
    assert len(set(admits)) < len(admits)  # a pass shares one admission time
