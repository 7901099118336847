"""Query server: host fallback lane and admission rules on the CPU (no device);
the device lanes are covered by test_server_gpu.py."""
import numpy as np
import pytest

import c2g
import oracle
from conftest import rmat_graph


def _graph(**kw):
    G = rmat_graph(12, 9, wmax=64, want_f=True)
    return G, c2g.Graph.from_csr(G["vlist"], G["elist"], G["wlist"], G["flist"], **kw)


def test_host_lane_serves_every_kind():
    G, g = _graph(no_device=True)
    with g, c2g.Server(g, host_lane=True, ads_labels=4, ads_iters=5) as s:
        ids = [(s.submit(c2g.Q_SSSP, v), c2g.Q_SSSP, v) for v in (0, 17, 300)]
        ids += [(s.submit(c2g.Q_PPR, v), c2g.Q_PPR, v) for v in (5, 9)]
        ids += [(s.submit(c2g.Q_ADS, 3), c2g.Q_ADS, 3)]
        for rid, kind, v in ids:
            info, res = s.poll(rid)
            assert info["status"] == 0 and info["lane"] == c2g.LANE_HOST and info["kind"] == kind
            assert info["t_done_ms"] >= info["t_admit_ms"] >= info["t_submit_ms"]
            if kind == c2g.Q_SSSP:
                assert np.array_equal(res, oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], v))
            elif kind == c2g.Q_PPR:
                po, _, _, _ = oracle.ppr_push(G["vlist"], G["elist"], G["flist"], v, alpha=0.15, eps=1e-7)
                assert np.array_equal(res, po)
            else:
                assert res.shape == (G["V"], 4) and res.max() <= 1.0 + 1e-6


def test_no_device_without_host_lane_is_refused():
    G, g = _graph(no_device=True)
    with g:
        with pytest.raises(c2g.C2GError) as e:
            c2g.Server(g, host_lane=False)
        assert e.value.code == c2g.EDEVICE


def test_inflight_limit_and_bad_requests():
    G, g = _graph(no_device=True)
    with g, c2g.Server(g, host_lane=True, max_inflight=2) as s:
        a = s.submit(c2g.Q_SSSP, 1)
        b = s.submit(c2g.Q_SSSP, 2)
        with pytest.raises(c2g.C2GError) as e:
            for _ in range(50):  # the first two may finish quickly; keep the queue full
                s.submit(c2g.Q_SSSP, 3)
        assert e.value.code == c2g.EBUSY
        with pytest.raises(c2g.C2GError):
            s.submit(c2g.Q_SSSP, G["V"])
        with pytest.raises(c2g.C2GError):
            s.submit(7, 0)
        s.poll(a)
        s.poll(b)


# ---- device lanes on the dry backend (host computations in place of
# kernels): admission by the HBM budget, lane failure and recovery,
# cancellation, deadlines and handle ownership, without a GPU.
# Roles: the reference moves work between its GPU streams and CPU worker by
# queue occupancy (src/sssp/purn-perk/Main.cu:128-145) and its worker polls a
# ring (src/sssp/purn-perk/SSSP.cpp:140-159); here the lanes do.

BIG = 1 << 50


def _dijkstra(G, v):
    return oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], v)


def _wait(cond, timeout=10.0):
    import time
    t = time.time()
    while not cond():
        assert time.time() - t < timeout, "timed out waiting"
        time.sleep(0.005)


def test_dry_lanes_serve_and_charge_the_budget():
    G, g = _graph(no_device=True)
    with g, c2g.Server(g, dry_device=True, fake_free_hbm=BIG, ads_labels=4, ads_iters=5) as s:
        st = s.stats()
        assert st["lanes_enabled"] == 15 and st["sssp_slots"] == 16
        assert st["charged"] == sum(st["lane_bytes"]) and all(b > 0 for b in st["lane_bytes"])
        rs = [(s.submit(c2g.Q_SSSP, v), v) for v in (0, 17, 300)]
        for rid, v in rs:
            info, d = s.poll(rid)
            assert info["lane"] == c2g.LANE_GPU and np.array_equal(d, _dijkstra(G, v))
        info, p = s.poll(s.submit(c2g.Q_PPR, 5))
        po, _, _, _ = oracle.ppr_push(G["vlist"], G["elist"], G["flist"], 5, alpha=0.15, eps=1e-7)
        assert info["lane"] == c2g.LANE_GPU and np.array_equal(p, po)
    big = st
    # half the 16-slot pool's bytes: the pool shrinks to 8 slots and nothing
    # else fits; kinds without a lane are refused (no host lane) ...
    with c2g.Graph.from_csr(G["vlist"], G["elist"], G["wlist"], G["flist"], no_device=True) as g2:
        budget = big["lane_bytes"][0] // 2 + 8192  # covers the pool's fixed part (about 9.6 KB)
        with c2g.Server(g2, dry_device=True, fake_free_hbm=budget, ads_labels=4, ads_iters=5) as s:
            st = s.stats()
            assert st["sssp_slots"] == 8 and st["lanes_enabled"] == c2g.SLANE_SLOTS
            assert st["charged"] <= st["budget"] == budget
            with pytest.raises(c2g.C2GError) as e:
                s.submit(c2g.Q_PPR, 5)
            assert e.value.code == c2g.EBUSY
            info, d = s.poll(s.submit(c2g.Q_SSSP, 9))
            assert info["lane"] == c2g.LANE_GPU and np.array_equal(d, _dijkstra(G, 9))
        # ... and go to the host lane when it is on
        with c2g.Server(g2, dry_device=True, fake_free_hbm=budget, host_lane=True, ads_labels=4, ads_iters=5) as s:
            info, p = s.poll(s.submit(c2g.Q_PPR, 5))
            assert info["lane"] == c2g.LANE_HOST
            assert s.stats()["to_host"] == 1
        # a budget below every lane: all kinds on the host lane
        with c2g.Server(g2, dry_device=True, fake_free_hbm=1, host_lane=True, ads_labels=4, ads_iters=5) as s:
            assert s.stats()["lanes_enabled"] == 0
            info, d = s.poll(s.submit(c2g.Q_SSSP, 11))
            assert info["lane"] == c2g.LANE_HOST and np.array_equal(d, _dijkstra(G, 11))


@pytest.mark.parametrize("lane,slots", [(c2g.SLANE_SLOTS, 0), (c2g.SLANE_GROUP, -1)])
def test_lane_failure_moves_running_requests_to_host(lane, slots):
    G, g = _graph(no_device=True)
    with g, c2g.Server(g, dry_device=True, fake_free_hbm=BIG, host_lane=True, fail_inject=lane, fail_after=1,
                       sssp_slots=slots, dry_step_ms=5) as s:
        # the first step succeeds and the second fails: the requests running in
        # it (at least the first of the second burst) end on the host lane
        for srcs, want in (([0], c2g.LANE_GPU), ([300, 41, 1000, 7], c2g.LANE_HOST)):
            ids = [s.submit(c2g.Q_SSSP, v) for v in srcs]
            for k, (rid, v) in enumerate(zip(ids, srcs)):
                info, d = s.poll(rid)
                assert info["status"] == 0 and np.array_equal(d, _dijkstra(G, v)), v
                if k == 0 or want == c2g.LANE_GPU:
                    assert info["lane"] == want, (v, info)
        st = s.stats()
        assert st["lane_failures"] == 1 and st["to_host"] >= 1 and st["failed"] == 0
        # the lane was reset and serves again
        info, d = s.poll(s.submit(c2g.Q_SSSP, 3))
        assert info["lane"] == c2g.LANE_GPU and np.array_equal(d, _dijkstra(G, 3))
        assert s.stats()["lanes_enabled"] & lane


def test_lane_failure_without_host_lane_fails_only_running_requests():
    G, g = _graph(no_device=True)
    with g, c2g.Server(g, dry_device=True, fake_free_hbm=BIG, fail_inject=c2g.SLANE_SLOTS, fail_after=0,
                       sssp_slots=1, max_batch=1, dry_step_ms=20) as s:
        a = s.submit(c2g.Q_SSSP, 0)
        with pytest.raises(c2g.C2GError) as e:
            s.poll(a)
        assert e.value.code == c2g.EDEVICE
        info, d = s.poll(s.submit(c2g.Q_SSSP, 17))  # not sticky: later requests succeed
        assert info["status"] == 0 and np.array_equal(d, _dijkstra(G, 17))


@pytest.mark.parametrize("lane,kind", [(c2g.SLANE_PPR, c2g.Q_PPR), (c2g.SLANE_ADS, c2g.Q_ADS)])
def test_aux_lane_failure_moves_pass_to_host(lane, kind):
    G, g = _graph(no_device=True)
    with g, c2g.Server(g, dry_device=True, fake_free_hbm=BIG, host_lane=True, fail_inject=lane, ads_labels=4,
                       ads_iters=5) as s:
        info, first = s.poll(s.submit(kind, 5))
        assert info["status"] == 0 and info["lane"] == c2g.LANE_HOST
        info, again = s.poll(s.submit(kind, 5))
        assert info["lane"] == c2g.LANE_GPU
        assert np.array_equal(first, again)  # host lane and dry lane run the same host code
        assert s.stats()["lane_failures"] == 1


def test_cancel_queued_and_running():
    G, g = _graph(no_device=True)
    with g:
        with c2g.Server(g, dry_device=True, fake_free_hbm=BIG, sssp_slots=1) as s:
            one_slot = s.stats()["lane_bytes"][0]
        # a budget of exactly one slot: one SSSP query runs, the rest queue
        with c2g.Server(g, dry_device=True, fake_free_hbm=one_slot, sssp_slots=1, dry_step_ms=300) as s:
            assert s.stats()["lanes_enabled"] == c2g.SLANE_SLOTS
            a = s.submit(c2g.Q_SSSP, 0)
            _wait(lambda: s.stats()["running"] == 1)
            b, c = s.submit(c2g.Q_SSSP, 17), s.submit(c2g.Q_SSSP, 300)
            assert s.cancel(c) == 1  # queued: finished at once
            with pytest.raises(c2g.C2GError) as e:
                s.poll(c, wait_ms=0)
            assert e.value.code == c2g.ETIMEOUT
            assert s.cancel(a) == 1  # running: finished at its lane's step boundary
            with pytest.raises(c2g.C2GError) as e:
                s.poll(a)
            assert e.value.code == c2g.ETIMEOUT
            info, d = s.poll(b)
            assert np.array_equal(d, _dijkstra(G, 17))
            with pytest.raises(c2g.C2GError) as e:  # b is gone after its poll
                s.cancel(b)
            assert e.value.code == c2g.EINVAL
            st = s.stats()
            assert st["cancelled"] == 2 and st["completed"] == 1


def test_deadline_expires_requests():
    G, g = _graph(no_device=True)
    with g, c2g.Server(g, dry_device=True, fake_free_hbm=BIG, deadline_ms=60, sssp_slots=1, max_batch=1,
                       dry_step_ms=120, host_lane=True) as s:
        ids = [s.submit(c2g.Q_SSSP, v) for v in (0, 17, 300)]
        for rid in ids:  # running ones expire at the step boundary, queued ones at admission
            with pytest.raises(c2g.C2GError) as e:
                s.poll(rid)
            assert e.value.code == c2g.ETIMEOUT
        assert s.stats()["cancelled"] == 3


def test_server_owns_the_handle():
    G, g = _graph(no_device=True)
    with g:
        s = c2g.Server(g, dry_device=True, fake_free_hbm=BIG)
        try:
            for call in (lambda: g.sssp(0), lambda: g.ppr(0), lambda: g.sssp_batch([0, 1]),
                         lambda: g.set_tune(chunk=16)):
                with pytest.raises(c2g.C2GError) as e:
                    call()
                assert e.value.code == c2g.EBUSY
            with pytest.raises(c2g.C2GError) as e:
                c2g.Server(g, dry_device=True, fake_free_hbm=BIG)
            assert e.value.code == c2g.EBUSY
            assert c2g.lib().c2g_close(g._h) == c2g.EBUSY
        finally:
            s.stop()
        g.set_tune(chunk=16)  # released at stop


def test_group_lane_charges_the_relabelled_copy():
    # the 64-lane group runs on a relabelled copy of the graph (row offsets,
    # edge entries and the new ids: 8 bytes per edge and 12 per vertex on a
    # graph without light rows or hidden lists); admission charges it only
    # while batch_relabel is on
    G = rmat_graph(12, 9, wmax=64, want_f=True)
    E, V = len(G["elist"]), G["V"]
    charged = {}
    for relabel in (1, 0):
        with c2g.Graph.from_csr(G["vlist"], G["elist"], G["wlist"], G["flist"], no_device=True) as g:
            g.set_tune(batch_relabel=relabel)
            with c2g.Server(g, dry_device=True, fake_free_hbm=BIG, ads_labels=4, ads_iters=5) as s:
                charged[relabel] = s.stats()["lane_bytes"][1]
    assert charged[1] - charged[0] == 8 * E + 12 * (V + 1)
