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
