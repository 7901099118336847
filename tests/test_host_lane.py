"""Host-core lane (the reference CPU worker, src/sssp/base-perk/SSSP.cpp:88-119
and src/ppr/w-base-perk/PPR.cpp:178-236) against the oracle."""
import numpy as np

import c2g
import oracle
from conftest import INT_MAX, rmat_graph


def test_host_sssp_equals_dijkstra():
    G = rmat_graph(16, 1, wmax=64)
    deg = np.diff(G["vlist"])
    with c2g.Graph.from_csr(G["vlist"], G["elist"], G["wlist"], no_device=True) as g:
        for s in [int(np.argmax(deg)), 5, int(np.flatnonzero(deg == 0)[0])]:
            d, st = g.host_sssp(s)
            assert st.lane == c2g.LANE_HOST
            assert np.array_equal(d, oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], s))


def test_host_bf_resumes_partial_state():
    # the -M handoff: any upper-bound dist with the right frontier converges
    G = rmat_graph(14, 2, wmax=64)
    s = int(np.argmax(np.diff(G["vlist"])))
    ref = oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], s)
    with c2g.Graph.from_csr(G["vlist"], G["elist"], G["wlist"], no_device=True) as g:
        d = np.full(G["V"], INT_MAX, np.int32)
        d[s] = 0
        d2, st = g.host_bellman_ford(d, np.array([s], np.int32))
        assert np.array_equal(d2, ref) and st.host_iterations > 0


def test_host_ppr_equals_oracle_push():
    G = rmat_graph(12, 3, wmax=0, symmetric=True, want_f=True)
    s = int(np.argmax(np.diff(G["vlist"])))
    with c2g.Graph.from_csr(G["vlist"], G["elist"], flist=G["flist"], no_device=True) as g:
        p, r, st = g.host_ppr(s, alpha=0.15, eps=1e-6)
    po, ro, _, _ = oracle.ppr_push(G["vlist"], G["elist"], G["flist"], s, alpha=0.15, eps=1e-6)
    # same serial algorithm and float order -> identical bits
    assert np.array_equal(p, po) and np.array_equal(r, ro)
