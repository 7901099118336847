"""Autotuner: picks a valid setting and never changes results."""
import numpy as np
import pytest

import c2g
import oracle
from conftest import rmat_graph

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("batched", [False, True])
def test_autotune_keeps_results(gpu, batched):
    G = rmat_graph(16, 1, wmax=64)
    deg = np.diff(G["vlist"])
    srcs = np.flatnonzero(deg > 0)[:70].astype(np.int32)
    with c2g.Graph.from_csr(G["vlist"], G["elist"], G["wlist"], device=0) as g:
        t = g.autotune(srcs[:8] if not batched else srcs, batched=batched)
        assert t.grid_per_cu in (2, 4, 8, 16) and t.iters_per_graph in (4, 8, 16)
        assert g.tune().as_dict() == t.as_dict()
        d, _ = g.sssp(int(srcs[0]))
        db, _ = g.sssp_batch(srcs[:5])
    ref = oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], int(srcs[0]))
    assert np.array_equal(d, ref) and np.array_equal(db[0], ref)


def test_autotune_ppr_switch_point(gpu):
    # the push/pull switch point is picked from timed candidates and keeps the
    # result within the FP32 summation-order bound
    G = rmat_graph(16, 3, wmax=0, symmetric=True, want_f=True)
    deg = np.diff(G["vlist"])
    srcs = np.flatnonzero(deg > 0)[:6].astype(np.int32)
    with c2g.Graph.from_csr(G["vlist"], G["elist"], flist=G["flist"], device=0) as g:
        p0, r0, _ = g.ppr(int(srcs[0]), eps=1e-9)
        t = g.autotune_ppr(srcs, eps=1e-9)
        assert t.ppr_pull in (0, 25, 50, 100, 200, 400, 800)
        assert g.tune().as_dict() == t.as_dict()
        p1, r1, _ = g.ppr(int(srcs[0]), eps=1e-9)
        with pytest.raises(c2g.C2GError):
            g.set_tune(ppr_pull=1001)
    assert np.abs(p1.astype(np.float64) - p0).sum() <= 2e-5
