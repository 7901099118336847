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
        assert t.grid_per_cu in (2, 4, 8, 16) and t.iters_per_graph in (4, 8, 16) and t.block in (128, 256, 512)
        assert t.block_single in (0, 128, 256, 512)
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
        assert t.ppr_pull in (0, 25, 50, 100, 200, 400, 800) and t.ppr_grid_per_cu in (1, 2, 4, 8)
        assert g.tune().as_dict() == t.as_dict()
        p1, r1, _ = g.ppr(int(srcs[0]), eps=1e-9)
        with pytest.raises(c2g.C2GError):
            g.set_tune(ppr_pull=1001)
    assert np.abs(p1.astype(np.float64) - p0).sum() <= 2e-5


@pytest.mark.parametrize("block", [128, 256, 512])
def test_block_shapes_keep_results(gpu, block):
    # every relax workgroup size the tuner may pick, single-source, batched
    # and on a pruned graph, gives Dijkstra's distances
    from conftest import random_graph
    G = rmat_graph(14, 1, wmax=64)
    deg = np.diff(G["vlist"])
    srcs = np.concatenate([[int(np.argmax(deg))], np.flatnonzero(deg == 1)[:3]]).astype(np.int32)
    refs = oracle.dijkstra_many(G["vlist"], G["elist"], G["wlist"], srcs)
    with c2g.Graph.from_csr(G["vlist"], G["elist"], G["wlist"], device=0) as g:
        g.set_tune(block=block, block_single=0)
        db, _ = g.sssp_batch(srcs)
        for q, s in enumerate(srcs):
            assert np.array_equal(g.sssp(int(s))[0], refs[q]) and np.array_equal(db[q], refs[q])
        g.set_tune(block=256, block_single=block)
        assert np.array_equal(g.sssp(int(srcs[0]))[0], refs[0])
        with pytest.raises(c2g.C2GError):
            g.set_tune(block=64)
    H = random_graph(200, 500, seed=4, wmax=9)
    P = c2g.prune_sssp_mem(H["vlist"], H["elist"], H["wlist"])
    with c2g.Graph.from_csr(P["vlist"], P["elist"], P["wlist"], hvlist=P["hvlist"], helist=P["helist"],
                            hwlist=P["hwlist"], variant=c2g.VARIANT_SSSP_PURN, device=0) as g:
        g.set_tune(block=block)
        ss = np.arange(0, 200, 7, dtype=np.int32)
        db, _ = g.sssp_batch(ss)
        for q, s in enumerate(ss):
            ref = oracle.dijkstra(H["vlist"], H["elist"], H["wlist"], int(s))
            assert np.array_equal(db[q], ref) and np.array_equal(g.sssp(int(s))[0], ref), int(s)
