"""Degenerate inputs and limits on the GPU: a graph without edges on every
path, and the largest batches the batched entry points accept (65,536
sources; 16,384 with delta-stepping), against the oracle."""
import numpy as np
import pytest

import c2g
import oracle
from conftest import INT_MAX, random_graph

pytestmark = pytest.mark.gpu


def test_edgeless_graph_on_every_path(gpu):
    V = 70
    vl = np.zeros(V + 1, np.int32)
    e = np.zeros(0, np.int32)
    with c2g.Graph.from_csr(vl, e, e, np.zeros(0, np.float32), device=0) as g:
        srcs = np.arange(V, dtype=np.int32)
        d, _ = g.sssp_batch(srcs)
        assert np.array_equal(d, np.where(np.eye(V, dtype=bool), 0, INT_MAX))
        ds, st = g.sssp(5)
        assert ds[5] == 0 and st.reached == 1
        # PPR: the source keeps alpha and has nothing to push (mass dropped)
        p, r, _ = g.ppr(3)
        assert p[3] == pytest.approx(0.15) and p.sum() == pytest.approx(0.15) and not r.any()
        P, R, _ = g.ppr_batch(srcs[:10])
        assert np.allclose(P, np.eye(10, V) * 0.15) and not R.any()
        # adsorption: no neighbours, so every iterate is p_inj * Y0
        Y, _ = g.adsorption(np.array([1, 2], np.int32), np.array([0, 3], np.int32), L=4, iters=5)
        Y0 = np.zeros((V, 4), np.float32)
        Y0[1, 0] = Y0[2, 3] = 1
        assert np.allclose(Y, 0.25 * Y0)
        with c2g.Server(g, ads_labels=4, ads_iters=3) as s:
            ids = [s.submit(c2g.Q_SSSP, 7), s.submit(c2g.Q_PPR, 8), s.submit(c2g.Q_ADS, 1)]
            (i0, r0), (i1, r1), (i2, r2) = [s.poll(i) for i in ids]
    assert i0["status"] == i1["status"] == i2["status"] == 0
    assert r0[7] == 0 and (r0 != INT_MAX).sum() == 1
    assert r1[8] == pytest.approx(0.15) and r1.sum() == pytest.approx(0.15)
    assert r2.sum() == pytest.approx(0.25 * max(1, V // 100))  # one seed per 100 vertices, p_inj each


def test_largest_batches(gpu):
    G = random_graph(40, 160, seed=9, wmax=9)
    srcs = (np.arange(65536) % 40).astype(np.int32)
    refs = np.stack([oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], s) for s in range(40)])
    with c2g.Graph.from_csr(G["vlist"], G["elist"], G["wlist"], G["flist"], device=0) as g:
        d, _ = g.sssp_batch(srcs)  # 1,024 groups of 64 lanes
        assert np.array_equal(d, refs[srcs])
        sums, reached, _ = g.sssp_batch_resident(srcs)
        fin = refs != INT_MAX
        assert np.array_equal(sums, np.where(fin, refs, 0).sum(axis=1)[srcs])
        assert np.array_equal(reached, fin.sum(axis=1)[srcs])
        P, R, _ = g.ppr_batch(srcs)
        for q in (0, 39, 64 * 511 + 5, 65535):
            p, r, _ = g.ppr(int(srcs[q]))
            assert np.abs(P[q].astype(np.float64) - p).sum() <= 2e-5, q
        for call in (g.sssp_batch_resident, g.ppr_batch):
            with pytest.raises(c2g.C2GError) as e:
                call(np.zeros(65537, np.int32))
            assert e.value.code == c2g.EINVAL
        g.set_tune(delta=4)
        d4, _ = g.sssp_batch(srcs[:16384])
        assert np.array_equal(d4, refs[srcs[:16384]])
        with pytest.raises(c2g.C2GError) as e:
            g.sssp_batch_resident(srcs[:16385])
        assert e.value.code == c2g.EINVAL


def test_allocation_failure_leaves_handle_usable(gpu):
    # a batch whose lane rows cannot fit in HBM (1,024 groups x 4 M vertices x
    # 256 bytes, about 1 TB) fails with C2G_ENOMEM; afterwards the same handle
    # runs smaller batches and the other paths correctly: no half-built
    # workspace of the failed call is used and no stale runtime error remains.
    # (The batched PPR's host outputs for that many sources would not fit in
    # host memory, so its own allocation path is not driven to failure here.)
    G = random_graph(1 << 22, 1 << 23, seed=11, wmax=50)
    big = np.arange(65536, dtype=np.int32)
    srcs = np.array([0, 1, 2, 3, 12345, 4194303], np.int32)
    with c2g.Graph.from_csr(G["vlist"], G["elist"], G["wlist"], G["flist"], device=0) as g:
        d0, _ = g.sssp_batch(srcs[:2])  # a workspace exists before the failure
        with pytest.raises(c2g.C2GError) as e:
            g.sssp_batch_resident(big)
        assert e.value.code == c2g.ENOMEM, e.value
        d, _ = g.sssp_batch(srcs)
        assert np.array_equal(d[:2], d0)
        for q in (0, 5):
            assert np.array_equal(d[q], oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], int(srcs[q]))), q
        with pytest.raises(c2g.C2GError) as e:
            g.sssp_batch_resident(big)
        assert e.value.code == c2g.ENOMEM, e.value
        dq, _ = g.sssp(int(srcs[5]))
        assert np.array_equal(dq, d[5])
        P, R, _ = g.ppr_batch(srcs)
        for q in (0, 5):
            p, r, _ = g.ppr(int(srcs[q]))
            assert np.abs(P[q].astype(np.float64) - p).sum() <= 2e-5, q
