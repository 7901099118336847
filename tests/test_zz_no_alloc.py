"""c2g_sssp_noalloc: the allocation-free variant of c2g_sssp (no reference
counterpart; the traversal of CalcuSSSP, src/sssp/base/Main.cu:109-356, and the
pruned seeding of src/sssp/purn/Main.cu:67-90, with its state in a static
device array).  Its distances must equal c2g_sssp's bit for bit, on consecutive
calls that share the scratch, the second on a smaller graph."""
import numpy as np
import pytest

import c2g
import oracle
from conftest import INT_MAX, rmat_graph

pytestmark = pytest.mark.gpu


def _noalloc(g, src, buf, garbage):
    buf.fill(np.full(g.V, garbage, np.int32))  # a stale buffer must not pass
    st = g.sssp_noalloc(src, buf)
    return buf.to_numpy(np.int32, g.V), st


def test_noalloc_equals_sssp_then_smaller_graph(gpu):
    big = rmat_graph(16, 1, wmax=64)
    small = rmat_graph(12, 3, wmax=255)
    with c2g.DeviceBuffer(4 * big["V"]) as buf:
        for G, nsrc in ((big, 3), (small, 2)):
            deg = np.diff(G["vlist"])
            srcs = [int(np.argmax(deg))] + [int(s) for s in np.flatnonzero(deg == 1)[: nsrc - 1]]
            with c2g.Graph.from_csr(G["vlist"], G["elist"], G["wlist"], device=0) as g:
                for k, s in enumerate(srcs):
                    ref, st_ref = g.sssp(s)
                    d, st = _noalloc(g, s, buf, garbage=-7 - k)
                    print(f"noalloc V={g.V} src={s} rounds={st.iterations} reached={st.reached}", flush=True)
                    assert np.array_equal(d, ref), (g.V, s, int((d != ref).sum()))
                    assert st.reached == int((ref != INT_MAX).sum())
                    assert st.iterations == st_ref.iterations and st.edges_relaxed == st_ref.edges_relaxed
                assert np.array_equal(d, oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], srcs[-1]))


def test_noalloc_pruned_equals_full_graph(gpu):
    G = rmat_graph(12, 2, wmax=255)
    P = c2g.prune_sssp_mem(G["vlist"], G["elist"], G["wlist"])
    srcs = [int(s) for s in P["porder"][:3]] + [int(np.argmax(np.diff(G["vlist"])))]
    with c2g.Graph.from_csr(P["vlist"], P["elist"], P["wlist"], hvlist=P["hvlist"], helist=P["helist"],
                            hwlist=P["hwlist"], variant=c2g.VARIANT_SSSP_PURN, device=0) as g, \
            c2g.DeviceBuffer(4 * g.V) as buf:
        for s in srcs:
            ref, _ = g.sssp(s)
            d, _ = _noalloc(g, s, buf, garbage=5)
            assert np.array_equal(d, ref), s
            assert np.array_equal(d, oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], s)), s


def test_noalloc_refuses_graphs_beyond_its_scratch(gpu):
    # chunk 1 makes the queue bound 2V + E items: 72 MiB at scale 18, over the 64 MiB scratch
    G = rmat_graph(18, 1, wmax=64)
    small = rmat_graph(12, 3, wmax=255)
    with c2g.Graph.from_csr(G["vlist"], G["elist"], G["wlist"], device=0) as g, \
            c2g.DeviceBuffer(4 * G["V"]) as buf:
        g.set_tune(chunk=1)
        with pytest.raises(c2g.C2GError) as ei:
            g.sssp_noalloc(0, buf)
        assert ei.value.code == c2g.EINVAL and "too large" in str(ei.value)
        g.set_tune(chunk=32)  # the default bound fits: the same graph runs
        s = int(np.argmax(np.diff(G["vlist"])))
        d, _ = _noalloc(g, s, buf, garbage=1)
        ref, _ = g.sssp(s)
        assert np.array_equal(d, ref)
    with c2g.Graph.from_csr(small["vlist"], small["elist"], small["wlist"], device=0) as g, \
            c2g.DeviceBuffer(4 * small["V"]) as buf:
        ref, _ = g.sssp(0)
        d, _ = _noalloc(g, 0, buf, garbage=3)
        assert np.array_equal(d, ref)
