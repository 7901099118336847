"""Multi-label adsorption on the GPU against the oracle's FP64 Jacobi.

Tolerance (FP32 accumulation over 20 iterations): max |Y_gpu - Y| <= 2e-6 and
mean |Y_gpu - Y| <= 2e-7 per (vertex, label).
"""
import numpy as np
import pytest

import c2g
import oracle
from conftest import no_dangling_graph, random_graph, rmat_graph

pytestmark = pytest.mark.gpu


def _seeds(V, L, frac, seed):
    rng = np.random.default_rng(seed)
    n = max(1, int(V * frac))
    sv = rng.choice(V, n, replace=False).astype(np.int32)
    sl = rng.integers(0, L, n).astype(np.int32)
    return sv, sl


def _check(Y, Yo):
    d = np.abs(Y.astype(np.float64) - Yo)
    assert d.max() <= 2e-6, d.max()
    assert d.mean() <= 2e-7, d.mean()


def test_rmat_16_labels(gpu):
    # C5's recipe at scale 14: 16 labels on 1% of vertices, 0.25/0.70/0.05, 20 iterations
    G = rmat_graph(14, 5, wmax=0, want_f=True)
    sv, sl = _seeds(G["V"], 16, 0.01, 5)
    with c2g.Graph.from_csr(G["vlist"], G["elist"], flist=G["flist"], device=0) as g:
        Y, st = g.adsorption(sv, sl, L=16, p_inj=0.25, p_cont=0.70, p_abnd=0.05, iters=20)
    Yo = oracle.adsorption(G["vlist"], G["elist"], G["flist"], sv, sl, L=16, p_inj=0.25, p_cont=0.70, iters=20)
    _check(Y, Yo)
    assert st.lane == c2g.LANE_GPU
    # mass bound: each row's label mass is at most 1
    assert Y.sum(axis=1).max() <= 1.0 + 1e-5


@pytest.mark.parametrize("L", [1, 5, 16, 17, 40, 64])
def test_label_counts_and_hub_rows(gpu, L):
    G = rmat_graph(12, 7, wmax=0, symmetric=True, want_f=True)  # hubs > 256 edges are split
    assert np.diff(G["vlist"]).max() > 256
    sv, sl = _seeds(G["V"], L, 0.05, L)
    with c2g.Graph.from_csr(G["vlist"], G["elist"], flist=G["flist"], device=0) as g:
        Y, _ = g.adsorption(sv, sl, L=L, iters=7)
    _check(Y, oracle.adsorption(G["vlist"], G["elist"], G["flist"], sv, sl, L=L, iters=7))


def test_zero_iterations_and_errors(gpu):
    G = random_graph(100, 300, seed=2)
    sv = np.array([3, 3, 9], np.int32)
    sl = np.array([1, 1, 0], np.int32)
    with c2g.Graph.from_csr(G["vlist"], G["elist"], flist=G["flist"], device=0) as g:
        Y, _ = g.adsorption(sv, sl, L=2, iters=0)
        assert Y[3, 1] == 1.0 and Y[9, 0] == 1.0 and Y.sum() == 2.0
        with pytest.raises(c2g.C2GError):
            g.adsorption(sv, sl, L=2, p_inj=0.5, p_cont=0.5, p_abnd=0.5)
        with pytest.raises(c2g.C2GError):
            g.adsorption(np.array([100], np.int32), np.array([0], np.int32), L=2)


def test_multi_chunk_result_copy(gpu):
    # 524,288 x 40 labels x 4 B = 80 MiB of Y: the result leaves the device in
    # two pinned 32 MiB chunks and a partial third that reuses the first buffer
    G = rmat_graph(19, 6, wmax=0, want_f=True)
    sv, sl = _seeds(G["V"], 40, 0.01, 6)
    with c2g.Graph.from_csr(G["vlist"], G["elist"], flist=G["flist"], device=0) as g:
        Y, _ = g.adsorption(sv, sl, L=40, iters=4)
    assert 2 * (32 << 20) < Y.nbytes < 3 * (32 << 20)
    _check(Y, oracle.adsorption(G["vlist"], G["elist"], G["flist"], sv, sl, L=40, iters=4))


def test_reference_push_mode(gpu):
    # the reference's own adsorption kernel (src/adsorption/w-base/Main.cu):
    # single-source push, alpha 0.2, r[s] = 10, p[s] = 0.8*10, threshold 1/V.
    # Against the oracle's serial restatement (L1 <= 2e-4: FP32 atomics in any
    # order, ten times the unit mass) and the push invariant
    # |10 pi_s - (p - 6 e_s)|_1 = sum(r) with pi from the FP64 power iteration.
    G = no_dangling_graph(3000, 30000, seed=3)
    V = G["V"]
    with c2g.Graph.from_csr(G["vlist"], G["elist"], flist=G["flist"], device=0) as g:
        for s in (0, 5, V - 1):
            p, r, st = g.adsorption_push(s)
            po, ro, rounds, _ = oracle.adsorption_push(G["vlist"], G["elist"], G["flist"], s)
            assert np.abs(p.astype(np.float64) - po).sum() <= 2e-4, s
            assert r.max() <= 1.0 / V + 1e-7
            pi = oracle.ppr_power(G["vlist"], G["elist"], G["flist"], s, alpha=0.2, tol=1e-15)
            q = p.astype(np.float64)
            q[s] -= 6.0
            assert abs(np.abs(10.0 * pi - q).sum() - r.astype(np.float64).sum()) <= 1e-3, s
            assert st.iterations > 1


def test_reference_push_mode_refuses_pruned_graph(gpu):
    G = random_graph(300, 900, seed=4)
    P = c2g.prune_ppr_mem(G["vlist"], G["elist"], G["flist"])
    with c2g.Graph.from_csr(P["vlist"], P["elist"], flist=P["wlist"], hvlist=P["hvlist"], helist=P["helist"],
                            hwf=P["hwlist"], variant=c2g.VARIANT_PPR_PURN, device=0) as g:
        with pytest.raises(c2g.C2GError) as e:
            g.adsorption_push(0)
        assert e.value.code == c2g.ENOTSUP


@pytest.mark.slow
def test_c5_full_scale_against_fp64(gpu):
    # the adsorption workload at full size: R-MAT scale 23, seed 5, 16 labels
    # on 1% of the vertices, 0.25/0.70/0.05, 20 iterations, against the FP64
    # Jacobi of the same recurrence; errors are printed for the report
    G = rmat_graph(23, 5, wmax=255, want_f=True)
    sv, sl = _seeds(G["V"], 16, 0.01, 5)
    with c2g.Graph.from_csr(G["vlist"], G["elist"], G["wlist"], G["flist"], device=0) as g:
        Y, _ = g.adsorption(sv, sl, L=16, p_inj=0.25, p_cont=0.70, p_abnd=0.05, iters=20)
    Yo = oracle.adsorption(G["vlist"], G["elist"], G["flist"], sv, sl, L=16, p_inj=0.25, p_cont=0.70, iters=20)
    d = np.abs(Y.astype(np.float64) - Yo)
    print(f"\nC5 scale 23: max-abs {d.max():.3g} mean-abs {d.mean():.3g} L1 per label {d.sum() / 16:.3g}")
    _check(Y, Yo)


@pytest.mark.parametrize("L", [16, 5, 40])
def test_relabelled_layout(gpu, L):
    # the run on the copy of the CSR renumbered by in-degree returns rows in the
    # caller's ids: against the oracle, and against the identity layout (rows
    # of at most 256 edges sum the same terms in the same order; longer rows
    # are split and added atomically)
    G = rmat_graph(12, 7, wmax=0, symmetric=True, want_f=True)
    sv, sl = _seeds(G["V"], L, 0.05, 3 + L)
    sv[0] = int(np.argmax(np.diff(G["vlist"])))  # a seed on a split hub row
    with c2g.Graph.from_csr(G["vlist"], G["elist"], flist=G["flist"], device=0) as g:
        Y0, _ = g.adsorption(sv, sl, L=L, iters=9)
        g.set_tune(ads_relabel=1)
        Y1, _ = g.adsorption(sv, sl, L=L, iters=9)
        Y2, _ = g.adsorption(sv, sl, L=L, iters=9)  # relabelled copy reused
        g.set_tune(ads_relabel=0)
        Y3, _ = g.adsorption(sv, sl, L=L, iters=9)  # and back
    _check(Y1, oracle.adsorption(G["vlist"], G["elist"], G["flist"], sv, sl, L=L, iters=9))
    assert np.abs(Y1 - Y0).max() <= 1e-6
    assert np.abs(Y2 - Y1).max() <= 1e-6
    assert np.abs(Y3 - Y0).max() <= 1e-6
    # no row above 256 edges: no atomics, so both layouts give the same bits
    H = random_graph(500, 4000, seed=L)
    hv, hl = _seeds(500, L, 0.05, L)
    with c2g.Graph.from_csr(H["vlist"], H["elist"], flist=H["flist"], device=0) as g:
        Z0, _ = g.adsorption(hv, hl, L=L, iters=9)
        g.set_tune(ads_relabel=1)
        Z1, _ = g.adsorption(hv, hl, L=L, iters=9)
    assert np.array_equal(Z0, Z1)
