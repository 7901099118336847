"""Pins the oracle (oracle/oracle.c) against independent facts, on the CPU.

A plausible slip in the oracle (a dropped term, a wrong sign or index, a
transposed operand) fails at least one of these."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as cg
import scipy.sparse.linalg as spla

import oracle
from conftest import INT_MAX, check_sssp_certificate, no_dangling_graph, random_graph, rmat_graph


def _scipy_dist(G, s):
    V = G["V"]
    A = sp.csr_matrix((G["wlist"].astype(np.float64) + 0.0, G["elist"], G["vlist"]), shape=(V, V))
    d = cg.dijkstra(A, indices=s)
    return np.where(np.isinf(d), INT_MAX, d).astype(np.int64)


@pytest.mark.parametrize("seed", range(6))
def test_dijkstra_vs_scipy(seed):
    G = random_graph(400, 3000, seed=seed, wmax=50)
    for s in (0, 7, 399):
        assert np.array_equal(oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], s), _scipy_dist(G, s))


def test_dijkstra_vs_scipy_rmat16():
    G = rmat_graph(16, 1, wmax=64)
    s = int(np.argmax(np.diff(G["vlist"])))
    d = oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], s)
    assert np.array_equal(d, _scipy_dist(G, s))
    check_sssp_certificate(G["vlist"], G["elist"], G["wlist"], s, d)


def _brute_bf(G, s):
    V = G["V"]
    d = np.full(V, np.iinfo(np.int64).max // 4, np.int64)
    d[s] = 0
    u = np.repeat(np.arange(V), np.diff(G["vlist"]))
    for _ in range(V):
        c = d[u] + G["wlist"]
        nd = d.copy()
        np.minimum.at(nd, G["elist"], c)
        if np.array_equal(nd, d):
            break
        d = nd
    return np.where(d >= np.iinfo(np.int64).max // 8, INT_MAX, d)


def test_dijkstra_vs_bruteforce_with_zero_weights():
    for seed in range(10):
        G = random_graph(30, 90, seed=100 + seed, wmax=3, zero_ok=True)
        for s in range(0, 30, 4):
            assert np.array_equal(oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], s), _brute_bf(G, s))


def test_dijkstra_closed_forms():
    # path 0->1->...->n-1 with weight i+1 on edge i: d[k] = k(k+1)/2
    n = 50
    vl = np.concatenate([np.arange(n), [n - 1]]).astype(np.int32)
    el = np.arange(1, n, dtype=np.int32)
    wl = np.arange(1, n, dtype=np.int32)
    d = oracle.dijkstra(vl, el, wl, 0)
    k = np.arange(n)
    assert np.array_equal(d, k * (k + 1) // 2)
    assert np.all(oracle.dijkstra(vl, el, wl, n - 1)[:-1] == INT_MAX)
    # shortcut beats the long path: 0->1 (1), 1->2 (1), 0->2 (5)
    vl = np.array([0, 2, 3, 3], np.int32)
    assert oracle.dijkstra(vl, np.array([1, 2, 2], np.int32), np.array([1, 5, 1], np.int32), 0).tolist() == [0, 1, 2]
    # overflow is reported, never wrapped
    vl = np.array([0, 1, 2, 2], np.int32)
    with pytest.raises(OverflowError):
        oracle.dijkstra(vl, np.array([1, 2], np.int32), np.array([2**30, 2**30], np.int32), 0)


def test_bellman_ford_lane_equals_dijkstra():
    G = rmat_graph(16, 1, wmax=64)
    deg = np.diff(G["vlist"])
    for s in [int(np.argmax(deg)), int(np.flatnonzero(deg == 1)[0]), 123]:
        d, rounds, rel = oracle.bellman_ford(G["vlist"], G["elist"], G["wlist"], s)
        assert np.array_equal(d, oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], s))
        reached = d != INT_MAX
        assert rel >= deg[reached].sum() or rounds == 0


# ---------------------------------------------------------------- PPR
def _ppr_direct(G, s, alpha):
    V = G["V"]
    P = sp.csr_matrix((G["flist"].astype(np.float64), G["elist"], G["vlist"]), shape=(V, V))
    A = sp.identity(V, format="csc") - (1 - alpha) * P.T.tocsc()
    b = np.zeros(V)
    b[s] = alpha
    return spla.spsolve(A, b)


def test_ppr_power_vs_direct_solve():
    for seed in range(4):
        G = random_graph(200, 1500, seed=seed)
        for s in (0, 50):
            x = oracle.ppr_power(G["vlist"], G["elist"], G["flist"], s, alpha=0.15)
            assert np.abs(x - _ppr_direct(G, s, 0.15)).max() < 1e-12


def test_ppr_power_cycle_closed_form():
    # directed n-cycle: pi_k = alpha (1-alpha)^k / (1 - (1-alpha)^n)
    n, a = 17, 0.15
    vl = np.arange(n + 1, dtype=np.int32)
    el = ((np.arange(n) + 1) % n).astype(np.int32)
    fl = np.ones(n, np.float32)
    x = oracle.ppr_power(vl, el, fl, 0, alpha=a)
    k = np.arange(n)
    assert np.abs(x - a * (1 - a) ** k / (1 - (1 - a) ** n)).max() < 1e-13


def test_ppr_dangling_mass_is_dropped():
    # 0 -> 1 only; 1 dangling: pi = [a, a(1-a)], total < 1
    x = oracle.ppr_power(np.array([0, 1, 1], np.int32), np.array([1], np.int32), np.ones(1, np.float32), 0, 0.2)
    assert np.allclose(x, [0.2, 0.16], atol=1e-15)


def test_forward_push_certificate():
    G = rmat_graph(14, 3, wmax=0, symmetric=True, want_f=True)
    s = int(np.argmax(np.diff(G["vlist"])))
    pi = oracle.ppr_power(G["vlist"], G["elist"], G["flist"], s, alpha=0.15)
    for eps in (1e-4, 1e-6, 1e-8):
        p, r, rounds, pushes = oracle.ppr_push(G["vlist"], G["elist"], G["flist"], s, alpha=0.15, eps=eps)
        deg = np.diff(G["vlist"])
        assert np.all(r <= eps * deg + 1e-12)            # termination rule
        l1 = np.abs(p - pi).sum()
        assert l1 <= r.sum() + 1e-5, (eps, l1, r.sum())  # |p - pi|_1 <= sum r (+ fp32)
        assert np.all(p <= pi + 1e-6)                     # push under-estimates
        # no dangling vertices (symmetric): mass is conserved at termination
        assert abs(p.sum() + r.sum() - 1.0) < 1e-4


# ---------------------------------------------------------- adsorption
def test_adsorption_vs_dense_matrix_power():
    G = random_graph(120, 700, seed=9)
    V, L = G["V"], 5
    rng = np.random.default_rng(1)
    sv = rng.choice(V, 12, replace=False).astype(np.int32)
    sl = rng.integers(0, L, 12).astype(np.int32)
    Y = oracle.adsorption(G["vlist"], G["elist"], G["flist"], sv, sl, L=L, p_inj=0.25, p_cont=0.7, iters=20)
    W = sp.csr_matrix((G["flist"].astype(np.float64), G["elist"], G["vlist"]), shape=(V, V)).toarray()
    Y0 = np.zeros((V, L))
    Y0[sv, sl] = 1
    Yt = Y0.copy()
    for _ in range(20):
        Yt = 0.25 * Y0 + 0.7 * W @ Yt
    assert np.abs(Y - Yt).max() < 1e-12
    # closed form of the limit: (I - 0.7 W)^-1 0.25 Y0 ; 20 steps are within 0.7^20
    Yinf = np.linalg.solve(np.eye(V) - 0.7 * W, 0.25 * Y0)
    assert np.abs(Y - Yinf).max() <= 0.7 ** 20 / 0.3 + 1e-12


def test_adsorption_push_invariant():
    # the reference's adsorption push (alpha 0.2, r[s] = 10, p[s] = 0.8*10,
    # constant threshold 1/V): with p' = p - (0.8 - 0.2)*10 e_s the forward-push
    # invariant gives |10 pi_s - p'|_1 = sum(r) on a graph without dangling
    # vertices, pi from the FP64 power iteration
    G = no_dangling_graph(400, 3000, seed=9)
    assert np.all(np.diff(G["vlist"]) > 0)
    for s in (0, 17, 399):
        p, r, rounds, pushes = oracle.adsorption_push(G["vlist"], G["elist"], G["flist"], s)
        pi = oracle.ppr_power(G["vlist"], G["elist"], G["flist"], s, alpha=0.2, tol=1e-15)
        q = p.astype(np.float64)
        q[s] -= (0.8 - 0.2) * 10.0
        l1 = np.abs(10.0 * pi - q).sum()
        assert abs(l1 - r.astype(np.float64).sum()) <= 1e-4 * 10, (s, l1, r.sum())
        assert r.max() <= 1.0 / 400 + 1e-7
        assert rounds > 1 and pushes > 0
    # a plausible mistake fails: crediting the source alpha*r0 instead of 0.8*r0
    q = p.astype(np.float64)
    assert abs(np.abs(10.0 * pi - q).sum() - r.astype(np.float64).sum()) > 1.0


def test_sssp_certificate_accepts_exact_and_rejects_each_perturbation():
    # the C certificate used on full-size GPU results: Dijkstra output passes
    # (scipy agrees with it above); every single slip fails with its own code
    G = rmat_graph(14, 1, wmax=64)
    vl, el, wl = G["vlist"], G["elist"], G["wlist"]
    s = int(np.argmax(np.diff(vl)))
    d = oracle.dijkstra(vl, el, wl, s)
    assert oracle.sssp_certificate(vl, el, wl, s, d) == 0
    fin = np.flatnonzero((d != INT_MAX) & (np.arange(G["V"]) != s))
    unr = np.flatnonzero(d == INT_MAX)
    rng = np.random.default_rng(0)
    v = int(rng.choice(fin))
    for delta, code in ((+1, None), (-1, None)):
        e = d.copy()
        e[v] += delta
        # +1: the tight in-edge of v now over-shoots (feasibility); -1: no tight in-edge
        assert oracle.sssp_certificate(vl, el, wl, s, e) in ((-2,) if delta > 0 else (-3, -2))
    e = d.copy()
    e[v] = INT_MAX  # a reachable vertex reported unreached
    assert oracle.sssp_certificate(vl, el, wl, s, e) == -2
    if len(unr):
        e = d.copy()
        e[int(unr[0])] = 5  # an unreachable vertex reported reached: no tight in-edge
        assert oracle.sssp_certificate(vl, el, wl, s, e) == -3
    e = d.copy()
    e[s] = 1
    assert oracle.sssp_certificate(vl, el, wl, s, e) == -1
    # every distance off by the same +k except the source: no tight edge out of s
    e = np.where(d == INT_MAX, d, d + 3).astype(np.int32)
    e[s] = 0
    assert oracle.sssp_certificate(vl, el, wl, s, e) != 0
    # a zero weight on an edge out of a reached vertex voids the proof: refused
    w0 = wl.copy()
    w0[vl[s]] = 0
    assert oracle.sssp_certificate(vl, el, w0, s, d) == -4


def test_dijkstra_many_equals_single():
    G = rmat_graph(14, 1, wmax=64)
    srcs = [0, 5, int(np.argmax(np.diff(G["vlist"])))]
    many = oracle.dijkstra_many(G["vlist"], G["elist"], G["wlist"], srcs, threads=3)
    for s, d in zip(srcs, many):
        assert np.array_equal(d, oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], s))


@pytest.mark.parametrize("mode", [0, 1])
def test_adsorption_push_modes_invariant(mode):
    # w-purn's threshold (mode 0, rmax 1/(V+1)) and w-base-perk's credit rule
    # (mode 1: p[s] = alpha*r0 at seeding, p[v] += (1-alpha)*r[v] at activation).
    # With A the total residual activated (the source's r0 included), the
    # push invariant |r0*pi_s - alpha*A|_1 = sum(r) holds for either rule;
    # A follows from p through the rule's credits.
    G = no_dangling_graph(400, 3000, seed=11)
    alpha, r0 = 0.2, 10.0
    V = G["V"]
    rmax = 1.0 / (V + 1) if mode == 0 else 1.0 / V
    for s in (3, 250):
        p, r, rounds, pushes = oracle.adsorption_push_mode(G["vlist"], G["elist"], G["flist"], s, mode, rmax)
        pi = oracle.ppr_power(G["vlist"], G["elist"], G["flist"], s, alpha=alpha, tol=1e-15)
        q = p.astype(np.float64)
        seed_credit = 0.8 if mode == 0 else alpha
        credit = alpha if mode == 0 else 1.0 - alpha
        q[s] -= seed_credit * r0
        aA = alpha * q / credit
        aA[s] += alpha * r0
        l1 = np.abs(r0 * pi - aA).sum()
        assert abs(l1 - r.astype(np.float64).sum()) <= 1e-3, (mode, s, l1, r.sum())
        assert r.max() <= rmax * (1 + 1e-6) and rounds > 1 and pushes > 0
        # the other rule's credit breaks the invariant
        wrong = alpha * q / (1.0 - credit)
        wrong[s] += alpha * r0
        assert abs(np.abs(r0 * pi - wrong).sum() - r.astype(np.float64).sum()) > 0.5
    # mode 0 with rmax 1/V is the w-base push exactly
    p0, r0v, _, _ = oracle.adsorption_push_mode(G["vlist"], G["elist"], G["flist"], 3, 0, 1.0 / V)
    pb, rb, _, _ = oracle.adsorption_push(G["vlist"], G["elist"], G["flist"], 3)
    assert np.array_equal(p0, pb) and np.array_equal(r0v, rb)
