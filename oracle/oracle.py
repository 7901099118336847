"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of liboracle.so (oracle.c).

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
as the checker; never by the product library.  See oracle.c for what each
function restates (reference file:line) and what pins it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

P = np.ctypeslib.ndpointer


def build():
    subprocess.run(["make", "-C", _HERE, "liboracle.so"], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        i32, f32, f64, i64 = (P(np.int32, flags="C"), P(np.float32, flags="C"), P(np.float64, flags="C"),
                              P(np.int64, flags="C"))
        L.orc_dijkstra.argtypes = [C.c_int32, i32, i32, i32, C.c_int32, i32]
        L.orc_sssp_certificate.argtypes = [C.c_int32, i32, i32, i32, C.c_int32, i32]
        L.orc_bellman_ford.argtypes = [C.c_int32, i32, i32, i32, i32, i32, C.c_int32, i64]
        L.orc_ppr_power.argtypes = [C.c_int32, i32, i32, f32, C.c_int32, C.c_double, C.c_double, C.c_int32, f64]
        L.orc_forward_push.argtypes = [C.c_int32, i32, i32, f32, C.c_float, C.c_float, f32, f32, i32,
                                       C.c_int32, C.c_int32, i64]
        L.orc_ppr_push.argtypes = [C.c_int32, i32, i32, f32, C.c_int32, C.c_float, C.c_float, f32, f32,
                                   C.c_int32, i64]
        L.orc_adsorption_push.argtypes = [C.c_int32, i32, i32, f32, C.c_int32, f32, f32, C.c_int32, i64]
        L.orc_adsorption_push_mode.argtypes = [C.c_int32, i32, i32, f32, C.c_int32, C.c_int32, C.c_float, f32, f32,
                                               C.c_int32, i64]
        L.orc_adsorption.argtypes = [C.c_int32, i32, i32, f32, i32, i32, C.c_int32, C.c_int32, C.c_double,
                                     C.c_double, C.c_int32, f64]
        L.orc_adsorption.restype = None
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dt)


def dijkstra(vlist, elist, wlist, src):
    V = len(vlist) - 1
    d = np.empty(V, np.int32)
    rc = lib().orc_dijkstra(V, _c(vlist, np.int32), _c(elist, np.int32), _c(wlist, np.int32), int(src), d)
    if rc == -1:
        raise OverflowError("a shortest distance does not fit in int32")
    if rc != 0:
        raise ValueError(f"orc_dijkstra rc={rc}")
    return d


def sssp_certificate(vlist, elist, wlist, src, dist):
    """0 if dist is the exact int32 distance vector from src (weights >= 1),
    else the negative code of the first check that fails (see oracle.c)."""
    V = len(vlist) - 1
    return lib().orc_sssp_certificate(V, _c(vlist, np.int32), _c(elist, np.int32), _c(wlist, np.int32), int(src),
                                      _c(dist, np.int32))


def dijkstra_many(vlist, elist, wlist, srcs, threads=None):
    """orc_dijkstra for several sources on host threads (ctypes releases the
    GIL during each call); returns a list of distance vectors."""
    from concurrent.futures import ThreadPoolExecutor

    vl, el, wl = _c(vlist, np.int32), _c(elist, np.int32), _c(wlist, np.int32)
    threads = threads or min(16, os.cpu_count() or 4, max(1, len(srcs)))
    with ThreadPoolExecutor(threads) as ex:
        return list(ex.map(lambda s: dijkstra(vl, el, wl, int(s)), srcs))


def bellman_ford(vlist, elist, wlist, src):
    """Reference CPU lane from {s:0}, frontier {s}.  Returns (dist, rounds, relaxations)."""
    V = len(vlist) - 1
    d = np.full(V, np.iinfo(np.int32).max, np.int32)
    d[src] = 0
    fr = np.array([src], np.int32)
    rel = np.zeros(1, np.int64)
    rounds = lib().orc_bellman_ford(V, _c(vlist, np.int32), _c(elist, np.int32), _c(wlist, np.int32), d, fr, 1, rel)
    return d, rounds, int(rel[0])


def ppr_power(vlist, elist, flist, src, alpha=0.15, tol=1e-15, max_iter=100000):
    V = len(vlist) - 1
    x = np.empty(V, np.float64)
    lib().orc_ppr_power(V, _c(vlist, np.int32), _c(elist, np.int32), _c(flist, np.float32), int(src),
                        float(alpha), float(tol), int(max_iter), x)
    return x


def ppr_push(vlist, elist, flist, src, alpha=0.15, eps=1e-7, max_rounds=1 << 30):
    V = len(vlist) - 1
    p = np.empty(V, np.float32)
    r = np.empty(V, np.float32)
    pushes = np.zeros(1, np.int64)
    rounds = lib().orc_ppr_push(V, _c(vlist, np.int32), _c(elist, np.int32), _c(flist, np.float32), int(src),
                                float(alpha), float(eps), p, r, int(max_rounds), pushes)
    return p, r, rounds, int(pushes[0])


def adsorption_push(vlist, elist, flist, src, max_rounds=1 << 30):
    """The reference's adsorption kernel (single-source push, alpha 0.2,
    r[s] = 10, constant threshold 1/V): returns p, r, rounds, pushes."""
    V = len(vlist) - 1
    p = np.empty(V, np.float32)
    r = np.empty(V, np.float32)
    pushes = np.zeros(1, np.int64)
    rounds = lib().orc_adsorption_push(V, _c(vlist, np.int32), _c(elist, np.int32), _c(flist, np.float32),
                                       int(src), p, r, int(max_rounds), pushes)
    return p, r, rounds, int(pushes[0])


def adsorption_push_mode(vlist, elist, flist, src, mode, rmax, max_rounds=1 << 30):
    """The reference's adsorption variants: mode 0 = w-base's push with the
    given rmax (w-purn: 1/(V+1) on the pruned CSR), mode 1 = w-base-perk's
    credit rule.  Returns p, r, rounds, pushes."""
    V = len(vlist) - 1
    p = np.empty(V, np.float32)
    r = np.empty(V, np.float32)
    pushes = np.zeros(1, np.int64)
    rounds = lib().orc_adsorption_push_mode(V, _c(vlist, np.int32), _c(elist, np.int32), _c(flist, np.float32),
                                            int(src), int(mode), float(rmax), p, r, int(max_rounds), pushes)
    return p, r, rounds, int(pushes[0])


def adsorption(vlist, elist, flist, seed_v, seed_l, L=16, p_inj=0.25, p_cont=0.70, iters=20):
    V = len(vlist) - 1
    Y = np.empty((V, L), np.float64)
    lib().orc_adsorption(V, _c(vlist, np.int32), _c(elist, np.int32), _c(flist, np.float32), _c(seed_v, np.int32),
                         _c(seed_l, np.int32), len(seed_v), int(L), float(p_inj), float(p_cont), int(iters), Y)
    return Y
