"""ctypes binding of libc2g.so (include/c2g.h) for tests, bench and scripts.

Mirrors the reference's per-path surface: a graph loaded from a dataset
directory (Graph(dir) in src/sssp/base/Graph.h:30-51) or from arrays, and one
call per query (CalcuSSSP / calcuatePPR launches, src/sssp/base/Main.cu:66-69,
src/ppr/base/Main.cu:79-83).  Every GPU call goes through the HIP library;
when the library or a device is missing the call raises C2GError.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libc2g.so")

OK, EINVAL, EIO, EFORMAT, ENOMEM, EDEVICE, ENOTSUP, EBUSY, ETIMEOUT = 0, -1, -2, -3, -4, -5, -6, -7, -8
VARIANT_ORIGIN, VARIANT_SSSP_PURN, VARIANT_PPR_PURN = 0, 1, 2
ADS_PUSH_BASE, ADS_PUSH_PERK = 0, 1
LANE_GPU, LANE_HOST, LANE_MIXED = 0, 1, 2
Q_SSSP, Q_PPR, Q_ADS = 0, 1, 2
INT_MAX = 2**31 - 1


class C2GError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"c2g error {code}: {msg}")
        self.code = code


class Stats(C.Structure):
    _fields_ = [("lane", C.c_int32), ("iterations", C.c_int32), ("host_iterations", C.c_int32),
                ("reached", C.c_int32), ("frontier_total", C.c_int64), ("edges_relaxed", C.c_int64),
                ("improvements", C.c_int64), ("reached_edges", C.c_int64), ("csr_edges", C.c_int64),
                ("items", C.c_int64), ("ms", C.c_double), ("kernel_ms", C.c_double),
                ("pull_rounds", C.c_int32), ("pad", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class Round(C.Structure):
    _fields_ = [("frontier", C.c_int64), ("edges", C.c_int64), ("lane_edges", C.c_int64),
                ("improvements", C.c_int64), ("items", C.c_int64), ("ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Opts(C.Structure):
    _fields_ = [("device", C.c_int32), ("load_flist", C.c_int32), ("verbose", C.c_int32),
                ("no_device", C.c_int32)]


class Tune(C.Structure):
    _fields_ = [("block", C.c_int32), ("grid_per_cu", C.c_int32), ("chunk", C.c_int32),
                ("batch_chunk", C.c_int32), ("iters_per_graph", C.c_int32), ("use_graph", C.c_int32),
                ("delta", C.c_int32), ("profile", C.c_int32), ("grid_per_cu_single", C.c_int32),
                ("ppr_pull", C.c_int32), ("ads_relabel", C.c_int32), ("block_single", C.c_int32),
                ("xcd_split", C.c_int32), ("light", C.c_int32), ("batch_order", C.c_int32),
                ("ppr_grid_per_cu", C.c_int32), ("batch_relabel", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerOpts(C.Structure):
    _fields_ = [("max_batch", C.c_int32), ("max_inflight", C.c_int32), ("hbm_reserve", C.c_int64),
                ("fake_free_hbm", C.c_int64), ("host_lane", C.c_int32), ("device_failed", C.c_int32),
                ("alpha", C.c_float), ("eps", C.c_float), ("ads_labels", C.c_int32),
                ("ads_iters", C.c_int32), ("p_inj", C.c_float), ("p_cont", C.c_float),
                ("p_abnd", C.c_float), ("sssp_slots", C.c_int32), ("ads_grid_per_cu", C.c_int32),
                ("deadline_ms", C.c_int32), ("fail_inject", C.c_int32), ("fail_after", C.c_int32),
                ("dry_device", C.c_int32), ("dry_step_ms", C.c_int32), ("reserved", C.c_int32 * 1)]


SLANE_SLOTS, SLANE_GROUP, SLANE_PPR, SLANE_ADS = 1, 2, 4, 8


class ServerStats(C.Structure):
    _fields_ = [("budget", C.c_int64), ("charged", C.c_int64), ("lane_bytes", C.c_int64 * 4),
                ("sssp_slots", C.c_int32), ("lanes_enabled", C.c_int32), ("submitted", C.c_int64),
                ("completed", C.c_int64), ("failed", C.c_int64), ("cancelled", C.c_int64),
                ("to_host", C.c_int64), ("lane_failures", C.c_int64), ("queued", C.c_int32),
                ("running", C.c_int32)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["lane_bytes"] = list(self.lane_bytes)
        return d


class ResultInfo(C.Structure):
    _fields_ = [("id", C.c_int64), ("kind", C.c_int32), ("status", C.c_int32), ("lane", C.c_int32),
                ("pad", C.c_int32), ("t_submit_ms", C.c_double), ("t_admit_ms", C.c_double),
                ("t_done_ms", C.c_double)]


_lib = None

P_I32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
P_F32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
P_I64 = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
P_U8 = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def lib():
    """Loads libc2g.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise C2GError(EDEVICE, f"{LIB_PATH} not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.c2g_last_error.restype = C.c_char_p
        L.c2g_version.restype = C.c_char_p
        L.c2g_read_i32.restype = C.c_int64
        L.c2g_read_f32.restype = C.c_int64
        L.c2g_server_submit.restype = C.c_int64
        L.c2g_rmat.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_uint64,
                               C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.POINTER(C.c_int32)),
                               C.POINTER(C.POINTER(C.c_int32)), C.POINTER(C.POINTER(C.c_int32)),
                               C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.c_int64)]
        L.c2g_free.argtypes = [C.c_void_p]
        L.c2g_server_submit.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
        L.c2g_server_stop.argtypes = [C.c_void_p]
        L.c2g_server_start.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]
        L.c2g_server_poll.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.POINTER(ResultInfo)]
        L.c2g_server_cancel.argtypes = [C.c_void_p, C.c_int64]
        L.c2g_server_get_stats.argtypes = [C.c_void_p, C.POINTER(ServerStats)]
        _lib = L
    return _lib


def last_error() -> str:
    return lib().c2g_last_error().decode(errors="replace")


def check(rc):
    if rc < 0:
        raise C2GError(rc, last_error())
    return rc


def device_available() -> bool:
    return bool(lib().c2g_device_available())


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


def _take(p, n, dtype):
    """Copies n elements from a library-allocated pointer and frees it."""
    if not p:
        return None
    arr = np.ctypeslib.as_array(p, shape=(max(int(n), 1),))[: int(n)].astype(dtype, copy=True)
    lib().c2g_free(C.cast(p, C.c_void_p))
    return arr


def rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1, symmetric=False, wmax=64,
         want_f=False):
    """Seeded R-MAT CSR: returns dict(V, E, vlist, elist, wlist|None, flist|None)."""
    L = lib()
    vl = C.POINTER(C.c_int32)()
    el = C.POINTER(C.c_int32)()
    wl = C.POINTER(C.c_int32)()
    fl = C.POINTER(C.c_float)()
    E = C.c_int64()
    check(L.c2g_rmat(scale, edge_factor, a, b, c, seed, int(symmetric), int(wmax), int(want_f),
                     C.byref(vl), C.byref(el), C.byref(wl) if wmax > 0 else None,
                     C.byref(fl) if want_f else None, C.byref(E)))
    V = 1 << scale
    return dict(V=V, E=E.value, vlist=_take(vl, V + 1, np.int32), elist=_take(el, E.value, np.int32),
                wlist=_take(wl, E.value, np.int32) if wmax > 0 else None,
                flist=_take(fl, E.value, np.float32) if want_f else None)


def write_dataset(ds, vlist, elist, wlist=None, flist=None):
    V = len(vlist) - 1
    check(lib().c2g_write_dataset(ds.encode(), V, _ptr(np.ascontiguousarray(vlist, np.int32), C.c_int32),
                                  _ptr(np.ascontiguousarray(elist, np.int32), C.c_int32),
                                  _ptr(None if wlist is None else np.ascontiguousarray(wlist, np.int32), C.c_int32),
                                  _ptr(None if flist is None else np.ascontiguousarray(flist, np.float32), C.c_float)))


def read_i32(path):
    n = check(lib().c2g_read_i32(path.encode(), None, 0))
    a = np.empty(n, np.int32)
    check(lib().c2g_read_i32(path.encode(), _ptr(a, C.c_int32), n))
    return a


def read_f32(path):
    n = check(lib().c2g_read_f32(path.encode(), None, 0))
    a = np.empty(n, np.float32)
    check(lib().c2g_read_f32(path.encode(), _ptr(a, C.c_float), n))
    return a


def prune_sssp(ds):
    st = np.zeros(3, np.int64)
    check(lib().c2g_prune_sssp(ds.encode(), _ptr(st, C.c_int64)))
    return st


def prune_ppr(ds):
    st = np.zeros(3, np.int64)
    check(lib().c2g_prune_ppr(ds.encode(), _ptr(st, C.c_int64)))
    return st


def prune_sssp_mem(vlist, elist, wlist):
    L = lib()
    outs = [C.POINTER(C.c_int32)() for _ in range(7)]
    npo = C.c_int64()
    V = len(vlist) - 1
    vlist = np.ascontiguousarray(vlist, np.int32)
    elist = np.ascontiguousarray(elist, np.int32)
    wlist = np.ascontiguousarray(wlist, np.int32)
    check(L.c2g_prune_sssp_mem(V, _ptr(vlist, C.c_int32), _ptr(elist, C.c_int32), _ptr(wlist, C.c_int32),
                               *[C.byref(o) for o in outs], C.byref(npo)))
    pv = _take(outs[0], V + 1, np.int32)
    E2 = int(pv[V])
    hv = _take(outs[3], V + 1, np.int32)
    HE = int(hv[V])
    return dict(vlist=pv, elist=_take(outs[1], E2, np.int32), wlist=_take(outs[2], E2, np.int32),
                hvlist=hv, helist=_take(outs[4], HE, np.int32), hwlist=_take(outs[5], HE, np.int32),
                porder=_take(outs[6], npo.value, np.int32))


def prune_ppr_mem(vlist, elist, flist):
    L = lib()
    V = len(vlist) - 1
    ov, oe, ohv, ohe = (C.POINTER(C.c_int32)() for _ in range(4))
    ow, ohw = C.POINTER(C.c_float)(), C.POINTER(C.c_float)()
    vlist = np.ascontiguousarray(vlist, np.int32)
    elist = np.ascontiguousarray(elist, np.int32)
    flist = np.ascontiguousarray(flist, np.float32)
    check(L.c2g_prune_ppr_mem(V, _ptr(vlist, C.c_int32), _ptr(elist, C.c_int32), _ptr(flist, C.c_float),
                              C.byref(ov), C.byref(oe), C.byref(ow), C.byref(ohv), C.byref(ohe), C.byref(ohw)))
    pv = _take(ov, V + 1, np.int32)
    hv = _take(ohv, V + 1, np.int32)
    return dict(vlist=pv, elist=_take(oe, int(pv[V]), np.int32), wlist=_take(ow, int(pv[V]), np.float32),
                hvlist=hv, helist=_take(ohe, int(hv[V]), np.int32), hwlist=_take(ohw, int(hv[V]), np.float32))


_hip = None


def _hip_rt():
    """The HIP runtime libc2g.so links (for caller-owned device buffers)."""
    global _hip
    if _hip is None:
        lib()
        try:
            H = C.CDLL("libamdhip64.so.7")
        except OSError:
            H = C.CDLL("/opt/rocm/lib/libamdhip64.so")
        H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        H.hipFree.argtypes = [C.c_void_p]
        H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        H.hipSetDevice.argtypes = [C.c_int]
        _hip = H
    return _hip


class DeviceBuffer:
    """A caller-owned device allocation (hipMalloc through the HIP runtime that
    libc2g.so uses), for entry points that write into device memory."""

    def __init__(self, nbytes, device=0):
        H = _hip_rt()
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        if H.hipSetDevice(int(device)) != 0 or H.hipMalloc(C.byref(self.ptr), max(self.nbytes, 1)) != 0:
            raise C2GError(ENOMEM, f"hipMalloc of {nbytes} bytes failed")

    def to_numpy(self, dtype, n):
        a = np.empty(int(n), dtype)
        if _hip_rt().hipMemcpy(a.ctypes.data_as(C.c_void_p), self.ptr, a.nbytes, 2) != 0:  # device to host
            raise C2GError(EDEVICE, "hipMemcpy device to host failed")
        return a

    def fill(self, arr):
        arr = np.ascontiguousarray(arr)
        if _hip_rt().hipMemcpy(self.ptr, arr.ctypes.data_as(C.c_void_p), arr.nbytes, 1) != 0:  # host to device
            raise C2GError(EDEVICE, "hipMemcpy host to device failed")

    def free(self):
        if self.ptr:
            _hip_rt().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.free()


class Graph:
    """Owning wrapper of a c2g_graph handle."""

    def __init__(self, handle):
        self._h = handle
        V = C.c_int32()
        E = C.c_int64()
        HE = C.c_int64()
        var = C.c_int32()
        check(lib().c2g_graph_info(self._h, C.byref(V), C.byref(E), C.byref(HE), C.byref(var)))
        self.V, self.E, self.HE, self.variant = V.value, E.value, HE.value, var.value

    @classmethod
    def open(cls, ds, variant=VARIANT_ORIGIN, device=-1, load_flist=False, no_device=False, verbose=False):
        h = C.c_void_p()
        o = Opts(device, int(load_flist), int(verbose), int(no_device))
        check(lib().c2g_open(ds.encode(), variant, C.byref(o), C.byref(h)))
        return cls(h)

    @classmethod
    def from_csr(cls, vlist, elist, wlist=None, flist=None, hvlist=None, helist=None, hwlist=None,
                 hwf=None, variant=VARIANT_ORIGIN, device=0, no_device=False):
        keep = []

        def arr(a, dt, ct):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dt)
            keep.append(a)
            return _ptr(a, ct)

        V = len(vlist) - 1
        h = C.c_void_p()
        o = Opts(device, 0, 0, int(no_device))
        check(lib().c2g_graph_from_csr(V, arr(vlist, np.int32, C.c_int32), arr(elist, np.int32, C.c_int32),
                                       arr(wlist, np.int32, C.c_int32), arr(flist, np.float32, C.c_float),
                                       arr(hvlist, np.int32, C.c_int32), arr(helist, np.int32, C.c_int32),
                                       arr(hwlist, np.int32, C.c_int32), arr(hwf, np.float32, C.c_float),
                                       variant, C.byref(o), C.byref(h)))
        return cls(h)

    def close(self):
        if self._h:
            lib().c2g_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- tuning
    def tune(self):
        t = Tune()
        check(lib().c2g_get_tune(self._h, C.byref(t)))
        return t

    def set_tune(self, **kw):
        t = self.tune()
        for k, v in kw.items():
            setattr(t, k, v)
        check(lib().c2g_set_tune(self._h, C.byref(t)))

    # ---- SSSP
    def sssp(self, src):
        d = np.empty(self.V, np.int32)
        st = Stats()
        check(lib().c2g_sssp(self._h, int(src), _ptr(d, C.c_int32), C.byref(st)))
        return d, st

    def sssp_rounds(self, src, cap=4096):
        """c2g_sssp_rounds: distances, stats and the per-round trace (list of dicts)."""
        d = np.empty(self.V, np.int32)
        rs = (Round * cap)()
        n = C.c_int32()
        st = Stats()
        check(lib().c2g_sssp_rounds(self._h, int(src), _ptr(d, C.c_int32), rs, cap, C.byref(n), C.byref(st)))
        return d, st, [rs[i].as_dict() for i in range(min(n.value, cap))]

    def sssp_batch_rounds(self, srcs, cap=4096):
        srcs = np.ascontiguousarray(srcs, np.int32)
        rs = (Round * cap)()
        n = C.c_int32()
        st = Stats()
        check(lib().c2g_sssp_batch_rounds(self._h, _ptr(srcs, C.c_int32), len(srcs), rs, cap, C.byref(n),
                                          C.byref(st)))
        return st, [rs[i].as_dict() for i in range(min(n.value, cap))]

    def sssp_noalloc(self, src, dbuf):
        """c2g_sssp_noalloc: distances into the DeviceBuffer dbuf (int32[V])."""
        if dbuf.nbytes < 4 * self.V:
            raise C2GError(EINVAL, "device buffer smaller than V int32")
        st = Stats()
        check(lib().c2g_sssp_noalloc(self._h, int(src), C.cast(dbuf.ptr, C.POINTER(C.c_int32)), C.byref(st)))
        return st

    def sssp_batch(self, srcs):
        srcs = np.ascontiguousarray(srcs, np.int32)
        d = np.empty((len(srcs), self.V), np.int32)
        st = Stats()
        check(lib().c2g_sssp_batch(self._h, _ptr(srcs, C.c_int32), len(srcs), _ptr(d, C.c_int32), C.byref(st)))
        return d, st

    def sssp_batch_resident(self, srcs):
        srcs = np.ascontiguousarray(srcs, np.int32)
        sums = np.zeros(len(srcs), np.int64)
        reached = np.zeros(len(srcs), np.int32)
        st = Stats()
        check(lib().c2g_sssp_batch_resident(self._h, _ptr(srcs, C.c_int32), len(srcs), _ptr(sums, C.c_int64),
                                            _ptr(reached, C.c_int32), C.byref(st)))
        return sums, reached, st

    def sssp_each_resident(self, srcs):
        srcs = np.ascontiguousarray(srcs, np.int32)
        sums = np.zeros(len(srcs), np.int64)
        reached = np.zeros(len(srcs), np.int32)
        st = Stats()
        check(lib().c2g_sssp_each_resident(self._h, _ptr(srcs, C.c_int32), len(srcs), _ptr(sums, C.c_int64),
                                           _ptr(reached, C.c_int32), C.byref(st)))
        return sums, reached, st

    def sssp_budget(self, src, budget):
        d = np.empty(self.V, np.int32)
        fr = np.empty(self.V, np.int32)
        cnt = C.c_int32()
        st = Stats()
        check(lib().c2g_sssp_budget(self._h, int(src), int(budget), _ptr(d, C.c_int32), _ptr(fr, C.c_int32),
                                    C.byref(cnt), C.byref(st)))
        return d, fr[: cnt.value].copy(), st

    def sssp_migrate(self, src, budget):
        d = np.empty(self.V, np.int32)
        st = Stats()
        check(lib().c2g_sssp_migrate(self._h, int(src), int(budget), _ptr(d, C.c_int32), C.byref(st)))
        return d, st

    def host_sssp(self, src):
        d = np.empty(self.V, np.int32)
        st = Stats()
        check(lib().c2g_host_sssp(self._h, int(src), _ptr(d, C.c_int32), C.byref(st)))
        return d, st

    def host_bellman_ford(self, dist, frontier, via_hidden=None):
        dist = np.ascontiguousarray(dist, np.int32)
        frontier = np.ascontiguousarray(frontier, np.int32)
        st = Stats()
        vh = None if via_hidden is None else np.ascontiguousarray(via_hidden, np.uint8)
        check(lib().c2g_host_bellman_ford(self._h, _ptr(dist, C.c_int32), _ptr(frontier, C.c_int32),
                                          len(frontier), _ptr(vh, C.c_uint8), C.byref(st)))
        return dist, st

    # ---- PPR
    def ppr(self, src, alpha=0.15, eps=1e-7):
        p = np.empty(self.V, np.float32)
        r = np.empty(self.V, np.float32)
        st = Stats()
        check(lib().c2g_ppr(self._h, int(src), C.c_float(alpha), C.c_float(eps), _ptr(p, C.c_float),
                            _ptr(r, C.c_float), C.byref(st)))
        return p, r, st

    def ppr_migrate(self, src, budget, handoff_min=0, alpha=0.15, eps=1e-7):
        """GPU rounds up to `budget`, then the host-core push finishes (c2g_ppr_migrate)."""
        p = np.empty(self.V, np.float32)
        r = np.empty(self.V, np.float32)
        st = Stats()
        check(lib().c2g_ppr_migrate(self._h, int(src), C.c_float(alpha), C.c_float(eps), int(budget),
                                    int(handoff_min), _ptr(p, C.c_float), _ptr(r, C.c_float), C.byref(st)))
        return p, r, st

    def ppr_batch(self, srcs, alpha=0.15, eps=1e-7):
        srcs = np.ascontiguousarray(srcs, np.int32)
        p = np.empty((len(srcs), self.V), np.float32)
        r = np.empty((len(srcs), self.V), np.float32)
        st = Stats()
        check(lib().c2g_ppr_batch(self._h, _ptr(srcs, C.c_int32), len(srcs), C.c_float(alpha), C.c_float(eps),
                                  _ptr(p, C.c_float), _ptr(r, C.c_float), C.byref(st)))
        return p, r, st

    def host_ppr(self, src, alpha=0.15, eps=1e-7):
        p = np.empty(self.V, np.float32)
        r = np.empty(self.V, np.float32)
        st = Stats()
        check(lib().c2g_host_ppr(self._h, int(src), C.c_float(alpha), C.c_float(eps), _ptr(p, C.c_float),
                                 _ptr(r, C.c_float), C.byref(st)))
        return p, r, st

    # ---- adsorption
    def adsorption_push(self, src):
        """The reference's single-source adsorption push (c2g_adsorption_push)."""
        p = np.empty(self.V, np.float32)
        r = np.empty(self.V, np.float32)
        st = Stats()
        check(lib().c2g_adsorption_push(self._h, int(src), _ptr(p, C.c_float), _ptr(r, C.c_float), C.byref(st)))
        return p, r, st

    def adsorption_push_mode(self, src, mode):
        """c2g_adsorption_push_mode: 0 = w-base (w-purn on a PPR_PURN graph), 1 = w-base-perk."""
        p = np.empty(self.V, np.float32)
        r = np.empty(self.V, np.float32)
        st = Stats()
        check(lib().c2g_adsorption_push_mode(self._h, int(src), int(mode), _ptr(p, C.c_float), _ptr(r, C.c_float),
                                             C.byref(st)))
        return p, r, st

    def adsorption(self, seed_v, seed_l, L=16, p_inj=0.25, p_cont=0.70, p_abnd=0.05, iters=20):
        seed_v = np.ascontiguousarray(seed_v, np.int32)
        seed_l = np.ascontiguousarray(seed_l, np.int32)
        y = np.empty((self.V, L), np.float32)
        st = Stats()
        check(lib().c2g_adsorption(self._h, _ptr(seed_v, C.c_int32), _ptr(seed_l, C.c_int32), len(seed_v),
                                   L, C.c_float(p_inj), C.c_float(p_cont), C.c_float(p_abnd), iters,
                                   _ptr(y, C.c_float), C.byref(st)))
        return y, st

    def autotune(self, srcs, batched=False):
        srcs = np.ascontiguousarray(srcs, np.int32)
        t = Tune()
        check(lib().c2g_autotune(self._h, _ptr(srcs, C.c_int32), len(srcs), int(batched), C.byref(t)))
        return t

    def autotune_ppr(self, srcs, alpha=0.15, eps=1e-7):
        srcs = np.ascontiguousarray(srcs, np.int32)
        t = Tune()
        check(lib().c2g_autotune_ppr(self._h, _ptr(srcs, C.c_int32), len(srcs), C.c_float(alpha), C.c_float(eps),
                                     C.byref(t)))
        return t


class Server:
    """Continuous-batching query server over one graph (c2g_server_*)."""

    def __init__(self, graph, max_batch=64, max_inflight=0, hbm_reserve=0, fake_free_hbm=0, host_lane=False,
                 device_failed=False, alpha=0.15, eps=1e-7, ads_labels=16, ads_iters=20, p_inj=0.25, p_cont=0.70,
                 p_abnd=0.05, sssp_slots=0, ads_grid_per_cu=0, deadline_ms=0, fail_inject=0, fail_after=0,
                 dry_device=False, dry_step_ms=0):
        self.graph = graph
        self.L = ads_labels
        o = ServerOpts()
        o.max_batch, o.max_inflight, o.hbm_reserve, o.fake_free_hbm = max_batch, max_inflight, hbm_reserve, fake_free_hbm
        o.host_lane, o.device_failed, o.alpha, o.eps = int(host_lane), int(device_failed), alpha, eps
        o.ads_labels, o.ads_iters, o.p_inj, o.p_cont, o.p_abnd = ads_labels, ads_iters, p_inj, p_cont, p_abnd
        o.sssp_slots = sssp_slots
        o.ads_grid_per_cu = ads_grid_per_cu
        o.deadline_ms, o.fail_inject, o.fail_after = deadline_ms, fail_inject, fail_after
        o.dry_device, o.dry_step_ms = int(dry_device), dry_step_ms
        self._s = C.c_void_p()
        check(lib().c2g_server_start(graph._h, C.byref(o), C.byref(self._s)))

    def submit(self, kind, src):
        return check(lib().c2g_server_submit(self._s, int(kind), int(src)))

    def poll(self, rid, wait_ms=-1, want_result=True):
        """Returns (info dict, result array) or None if not done within wait_ms."""
        info = ResultInfo()
        out = None
        V = self.graph.V
        if want_result:
            out = np.empty(V * self.L, np.float32)  # large enough for every kind
        rc = lib().c2g_server_poll(self._s, rid, int(wait_ms),
                                   out.ctypes.data_as(C.c_void_p) if out is not None else None, C.byref(info))
        check(rc)
        if rc == 0:
            return None
        d = {k: getattr(info, k) for k, _ in ResultInfo._fields_ if k != "pad"}
        res = None
        if out is not None:
            if info.kind == Q_SSSP:
                res = out[:V].view(np.int32).copy()
            elif info.kind == Q_PPR:
                res = out[:V].copy()
            else:
                res = out[: V * self.L].reshape(V, self.L).copy()
        return d, res

    def cancel(self, rid):
        """1 if the request was cancelled, 0 if it had already finished."""
        return check(lib().c2g_server_cancel(self._s, rid))

    def stats(self):
        st = ServerStats()
        check(lib().c2g_server_get_stats(self._s, C.byref(st)))
        return st.as_dict()

    def stop(self):
        if self._s:
            lib().c2g_server_stop(self._s)
            self._s = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.stop()
