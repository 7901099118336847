"""The C-ABI library loads, exports every function include/c2g.h declares, and
reports errors as codes + messages (never exit())."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import c2g

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = []
    for hdr in os.listdir(os.path.join(ROOT, "include")):
        if hdr.endswith(".h"):
            src = open(os.path.join(ROOT, "include", hdr)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names += re.findall(r"\b(c2g_\w+)\s*\(", src)
    return sorted(set(names))


def test_every_declared_symbol_is_exported():
    names = _declared()
    assert len(names) >= 30
    L = c2g.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_version_and_errors(tmp_path):
    L = c2g.lib()
    assert b"gfx950" in L.c2g_version()
    with pytest.raises(c2g.C2GError) as e:
        c2g.Graph.open(str(tmp_path / "nope"), no_device=True)
    assert e.value.code == c2g.EIO and "not exist" in str(e.value)


def test_dataset_roundtrip_and_validation(tmp_path):
    G = c2g.rmat(10, 8, seed=1, wmax=64, want_f=True)
    ds = str(tmp_path / "ds")
    c2g.write_dataset(ds, G["vlist"], G["elist"], G["wlist"], G["flist"])
    assert np.array_equal(c2g.read_i32(ds + "/origin/vlist.bin"), G["vlist"])
    assert np.array_equal(c2g.read_f32(ds + "/origin/flist.bin"), G["flist"])
    with c2g.Graph.open(ds, no_device=True, load_flist=True) as g:
        assert (g.V, g.E) == (G["V"], G["E"])
    # truncated elist -> format error
    el = c2g.read_i32(ds + "/origin/elist.bin")
    el[:-1].tofile(ds + "/origin/elist.bin")
    with pytest.raises(c2g.C2GError) as e:
        c2g.Graph.open(ds, no_device=True)
    assert e.value.code == c2g.EFORMAT
    # out-of-range destination
    el[3] = G["V"]
    el.tofile(ds + "/origin/elist.bin")
    with pytest.raises(c2g.C2GError) as e:
        c2g.Graph.open(ds, no_device=True)
    assert e.value.code == c2g.EFORMAT


def test_gpu_calls_fail_loudly_without_device():
    G = c2g.rmat(8, 4, seed=1, wmax=8)
    with c2g.Graph.from_csr(G["vlist"], G["elist"], G["wlist"], no_device=True) as g:
        with pytest.raises(c2g.C2GError) as e:
            g.sssp(0)
        assert e.value.code == c2g.EDEVICE
        with pytest.raises(c2g.C2GError) as e:
            g.sssp_batch(np.array([0, 1]))
        assert e.value.code == c2g.EDEVICE


def test_tune_validation():
    G = c2g.rmat(8, 4, seed=1, wmax=8)
    with c2g.Graph.from_csr(G["vlist"], G["elist"], G["wlist"], no_device=True) as g:
        g.set_tune(chunk=32, iters_per_graph=4)
        assert g.tune().chunk == 32
        with pytest.raises(c2g.C2GError):
            g.set_tune(iters_per_graph=3)
        with pytest.raises(c2g.C2GError):
            g.set_tune(chunk=0)


@pytest.mark.parametrize("wmax,want", [(1023, 32), (255, 8), (64, 2), (1, 0)])
def test_default_light_bound(wmax, want):
    # the light-edge bound starts at about 1/16 of the mean weight (mean of
    # U[1,wmax] is (wmax+1)/2), and off when under 1 % of the edges would be
    # light (unit weights: none is)
    G = c2g.rmat(12, 16, seed=3, wmax=wmax)
    with c2g.Graph.from_csr(G["vlist"], G["elist"], G["wlist"], no_device=True) as g:
        assert g.tune().light == want and g.tune().batch_order == 1 and g.tune().ppr_grid_per_cu == 1
        assert g.tune().batch_relabel == 1
        g.set_tune(light=5, batch_relabel=0)
        assert g.tune().light == 5 and g.tune().batch_relabel == 0
        with pytest.raises(c2g.C2GError):
            g.set_tune(light=-1)
        with pytest.raises(c2g.C2GError):
            g.set_tune(batch_relabel=2)
