"""Command-line tools keep the reference's contracts (positional dataset_path,
source, migration / iter; dataset directory layout)."""
import os
import subprocess

import numpy as np
import pytest

import c2g
import oracle
from conftest import INT_MAX, ROOT

BIN = os.path.join(ROOT, "c-2graph_amd", "bin")


def run(*args, check=True):
    p = subprocess.run([os.path.join(BIN, args[0])] + [str(a) for a in args[1:]], capture_output=True, text=True,
                       timeout=300)
    if check and p.returncode != 0:
        raise AssertionError(f"{args}: rc={p.returncode}\n{p.stdout}\n{p.stderr}")
    return p


@pytest.fixture
def ds(tmp_path):
    d = str(tmp_path / "ds")
    run("rmat-gen", d, 12, "--seed", 3, "--wmax", 64)
    return d


def test_rmat_gen_and_prune_tools(ds):
    vl = c2g.read_i32(ds + "/origin/vlist.bin")
    G = c2g.rmat(12, 16, seed=3, wmax=64, want_f=True)
    assert np.array_equal(vl, G["vlist"])
    assert np.array_equal(c2g.read_f32(ds + "/origin/flist.bin"), G["flist"])
    out = run("purn-sssp", ds).stdout
    assert "the ratio is" in out
    run("purn-ppr", ds)
    P = c2g.prune_sssp_mem(G["vlist"], G["elist"], G["wlist"])
    assert np.array_equal(c2g.read_i32(ds + "/sssp/elist.bin"), P["elist"])
    assert np.array_equal(c2g.read_i32(ds + "/sssp/hvlist.bin"), P["hvlist"])
    assert np.array_equal(c2g.read_i32(ds + "/sssp/porder.bin"), P["porder"])


def test_convert2binary_tool(tmp_path):
    txt = tmp_path / "g.txt"
    txt.write_text("0 1:2:2:0\n2 0\nbad\n5 3:1\n")
    d = str(tmp_path / "c")
    run("convert2binary", str(txt), d)
    assert c2g.read_i32(d + "/origin/vlist.bin").tolist() == [0, 2, 2, 3, 3, 3, 5]
    assert c2g.read_i32(d + "/origin/elist.bin").tolist() == [1, 2, 0, 1, 3]


@pytest.mark.gpu
def test_sssp_base_prints_rounds(gpu, ds):
    # --rounds: the per-iteration frontier sizes (iteration_act_num)
    out = run("sssp-base", ds, 0, "--reps", 1, "--rounds").stdout
    lines = out[out.index("iteration\tfrontier"):].strip().splitlines()[1:]
    assert len(lines) >= 2 and lines[0].split("\t")[1] == "1"


def test_usage_errors():
    assert run("sssp-base", check=False).returncode == 2
    assert run("purn-sssp", check=False).returncode == 2
    p = run("sssp-base", "/nonexistent", 0, check=False)
    assert p.returncode == 1 and "not exist" in p.stderr


@pytest.mark.gpu
def test_sssp_tools(gpu, ds, tmp_path):
    G = c2g.rmat(12, 16, seed=3, wmax=64)
    deg = np.diff(G["vlist"])
    run("purn-sssp", ds)
    for tool in ("sssp-base", "sssp-purn"):
        out = str(tmp_path / f"{tool}.bin")
        p = run(tool, ds, 0, "--reps", 2, "--out", out)
        assert "gpu runtime:" in p.stdout and "源顶点source = " in p.stdout
        s = int(p.stdout.split("源顶点source = ")[1].split()[0])
        # skip rule of sssp/base/Main.cu:61-64, on the graph the tool loads
        dg = deg if tool == "sssp-base" else np.diff(c2g.read_i32(ds + "/sssp/vlist.bin"))
        assert dg[s] >= 1 and np.all(dg[:s] < 1)
        d = c2g.read_i32(out)
        assert np.array_equal(d, oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], s))
        first = [int(l.split("\t")[1]) for l in p.stdout.splitlines() if l.count("\t") == 1 and l.split("\t")[0].isdigit()]
        assert first[:10] == d[:10].tolist()
    out = str(tmp_path / "m.bin")
    p = run("sssp-base-M", ds, 0, 2, "--reps", 1, "--out", out)
    s = int(p.stdout.split("源顶点source = ")[1].split()[0])
    indeg = np.bincount(G["elist"], minlength=G["V"])
    assert deg[s] >= 10 and indeg[s] >= 10
    assert np.array_equal(c2g.read_i32(out), oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], s))
    # sssp-purn-M: pruned graph, GPU budget of 2 rounds then the host lane;
    # source rule of sssp/purn-perk/Main.cu (10 <= |hidden list| <= 1000)
    hv = np.diff(c2g.read_i32(ds + "/sssp/hvlist.bin"))
    assert np.any((hv >= 10) & (hv <= 1000))
    p = run("sssp-purn-M", ds, 0, 2, "--reps", 1, "--out", out)
    s = int(p.stdout.split("源顶点source = ")[1].split()[0])
    assert 10 <= hv[s] <= 1000 and np.all((hv[:s] < 10) | (hv[:s] > 1000))
    assert np.array_equal(c2g.read_i32(out), oracle.dijkstra(G["vlist"], G["elist"], G["wlist"], s))


@pytest.mark.gpu
def test_ppr_tools(gpu, ds, tmp_path):
    G = c2g.rmat(12, 16, seed=3, wmax=64, want_f=True)
    run("purn-ppr", ds)
    out = str(tmp_path / "p.bin")
    p = run("ppr-base", ds, 7, "--reps", 1, "--out", out)
    assert p.stdout.count("pagerank") == 11
    po, _, _, _ = oracle.ppr_push(G["vlist"], G["elist"], G["flist"], 7, alpha=0.15, eps=1e-7)
    assert np.abs(c2g.read_f32(out).astype(np.float64) - po).sum() <= 2e-5
    p = run("ppr-base-M", ds, 7, 2, "--reps", 1, "--handoff", 0, "--out", out)
    assert np.abs(c2g.read_f32(out).astype(np.float64) - po).sum() <= 2e-5
    run("ppr-purn", ds, 7, "--reps", 1)
    run("ppr-purn-M", ds, 7, 2, "--reps", 1)
    p = run("adsorption", ds, "--labels", 4, "--iters", 3)
    assert "gpu runtime" in p.stdout
    # the reference's adsorption main: dataset, source; first 11 p/r lines
    txt = str(tmp_path / "push.txt")
    p = run("adsorption-push", ds, 7, "--out", txt)
    assert p.stdout.count("\tpagerank\t") == 11 and "gpu runtime:" in p.stdout
    rows = np.loadtxt(txt)
    po, ro, _, _ = oracle.adsorption_push(G["vlist"], G["elist"], G["flist"], 7)
    assert np.array_equal(rows[:, 0], np.arange(G["V"]))
    assert np.abs(rows[:, 1] - po).sum() <= 2e-3  # printed with 6 significant digits


def _adj_texts(tmp_path):
    """Adjacency texts in the reference's input grammar ("src d1:d2:..."),
    with duplicates, self-loops, lines that do not parse, out-of-order rows,
    a vertex that appears only as a target, and one large random graph."""
    rng = np.random.default_rng(7)
    texts = {"tiny": "0 1:2:2:0\n2 0\nbad\n5 3:1\n7\n3 3:3:4\n"}
    lines = []
    for u in rng.permutation(300):
        k = int(rng.integers(1, 12))
        d = rng.integers(0, 320, k)  # ids up to 319: some vertices only appear as targets
        if rng.random() < 0.2:
            d = np.append(d, u)  # self-loop
        if rng.random() < 0.3:
            d = np.append(d, d[:2])  # duplicates
        lines.append(f"{u} " + ":".join(map(str, d)))
        if rng.random() < 0.05:
            lines.append("# comment line")
    texts["mixed"] = "\n".join(lines) + "\n"
    lines = [f"{u} " + ":".join(map(str, rng.integers(0, 5000, int(rng.integers(1, 40)))))
             for u in range(5000) if rng.random() < 0.9]
    texts["large"] = "\n".join(lines) + "\n"
    texts["dup_src"] = "4 1:2\n4 3:1\n0 4\n"  # a source on two lines: both lines' targets kept
    out = {}
    for name, t in texts.items():
        p = tmp_path / f"{name}.txt"
        p.write_text(t)
        out[name] = str(p)
    return out


def test_convert2binary_byte_identical_to_reference(tmp_path):
    # utils/convert2binary.cpp:10-135, compiled from the reference's source
    # into oracle/_ref: all four origin files must match byte for byte
    ref = os.path.join(ROOT, "oracle", "_ref", "convert2binary")
    if not os.path.exists(ref):
        pytest.skip("oracle/_ref/convert2binary not built (reference tree absent)")
    for name, txt in _adj_texts(tmp_path).items():
        a, b = str(tmp_path / f"{name}_ref"), str(tmp_path / f"{name}_ours")
        os.makedirs(a + "/origin")  # the reference writes into an existing origin/
        subprocess.run([ref, txt, a], check=True, capture_output=True, timeout=120)
        run("convert2binary", txt, b)
        for f in ("vlist", "elist", "wlist", "flist"):
            ra = open(f"{a}/origin/{f}.bin", "rb").read()
            rb = open(f"{b}/origin/{f}.bin", "rb").read()
            assert ra == rb, (name, f, len(ra), len(rb))


def test_convert2binary_weights_follow_libc_rand(tmp_path):
    # the weights are (rand() % 100) + 1 in CSR order from the C library's
    # unseeded generator (convert2binary.cpp:108): compare with libc itself
    import ctypes
    libc = ctypes.CDLL("libc.so.6")
    txt = _adj_texts(tmp_path)["large"]
    d = str(tmp_path / "w")
    run("convert2binary", txt, d)
    w = c2g.read_i32(d + "/origin/wlist.bin")
    libc.srand(1)  # rand() before any srand() behaves as srand(1)
    expect = np.array([libc.rand() % 100 + 1 for _ in range(len(w))], np.int32)
    assert len(w) > 50000 and np.array_equal(w, expect)
    vl = c2g.read_i32(d + "/origin/vlist.bin")
    f = c2g.read_f32(d + "/origin/flist.bin")
    row = np.repeat(np.arange(len(vl) - 1), np.diff(vl))
    sums = np.bincount(row, weights=w, minlength=len(vl) - 1)
    assert np.array_equal(f, (w / sums[row]).astype(np.float32))  # 1.0*w/sum in double, stored as float
