"""bench.py's traffic provenance on the CPU: the committed counter summary is
used only for the kernel source it was measured on, and its bytes per
traversal are spread over the run's own launches."""
import hashlib
import importlib.util
import json
import os
import types

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


C4 = types.SimpleNamespace(scale=24, seed=4, queries=256, wmax=1023)


def test_committed_summary_is_used_only_for_its_kernel_source(bench):
    with open(bench.TRAFFIC_JSON) as f:
        t = json.load(f)
    assert t["traffic_bytes_per_launch"] == t["read_bytes_per_launch"] + t["write_bytes_per_launch"]
    with open(bench.KERNEL_SRC, "rb") as f:
        same = t["kernel_source_sha256"] == hashlib.sha256(f.read()).hexdigest()
    traffic, prov = bench.measured_traffic(t["kernel"], C4, {"light": t["light"]})
    if same:
        assert traffic == t["traffic_bytes_per_launch"] and "note" not in prov
    else:
        assert traffic is None and prov["note"].startswith("stale")


def test_stale_or_foreign_summaries_give_no_traffic(bench, tmp_path, monkeypatch):
    with open(bench.TRAFFIC_JSON) as f:
        t = json.load(f)
    # another light bound, another kernel, another workload
    assert bench.measured_traffic(t["kernel"], C4, {"light": t["light"] + 1})[0] is None
    assert bench.measured_traffic("relax_b_k<false, 512>", C4, {"light": t["light"]})[0] is None
    other = types.SimpleNamespace(scale=20, seed=4, queries=256, wmax=1023)
    assert bench.measured_traffic(t["kernel"], other, {"light": t["light"]})[0] is None
    # a kernel source that changed since the counter run
    src = tmp_path / "sssp_batch.hip"
    src.write_bytes(open(bench.KERNEL_SRC, "rb").read() + b"\n")
    monkeypatch.setattr(bench, "KERNEL_SRC", str(src))
    traffic, prov = bench.measured_traffic(t["kernel"], C4, {"light": t["light"]})
    assert traffic is None and prov["note"].startswith("stale")


def test_traffic_is_spread_over_the_runs_launches(bench):
    prov = {"dispatches": 28}
    traffic, p2 = bench.traffic_for_run(14_000, prov, 26)
    assert p2["traffic_per_traversal"] == 14_000 * 28 and traffic == 14_000 * 28 // 26
    assert "traffic_per_traversal" not in prov  # the caller's dict is left alone
    assert bench.traffic_for_run(None, {"note": "x"}, 26) == (None, {"note": "x"})
