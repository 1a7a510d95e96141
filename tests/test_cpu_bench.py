"""The bench contract pieces that need no GPU: the multi-GPU refusal, the
traffic summary the roofline reads, the kernel names it is keyed by, and the
CPU baseline's sampling."""
import json
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_template_names_match_the_profiled_instantiations():
    assert bench.template_name(0, 2, 5) == "k_reduce<float, float, true, 2>"
    assert bench.template_name(1, 1, 5) == "k_reduce<float, float, false, 1>"
    assert bench.template_name(2, 2, 5) == "k_scatter<float, float, 2, false>"
    assert bench.template_name(2, 4, 5) == "k_scatter<float, float, 4, true>"


def test_committed_traffic_covers_the_c2_passes():
    line = [l for l in open(os.path.join(ROOT, "profiles", "c2_hbm_traffic.md"))
            if l.startswith("TRAFFIC_JSON ")]
    assert len(line) == 1
    t = json.loads(line[0][len("TRAFFIC_JSON "):])
    for k in ("k_reduce<float, float, true, 4>", "k_reduce<float, float, true, 3>",
              "k_reduce<float, float, false, 1>", "k_scatter<float, float, 2, false>",
              "k_scatter<float, float, 1, false>"):
        assert k in t and t[k]["fetch_bytes"] > 0 and t[k]["write_bytes"] > 0
        assert bench.measured_traffic(2, k) == t[k]["fetch_bytes"] + t[k]["write_bytes"]
    # the memo levels written beside the leaf-to-root pass (56.4M + 19.0M rows
    # x 128 B) show up
    assert t["k_reduce<float, float, true, 4>"]["write_bytes"] > 9.0e9


def test_json_line_holds_no_nan():
    out = bench.finite({"fit": float("nan"), "errs": [1e-7, float("inf")], "cfg": {"n": 3}})
    assert out == {"fit": None, "errs": [1e-7, None], "cfg": {"n": 3}}
    json.dumps(out, allow_nan=False)


def test_cpu_baseline_is_measured_not_scaled():
    """The CPU oracle is timed on the whole workload (the tensor the GPU legs
    used) or, when asked, on a prefix of its draws -- never scaled -- and the
    single-thread run reports its own sample."""
    import types
    import adatm_amd as A
    t = A.synth_tensor(1, 40_000)
    whole = bench.cpu_baseline(types.SimpleNamespace(config=1, cpu_draws=0, cpu1_draws=10_000,
                                                     ndraws=40_000), 8, A, t)
    assert whole["unit"] == "ms/sweep" and whole["sample"].startswith(
        f"the whole workload ({t.nnz} nnz)")
    one = whole["single_thread"]
    assert one["cores"] == 1 and one["sample"].startswith("first 10000 draws")
    part = bench.cpu_baseline(types.SimpleNamespace(config=1, cpu_draws=10_000, cpu1_draws=0,
                                                    ndraws=40_000), 8, A)
    assert part["unit"] == "ms/sweep" and part["sample"].startswith("first 10000 draws")
    assert "single_thread" not in part
    assert whole["value"] > 0 and part["value"] > 0 and one["value"] > 0
