"""CPU checks of bench.py's contract: it refuses more than one GPU with a
message, and the committed counter summaries it folds into its JSON line match
the default workload (so `traffic` and `cache` are never silently null)."""
import importlib.util
import json
import os
import subprocess
import sys
import types

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert p.returncode != 0
    msg = json.loads(p.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_committed_counters_match_the_default_workload(monkeypatch):
    b = _bench()
    monkeypatch.setattr(sys, "argv", ["bench.py"])
    args = b.parse()
    ref_bp = args.ref_bp
    a = types.SimpleNamespace(ref_bp=ref_bp)
    nq, nl = args.reads + args.random_reads, args.locate_reads
    for kernel in ("k_mem_persist", "k_locate_search"):
        traffic = b.measured_traffic(kernel, a, nq, nl)
        cache = b.measured_cache(kernel, a, nq, nl)
        assert traffic and traffic > 0
        assert cache and 0.0 < cache["l2_hit_rate"] < 1.0 and cache["l2_read_requests_per_probe"] >= 1.0
    # a different workload gets no counters rather than another workload's
    assert b.measured_traffic("k_mem_persist", types.SimpleNamespace(ref_bp=ref_bp // 2), nq, nl) is None
    assert b.measured_cache("k_mem_persist", a, nq + 1, nl) is None


def test_profile_summaries_name_their_commands():
    prof = os.path.join(ROOT, "profiles")
    for name in os.listdir(prof):
        if name.endswith(".md"):
            text = open(os.path.join(prof, name)).read()
            assert "rocprofv3" in text or "bench.py" in text or "hipcc" in text, name
