"""bench.py host-side logic (no GPU): the multi-GPU refusal, the kernel names
behind the timing stages, and the measured-traffic table it reports."""
import importlib.util
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_kernel_names():
    b = _bench()
    assert b.kernel_name("fp64", "density", 2) == "density_fused_f64"
    assert b.kernel_name("mixed", "density", 0) == "density_x2_i8"  # the default shape
    assert b.kernel_name("mixed", "density", 65) == "density_x2_i8"
    assert b.kernel_name("mixed", "density", 71) == "density_x2_i8"
    assert b.kernel_name("mixed", "vgemm", 65) == "vgemm_i8"


def test_traffic_table():
    """Every entry names a profile summary that exists, and the lookup keys
    density kernels on the shape and V kernels on the split."""
    b = _bench()
    with open(b.TRAFFIC_FILE) as f:
        entries = json.load(f)
    assert entries
    for e in entries:
        assert os.path.exists(os.path.join(ROOT, e["source"]))
        for k, v in e["kernels"].items():
            assert "<" not in k and v["bytes_per_launch"] > 0 and v["launches"] > 0
    e = entries[0]
    dens = [k for k in e["kernels"] if k.startswith("density")][0]
    t, src = b.measured_traffic(e["workload"], e["precision"], e["split"], e["shape"], dens)
    assert t == e["kernels"][dens]["bytes_per_launch"] and src == e["source"]
    t, _ = b.measured_traffic(e["workload"], e["precision"], e["split"], e["shape"] + 1, dens)
    assert t is None
    t, _ = b.measured_traffic("other", e["precision"], e["split"], e["shape"], dens)
    assert t is None
