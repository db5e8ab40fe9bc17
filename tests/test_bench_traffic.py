"""bench.py's roofline.traffic: the measured bytes per launch it reports are
looked up in the committed profile for the launch shape and block order the
run used, and are absent (None) for a shape the profile did not measure."""
from __future__ import annotations

import importlib.util
import json
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_traffic_follows_the_launch_shape():
    bench = _bench()
    wl = "bidisperse-sheared-4M-phi0.60"
    with open(os.path.join(ROOT, "profiles", f"traffic_{wl}.json")) as f:
        prof = json.load(f)
    seen = set()
    for e in prof["launches"]:
        key = (e["shape"], e["block_map"])
        assert key not in seen
        seen.add(key)
        for k in ("pass1", "pass2"):
            total = e[f"{k}_read_bytes"] + e[f"{k}_write_bytes"]
            assert abs(e[f"{k}_bytes_per_launch"] - total) <= 1.0
            assert bench.traffic_from_profiles(wl, k, *key) == e[f"{k}_bytes_per_launch"]
    assert bench.traffic_from_profiles(wl, "pass1", 2, 0) is None
    assert bench.traffic_from_profiles("no-such-workload", "pass1", 0, 0) is None
