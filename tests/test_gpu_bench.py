"""bench.py's one JSON line, on a packed bed of 6 000 spheres (and the
crystal leg on 12 x 12 x 6 rock-salt cells): the keys of the benchmark line,
value and ms_per_step consistent with the sphere count, the roofline fraction
equal to achieved / peak, and the CPU baseline leg on a bounded sample.  A
request for several GPUs is refused with exit status 2."""
from __future__ import annotations

import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def _bench(*args, timeout=240):
    return subprocess.run([sys.executable, BENCH, *args], capture_output=True, text=True,
                          timeout=timeout)


def test_bench_line_on_a_small_bed():
    p = _bench("--n", "6000", "--settle-steps", "50", "--cells", "12,12,6", "--crystal-steps",
               "20", "--steps", "20", "--warmup", "5", "--tune-steps", "2", "--ens-steps", "0",
               "--serve-requests", "0", "--cpu-steps", "1")
    assert p.returncode == 0, p.stderr[-3000:]
    d = json.loads(p.stdout.strip().splitlines()[-1])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
              "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
              "roofline", "cpu_baseline"):
        assert k in d, k
    assert (d["n_gpus"], d["steps"], d["warmup"]) == (1, 20, 5)
    assert d["unit"] == "particle-steps/s" and d["higher_is_better"] is True
    assert d["dtype"] == "f64" and d["data"] == "synthetic"
    c = d["config"]
    n = 6000
    assert c["spheres"] == n and c["workload"] == "bidisperse-sheared-6000-phi0.60"
    assert 0.55 < c["packing_fraction_bulk"] < 0.66
    assert d["crystal"]["spheres"] == 8 * 12 * 12 * 6 and d["crystal"]["value"] > 0.0
    # value = spheres x steps / seconds of the timed steps
    assert d["value"] > 0.0
    assert abs(d["value"] * d["ms_per_step"] * 1e-3 - n) <= 1e-9 * n
    r = d["roofline"]
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert r["achieved"] > 0.0 and abs(r["frac"] - r["achieved"] / r["peak"]) <= 1e-12
    # no counter profile is committed for this bed
    assert r["traffic"] is None
    b = d["cpu_baseline"]
    assert b["value"] > 0.0 and b["cores"] >= 1 and b["kind"] in ("reference", "port")
    assert b["unit"] == "particle-steps/s" and b["sample"]
    m = d["mixed"]
    assert m["value"] > 0.0 and m["pair_contacts"] > 0


def test_bench_refuses_several_gpus():
    p = _bench("--gpus", "2", timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr
