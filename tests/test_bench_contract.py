"""bench.py's contract where it needs no GPU: the one-GPU refusal, the
committed profiles its roofline `traffic` is read from, and the staleness
rule that keeps an old profile out of a new line."""
from __future__ import annotations

import glob
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TRAFFIC = sorted(glob.glob(os.path.join(ROOT, "profiles", "*_traffic.json")))


def test_bench_refuses_more_than_one_gpu():
    env = {k: v for k, v in os.environ.items() if k not in ("RANK", "LOCAL_RANK", "WORLD_SIZE")}
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "8"], capture_output=True,
                       text=True, env=env, timeout=120)
    assert r.returncode == 2
    assert "measures one GPU" in r.stderr
    assert r.stdout.strip() == ""          # no JSON line for a configuration it did not measure


@pytest.mark.parametrize("path", TRAFFIC, ids=[os.path.basename(p) for p in TRAFFIC])
def test_committed_traffic_profiles_are_whole(path):
    """Each traffic file sums its kernels' calibrated bytes, has its summary
    beside it, and was made by one kernel-trace run and two single-counter
    runs of bench.py (no counter run combined with a trace domain)."""
    with open(path) as f:
        prof = json.load(f)
    assert os.path.exists(path.replace("_traffic.json", ".md"))
    total = sum(k["read_bytes"] + k["write_bytes"] for k in prof["kernels"].values())
    assert abs(total - prof["traffic_bytes"]) <= 1e-9 * total
    assert prof["calibration"]["fetch"] == pytest.approx(2.0, rel=1e-3)
    assert prof["calibration"]["write"] == pytest.approx(1.0, rel=1e-3)
    kt, fetch, write = prof["commands"]
    for c in (kt, fetch, write):
        assert c.startswith("rocprofv3 ") and " -- python3 bench.py " in c
    assert "--kernel-trace" in kt and "--pmc" not in kt
    for c, counter in ((fetch, "FETCH_SIZE"), (write, "WRITE_SIZE")):
        assert f"--pmc {counter} " in c
        assert "trace" not in c.split(" -- ")[0]


def test_traffic_is_taken_only_from_a_profile_that_matches_the_run():
    sys.path.insert(0, ROOT)
    import bench

    tag = "assembly_p2_nh_cube64"
    with open(os.path.join(ROOT, "profiles", f"{tag}_traffic.json")) as f:
        prof = json.load(f)
    ke = prof["kernels"]["element_kernel"]["avg_ms"]
    kg = prof["kernels"]["gather_kernel"]["avg_ms"]
    traffic, src = bench.load_traffic(tag, 1.05 * ke, 0.95 * kg)
    assert traffic == prof["traffic_bytes"] and src == f"profiles/{tag}_traffic.json"
    traffic, src = bench.load_traffic(tag, 1.3 * ke, kg)
    assert traffic is None and src.startswith("stale profile")
    assert bench.load_traffic("no_such_workload", ke, kg) == (None, "no profile")
    assert bench.profile_tag(64, 2) == tag
    assert bench.profile_tag(64, 2, 32) == tag + "_f32tangent"
    assert bench.profile_tag(64, 1, 64, "bar") == "assembly_p1_j2_bar384"
