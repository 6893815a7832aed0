"""bench.py host logic: the traffic lookup keyed by kernel symbol and the
one-GPU guard (no GPU needed)."""
import json
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)

bench = pytest.importorskip("bench")


def test_traffic_lookup_matches_symbol_and_sizing():
    sym = "_ZN4bmfp12k_bcsr_stageIddLi1ELi1ELb0ELi4EEEvNS_8BcsrArgsIT_T0_EE"
    t = bench.profiled_traffic(sym, "hbm")
    assert t is not None and 7.0e9 < t < 7.5e9
    assert bench.profiled_traffic(sym, "literal") is None        # no row for that sizing
    assert bench.profiled_traffic("_ZN4bmfp3nopEv", "hbm") is None


def test_every_traffic_row_parses():
    rows = 0
    with open(bench.TRAFFIC_PROFILE) as f:
        for line in f:
            if line.startswith("| `_Z"):
                cells = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
                assert bench.profiled_traffic(cells[0], cells[1]) == float(cells[6])
                rows += 1
    assert rows >= 20


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in json.loads(p.stdout.strip().splitlines()[-1])["error"]
