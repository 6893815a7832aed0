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
    # the bs = 1 sweep matrix: 6,256,665,396 algorithmic bytes; the measured
    # traffic is at least that and re-reads part of x
    assert t is not None and 6256665396 <= t < 1.2 * 6256665396
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


def test_serving_byte_model_by_slot():
    # step 0 alone moves no matrix: x' written, b read, r32 written (20 B per
    # row and right-hand side); every further refinement step adds one FP64
    # matrix read (values + block columns + row pointers) and 40 B of vectors;
    # every inner iteration two inner-matrix reads (FP16 values + one FP32
    # scale per block row at bs >= 6) and 76 B of vectors
    r = {"side": 20, "bs": 6, "nrhs": 3}
    nodes, nnzb, n = 400, 58 * 58, 400 * 6
    idx = 4 * nnzb + 4 * (nodes + 1)
    b0 = bench.request_bytes_by_slot(r, 1, 0)
    assert sum(b0.values()) == 20 * n * 3 and b0["R"] == 12 * n * 3
    d_step = bench.request_bytes_moved(r, 3, 0) - bench.request_bytes_moved(r, 2, 0)
    assert d_step == 8 * nnzb * 36 + idx + 40 * n * 3
    d_it = bench.request_bytes_moved(r, 1, 5) - bench.request_bytes_moved(r, 1, 4)
    assert d_it == 2 * (2 * nnzb * 36 + 4 * nodes + idx) + 76 * n * 3
    fp32 = bench.request_bytes_moved(r, 1, 1, f16_min_bs=0)
    assert fp32 - bench.request_bytes_moved(r, 1, 1) == 2 * (2 * nnzb * 36 - 4 * nodes)
    jac = bench.request_bytes_moved(r, 2, 7, jacobi=True) - bench.request_bytes_moved(r, 2, 7)
    assert jac == 7 * 12 * n
