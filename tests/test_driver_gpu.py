"""The drop-in benchmark driver (bmfp_main, the reference's main.cpp flow) on
the GPU: every measurement row of the reference is present in result.json,
including the run-time-specialised "jit" row, and no GPU result disagrees
with the CPU CSR reference (compare_results prints "ERROR" on a mismatch)."""
import json
import os
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import block_matrix_format_performance_amd as bm
import oracle
from block_matrix_format_performance_amd import report

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("typ,bs", [("float", "2,5"), ("double", "4,7")])
def test_driver_rows_and_agreement(tmp_path, typ, bs):
    print(f"\n[gpu-case] driver --type {typ} --bs {bs}", flush=True)
    out, gbs = tmp_path / "result.json", tmp_path / "gbs.json"
    p = subprocess.run([bm.MAIN_PATH, "--type", typ, "--bs", bs, "--n-rows", "3000", "--bpr", "5", "--reps", "2",
                        "--x", "random", "--out", str(out), "--out-gbs", str(gbs)],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    assert "ERROR" not in p.stdout + p.stderr, p.stdout + p.stderr
    res = json.loads(out.read_text())
    assert sorted(res) == sorted(bs.split(","))
    for b, rows in res.items():
        for name in ["CPU CSR", "GPU CSR", "GPU CSR-Vector", "GPU Merge SpMV", "jit"] + bm_rows():
            assert name in rows, (b, name)
        assert rows["jit"] > 0
    # the report: speed-up over CPU CSR and the fraction of the HBM roofline
    bw = json.loads(gbs.read_text())
    assert sorted(bw) == sorted(res)
    for b in bw:
        assert 0 < bw[b]["GPU Merge SpMV"] < 8000 and 0 < bw[b]["GPU CSR"] < 8000
    sp = report.speedup(report.load(out))
    # rows a family cannot run (null in result.json) come out as NaN
    assert all(v > 0 for rows in sp.values() for v in rows.values() if v == v)
    assert all(sp[int(b)]["GPU Merge SpMV"] > 0 and sp[int(b)]["jit"] > 0 for b in res)
    assert report.main([str(out), "--gbs", str(gbs)]) == 0


def bm_rows():
    return ["GPU BCSR (row major, thread per row)", "GPU BCSR (column major, thread per row, coal x)",
            "GPU BSR (rocSPARSE, row major)"]


def test_driver_matrix_market_input(tmp_path):
    print("\n[gpu-case] driver --mm (4x4-block stencil file)", flush=True)
    m = oracle.gen_stencil9(30, 20, 4, 0, 5)
    c = oracle.bcsr_to_csr(m)
    path = tmp_path / "s4.mtx"
    bm.write_mm(path, c.row_ptr, c.columns, c.values, c.n_rows, c.n_cols)
    out = tmp_path / "result.json"
    p = subprocess.run([bm.MAIN_PATH, "--mm", str(path), "--type", "double", "--reps", "2", "--x", "random",
                        "--out", str(out)], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    assert "detected block size 4" in p.stdout
    assert "ERROR" not in p.stdout + p.stderr, p.stdout + p.stderr
    res = json.loads(out.read_text())
    assert list(res) == ["4"]
    for name in ["CPU CSR", "GPU CSR", "GPU CSR-Vector", "GPU Merge SpMV", "jit"] + bm_rows():
        assert name in res["4"], name


def vtk_points(path):
    lines = path.read_text().splitlines()
    i = next(k for k, l in enumerate(lines) if l.startswith("POINTS"))
    n = int(lines[i].split()[1])
    return np.array([[float(v) for v in l.split()[:2]] for l in lines[i + 1:i + 1 + n]])


def test_driver_bridge_solve(tmp_path):
    # the suspension-bridge truss solved by the Jacobi-preconditioned FP64
    # BiCGStab through the drop-in solver class; checked against the matrix
    # the driver wrote (matrix.mtx, exact %.20g) and a sparse direct solve
    print("\n[gpu-case] driver --bridge-solve (1280 m main span)", flush=True)
    p = subprocess.run([bm.MAIN_PATH, "--bridge-solve"], capture_output=True, text=True, timeout=300, cwd=tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    rp, cols, vals, n, m = bm.read_mm(tmp_path / "matrix.mtx")
    K = sp.csr_matrix((vals, cols, rp), shape=(n, m))
    f = np.loadtxt(tmp_path / "bridge_rhs.txt")
    u = np.loadtxt(tmp_path / "bridge_solution.txt")
    res = np.linalg.norm(K @ u - f) / np.linalg.norm(f)
    ref = spla.spsolve(K.tocsc(), f)
    err = np.linalg.norm(u - ref) / np.linalg.norm(ref)
    print(f"bridge: true relative residual {res:.2e}, error vs sparse direct solve {err:.2e}")
    assert res < 3e-8
    assert err < 1e-6
    assert u[1::2].min() < 0                                   # the deck sags under the load
    moved = vtk_points(tmp_path / "output_2.vtk") - vtk_points(tmp_path / "output_1.vtk")
    assert np.allclose(moved.ravel(), u, rtol=1e-9, atol=1e-12)


def test_driver_bridge_spmv_rows(tmp_path):
    print("\n[gpu-case] driver --bridge (SpMV rows)", flush=True)
    out = tmp_path / "result.json"
    p = subprocess.run([bm.MAIN_PATH, "--bridge", "--type", "double", "--reps", "2", "--x", "random", "--out",
                        str(out)], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    assert "ERROR" not in p.stdout + p.stderr, p.stdout + p.stderr
    res = json.loads(out.read_text())
    assert list(res) == ["2"] and "GPU Merge SpMV" in res["2"] and "GPU BCSR (row major, thread per row)" in res["2"]


def test_driver_solve():
    # the 256x256 bs=2 system of the drop-in driver (b = A*1, eps 1e-8):
    # the printed max|x-1| is parsed and bounded
    print("\n[gpu-case] driver --solve", flush=True)
    p = subprocess.run([bm.MAIN_PATH, "--bs", "2", "--n-rows", "500", "--reps", "1", "--out", os.devnull, "--solve"],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    line = next(l for l in p.stdout.splitlines() if l.startswith("BiCGStab 256x256 bs=2"))
    err = float(line.split("max|x-1| ")[1].split(",")[0])
    true_res = float(line.split(", true ")[1].split(",")[0])
    print(line)
    assert true_res < 2e-8
    assert err < 2e-5   # the oracle's solve at eps 1e-8 has max|x-1| = 8.2e-6


@pytest.mark.parametrize("typ,index", [("double", "int"), ("double", "unsigned"), ("float", "int"),
                                       ("float", "unsigned")])
def test_driver_solve_through_reference_csr_signature(tmp_path, typ, index):
    # gpu_bicgstab<T, C>(const csr_matrix_class<T, C>&, bool) + solve(csr, ...)
    # (bicgstab.h:17-20), as the reference's caller uses it (main.cpp:372-373),
    # for every instantiation of bicgstab.cu:528-531; x against the oracle's
    # BiCGStab on the same matrix (FP32: the matrix and b rounded to float)
    side, bs = 64, 2
    eps = 1e-10 if typ == "double" else 1e-5
    print(f"\n[gpu-case] driver --solve-csr {typ} {index}", flush=True)
    xfile = tmp_path / "x.bin"
    p = subprocess.run([bm.MAIN_PATH, "--solve-csr", "--type", typ, "--index", index, "--solve-side", str(side),
                        "--solve-bs", str(bs), "--solve-eps", str(eps), "--x-out", str(xfile)],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    assert f"BiCGStab CSR<{typ}, {index}>" in p.stdout
    x = np.fromfile(xfile, dtype=np.float64)
    m = oracle.gen_stencil9(side, side, bs, 0, 0)
    c = oracle.bcsr_to_csr(m)
    if typ == "float":
        c = oracle.CSR(c.n_rows, c.n_cols, c.row_ptr, c.columns, c.values.astype(np.float32).astype(np.float64))
    b = oracle.csr_spmv(c, np.ones(c.n_rows))
    if typ == "float":
        b = b.astype(np.float32).astype(np.float64)
    xo, oi = oracle.bicgstab(c, b, eps, 100000)
    assert oi.status == 0
    err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    res = np.linalg.norm(b - oracle.csr_spmv(c, x)) / np.linalg.norm(b)
    print(p.stdout.strip().splitlines()[-1], f"rel(x, x_oracle) {err:.2e}, FP64 residual {res:.2e}")
    if typ == "double":
        assert err < 1e-9 and res < 2e-10
    else:
        # FP32 vectors: the recurrence residual reaches 1e-5, the true one
        # stalls near the FP32 rounding level of A x
        assert err < 1e-4 and res < 1e-4
