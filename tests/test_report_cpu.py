"""The result report (block_matrix_format_performance_amd/report.py): speed-up
vs "CPU CSR" as the reference's plot/plot.py computes it, roofline fractions
from the driver's bandwidth file, text tables and PNG plots.  CPU only."""
import json
import math

import pytest

from block_matrix_format_performance_amd import report

RESULT = {"2": {"CPU CSR": 0.5, "GPU CSR": 0.01, "GPU BCSR (row major, thread per row)": 0.004},
          "16": {"CPU CSR": 2.0, "GPU CSR": 0.02, "GPU BCSR (row major, thread per row)": 0.008, "jit": 0.1}}
GBS = {"2": {"CPU CSR": 9.0, "GPU CSR": 3000.0, "GPU BCSR (row major, thread per row)": 4800.0},
       "16": {"CPU CSR": 11.0, "GPU CSR": 2500.0, "GPU BCSR (row major, thread per row)": 6000.0, "jit": 400.0}}


@pytest.fixture
def files(tmp_path):
    r, g = tmp_path / "result.json", tmp_path / "gbs.json"
    r.write_text(json.dumps(RESULT))
    g.write_text(json.dumps(GBS))
    return r, g


def test_speedup_is_cpu_time_over_format_time(files):
    sp = report.speedup(report.load(files[0]))
    assert list(sp) == [2, 16]                       # numeric block-size order
    assert sp[2]["GPU CSR"] == pytest.approx(50.0)
    assert sp[16]["GPU BCSR (row major, thread per row)"] == pytest.approx(250.0)
    assert sp[16]["CPU CSR"] == 1.0
    assert "jit" not in sp[2]


def test_missing_base_row_is_an_error(tmp_path):
    p = tmp_path / "r.json"
    p.write_text(json.dumps({"4": {"GPU CSR": 1.0}}))
    with pytest.raises(KeyError):
        report.speedup(report.load(p))


def test_roofline_fraction_and_tables(files):
    roof = report.roofline_fraction(report.load(files[1]), 8000.0)
    assert roof[16]["GPU BCSR (row major, thread per row)"] == pytest.approx(0.75)
    text = report.render(roof, "fraction", "{:.3f}")
    assert "0.750" in text and "jit" in text
    lines = report.render(report.speedup(report.load(files[0])), "speed-up").splitlines()
    assert lines[0].split()[-2:] == ["2", "16"]
    assert any(l.startswith("jit") and l.split()[-2] == "-" for l in lines)   # absent at bs 2


def test_cli_writes_plots(files, tmp_path, capsys):
    prefix = tmp_path / "rep"
    assert report.main([str(files[0]), "--gbs", str(files[1]), "--png-prefix", str(prefix)]) == 0
    out = capsys.readouterr().out
    assert "speed-up vs CPU CSR" in out and "fraction of 8000 GB/s" in out
    for suffix in ("_speedup.png", "_roofline.png"):
        data = (tmp_path / f"rep{suffix}").read_bytes()
        assert data[:8] == b"\x89PNG\r\n\x1a\n" and len(data) > 1000
