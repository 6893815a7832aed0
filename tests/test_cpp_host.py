"""The C++ drop-in host layer (include/bmfp/*.h) against the reference's own
host code: host_selftest prints the same golden layout that oracle/_ref
produced from the reference sources (tests/golden/ref_host.json); every field
must match exactly.  CPU only."""
import json
import os
import subprocess

import numpy as np
import pytest

import block_matrix_format_performance_amd as bm
import oracle

SELFTEST = os.path.join(bm.PKG_DIR, "host_selftest")
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_host.json")


@pytest.fixture(scope="module")
def selftest():
    if not os.path.exists(SELFTEST):
        pytest.fail("host_selftest not built (run __graft_entry__.build())")
    out = subprocess.run([SELFTEST], check=True, capture_output=True, text=True).stdout
    return json.loads(out)


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def test_converters_match_reference_exactly(selftest, golden):
    assert len(selftest["matrices"]) == len(golden["matrices"])
    for mine, ref in zip(selftest["matrices"], golden["matrices"]):
        for key in ("n_rows", "blocks_per_row", "bs", "nnzb", "bcsr_row_ptr", "bcsr_columns", "csr_row_ptr",
                    "csr_columns", "bcsr_values", "bcsr_values_transposed", "csr_values", "block_lookup_mid"):
            assert mine[key] == ref[key], key


def test_write_mm_text_matches_reference_exactly(selftest, golden):
    # the reference's write_mm (through common/mmio.c) vs matrix_converters.h's
    # write_mm and bmfp_mm_write, byte for byte; reading it back is exact
    for mine, ref in zip(selftest["matrices"], golden["matrices"]):
        assert mine["mm_text"] == ref["mm_text"]
        assert mine["mm_text_abi"] == ref["mm_text"]
        assert mine["mm_readback_exact"] is True


def test_csr_to_bcsr_round_trip(selftest):
    for m in selftest["matrices"]:
        assert m["roundtrip_row_ptr"] == m["bcsr_row_ptr"]
        assert m["roundtrip_columns"] == m["bcsr_columns"]
        assert m["roundtrip_values"] == m["bcsr_values"]


def test_measurement_class_matches_reference(selftest, golden):
    assert selftest["measurement"] == golden["measurement"]


def test_compare_results_text_matches_reference(selftest, golden):
    assert selftest["compare"] == golden["compare"]


def test_stencil_template_matches_oracle(selftest):
    s = selftest["stencil"]
    m = oracle.gen_stencil9(s["nx"], s["ny"], s["bs"], s["kind"], s["seed"])
    np.testing.assert_array_equal(s["row_ptr"], m.row_ptr)
    np.testing.assert_array_equal(s["columns"], m.columns)
    assert np.array_equal(np.array(s["values"]), m.values)


def test_driver_and_cpp_library_built():
    assert os.path.exists(bm.MAIN_PATH)
    assert os.path.exists(bm.CPP_LIB_PATH)
