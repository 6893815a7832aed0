"""Writes tests/golden/ref_host.json: golden vectors produced by the
reference's own host code (common/matrix_converters.h and
common/measurement_class.{h,cpp}), compiled from the reference sources by
oracle/Makefile's ``ref`` target into oracle/_ref/ and run here.

Run in a container that has the reference sources mounted:

    python tests/golden/make_golden.py [--ref /root/reference]

The fixture is data only (inputs and outputs); no reference source is copied.
"""
import argparse
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    args = ap.parse_args()
    if not os.path.isdir(args.ref):
        sys.exit(f"reference sources not found at {args.ref}")
    oracle = os.path.join(REPO, "oracle")
    subprocess.run(["make", "-C", oracle, "ref", f"REF={args.ref}"], check=True)
    out = subprocess.run([os.path.join(oracle, "_ref", "ref_harness")],
                         check=True, capture_output=True, text=True).stdout
    data = json.loads(out)
    data["_source"] = ("oracle/ref_harness.cpp driving matrix_converters.h:13-196 and "
                       "measurement_class.{h,cpp} of the reference")
    path = os.path.join(HERE, "ref_host.json")
    with open(path, "w") as f:
        json.dump(data, f, separators=(",", ":"))
        f.write("\n")
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
