"""Extracts the Stata-generated input vectors that the reference's own tests embed
(tests/testthat/test-ci.R:89 u1 seed 12345678, :227 u seed 12345678, :301 u (bigger-test),
:304 u1 seed 1) into tests/golden/test_ci_vectors.json.  These are data fixtures, not code.
Run in a container that has the reference mounted:  python tests/golden/make_test_ci_vectors.py
"""
import json
import os
import re

SRC = "/root/reference/tests/testthat/test-ci.R"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "test_ci_vectors.json")


def main():
    lines = open(SRC).read().splitlines()
    picks = {"u1_seed12345678": 89, "u_seed12345678": 227, "u_bigger": 301, "u1_seed1": 304}
    out = {}
    for name, lineno in picks.items():
        m = re.search(r"c\((.*)\)", lines[lineno - 1])
        out[name] = [float(v) for v in m.group(1).split(",")]
    with open(OUT, "w") as f:
        json.dump(out, f)
    print({k: len(v) for k, v in out.items()})


if __name__ == "__main__":
    main()
