"""Writes tests/golden/solve_jl.json from the reference's known-answer test.

The reference's only golden vectors for this path are the literals in
test/solve.jl:3-7 (p1, p2, x and the two-way OLS residuals r_ols) plus the
interaction vector at :58.  This script reads that file as text and stores
the numbers as data; it is run once in the build container, where the
reference is mounted.
"""
import json
import os
import re
import sys

SRC = "/root/reference/test/solve.jl"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "solve_jl.json")


def _vector(text: str, name: str):
    m = re.search(r"^\s*" + re.escape(name) + r"\s*=\s*\[([^\]]*)\]", text, re.M)
    if not m:
        sys.exit(f"{name} not found in {SRC}")
    return [float(t) for t in m.group(1).replace(",", " ").split()]


def main():
    text = open(SRC).read()
    x = _vector(text, "x")
    r_ols = _vector(text, "r_ols")
    interaction = _vector(text, "interaction")
    # p1 = repeat(1:5, inner = 2), p2 = repeat(1:5, outer = 2)  (:3-4)
    assert "repeat(1:5, inner = 2)" in text and "repeat(1:5, outer = 2)" in text
    p1 = [k for k in range(1, 6) for _ in range(2)]
    p2 = [k for _ in range(2) for k in range(1, 6)]
    fixture = {
        "source": "test/solve.jl:3-7,58",
        "p1": p1,
        "p2": p2,
        "x": x,
        "r_ols": r_ols,
        "interaction": interaction,
        "weights_update_1": list(range(1, 11)),
        "weights_update_2": list(range(10, 0, -1)),
        "rtol_float64": 1.4901161193847656e-08,
    }
    with open(OUT, "w") as f:
        json.dump(fixture, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
