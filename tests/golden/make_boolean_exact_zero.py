"""Extracts the golden vectors of the reference test
tests/BooleanOperationExactZero (result.hpp: run start indices and values
after a UNION of a box mask and a plane substrate, 2D, x reflective
[-10, 10], y infinite, gridDelta 1) into boolean_exact_zero.json.

Run once in a container that has the reference tree mounted; the JSON is the
committed fixture (data only)."""
import json
import os
import re
import sys

SRC = "/root/reference/tests/BooleanOperationExactZero/result.hpp"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "boolean_exact_zero.json")


def main():
    text = open(SRC).read()
    idx = [[int(a), int(b)] for a, b in re.findall(r"VectorType<double, 2>\{\s*(-?\d+),\s*(-?\d+)\}", text)]
    vals_block = text[text.index("resultValues[]"):]
    vals = [float(v) for v in re.findall(r"(-?\d+(?:\.\d+)?(?:e[+-]?\d+)?)\s*,?\s*\n", vals_block)]
    assert len(idx) == len(vals), (len(idx), len(vals))
    json.dump({"source": "tests/BooleanOperationExactZero/result.hpp", "indices": idx, "values": vals},
              open(OUT, "w"), indent=0)
    print(f"wrote {len(idx)} runs to {OUT}")


if __name__ == "__main__":
    sys.exit(main())
