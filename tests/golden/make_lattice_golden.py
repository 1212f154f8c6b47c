"""Writes lattice_d3.npz: exact-arithmetic kNN answers on an integer lattice.

Inputs are small integer coordinates, so every FP32 squared distance is an
exact integer; the expected (index, distance) rows come from int64 arithmetic
and the reference's key order (distance bits, then index), independent of the
oracle and of the HIP library.  Regenerate with:
    python tests/golden/make_lattice_golden.py
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from tests.pins import exact_lattice_knn, lattice_case  # noqa: E402


def main():
    pts, qs = lattice_case(2024, 240, 48, 3, side=5)
    cases = [(np.inf, False, 12), (4.0, False, 12), (0.0, True, 5), (2.0, True, 20), (-1.0, False, 250)]
    out = {"points": pts, "queries": qs, "ncases": np.int32(len(cases))}
    for c, (md, heap, k) in enumerate(cases):
        i, d = exact_lattice_knn(pts, qs, k, md, heap)
        out[f"idx_{c}"] = i
        out[f"dist_{c}"] = d
        out[f"maxd2_{c}"] = np.float32(md)
        out[f"heap_{c}"] = np.int32(heap)
        out[f"k_{c}"] = np.int32(k)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lattice_d3.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
