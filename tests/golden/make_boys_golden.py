#!/usr/bin/env python3
"""Write boys_ref.json: golden vectors of the reference's own Boys lookup.

Runs in the build container only (needs the reference tree): it loads
oracle/_ref/libboys_ref.so, compiled by `make -C oracle ref` from the
reference's src/lookup/boys.cpp (Lookup::boys, boys.cpp:1592-1618), and records
F_n(x) for n = 0..32 at every 7th table grid point, at seeded off-grid points
in [0, 157.1) and beyond the table (asymptotic branch).
"""
import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libboys_ref.so"))
    lib.ref_lookup_boys.restype = ctypes.c_double
    lib.ref_lookup_boys.argtypes = [ctypes.c_int, ctypes.c_double]
    rng = np.random.default_rng(20261202)
    xs = [i * 0.1 for i in range(0, 1571, 7)]
    xs += list(rng.uniform(0.0, 157.1, 300))
    xs += list(rng.uniform(157.1, 400.0, 40))
    xs += [1e-14, 1e-8, 1e-3, 0.05, 157.09999, 157.1]
    rows = []
    for x in xs:
        rows.append([float(x)] + [lib.ref_lookup_boys(n, float(x)) for n in range(33)])
    with open(os.path.join(HERE, "boys_ref.json"), "w") as f:
        json.dump({"source": "reference Lookup::boys (src/lookup/boys.cpp:1592-1618) via oracle/_ref",
                   "orders": 33, "rows": rows}, f)
    print("wrote", len(rows), "rows")


if __name__ == "__main__":
    main()
