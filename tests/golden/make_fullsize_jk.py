"""Writes the full-size J/K fixtures: oracle J and K at a stratified sample of
elements for the alkane workloads (n-C20H42 / cc-pVDZ, 510 functions;
n-C10H22 / cc-pVTZ, 680 functions), Schwarz 1e-12, engine os.

The oracle's full build at these sizes takes hours on the CPU; its element
sampler (pkref_jk_elements) sums every AO quartet of one element on its own
with the screening of the full build, so a sample of elements costs minutes.

Two sets of values per density: J_<kind>, K_<kind> with the oracle's accurate
Boys function (pkref_set_boys_exact, ~1e-17 relative; the bar of the tests),
and J_<kind>_lookup, K_<kind>_lookup with the reference's own table lookup
(src/lookup/boys.cpp:1592-1618, ~1e-12 absolute), whose error the high-L
integrals of cc-pVTZ amplify to a few 1e-8 in some d/f elements; the tests
report that difference beside their check.
Two densities per molecule, regenerated from their seeds by the tests:
  signed:   P = 0.05 (A + A^T),        A ~ N(0, 1)  (numpy default_rng(seed))
  positive: P = 0.05 (|A| + |A|^T)     (every product of P adds up with one sign,
            the worst case for small dropped terms such as screened primitives)
Elements: for every pair of shell types (s, p, d, f) of the two functions a
fixed number of random (i, j), plus random diagonal elements.

Run in the build container (CPU only): python tests/golden/make_fullsize_jk.py [workload ...]
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "planck-refactored_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from pkfock.molecules import ANGSTROM_TO_BOHR, alkane  # noqa: E402
from refcpu import Ref, exact_boys  # noqa: E402

WORKLOADS = {"c20h42_ccpvdz": (20, "cc-pvdz"), "c10h22_ccpvtz": (10, "cc-pvtz")}
SEEDS = {"signed": 101, "positive": 102}
PER_TYPE_PAIR = 24
DIAGONAL = 12


def density(n, kind):
    A = np.random.default_rng(SEEDS[kind]).normal(size=(n, n))
    if kind == "positive":
        A = np.abs(A)
    return 0.05 * (A + A.T)


def function_l(ref):
    _, ls, _, _, _, _ = ref.shells()
    return np.concatenate([[l] * ((l + 1) * (l + 2) // 2) for l in ls])


def elements(fl, seed=5):
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    types = sorted(set(fl.tolist()))
    for a in types:
        for b in types:
            if b < a:
                continue
            ia, ib = np.flatnonzero(fl == a), np.flatnonzero(fl == b)
            for _ in range(PER_TYPE_PAIR):
                rows.append(int(rng.choice(ia)))
                cols.append(int(rng.choice(ib)))
    d = rng.choice(len(fl), size=DIAGONAL, replace=False)
    rows += [int(x) for x in d]
    cols += [int(x) for x in d]
    return np.array(rows, dtype=np.int32), np.array(cols, dtype=np.int32)


def main(names):
    for name in names:
        ncarb, basis = WORKLOADS[name]
        Z, X = alkane(ncarb)
        ref = Ref(basis, Z, X * ANGSTROM_TO_BOHR, engine="os", tol_eri=1e-12)
        n = ref.nbasis
        fl = function_l(ref)
        rows, cols = elements(fl)
        out = {"rows": rows, "cols": cols}
        meta = {"workload": name, "basis": basis, "carbons": ncarb, "nbasis": n, "tol_eri": 1e-12,
                "engine": "os", "seeds": SEEDS, "scale": 0.05, "boys": "accurate; *_lookup: reference lookup"}
        for kind in SEEDS:
            P = density(n, kind)
            meta["P_checksum_" + kind] = [float(P.sum()), float((P * P).sum())]
            t = time.perf_counter()
            J, K = ref.jk_elements(P, rows, cols)
            out["J_%s_lookup" % kind] = J
            out["K_%s_lookup" % kind] = K
            with exact_boys():
                J, K = ref.jk_elements(P, rows, cols)
            print("%s %s: %d elements in %.1f s" % (name, kind, len(rows), time.perf_counter() - t), flush=True)
            out["J_" + kind] = J
            out["K_" + kind] = K
        out["meta"] = np.array(json.dumps(meta))
        np.savez_compressed(os.path.join(HERE, "jk_elements_%s.npz" % name), **out)


if __name__ == "__main__":
    main(sys.argv[1:] or sorted(WORKLOADS))
