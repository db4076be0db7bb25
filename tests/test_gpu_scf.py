"""GPU RHF SCF (device driver + HIP Fock builds) against the reference's own
energy pins (tests/golden/scf_kat.json) and against the oracle's SCF.

Bar: total energy within 1e-9 Eh of the reference-held value.
"""
import json
import os

import numpy as np
import pytest

from pkfock.host import Molecule

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "scf_kat.json")))["cases"]


@pytest.mark.parametrize("case", KAT, ids=[c["id"] for c in KAT])
def test_reference_energy(case):
    Z = [a[0] for a in case["atoms"]]
    X = np.array([a[1:] for a in case["atoms"]])
    mol = Molecule(Z, X, case["basis"])
    r = mol.rhf(engine=case["engine"], tol_eri=1e-10, max_cycles=case.get("max_cycles", 0))
    print("%s: E = %.10f (ref %.10f) iterations %d wall %.2fs" % (
        case["id"], r["energy"], case["energy"], r["iterations"], r["wall_s"]))
    assert r["converged"] or case.get("degenerate_homo")
    assert abs(r["energy"] - case["energy"]) < 1e-9
    if "iterations" in case:
        assert abs(r["iterations"] - case["iterations"]) <= 2
