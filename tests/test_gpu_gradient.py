"""GPU parity of the two-electron gradient (pk_eri_gradient).

The device contracts ERI first derivatives with the two-particle density
(src/gradient/gradient.cpp:209-232 over src/integrals/os.cpp:1296-1420); the
oracle restates that loop per AO quartet (oracle/pk_ref.cpp, eri_gradient),
pinned by finite differences of its own two-electron energy
(tests/test_oracle.py).  Bar: 1e-9 Eh/Bohr per atom component (the Fock bar).

Cases: water 6-31G* and ethylene 6-31G* against the live oracle; ethylene
cc-pVDZ and water cc-pVTZ (every f class up to (ff|ff)) against the committed
oracle values of tests/golden/eri_gradient_*.npz.  Closed-shell form
(Pj = Pk = P, kx = 1), open-shell form (Pj = Pa + Pb, Pk = Pa, Pb, kx = 2) and
a hybrid exchange fraction; engine os and rys lists, mixed-precision lists
(the gradient runs FP64 over both), translation invariance.
"""
import json
import os

import numpy as np
import pytest

from pkfock.engine import FockEngine, Shells, PREC_MIXED
from pkfock.molecules import ANGSTROM_TO_BOHR, water
from refcpu import Ref

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-9
ETHYLENE = ([6, 6, 1, 1, 1, 1],
            np.array([[-3.438338, 0.906062, 0.0], [-3.438338, 2.230667, -0.000001],
                      [-2.836048, 0.343683, -0.709301], [-4.040627, 0.343684, 0.709301],
                      [-2.836049, 2.793045, -0.709302], [-4.040627, 2.793046, 0.709300]]))


def _density(n, seed, scale=0.1):
    A = np.random.default_rng(seed).normal(size=(n, n)) * scale
    return A + A.T


def _mol(name):
    Z, X = ETHYLENE if name == "ethylene" else water()
    return np.asarray(Z), np.asarray(X, dtype=np.float64) * ANGSTROM_TO_BOHR


@pytest.mark.parametrize("mol", ["water", "ethylene"])
def test_gradient_matches_oracle_631gs(mol):
    Z, Xb = _mol(mol)
    ref = Ref("6-31g*", Z, Xb, engine="os", tol_eri=1e-12)
    eng = FockEngine(Shells(*ref.shells()), engine="os", tol_eri=1e-12)
    n = ref.nbasis
    Pa, Pb = _density(n, 1), _density(n, 2)
    for label, args in (("rhf", (Pa, Pa, None, 1.0)), ("uhf", (Pa + Pb, Pa, Pb, 2.0)),
                        ("hybrid 0.25", (Pa, Pa, None, 0.25))):
        g_ref = ref.eri_gradient(*args)
        g = eng.eri_gradient_atoms(Xb, *args)
        d = np.abs(g - g_ref).max()
        print("%s/6-31G* %s: max|dg| = %.2e (|g| up to %.2f), sum %.1e" % (
            mol, label, d, np.abs(g_ref).max(), np.abs(g.sum(axis=0)).max()))
        assert d < TOL
        assert np.abs(g.sum(axis=0)).max() < 1e-10


@pytest.mark.parametrize("name", ["ethylene_ccpvdz", "water_ccpvtz"])
@pytest.mark.parametrize("engine", ["os", "rys"])
def test_gradient_matches_oracle_fixtures(name, engine):
    d = np.load(os.path.join(GOLDEN, "eri_gradient_%s.npz" % name), allow_pickle=False)
    meta = json.loads(str(d["meta"]))
    Z = np.array(meta["Z"])
    Xb = np.array(meta["xyz_angstrom"]) * ANGSTROM_TO_BOHR
    ref = Ref(meta["basis"], Z, Xb, engine="os", tol_eri=meta["tol_eri"])
    eng = FockEngine(Shells(*ref.shells()), engine=engine, tol_eri=meta["tol_eri"])
    n = meta["nbasis"]
    assert eng.nbasis == n
    Pa = _density(n, meta["seeds"][0], meta["scale"])
    Pb = _density(n, meta["seeds"][1], meta["scale"])
    assert np.allclose([Pa.sum(), (Pa * Pa).sum(), Pb.sum(), (Pb * Pb).sum()], meta["P_checksum"], rtol=1e-12)
    for label, args in (("rhf", (Pa, Pa, None, 1.0)), ("uhf", (Pa + Pb, Pa, Pb, 2.0))):
        g = eng.eri_gradient_atoms(Xb, *args)
        dg = np.abs(g - d["g_" + label]).max()
        print("%s engine %s %s: max|dg| = %.2e (|g| up to %.2f)" % (name, engine, label, dg,
                                                                   np.abs(d["g_" + label]).max()))
        assert dg < TOL


def test_gradient_with_mixed_precision_lists():
    """In mixed precision the energy build splits each class list at the FP32
    cutoff; the gradient covers both halves in FP64, so it is unchanged."""
    Z, Xb = _mol("water")
    ref = Ref("cc-pvdz", Z, Xb, engine="os", tol_eri=1e-12)
    eng = FockEngine(Shells(*ref.shells()), engine="os", tol_eri=1e-12)
    P = _density(ref.nbasis, 5)
    g64 = eng.eri_gradient_atoms(Xb, P)
    eng.set("os", 1e-12, precision=PREC_MIXED, fp32_cutoff=1e-3)
    g = eng.eri_gradient_atoms(Xb, P)
    assert np.abs(g - g64).max() < 1e-12
    assert np.abs(g - ref.eri_gradient(P)).max() < TOL
