"""One-electron integrals S, T, V on the device (pk_one_electron) against the
oracle (oracle/pk_ref.cpp, the reference's os.cpp:248-602 restated) and
against what the mathematics fixes.

Bars: S and T within 1e-13 absolute of the oracle (same quantities by another
evaluation order: closed-form moments here, recurrences there).  V within
1e-12 of the oracle run with its accurate Boys series, and within 3e-10 of the
oracle with the reference's table + 6-term Taylor lookup (that lookup's
~1e-12 error per call, summed over nuclei and primitive pairs, moves V by up to
1.0e-10 on the 8-water case: measured as the difference of the two oracle
modes).
Invariants: unit diagonal of S (the reference's normalisation makes every
Cartesian component a unit function), exact symmetry (one thread writes both
triangles), translation invariance of S, T and V to 1e-12.
"""
import numpy as np
import pytest

import refcpu
from pkfock.host import Molecule
from pkfock.molecules import ANGSTROM_TO_BOHR, alkane, water, water_cluster

pytestmark = pytest.mark.gpu

CASES = [("sto-3g", water()), ("6-31g*", water()), ("cc-pvdz", alkane(2)), ("cc-pvtz", water()),
         ("cc-pvtz", alkane(3)), ("6-31g*", water_cluster(3, n=3)), ("6-31g*", water_cluster(7, n=8))]


@pytest.mark.parametrize("basis,mol", CASES, ids=[c[0] + "-%d" % len(c[1][0]) for c in CASES])
def test_one_electron_matches_oracle(basis, mol):
    Z, X = mol
    print("case %s natoms %d" % (basis, len(Z)), flush=True)
    m = Molecule(Z, X, basis)
    r = refcpu.Ref(basis, Z, X * ANGSTROM_TO_BOHR)
    S, T, V = m.one_electron()
    dS = np.abs(S - r.matrix("overlap")).max()
    dT = np.abs(T - r.matrix("kinetic")).max()
    dV_table = np.abs(V - r.matrix("nuclear")).max()
    with refcpu.exact_boys():
        dV = np.abs(V - r.matrix("nuclear")).max()
    print("nb %d: max |dS| %.2e |dT| %.2e |dV| %.2e (table lookup oracle %.2e)" % (m.nbasis, dS, dT, dV, dV_table))
    assert dS < 1e-13 and dT < 1e-13 and dV < 1e-12 and dV_table < 3e-10
    assert np.array_equal(S, S.T) and np.array_equal(T, T.T) and np.array_equal(V, V.T)
    assert np.abs(np.diag(S) - 1.0).max() < 1e-13


def test_one_electron_translation_invariance():
    Z, X = alkane(3)
    S0, T0, V0 = Molecule(Z, X, "cc-pvtz").one_electron()
    S1, T1, V1 = Molecule(Z, X + np.array([1.7, -2.3, 0.9]), "cc-pvtz").one_electron()
    assert np.abs(S1 - S0).max() < 1e-12
    assert np.abs(T1 - T0).max() < 1e-12
    assert np.abs(V1 - V0).max() < 1e-11 * max(1.0, np.abs(V0).max())


def test_one_electron_single_atom_closed_forms():
    """Hydrogen-atom STO-3G 1s (one function): S = 1, and T and V are the
    textbook closed forms of a contraction of s Gaussians at the nucleus,
        T = sum c_i c_j 3 a_i a_j / (a_i + a_j) * (pi / (a_i + a_j))^1.5,
        V = -sum c_i c_j 2 pi / (a_i + a_j),
    with c_i the normalised contraction coefficients."""
    m = Molecule([1], np.zeros((1, 3)), "sto-3g")
    S, T, V = m.one_electron()
    sh = m.shells()
    a = sh.exponent
    c = sh.coef * sh.prim_norm
    A, B = np.meshgrid(a, a, indexing="ij")
    cc = np.outer(c, c)
    z = A + B
    T_ref = (cc * 3 * A * B / z * (np.pi / z) ** 1.5).sum()
    V_ref = -(cc * 2 * np.pi / z).sum()
    assert abs(S[0, 0] - 1.0) < 1e-14
    assert abs(T[0, 0] - T_ref) < 1e-13
    assert abs(V[0, 0] - V_ref) < 1e-13
