"""ERI-level parity of the J/K kernels themselves (not only of their sums).

pk_eri_tensor runs the cooperative kernel; the per-class OS and Rys templates
that do most of a build never write an integral out.  Their integrals are read
back through J instead: for the unit pair density P = e_k e_l^T + e_l e_k^T
(k != l; e_k e_k^T for k = l), J[P]_ij = sum_ls (ij|ls) P_ls equals
2 (ij|kl) (resp. (ij|kk)), so one build per pair (k, l) yields the column
(.. | kl) of the ERI tensor as computed and digested by the hot kernels.  The
columns are compared with the oracle's tensor (src/integrals/os.cpp:1755-1829
semantics, engine os) at the 1e-11 ERI bar; the FP32 instantiations (used by
mixed precision below its cutoff) at a relative bar.
"""
import numpy as np
import pytest

from pkfock.engine import FockEngine, Shells, PREC_MIXED
from pkfock.molecules import ANGSTROM_TO_BOHR, water
from refcpu import Ref

pytestmark = pytest.mark.gpu

ERI_TOL = 1e-11
ETHYLENE = ([6, 6, 1, 1, 1, 1],
            np.array([[-3.438338, 0.906062, 0.0], [-3.438338, 2.230667, -0.000001],
                      [-2.836048, 0.343683, -0.709301], [-4.040627, 0.343684, 0.709301],
                      [-2.836049, 2.793045, -0.709302], [-4.040627, 2.793046, 0.709300]]))
_CACHE = {}


def _setup(basis, mol):
    key = (basis, mol)
    if key not in _CACHE:
        Z, X = water() if mol == "water" else ETHYLENE
        ref = Ref(basis, Z, X * ANGSTROM_TO_BOHR, engine="os", tol_eri=1e-12)
        _CACHE[key] = (ref, ref.eri_full())
    ref, T = _CACHE[key]
    eng = FockEngine(Shells(*ref.shells()), engine="os", tol_eri=1e-12)
    return ref, T, eng


def _columns(eng, n):
    """ERI tensor of the J/K kernels: out[:, :, k, l] from unit pair densities."""
    out = np.zeros((n, n, n, n))
    P = np.zeros((n, n))
    for k in range(n):
        for l in range(k, n):
            P[k, l] = P[l, k] = 1.0
            J, _ = eng.jk(P)
            P[k, l] = P[l, k] = 0.0
            col = J if k == l else 0.5 * J
            out[:, :, k, l] = col
            out[:, :, l, k] = col
    return out


@pytest.mark.parametrize("engine", ["os", "rys"])
@pytest.mark.parametrize("basis,mol", [("6-31g*", "water"), ("cc-pvdz", "ethylene"), ("cc-pvtz", "water")])
def test_template_eris_match_oracle(basis, mol, engine):
    ref, T, eng = _setup(basis, mol)
    eng.set(engine, 1e-12)
    n = ref.nbasis
    E = _columns(eng, n)
    d = np.abs(E - T).max()
    kinds = sorted({(c["la"], c["lb"], c["lc"], c["ld"]) for c in eng.class_stats()})
    print("%s/%s engine %s nb=%d: %d classes  max|dERI| = %.3e" % (mol, basis, engine, n, len(kinds), d))
    assert d < ERI_TOL


@pytest.mark.parametrize("basis,mol", [("6-31g*", "water"), ("cc-pvtz", "water")])
def test_fp32_template_eris(basis, mol):
    """Every per-lane class kernel in its FP32 instantiation (cutoff above every
    Schwarz bound).  The mixed-precision path sends a quartet to FP32 when its
    Schwarz bound Q_ij Q_kl lies below the cutoff, so the error that matters is
    relative to that bound: every integral within 2e-5 Q_ij Q_kl (FP32
    cancellation in the transfers makes the error relative to the integral
    itself larger for integrals far below their bound)."""
    ref, T, eng = _setup(basis, mol)
    eng.set("os", 1e-12, precision=PREC_MIXED, fp32_cutoff=1e30)
    n = ref.nbasis
    E = _columns(eng, n)
    n32 = sum(c["quartets"] for c in eng.class_stats() if c["precision"])
    Q = ref.matrix("schwarz")
    bound = np.einsum("ij,kl->ijkl", Q, Q)
    rel = np.abs(E - T) / np.maximum(bound, 1e-30)
    print("%s/%s FP32: %d of %d quartets in FP32  max|dERI| = %.3e  max |dERI| / (Q_ij Q_kl) = %.3e" % (
        mol, basis, n32, eng.stats()["quartets"], np.abs(E - T).max(), rel.max()))
    assert n32 > 0
    assert rel.max() < 2e-5
