"""GPU parity: the HIP engine against the CPU oracle on the same inputs.

Bars (from the engine's contract): contracted ERIs within 1e-11 absolute,
J and K (so the Fock matrix) within 1e-9 absolute.
"""
import numpy as np
import pytest

from pkfock.engine import FockEngine, Shells
from pkfock.molecules import ANGSTROM_TO_BOHR, alkane, water
from refcpu import Ref

pytestmark = pytest.mark.gpu

ERI_TOL = 1e-11
FOCK_TOL = 1e-9

ETHYLENE = ([6, 6, 1, 1, 1, 1],
            np.array([[-3.438338, 0.906062, 0.0], [-3.438338, 2.230667, -0.000001],
                      [-2.836048, 0.343683, -0.709301], [-4.040627, 0.343684, 0.709301],
                      [-2.836049, 2.793045, -0.709302], [-4.040627, 2.793046, 0.709300]]))
# carbon monoxide, 1.128 A: f shells on two centres (every f class occurs)
CO = ([6, 8], np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.128]]))


def _molecule(mol):
    if mol == "water":
        return water()
    if mol == "ethylene":
        return ETHYLENE
    if mol == "co":
        return CO
    return alkane(3)


def _pair(basis, Z, X, engine="os", tol=1e-10):
    ref = Ref(basis, Z, X * ANGSTROM_TO_BOHR, engine=engine, tol_eri=tol)
    eng = FockEngine(Shells(*ref.shells()), engine=engine, tol_eri=tol)
    assert eng.nbasis == ref.nbasis
    return ref, eng


def _density(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n, n)) * 0.2
    return A + A.T


@pytest.mark.parametrize("engine", ["os", "rys"])
@pytest.mark.parametrize("basis,mol", [("sto-3g", "water"), ("6-31g*", "water"), ("cc-pvdz", "ethylene"),
                                      ("cc-pvtz", "water")])
def test_schwarz_and_eri_tensor(basis, mol, engine):
    Z, X = _molecule(mol)
    ref, eng = _pair(basis, Z, X, tol=1e-12)
    eng.set(engine, 1e-12)
    dq = np.abs(eng.schwarz() - ref.matrix("schwarz")).max()
    print("%s/%s nb=%d  max|dQ| = %.3e" % (mol, basis, ref.nbasis, dq))
    assert dq < ERI_TOL
    d = np.abs(eng.eri_tensor() - ref.eri_full()).max()
    print("%s/%s  max|dERI| = %.3e" % (mol, basis, d))
    assert d < ERI_TOL


@pytest.mark.parametrize("engine", ["os", "rys", "auto"])
@pytest.mark.parametrize("basis,mol,tol", [("6-31g*", "water", 1e-12), ("cc-pvdz", "ethylene", 1e-10),
                                           ("cc-pvdz", "propane", 1e-12), ("cc-pvtz", "water", 1e-12),
                                           ("cc-pvtz", "co", 1e-12)])
def test_jk_matches_oracle(basis, mol, tol, engine):
    Z, X = _molecule(mol)
    ref, eng = _pair(basis, Z, X, tol=tol)
    eng.set(engine, tol)
    P = _density(ref.nbasis, 7)
    J, K = eng.jk(P)
    Jr, Kr = ref.jk(P)
    dj, dk = np.abs(J - Jr).max(), np.abs(K - Kr).max()
    print("%s/%s nb=%d tol=%g  max|dJ| = %.3e  max|dK| = %.3e  |J|max %.2f" % (
        mol, basis, ref.nbasis, tol, dj, dk, np.abs(Jr).max()))
    assert dj < FOCK_TOL and dk < FOCK_TOL
    G = eng.g(P)
    assert np.abs(G - (Jr - 0.5 * Kr)).max() < FOCK_TOL


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("basis,mol", [("6-31g*", "water"), ("cc-pvdz", "ethylene")])
def test_uhf_jk_matches_oracle(basis, mol, graph):
    """UHF variant (_compute_2e_fock_uhf, base.h:75-93): Jt = J[Pa + Pb],
    Ka = K[Pa], Kb = K[Pb] against the oracle's J and K of each spin density;
    host arrays, then device arrays (torch), eager and hipGraph-replayed."""
    import torch
    Z, X = _molecule(mol)
    ref, eng = _pair(basis, Z, X, tol=1e-12)
    eng.set("os", 1e-12)
    eng.set_graph(graph)
    n = ref.nbasis
    Pa, Pb = _density(n, 21), 0.5 * _density(n, 22)
    Jt, Ka, Kb = eng.jk_uhf(Pa, Pb)
    Ja_r, Ka_r = ref.jk(Pa)
    Jb_r, Kb_r = ref.jk(Pb)
    d = [np.abs(Jt - (Ja_r + Jb_r)).max(), np.abs(Ka - Ka_r).max(), np.abs(Kb - Kb_r).max()]
    print("%s/%s nb=%d graph=%s  max|dJt| %.2e |dKa| %.2e |dKb| %.2e" % (mol, basis, n, graph, *d))
    assert max(d) < FOCK_TOL
    dev = torch.device("cuda")
    tPa, tPb = torch.tensor(Pa, device=dev), torch.tensor(Pb, device=dev)
    out = [torch.empty_like(tPa) for _ in range(3)]
    for _ in range(2):  # two calls in a row (graph replay on the second)
        eng.jk_uhf_device(tPa.data_ptr(), tPb.data_ptr(), *(o.data_ptr() for o in out))
        torch.cuda.synchronize()
        for o, r in zip(out, (Jt, Ka, Kb)):
            assert np.abs(o.cpu().numpy() - r).max() < 1e-12


@pytest.mark.parametrize("cutoff", [1e-4, 1e-6])
@pytest.mark.parametrize("basis,mol", [("cc-pvdz", "propane"), ("cc-pvtz", "propane")])
def test_mixed_precision_jk(basis, mol, cutoff):
    """Mixed precision: quartets with Schwarz bound below the cutoff in FP32.
    Each FP32 integral carries a relative error of a few 1e-7 on a value below
    the cutoff; the bar is 1e-5 * cutoff * nb^2 / 100 on J and K.  With cutoff
    0 (FP64 quartets only) the FP64 bar holds.  Reference: the oracle for
    cc-pVDZ; for cc-pVTZ (225 functions, f shells) the FP64 engine, itself
    pinned against the oracle on the f-shell cases above."""
    from pkfock.engine import PREC_MIXED
    Z, X = _molecule(mol)
    tol = 1e-12
    if basis == "cc-pvdz":
        ref, eng = _pair(basis, Z, X, tol=tol)
        nb = ref.nbasis
        P = _density(nb, 11)
        Jr, Kr = ref.jk(P)
    else:
        from pkfock.host import Molecule
        eng = FockEngine(Molecule(Z, X, basis).shells(), "os", tol)
        nb = eng.nbasis
        P = _density(nb, 11)
        Jr, Kr = eng.jk(P)
    eng.set("os", tol, precision=PREC_MIXED, fp32_cutoff=0.0)
    J, K = eng.jk(P)
    assert max(np.abs(J - Jr).max(), np.abs(K - Kr).max()) < FOCK_TOL
    eng.set("os", tol, precision=PREC_MIXED, fp32_cutoff=cutoff)
    J, K = eng.jk(P)
    dj, dk = np.abs(J - Jr).max(), np.abs(K - Kr).max()
    n32 = sum(c["quartets"] for c in eng.class_stats() if c["precision"])
    print("%s/%s cutoff %g: %d FP32 quartets of %d  max|dJ| = %.3e  max|dK| = %.3e" % (
        mol, basis, cutoff, n32, eng.stats()["quartets"], dj, dk))
    assert n32 > 0
    bar = 1e-5 * cutoff * nb ** 2 / 100
    assert dj < bar and dk < bar
