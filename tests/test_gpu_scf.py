"""GPU RHF SCF (device driver + HIP Fock builds) against the reference's own
energy pins (tests/golden/scf_kat.json) and against the oracle's SCF.

Bar: total energy within 1e-9 Eh of the reference-held value.
"""
import json
import os
import sys

import numpy as np
import pytest

from pkfock.host import Molecule
from pkfock.molecules import ANGSTROM_TO_BOHR, water

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "scf_kat.json")))["cases"]


# Twisted (D2d) ethylene / STO-3G: the two electrons of the degenerate e pair
# doubly occupy one orbital of the pair, and rotating that orbital inside the
# pair leaves the energy unchanged (the oracle's SCF from seven rotation
# angles lands on -76.852246554536 +- 3e-13).  Along that flat direction the
# density drifts with the rounding of each build (the device build accumulates
# J and K with atomics, so its rounding differs run to run), and whether the
# density test passes within the reference's cycle limit is a matter of
# chance (the same SCF algorithm on the oracle's integrals, with symmetric
# 1e-14 noise added to H, met it within 50 cycles in 2 of 8 runs).  The
# energy is what the case pins: from the oracle's converged density the
# device SCF must stay at the reference energy; from the core guess it must
# either converge to it or report non-convergence, never a wrong energy.
FLAT_MODE = {"ethylene_cas44_sto3g_sa2_os_nosymm"}


@pytest.mark.parametrize("case", KAT, ids=[c["id"] for c in KAT])
def test_reference_energy(case):
    Z = [a[0] for a in case["atoms"]]
    X = np.array([a[1:] for a in case["atoms"]])
    mol = Molecule(Z, X, case["basis"])
    r = mol.rhf(engine=case["engine"], tol_eri=1e-10, max_cycles=case.get("max_cycles", 0))
    print("%s: E = %.10f (ref %.10f) iterations %d wall %.2fs" % (
        case["id"], r["energy"], case["energy"], r["iterations"], r["wall_s"]))
    if case["id"] in FLAT_MODE:
        import refcpu
        if r["converged"]:
            assert abs(r["energy"] - case["energy"]) < 1e-9
        P0 = refcpu.Ref(case["basis"], Z, X * ANGSTROM_TO_BOHR, engine=case["engine"]).rhf()["P"]
        r = mol.rhf(engine=case["engine"], tol_eri=1e-10, max_cycles=case.get("max_cycles", 0), guess=P0)
        print("  from the oracle's converged density: E = %.10f iterations %d" % (r["energy"], r["iterations"]))
        assert r["converged"] and r["iterations"] <= 5
        assert abs(r["energy"] - case["energy"]) < 1e-9
        return
    assert r["converged"] or case.get("degenerate_homo")
    assert abs(r["energy"] - case["energy"]) < 1e-9
    if "iterations" in case:
        assert abs(r["iterations"] - case["iterations"]) <= 2


# ---------------------------------------------------------------- SAD guess
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import sad_ref  # noqa: E402

_ETHYLENE_CCPVDZ = ([6, 6, 1, 1, 1, 1], np.array([[-3.438338, 0.906062, 0.0], [-3.438338, 2.230667, -0.000001],
                                                  [-2.836048, 0.343683, -0.709301], [-4.040627, 0.343684, 0.709301],
                                                  [-2.836049, 2.793045, -0.709302], [-4.040627, 2.793046, 0.709300]]))


def _sad_oracle(basis, Z, X):
    from test_oracle import sad_parts
    ref, Praw = sad_parts(basis, Z, np.asarray(X) * ANGSTROM_TO_BOHR)
    return sad_ref.project(Praw, ref.matrix("overlap"), int(sum(Z)))[0]


@pytest.mark.parametrize("case", ["water", "water3", "ethylene"])
def test_sad_density_matches_oracle(case):
    """Device SCF's SAD initial density (atomic UHF on the engine, host
    averaging, device projection) against the oracle's restatement; the cases
    have non-degenerate natural occupations at n_occ, so the projection is
    unique.  Bar 1e-6: both atomic UHFs stop when the density change per
    iteration is below 1e-10, along different DIIS / eigensolver paths (the
    carbon atom's occupied p orbital is a choice inside a degenerate shell);
    measured 2e-9 (water) to 7e-8 (ethylene cc-pVDZ)."""
    from pkfock.molecules import water_cluster
    if case == "water":
        (Z, X), basis = water(), "6-31g*"
    elif case == "water3":
        (Z, X), basis = water_cluster(7, n=3), "6-31g*"
    else:
        (Z, X), basis = _ETHYLENE_CCPVDZ, "cc-pvdz"
    mol = Molecule(Z, X, basis)
    r = mol.rhf(tol_eri=1e-12, sad=True, guess_only=True)
    P_ref = _sad_oracle(basis, Z, X)
    d = np.abs(r["P"] - P_ref).max()
    print("%s/%s SAD max|dP| = %.2e" % (case, basis, d))
    assert d < 1e-6


def test_scf_from_sad_reaches_reference_energies():
    """SCF from the SAD guess reaches the reference energies where the SAD
    projection is unique: ethylene cc-pVDZ (regression_cases.json,
    -78.040190575) and water 6-31G* (the core-guess pin of this file).  Not
    asserted: ethylene STO-3G at the engine_symmetry geometry, whose raw SAD
    density has a degenerate natural-occupation pair exactly at n_occ, so the
    projected guess (and the SCF solution reached from it) depends on the
    eigensolver, in the reference as here."""
    Z, X = _ETHYLENE_CCPVDZ
    r = Molecule(Z, X, "cc-pvdz").rhf(engine="auto", tol_eri=1e-10, sad=True, max_cycles=100)
    print("ethylene cc-pVDZ from SAD: %.10f in %d iterations" % (r["energy"], r["iterations"]))
    assert r["rc"] == 0 and abs(r["energy"] - (-78.040190575)) < 1e-9
    Z, X = water()
    core = Molecule(Z, X, "6-31g*").rhf(tol_eri=1e-12)
    r = Molecule(Z, X, "6-31g*").rhf(tol_eri=1e-12, sad=True)
    print("water 6-31G* from SAD: %.10f in %d iterations (core guess: %d)" % (r["energy"], r["iterations"],
                                                                            core["iterations"]))
    assert r["rc"] == 0 and abs(r["energy"] - core["energy"]) < 1e-9


def test_batched_scf_matches_single_molecule_scf():
    """Batched RHF (one batched Fock build per iteration for all molecules)
    against the single-molecule device RHF of each: same energies (1e-9 Eh)
    and iteration counts within one (the builds differ only in the order of
    the FP64 atomic sums); a learned-guess density supplied per molecule."""
    from pkfock.molecules import water_cluster
    mols = [Molecule(*water_cluster(300 + s, n_min=1, n_max=5), "6-31g*") for s in range(6)]
    mols.append(Molecule(*_ETHYLENE_CCPVDZ, "cc-pvdz"))
    scf = dict(tol_eri=1e-12, tol_energy=1e-9, tol_density=1e-7)
    single = [m.rhf(**scf) for m in mols]
    batch = Molecule.rhf_batch(mols, **scf)
    for s, b in zip(single, batch):
        print("nb %d: single %.10f (%d it)  batch %.10f (%d it)" % (
            b["P"].shape[0], s["energy"], s["iterations"], b["energy"], b["iterations"]))
        assert b["rc"] == 0 and abs(s["energy"] - b["energy"]) < 1e-9
        assert abs(s["iterations"] - b["iterations"]) <= 1
    # converged densities as guesses: every molecule stops after two iterations
    again = Molecule.rhf_batch(mols, guesses=[b["P"] for b in batch], **scf)
    assert all(a["iterations"] <= 3 and abs(a["energy"] - b["energy"]) < 1e-9 for a, b in zip(again, batch))


def test_batched_scf_guess_fallback():
    """Batched RHF with guess_fallback: a deliberately poor guess (a scaled
    identity) for one molecule falls back to the core guess, a converged
    density is kept for the others; every energy equals the plain batched
    result within 1e-9 Eh."""
    from pkfock.molecules import water_cluster
    mols = [Molecule(*water_cluster(310 + s, n_min=2, n_max=4), "6-31g*") for s in range(4)]
    scf = dict(tol_eri=1e-12, tol_energy=1e-9, tol_density=1e-7)
    plain = Molecule.rhf_batch(mols, **scf)
    guesses = [b["P"] for b in plain]
    guesses[1] = 3.0 * np.eye(mols[1].nbasis)
    guesses[2] = None
    fb = Molecule.rhf_batch(mols, guesses=guesses, guess_fallback=True, **scf)
    for i, (p, f) in enumerate(zip(plain, fb)):
        print("molecule %d: guess_used %d  errors %.3e / %.3e  %d it  dE %.2e" % (
            i, f["guess_used"], f["guess_diis_error"], f["core_diis_error"], f["iterations"],
            f["energy"] - p["energy"]))
        assert f["rc"] == 0 and abs(f["energy"] - p["energy"]) < 1e-9
    assert fb[1]["guess_used"] == 0 and fb[1]["core_diis_error"] < fb[1]["guess_diis_error"]
    assert fb[0]["guess_used"] == 1 and fb[3]["guess_used"] == 1 and fb[0]["iterations"] <= 3
    assert fb[2]["guess_used"] == 0
