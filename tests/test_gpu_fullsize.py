"""GPU parity at the full sizes of the benchmark workloads.

n-C20H42 / cc-pVDZ (510 functions) and n-C10H22 / cc-pVTZ (680 functions,
f shells): the engine's J and K (the path of src/integrals/os.cpp:1435-1523)
against oracle values committed in tests/golden/jk_elements_*.npz (written by
tests/golden/make_fullsize_jk.py: every AO quartet of each sampled element
summed on the CPU, no primitive screening).  The bar applies to the oracle with
its accurate Boys function; the same elements with the reference's Boys table
lookup (~1e-12 absolute in F_m, amplified by the high-L d/f integrals of
cc-pVTZ) differ from the accurate values by up to 4e-8, and that difference is
printed beside each check.  Two seeded
densities each; one of them positive everywhere, so that small terms the
engine drops (primitive pairs below its prefactor bound) add up with one sign.

Bars: FP64 J/K within 1e-9 absolute (the Fock bar).  Mixed precision: every
integral of a quartet whose Schwarz bound lies below the cutoff c is computed
in FP32 within 2e-5 of that bound (tests/test_gpu_eri_templates.py measures
1e-5 at most), so an element moves by less than 2e-5 c sum_kl |P_kl|; the test
uses that bound.
"""
import json
import os

import numpy as np
import pytest

from pkfock.engine import PREC_MIXED
from pkfock.host import Molecule
from pkfock.molecules import alkane

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FOCK_TOL = 1e-9
_ENGINES = {}


def _fixture(name):
    d = np.load(os.path.join(GOLDEN, "jk_elements_%s.npz" % name), allow_pickle=False)
    meta = json.loads(str(d["meta"]))
    return d, meta


def _density(n, seed, scale, kind):
    A = np.random.default_rng(seed).normal(size=(n, n))
    if kind == "positive":
        A = np.abs(A)
    return scale * (A + A.T)


def _engine(meta):
    key = meta["workload"]
    if key not in _ENGINES:
        Z, X = alkane(meta["carbons"])
        eng = Molecule(Z, X, meta["basis"]).engine(meta["engine"], tol_eri=meta["tol_eri"])
        assert eng.nbasis == meta["nbasis"]
        _ENGINES[key] = eng
    return _ENGINES[key]


def _densities(meta):
    out = {}
    for kind, seed in meta["seeds"].items():
        P = _density(meta["nbasis"], seed, meta["scale"], kind)
        # the generator must reproduce the density the oracle saw
        assert np.allclose([P.sum(), (P * P).sum()], meta["P_checksum_" + kind], rtol=1e-12, atol=1e-9)
        out[kind] = P
    return out


@pytest.mark.parametrize("name", ["c20h42_ccpvdz", "c10h22_ccpvtz"])
def test_fullsize_fp64_jk_matches_oracle(name):
    d, meta = _fixture(name)
    eng = _engine(meta)
    eng.set("os", meta["tol_eri"])
    r, c = d["rows"], d["cols"]
    for kind, P in _densities(meta).items():
        J, K = eng.jk(P)
        dj = np.abs(J[r, c] - d["J_" + kind]).max()
        dk = np.abs(K[r, c] - d["K_" + kind]).max()
        lj = np.abs(J[r, c] - d["J_%s_lookup" % kind]).max()
        lk = np.abs(K[r, c] - d["K_%s_lookup" % kind]).max()
        print("%s %s: %d elements  max|dJ| %.2e  max|dK| %.2e  (|J| up to %.1f, |K| up to %.1f); "
              "vs the reference-lookup oracle %.2e / %.2e" % (
                  name, kind, len(r), dj, dk, np.abs(d["J_" + kind]).max(), np.abs(d["K_" + kind]).max(), lj, lk))
        assert dj < FOCK_TOL and dk < FOCK_TOL
        assert np.array_equal(J, J.T) and np.array_equal(K, K.T)


@pytest.mark.parametrize("cutoff", [1e-4, 1e-5, 1e-6, 1e-7])
@pytest.mark.parametrize("name", ["c20h42_ccpvdz", "c10h22_ccpvtz"])
def test_fullsize_mixed_jk_matches_oracle(name, cutoff):
    d, meta = _fixture(name)
    eng = _engine(meta)
    eng.set("os", meta["tol_eri"], precision=PREC_MIXED, fp32_cutoff=cutoff)
    r, c = d["rows"], d["cols"]
    try:
        for kind, P in _densities(meta).items():
            J, K = eng.jk(P)
            dj = np.abs(J[r, c] - d["J_" + kind]).max()
            dk = np.abs(K[r, c] - d["K_" + kind]).max()
            bar = max(FOCK_TOL, 2e-5 * cutoff * np.abs(P).sum(axis=1).max())
            n32 = sum(s["quartets"] for s in eng.class_stats() if s["precision"])
            print("%s mixed %g %s: FP32 quartets %.1f%%  max|dJ| %.2e  max|dK| %.2e  bar %.1e" % (
                name, cutoff, kind, 100.0 * n32 / eng.stats()["quartets"], dj, dk, bar))
            assert n32 > 0
            assert dj < bar and dk < bar
    finally:
        eng.set("os", meta["tol_eri"])


def test_fullsize_graph_replays_with_side_queues_match_eager():
    """n-C10H22 / cc-pVTZ holds more than 2^20 shell quartets, so a build
    spreads its class kernels over the context stream and three side streams
    (fork and join by events); captured into a hipGraph, the fork and join are
    part of the graph.  Four replays, with eager builds between them, against
    the eager build of the same density."""
    d, meta = _fixture("c10h22_ccpvtz")
    eng = _engine(meta)
    eng.set("os", meta["tol_eri"])
    P = _densities(meta)["signed"]
    assert eng.stats()["quartets"] >= 1 << 20
    J0, K0 = eng.jk(P)
    try:
        eng.set_graph(True)
        for rep in range(4):
            Jg, Kg = eng.jk(P)
            eng.set_graph(False)
            Je, Ke = eng.jk(P)  # eager work between the replays
            eng.set_graph(True)
            dj = max(np.abs(Jg - J0).max(), np.abs(Je - J0).max())
            dk = max(np.abs(Kg - K0).max(), np.abs(Ke - K0).max())
            print("replay %d: max|dJ| %.2e  max|dK| %.2e" % (rep, dj, dk))
            # the FP64 atomics add in a different order on every build
            assert dj < 1e-12 * max(1.0, np.abs(J0).max()) and dk < 1e-12 * max(1.0, np.abs(K0).max())
    finally:
        eng.set_graph(False)
