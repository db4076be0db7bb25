"""Batched Fock builds (many molecules in one context), hipGraph replay and the
launch-shape autotuner, against per-molecule builds and the CPU oracle.

The batched build merges the quartet lists of all molecules per angular class;
every molecule's J and K must equal its own single-molecule build (the
summation order of the FP64 atomics differs, so within 1e-12, not bitwise) and
the oracle (1e-9).
"""
import numpy as np
import pytest

from pkfock.engine import FockEngine, PREC_MIXED
from pkfock.host import Molecule
from pkfock.molecules import ANGSTROM_TO_BOHR, water, water_cluster
from refcpu import Ref

pytestmark = pytest.mark.gpu

SAME = 1e-12
FOCK_TOL = 1e-9


def _density(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n, n)) * 0.2
    return A + A.T


def _batch():
    # water clusters of 1-3 molecules (6-31G*) and one cc-pVDZ water: uneven
    # sizes and different class sets per molecule
    mols = [(*water_cluster(s, n_min=1, n_max=3), "6-31g*") for s in range(5)]
    mols.append((*water(), "cc-pvdz"))
    return mols


def test_batch_matches_single_and_oracle():
    mols = _batch()
    shells = [Molecule(Z, X, b).shells() for Z, X, b in mols]
    eng = FockEngine(shells, "os", 1e-12)
    assert eng.nmol == len(mols)
    Ps = [_density(s.nbasis, 3 + i) for i, s in enumerate(shells)]
    Jb, Kb = eng.jk_batch(Ps)
    for i, (s, P) in enumerate(zip(shells, Ps)):
        single = FockEngine(s, "os", 1e-12)
        J, K = single.jk(P)
        d = max(np.abs(J - Jb[i]).max(), np.abs(K - Kb[i]).max())
        print("molecule %d nb=%d  batch vs single %.2e" % (i, s.nbasis, d))
        assert d < SAME
    for i in (0, len(mols) - 1):
        Z, X, b = mols[i]
        ref = Ref(b, Z, X * ANGSTROM_TO_BOHR, engine="os", tol_eri=1e-12)
        Jr, Kr = ref.jk(Ps[i])
        d = max(np.abs(Jr - Jb[i]).max(), np.abs(Kr - Kb[i]).max())
        print("molecule %d  batch vs oracle %.2e" % (i, d))
        assert d < FOCK_TOL
    Q = eng.schwarz()
    for i, s in enumerate(shells):
        assert np.abs(Q[i] - FockEngine(s, "os", 1e-12).schwarz()).max() < SAME


def test_graph_replays_match_eager():
    mols = _batch()
    shells = [Molecule(Z, X, b).shells() for Z, X, b in mols]
    eager = FockEngine(shells, "os", 1e-12)
    graph = FockEngine(shells, "os", 1e-12)
    graph.set_graph(True)
    for rep in range(4):
        Ps = [_density(s.nbasis, 100 * rep + i) for i, s in enumerate(shells)]
        Je, Ke = eager.jk_batch(Ps)   # eager work between the replays
        Jg, Kg = graph.jk_batch(Ps)
        d = max(max(np.abs(a - b).max() for a, b in zip(Je, Jg)), max(np.abs(a - b).max() for a, b in zip(Ke, Kg)))
        print("replay %d: graph vs eager %.2e" % (rep, d))
        assert d < SAME


def test_autotune_keeps_results():
    Z, X = water_cluster(7, n=4)
    s = Molecule(Z, X, "6-31g*").shells()
    eng = FockEngine(s, "os", 1e-12)
    P = _density(s.nbasis, 5)
    J0, K0 = eng.jk(P)
    eng.autotune(reps=1)
    shapes = [(c["block"], c["remap"]) for c in eng.class_stats()]
    print("tuned shapes", shapes)
    assert all(b in (64, 128, 256) and r in (0, 1) for b, r in shapes)
    J1, K1 = eng.jk(P)
    assert max(np.abs(J1 - J0).max(), np.abs(K1 - K0).max()) < SAME
    eng.set_graph(True)
    J2, K2 = eng.jk(P)
    assert max(np.abs(J2 - J0).max(), np.abs(K2 - K0).max()) < SAME


def test_batch_mixed_precision():
    mols = _batch()
    shells = [Molecule(Z, X, b).shells() for Z, X, b in mols]
    eng = FockEngine(shells, "os", 1e-12)
    Ps = [_density(s.nbasis, 40 + i) for i, s in enumerate(shells)]
    J64, K64 = eng.jk_batch(Ps)
    eng.set("os", 1e-12, precision=PREC_MIXED, fp32_cutoff=1e-6)
    J32, K32 = eng.jk_batch(Ps)
    d = max(max(np.abs(a - b).max() for a, b in zip(J64, J32)), max(np.abs(a - b).max() for a, b in zip(K64, K32)))
    n32 = sum(c["quartets"] for c in eng.class_stats() if c["precision"])
    print("batch mixed 1e-6: %d FP32 quartets, max diff vs FP64 %.2e" % (n32, d))
    assert n32 > 0 and d < 1e-7


def test_all_256_water_clusters_batched():
    """The batching workload at its full size: the 256 seeded water clusters of
    1-16 molecules (6-31G*, 19-304 functions) in one batched, graph-replayed
    build equal their 256 single-molecule builds within 1e-12, and 8 of them
    (spread over the sizes) equal the oracle's J and K within 1e-9 on 48
    sampled elements each."""
    import torch
    mols = [water_cluster(s, n_min=1, n_max=16) for s in range(256)]
    shells = [Molecule(Z, X, "6-31g*").shells() for Z, X in mols]
    eng = FockEngine(shells, "os", 1e-12)
    rng = np.random.default_rng(17)
    Ps = []
    for n in eng.nbs:
        A = rng.normal(size=(int(n), int(n))) * 0.05
        Ps.append(A + A.T)
    dev = torch.device("cuda")
    P = torch.tensor(eng.pack(Ps), device=dev)
    J, K = torch.empty_like(P), torch.empty_like(P)
    eng.set_stream(torch.cuda.current_stream().cuda_stream)
    eng.set_graph(True)
    for _ in range(2):  # capture, then replay
        eng.jk_device(P.data_ptr(), J.data_ptr(), K.data_ptr())
    torch.cuda.synchronize()
    Jb, Kb = eng.unpack(J.cpu().numpy()), eng.unpack(K.cpu().numpy())
    worst = 0.0
    for i, sh in enumerate(shells):
        single = FockEngine(sh, "os", 1e-12)
        Js, Ks = single.jk(Ps[i])
        worst = max(worst, np.abs(Js - Jb[i]).max(), np.abs(Ks - Kb[i]).max())
        single.close()
    print("256 clusters, %d functions in all: max |batch - single| = %.2e" % (eng.nbasis, worst))
    assert worst < SAME
    sizes = np.array([len(Z) // 3 for Z, _ in mols])
    pick = [int(np.flatnonzero(sizes == s)[0]) for s in (1, 3, 5, 7, 9, 11, 14, 16) if (sizes == s).any()]
    for i in pick:
        Z, X = mols[i]
        ref = Ref("6-31g*", Z, X * ANGSTROM_TO_BOHR, engine="os", tol_eri=1e-12)
        n = ref.nbasis
        r = rng.integers(0, n, 48)
        c = rng.integers(0, n, 48)
        Je, Ke = ref.jk_elements(Ps[i], r, c)
        dj, dk = np.abs(Jb[i][r, c] - Je).max(), np.abs(Kb[i][r, c] - Ke).max()
        print("cluster %d (%d waters, nb %d): max|dJ| %.2e max|dK| %.2e" % (i, sizes[i], n, dj, dk))
        assert dj < FOCK_TOL and dk < FOCK_TOL
