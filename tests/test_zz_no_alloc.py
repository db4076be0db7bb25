"""pk_fock_jk_noalloc: the J/K build into the library's static device scratch
(no device allocation in the call), checked against pk_fock_jk on the same
context and density, on two calls in a row on one GPU (the second with a
smaller molecule, so the shared scratch holds the first call's larger
contents), and its size limit."""
import ctypes

import numpy as np
import pytest
import torch

from pkfock import engine as E
from pkfock.engine import FockEngine
from pkfock.host import Molecule
from pkfock.molecules import alkane, water, water_cluster

pytestmark = pytest.mark.gpu


def _lib():
    lib = E.load()
    lib.pk_fock_jk_noalloc.argtypes = [ctypes.c_void_p] * 4
    return lib


def _density(n, seed):
    A = np.random.default_rng(seed).normal(size=(n, n)) * 0.05
    return A + A.T


def _check(eng, seed):
    dev = torch.device("cuda:0")
    n = eng.nbasis
    P = torch.tensor(_density(n, seed), device=dev)
    J0, K0 = torch.empty_like(P), torch.empty_like(P)
    stream = torch.cuda.current_stream(dev)
    eng.set_stream(stream.cuda_stream)
    eng.jk_device(P.data_ptr(), J0.data_ptr(), K0.data_ptr())  # builds the work lists
    J, K = torch.full_like(P, np.nan), torch.full_like(P, np.nan)
    rc = _lib().pk_fock_jk_noalloc(eng.ctx, ctypes.c_void_p(P.data_ptr()), ctypes.c_void_p(J.data_ptr()),
                                   ctypes.c_void_p(K.data_ptr()))
    assert rc == 0, E.load().pk_last_error().decode()
    torch.cuda.synchronize()
    dj = (J - J0).abs().max().item()
    dk = (K - K0).abs().max().item()
    print("nb=%d: max|J - J_alloc| %.2e  max|K - K_alloc| %.2e" % (n, dj, dk), flush=True)
    assert dj < 1e-12 and dk < 1e-12


def test_noalloc_matches_allocating_build_twice():
    Z, X = alkane(4)
    big = Molecule(Z, X, "cc-pvdz").engine("os", tol_eri=1e-12)
    Z, X = water()
    small = Molecule(Z, X, "6-31g*").engine("os", tol_eri=1e-12)
    assert small.nbasis < big.nbasis
    _check(big, 1)
    _check(small, 2)


def test_noalloc_refuses_unbuilt_lists_and_oversize():
    Z, X = water()
    eng = Molecule(Z, X, "6-31g*").engine("os", tol_eri=1e-12)
    eng.set("os", 1e-11)  # work lists of the new settings not built yet
    dev = torch.device("cuda:0")
    P = torch.zeros(eng.nbasis, eng.nbasis, dtype=torch.float64, device=dev)
    rc = _lib().pk_fock_jk_noalloc(eng.ctx, ctypes.c_void_p(P.data_ptr()), ctypes.c_void_p(P.data_ptr()),
                                   ctypes.c_void_p(P.data_ptr()))
    assert rc == -5 and b"work lists" in E.load().pk_last_error()
    # packed batch beyond the scratch limit: refused before any launch
    mols = [water_cluster(s, n_min=16, n_max=16) for s in range(12)]
    shells = [Molecule(Zc, Xc, "6-31g*").shells() for Zc, Xc in mols]
    batch = FockEngine(shells, "os", 1e-12)
    assert batch.packed_size > 1 << 20
    rc = _lib().pk_fock_jk_noalloc(batch.ctx, ctypes.c_void_p(P.data_ptr()), ctypes.c_void_p(P.data_ptr()),
                                   ctypes.c_void_p(P.data_ptr()))
    assert rc == -1 and b"scratch limit" in E.load().pk_last_error()
