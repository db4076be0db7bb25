"""Host layer (lib/libpkscf.so) on the CPU: the product's own basis loader
against the oracle, the C ABI's exported symbols, and the device-only entry
points refusing to run without a device.

Bars: basis data to 1e-13 (same formulas up to rounding); nuclear repulsion
to 1e-12.  The one-electron integrals run on the device only
(tests/test_gpu_one_electron.py).
"""
import ctypes
import os
import re

import numpy as np
import pytest

import refcpu
from pkfock import engine as eng_mod
from pkfock import mlp as mlp_mod
from pkfock import host
from pkfock.molecules import ANGSTROM_TO_BOHR, alkane, water, water_cluster

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CASES = [("sto-3g", water()), ("6-31g*", water()), ("cc-pvdz", alkane(2)), ("cc-pvtz", water()),
         ("6-31g*", water_cluster(3, n=3))]


@pytest.mark.parametrize("basis,mol", CASES, ids=[c[0] + "-%d" % len(c[1][0]) for c in CASES])
def test_basis_matches_oracle(basis, mol):
    Z, X = mol
    m = host.Molecule(Z, X, basis)
    r = refcpu.Ref(basis, Z, X * ANGSTROM_TO_BOHR)
    assert m.nbasis == r.nbasis
    sh = m.shells()
    rc, rl, rn, re_, rco, rno = r.shells()
    assert np.array_equal(sh.l, rl) and np.array_equal(sh.nprim, rn)
    assert np.abs(sh.center - rc).max() < 1e-15
    assert np.abs(sh.exponent - re_).max() == 0.0
    assert np.abs(sh.coef - rco).max() <= 1e-13 * np.abs(rco).max()
    assert np.abs(sh.prim_norm - rno).max() <= 1e-13 * np.abs(rno).max()
    assert abs(m.nuclear_repulsion() - r.enuc()) < 1e-12


def test_one_electron_refuses_without_device():
    """The one-electron integrals run on the device only (no CPU fallback)."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a HIP device is present")
    except ImportError:
        pass
    m = host.Molecule(*water(), "sto-3g")
    with pytest.raises(eng_mod.PkError, match="no HIP device"):
        m.one_electron()


def test_basis_errors_are_reported():
    with pytest.raises(eng_mod.PkError, match="cannot open basis file"):
        host.Molecule([1, 1], np.zeros((2, 3)) + [[0, 0, 0], [0, 0, 0.74]], "no-such-basis")
    with pytest.raises(eng_mod.PkError, match="Unsupported Element"):
        host.Molecule([30, 1], [[0, 0, 0], [0, 0, 1.5]], "sto-3g")  # Zn not in the shipped files


def test_unit_constant_is_the_reference_one():
    # ANGSTROM_TO_BOHR is not 1/BOHR_TO_ANGSTROM in the reference (src/base/tables.h:7-8)
    assert ANGSTROM_TO_BOHR == 1.8897259886
    assert abs(1.0 / 0.529177210903 - ANGSTROM_TO_BOHR) > 1e-8


def _declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b((?:pk|pkref)_[a-z0-9_]+)\s*\(", src)))


@pytest.mark.parametrize("header,lib", [("pk_fock.h", eng_mod.LIB_PATH), ("pk_scf.h", host.LIB_PATH),
                                        ("pk_mlp.h", mlp_mod.LIB_PATH)])
def test_c_abi_exports_every_declared_symbol(header, lib):
    so = ctypes.CDLL(lib)
    names = _declared(header)
    assert len(names) >= 8
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_network_refuses_without_device():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a HIP device is present")
    except ImportError:
        pass
    with pytest.raises(mlp_mod.PkMlpError, match="no HIP device"):
        mlp_mod.Net(4, 2, 8)


def test_engine_refuses_without_device():
    """pk_ctx_create must fail loudly (no CPU fallback) when no HIP device is visible."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a HIP device is present")
    except ImportError:
        pass
    m = host.Molecule(*water(), "sto-3g")
    with pytest.raises(eng_mod.PkError, match="no HIP device"):
        m.engine()


def test_molecule_generators():
    Z, X = alkane(20)
    assert len(Z) == 62 and Z.count(6) == 20
    d = np.linalg.norm(X[:, None] - X[None], axis=-1)
    cc = d[:20, :20][np.triu_indices(20, 1)]
    assert abs(np.sort(cc)[:19] - 1.54).max() < 1e-12
    Zc, Xc = water_cluster(7)
    Zc2, Xc2 = water_cluster(7)
    assert Zc == Zc2 and np.array_equal(Xc, Xc2)
    O = Xc[::3]
    if len(O) > 1:
        dO = np.linalg.norm(O[:, None] - O[None], axis=-1)
        assert dO[np.triu_indices(len(O), 1)].min() >= 2.6
