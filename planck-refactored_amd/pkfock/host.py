"""ctypes binding of the host layer (lib/libpkscf.so, include/pk_scf.h):
basis construction, one-electron integrals and the device RHF SCF driver.

Mirrors the reference's host path: read_gbs_basis (src/basis/gaussian.cpp:183),
_compute_1e / _compute_nuclear_attraction (src/integrals/base.h:10-27) and
SCF::run_rhf (src/scf/scf.cpp:175-441).
"""
import ctypes
import os

import numpy as np

from . import engine as _eng
from .molecules import ANGSTROM_TO_BOHR

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libpkscf.so")
BASIS_DIR = os.path.join(os.path.dirname(_HERE), "basis-sets")


class ScfOptions(ctypes.Structure):
    _fields_ = [("tol_energy", ctypes.c_double), ("tol_density", ctypes.c_double), ("tol_eri", ctypes.c_double),
                ("diis_dim", ctypes.c_int32), ("max_cycles", ctypes.c_int32), ("engine", ctypes.c_int32),
                ("charge", ctypes.c_int32), ("use_guess", ctypes.c_int32),
                ("guess_density", ctypes.POINTER(ctypes.c_double)), ("precision", ctypes.c_int32),
                ("fp32_cutoff", ctypes.c_double), ("guess_fallback", ctypes.c_int32), ("sad_guess", ctypes.c_int32),
                ("guess_only", ctypes.c_int32)]


class ScfResult(ctypes.Structure):
    _fields_ = [("energy", ctypes.c_double), ("nuclear_repulsion", ctypes.c_double), ("iterations", ctypes.c_int32),
                ("converged", ctypes.c_int32), ("first_diis_error", ctypes.c_double), ("wall_s", ctypes.c_double),
                ("fock_ms", ctypes.c_double), ("guess_used", ctypes.c_int32), ("guess_diis_error", ctypes.c_double),
                ("core_diis_error", ctypes.c_double)]


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise _eng.PkError("host library not built: %s is missing (run __graft_entry__.build())" % LIB_PATH)
    _eng.load()
    lib = ctypes.CDLL(LIB_PATH)
    vp, ip = ctypes.c_void_p, ctypes.c_int
    lib.pk_basis_load.argtypes = [ctypes.c_char_p, ip, vp, vp, ctypes.POINTER(vp)]
    lib.pk_basis_free.argtypes = [vp]
    lib.pk_basis_nbasis.argtypes = [vp]
    lib.pk_basis_describe.argtypes = [vp, ctypes.POINTER(_eng.BasisDesc)]
    lib.pk_one_electron.argtypes = [vp, vp, vp, vp]
    lib.pk_nuclear_repulsion.argtypes = [vp]
    lib.pk_nuclear_repulsion.restype = ctypes.c_double
    lib.pk_scf_default_options.argtypes = [ctypes.POINTER(ScfOptions)]
    lib.pk_scf_rhf.argtypes = [vp, ctypes.POINTER(ScfOptions), ctypes.POINTER(ScfResult), vp, vp, ip]
    lib.pk_scf_rhf_batch.argtypes = [vp, ip, ctypes.POINTER(ScfOptions), vp, vp, vp]
    lib.pk_host_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def basis_path(name):
    return os.path.join(BASIS_DIR, name.lower().replace("*", "s"))


class Molecule:
    """Atoms Z and coordinates (Angstrom in, Bohr kept) with a basis set."""

    def __init__(self, Z, xyz_angstrom, basis, charge=0, bohr=False):
        lib = load()
        self.Z = np.ascontiguousarray(Z, dtype=np.int32)
        X = np.asarray(xyz_angstrom, dtype=np.float64).reshape(-1, 3)
        self.xyz = np.ascontiguousarray(X if bohr else X * ANGSTROM_TO_BOHR)
        self.charge = charge
        self.basis_name = basis
        self.handle = ctypes.c_void_p()
        rc = lib.pk_basis_load(basis_path(basis).encode(), len(self.Z), self.Z.ctypes.data, self.xyz.ctypes.data,
                               ctypes.byref(self.handle))
        if rc != 0:
            raise _eng.PkError("pk_basis_load: " + lib.pk_host_last_error().decode())
        self.nbasis = lib.pk_basis_nbasis(self.handle)

    def __del__(self):
        try:
            if self.handle:
                load().pk_basis_free(self.handle)
        except Exception:
            pass

    def shells(self):
        d = _eng.BasisDesc()
        load().pk_basis_describe(self.handle, ctypes.byref(d))
        ns = d.nshells
        npr = np.ctypeslib.as_array(d.nprim, shape=(ns,)).copy()
        tot = int(npr.sum())
        return _eng.Shells(np.ctypeslib.as_array(d.center, shape=(3 * ns,)).copy(),
                           np.ctypeslib.as_array(d.l, shape=(ns,)).copy(), npr,
                           np.ctypeslib.as_array(d.exponent, shape=(tot,)).copy(),
                           np.ctypeslib.as_array(d.coef, shape=(tot,)).copy(),
                           np.ctypeslib.as_array(d.prim_norm, shape=(tot,)).copy())

    def one_electron(self):
        """S, T, V (computed on the device; raises PkError without one)."""
        n = self.nbasis
        S, T, V = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
        lib = load()
        if lib.pk_one_electron(self.handle, S.ctypes.data, T.ctypes.data, V.ctypes.data) != 0:
            raise _eng.PkError("pk_one_electron: " + lib.pk_host_last_error().decode())
        return S, T, V

    def nuclear_repulsion(self):
        return load().pk_nuclear_repulsion(self.handle)

    def engine(self, engine="os", tol_eri=1e-10):
        return _eng.FockEngine(self.shells(), engine=engine, tol_eri=tol_eri)

    @staticmethod
    def rhf_batch(mols, engine="os", tol_eri=1e-10, tol_energy=1e-10, tol_density=1e-10, diis_dim=8, max_cycles=0,
                  guesses=None, fp32_cutoff=None, guess_fallback=False):
        """Batched RHF of independent molecules (pk_scf_rhf_batch): one batched
        Fock build per iteration for all of them.  guess_fallback: each
        molecule given a guess starts from the core guess instead when that
        has the smaller first DIIS error.  Returns a list of result dicts
        (with "P")."""
        lib = load()
        o = ScfOptions()
        lib.pk_scf_default_options(ctypes.byref(o))
        o.tol_energy, o.tol_density, o.tol_eri = tol_energy, tol_density, tol_eri
        o.diis_dim, o.max_cycles = diis_dim, max_cycles
        o.engine = _eng.ENGINE[engine] if isinstance(engine, str) else int(engine)
        o.charge = 0
        if fp32_cutoff is not None:
            o.precision, o.fp32_cutoff = _eng.PREC_MIXED, float(fp32_cutoff)
        o.guess_fallback = int(bool(guess_fallback))
        nm = len(mols)
        res = (ScfResult * nm)()
        Ps = [np.zeros((m.nbasis, m.nbasis)) for m in mols]
        pout = (ctypes.c_void_p * nm)(*[p.ctypes.data for p in Ps])
        keep = None
        gin = None
        if guesses is not None:
            keep = [None if g is None else np.ascontiguousarray(g, dtype=np.float64) for g in guesses]
            gin = (ctypes.c_void_p * nm)(*[None if g is None else g.ctypes.data for g in keep])
        bases = (ctypes.c_void_p * nm)(*[m.handle.value for m in mols])
        rc = lib.pk_scf_rhf_batch(bases, nm, ctypes.byref(o), res, pout, gin)
        if rc < 0:
            raise _eng.PkError("pk_scf_rhf_batch: " + lib.pk_host_last_error().decode())
        out = []
        for i in range(nm):
            d = {k: getattr(res[i], k) for k, _ in ScfResult._fields_}
            d["P"] = Ps[i]
            d["rc"] = 0 if res[i].converged else 1
            out.append(d)
        return out

    def rhf(self, engine="os", tol_eri=1e-10, tol_energy=1e-10, tol_density=1e-10, diis_dim=8, max_cycles=0,
            guess=None, fp32_cutoff=None, guess_fallback=False, sad=False, guess_only=False):
        """Device RHF (pk_scf_rhf).  fp32_cutoff: mixed precision, quartets
        with Schwarz bound below it in FP32 (None: FP64 throughout)."""
        lib = load()
        o = ScfOptions()
        lib.pk_scf_default_options(ctypes.byref(o))
        o.tol_energy, o.tol_density, o.tol_eri = tol_energy, tol_density, tol_eri
        if fp32_cutoff is not None:
            o.precision, o.fp32_cutoff = _eng.PREC_MIXED, float(fp32_cutoff)
        o.sad_guess, o.guess_only = int(bool(sad)), int(bool(guess_only))
        o.diis_dim, o.max_cycles = diis_dim, max_cycles
        o.engine = _eng.ENGINE[engine] if isinstance(engine, str) else int(engine)
        o.charge = self.charge
        g = None
        if guess is not None:
            g = np.ascontiguousarray(guess, dtype=np.float64)
            o.use_guess = 1
            o.guess_density = g.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            o.guess_fallback = int(bool(guess_fallback))
        r = ScfResult()
        n = self.nbasis
        P = np.zeros((n, n))
        hist = np.zeros(1000)
        rc = lib.pk_scf_rhf(self.handle, ctypes.byref(o), ctypes.byref(r), P.ctypes.data, hist.ctypes.data, len(hist))
        if rc < 0:
            raise _eng.PkError("pk_scf_rhf: " + lib.pk_host_last_error().decode())
        out = {k: getattr(r, k) for k, _ in ScfResult._fields_}
        out["P"] = P
        out["history"] = hist[:r.iterations].copy()
        out["rc"] = rc
        return out
