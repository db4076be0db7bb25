"""ctypes binding of the CPU oracle (oracle/libpk_ref.so) for the tests.

The oracle is test infrastructure: a plain-C++ restatement of the reference's
FP64 Fock-build path (see oracle/pk_ref.cpp for the reference file:line map).
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "libpk_ref.so")
BASIS_DIR = os.path.join(ROOT, "planck-refactored_amd", "basis-sets")
ENGINE = {"os": 0, "rys": 1, "auto": 2}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(LIB)
        vp, ip, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.pkref_create.restype = vp
        L.pkref_create.argtypes = [ctypes.c_char_p, ip, vp, vp, ip, d]
        L.pkref_destroy.argtypes = [vp]
        L.pkref_last_error.restype = ctypes.c_char_p
        L.pkref_nbasis.argtypes = [vp]
        L.pkref_nshells.argtypes = [vp]
        L.pkref_shell_info.argtypes = [vp, ip, vp, vp, vp, vp, vp, vp, vp]
        L.pkref_nuclear_repulsion.restype = d
        L.pkref_nuclear_repulsion.argtypes = [vp]
        for f in ("pkref_overlap", "pkref_kinetic", "pkref_nuclear", "pkref_schwarz"):
            getattr(L, f).argtypes = [vp, vp]
        L.pkref_eri.restype = d
        L.pkref_eri.argtypes = [vp, ip, ip, ip, ip]
        L.pkref_eri_engine.restype = d
        L.pkref_eri_engine.argtypes = [vp, ip, ip, ip, ip, ip]
        L.pkref_eri_full.argtypes = [vp, vp]
        L.pkref_fock_jk.argtypes = [vp, vp, vp, vp, ip]
        L.pkref_fock_jk_sample.argtypes = [vp, vp, vp, vp, ip, ctypes.c_long, ctypes.c_long, vp]
        L.pkref_jk_elements.argtypes = [vp, vp, ip, vp, vp, vp, vp, ip]
        L.pkref_rhf.argtypes = [vp, ip, d, d, ip, ip, vp, vp, vp, vp, ip]
        L.pkref_boys.restype = d
        L.pkref_boys.argtypes = [ip, d]
        L.pkref_set_boys_exact.argtypes = [ip]
        L.pkref_rys_rule.restype = ip
        L.pkref_rys_rule.argtypes = [ip, d, vp, vp]
        L.pkref_auto_prefers_rys.argtypes = [vp, vp, vp, vp]
        L.pkref_syev.argtypes = [ip, vp, vp, vp]
        L.pkref_eri_deriv.argtypes = [vp, ip, ip, ip, ip, vp]
        L.pkref_eri_gradient.argtypes = [vp, vp, vp, vp, d, vp, ip]
        L.pkref_screen_primitives.argtypes = [vp, d, ip]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def basis_file(name):
    n = name.lower().replace("*", "s")
    return os.path.join(BASIS_DIR, n)


class Ref:
    """Oracle context for one molecule (coordinates in Bohr)."""

    def __init__(self, basis, Z, xyz_bohr, engine="os", tol_eri=1e-10):
        self.Z = np.ascontiguousarray(Z, dtype=np.int32)
        self.xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1, 3)
        self.ctx = lib().pkref_create(basis_file(basis).encode(), len(self.Z), _p(self.Z), _p(self.xyz),
                                      ENGINE[engine], float(tol_eri))
        if not self.ctx:
            raise RuntimeError(lib().pkref_last_error().decode())
        self.nbasis = lib().pkref_nbasis(self.ctx)

    def __del__(self):
        try:
            if self.ctx:
                lib().pkref_destroy(self.ctx)
        except Exception:
            pass

    def shells(self):
        """(center, l, nprim, exponent, coef, prim_norm) arrays of the oracle's basis."""
        L = lib()
        ns = L.pkref_nshells(self.ctx)
        cen, ls, nps, ex, co, no = [], [], [], [], [], []
        for s in range(ns):
            l, npr, at = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            c = np.zeros(3)
            e = np.zeros(64)
            cf = np.zeros(64)
            nm = np.zeros(64)
            L.pkref_shell_info(self.ctx, s, ctypes.byref(l), ctypes.byref(npr), ctypes.byref(at), _p(c), _p(e),
                               _p(cf), _p(nm))
            k = npr.value
            cen.append(c)
            ls.append(l.value)
            nps.append(k)
            ex += list(e[:k])
            co += list(cf[:k])
            no += list(nm[:k])
        return np.array(cen), np.array(ls), np.array(nps), np.array(ex), np.array(co), np.array(no)

    def matrix(self, which):
        M = np.zeros((self.nbasis, self.nbasis))
        getattr(lib(), "pkref_" + which)(self.ctx, _p(M))
        return M

    def enuc(self):
        return lib().pkref_nuclear_repulsion(self.ctx)

    def eri(self, i, j, k, l, engine=None):
        if engine is None:
            return lib().pkref_eri(self.ctx, i, j, k, l)
        return lib().pkref_eri_engine(self.ctx, ENGINE[engine], i, j, k, l)

    def eri_full(self):
        n = self.nbasis
        out = np.zeros((n, n, n, n))
        lib().pkref_eri_full(self.ctx, _p(out))
        return out

    def jk(self, P, nthreads=0):
        P = np.ascontiguousarray(P, dtype=np.float64)
        J = np.zeros_like(P)
        K = np.zeros_like(P)
        lib().pkref_fock_jk(self.ctx, _p(P), _p(J), _p(K), nthreads)
        return J, K

    def jk_sample(self, P, stride, offset=0, nthreads=0):
        P = np.ascontiguousarray(P, dtype=np.float64)
        J = np.zeros_like(P)
        K = np.zeros_like(P)
        nq = ctypes.c_double()
        lib().pkref_fock_jk_sample(self.ctx, _p(P), _p(J), _p(K), nthreads, stride, offset, ctypes.byref(nq))
        return J, K, nq.value

    def jk_elements(self, P, rows, cols, nthreads=0):
        """J[rows[e], cols[e]] and K[rows[e], cols[e]] of density P."""
        P = np.ascontiguousarray(P, dtype=np.float64)
        r = np.ascontiguousarray(rows, dtype=np.int32)
        c = np.ascontiguousarray(cols, dtype=np.int32)
        J = np.zeros(len(r))
        K = np.zeros(len(r))
        if lib().pkref_jk_elements(self.ctx, _p(P), len(r), _p(r), _p(c), _p(J), _p(K), nthreads) != 0:
            raise ValueError(lib().pkref_last_error().decode())
        return J, K

    def screen_primitives(self, cut, rule=1):
        """Drop primitive pairs as the engine does (rule 0: prefactor bound,
        rule 1: primitive Schwarz bound); returns the number dropped."""
        return lib().pkref_screen_primitives(self.ctx, float(cut), int(rule))

    def eri_deriv(self, i, j, k, l):
        """d(ij|kl)/dX_q for the centres of i, j, k, l (i <= j, k <= l): a 4 x 3 array."""
        out = np.zeros(12)
        if lib().pkref_eri_deriv(self.ctx, i, j, k, l, _p(out)) != 0:
            raise ValueError(lib().pkref_last_error().decode())
        return out.reshape(4, 3)

    def eri_gradient(self, Pj, Pk1=None, Pk2=None, kx=1.0, nthreads=0):
        """Two-electron gradient (natoms x 3): 1/4 sum Gamma d(ij|kl)/dR with
        Gamma = 2 Pj_ij Pj_kl - kx (Pk1_ik Pk1_jl + Pk2_ik Pk2_jl)."""
        Pj = np.ascontiguousarray(Pj, dtype=np.float64)
        Pk1 = Pj if Pk1 is None else np.ascontiguousarray(Pk1, dtype=np.float64)
        Pk2 = None if Pk2 is None else np.ascontiguousarray(Pk2, dtype=np.float64)
        g = np.zeros((len(self.Z), 3))
        rc = lib().pkref_eri_gradient(self.ctx, _p(Pj), _p(Pk1), None if Pk2 is None else _p(Pk2), float(kx),
                                      _p(g), nthreads)
        if rc != 0:
            raise ValueError(lib().pkref_last_error().decode())
        return g

    def rhf(self, charge=0, tol_energy=1e-10, tol_density=1e-10, diis_dim=8, max_cycles=0):
        E, it = ctypes.c_double(), ctypes.c_int()
        n = self.nbasis
        P = np.zeros((n, n))
        hist = np.zeros(1000)
        rc = lib().pkref_rhf(self.ctx, charge, tol_energy, tol_density, diis_dim, max_cycles, ctypes.byref(E),
                             ctypes.byref(it), _p(P), _p(hist), len(hist))
        return {"rc": rc, "energy": E.value, "iterations": it.value, "P": P, "history": hist[:it.value]}


class exact_boys:
    """Context manager: the oracle's Boys function switches from the
    reference's table lookup (~1e-12 absolute) to an accurate series
    (~1e-17 relative) inside the block."""

    def __enter__(self):
        lib().pkref_set_boys_exact(1)
        return self

    def __exit__(self, *exc):
        lib().pkref_set_boys_exact(0)
        return False


def boys(n, x):
    return lib().pkref_boys(n, x)


def rys_rule(n, T):
    r = np.zeros(max(n, 1))
    w = np.zeros(max(n, 1))
    if lib().pkref_rys_rule(n, T, _p(r), _p(w)) != 0:
        raise ValueError(lib().pkref_last_error().decode())
    return r, w
