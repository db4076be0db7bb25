"""ctypes binding of the HIP Fock-build engine (lib/libpkfock.so, include/pk_fock.h).

The engine is the product: there is no CPU fallback.  Loading fails loudly when
the shared library is missing, and every call raises PkError on a non-zero
status with the engine's own message.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libpkfock.so")

ENGINE = {"os": 0, "obara-saika": 0, "rys": 1, "rys-quadrature": 1, "auto": 2}
PREC_FP64, PREC_MIXED = 0, 1
MEM_HOST, MEM_DEVICE = 0, 1


class PkError(RuntimeError):
    pass


class BasisDesc(ctypes.Structure):
    _fields_ = [("nshells", ctypes.c_int32),
                ("center", ctypes.POINTER(ctypes.c_double)),
                ("l", ctypes.POINTER(ctypes.c_int32)),
                ("nprim", ctypes.POINTER(ctypes.c_int32)),
                ("exponent", ctypes.POINTER(ctypes.c_double)),
                ("coef", ctypes.POINTER(ctypes.c_double)),
                ("prim_norm", ctypes.POINTER(ctypes.c_double))]


class Stats(ctypes.Structure):
    _fields_ = [("nbasis", ctypes.c_int64), ("nshells", ctypes.c_int64), ("shell_pairs", ctypes.c_int64),
                ("quartets", ctypes.c_int64), ("primitive_quartets", ctypes.c_int64),
                ("model_flops", ctypes.c_double), ("last_build_ms", ctypes.c_double),
                ("last_eri_kernel_ms", ctypes.c_double), ("kernel_launches", ctypes.c_int64)]


class ClassStats(ctypes.Structure):
    _fields_ = [("la", ctypes.c_int32), ("lb", ctypes.c_int32), ("lc", ctypes.c_int32), ("ld", ctypes.c_int32),
                ("kernels", ctypes.c_int32), ("quartets", ctypes.c_int64), ("primitive_quartets", ctypes.c_int64),
                ("model_flops", ctypes.c_double), ("ms_total", ctypes.c_double), ("builds", ctypes.c_int64),
                ("precision", ctypes.c_int32), ("block", ctypes.c_int32), ("remap", ctypes.c_int32),
                ("waves", ctypes.c_int64), ("alt", ctypes.c_int32), ("kind", ctypes.c_int32)]


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PkError("HIP engine not built: %s is missing (run __graft_entry__.build())" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int
    lib.pk_ctx_create.argtypes = [ctypes.POINTER(BasisDesc), ctypes.POINTER(vp)]
    lib.pk_ctx_destroy.argtypes = [vp]
    lib.pk_ctx_set.argtypes = [vp, ip, ctypes.c_double, ip, ctypes.c_double]
    lib.pk_ctx_set_stream.argtypes = [vp, vp]
    lib.pk_ctx_nbasis.argtypes = [vp]
    lib.pk_ctx_stats.argtypes = [vp, ctypes.POINTER(Stats)]
    lib.pk_fock_jk.argtypes = [vp, vp, vp, vp, ip]
    lib.pk_fock_2e.argtypes = [vp, vp, vp, ip]
    lib.pk_fock_jk_uhf.argtypes = [vp, vp, vp, vp, vp, vp, ip]
    lib.pk_schwarz.argtypes = [vp, vp, ip]
    lib.pk_eri_tensor.argtypes = [vp, vp]
    lib.pk_ctx_profile.argtypes = [vp, ip]
    lib.pk_ctx_class_count.argtypes = [vp]
    lib.pk_ctx_class_stats.argtypes = [vp, ip, ctypes.POINTER(ClassStats)]
    lib.pk_batch_create.argtypes = [ctypes.POINTER(BasisDesc), ip, ctypes.POINTER(vp)]
    lib.pk_ctx_nmol.argtypes = [vp]
    lib.pk_ctx_mol_layout.argtypes = [vp, vp, vp]
    lib.pk_ctx_set_graph.argtypes = [vp, ip]
    lib.pk_ctx_autotune.argtypes = [vp, ip]
    lib.pk_eri_gradient.argtypes = [vp, vp, vp, vp, ctypes.c_double, vp, ip]
    lib.pk_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def _check(rc, what):
    if rc != 0:
        msg = load().pk_last_error().decode(errors="replace")
        raise PkError("%s failed (%d): %s" % (what, rc, msg))


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Shells:
    """Contracted Cartesian shells in basis order (the pk_basis_desc arrays)."""

    def __init__(self, center, l, nprim, exponent, coef, prim_norm):
        self.center = np.ascontiguousarray(center, dtype=np.float64).reshape(-1, 3)
        self.l = np.ascontiguousarray(l, dtype=np.int32)
        self.nprim = np.ascontiguousarray(nprim, dtype=np.int32)
        self.exponent = np.ascontiguousarray(exponent, dtype=np.float64)
        self.coef = np.ascontiguousarray(coef, dtype=np.float64)
        self.prim_norm = np.ascontiguousarray(prim_norm, dtype=np.float64)

    @property
    def nbasis(self):
        return int(sum((l + 1) * (l + 2) // 2 for l in self.l))

    def desc(self):
        d = BasisDesc()
        d.nshells = len(self.l)
        dbl = ctypes.POINTER(ctypes.c_double)
        i32 = ctypes.POINTER(ctypes.c_int32)
        d.center = self.center.ctypes.data_as(dbl)
        d.l = self.l.ctypes.data_as(i32)
        d.nprim = self.nprim.ctypes.data_as(i32)
        d.exponent = self.exponent.ctypes.data_as(dbl)
        d.coef = self.coef.ctypes.data_as(dbl)
        d.prim_norm = self.prim_norm.ctypes.data_as(dbl)
        return d


def shells_to_atoms(gshell, centers, xyz_bohr):
    """Sum per-shell rows onto atoms: shell s belongs to the atom whose
    position equals its centre within 1e-5 Bohr (the reference's
    build_shell_atom_map, src/gradient/gradient.cpp:20-46)."""
    xyz = np.asarray(xyz_bohr, dtype=np.float64).reshape(-1, 3)
    out = np.zeros((len(xyz), gshell.shape[1]))
    for s, c in enumerate(np.asarray(centers).reshape(-1, 3)):
        d2 = ((xyz - c) ** 2).sum(axis=1)
        a = int(np.argmin(d2))
        if d2[a] >= 1e-10:
            raise PkError("shell %d does not match any atom" % s)
        out[a] += gshell[s]
    return out


class FockEngine:
    """One engine context per geometry (pk_ctx).  A list of Shells makes a
    batch context (pk_batch_create): matrices are then packed, molecule m's
    block at element offsets[m] with side nbs[m] (see pack / unpack)."""

    def __init__(self, shells, engine="os", tol_eri=1e-10):
        lib = load()
        self._keep = shells
        self.ctx = ctypes.c_void_p()
        if isinstance(shells, (list, tuple)):
            descs = (BasisDesc * len(shells))(*[s.desc() for s in shells])
            _check(lib.pk_batch_create(descs, len(shells), ctypes.byref(self.ctx)), "pk_batch_create")
        else:
            _check(lib.pk_ctx_create(ctypes.byref(shells.desc()), ctypes.byref(self.ctx)), "pk_ctx_create")
        self.nbasis = lib.pk_ctx_nbasis(self.ctx)
        self.nmol = lib.pk_ctx_nmol(self.ctx)
        nbs = np.zeros(self.nmol, dtype=np.int32)
        offs = np.zeros(self.nmol, dtype=np.int64)
        _check(lib.pk_ctx_mol_layout(self.ctx, _ptr(nbs), _ptr(offs)), "pk_ctx_mol_layout")
        self.nbs, self.offsets = nbs, offs
        self.packed_size = int(offs[-1] + nbs[-1] ** 2)
        self.set(engine, tol_eri)

    def pack(self, mats):
        """List of nb_m x nb_m matrices -> packed vector."""
        out = np.empty(self.packed_size)
        for m, a in enumerate(mats):
            n = int(self.nbs[m])
            out[self.offsets[m]:self.offsets[m] + n * n] = np.ascontiguousarray(a, dtype=np.float64).ravel()
        return out

    def unpack(self, vec):
        return [vec[o:o + int(n) * int(n)].reshape(int(n), int(n)) for o, n in zip(self.offsets, self.nbs)]

    def jk_batch(self, Ps):
        """J, K of every molecule of a batch context (lists of matrices)."""
        P = self.pack(Ps)
        J, K = np.zeros_like(P), np.zeros_like(P)
        _check(load().pk_fock_jk(self.ctx, _ptr(P), _ptr(J), _ptr(K), MEM_HOST), "pk_fock_jk")
        return self.unpack(J), self.unpack(K)

    def set_graph(self, enable=True):
        _check(load().pk_ctx_set_graph(self.ctx, int(bool(enable))), "pk_ctx_set_graph")

    def autotune(self, reps=2):
        _check(load().pk_ctx_autotune(self.ctx, int(reps)), "pk_ctx_autotune")

    def set(self, engine="os", tol_eri=1e-10, precision=PREC_FP64, fp32_cutoff=0.0):
        e = ENGINE[engine] if isinstance(engine, str) else int(engine)
        _check(load().pk_ctx_set(self.ctx, e, float(tol_eri), int(precision), float(fp32_cutoff)), "pk_ctx_set")

    def close(self):
        if self.ctx:
            load().pk_ctx_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stats(self):
        s = Stats()
        _check(load().pk_ctx_stats(self.ctx, ctypes.byref(s)), "pk_ctx_stats")
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def profile(self, enable=True):
        _check(load().pk_ctx_profile(self.ctx, int(bool(enable))), "pk_ctx_profile")

    def class_stats(self):
        out = []
        for i in range(load().pk_ctx_class_count(self.ctx)):
            s = ClassStats()
            _check(load().pk_ctx_class_stats(self.ctx, i, ctypes.byref(s)), "pk_ctx_class_stats")
            out.append({k: getattr(s, k) for k, _ in ClassStats._fields_})
        return out

    def schwarz(self):
        Q = np.zeros(self.packed_size)
        _check(load().pk_schwarz(self.ctx, _ptr(Q), MEM_HOST), "pk_schwarz")
        return Q.reshape(self.nbasis, self.nbasis) if self.nmol == 1 else self.unpack(Q)

    def eri_tensor(self):
        n = self.nbasis
        out = np.zeros((n, n, n, n))
        _check(load().pk_eri_tensor(self.ctx, _ptr(out)), "pk_eri_tensor")
        return out

    def jk(self, P):
        P = np.ascontiguousarray(P, dtype=np.float64)
        J = np.zeros_like(P)
        K = np.zeros_like(P)
        _check(load().pk_fock_jk(self.ctx, _ptr(P), _ptr(J), _ptr(K), MEM_HOST), "pk_fock_jk")
        return J, K

    def jk_uhf(self, Pa, Pb):
        """UHF: (J[Pa + Pb], K[Pa], K[Pb]) (pk_fock_jk_uhf, host arrays)."""
        Pa = np.ascontiguousarray(Pa, dtype=np.float64)
        Pb = np.ascontiguousarray(Pb, dtype=np.float64)
        Jt, Ka, Kb = np.zeros_like(Pa), np.zeros_like(Pa), np.zeros_like(Pa)
        _check(load().pk_fock_jk_uhf(self.ctx, _ptr(Pa), _ptr(Pb), _ptr(Jt), _ptr(Ka), _ptr(Kb), MEM_HOST),
               "pk_fock_jk_uhf")
        return Jt, Ka, Kb

    def jk_uhf_device(self, Pa_ptr, Pb_ptr, Jt_ptr, Ka_ptr, Kb_ptr):
        """Device pointers (ints) for Pa, Pb, Jt, Ka, Kb."""
        v = ctypes.c_void_p
        _check(load().pk_fock_jk_uhf(self.ctx, v(Pa_ptr), v(Pb_ptr), v(Jt_ptr), v(Ka_ptr), v(Kb_ptr), MEM_DEVICE),
               "pk_fock_jk_uhf")

    def eri_gradient(self, Pj, Pk1=None, Pk2=None, kx=1.0):
        """Two-electron gradient per shell centre (nshells x 3, host arrays):
        dE2/dR with E2 = 1/4 sum Gamma_ijkl (ij|kl),
        Gamma = 2 Pj_ij Pj_kl - kx (Pk1_ik Pk1_jl + Pk2_ik Pk2_jl) (pk_eri_gradient).
        RHF: eri_gradient(P); UHF: eri_gradient(Pa + Pb, Pa, Pb, 2.0)."""
        Pj = np.ascontiguousarray(Pj, dtype=np.float64)
        Pk1 = Pj if Pk1 is None else np.ascontiguousarray(Pk1, dtype=np.float64)
        Pk2 = None if Pk2 is None else np.ascontiguousarray(Pk2, dtype=np.float64)
        for M in (Pk1, Pk2):
            if M is not None and M.shape != Pj.shape:
                raise ValueError("eri_gradient: densities of different shapes")
        if self.nmol != 1:
            raise ValueError("eri_gradient: one molecule per context")
        if Pj.shape != (self.nbasis, self.nbasis):
            raise ValueError("eri_gradient: density must be nbasis x nbasis")
        ns = len(self._keep.l)
        g = np.zeros((ns, 3))
        _check(load().pk_eri_gradient(self.ctx, _ptr(Pj), _ptr(Pk1), None if Pk2 is None else _ptr(Pk2), float(kx),
                                      _ptr(g), MEM_HOST), "pk_eri_gradient")
        return g

    def eri_gradient_atoms(self, xyz_bohr, Pj, Pk1=None, Pk2=None, kx=1.0):
        """eri_gradient summed over the shells of each atom (natoms x 3): a
        shell belongs to the atom at its centre (within 1e-5 Bohr)."""
        return shells_to_atoms(self.eri_gradient(Pj, Pk1, Pk2, kx), self._keep.center, xyz_bohr)

    def g(self, P):
        P = np.ascontiguousarray(P, dtype=np.float64)
        G = np.zeros_like(P)
        _check(load().pk_fock_2e(self.ctx, _ptr(P), _ptr(G), MEM_HOST), "pk_fock_2e")
        return G

    def jk_device(self, P_ptr, J_ptr, K_ptr):
        """P, J, K are device pointers (ints), e.g. torch tensor .data_ptr()."""
        _check(load().pk_fock_jk(self.ctx, ctypes.c_void_p(P_ptr), ctypes.c_void_p(J_ptr),
                                 ctypes.c_void_p(K_ptr), MEM_DEVICE), "pk_fock_jk")

    def set_stream(self, stream_ptr):
        _check(load().pk_ctx_set_stream(self.ctx, ctypes.c_void_p(stream_ptr)), "pk_ctx_set_stream")
