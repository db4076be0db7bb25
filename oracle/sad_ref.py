"""CPU oracle of the SAD initial density (test infrastructure only).

numpy restatement of the reference's compute_sad_guess_rhf
(src/scf/sad.cpp:589-643) on top of the oracle's integrals (oracle/pk_ref.cpp
through a context factory):

* per element, an atomic UHF at the origin (sad.cpp:284-356): multiplicity 1
  for even Z, 2 for odd Z (sad.cpp:217-220); core-Hamiltonian guess per spin
  with the Loewdin orthogonaliser (scf.cpp:499-505); F_s = H + J[Pa + Pb] - K[P_s];
  DIIS per spin on X^T (F P S - S P F) X, at most 6 vectors, extrapolation from
  2 on (types.h:759-837), subspace restart when the error grows by more than 2x
  after iteration 2 (scf.cpp:636-644); converged when iteration > 1 and
  |dE|, rms and max of the change of Pa + Pb are below 1e-10 (scf.cpp:89-135);
* spherical average (sad.cpp:367-430): shell populations N_l = sum over the
  shell of (P S)_mumu, cross-shell blocks zeroed, each shell block
  (N_l / g) pinv(S_shell, 1e-10);
* block-diagonal assembly in molecular AO order (sad.cpp:436-470);
* projection on an RHF density (sad.cpp:476-505): X = S^-1/2 thresholded at
  1e-8, the n_occ eigenvectors of X^T P X with the largest eigenvalues,
  P = 2 C C^T with C = X U.

Pinned in tests/test_oracle.py against the reference's own log of ethylene /
STO-3G with the SAD guess (tests/benchmarks/scf/engine_symmetry/
ethylene_cas44_sto3g_sa2_os_nosymm_sad.log): with the diagonal-block shell
populations the log was written with, one rotation inside the doubly
degenerate natural-occupation pair at n_occ reproduces the printed iteration-1
energy, rms and max density change and DIIS error together.  The projected
density is otherwise defined only up to that rotation (eigensolver-dependent in
the reference too); the raw averaged density is unique.  Also pinned: the
textbook STO-3G hydrogen-atom energy, trace and idempotency invariants.
"""
import numpy as np


def _inv_sqrt(S, thresh):
    w, U = np.linalg.eigh(S)
    inv = np.where(w > thresh, 1.0 / np.sqrt(np.maximum(w, thresh)), 0.0)
    return (U * inv) @ U.T


def _pinv_sym(A, thresh):
    w, U = np.linalg.eigh(A)
    inv = np.where(w > thresh, 1.0 / np.where(w > thresh, w, 1.0), 0.0)
    return (U * inv) @ U.T


class _DIIS:
    def __init__(self, m):
        self.m, self.F, self.E = m, [], []

    def push(self, F, e):
        self.F.append(F)
        self.E.append(e)
        if len(self.F) > self.m:
            self.F.pop(0)
            self.E.pop(0)

    def clear(self):
        self.F, self.E = [], []

    def extrapolate(self):
        m = len(self.F)
        B = np.zeros((m + 1, m + 1))
        for i in range(m):
            for j in range(i, m):
                B[i, j] = B[j, i] = (self.E[i] * self.E[j]).sum()
            B[i, m] = B[m, i] = -1.0
        rhs = np.zeros(m + 1)
        rhs[m] = -1.0
        c = np.linalg.lstsq(B, rhs, rcond=None)[0]
        return sum(ci * Fi for ci, Fi in zip(c[:m], self.F))


def atomic_uhf(ctx, Z, tol=1e-10, max_iter=200, diis_dim=6):
    """ctx: oracle context of one atom at the origin.  Returns (Pa, Pb, S)."""
    S, H = ctx.matrix("overlap"), ctx.matrix("kinetic") + ctx.matrix("nuclear")
    n_unp = 0 if Z % 2 == 0 else 1
    na, nb = (Z + n_unp) // 2, (Z - n_unp) // 2
    X = _inv_sqrt(S, 0.0)

    def dens(F, nocc):
        _, V = np.linalg.eigh(X.T @ F @ X)
        C = X @ V
        return C[:, :nocc] @ C[:, :nocc].T

    Pa, Pb = dens(H, na), dens(H, nb)
    da, db = _DIIS(diis_dim), _DIIS(diis_dim)
    E_prev, err_prev = 0.0, np.inf
    for it in range(1, max_iter + 1):
        Ja, Ka = ctx.jk(Pa)
        Jb, Kb = ctx.jk(Pb)
        Fa, Fb = H + Ja + Jb - Ka, H + Ja + Jb - Kb
        E = 0.5 * ((Pa * (H + Fa)).sum() + (Pb * (H + Fb)).sum())
        ea = X.T @ (Fa @ Pa @ S - S @ Pa @ Fa) @ X
        eb = X.T @ (Fb @ Pb @ S - S @ Pb @ Fb) @ X
        rms = lambda m: np.sqrt((m * m).sum() / m.size)  # noqa: E731
        cur = max(rms(ea), rms(eb))
        if it > 2 and cur > 2.0 * err_prev:
            da.clear()
            db.clear()
        da.push(Fa, ea)
        db.push(Fb, eb)
        err_prev = cur
        Fa_d = da.extrapolate() if len(da.F) >= 2 else Fa
        Fb_d = db.extrapolate() if len(db.F) >= 2 else Fb
        Pa_n, Pb_n = dens(Fa_d, na), dens(Fb_d, nb)
        d = (Pa_n + Pb_n) - (Pa + Pb)
        Pa, Pb = Pa_n, Pb_n
        dE, E_prev = abs(E - E_prev), E
        if it > 1 and dE < tol and np.sqrt((d * d).mean()) < tol and np.abs(d).max() < tol:
            return Pa, Pb, S
    raise RuntimeError("atomic UHF for Z=%d did not converge" % Z)


def spherical_average(P, S, shell_slices, thresh=1e-10, populations="mulliken"):
    """populations "mulliken": N_l = sum over the shell's rows of (P S)_mumu
    (sad.cpp:393-407, the current source); "block": N_l = Tr(P_ll S_ll), the
    shell's diagonal block only (the formula the reference's SAD logs were
    produced with, before the Mulliken partition)."""
    PS = P @ S
    out = np.zeros_like(P)
    for sl in shell_slices:
        g = sl.stop - sl.start
        N = np.trace(PS[sl, sl]) if populations == "mulliken" else np.trace(P[sl, sl] @ S[sl, sl])
        out[sl, sl] = (N / g) * _pinv_sym(S[sl, sl], thresh)
    return 0.5 * (out + out.T)


def sad_raw(Z, ao_atom, make_atom_ctx, populations="mulliken"):
    """Block-diagonal spherically averaged atomic densities (before the
    projection; unique, unlike the projection when natural occupations are
    degenerate at n_occ)."""
    cache = {}
    for z in sorted(set(Z)):
        ctx, shells = make_atom_ctx(z)
        Pa, Pb, S = atomic_uhf(ctx, z)
        cache[z] = spherical_average(Pa + Pb, S, shells, populations=populations)
    n = len(ao_atom)
    P = np.zeros((n, n))
    ao_atom = np.asarray(ao_atom)
    for a, z in enumerate(Z):
        idx = np.nonzero(ao_atom == a)[0]
        P[np.ix_(idx, idx)] = cache[z]
    return 0.5 * (P + P.T)


def project(P, S_mol, n_el):
    """Natural-orbital projection on an RHF density (sad.cpp:476-505); also
    returns the eigenpairs of X^T P X for callers that resolve degeneracy."""
    X = _inv_sqrt(S_mol, 1e-8)
    Pbar = X.T @ P @ X
    Pbar = 0.5 * (Pbar + Pbar.T)
    w, U = np.linalg.eigh(Pbar)
    C = X @ U[:, -(n_el // 2):]
    P = 2.0 * C @ C.T
    return 0.5 * (P + P.T), w, U, X


def sad_density(Z, ao_atom, make_atom_ctx, S_mol, charge=0, populations="mulliken"):
    """Z: atomic numbers; ao_atom: atom index of every molecular AO (basis
    order); make_atom_ctx(z) -> (oracle context of atom z at the origin, list of
    AO slices of its shells).  Returns the projected RHF SAD density."""
    P = sad_raw(Z, ao_atom, make_atom_ctx, populations)
    return project(P, S_mol, int(sum(Z)) - charge)[0]
