"""Pinning the CPU oracle (oracle/pk_ref.cpp) before it is trusted.

What pins each part (none of these re-types the oracle's own formula):
  Boys lookup ....... golden vectors of the reference's compiled Lookup::boys
                      (tests/golden/boys_ref.json) and 40-digit mpmath values
  primitive ERIs .... closed forms (ss|ss) and (p s|ss) evaluated with mpmath
  contracted ERIs ... permutation symmetry, translation invariance, rotation
                      covariance, the Schwarz inequality, and two independent
                      algorithms (Obara-Saika and Rys) agreeing
  J/K digestion ..... equal to the dense contraction of the oracle's own nb^4 tensor
  basis + 1e + SCF .. the reference's RHF energy pins (tests/golden/scf_kat.json)
  auto routing ...... the per-class Rys fractions of the reference cost model
  Rys rule .......... moment conditions sum_r w_r u_r^k = F_k(T), k < 2n (n <= 5;
                      the reference's moment method degrades from n = 7 on)
  ERI derivatives ... central finite differences (Richardson-extrapolated) of the
  and 2e gradient     oracle's own two-electron energy 1/2 Tr[Pj J] - kx/4 sum Tr[Pk K]
                      at displaced geometries, fixed densities, accurate Boys;
                      translation invariance of every derivative quartet
"""
import json
import math
import os
import sys
import tempfile

import mpmath
import numpy as np
import pytest

import refcpu
from pkfock.molecules import ANGSTROM_TO_BOHR, alkane, water

HERE = os.path.dirname(os.path.abspath(__file__))
mpmath.mp.dps = 40


def F_exact(n, x):
    if x == 0:
        return mpmath.mpf(1) / (2 * n + 1)
    x = mpmath.mpf(x)
    return mpmath.gammainc(n + mpmath.mpf(1) / 2, 0, x) / (2 * x ** (n + mpmath.mpf(1) / 2))


def test_boys_matches_reference_lookup():
    g = json.load(open(os.path.join(HERE, "golden", "boys_ref.json")))
    worst = 0.0
    for row in g["rows"]:
        x = row[0]
        for n in range(g["orders"]):
            worst = max(worst, abs(refcpu.boys(n, x) - row[1 + n]))
    # the reference table holds 62 entries (of 103686) that are not the
    # correctly rounded 15-decimal value; the oracle's table is correctly rounded
    assert worst <= 1.01e-15, worst


def test_boys_against_mpmath():
    rng = np.random.default_rng(5)
    worst = 0.0
    for x in list(rng.uniform(0, 157, 60)) + [0.0, 0.05, 1e-9, 200.0, 1000.0]:
        for n in (0, 1, 3, 8, 12):
            worst = max(worst, abs(refcpu.boys(n, float(x)) - float(F_exact(n, x))))
    assert worst < 1e-12, worst  # table + 6-term Taylor accuracy of the reference lookup


def test_exact_boys_mode_against_mpmath():
    """The oracle's accurate Boys mode (used to pin device integrals past the
    lookup's error) against 40-digit mpmath, relative 1e-15."""
    rng = np.random.default_rng(6)
    worst = 0.0
    with refcpu.exact_boys():
        for x in list(rng.uniform(0, 160, 80)) + [0.0, 1e-9, 0.05, 30.0, 157.05, 157.2, 400.0]:
            for n in (0, 1, 2, 4, 7, 12, 20):
                ex = float(F_exact(n, x))
                worst = max(worst, abs(refcpu.boys(n, float(x)) - ex) / ex)
    assert refcpu.boys(0, 0.5) != 0.0
    assert worst < 1e-15, worst


def _custom_basis(shells):
    """Temporary Gaussian94 file: element -> list of (label, exponent)."""
    f = tempfile.NamedTemporaryFile("w", suffix=".gbs", delete=False)
    for el, sh in shells.items():
        f.write("%s     0\n" % el)
        for lab, ex in sh:
            f.write("%s   1   1.00\n      %.10E           1.0000000\n" % (lab, ex))
        f.write("****\n")
    f.close()
    return f.name


def _norm_s(a):
    return (2 * a / mpmath.pi) ** mpmath.mpf(0.75)


def test_primitive_ss_and_ps_closed_forms():
    ex = {"H": 1.3, "He": 0.45, "Li": 2.1, "Be": 0.8}
    path = _custom_basis({"H": [("S", ex["H"])], "He": [("S", ex["He"])], "Li": [("P", ex["Li"])],
                          "Be": [("S", ex["Be"])]})
    X = np.array([[0.0, 0.1, -0.2], [1.1, 0.3, 0.4], [-0.5, 1.2, 0.7], [0.9, -0.8, 1.5]])  # Bohr
    try:
        r = refcpu.Ref.__new__(refcpu.Ref)
        r.Z = np.array([1, 2, 3, 4], dtype=np.int32)
        r.xyz = np.ascontiguousarray(X)
        lib = refcpu.lib()
        r.ctx = lib.pkref_create(path.encode(), 4, r.Z.ctypes.data, r.xyz.ctypes.data, 0, 0.0)
        assert r.ctx, lib.pkref_last_error()
        r.nbasis = lib.pkref_nbasis(r.ctx)
        assert r.nbasis == 6  # s, s, px py pz, s
        A, B, C, D = [mpmath.matrix(list(map(mpmath.mpf, X[i]))) for i in range(4)]
        a, b, c, d = ex["H"], ex["He"], ex["Be"], ex["Be"]

        def ss(a, A, b, B, c, C, d, D, order=0):
            p, q = a + b, c + d
            P = (a * A + b * B) / p
            Q = (c * C + d * D) / q
            AB2 = sum((A[i] - B[i]) ** 2 for i in range(3))
            CD2 = sum((C[i] - D[i]) ** 2 for i in range(3))
            PQ2 = sum((P[i] - Q[i]) ** 2 for i in range(3))
            rho = p * q / (p + q)
            pre = 2 * mpmath.pi ** 2.5 / (p * q * mpmath.sqrt(p + q)) * mpmath.exp(-a * b / p * AB2 - c * d / q * CD2)
            return pre, P, Q, rho, F_exact(order, rho * PQ2)

        # (s_H s_He | s_Be s_Be), functions 0, 1, 5, 5
        pre, P, Q, rho, F0 = ss(a, A, b, B, c, D, d, D)
        want = pre * F0 * _norm_s(a) * _norm_s(b) * _norm_s(c) * _norm_s(d)
        got = r.eri(0, 1, 5, 5)
        assert abs(got - float(want)) < 1e-12 * max(1.0, abs(got)), (got, want)
        # (p_x s_H | s_Be s_Be): PA_x [ss|ss]^0 + WP_x [ss|ss]^1, functions 2, 0, 5, 5
        al = ex["Li"]
        pre, P, Q, rho, F0 = ss(al, C, a, A, c, D, d, D)
        F1 = F_exact(1, rho * sum((P[i] - Q[i]) ** 2 for i in range(3)))
        p_ = al + a
        W = (p_ * P + (c + d) * Q) / (p_ + c + d)
        Np = _norm_s(al) * 2 * mpmath.sqrt(al)
        want = pre * ((P[0] - C[0]) * F0 + (W[0] - P[0]) * F1) * Np * _norm_s(a) * _norm_s(c) * _norm_s(d)
        got = r.eri(2, 0, 5, 5)
        assert abs(got - float(want)) < 1e-12, (got, want)
    finally:
        os.unlink(path)


@pytest.fixture(scope="module")
def ethylene():
    Z = [6, 6, 1, 1, 1, 1]
    X = np.array([[-3.438338, 0.906062, 0.0], [-3.438338, 2.230667, -0.000001],
                  [-2.836048, 0.343683, -0.709301], [-4.040627, 0.343684, 0.709301],
                  [-2.836049, 2.793045, -0.709302], [-4.040627, 2.793046, 0.709300]])
    return Z, X


def test_eri_symmetries_and_engines(ethylene):
    Z, X = ethylene
    ref = refcpu.Ref("cc-pvdz", Z, X * ANGSTROM_TO_BOHR, engine="os")
    moved = refcpu.Ref("cc-pvdz", Z, (X + np.array([0.7, -1.9, 3.3])) * ANGSTROM_TO_BOHR, engine="os")
    Q = ref.matrix("schwarz")
    rng = np.random.default_rng(11)
    n = ref.nbasis
    for _ in range(40):
        i, j, k, l = (int(v) for v in rng.integers(0, n, 4))
        v = ref.eri(i, j, k, l)
        for perm in ((j, i, k, l), (i, j, l, k), (k, l, i, j), (l, k, j, i)):
            assert abs(ref.eri(*perm) - v) < 1e-14
        assert abs(moved.eri(i, j, k, l) - v) < 1e-11
        assert abs(v) <= Q[i, j] * Q[k, l] * (1 + 1e-12) + 1e-14
        assert abs(ref.eri(i, j, k, l, engine="rys") - v) < 1e-11  # L <= 8 here: n <= 5 roots


def test_rotation_covariance():
    Z, X = water()
    ref = refcpu.Ref("6-31g*", Z, X * ANGSTROM_TO_BOHR)
    R = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])  # x -> y, y -> -x
    rot = refcpu.Ref("6-31g*", Z, (X @ R.T) * ANGSTROM_TO_BOHR)
    # O functions: 0 s, 1 s, 2 px 3 py 4 pz, 5 s, 6 px 7 py 8 pz, 9.. d
    for (i, j, k, l), (ri, rj, rk, rl), sign in (((2, 0, 2, 0), (3, 0, 3, 0), 1.0), ((2, 6, 3, 7), (3, 7, 2, 6), 1.0),
                                                 ((3, 0, 2, 0), (2, 0, 3, 0), -1.0),
                                                 ((4, 4, 2, 2), (4, 4, 3, 3), 1.0)):
        assert abs(rot.eri(ri, rj, rk, rl) - sign * ref.eri(i, j, k, l)) < 1e-12


def test_direct_jk_equals_tensor_contraction():
    Z, X = water()
    ref = refcpu.Ref("6-31g*", Z, X * ANGSTROM_TO_BOHR, tol_eri=1e-12)
    n = ref.nbasis
    T = ref.eri_full()
    rng = np.random.default_rng(2)
    A = rng.normal(size=(n, n))
    P = A + A.T
    J, K = ref.jk(P)
    Jd = np.einsum("ijkl,kl->ij", T, P)
    Kd = np.einsum("ikjl,kl->ij", T, P)
    assert np.abs(J - Jd).max() < 1e-12
    assert np.abs(K - Kd).max() < 1e-12


def test_syev_against_numpy():
    rng = np.random.default_rng(4)
    for n in (1, 2, 7, 40):
        A = rng.normal(size=(n, n))
        A = A + A.T
        w = np.zeros(n)
        V = np.zeros((n, n))
        refcpu.lib().pkref_syev(n, A.ctypes.data, w.ctypes.data, V.ctypes.data)
        assert np.allclose(w, np.linalg.eigvalsh(A), atol=1e-12)
        assert np.allclose(A @ V, V * w, atol=1e-11)


def test_rys_rule_moments():
    for n in (2, 3, 4, 5):
        for T in (1e-4, 0.7, 3.0, 11.0, 27.5, 60.0):
            r, w = refcpu.rys_rule(n, T)
            for k in range(2 * n):
                ex = float(F_exact(k, T))
                assert abs(float(np.sum(w * r ** k)) - ex) <= 1e-11 * ex, (n, T, k)


def test_rys_rule_refused_beyond_its_pins():
    """The reference's moment method loses accuracy from n = 6 roots: the
    oracle refuses it there and its Rys engine falls back to OS, so a (dd|dd)
    ERI (n = 5) stays Rys and an (fd|dd) component (n = 6) equals OS."""
    with pytest.raises(ValueError):
        refcpu.rys_rule(6, 1.0)
    with pytest.raises(ValueError):
        refcpu.rys_rule(0, 1.0)
    Z, X = [6, 8], np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 2.1316]])
    ros = refcpu.Ref("cc-pvtz", Z, X, engine="os")
    rrys = refcpu.Ref("cc-pvtz", Z, X, engine="rys")
    n = ros.nbasis
    # the last functions of each atom's list are its f shells (file order)
    f_c = [i for i in range(n) if i < 35][-10:]
    for (i, j, k, l) in [(f_c[0], f_c[3], f_c[5], f_c[9]), (f_c[1], 0, f_c[2], 1)]:
        a, b = ros.eri(i, j, k, l), rrys.eri(i, j, k, l)
        assert abs(a - b) <= 1e-13 * max(1.0, abs(a)), (i, j, k, l, a, b)


def test_auto_routing_fractions():
    """Per-class share of Cartesian components the reference's cost model sends
    to Rys (orientation la >= lb, lc >= ld)."""
    import ctypes

    def comps(l):
        return [(x, y, l - x - y) for x in range(l, -1, -1) for y in range(l - x, -1, -1)]

    def frac(la, lb, lc, ld):
        tot = rys = 0
        for a in comps(la):
            for b in comps(lb):
                for c in comps(lc):
                    for d in comps(ld):
                        arr = [(ctypes.c_int * 3)(*v) for v in (a, b, c, d)]
                        rys += refcpu.lib().pkref_auto_prefers_rys(*arr)
                        tot += 1
        return rys / tot

    assert frac(1, 1, 1, 1) == 0.0  # L <= 4 -> Obara-Saika everywhere
    assert abs(frac(2, 1, 1, 1) - 0.61) < 0.01
    assert abs(frac(2, 2, 1, 0) - 0.33) < 0.01
    assert abs(frac(2, 1, 2, 1) - 0.75) < 0.01
    assert abs(frac(2, 2, 2, 2) - 0.993) < 0.002


KAT = json.load(open(os.path.join(HERE, "golden", "scf_kat.json")))["cases"]


@pytest.mark.parametrize("case", KAT, ids=[c["id"] for c in KAT])
def test_oracle_rhf_reference_energies(case):
    Z = [a[0] for a in case["atoms"]]
    X = np.array([a[1:] for a in case["atoms"]])
    ref = refcpu.Ref(case["basis"], Z, X * ANGSTROM_TO_BOHR, engine=case["engine"], tol_eri=1e-10)
    r = ref.rhf(max_cycles=case.get("max_cycles", 0))
    # D2d ethylene has a doubly degenerate HOMO holding two electrons: the SCF
    # settles on the reference energy but the density may keep rotating inside
    # the degenerate pair, so only the energy is pinned there
    assert r["rc"] == 0 or case.get("degenerate_homo")
    assert abs(r["energy"] - case["energy"]) < 1e-9, (r["energy"], case["energy"])
    if "iterations" in case:
        assert r["iterations"] == case["iterations"]


# ---------------------------------------------------------------- SAD guess
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import sad_ref  # noqa: E402


def _shell_slices(ref):
    cen, ls, _, _, _, _ = ref.shells()
    out, f = [], 0
    for l in ls:
        n = (l + 1) * (l + 2) // 2
        out.append(slice(f, f + n))
        f += n
    return cen, out


def sad_parts(basis, Z, X_bohr, populations="mulliken", engine="os"):
    ref = refcpu.Ref(basis, Z, X_bohr, engine=engine, tol_eri=1e-12)
    cen, slices = _shell_slices(ref)
    ao_atom = np.zeros(ref.nbasis, int)
    for c, sl in zip(cen, slices):
        ao_atom[sl] = int(np.argmin(np.linalg.norm(X_bohr - c, axis=1)))

    def make_atom(z):
        r = refcpu.Ref(basis, [z], np.zeros((1, 3)), engine=engine, tol_eri=1e-12)
        return r, _shell_slices(r)[1]

    return ref, sad_ref.sad_raw(list(Z), ao_atom, make_atom, populations)


def _first_iteration(ref, P, nocc):
    """Energy, rms / max change of the next density and DIIS error of the first
    RHF iteration from P (scf.cpp:281-360)."""
    S = ref.matrix("overlap")
    H = ref.matrix("kinetic") + ref.matrix("nuclear")
    J, K = ref.jk(P)
    F = H + J - 0.5 * K
    E1 = 0.5 * (P * (H + F)).sum() + ref.enuc()
    w, U = np.linalg.eigh(S)
    Xo = (U / np.sqrt(w)) @ U.T
    e = Xo @ (F @ P @ S - S @ P @ F) @ Xo
    _, V = np.linalg.eigh(Xo @ F @ Xo)
    C = Xo @ V[:, :nocc]
    d = 2 * C @ C.T - P
    return E1, np.sqrt((d * d).mean()), np.abs(d).max(), np.sqrt((e * e).mean())


def test_sad_oracle_reproduces_reference_first_iteration():
    """Ethylene / STO-3G, SAD guess, no symmetry: the reference's log prints the
    first SCF iteration from the SAD density: E -65.3840017, rms(D) 1.098,
    max(D) 4.285, DIIS 5.955e-01.  The natural occupations of the raw SAD
    density are doubly degenerate exactly at n_occ, so the projected density
    is fixed only up to a rotation inside that pair (the reference's result is
    whatever its eigensolver returned); and the log was written with the
    diagonal-block shell populations.  The pin: some rotation angle of the
    pair reproduces all four printed numbers at once."""
    Z = [6, 6, 1, 1, 1, 1]
    X = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.34], [0.92, 0.0, -0.54], [-0.92, 0.0, -0.54],
                  [0.0, 0.92, 1.88], [0.0, -0.92, 1.88]]) * ANGSTROM_TO_BOHR
    ref, Praw = sad_parts("sto-3g", Z, X, populations="block")
    S = ref.matrix("overlap")
    Pp, w, U, Xs = sad_ref.project(Praw, S, 16)
    assert abs(np.trace(Pp @ S) - 16.0) < 1e-10
    assert abs(w[6] - w[5]) < 1e-10 and w[7] - w[6] > 0.1  # degenerate pair at the boundary

    def at(t):
        v = np.cos(t) * U[:, 5] + np.sin(t) * U[:, 6]
        C = Xs @ np.column_stack([v] + [U[:, k] for k in range(7, 14)])
        return _first_iteration(ref, 2 * C @ C.T, 8)

    ts = np.linspace(0.0, np.pi, 721)
    E = np.array([at(t)[0] for t in ts])
    i = int(np.argmin(np.abs(E + 65.3840017)))
    lo, hi = ts[max(i - 1, 0)], ts[min(i + 1, len(ts) - 1)]
    for _ in range(60):  # bisection on E(t) + 65.3840017 around the crossing
        mid = 0.5 * (lo + hi)
        if (at(lo)[0] + 65.3840017) * (at(mid)[0] + 65.3840017) <= 0:
            hi = mid
        else:
            lo = mid
    E1, rms, dmax, err = at(0.5 * (lo + hi))
    print("theta %.4f: E1 %.8f rms(D) %.4f max(D) %.4f DIIS %.5f" % (0.5 * (lo + hi), E1, rms, dmax, err))
    assert abs(E1 + 65.3840017) < 1e-7
    assert abs(rms - 1.098) < 5e-4 and abs(dmax - 4.285) < 5e-4 and abs(err - 0.5955) < 5e-5


def test_sad_oracle_invariants():
    """Raw SAD: trace = electron count, block-diagonal per atom and shell;
    projected: idempotent in the S metric with trace N (sad.cpp:476-505); H atom
    UHF in STO-3G gives the textbook -0.46658185 Eh."""
    Z, X = water()
    Xb = X * ANGSTROM_TO_BOHR
    ref, Praw = sad_parts("6-31g*", Z, Xb)
    S = ref.matrix("overlap")
    assert abs(np.trace(Praw @ S) - 10.0) < 1e-9
    P = sad_ref.project(Praw, S, 10)[0]
    assert abs(np.trace(P @ S) - 10.0) < 1e-9
    assert np.abs(P @ S @ P - 2.0 * P).max() < 1e-9
    h = refcpu.Ref("sto-3g", [1], np.zeros((1, 3)), engine="os", tol_eri=1e-12)
    Pa, Pb, Sh = sad_ref.atomic_uhf(h, 1)
    Hh = h.matrix("kinetic") + h.matrix("nuclear")
    assert abs((Pa * Hh).sum() - (-0.46658185)) < 1e-8


def _e2(basis, Z, X, Pj, Pks, kx):
    r = refcpu.Ref(basis, Z, X, engine="os", tol_eri=0.0)
    J, _ = r.jk(Pj)
    e = 0.5 * np.sum(Pj * J)
    for Pk in Pks:
        _, K = r.jk(Pk)
        e -= 0.25 * kx * np.sum(Pk * K)
    return e


@pytest.mark.parametrize("basis", ["6-31g*", "cc-pvdz"])
def test_eri_gradient_against_finite_differences(basis):
    """The oracle's 2e gradient (gradient.cpp:209-232 with os.cpp:1296-1420)
    is the derivative of the oracle's own 2e energy at fixed densities: RHF
    form (Pj = Pk = P, kx = 1) and UHF form (Pj = Pa + Pb, Pk = Pa, Pb, kx = 2).
    A dropped 2 alpha weight, a wrong sign of the lowering term, a missed
    permutation or a wrong centre all move it by O(1)."""
    Z, X = water()
    X = X * ANGSTROM_TO_BOHR
    ref = refcpu.Ref(basis, Z, X, engine="os", tol_eri=0.0)
    n = ref.nbasis
    rng = np.random.default_rng(1)
    A = rng.normal(size=(n, n)) * 0.2
    Pa = A + A.T
    B = rng.normal(size=(n, n)) * 0.2
    Pb = B + B.T
    with refcpu.exact_boys():
        g_rhf = ref.eri_gradient(Pa, Pa, None, 1.0)
        g_uhf = ref.eri_gradient(Pa + Pb, Pa, Pb, 2.0)
        for g, Pj, Pks, kx in ((g_rhf, Pa, [Pa], 1.0), (g_uhf, Pa + Pb, [Pa, Pb], 2.0)):
            def d(a, q, h):
                Xp, Xm = X.copy(), X.copy()
                Xp[a, q] += h
                Xm[a, q] -= h
                return (_e2(basis, Z, Xp, Pj, Pks, kx) - _e2(basis, Z, Xm, Pj, Pks, kx)) / (2 * h)
            fd = np.array([[(4 * d(a, q, 5e-4) - d(a, q, 1e-3)) / 3 for q in range(3)] for a in range(len(Z))])
            err = np.abs(g - fd).max()
            print("%s kx=%g: max|g - fd| = %.2e (|g| up to %.2f)" % (basis, kx, err, np.abs(g).max()))
            assert err < 2e-9 * max(1.0, np.abs(g).max())
            assert np.abs(g.sum(axis=0)).max() < 1e-11  # translation invariance


def test_eri_derivative_translation_invariance():
    """sum over the four centres of d(ij|kl)/dX_q vanishes for every quartet
    (the integral depends on differences of centres only); d/dB of (ij|kl) is
    d/dA of (ji|kl) up to the centre relabelling."""
    Z, X = water()
    ref = refcpu.Ref("6-31g*", Z, X * ANGSTROM_TO_BOHR, engine="os", tol_eri=0.0)
    n = ref.nbasis
    rng = np.random.default_rng(2)
    worst = 0.0
    for _ in range(200):
        i, j, k, l = rng.integers(0, n, 4)
        i, j = min(i, j), max(i, j)
        k, l = min(k, l), max(k, l)
        D = ref.eri_deriv(int(i), int(j), int(k), int(l))
        worst = max(worst, np.abs(D.sum(axis=0)).max() / max(1e-3, np.abs(D).max()))
        # bra-ket swap: d(kl|ij) on the centres of k, l, i, j
        D2 = ref.eri_deriv(int(k), int(l), int(i), int(j))
        assert np.allclose(D2[[2, 3, 0, 1]], D, rtol=1e-12, atol=1e-14)
    assert worst < 1e-10
