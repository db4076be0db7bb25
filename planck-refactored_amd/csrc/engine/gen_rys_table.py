#!/usr/bin/env python3
"""Generate the Rys quadrature tables used by the HIP ERI kernels.

For n = 1..NMAX roots the n-point rule of the measure e^{-T u} / (2 sqrt(u)) du on
[0, 1] (u = t^2) satisfies  sum_r w_r u_r^k = F_k(T),  k < 2n,  F_k the Boys
function.  For T < TMAX(n) the roots u_r(T) and weights w_r(T) are tabulated as
Chebyshev expansions on intervals of width H; beyond TMAX(n) the rule is the
large-T limit u_r = x_r^2 / T, w_r = W_r / sqrt(T) with x_r, W_r the positive
nodes / weights of the 2n-point Gauss-Hermite rule.

Exact rules are computed in 60-digit arithmetic (mpmath): Boys moments from the
incomplete gamma function and downward recursion, the Jacobi matrix from the
Cholesky factor of the Hankel moment matrix (Golub & Welsch 1969), and its
eigen-decomposition.  The script checks every fit at the interval midpoints
between Chebyshev nodes and refuses to write a table whose worst relative error
exceeds 2e-14.

Output: rys_table.inc (C header, hexadecimal literals).  Run:
    python gen_rys_table.py            # writes rys_table.inc next to this file
"""
import multiprocessing as mp
import os
import sys

import mpmath

NMAX = 7        # up to (ff|ff): L = 12 -> n = 7
H = 1.0         # interval width in T
DEG = 11        # Chebyshev degree (DEG + 1 coefficients per function)
DPS = 60
# T beyond which the Gauss-Hermite limit is exact to double precision, per n
TMAX = {1: 40.0, 2: 50.0, 3: 55.0, 4: 60.0, 5: 65.0, 6: 75.0, 7: 80.0}


def boys_moments(nm, T):
    """F_0..F_{nm-1}(T) to DPS digits."""
    T = mpmath.mpf(T)
    top = nm - 1
    if T == 0:
        return [mpmath.mpf(1) / (2 * k + 1) for k in range(nm)]
    F = [mpmath.mpf(0)] * nm
    F[top] = mpmath.gammainc(top + mpmath.mpf(1) / 2, 0, T) / (2 * T ** (top + mpmath.mpf(1) / 2))
    eT = mpmath.exp(-T)
    for k in range(top, 0, -1):
        F[k - 1] = (2 * T * F[k] + eT) / (2 * k - 1)
    return F


def rys_rule(n, T):
    """(roots u_r ascending, weights w_r) as mpf lists."""
    mu = boys_moments(2 * n + 1, T)
    Hk = mpmath.matrix(n + 1, n + 1)
    for i in range(n + 1):
        for j in range(n + 1):
            Hk[i, j] = mu[i + j]
    R = mpmath.cholesky(Hk).T  # upper triangular, H = R^T R
    alpha, beta = [], []
    for j in range(n):
        a = R[j, j + 1] / R[j, j]
        if j > 0:
            a -= R[j - 1, j] / R[j - 1, j - 1]
        alpha.append(a)
        if j < n - 1:
            beta.append(R[j + 1, j + 1] / R[j, j])
    Jm = mpmath.matrix(n, n)
    for i in range(n):
        Jm[i, i] = alpha[i]
        if i < n - 1:
            Jm[i, i + 1] = Jm[i + 1, i] = beta[i]
    E, Q = mpmath.eigsy(Jm)
    pairs = sorted((E[i], mu[0] * Q[0, i] ** 2) for i in range(n))
    return [p[0] for p in pairs], [p[1] for p in pairs]


def hermite_half(n):
    """Positive nodes x_r (ascending) and weights of the 2n-point Gauss-Hermite rule."""
    m = 2 * n
    Jm = mpmath.matrix(m, m)
    for i in range(m - 1):
        Jm[i, i + 1] = Jm[i + 1, i] = mpmath.sqrt(mpmath.mpf(i + 1) / 2)
    E, Q = mpmath.eigsy(Jm)
    pts = sorted((E[i], mpmath.sqrt(mpmath.pi) * Q[0, i] ** 2) for i in range(m))
    pos = [p for p in pts if p[0] > 0]
    return [p[0] for p in pos], [p[1] for p in pos]


def cheb_nodes(D):
    return [mpmath.cos(mpmath.pi * (k + mpmath.mpf(1) / 2) / (D + 1)) for k in range(D + 1)]


def fit_interval(args):
    n, i = args
    mpmath.mp.dps = DPS
    T0 = mpmath.mpf(i) * H
    xs = cheb_nodes(DEG)
    vals = []
    for x in xs:
        T = T0 + (x + 1) * H / 2
        u, w = rys_rule(n, T)
        vals.append(u + w)
    nf = 2 * n
    coefs = []
    for f in range(nf):
        c = []
        for j in range(DEG + 1):
            s = mpmath.fsum(vals[k][f] * mpmath.cos(mpmath.pi * j * (k + mpmath.mpf(1) / 2) / (DEG + 1))
                            for k in range(DEG + 1))
            c.append(s * (1 if j == 0 else 2) / (DEG + 1))
        coefs.append(c)
    # check at points between the nodes
    worst = 0.0
    for k in range(DEG + 2):
        x = mpmath.cos(mpmath.pi * mpmath.mpf(k) / (DEG + 1))
        T = T0 + (x + 1) * H / 2
        u, w = rys_rule(n, T)
        ref = u + w
        for f in range(nf):
            c = coefs[f]
            b1 = b2 = mpmath.mpf(0)
            for j in range(DEG, 0, -1):
                b1, b2 = 2 * x * b1 - b2 + c[j], b1
            v = x * b1 - b2 + c[0]
            worst = max(worst, float(abs(v - ref[f]) / abs(ref[f])))
    return n, i, [[float(v) for v in c] for c in coefs], worst


def asym_check(n):
    mpmath.mp.dps = DPS
    x, W = hermite_half(n)
    T = mpmath.mpf(TMAX[n])
    u, w = rys_rule(n, T)
    err = 0.0
    for r in range(n):
        err = max(err, float(abs(x[r] ** 2 / T - u[r]) / u[r]), float(abs(W[r] / mpmath.sqrt(T) - w[r]) / w[r]))
    return n, [float(v ** 2) for v in x], [float(v) for v in W], err


def main():
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rys_table.inc")
    jobs = [(n, i) for n in range(1, NMAX + 1) for i in range(int(round(TMAX[n] / H)))]
    with mp.Pool(min(8, os.cpu_count() or 1)) as pool:
        fits = pool.map(fit_interval, jobs, chunksize=4)
        asym = pool.map(asym_check, range(1, NMAX + 1))
    worst = max(f[3] for f in fits)
    worst_asym = max(a[3] for a in asym)
    print("worst fit rel err %.3e, worst large-T limit rel err %.3e" % (worst, worst_asym))
    if worst > 2e-14 or worst_asym > 2e-14:
        sys.exit("table not accurate enough; adjust H / DEG / TMAX")
    fits.sort(key=lambda f: (f[0], f[1]))
    lines = ["// Generated by gen_rys_table.py -- do not edit.",
             "// Rys roots u_r(T) and weights w_r(T): Chebyshev fits of degree %d on T intervals of"
             " width %g, worst relative error %.2e; Gauss-Hermite limit beyond TMAX(n), worst"
             " relative error %.2e." % (DEG, H, worst, worst_asym),
             "#pragma once",
             "#define PK_RYS_NMAX %d" % NMAX,
             "#define PK_RYS_DEG %d" % DEG,
             "#define PK_RYS_INV_H %r" % (1.0 / H)]
    base, off = [], 0
    for n in range(1, NMAX + 1):
        base.append(off)
        off += int(round(TMAX[n] / H)) * 2 * n * (DEG + 1)
    lines.append("static constexpr double pk_rys_tmax[%d] = {0.0, %s};" % (NMAX + 1, ", ".join(repr(TMAX[n]) for n in range(1, NMAX + 1))))
    lines.append("static constexpr int pk_rys_base[%d] = {0, %s};" % (NMAX + 1, ", ".join(str(b) for b in base)))
    ax, aw = [], []
    abase = []
    for n, x2, W, _ in sorted(asym):
        abase.append(len(ax))
        ax += x2
        aw += W
    lines.append("static constexpr int pk_rys_abase[%d] = {0, %s};" % (NMAX + 1, ", ".join(str(b) for b in abase)))
    lines.append("static constexpr double pk_rys_ax2[%d] = {%s};" % (len(ax), ", ".join(v.hex() for v in ax)))
    lines.append("static constexpr double pk_rys_aw[%d] = {%s};" % (len(aw), ", ".join(v.hex() for v in aw)))
    lines.append("// layout: [n][interval][u_0..u_{n-1}, w_0..w_{n-1}][c_0..c_DEG]")
    lines.append("#ifndef PK_RYS_COEF_QUALIFIER")
    lines.append("#define PK_RYS_COEF_QUALIFIER static const")
    lines.append("#endif")
    lines.append("PK_RYS_COEF_QUALIFIER double pk_rys_coef[%d] = {" % off)
    for n, i, coefs, _ in fits:
        for c in coefs:
            lines.append("  " + ", ".join(v.hex() for v in c) + ",")
    lines.append("};")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("wrote", out, off, "coefficients")


if __name__ == "__main__":
    main()
