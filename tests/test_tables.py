"""The engine's generated tables, evaluated on the CPU exactly as the kernels do.

Rys table (rys_table.inc): for n = 1..7 the rule reproduces the Boys moments,
sum_r w_r u_r^k = F_k(T) for k < 2n, at seeded T across every interval and the
large-T limit; bar 1e-13 relative (the kernels' FP64 ERIs need ~1e-14).
Boys table (boys_table.inc): the kernels' 8-term Taylor + downward recursion
(and the large-T upward recursion) against 40-digit values; bar 4e-15 relative.
"""
import math
import os
import re

import mpmath
import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENG = os.path.join(ROOT, "planck-refactored_amd", "csrc", "engine")
mpmath.mp.dps = 40


def _arr(src, name):
    body = src.split(name, 1)[1].split("{", 1)[1].split("}", 1)[0]
    return [float.fromhex(t) if "0x" in t else float(t) for t in re.findall(r"-?0x[0-9a-fA-Fp.+-]+|-?\d+\.\d+|\b\d+\b", body)]


@pytest.fixture(scope="module")
def rys():
    src = open(os.path.join(ENG, "rys_table.inc")).read()
    deg = int(re.search(r"#define PK_RYS_DEG (\d+)", src).group(1))
    return {"deg": deg, "tmax": _arr(src, "pk_rys_tmax"), "base": [int(v) for v in _arr(src, "pk_rys_base")],
            "abase": [int(v) for v in _arr(src, "pk_rys_abase")], "ax2": _arr(src, "pk_rys_ax2"),
            "aw": _arr(src, "pk_rys_aw"), "coef": np.array(_arr(src, "pk_rys_coef["))}


def rule(t, n, T):
    if T >= t["tmax"][n]:
        b = t["abase"][n]
        return ([t["ax2"][b + r] / T for r in range(n)], [t["aw"][b + r] / math.sqrt(T) for r in range(n)])
    i = int(T)
    x = 2.0 * (T - i) - 1.0
    D = t["deg"] + 1
    c = t["coef"][t["base"][n] + i * 2 * n * D:]
    vals = []
    for f in range(2 * n):
        cf = c[f * D:(f + 1) * D]
        b1 = b2 = 0.0
        for j in range(D - 1, 0, -1):
            b1, b2 = 2 * x * b1 - b2 + cf[j], b1
        vals.append(x * b1 - b2 + cf[0])
    return vals[:n], vals[n:]


def F(k, T):
    if T == 0:
        return mpmath.mpf(1) / (2 * k + 1)
    T = mpmath.mpf(T)
    return mpmath.gammainc(k + mpmath.mpf(1) / 2, 0, T) / (2 * T ** (k + mpmath.mpf(1) / 2))


@pytest.mark.parametrize("n", range(1, 8))
def test_rys_rule_moments(rys, n):
    rng = np.random.default_rng(100 + n)
    Ts = list(rng.uniform(0, rys["tmax"][n] + 5, 25)) + [0.0, 1e-12, rys["tmax"][n] - 1e-9, rys["tmax"][n], 500.0]
    worst = 0.0
    for T in Ts:
        u, w = rule(rys, n, float(T))
        u = [mpmath.mpf(v) for v in u]
        w = [mpmath.mpf(v) for v in w]
        for k in range(2 * n):
            ex = F(k, T)
            worst = max(worst, float(abs(sum(w[r] * u[r] ** k for r in range(n)) - ex) / ex))
    assert worst < 1e-13, worst


def test_boys_device_scheme():
    src = open(os.path.join(ENG, "boys_table.inc")).read()
    cols = int(re.search(r"#define PK_BOYS_COLS (\d+)", src).group(1))
    tab = np.array(_arr(src, "pk_boys_table[")).reshape(-1, cols)
    etab = np.array(_arr(src, "pk_boys_exp["))

    xmax = int(re.search(r"#define PK_BOYS_XMAX (\d+)", src).group(1))

    def boys(M, T):
        eT = math.exp(-T)
        if T < xmax:
            i = int(T * 16 + 0.5)
            d = T - i / 16.0
            row = tab[i, M:M + 8]
            s = row[7]
            for k in range(7, 0, -1):
                s = s * (-d / k) + row[k - 1]
            e = 0.0
            for k in range(8, -1, -1):
                e = e * (-d) + 1.0 / math.factorial(k)
            eT = etab[i] * e
            out = [0.0] * (M + 1)
            out[M] = s
            for m in range(M, 0, -1):
                out[m - 1] = (2 * T * out[m] + eT) / (2 * m - 1)
            return out
        # large T: the kernels drop e^{-T} for M <= 4
        if M <= 4:
            eT = 0.0
        out = [0.88622692545275801365 * math.sqrt(1.0 / T)]
        for m in range(1, M + 1):
            out.append(((2 * m - 1) * out[-1] - eT) * (0.5 / T))
        return out

    rng = np.random.default_rng(9)
    worst = 0.0
    for T in list(rng.uniform(0, 55, 160)) + [0.0, 1e-13, 35.9999, 36.0, 49.99999, 50.0, 50.5, 80.0, 1e4]:
        for M in (0, 1, 2, 3, 4, 8, 12):
            got = boys(M, float(T))
            for m in range(M + 1):
                ex = F(m, T)
                worst = max(worst, float(abs(got[m] - ex) / ex))
    assert worst < 4e-15, worst
