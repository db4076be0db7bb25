// pk_ref.cpp — CPU oracle for Planck's FP64 direct-SCF Fock build.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, from the mathematics and the
// reference's documented semantics, the algorithm of HemanthHaridas/planck-refactored
// for the two-electron Fock-build path.  It contains no reference source.  Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
//
// Reference functions restated here (paths relative to the reference tree):
//   GBS reader + SP split ............ src/basis/gaussian.cpp:54-163, 183-302
//   Cartesian order, normalization ... src/basis/basis.cpp:20-75
//   primitive / shell pairs .......... src/base/types.h:606-675, src/integrals/shellpair.cpp:7-22
//   Boys table + Taylor lookup ....... src/lookup/boys.cpp:7-15, 1592-1618
//   OS ERI VRR / HRR / contraction ... src/integrals/os.cpp:618-1023, 415-459
//   Rys rule, VRR, HRR, contraction .. src/integrals/rys_roots.cpp:88-235, rys.cpp:250-505
//   auto routing cost model .......... src/integrals/rys.cpp:195-236
//   Schwarz table .................... src/integrals/os.cpp:1703-1752, rys.cpp:509-556
//   Fock G = J - K/2 ................. src/integrals/os.cpp:1435-1523 (here true direct, no nb^4)
//   1e integrals S, T, V ............. src/integrals/os.cpp:248-602
//   nuclear repulsion ................ src/base/types.h:934-964
//   RHF SCF + DIIS ................... src/scf/scf.cpp:15-45, 89-135, 175-441; types.h:759-837
//
// Parity status of each part is stated in the header of tests/test_oracle.py.

#include "pk_ref.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include <omp.h>
#include <quadmath.h>

namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;
thread_local std::string g_err;

// ---------------------------------------------------------------- elements
const char *kSymbols[] = {"X",  "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",
                          "Ne", "Na", "Mg", "Al", "Si", "P",  "S",  "Cl", "Ar", "K",
                          "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu",
                          "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr"};
constexpr int kMaxZ = 36;

// ---------------------------------------------------------------- basis data
struct Shell {
    double c[3];
    int L;
    int atom;
    std::vector<double> e, coef, norm;  // coef already multiplied by Nc
};
struct Func {
    int shell;
    int l[3];
    double cnorm;
};
struct PrimPair {
    double alpha, beta, zeta, inv_zeta, prefactor, coeff;
    double P[3], PA[3], PB[3];
};
struct Pair {
    int i, j;
    double R[3];
    std::vector<PrimPair> pp;
};

struct GbsShell {
    std::string label;
    std::vector<std::pair<double, double>> prims;
};

// gaussian.cpp:54-163: '!' comments and blank lines skipped, "****" closes an
// element, "SYM charge" opens one, "L nprim scale" opens a shell, SP splits in
// S then P, Fortran 'D' exponents read as 'E'.
bool read_gbs(const std::string &path, std::map<std::string, std::vector<GbsShell>> &out) {
    std::ifstream in(path);
    if (!in) {
        g_err = "cannot open basis file: " + path;
        return false;
    }
    std::string line, cur;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty() || line[0] == '!') continue;
        if (line == "****") {
            cur.clear();
            continue;
        }
        {
            std::istringstream h(line);
            std::string sym;
            int charge;
            if ((h >> sym >> charge) && h.eof()) {
                cur = sym;
                out[sym];
                continue;
            }
        }
        if (cur.empty()) {
            g_err = "shell before element header";
            return false;
        }
        std::istringstream iss(line);
        std::string label;
        size_t nprim = 0;
        double scale = 1.0;
        iss >> label >> nprim >> scale;
        auto fix = [](std::string s) {
            for (char &ch : s)
                if (ch == 'D' || ch == 'd') ch = 'E';
            return s;
        };
        if (label == "SP") {
            GbsShell s{"S", {}}, p{"P", {}};
            for (size_t i = 0; i < nprim; ++i) {
                std::getline(in, line);
                std::istringstream pr(fix(line));
                double ex, cs, cp;
                pr >> ex >> cs >> cp;
                s.prims.push_back({ex, cs * scale});
                p.prims.push_back({ex, cp * scale});
            }
            out[cur].push_back(s);
            out[cur].push_back(p);
        } else {
            GbsShell sh{label, {}};
            for (size_t i = 0; i < nprim; ++i) {
                std::getline(in, line);
                std::istringstream pr(fix(line));
                double ex, cs;
                pr >> ex >> cs;
                sh.prims.push_back({ex, cs * scale});
            }
            out[cur].push_back(sh);
        }
    }
    return true;
}

int label_to_L(const std::string &s) {
    const char *labels[] = {"S", "P", "D", "F", "G", "H"};
    for (int i = 0; i < 6; ++i)
        if (s == labels[i]) return i;
    return -1;
}

int dfact(int n) {
    int r = 1;
    while (n > 0) {
        r *= n;
        n -= 2;
    }
    return r;
}

// ---------------------------------------------------------------- Boys table
// boys.cpp:7-15: F_n(0.1 i) for i < 1571, n < 66, each rounded to 15 fixed
// decimals.  Regenerated here from the series
//   F_n(x) = e^{-x} sum_k (2x)^k / ((2n+1)(2n+3)...(2n+2k+1))
// in binary128 (top order) and downward recursion F_{n-1} = (2x F_n + e^{-x})/(2n-1).
constexpr int kBoysRows = 1571, kBoysCols = 66;
constexpr double kBoysStep = 0.1;
double g_boys[kBoysRows][kBoysCols];
bool g_boys_ready = false;

void init_boys_table() {
    if (g_boys_ready) return;
#pragma omp critical(pkref_boys_init)
    {
        if (!g_boys_ready) {
            for (int i = 0; i < kBoysRows; ++i) {
                __float128 x = (__float128)i / 10;
                __float128 ex = expq(-x);
                const int nt = kBoysCols - 1;
                __float128 term = (__float128)1 / (2 * nt + 1), sum = term;
                for (int k = 1; k < 5000; ++k) {
                    term *= 2 * x / (2 * nt + 2 * k + 1);
                    sum += term;
                    if (term < sum * (__float128)1e-40) break;
                }
                __float128 F[kBoysCols];
                F[nt] = ex * sum;
                for (int n = nt; n > 0; --n) F[n - 1] = (2 * x * F[n] + ex) / (2 * n - 1);
                for (int n = 0; n < kBoysCols; ++n) {
                    char buf[128];
                    quadmath_snprintf(buf, sizeof buf, "%.15Qf", F[n]);
                    g_boys[i][n] = std::strtod(buf, nullptr);
                }
            }
            g_boys_ready = true;
        }
    }
}

// Not the reference: F_n(x) to ~1e-17 relative, for checks that must see past
// the lookup's ~1e-12 error.  Below 157.1 the series
//   F_n(x) = e^{-x} sum_k (2x)^k / ((2n+1)(2n+3)...(2n+2k+1))
// in long double (x87, 64-bit mantissa; all terms positive, so the sum is
// well conditioned); beyond, the asymptotic form (its neglected part is below
// e^{-157}).  Selected by pkref_set_boys_exact(1).
int g_boys_exact = 0;
double boys_exact(int n, double x) {
    if (x >= kBoysRows * kBoysStep) return std::tgamma(n + 0.5) / (2.0 * std::pow(x, n + 0.5));
    const long double X = x;
    long double term = 1.0L / (2 * n + 1), sum = term;
    for (int k = 1; k < 4000; ++k) {
        term *= 2 * X / (2 * n + 2 * k + 1);
        sum += term;
        if (term < sum * 1e-22L) break;
    }
    return (double)(expl(-X) * sum);
}

// boys.cpp:1592-1618: asymptotic form at and beyond 157.1, else a 6-term Taylor
// expansion from the grid point below x.
double boys(int n, double x) {
    if (n < 0 || n >= kBoysCols) return std::nan("");
    if (g_boys_exact) return boys_exact(n, x);
    const double x_max = kBoysRows * kBoysStep;
    if (x >= x_max) return std::tgamma(n + 0.5) / (2.0 * std::pow(x, n + 0.5));
    static const double fact[7] = {1.0, 1.0, 2.0, 6.0, 24.0, 120.0, 720.0};
    const int xi = static_cast<int>(x / kBoysStep);
    const double d = x - xi * kBoysStep;
    const int kmax = std::min(6, kBoysCols - 1 - n);
    double r = g_boys[xi][n], dk = d, sgn = -1.0;
    for (int k = 1; k <= kmax; ++k, dk *= d, sgn = -sgn) r += sgn * g_boys[xi][n + k] / fact[k] * dk;
    return r;
}

// ---------------------------------------------------------------- tridiagonal / dense eigen
// Householder reduction to tridiagonal form followed by implicit QL with shifts
// (the EISPACK tred2/tql2 pair).  V holds eigenvectors as columns on return,
// eigenvalues ascending.
void tred2(int n, std::vector<double> &V, std::vector<double> &d, std::vector<double> &e) {
    auto v = [&](int i, int j) -> double & { return V[(size_t)i * n + j]; };
    for (int j = 0; j < n; ++j) d[j] = v(n - 1, j);
    for (int i = n - 1; i > 0; --i) {
        double scale = 0.0, h = 0.0;
        for (int k = 0; k < i; ++k) scale += std::fabs(d[k]);
        if (scale == 0.0) {
            e[i] = d[i - 1];
            for (int j = 0; j < i; ++j) {
                d[j] = v(i - 1, j);
                v(i, j) = 0.0;
                v(j, i) = 0.0;
            }
        } else {
            for (int k = 0; k < i; ++k) {
                d[k] /= scale;
                h += d[k] * d[k];
            }
            double f = d[i - 1];
            double g = std::sqrt(h);
            if (f > 0) g = -g;
            e[i] = scale * g;
            h = h - f * g;
            d[i - 1] = f - g;
            for (int j = 0; j < i; ++j) e[j] = 0.0;
            for (int j = 0; j < i; ++j) {
                f = d[j];
                v(j, i) = f;
                g = e[j] + v(j, j) * f;
                for (int k = j + 1; k <= i - 1; ++k) {
                    g += v(k, j) * d[k];
                    e[k] += v(k, j) * f;
                }
                e[j] = g;
            }
            f = 0.0;
            for (int j = 0; j < i; ++j) {
                e[j] /= h;
                f += e[j] * d[j];
            }
            double hh = f / (h + h);
            for (int j = 0; j < i; ++j) e[j] -= hh * d[j];
            for (int j = 0; j < i; ++j) {
                f = d[j];
                g = e[j];
                for (int k = j; k <= i - 1; ++k) v(k, j) -= (f * e[k] + g * d[k]);
                d[j] = v(i - 1, j);
                v(i, j) = 0.0;
            }
        }
        d[i] = h;
    }
    for (int i = 0; i < n - 1; ++i) {
        v(n - 1, i) = v(i, i);
        v(i, i) = 1.0;
        double h = d[i + 1];
        if (h != 0.0) {
            for (int k = 0; k <= i; ++k) d[k] = v(k, i + 1) / h;
            for (int j = 0; j <= i; ++j) {
                double g = 0.0;
                for (int k = 0; k <= i; ++k) g += v(k, i + 1) * v(k, j);
                for (int k = 0; k <= i; ++k) v(k, j) -= g * d[k];
            }
        }
        for (int k = 0; k <= i; ++k) v(k, i + 1) = 0.0;
    }
    for (int j = 0; j < n; ++j) {
        d[j] = v(n - 1, j);
        v(n - 1, j) = 0.0;
    }
    v(n - 1, n - 1) = 1.0;
    e[0] = 0.0;
}

bool tql2(int n, std::vector<double> &V, std::vector<double> &d, std::vector<double> &e) {
    auto v = [&](int i, int j) -> double & { return V[(size_t)i * n + j]; };
    for (int i = 1; i < n; ++i) e[i - 1] = e[i];
    e[n - 1] = 0.0;
    double f = 0.0, tst1 = 0.0;
    const double eps = std::ldexp(1.0, -52);
    for (int l = 0; l < n; ++l) {
        tst1 = std::max(tst1, std::fabs(d[l]) + std::fabs(e[l]));
        int m = l;
        while (m < n) {
            if (std::fabs(e[m]) <= eps * tst1) break;
            ++m;
        }
        if (m > l) {
            int iter = 0;
            do {
                if (++iter > 300) return false;
                double g = d[l];
                double p = (d[l + 1] - g) / (2.0 * e[l]);
                double r = std::hypot(p, 1.0);
                if (p < 0) r = -r;
                d[l] = e[l] / (p + r);
                d[l + 1] = e[l] * (p + r);
                double dl1 = d[l + 1];
                double h = g - d[l];
                for (int i = l + 2; i < n; ++i) d[i] -= h;
                f += h;
                p = d[m];
                double c = 1.0, c2 = c, c3 = c, el1 = e[l + 1], s = 0.0, s2 = 0.0;
                for (int i = m - 1; i >= l; --i) {
                    c3 = c2;
                    c2 = c;
                    s2 = s;
                    g = c * e[i];
                    h = c * p;
                    r = std::hypot(p, e[i]);
                    e[i + 1] = s * r;
                    s = e[i] / r;
                    c = p / r;
                    p = c * d[i] - s * g;
                    d[i + 1] = h + s * (c * g + s * d[i]);
                    for (int k = 0; k < n; ++k) {
                        h = v(k, i + 1);
                        v(k, i + 1) = s * v(k, i) + c * h;
                        v(k, i) = c * v(k, i) - s * h;
                    }
                }
                p = -s * s2 * c3 * el1 * e[l] / dl1;
                e[l] = s * p;
                d[l] = c * p;
            } while (std::fabs(e[l]) > eps * tst1);
        }
        d[l] = d[l] + f;
        e[l] = 0.0;
    }
    // sort ascending
    for (int i = 0; i < n - 1; ++i) {
        int k = i;
        double p = d[i];
        for (int j = i + 1; j < n; ++j)
            if (d[j] < p) {
                k = j;
                p = d[j];
            }
        if (k != i) {
            d[k] = d[i];
            d[i] = p;
            for (int j = 0; j < n; ++j) std::swap(v(j, i), v(j, k));
        }
    }
    return true;
}

// A row-major symmetric n x n; w ascending; Vout row-major with eigenvectors in columns
bool syev(int n, const double *A, double *w, double *Vout) {
    std::vector<double> V(A, A + (size_t)n * n), d(n), e(n);
    if (n == 1) {
        w[0] = A[0];
        Vout[0] = 1.0;
        return true;
    }
    tred2(n, V, d, e);
    if (!tql2(n, V, d, e)) return false;
    std::copy(d.begin(), d.end(), w);
    std::copy(V.begin(), V.end(), Vout);
    return true;
}

// ---------------------------------------------------------------- small dense helpers
using Mat = std::vector<double>;  // row-major n x n
Mat matmul(int n, const Mat &A, const Mat &B) {
    Mat C((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) {
            const double a = A[(size_t)i * n + k];
            if (a == 0.0) continue;
            const double *b = &B[(size_t)k * n];
            double *c = &C[(size_t)i * n];
            for (int j = 0; j < n; ++j) c[j] += a * b[j];
        }
    return C;
}
Mat transpose(int n, const Mat &A) {
    Mat T((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) T[(size_t)j * n + i] = A[(size_t)i * n + j];
    return T;
}

// Column-pivoted Householder QR solve of a small square system (the DIIS solve,
// types.h:813 colPivHouseholderQr).  Rank cut at eps * size * |max pivot|; the
// components beyond the rank are set to zero.
std::vector<double> qr_colpiv_solve(int n, std::vector<double> A, std::vector<double> b) {
    std::vector<int> perm(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    std::vector<double> colnorm(n);
    auto a = [&](int i, int j) -> double & { return A[(size_t)i * n + j]; };
    for (int j = 0; j < n; ++j) {
        double s = 0;
        for (int i = 0; i < n; ++i) s += a(i, j) * a(i, j);
        colnorm[j] = s;
    }
    std::vector<double> diag(n, 0.0);
    double maxpivot = 0.0;
    for (int k = 0; k < n; ++k) {
        int best = k;
        for (int j = k + 1; j < n; ++j)
            if (colnorm[j] > colnorm[best]) best = j;
        if (best != k) {
            for (int i = 0; i < n; ++i) std::swap(a(i, k), a(i, best));
            std::swap(colnorm[k], colnorm[best]);
            std::swap(perm[k], perm[best]);
        }
        double nrm = 0;
        for (int i = k; i < n; ++i) nrm += a(i, k) * a(i, k);
        nrm = std::sqrt(nrm);
        if (nrm == 0.0) {
            diag[k] = 0.0;
            continue;
        }
        double alpha = a(k, k) > 0 ? -nrm : nrm;
        std::vector<double> v(n, 0.0);
        for (int i = k; i < n; ++i) v[i] = a(i, k);
        v[k] -= alpha;
        double vn = 0;
        for (int i = k; i < n; ++i) vn += v[i] * v[i];
        if (vn > 0) {
            for (int j = k; j < n; ++j) {
                double s = 0;
                for (int i = k; i < n; ++i) s += v[i] * a(i, j);
                s = 2 * s / vn;
                for (int i = k; i < n; ++i) a(i, j) -= s * v[i];
            }
            double s = 0;
            for (int i = k; i < n; ++i) s += v[i] * b[i];
            s = 2 * s / vn;
            for (int i = k; i < n; ++i) b[i] -= s * v[i];
        }
        diag[k] = a(k, k);
        maxpivot = std::max(maxpivot, std::fabs(diag[k]));
        for (int j = k + 1; j < n; ++j) {
            double s = 0;
            for (int i = k + 1; i < n; ++i) s += a(i, j) * a(i, j);
            colnorm[j] = s;
        }
    }
    const double thr = std::ldexp(1.0, -52) * n * maxpivot;
    int rank = 0;
    for (int k = 0; k < n; ++k)
        if (std::fabs(diag[k]) > thr) ++rank;
    std::vector<double> y(n, 0.0);
    for (int k = rank - 1; k >= 0; --k) {
        double s = b[k];
        for (int j = k + 1; j < rank; ++j) s -= a(k, j) * y[j];
        y[k] = s / a(k, k);
    }
    std::vector<double> x(n, 0.0);
    for (int k = 0; k < n; ++k) x[perm[k]] = y[k];
    return x;
}

}  // namespace

// ====================================================================== context
struct pkref_ctx {
    std::vector<Shell> shells;
    std::vector<Func> funcs;
    std::vector<Pair> pairs;  // i <= j, row-major upper triangle (shellpair.cpp:7-22)
    std::vector<int> Z;
    std::vector<double> xyz;  // natoms*3 Bohr
    int nb = 0;
    int engine = 0;
    double tol = 1e-10;
    std::vector<double> Q;  // nb*nb Schwarz
    bool have_q = false;
};

namespace {

// ---------------------------------------------------------------- OS ERI (os.cpp:618-1010)
struct EriScratch {
    int d[7];
    std::vector<double> v, h;
    void resize(int ax, int ay, int az, int cx, int cy, int cz, int m) {
        d[0] = ax + 1; d[1] = ay + 1; d[2] = az + 1;
        d[3] = cx + 1; d[4] = cy + 1; d[5] = cz + 1; d[6] = m + 1;
        size_t sp = (size_t)d[0] * d[1] * d[2] * d[3] * d[4] * d[5];
        v.assign(sp * d[6], 0.0);
        h.resize(sp);
    }
    size_t si(int ax, int ay, int az, int cx, int cy, int cz) const {
        return ((((((size_t)ax * d[1] + ay) * d[2] + az) * d[3] + cx) * d[4] + cy) * d[5] + cz);
    }
    double &V(int ax, int ay, int az, int cx, int cy, int cz, int m) {
        return v[si(ax, ay, az, cx, cy, cz) * d[6] + m];
    }
    double &H(int ax, int ay, int az, int cx, int cy, int cz) { return h[si(ax, ay, az, cx, cy, cz)]; }
};
thread_local EriScratch t_eri;

// 3-D HRR (a|b) from (a+b|0) on a [13][13][13] slab, z then y then x (os.cpp:415-459)
double hrr3(double W[13][13][13], const int a[3], const int b[3], const double AB[3]) {
    for (int k = 0; k < b[2]; ++k)
        for (int ix = 0; ix <= a[0] + b[0]; ++ix)
            for (int iy = 0; iy <= a[1] + b[1]; ++iy)
                for (int iz = 0; iz <= a[2] + b[2] - k - 1; ++iz) W[ix][iy][iz] = W[ix][iy][iz + 1] + AB[2] * W[ix][iy][iz];
    for (int k = 0; k < b[1]; ++k)
        for (int ix = 0; ix <= a[0] + b[0]; ++ix)
            for (int iy = 0; iy <= a[1] + b[1] - k - 1; ++iy)
                for (int iz = 0; iz <= a[2]; ++iz) W[ix][iy][iz] = W[ix][iy + 1][iz] + AB[1] * W[ix][iy][iz];
    for (int k = 0; k < b[0]; ++k)
        for (int ix = 0; ix <= a[0] + b[0] - k - 1; ++ix)
            for (int iy = 0; iy <= a[1]; ++iy)
                for (int iz = 0; iz <= a[2]; ++iz) W[ix][iy][iz] = W[ix + 1][iy][iz] + AB[0] * W[ix][iy][iz];
    return W[a[0]][a[1]][a[2]];
}

double os_eri_primitive(const PrimPair &ab, const PrimPair &cd, const int la[3], const int lb[3], const int lc[3],
                        const int ld[3], const double AB[3], const double CD[3]) {
    int le[3], lf[3];
    for (int q = 0; q < 3; ++q) {
        le[q] = la[q] + lb[q];
        lf[q] = lc[q] + ld[q];
    }
    const int M = le[0] + le[1] + le[2] + lf[0] + lf[1] + lf[2];
    EriScratch &s = t_eri;
    s.resize(le[0], le[1], le[2], lf[0], lf[1], lf[2], M);
    const double z = ab.zeta, e = cd.zeta, dl = z + e, rho = z * e / dl;
    const double i2z = 0.5 / z, i2e = 0.5 / e, i2d = 0.5 / dl, roz = rho / z, roe = rho / e;
    double W[3], WP[3], WQ[3], PQ[3];
    for (int q = 0; q < 3; ++q) {
        W[q] = (z * ab.P[q] + e * cd.P[q]) / dl;
        WP[q] = W[q] - ab.P[q];
        WQ[q] = W[q] - cd.P[q];
        PQ[q] = ab.P[q] - cd.P[q];
    }
    const double T = rho * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
    const double pref = ab.prefactor * cd.prefactor * 2.0 * std::sqrt(rho / kPi);
    for (int m = 0; m <= M; ++m) s.V(0, 0, 0, 0, 0, 0, m) = pref * boys(m, T);
    // bra side: x, then y, then z
    for (int ax = 1; ax <= le[0]; ++ax)
        for (int m = 0; m <= M - ax; ++m) {
            double &t = s.V(ax, 0, 0, 0, 0, 0, m);
            t = ab.PA[0] * s.V(ax - 1, 0, 0, 0, 0, 0, m) + WP[0] * s.V(ax - 1, 0, 0, 0, 0, 0, m + 1);
            if (ax > 1) t += (ax - 1) * i2z * (s.V(ax - 2, 0, 0, 0, 0, 0, m) - roz * s.V(ax - 2, 0, 0, 0, 0, 0, m + 1));
        }
    for (int ax = 0; ax <= le[0]; ++ax)
        for (int ay = 1; ay <= le[1]; ++ay)
            for (int m = 0; m <= M - ax - ay; ++m) {
                double &t = s.V(ax, ay, 0, 0, 0, 0, m);
                t = ab.PA[1] * s.V(ax, ay - 1, 0, 0, 0, 0, m) + WP[1] * s.V(ax, ay - 1, 0, 0, 0, 0, m + 1);
                if (ay > 1) t += (ay - 1) * i2z * (s.V(ax, ay - 2, 0, 0, 0, 0, m) - roz * s.V(ax, ay - 2, 0, 0, 0, 0, m + 1));
            }
    for (int ax = 0; ax <= le[0]; ++ax)
        for (int ay = 0; ay <= le[1]; ++ay)
            for (int az = 1; az <= le[2]; ++az)
                for (int m = 0; m <= M - ax - ay - az; ++m) {
                    double &t = s.V(ax, ay, az, 0, 0, 0, m);
                    t = ab.PA[2] * s.V(ax, ay, az - 1, 0, 0, 0, m) + WP[2] * s.V(ax, ay, az - 1, 0, 0, 0, m + 1);
                    if (az > 1)
                        t += (az - 1) * i2z * (s.V(ax, ay, az - 2, 0, 0, 0, m) - roz * s.V(ax, ay, az - 2, 0, 0, 0, m + 1));
                }
    // ket side: x, y, z with the bra-coupling term
    for (int ax = 0; ax <= le[0]; ++ax)
        for (int ay = 0; ay <= le[1]; ++ay)
            for (int az = 0; az <= le[2]; ++az) {
                const int A = ax + ay + az;
                for (int cx = 1; cx <= lf[0]; ++cx)
                    for (int m = 0; m <= M - A - cx; ++m) {
                        double &t = s.V(ax, ay, az, cx, 0, 0, m);
                        t = cd.PA[0] * s.V(ax, ay, az, cx - 1, 0, 0, m) + WQ[0] * s.V(ax, ay, az, cx - 1, 0, 0, m + 1);
                        if (cx > 1)
                            t += (cx - 1) * i2e *
                                 (s.V(ax, ay, az, cx - 2, 0, 0, m) - roe * s.V(ax, ay, az, cx - 2, 0, 0, m + 1));
                        if (ax > 0) t += ax * i2d * s.V(ax - 1, ay, az, cx - 1, 0, 0, m + 1);
                    }
            }
    for (int ax = 0; ax <= le[0]; ++ax)
        for (int ay = 0; ay <= le[1]; ++ay)
            for (int az = 0; az <= le[2]; ++az)
                for (int cx = 0; cx <= lf[0]; ++cx) {
                    const int A = ax + ay + az + cx;
                    for (int cy = 1; cy <= lf[1]; ++cy)
                        for (int m = 0; m <= M - A - cy; ++m) {
                            double &t = s.V(ax, ay, az, cx, cy, 0, m);
                            t = cd.PA[1] * s.V(ax, ay, az, cx, cy - 1, 0, m) +
                                WQ[1] * s.V(ax, ay, az, cx, cy - 1, 0, m + 1);
                            if (cy > 1)
                                t += (cy - 1) * i2e *
                                     (s.V(ax, ay, az, cx, cy - 2, 0, m) - roe * s.V(ax, ay, az, cx, cy - 2, 0, m + 1));
                            if (ay > 0) t += ay * i2d * s.V(ax, ay - 1, az, cx, cy - 1, 0, m + 1);
                        }
                }
    for (int ax = 0; ax <= le[0]; ++ax)
        for (int ay = 0; ay <= le[1]; ++ay)
            for (int az = 0; az <= le[2]; ++az)
                for (int cx = 0; cx <= lf[0]; ++cx)
                    for (int cy = 0; cy <= lf[1]; ++cy) {
                        const int A = ax + ay + az + cx + cy;
                        for (int cz = 1; cz <= lf[2]; ++cz)
                            for (int m = 0; m <= M - A - cz; ++m) {
                                double &t = s.V(ax, ay, az, cx, cy, cz, m);
                                t = cd.PA[2] * s.V(ax, ay, az, cx, cy, cz - 1, m) +
                                    WQ[2] * s.V(ax, ay, az, cx, cy, cz - 1, m + 1);
                                if (cz > 1)
                                    t += (cz - 1) * i2e *
                                         (s.V(ax, ay, az, cx, cy, cz - 2, m) -
                                          roe * s.V(ax, ay, az, cx, cy, cz - 2, m + 1));
                                if (az > 0) t += az * i2d * s.V(ax, ay, az - 1, cx, cy, cz - 1, m + 1);
                            }
                    }
    // m = 0 slice, then A->B transfer on the 6-D slab (os.cpp:837-879)
    for (int ax = 0; ax <= le[0]; ++ax)
        for (int ay = 0; ay <= le[1]; ++ay)
            for (int az = 0; az <= le[2]; ++az)
                for (int cx = 0; cx <= lf[0]; ++cx)
                    for (int cy = 0; cy <= lf[1]; ++cy)
                        for (int cz = 0; cz <= lf[2]; ++cz)
                            s.H(ax, ay, az, cx, cy, cz) = s.V(ax, ay, az, cx, cy, cz, 0);
    for (int k = 0; k < lb[2]; ++k)
        for (int ax = 0; ax <= le[0]; ++ax)
            for (int ay = 0; ay <= le[1]; ++ay)
                for (int az = 0; az <= la[2] + lb[2] - k - 1; ++az)
                    for (int cx = 0; cx <= lf[0]; ++cx)
                        for (int cy = 0; cy <= lf[1]; ++cy)
                            for (int cz = 0; cz <= lf[2]; ++cz)
                                s.H(ax, ay, az, cx, cy, cz) = s.H(ax, ay, az + 1, cx, cy, cz) + AB[2] * s.H(ax, ay, az, cx, cy, cz);
    for (int k = 0; k < lb[1]; ++k)
        for (int ax = 0; ax <= le[0]; ++ax)
            for (int ay = 0; ay <= la[1] + lb[1] - k - 1; ++ay)
                for (int az = 0; az <= la[2]; ++az)
                    for (int cx = 0; cx <= lf[0]; ++cx)
                        for (int cy = 0; cy <= lf[1]; ++cy)
                            for (int cz = 0; cz <= lf[2]; ++cz)
                                s.H(ax, ay, az, cx, cy, cz) = s.H(ax, ay + 1, az, cx, cy, cz) + AB[1] * s.H(ax, ay, az, cx, cy, cz);
    for (int k = 0; k < lb[0]; ++k)
        for (int ax = 0; ax <= la[0] + lb[0] - k - 1; ++ax)
            for (int ay = 0; ay <= la[1]; ++ay)
                for (int az = 0; az <= la[2]; ++az)
                    for (int cx = 0; cx <= lf[0]; ++cx)
                        for (int cy = 0; cy <= lf[1]; ++cy)
                            for (int cz = 0; cz <= lf[2]; ++cz)
                                s.H(ax, ay, az, cx, cy, cz) = s.H(ax + 1, ay, az, cx, cy, cz) + AB[0] * s.H(ax, ay, az, cx, cy, cz);
    // C->D transfer on the 3-D slice at (la) (os.cpp:918-923)
    static thread_local double W3[13][13][13];
    for (int cx = 0; cx <= lf[0]; ++cx)
        for (int cy = 0; cy <= lf[1]; ++cy)
            for (int cz = 0; cz <= lf[2]; ++cz) W3[cx][cy][cz] = s.H(la[0], la[1], la[2], cx, cy, cz);
    return hrr3(W3, lc, ld, CD);
}

// ---------------------------------------------------------------- Rys (rys_roots.cpp / rys.cpp)
long double boys_moment(int m, long double T) {
    if (T < 1.0e-14L) return 1.0L / (2 * m + 1);
    if (T < 1.0L) {
        long double sum = 0.0L, coeff = 1.0L;
        for (int n = 0; n < 256; ++n) {
            const long double term = coeff / (long double)(2 * m + 2 * n + 1);
            sum += term;
            if (std::fabs(term) < 1.0e-28L * std::fabs(sum)) break;
            coeff *= -T / (long double)(n + 1);
        }
        return sum;
    }
    const long double sT = std::sqrt(T);
    long double F = 0.5L * std::sqrt(3.14159265358979323846264338327950288L / T) * erfl(sT);
    if (m == 0) return F;
    const long double eT = std::exp(-T);
    for (int k = 1; k <= m; ++k) F = ((2 * k - 1) * F - eT) / (2.0L * T);
    return F;
}

void gauss_legendre01(int n, double *r, double *w) {
    // n-point Gauss-Legendre on [0,1] by Newton on P_n (used only as the
    // degenerate-Jacobi fallback, rys_roots.cpp:17-86)
    for (int i = 0; i < n; ++i) {
        double x = std::cos(kPi * (i + 0.75) / (n + 0.5)), dp = 0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1, p1 = x;
            for (int k = 2; k <= n; ++k) {
                double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) p0 = 1;
            dp = n * (x * p1 - p0) / (x * x - 1);
            double dx = p1 / dp;
            x -= dx;
            if (std::fabs(dx) < 1e-16) break;
        }
        r[n - 1 - i] = 0.5 * (x + 1);
        w[n - 1 - i] = 1.0 / ((1 - x * x) * dp * dp);
    }
}

void rys_rule(int n, double T, double *roots, double *weights) {
    if (n == 1) {  // rys_roots.h:28-43 closed form
        if (T < 1.0e-14) {
            roots[0] = 1.0 / 3.0;
            weights[0] = 1.0;
        } else {
            const double sT = std::sqrt(T);
            const double F0 = 0.5 * std::sqrt(kPi / T) * std::erf(sT);
            const double F1 = (F0 - std::exp(-T)) / (2.0 * T);
            roots[0] = F1 / F0;
            weights[0] = F0;
        }
        return;
    }
    constexpr int MX = 11;
    long double mom[2 * MX + 1] = {};
    for (int k = 0; k <= 2 * n; ++k) mom[k] = boys_moment(k, (long double)T);
    long double poly[MX + 1][MX + 1] = {}, alpha[MX] = {}, beta[MX] = {};
    auto inner = [&](const long double *a, int da, const long double *b, int db) {
        long double s = 0;
        for (int i = 0; i <= da; ++i)
            for (int j = 0; j <= db; ++j) s += a[i] * b[j] * mom[i + j];
        return s;
    };
    poly[0][0] = 1.0L / std::sqrt(mom[0]);
    bool bad = false;
    for (int k = 0; k < n; ++k) {
        long double xp[MX + 1] = {};
        for (int i = 0; i <= k; ++i) xp[i + 1] = poly[k][i];
        alpha[k] = inner(xp, k + 1, poly[k], k);
        if (k == n - 1) break;
        long double q[MX + 1] = {};
        for (int i = 0; i <= k + 1; ++i) q[i] += xp[i];
        for (int i = 0; i <= k; ++i) q[i] -= alpha[k] * poly[k][i];
        if (k > 0)
            for (int i = 0; i <= k - 1; ++i) q[i] -= beta[k - 1] * poly[k - 1][i];
        long double nn = inner(q, k + 1, q, k + 1);
        if (nn < 0.0L && std::fabs(nn) < 1.0e-24L) nn = 0.0L;
        beta[k] = std::sqrt(nn);
        if (!(beta[k] > 0.0L)) {
            bad = true;
            break;
        }
        for (int i = 0; i <= k + 1; ++i) poly[k + 1][i] = q[i] / beta[k];
    }
    std::vector<double> J((size_t)n * n, 0.0), w(n), V((size_t)n * n);
    if (!bad) {
        for (int i = 0; i < n; ++i) {
            J[(size_t)i * n + i] = (double)alpha[i];
            if (i + 1 < n) J[(size_t)i * n + i + 1] = J[(size_t)(i + 1) * n + i] = (double)beta[i];
        }
        bad = !syev(n, J.data(), w.data(), V.data());
    }
    if (bad) {
        gauss_legendre01(n, roots, weights);
        return;
    }
    const double m0 = (double)mom[0];
    for (int i = 0; i < n; ++i) {
        roots[i] = std::clamp(w[i], 0.0, 1.0);
        weights[i] = std::max(0.0, m0 * V[i] * V[i]);  // first component of eigenvector i
    }
}

void rys_vrr_1d(double I[13][13], int le, int lf, double C00, double D00, double B00, double B10, double B01) {
    I[0][0] = 1.0;
    for (int a = 1; a <= le; ++a) {
        I[a][0] = C00 * I[a - 1][0];
        if (a >= 2) I[a][0] += (a - 1) * B10 * I[a - 2][0];
    }
    for (int c = 1; c <= lf; ++c) {
        I[0][c] = D00 * I[0][c - 1];
        if (c >= 2) I[0][c] += (c - 1) * B01 * I[0][c - 2];
        for (int a = 1; a <= le; ++a) {
            I[a][c] = D00 * I[a][c - 1] + a * B00 * I[a - 1][c - 1];
            if (c >= 2) I[a][c] += (c - 1) * B01 * I[a][c - 2];
        }
    }
}

double rys_eri_primitive(const PrimPair &ab, const PrimPair &cd, const int la[3], const int lb[3], const int lc[3],
                         const int ld[3], const double AB[3], const double CD[3]) {
    const double z = ab.zeta, e = cd.zeta, dl = z + e, idl = 1.0 / dl, rho = z * e * idl;
    const double roz = rho * ab.inv_zeta, roe = rho * cd.inv_zeta;
    double PQ[3], W[3], WP[3], WQ[3];
    for (int q = 0; q < 3; ++q) {
        PQ[q] = ab.P[q] - cd.P[q];
        W[q] = (z * ab.P[q] + e * cd.P[q]) * idl;
        WP[q] = W[q] - ab.P[q];
        WQ[q] = W[q] - cd.P[q];
    }
    const double T = rho * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
    const double pref = ab.prefactor * cd.prefactor * 2.0 * std::sqrt(rho / kPi);
    int le[3], lf[3];
    for (int q = 0; q < 3; ++q) {
        le[q] = la[q] + lb[q];
        lf[q] = lc[q] + ld[q];
    }
    const int L = le[0] + le[1] + le[2] + lf[0] + lf[1] + lf[2];
    const int n = L / 2 + 1;
    double t2[11], wt[11];
    rys_rule(n, T, t2, wt);
    // 6-D accumulation slab [le+1]^3 x [lf+1]^3
    const int d0 = le[0] + 1, d1 = le[1] + 1, d2 = le[2] + 1, d3 = lf[0] + 1, d4 = lf[1] + 1, d5 = lf[2] + 1;
    std::vector<double> S((size_t)d0 * d1 * d2 * d3 * d4 * d5, 0.0);
    auto idx = [&](int ax, int ay, int az, int cx, int cy, int cz) {
        return ((((((size_t)ax * d1 + ay) * d2 + az) * d3 + cx) * d4 + cy) * d5 + cz);
    };
    double Ix[13][13], Iy[13][13], Iz[13][13];
    for (int r = 0; r < n; ++r) {
        const double u = t2[r], wr = wt[r];
        const double B00 = 0.5 * idl * u;
        const double B10 = 0.5 * ab.inv_zeta * (1.0 - roz * u);
        const double B01 = 0.5 * cd.inv_zeta * (1.0 - roe * u);
        rys_vrr_1d(Ix, le[0], lf[0], ab.PA[0] + u * WP[0], cd.PA[0] + u * WQ[0], B00, B10, B01);
        rys_vrr_1d(Iy, le[1], lf[1], ab.PA[1] + u * WP[1], cd.PA[1] + u * WQ[1], B00, B10, B01);
        rys_vrr_1d(Iz, le[2], lf[2], ab.PA[2] + u * WP[2], cd.PA[2] + u * WQ[2], B00, B10, B01);
        for (int ax = 0; ax < d0; ++ax)
            for (int ay = 0; ay < d1; ++ay)
                for (int az = 0; az < d2; ++az)
                    for (int cx = 0; cx < d3; ++cx)
                        for (int cy = 0; cy < d4; ++cy)
                            for (int cz = 0; cz < d5; ++cz)
                                S[idx(ax, ay, az, cx, cy, cz)] += wr * Ix[ax][cx] * Iy[ay][cy] * Iz[az][cz];
    }
    for (double &x : S) x *= pref;
    for (int k = 0; k < lb[2]; ++k)
        for (int ax = 0; ax < d0; ++ax)
            for (int ay = 0; ay < d1; ++ay)
                for (int az = 0; az <= la[2] + lb[2] - k - 1; ++az)
                    for (int cx = 0; cx < d3; ++cx)
                        for (int cy = 0; cy < d4; ++cy)
                            for (int cz = 0; cz < d5; ++cz)
                                S[idx(ax, ay, az, cx, cy, cz)] = S[idx(ax, ay, az + 1, cx, cy, cz)] + AB[2] * S[idx(ax, ay, az, cx, cy, cz)];
    for (int k = 0; k < lb[1]; ++k)
        for (int ax = 0; ax < d0; ++ax)
            for (int ay = 0; ay <= la[1] + lb[1] - k - 1; ++ay)
                for (int az = 0; az <= la[2]; ++az)
                    for (int cx = 0; cx < d3; ++cx)
                        for (int cy = 0; cy < d4; ++cy)
                            for (int cz = 0; cz < d5; ++cz)
                                S[idx(ax, ay, az, cx, cy, cz)] = S[idx(ax, ay + 1, az, cx, cy, cz)] + AB[1] * S[idx(ax, ay, az, cx, cy, cz)];
    for (int k = 0; k < lb[0]; ++k)
        for (int ax = 0; ax <= la[0] + lb[0] - k - 1; ++ax)
            for (int ay = 0; ay <= la[1]; ++ay)
                for (int az = 0; az <= la[2]; ++az)
                    for (int cx = 0; cx < d3; ++cx)
                        for (int cy = 0; cy < d4; ++cy)
                            for (int cz = 0; cz < d5; ++cz)
                                S[idx(ax, ay, az, cx, cy, cz)] = S[idx(ax + 1, ay, az, cx, cy, cz)] + AB[0] * S[idx(ax, ay, az, cx, cy, cz)];
    // C->D (rys.cpp:334-363): x first, then y on the cx = lc slab, then z
    static thread_local double W3[13][13][13];
    for (int cx = 0; cx < d3; ++cx)
        for (int cy = 0; cy < d4; ++cy)
            for (int cz = 0; cz < d5; ++cz) W3[cx][cy][cz] = S[idx(la[0], la[1], la[2], cx, cy, cz)];
    for (int k = 0; k < ld[0]; ++k)
        for (int cx = 0; cx <= lc[0] + ld[0] - k - 1; ++cx)
            for (int cy = 0; cy <= lc[1] + ld[1]; ++cy)
                for (int cz = 0; cz <= lc[2] + ld[2]; ++cz) W3[cx][cy][cz] = W3[cx + 1][cy][cz] + CD[0] * W3[cx][cy][cz];
    for (int k = 0; k < ld[1]; ++k)
        for (int cy = 0; cy <= lc[1] + ld[1] - k - 1; ++cy)
            for (int cz = 0; cz <= lc[2] + ld[2]; ++cz) W3[lc[0]][cy][cz] = W3[lc[0]][cy + 1][cz] + CD[1] * W3[lc[0]][cy][cz];
    for (int k = 0; k < ld[2]; ++k)
        for (int cz = 0; cz <= lc[2] + ld[2] - k - 1; ++cz)
            W3[lc[0]][lc[1]][cz] = W3[lc[0]][lc[1]][cz + 1] + CD[2] * W3[lc[0]][lc[1]][cz];
    return W3[lc[0]][lc[1]][lc[2]];
}

// rys.cpp:195-220 per-component cost model
bool auto_prefers_rys(const int la[3], const int lb[3], const int lc[3], const int ld[3]) {
    int le[3], lf[3];
    for (int q = 0; q < 3; ++q) {
        le[q] = la[q] + lb[q];
        lf[q] = lc[q] + ld[q];
    }
    const int L = le[0] + le[1] + le[2] + lf[0] + lf[1] + lf[2];
    const int nr = L / 2 + 1;
    const double six = (double)(le[0] + 1) * (le[1] + 1) * (le[2] + 1) * (lf[0] + 1) * (lf[1] + 1) * (lf[2] + 1);
    const int lbd = lb[0] + lb[1] + lb[2] + ld[0] + ld[1] + ld[2];
    const double os = six * (L + 1) + (lbd + 1) * six * 0.25;
    const double ry = six * nr + (lbd + 1) * six * 0.20 + 24.0 * nr;
    return ry < os;
}

enum { ENG_OS = 0, ENG_RYS = 1, ENG_AUTO = 2 };

constexpr int kRysMaxOracleRoots = 5;  // rys_rule is pinned by moments up to n = 5

int total_l(const int *a, const int *b, const int *c, const int *d) {
    return a[0] + a[1] + a[2] + b[0] + b[1] + b[2] + c[0] + c[1] + c[2] + d[0] + d[1] + d[2];
}

double contracted_eri(const pkref_ctx &c, int eng, const Pair &ab, const Pair &cd) {
    const Func &A = c.funcs[ab.i], &B = c.funcs[ab.j], &C = c.funcs[cd.i], &D = c.funcs[cd.j];
    bool rys = eng == ENG_RYS || (eng == ENG_AUTO && auto_prefers_rys(A.l, B.l, C.l, D.l));
    // The reference's moment/Stieltjes/Jacobi rule loses accuracy from n = 6
    // roots (root errors 9e-9 at n = 7, 6e-3 at n = 9): the oracle does not
    // use it there and computes those components by Obara-Saika instead.
    if (rys && total_l(A.l, B.l, C.l, D.l) / 2 + 1 > kRysMaxOracleRoots) rys = false;
    double v = 0.0;
    for (const PrimPair &p : ab.pp)
        for (const PrimPair &q : cd.pp)
            v += p.coeff * q.coeff *
                 (rys ? rys_eri_primitive(p, q, A.l, B.l, C.l, D.l, ab.R, cd.R)
                      : os_eri_primitive(p, q, A.l, B.l, C.l, D.l, ab.R, cd.R));
    return v;
}

size_t pair_index(int nb, int i, int j) {
    if (i > j) std::swap(i, j);
    // row-major upper triangle: pairs (i, j>=i)
    return (size_t)i * nb - (size_t)i * (i - 1) / 2 + (j - i);
}

void ensure_schwarz(pkref_ctx &c) {
    if (c.have_q) return;
    const int nb = c.nb;
    c.Q.assign((size_t)nb * nb, 0.0);
    // engine os -> OS diagonal; rys and auto -> Rys diagonal (rys.cpp:509-556)
    const int eng = c.engine == ENG_OS ? ENG_OS : ENG_RYS;
    const long np = (long)c.pairs.size();
#pragma omp parallel for schedule(dynamic)
    for (long p = 0; p < np; ++p) {
        const Pair &pr = c.pairs[p];
        const double q = std::sqrt(std::fabs(contracted_eri(c, eng, pr, pr)));
        c.Q[(size_t)pr.i * nb + pr.j] = q;
        c.Q[(size_t)pr.j * nb + pr.i] = q;
    }
    c.have_q = true;
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

const char *pkref_last_error(void) { return g_err.c_str(); }

pkref_ctx *pkref_create(const char *gbs_path, int natoms, const int *Z, const double *xyz, int engine, double tol) {
    init_boys_table();
    std::map<std::string, std::vector<GbsShell>> gbs;
    if (!read_gbs(gbs_path, gbs)) return nullptr;
    auto *c = new pkref_ctx;
    c->engine = engine;
    c->tol = tol;
    c->Z.assign(Z, Z + natoms);
    c->xyz.assign(xyz, xyz + 3 * natoms);
    for (int a = 0; a < natoms; ++a) {
        if (Z[a] < 1 || Z[a] > kMaxZ) {
            g_err = "unsupported element Z=" + std::to_string(Z[a]);
            delete c;
            return nullptr;
        }
        auto it = gbs.find(kSymbols[Z[a]]);
        if (it == gbs.end()) {
            g_err = std::string("Unsupported Element: ") + kSymbols[Z[a]];
            delete c;
            return nullptr;
        }
        for (const GbsShell &g : it->second) {
            Shell s;
            s.c[0] = xyz[3 * a];
            s.c[1] = xyz[3 * a + 1];
            s.c[2] = xyz[3 * a + 2];
            s.L = label_to_L(g.label);
            s.atom = a;
            if (s.L < 0) {
                g_err = "unknown shell label " + g.label;
                delete c;
                return nullptr;
            }
            const int L = s.L, np = (int)g.prims.size();
            for (auto &pr : g.prims) {
                s.e.push_back(pr.first);
                s.coef.push_back(pr.second);
            }
            // basis.cpp:39-48 primitive norm; basis.cpp:50-75 contracted norm
            const double pf = std::pow(2.0, 2.0 * L + 1.5) / std::pow(kPi, 1.5);
            for (int i = 0; i < np; ++i) s.norm.push_back(std::sqrt(std::pow(s.e[i], L + 1.5) * pf));
            const double pif = std::pow(kPi, 1.5) / std::pow(2.0, L);
            double sum = 0.0;
            for (int i = 0; i < np; ++i)
                for (int j = 0; j < np; ++j)
                    sum += s.coef[i] * s.coef[j] * s.norm[i] * s.norm[j] * (pif / std::pow(s.e[i] + s.e[j], L + 1.5));
            const double Nc = 1.0 / std::sqrt(sum);
            for (double &x : s.coef) x *= Nc;
            const int sidx = (int)c->shells.size();
            c->shells.push_back(s);
            for (int lx = L; lx >= 0; --lx)
                for (int ly = L - lx; ly >= 0; --ly) {
                    Func f;
                    f.shell = sidx;
                    f.l[0] = lx;
                    f.l[1] = ly;
                    f.l[2] = L - lx - ly;
                    f.cnorm = 1.0 / std::sqrt((double)(dfact(2 * lx - 1) * dfact(2 * ly - 1) * dfact(2 * f.l[2] - 1)));
                    c->funcs.push_back(f);
                }
        }
    }
    c->nb = (int)c->funcs.size();
    // types.h:606-675 pair data for every AO pair i <= j
    const int nb = c->nb;
    c->pairs.reserve((size_t)nb * (nb + 1) / 2);
    for (int i = 0; i < nb; ++i)
        for (int j = i; j < nb; ++j) {
            Pair p;
            p.i = i;
            p.j = j;
            const Func &fa = c->funcs[i], &fb = c->funcs[j];
            const Shell &A = c->shells[fa.shell], &B = c->shells[fb.shell];
            for (int q = 0; q < 3; ++q) p.R[q] = A.c[q] - B.c[q];
            const double R2 = p.R[0] * p.R[0] + p.R[1] * p.R[1] + p.R[2] * p.R[2];
            for (size_t ia = 0; ia < A.e.size(); ++ia)
                for (size_t ib = 0; ib < B.e.size(); ++ib) {
                    PrimPair pp;
                    pp.alpha = A.e[ia];
                    pp.beta = B.e[ib];
                    pp.zeta = pp.alpha + pp.beta;
                    pp.inv_zeta = 1.0 / pp.zeta;
                    pp.coeff = A.coef[ia] * B.coef[ib] * A.norm[ia] * B.norm[ib] * fa.cnorm * fb.cnorm;
                    pp.prefactor = std::pow(kPi * pp.inv_zeta, 1.5) * std::exp(-(pp.alpha * pp.beta * pp.inv_zeta) * R2);
                    for (int q = 0; q < 3; ++q) {
                        pp.P[q] = (pp.alpha * A.c[q] + pp.beta * B.c[q]) * pp.inv_zeta;
                        pp.PA[q] = pp.P[q] - A.c[q];
                        pp.PB[q] = pp.P[q] - B.c[q];
                    }
                    p.pp.push_back(pp);
                }
            c->pairs.push_back(std::move(p));
        }
    return c;
}

void pkref_destroy(pkref_ctx *c) { delete c; }
int pkref_nbasis(const pkref_ctx *c) { return c->nb; }
int pkref_nshells(const pkref_ctx *c) { return (int)c->shells.size(); }

int pkref_shell_info(const pkref_ctx *c, int s, int *L, int *nprim, int *atom, double *center, double *exps,
                     double *coefs, double *norms) {
    if (s < 0 || s >= (int)c->shells.size()) return -1;
    const Shell &sh = c->shells[s];
    *L = sh.L;
    *nprim = (int)sh.e.size();
    *atom = sh.atom;
    for (int q = 0; q < 3; ++q) center[q] = sh.c[q];
    if (exps) std::copy(sh.e.begin(), sh.e.end(), exps);
    if (coefs) std::copy(sh.coef.begin(), sh.coef.end(), coefs);
    if (norms) std::copy(sh.norm.begin(), sh.norm.end(), norms);
    return 0;
}

double pkref_nuclear_repulsion(const pkref_ctx *c) {
    const int n = (int)c->Z.size();
    double E = 0.0;
    for (int a = 0; a < n; ++a)
        for (int b = a + 1; b < n; ++b) {
            const double dx = c->xyz[3 * a] - c->xyz[3 * b], dy = c->xyz[3 * a + 1] - c->xyz[3 * b + 1],
                         dz = c->xyz[3 * a + 2] - c->xyz[3 * b + 2];
            E += (double)c->Z[a] * (double)c->Z[b] / std::sqrt(dx * dx + dy * dy + dz * dz);
        }
    return E;
}

// os.cpp:248-292 1-D overlap recursion
static double os_1d(double g, double pa, double pb, int la, int lb) {
    double S[16][16] = {};
    S[0][0] = 1.0;
    for (int i = 1; i <= la; ++i) {
        S[i][0] = pa * S[i - 1][0];
        if (i > 1) S[i][0] = (i - 1) * g * S[i - 2][0] + S[i][0];
    }
    for (int j = 1; j <= lb; ++j) {
        S[0][j] = pb * S[0][j - 1];
        if (j > 1) S[0][j] = (j - 1) * g * S[0][j - 2] + S[0][j];
    }
    for (int i = 1; i <= la; ++i)
        for (int j = 1; j <= lb; ++j) {
            S[i][j] = pa * S[i - 1][j] + j * g * S[i - 1][j - 1];
            if (i > 1) S[i][j] += (i - 1) * g * S[i - 2][j];
        }
    return S[la][lb];
}

static void overlap_kinetic(const pkref_ctx *c, const Pair &p, double &S, double &T) {
    const int *la = c->funcs[p.i].l, *lb = c->funcs[p.j].l;
    S = T = 0.0;
    for (const PrimPair &pp : p.pp) {
        const double g = pp.inv_zeta * 0.5, b = pp.beta, b2 = b * b, sc = pp.prefactor * pp.coeff;
        double s[3], sp[3], sm[3], t[3];
        for (int q = 0; q < 3; ++q) {
            s[q] = os_1d(g, pp.PA[q], pp.PB[q], la[q], lb[q]);
            sp[q] = os_1d(g, pp.PA[q], pp.PB[q], la[q], lb[q] + 2);
            sm[q] = lb[q] >= 2 ? os_1d(g, pp.PA[q], pp.PB[q], la[q], lb[q] - 2) : 0.0;
            t[q] = -0.5 * lb[q] * (lb[q] - 1) * sm[q] + b * (2 * lb[q] + 1) * s[q] - 2.0 * b2 * sp[q];
        }
        S += s[0] * s[1] * s[2] * sc;
        T += (t[0] * s[1] * s[2] + s[0] * t[1] * s[2] + s[0] * s[1] * t[2]) * sc;
    }
}

void pkref_overlap(const pkref_ctx *c, double *S) {
    const int nb = c->nb;
    for (const Pair &p : c->pairs) {
        double s, t;
        overlap_kinetic(c, p, s, t);
        S[(size_t)p.i * nb + p.j] = S[(size_t)p.j * nb + p.i] = s;
    }
}
void pkref_kinetic(const pkref_ctx *c, double *T) {
    const int nb = c->nb;
    for (const Pair &p : c->pairs) {
        double s, t;
        overlap_kinetic(c, p, s, t);
        T[(size_t)p.i * nb + p.j] = T[(size_t)p.j * nb + p.i] = t;
    }
}

// os.cpp:462-571 nuclear attraction primitive; os.cpp:520-560 accumulation -Z * sum
static double nuclear_primitive(const PrimPair &pp, const int la[3], const int lb[3], const double AB[3], const double *C) {
    const int L = la[0] + la[1] + la[2] + lb[0] + lb[1] + lb[2];
    double pc[3];
    for (int q = 0; q < 3; ++q) pc[q] = pp.P[q] - C[q];
    const double T = pp.zeta * (pc[0] * pc[0] + pc[1] * pc[1] + pc[2] * pc[2]);
    const double np = pp.prefactor * 2.0 * std::sqrt(pp.zeta / kPi);
    static thread_local double V[13][13][13][15];
    const int lx = la[0] + lb[0], ly = la[1] + lb[1], lz = la[2] + lb[2];
    for (int ix = 0; ix <= lx; ++ix)
        for (int iy = 0; iy <= ly; ++iy)
            for (int iz = 0; iz <= lz; ++iz)
                for (int m = 0; m <= L + 1; ++m) V[ix][iy][iz][m] = 0.0;
    for (int m = 0; m <= L; ++m) V[0][0][0][m] = np * boys(m, T);
    const double hiz = pp.inv_zeta * 0.5;
    for (int ix = 1; ix <= lx; ++ix)
        for (int m = 0; m <= L - ix; ++m) {
            V[ix][0][0][m] = pp.PA[0] * V[ix - 1][0][0][m] - pc[0] * V[ix - 1][0][0][m + 1];
            if (ix > 1) V[ix][0][0][m] += (ix - 1) * hiz * (V[ix - 2][0][0][m] - V[ix - 2][0][0][m + 1]);
        }
    for (int ix = 0; ix <= lx; ++ix)
        for (int iy = 1; iy <= ly; ++iy)
            for (int m = 0; m <= L - ix - iy; ++m) {
                V[ix][iy][0][m] = pp.PA[1] * V[ix][iy - 1][0][m] - pc[1] * V[ix][iy - 1][0][m + 1];
                if (iy > 1) V[ix][iy][0][m] += (iy - 1) * hiz * (V[ix][iy - 2][0][m] - V[ix][iy - 2][0][m + 1]);
            }
    for (int ix = 0; ix <= lx; ++ix)
        for (int iy = 0; iy <= ly; ++iy)
            for (int iz = 1; iz <= lz; ++iz)
                for (int m = 0; m <= L - ix - iy - iz; ++m) {
                    V[ix][iy][iz][m] = pp.PA[2] * V[ix][iy][iz - 1][m] - pc[2] * V[ix][iy][iz - 1][m + 1];
                    if (iz > 1) V[ix][iy][iz][m] += (iz - 1) * hiz * (V[ix][iy][iz - 2][m] - V[ix][iy][iz - 2][m + 1]);
                }
    static thread_local double W[13][13][13];
    for (int ix = 0; ix <= lx; ++ix)
        for (int iy = 0; iy <= ly; ++iy)
            for (int iz = 0; iz <= lz; ++iz) W[ix][iy][iz] = V[ix][iy][iz][0];
    return hrr3(W, la, lb, AB);
}

void pkref_nuclear(const pkref_ctx *c, double *Vm) {
    const int nb = c->nb, na = (int)c->Z.size();
    for (const Pair &p : c->pairs) {
        const int *la = c->funcs[p.i].l, *lb = c->funcs[p.j].l;
        double v = 0.0;
        for (int a = 0; a < na; ++a) {
            double vp = 0.0;
            for (const PrimPair &pp : p.pp) vp += nuclear_primitive(pp, la, lb, p.R, &c->xyz[3 * a]) * pp.coeff;
            v -= (double)c->Z[a] * vp;
        }
        Vm[(size_t)p.i * nb + p.j] = Vm[(size_t)p.j * nb + p.i] = v;
    }
}

void pkref_schwarz(pkref_ctx *c, double *Q) {
    ensure_schwarz(*c);
    std::copy(c->Q.begin(), c->Q.end(), Q);
}

double pkref_eri(pkref_ctx *c, int i, int j, int k, int l) {
    const Pair &ab = c->pairs[pair_index(c->nb, i, j)];
    const Pair &cd = c->pairs[pair_index(c->nb, k, l)];
    // the stored pair is (min, max); the ERI is symmetric under i<->j, k<->l
    return contracted_eri(*c, c->engine, ab, cd);
}

double pkref_eri_engine(pkref_ctx *c, int e, int i, int j, int k, int l) {
    const Pair &ab = c->pairs[pair_index(c->nb, i, j)];
    const Pair &cd = c->pairs[pair_index(c->nb, k, l)];
    return contracted_eri(*c, e, ab, cd);
}

int pkref_eri_full(pkref_ctx *c, double *out) {
    ensure_schwarz(*c);
    const size_t nb = c->nb, nb2 = nb * nb, nb3 = nb2 * nb;
    std::fill(out, out + nb3 * nb, 0.0);
    const long np = (long)c->pairs.size();
#pragma omp parallel for schedule(dynamic)
    for (long p = 0; p < np; ++p) {
        const Pair &ab = c->pairs[p];
        for (long q = p; q < np; ++q) {
            const Pair &cd = c->pairs[q];
            const size_t i = ab.i, j = ab.j, k = cd.i, l = cd.j;
            if (c->Q[i * nb + j] * c->Q[k * nb + l] < c->tol) continue;
            const double v = contracted_eri(*c, c->engine, ab, cd);
            out[i * nb3 + j * nb2 + k * nb + l] = v;
            out[j * nb3 + i * nb2 + k * nb + l] = v;
            out[i * nb3 + j * nb2 + l * nb + k] = v;
            out[j * nb3 + i * nb2 + l * nb + k] = v;
            out[k * nb3 + l * nb2 + i * nb + j] = v;
            out[l * nb3 + k * nb2 + i * nb + j] = v;
            out[k * nb3 + l * nb2 + j * nb + i] = v;
            out[l * nb3 + k * nb2 + j * nb + i] = v;
        }
    }
    return 0;
}

// Direct J/K: every unique AO quartet (p <= q over AO pairs, each pair i <= j)
// passing Q_ij * Q_kl >= tol contributes its 8 permutations with degeneracy d:
//   A_ij += 2 d v P_kl, A_kl += 2 d v P_ij            J = A + A^T
//   B_ik, B_jk, B_il, B_jl += d v P_jl, P_il, P_jk, P_ik   K = B + B^T
// This equals the reference's nb^4 scatter + contraction (os.cpp:1452-1522).
int pkref_fock_jk_sample(pkref_ctx *c, const double *P, double *J, double *K, int nthreads, long stride,
                         long offset, double *n_quartets) {
    ensure_schwarz(*c);
    const int nb = c->nb;
    const long np = (long)c->pairs.size();
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
    // The sampled bra pairs are dealt round-robin to a fixed number of slots,
    // each summed serially into its own accumulators and the slots added in
    // order: the result is the same for every thread count and every run.
    const long nbra = np > offset ? (np - offset + stride - 1) / stride : 0;
    const int nslot = (int)std::min<long>(64, std::max<long>(1, nbra));
    std::vector<std::vector<double>> Aacc(nslot), Bacc(nslot);
    double count = 0.0;
#pragma omp parallel for num_threads(nt) schedule(dynamic) reduction(+ : count)
    for (int sl = 0; sl < nslot; ++sl) {
        std::vector<double> &A = Aacc[sl], &B = Bacc[sl];
        A.assign((size_t)nb * nb, 0.0);
        B.assign((size_t)nb * nb, 0.0);
        for (long ib = sl; ib < nbra; ib += nslot) {
            const long p = offset + ib * stride;
            const Pair &ab = c->pairs[p];
            const size_t i = ab.i, j = ab.j;
            const double qij = c->Q[i * nb + j];
            for (long q = p; q < np; ++q) {
                const Pair &cd = c->pairs[q];
                const size_t k = cd.i, l = cd.j;
                if (qij * c->Q[k * nb + l] < c->tol) continue;
                count += 1.0;
                const double v = contracted_eri(*c, c->engine, ab, cd);
                double d = 1.0;
                if (i == j) d *= 0.5;
                if (k == l) d *= 0.5;
                if (p == q) d *= 0.5;
                const double dv = d * v;
                A[i * nb + j] += 2.0 * dv * P[k * nb + l];
                A[k * nb + l] += 2.0 * dv * P[i * nb + j];
                B[i * nb + k] += dv * P[j * nb + l];
                B[j * nb + k] += dv * P[i * nb + l];
                B[i * nb + l] += dv * P[j * nb + k];
                B[j * nb + l] += dv * P[i * nb + k];
            }
        }
    }
    if (nbra == 0) {
        Aacc[0].assign((size_t)nb * nb, 0.0);
        Bacc[0].assign((size_t)nb * nb, 0.0);
    }
    for (size_t x = 0; x < (size_t)nb * nb; ++x) {
        double a = 0, b = 0;
        for (int t = 0; t < nslot; ++t) {
            a += Aacc[t][x];
            b += Bacc[t][x];
        }
        Aacc[0][x] = a;
        Bacc[0][x] = b;
    }
    for (int i = 0; i < nb; ++i)
        for (int j = 0; j < nb; ++j) {
            J[(size_t)i * nb + j] = Aacc[0][(size_t)i * nb + j] + Aacc[0][(size_t)j * nb + i];
            K[(size_t)i * nb + j] = Bacc[0][(size_t)i * nb + j] + Bacc[0][(size_t)j * nb + i];
        }
    if (n_quartets) *n_quartets = count;
    return 0;
}

int pkref_fock_jk(pkref_ctx *c, const double *P, double *J, double *K, int nthreads) {
    return pkref_fock_jk_sample(c, P, J, K, nthreads, 1, 0, nullptr);
}

// Single elements of J and K, each from its own sums over every AO quartet
// that passes the same screening test as the full build (Q_ij Q_kl is
// invariant under the 8 permutations, so each term below is kept exactly
// when its unique quartet is):
//   J_ij = sum_{k,l} (ij|kl) P_kl,   K_ij = sum_{k,l} (ik|jl) P_kl.
// Equal to the corresponding elements of pkref_fock_jk (os.cpp:1509-1522),
// at a cost of O(nb^2) ERIs per element instead of the whole build, so large
// molecules can be checked on a sample of elements.  Summed serially in long
// double per element: the result does not depend on the thread count.
int pkref_jk_elements(pkref_ctx *c, const double *P, int nel, const int *ei, const int *ej, double *J, double *K,
                      int nthreads) {
    ensure_schwarz(*c);
    const int nb = c->nb;
    for (int e = 0; e < nel; ++e)
        if (ei[e] < 0 || ei[e] >= nb || ej[e] < 0 || ej[e] >= nb) {
            g_err = "pkref_jk_elements: element index out of range";
            return -1;
        }
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for num_threads(nt) schedule(dynamic)
    for (int e = 0; e < nel; ++e) {
        const int i = ei[e], j = ej[e];
        const Pair &ij = c->pairs[pair_index(nb, i, j)];
        const double qij = c->Q[(size_t)i * nb + j];
        long double sj = 0.0L, sk = 0.0L;
        for (int k = 0; k < nb; ++k)
            for (int l = k; l < nb; ++l) {
                if (qij * c->Q[(size_t)k * nb + l] < c->tol) continue;
                const double v = contracted_eri(*c, c->engine, ij, c->pairs[pair_index(nb, k, l)]);
                sj += (long double)v * (k == l ? P[(size_t)k * nb + l] : P[(size_t)k * nb + l] + P[(size_t)l * nb + k]);
            }
        for (int k = 0; k < nb; ++k) {
            const Pair &ik = c->pairs[pair_index(nb, i, k)];
            const double qik = c->Q[(size_t)i * nb + k];
            for (int l = 0; l < nb; ++l) {
                if (qik * c->Q[(size_t)j * nb + l] < c->tol) continue;
                const double v = contracted_eri(*c, c->engine, ik, c->pairs[pair_index(nb, j, l)]);
                sk += (long double)v * P[(size_t)k * nb + l];
            }
        }
        J[e] = (double)sj;
        K[e] = (double)sk;
    }
    return 0;
}

// ERI first derivatives of one contracted Cartesian quartet (ij|kl), restating
// ObaraSaika::_compute_eri_deriv_elem (src/integrals/os.cpp:1296-1420): for
// centre X in (A, B, C, D) of the functions (i, j, k, l) and axis q,
//   d(ij|kl)/dX_q = sum_prims 2 x_exp c_ij c_kl [.. x+1_q ..] - x_q (.. x-1_q ..)
// where x_exp is the primitive exponent on that centre (alpha, beta of the bra
// pair, alpha, beta of the ket pair), [..] the primitive OS ERI with the
// centre's Cartesian exponent raised along q and (..) the contracted ERI with
// it lowered (absent when x_q = 0).  out[cen * 3 + q], cen 0..3 = A, B, C, D.
static void eri_deriv_elem(const pkref_ctx &c, const Pair &ab, const Pair &cd, double out[12]) {
    const Func &A = c.funcs[ab.i], &B = c.funcs[ab.j], &C = c.funcs[cd.i], &D = c.funcs[cd.j];
    for (int x = 0; x < 12; ++x) out[x] = 0.0;
    for (int cen = 0; cen < 4; ++cen)
        for (int q = 0; q < 3; ++q) {
            int l[4][3];
            for (int t = 0; t < 3; ++t) {
                l[0][t] = A.l[t];
                l[1][t] = B.l[t];
                l[2][t] = C.l[t];
                l[3][t] = D.l[t];
            }
            const int lq = l[cen][q];
            l[cen][q] = lq + 1;
            double up = 0.0;
            for (const PrimPair &p : ab.pp)
                for (const PrimPair &r : cd.pp) {
                    const double w = cen == 0 ? p.alpha : cen == 1 ? p.beta : cen == 2 ? r.alpha : r.beta;
                    up += 2.0 * w * p.coeff * r.coeff * os_eri_primitive(p, r, l[0], l[1], l[2], l[3], ab.R, cd.R);
                }
            double down = 0.0;
            if (lq > 0) {
                l[cen][q] = lq - 1;
                for (const PrimPair &p : ab.pp)
                    for (const PrimPair &r : cd.pp)
                        down += p.coeff * r.coeff * os_eri_primitive(p, r, l[0], l[1], l[2], l[3], ab.R, cd.R);
            }
            out[cen * 3 + q] = up - (double)lq * down;
        }
}

// Two-electron part of the nuclear gradient, restating the ERI loop of
// compute_closed_shell_gradient_from_density / compute_uhf_gradient
// (src/gradient/gradient.cpp:209-232, 436-463) with
// accumulate_eri_gradient_permutations (gradient.cpp:70-118): every ordered
// pair of AO pairs (i <= j, k <= l) passing Q_ij Q_kl >= tol contributes
// 0.25 Gamma d(ij|kl)/dX to the atom of each centre, for the index
// permutations (ij|kl), (ij|lk), (ji|kl), (ji|lk) that are distinct, each
// skipped when |Gamma| < 1e-14 as in the reference, with
//   Gamma_ijkl = 2 Pj_ij Pj_kl - kx (Pk1_ik Pk1_jl + Pk2_ik Pk2_jl)
// (RHF: Pj = Pk1 = P, no Pk2, kx = 1; RKS: kx = the exact-exchange fraction;
// UHF: Pj = Pa + Pb, Pk1 = Pa, Pk2 = Pb, kx = 2).  grad is natoms x 3
// (row-major), overwritten.  Deterministic: fixed slots summed in order.
static int eri_gradient(pkref_ctx *c, const double *Pj, const double *Pk1, const double *Pk2, double kx,
                        double *grad, int nthreads) {
    ensure_schwarz(*c);
    const int nb = c->nb, nat = (int)c->Z.size();
    const long np = (long)c->pairs.size();
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
    auto gamma = [&](size_t i, size_t j, size_t k, size_t l) {
        double g = 2.0 * Pj[i * nb + j] * Pj[k * nb + l] - kx * Pk1[i * nb + k] * Pk1[j * nb + l];
        if (Pk2) g -= kx * Pk2[i * nb + k] * Pk2[j * nb + l];
        return g;
    };
    const int nslot = (int)std::min<long>(64, std::max<long>(1, np));
    std::vector<std::vector<double>> acc(nslot, std::vector<double>((size_t)nat * 3, 0.0));
#pragma omp parallel for num_threads(nt) schedule(dynamic)
    for (int sl = 0; sl < nslot; ++sl) {
        std::vector<double> &g = acc[sl];
        for (long p = sl; p < np; p += nslot) {
            const Pair &ab = c->pairs[p];
            const size_t i = ab.i, j = ab.j;
            const int atA = c->shells[c->funcs[i].shell].atom, atB = c->shells[c->funcs[j].shell].atom;
            for (long q = 0; q < np; ++q) {
                const Pair &cd = c->pairs[q];
                const size_t k = cd.i, l = cd.j;
                if (c->Q[i * nb + j] * c->Q[k * nb + l] < c->tol) continue;
                const int atC = c->shells[c->funcs[k].shell].atom, atD = c->shells[c->funcs[l].shell].atom;
                double dI[12];
                eri_deriv_elem(*c, ab, cd, dI);
                auto perm = [&](double gm, bool sab, bool scd) {
                    if (std::fabs(gm) < 1e-14) return;
                    const int da = sab ? 1 : 0, db = sab ? 0 : 1, dc = scd ? 3 : 2, dd = scd ? 2 : 3;
                    const int xa = sab ? atB : atA, xb = sab ? atA : atB, xc = scd ? atD : atC, xd = scd ? atC : atD;
                    const double f = 0.25 * gm;
                    for (int t = 0; t < 3; ++t) {
                        g[xa * 3 + t] += f * dI[da * 3 + t];
                        g[xb * 3 + t] += f * dI[db * 3 + t];
                        g[xc * 3 + t] += f * dI[dc * 3 + t];
                        g[xd * 3 + t] += f * dI[dd * 3 + t];
                    }
                };
                perm(gamma(i, j, k, l), false, false);
                if (k != l) perm(gamma(i, j, l, k), false, true);
                if (i != j) perm(gamma(j, i, k, l), true, false);
                if (i != j && k != l) perm(gamma(j, i, l, k), true, true);
            }
        }
    }
    for (int x = 0; x < nat * 3; ++x) {
        double s = 0.0;
        for (int t = 0; t < nslot; ++t) s += acc[t][x];
        grad[x] = s;
    }
    return 0;
}

// Primitive-pair screening of the engine (not reference behaviour: the
// reference keeps every primitive pair), restated so that its effect can be
// measured against the unscreened oracle.  A primitive pair is a primitive of
// one shell times a primitive of another (or the same) shell; it is dropped
// from every function pair of those two shells when
//   rule 0: |kc_p| * max_q |kc_q| / sqrt(zeta_q) < cut,
//           kc = (pi/zeta)^1.5 exp(-alpha beta R^2/zeta) c_a N_a c_b N_b sqrt(zeta) (4/pi)^(1/4)
//           (a prefactor bound that leaves out the angular factors);
//   rule 1: Q_p * max_q Q_q < cut, Q_p = max over the components a, b of the
//           two shells of sqrt|[ab|ab]| of the single primitive pair, with
//           c_a N_a c_b N_b and no component norms (the primitive Schwarz
//           bound: no primitive quartet moves a contracted integral by more
//           than Q_p Q_q).
// The Schwarz table is recomputed from the remaining primitives.  Returns
// the number of primitive pairs (of shell pairs) dropped.
int pkref_screen_primitives(pkref_ctx *c, double cut, int rule) {
    struct Key {
        int s0, s1, p0, p1;
        bool operator<(const Key &o) const {
            return s0 != o.s0 ? s0 < o.s0 : s1 != o.s1 ? s1 < o.s1 : p0 != o.p0 ? p0 < o.p0 : p1 < o.p1;
        }
    };
    auto key = [&](const Pair &pr, size_t k) {
        const Func &fa = c->funcs[pr.i], &fb = c->funcs[pr.j];
        const int na = (int)c->shells[fa.shell].e.size(), nbp = (int)c->shells[fb.shell].e.size();
        (void)na;
        int ia = (int)(k / nbp), ib = (int)(k % nbp), sa = fa.shell, sb = fb.shell;
        if (sa > sb || (sa == sb && ia > ib)) {
            std::swap(sa, sb);
            std::swap(ia, ib);
        }
        return Key{sa, sb, ia, ib};
    };
    std::map<Key, double> bound;  // per primitive pair of shells
    for (const Pair &pr : c->pairs) {
        const Func &fa = c->funcs[pr.i], &fb = c->funcs[pr.j];
        for (size_t k = 0; k < pr.pp.size(); ++k) {
            const PrimPair &pp = pr.pp[k];
            const double raw = pp.coeff / (fa.cnorm * fb.cnorm);
            double v;
            if (rule == 0) {
                v = std::fabs(pp.prefactor * raw) * std::sqrt(pp.zeta) * 1.0622519320271969145;
            } else {
                PrimPair q = pp;
                v = std::sqrt(std::fabs(raw * raw * os_eri_primitive(q, q, fa.l, fb.l, fa.l, fb.l, pr.R, pr.R)));
            }
            double &b = bound[key(pr, k)];
            if (rule == 0) {
                b = v;  // the same for every component pair
            } else {
                b = std::max(b, v);
            }
        }
    }
    double partner = 0.0;  // max over primitive pairs of the partner factor
    for (const Pair &pr : c->pairs)
        for (size_t k = 0; k < pr.pp.size(); ++k) {
            const double b = bound[key(pr, k)];
            partner = std::max(partner, rule == 0 ? b / std::sqrt(pr.pp[k].zeta) : b);
        }
    int dropped = 0;
    for (const auto &kv : bound)
        if (kv.second * partner < cut) ++dropped;
    for (Pair &pr : c->pairs) {
        std::vector<PrimPair> keep;
        for (size_t k = 0; k < pr.pp.size(); ++k)
            if (bound[key(pr, k)] * partner >= cut) keep.push_back(pr.pp[k]);
        pr.pp.swap(keep);
    }
    c->have_q = false;
    return dropped;
}

int pkref_eri_deriv(pkref_ctx *c, int i, int j, int k, int l, double *out12) {
    const int nb = c->nb;
    if (i < 0 || j < 0 || k < 0 || l < 0 || i >= nb || j >= nb || k >= nb || l >= nb || i > j || k > l) {
        g_err = "pkref_eri_deriv: indices must satisfy 0 <= i <= j < nb, 0 <= k <= l < nb";
        return -1;
    }
    eri_deriv_elem(*c, c->pairs[pair_index(nb, i, j)], c->pairs[pair_index(nb, k, l)], out12);
    return 0;
}

int pkref_eri_gradient(pkref_ctx *c, const double *Pj, const double *Pk1, const double *Pk2, double kx,
                       double *grad, int nthreads) {
    if (!Pj || !Pk1 || !grad) {
        g_err = "pkref_eri_gradient: null density or output";
        return -1;
    }
    return eri_gradient(c, Pj, Pk1, Pk2, kx, grad, nthreads);
}

double pkref_boys(int n, double x) {
    init_boys_table();
    return boys(n, x);
}

void pkref_set_boys_exact(int on) { g_boys_exact = on ? 1 : 0; }

int pkref_rys_rule(int n, double T, double *roots, double *weights) {
    // refused beyond the roots its moment test pins (see contracted_eri)
    if (n < 1 || n > kRysMaxOracleRoots) {
        g_err = "pkref_rys_rule: n = " + std::to_string(n) + " is outside 1.." + std::to_string(kRysMaxOracleRoots) +
                " (the moment method is inaccurate beyond)";
        for (int i = 0; i < n && i < 11; ++i) roots[i] = weights[i] = std::nan("");
        return -1;
    }
    rys_rule(n, T, roots, weights);
    return 0;
}

int pkref_auto_prefers_rys(const int *lA, const int *lB, const int *lC, const int *lD) {
    return auto_prefers_rys(lA, lB, lC, lD) ? 1 : 0;
}

int pkref_syev(int n, const double *A, double *w, double *V) { return syev(n, A, w, V) ? 0 : -1; }

// scf.cpp:175-441 run_rhf with the core-Hamiltonian guess and DIIS.
int pkref_rhf(pkref_ctx *c, int charge, double tol_e, double tol_d, int diis_dim, int max_cycles, double *energy,
              int *iterations, double *P_out, double *hist, int hist_len) {
    const int n = c->nb;
    int nel = -charge;
    for (int z : c->Z) nel += z;
    if (nel % 2 != 0) {
        g_err = "RHF requires an even number of electrons (closed shell)";
        return -1;
    }
    const int nocc = nel / 2;
    if (max_cycles <= 0) max_cycles = n > 1000 ? 300 : n > 500 ? 200 : n > 250 ? 100 : 50;
    Mat S((size_t)n * n), T((size_t)n * n), V((size_t)n * n), H((size_t)n * n);
    pkref_overlap(c, S.data());
    pkref_kinetic(c, T.data());
    pkref_nuclear(c, V.data());
    for (size_t x = 0; x < H.size(); ++x) H[x] = T[x] + V[x];
    const double Enuc = pkref_nuclear_repulsion(c);
    // X = U s^-1/2 U^T (scf.cpp:15-29)
    std::vector<double> w(n);
    Mat U((size_t)n * n);
    if (!syev(n, S.data(), w.data(), U.data())) {
        g_err = "Overlap matrix diagonalization failed";
        return -1;
    }
    if (w[0] < 1e-8) {
        g_err = "Overlap matrix is near-singular";
        return -1;
    }
    Mat X((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0;
            for (int k = 0; k < n; ++k) s += U[(size_t)i * n + k] / std::sqrt(w[k]) * U[(size_t)j * n + k];
            X[(size_t)i * n + j] = s;
        }
    const Mat Xt = transpose(n, X);
    auto density_from = [&](const Mat &F) {
        Mat Fp = matmul(n, Xt, matmul(n, F, X));
        std::vector<double> e(n);
        Mat Cp((size_t)n * n);
        syev(n, Fp.data(), e.data(), Cp.data());
        Mat C = matmul(n, X, Cp);
        Mat P((size_t)n * n, 0.0);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                double s = 0;
                for (int k = 0; k < nocc; ++k) s += C[(size_t)i * n + k] * C[(size_t)j * n + k];
                P[(size_t)i * n + j] = 2.0 * s;
            }
        return P;
    };
    Mat P = density_from(H);  // scf.cpp:33-45
    std::vector<Mat> fh, eh;
    double Eprev = 0.0;
    Mat J((size_t)n * n), K((size_t)n * n);
    for (int it = 1; it <= max_cycles; ++it) {
        pkref_fock_jk(c, P.data(), J.data(), K.data(), 0);
        Mat F((size_t)n * n);
        for (size_t x = 0; x < F.size(); ++x) F[x] = H[x] + J[x] - 0.5 * K[x];
        double Eg = 0.0;
        for (size_t x = 0; x < F.size(); ++x) Eg += P[x] * (H[x] + F[x]);
        const double Etot = 0.5 * Eg + Enuc;
        if (hist && it - 1 < hist_len) hist[it - 1] = Etot;
        // e = X^T (FPS - SPF) X
        Mat FPS = matmul(n, matmul(n, F, P), S), SPF = matmul(n, matmul(n, S, P), F);
        for (size_t x = 0; x < FPS.size(); ++x) FPS[x] -= SPF[x];
        Mat e = matmul(n, Xt, matmul(n, FPS, X));
        fh.push_back(F);
        eh.push_back(e);
        if ((int)fh.size() > diis_dim) {
            fh.erase(fh.begin());
            eh.erase(eh.begin());
        }
        Mat Fd = F;
        if (fh.size() >= 2) {
            const int m = (int)fh.size();
            std::vector<double> B((size_t)(m + 1) * (m + 1), 0.0), rhs(m + 1, 0.0);
            for (int i = 0; i < m; ++i) {
                for (int j = i; j < m; ++j) {
                    double s = 0;
                    for (size_t x = 0; x < eh[i].size(); ++x) s += eh[i][x] * eh[j][x];
                    B[(size_t)i * (m + 1) + j] = B[(size_t)j * (m + 1) + i] = s;
                }
                B[(size_t)i * (m + 1) + m] = B[(size_t)m * (m + 1) + i] = -1.0;
            }
            rhs[m] = -1.0;
            std::vector<double> cc = qr_colpiv_solve(m + 1, B, rhs);
            std::fill(Fd.begin(), Fd.end(), 0.0);
            for (int i = 0; i < m; ++i)
                for (size_t x = 0; x < Fd.size(); ++x) Fd[x] += cc[i] * fh[i][x];
        }
        Mat Pn = density_from(Fd);
        double dmax = 0, dss = 0;
        for (size_t x = 0; x < Pn.size(); ++x) {
            const double d = Pn[x] - P[x];
            dmax = std::max(dmax, std::fabs(d));
            dss += d * d;
        }
        const double drms = std::sqrt(dss / (double)Pn.size());
        const double dE = std::fabs(Etot - Eprev);
        P = Pn;
        Eprev = Etot;
        if (it > 1 && dE < tol_e && drms < tol_d && dmax < tol_d) {
            *energy = Etot;
            *iterations = it;
            if (P_out) std::copy(P.begin(), P.end(), P_out);
            return 0;
        }
    }
    *energy = Eprev;
    *iterations = max_cycles;
    if (P_out) std::copy(P.begin(), P.end(), P_out);
    g_err = "SCF did not converge";
    return 1;
}

}  // extern "C"
