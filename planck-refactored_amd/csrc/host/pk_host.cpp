// pk_host.cpp — basis construction, one-electron integrals and nuclear
// repulsion for the host side of the engine (include/pk_scf.h).
//
// Reference behaviour followed:
//   Gaussian94 reader, SP split, D exponents ........ src/basis/gaussian.cpp:54-163
//   per-atom shells in file order, Nc scaling ...... src/basis/gaussian.cpp:183-302
//   primitive / contracted normalisation ........... src/basis/basis.cpp:39-75
//   Cartesian component order and norms ............ src/basis/basis.cpp:20-37, gaussian.cpp:286-298
//   nuclear repulsion ............................... src/base/types.h:934-964
// The one-electron integrals are computed on the device (pk_one.hip).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "pk_host_internal.h"

namespace pkh {

thread_local std::string g_err;

int fail(int code, const std::string &m) {
    g_err = m;
    return code;
}

namespace {

const char *kSym[] = {"",   "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg",
                      "Al", "Si", "P",  "S",  "Cl", "Ar", "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn",
                      "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr"};
constexpr int kMaxZ = 36;

struct RawShell {
    int l;
    std::vector<double> e, c;
};

std::string fortran_exp(std::string s) {
    for (char &ch : s)
        if (ch == 'D' || ch == 'd') ch = 'E';
    return s;
}

int shell_l(const std::string &lab) {
    static const char *L = "SPDFGH";
    if (lab.size() != 1) return -1;
    const char *p = std::strchr(L, lab[0]);
    return p ? (int)(p - L) : -1;
}

bool parse_gbs(const char *path, std::map<std::string, std::vector<RawShell>> &out) {
    std::ifstream f(path);
    if (!f) {
        g_err = std::string("cannot open basis file: ") + path;
        return false;
    }
    std::string line, elem;
    while (std::getline(f, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty() || line[0] == '!') continue;
        if (line == "****") {
            elem.clear();
            continue;
        }
        std::istringstream hs(line);
        std::string tok;
        int charge = 0;
        if ((hs >> tok >> charge) && hs.eof()) {
            elem = tok;
            out[elem];
            continue;
        }
        if (elem.empty()) {
            g_err = "basis file: shell before element header";
            return false;
        }
        std::istringstream ss(line);
        std::string lab;
        int np = 0;
        double scale = 1.0;
        ss >> lab >> np >> scale;
        const bool sp = lab == "SP";
        if (!sp && shell_l(lab) < 0) {
            g_err = "basis file: malformed shell line: " + line;
            return false;
        }
        RawShell a{sp ? 0 : shell_l(lab), {}, {}}, b{1, {}, {}};
        for (int i = 0; i < np; ++i) {
            if (!std::getline(f, line)) {
                g_err = "basis file: truncated shell";
                return false;
            }
            std::istringstream ps(fortran_exp(line));
            double ex, c1, c2 = 0.0;
            ps >> ex >> c1;
            if (sp) ps >> c2;
            a.e.push_back(ex);
            a.c.push_back(c1 * scale);
            if (sp) {
                b.e.push_back(ex);
                b.c.push_back(c2 * scale);
            }
        }
        out[elem].push_back(a);
        if (sp) out[elem].push_back(b);
    }
    return true;
}

int dfac(int n) {
    int r = 1;
    for (; n > 0; n -= 2) r *= n;
    return r;
}

int ncart(int l) { return (l + 1) * (l + 2) / 2; }
void cart(int l, int k, int &x, int &y, int &z) {
    int lx = l;
    while (k > l - lx) {
        k -= l - lx + 1;
        --lx;
    }
    x = lx;
    z = k;
    y = l - lx - k;
}

}  // namespace

int load_basis(const char *path, int natoms, const int32_t *Z, const double *xyz, Basis &b) {
    std::map<std::string, std::vector<RawShell>> gbs;
    if (!parse_gbs(path, gbs)) return PK_ERR_ARG;
    b = Basis{};
    b.Z.assign(Z, Z + natoms);
    b.xyz.assign(xyz, xyz + 3 * natoms);
    for (int a = 0; a < natoms; ++a) {
        if (Z[a] < 1 || Z[a] > kMaxZ) return fail(PK_ERR_ARG, "unsupported element Z=" + std::to_string(Z[a]));
        auto it = gbs.find(kSym[Z[a]]);
        if (it == gbs.end()) return fail(PK_ERR_ARG, std::string("Unsupported Element: ") + kSym[Z[a]]);
        for (const RawShell &r : it->second) {
            const int l = r.l, np = (int)r.e.size();
            const double pi15 = std::pow(M_PI, 1.5);
            std::vector<double> nrm(np);
            for (int i = 0; i < np; ++i) nrm[i] = std::sqrt(std::pow(2.0, 2 * l + 1.5) * std::pow(r.e[i], l + 1.5) / pi15);
            double s = 0.0;
            for (int i = 0; i < np; ++i)
                for (int j = 0; j < np; ++j)
                    s += r.c[i] * r.c[j] * nrm[i] * nrm[j] * pi15 / (std::pow(2.0, l) * std::pow(r.e[i] + r.e[j], l + 1.5));
            const double Nc = 1.0 / std::sqrt(s);
            for (int q = 0; q < 3; ++q) b.center.push_back(xyz[3 * a + q]);
            b.l.push_back(l);
            b.nprim.push_back(np);
            b.atom.push_back(a);
            for (int i = 0; i < np; ++i) {
                b.exponent.push_back(r.e[i]);
                b.coef.push_back(r.c[i] * Nc);
                b.prim_norm.push_back(nrm[i]);
            }
            for (int k = 0; k < ncart(l); ++k) {
                int x, y, z;
                cart(l, k, x, y, z);
                b.fnorm.push_back(1.0 / std::sqrt((double)(dfac(2 * x - 1) * dfac(2 * y - 1) * dfac(2 * z - 1))));
            }
        }
    }
    b.nb = (int)b.fnorm.size();
    return PK_OK;
}

double nuclear_repulsion(const Basis &b) {
    double E = 0.0;
    const int n = (int)b.Z.size();
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            const double dx = b.xyz[3 * i] - b.xyz[3 * j], dy = b.xyz[3 * i + 1] - b.xyz[3 * j + 1],
                         dz = b.xyz[3 * i + 2] - b.xyz[3 * j + 2];
            E += (double)b.Z[i] * b.Z[j] / std::sqrt(dx * dx + dy * dy + dz * dz);
        }
    return E;
}

}  // namespace pkh

// ====================================================================== C ABI
namespace pkh {

Basis atom_basis(const Basis &b, int a) {
    Basis o;
    o.Z = {b.Z[a]};
    o.xyz = {0.0, 0.0, 0.0};
    size_t poff = 0, f = 0;
    for (size_t s = 0; s < b.l.size(); ++s) {
        const int np = b.nprim[s], nc = (b.l[s] + 1) * (b.l[s] + 2) / 2;
        if (b.atom[s] == a) {
            o.l.push_back(b.l[s]);
            o.nprim.push_back(np);
            o.atom.push_back(0);
            for (int q = 0; q < 3; ++q) o.center.push_back(0.0);
            for (int k = 0; k < np; ++k) {
                o.exponent.push_back(b.exponent[poff + k]);
                o.coef.push_back(b.coef[poff + k]);
                o.prim_norm.push_back(b.prim_norm[poff + k]);
            }
            for (int k = 0; k < nc; ++k) o.fnorm.push_back(b.fnorm[f + k]);
            o.nb += nc;
        }
        poff += np;
        f += nc;
    }
    return o;
}

void jacobi_eigh(int n, std::vector<double> A, std::vector<double> &w, std::vector<double> &V) {
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                const double v = A[(size_t)i * n + j] * A[(size_t)i * n + j];
                tot += v;
                if (i != j) off += v;
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[(size_t)p * n + q];
                if (apq == 0.0) continue;
                const double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {  // columns p, q
                    const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = c * akp - s * akq;
                    A[(size_t)k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {  // rows p, q
                    const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = c * apk - s * aqk;
                    A[(size_t)q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
                    V[(size_t)k * n + p] = c * vkp - s * vkq;
                    V[(size_t)k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    std::vector<int> ord(n);
    for (int i = 0; i < n; ++i) ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&](int x, int y) { return A[(size_t)x * n + x] < A[(size_t)y * n + y]; });
    w.resize(n);
    std::vector<double> Vs((size_t)n * n);
    for (int k = 0; k < n; ++k) {
        w[k] = A[(size_t)ord[k] * n + ord[k]];
        for (int i = 0; i < n; ++i) Vs[(size_t)i * n + k] = V[(size_t)i * n + ord[k]];
    }
    V.swap(Vs);
}

}  // namespace pkh

extern "C" {

struct pk_basis {
    pkh::Basis b;
};

const char *pk_host_last_error(void) { return pkh::g_err.c_str(); }

int pk_basis_load(const char *path, int natoms, const int32_t *Z, const double *xyz, pk_basis **out) {
    if (!path || !Z || !xyz || !out || natoms <= 0) return pkh::fail(PK_ERR_ARG, "pk_basis_load: bad arguments");
    auto *p = new pk_basis;
    const int rc = pkh::load_basis(path, natoms, Z, xyz, p->b);
    if (rc != PK_OK) {
        delete p;
        return rc;
    }
    for (int l : p->b.l)
        if (l > 3) {
            delete p;
            return pkh::fail(PK_ERR_UNSUPPORTED, "pk_basis_load: shells above f are not supported");
        }
    *out = p;
    return PK_OK;
}

void pk_basis_free(pk_basis *b) { delete b; }

int pk_basis_nbasis(const pk_basis *b) { return b ? b->b.nb : -1; }

int pk_basis_describe(const pk_basis *b, pk_basis_desc *d) {
    if (!b || !d) return pkh::fail(PK_ERR_ARG, "pk_basis_describe: null argument");
    d->nshells = (int32_t)b->b.l.size();
    d->center = b->b.center.data();
    d->l = b->b.l.data();
    d->nprim = b->b.nprim.data();
    d->exponent = b->b.exponent.data();
    d->coef = b->b.coef.data();
    d->prim_norm = b->b.prim_norm.data();
    return PK_OK;
}

int pk_one_electron(const pk_basis *b, double *S, double *T, double *V) {
    if (!b) return pkh::fail(PK_ERR_ARG, "pk_one_electron: null basis");
    return pkh::one_electron(b->b, S, T, V);
}

double pk_nuclear_repulsion(const pk_basis *b) { return b ? pkh::nuclear_repulsion(b->b) : 0.0; }

}  // extern "C"

const pkh::Basis &pkh_basis_of(const pk_basis *b) { return b->b; }
