// pk_scf.hip — RHF direct SCF driver on the device (include/pk_scf.h).
//
// Restates SCF::run_rhf (src/scf/scf.cpp:175-441): Loewdin orthogonaliser
// X = S^{-1/2} (scf.cpp:15-29, near-singular below 1e-8), core-Hamiltonian guess
// (scf.cpp:33-45), per iteration G = J - K/2 from the Fock engine, F = H + G,
// E = 1/2 Tr[P (H + F)] + E_nuc, DIIS on e = X^T (F P S - S P F) X with at most
// diis_dim vectors and extrapolation from two vectors on (types.h:759-837),
// diagonalisation of X^T F X, P = 2 C_occ C_occ^T, convergence when
// iteration > 1 and dE, rms dP, max dP are all below their tolerances
// (scf.cpp:89-135).  Every nb x nb matrix stays in device memory: GEMMs by
// rocBLAS, the symmetric eigensolver by rocSOLVER (syevd); only the DIIS
// system (at most 9 x 9) is solved on the host.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "pk_host_internal.h"

namespace {

__global__ void k_add(const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ c, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c[i] = a[i] + b[i];
}

// D = M - M^T (column-major n x n)
__global__ void k_antisym(const double *__restrict__ M, double *__restrict__ D, int n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * n) return;
    const int r = (int)(i % n), c = (int)(i / n);
    D[i] = M[i] - M[(size_t)r * n + c];
}

// U[:, k] *= 1 / sqrt(w[k])
__global__ void k_scale_cols(double *__restrict__ U, const double *__restrict__ w, int n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * n) return;
    U[i] /= sqrt(w[i / n]);
}

// per-block max |a - b| and sum (a - b)^2
__global__ void k_diff(const double *__restrict__ a, const double *__restrict__ b, size_t n, double *__restrict__ out) {
    __shared__ double smax[256], ssum[256];
    double m = 0.0, s = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double d = a[i] - b[i];
        m = fmax(m, fabs(d));
        s += d * d;
    }
    smax[threadIdx.x] = m;
    ssum[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + o]);
            ssum[threadIdx.x] += ssum[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = smax[0];
        out[2 * blockIdx.x + 1] = ssum[0];
    }
}

// ---- molecule-table kernels of the batched SCF: blockIdx.y selects an entry
// of the tables (one molecule, or one molecule and DIIS slot), blockIdx.x and
// the threads stride over its n x n elements; reductions leave one partial
// per block, summed on the host.
__device__ __forceinline__ void block_sum2(double a, double b, double *out) {
    __shared__ double sa[256], sb[256];
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sa[threadIdx.x] += sa[threadIdx.x + o];
            sb[threadIdx.x] += sb[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sa[0];
        out[1] = sb[0];
    }
}

// F = H + G; partials of sum P.H and sum P.F
__global__ void kb_fock(const double *const *H, const double *const *G, double *const *F, const double *const *P,
                        const long long *n2, double *part) {
    const int y = blockIdx.y;
    const size_t n = (size_t)n2[y];
    const double *h = H[y], *g = G[y], *p = P[y];
    double *f = F[y];
    double eh = 0.0, ef = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double fi = h[i] + g[i];
        f[i] = fi;
        eh = fma(p[i], h[i], eh);
        ef = fma(p[i], fi, ef);
    }
    block_sum2(eh, ef, part + 2 * ((size_t)y * gridDim.x + blockIdx.x));
}

// D = M - M^T (column-major n x n)
__global__ void kb_antisym(const double *const *M, double *const *D, const int *nn) {
    const int y = blockIdx.y, n = nn[y];
    const double *m = M[y];
    double *d = D[y];
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < (size_t)n * n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(i % n), c = (int)(i / n);
        d[i] = m[i] - m[(size_t)r * n + c];
    }
}

// partials of a . b (second partial unused)
__global__ void kb_dots(const double *const *A, const double *const *B, const long long *n2, double *part) {
    const int y = blockIdx.y;
    const size_t n = (size_t)n2[y];
    const double *a = A[y], *b = B[y];
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        s = fma(a[i], b[i], s);
    block_sum2(s, 0.0, part + 2 * ((size_t)y * gridDim.x + blockIdx.x));
}

// out = sum_k c[kmax y + k] src[kmax y + k], k < ns[y]
__global__ void kb_combine(const double *const *src, const double *coef, const int *ns, int kmax,
                           double *const *out, const long long *n2) {
    const int y = blockIdx.y, k1 = ns[y];
    const size_t n = (size_t)n2[y];
    double *o = out[y];
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double v = 0.0;
        for (int k = 0; k < k1; ++k) v = fma(coef[(size_t)kmax * y + k], src[(size_t)kmax * y + k][i], v);
        o[i] = v;
    }
}

// partials of max |Pn - P| and sum (Pn - P)^2, then P = Pn
__global__ void kb_diff(const double *const *Pn, double *const *P, const long long *n2, double *part) {
    __shared__ double smax[256], ssum[256];
    const int y = blockIdx.y;
    const size_t n = (size_t)n2[y];
    const double *a = Pn[y];
    double *b = P[y];
    double m = 0.0, s = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double d = a[i] - b[i];
        m = fmax(m, fabs(d));
        s = fma(d, d, s);
        b[i] = a[i];
    }
    smax[threadIdx.x] = m;
    ssum[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + o]);
            ssum[threadIdx.x] += ssum[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * ((size_t)y * gridDim.x + blockIdx.x)] = smax[0];
        part[2 * ((size_t)y * gridDim.x + blockIdx.x) + 1] = ssum[0];
    }
}

// small column-pivoted Householder QR solve (the DIIS system, types.h:813)
std::vector<double> qr_solve(int n, std::vector<double> A, std::vector<double> b) {
    auto a = [&](int i, int j) -> double & { return A[(size_t)i * n + j]; };
    std::vector<int> perm(n);
    std::vector<double> cn(n), diag(n, 0.0);
    for (int j = 0; j < n; ++j) {
        perm[j] = j;
        double s = 0;
        for (int i = 0; i < n; ++i) s += a(i, j) * a(i, j);
        cn[j] = s;
    }
    double maxp = 0;
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int j = k + 1; j < n; ++j)
            if (cn[j] > cn[p]) p = j;
        if (p != k) {
            for (int i = 0; i < n; ++i) std::swap(a(i, k), a(i, p));
            std::swap(cn[k], cn[p]);
            std::swap(perm[k], perm[p]);
        }
        double nr = 0;
        for (int i = k; i < n; ++i) nr += a(i, k) * a(i, k);
        nr = std::sqrt(nr);
        if (nr > 0) {
            const double alpha = a(k, k) > 0 ? -nr : nr;
            std::vector<double> v(n, 0.0);
            for (int i = k; i < n; ++i) v[i] = a(i, k);
            v[k] -= alpha;
            double vv = 0;
            for (int i = k; i < n; ++i) vv += v[i] * v[i];
            if (vv > 0) {
                for (int j = k; j < n; ++j) {
                    double s = 0;
                    for (int i = k; i < n; ++i) s += v[i] * a(i, j);
                    s *= 2 / vv;
                    for (int i = k; i < n; ++i) a(i, j) -= s * v[i];
                }
                double s = 0;
                for (int i = k; i < n; ++i) s += v[i] * b[i];
                s *= 2 / vv;
                for (int i = k; i < n; ++i) b[i] -= s * v[i];
            }
        }
        diag[k] = a(k, k);
        maxp = std::max(maxp, std::fabs(diag[k]));
        for (int j = k + 1; j < n; ++j) {
            double s = 0;
            for (int i = k + 1; i < n; ++i) s += a(i, j) * a(i, j);
            cn[j] = s;
        }
    }
    const double thr = std::ldexp(1.0, -52) * n * maxp;
    int rank = 0;
    for (int k = 0; k < n; ++k)
        if (std::fabs(diag[k]) > thr) ++rank;
    std::vector<double> y(n, 0.0), x(n, 0.0);
    for (int k = rank - 1; k >= 0; --k) {
        double s = b[k];
        for (int j = k + 1; j < rank; ++j) s -= a(k, j) * y[j];
        y[k] = s / a(k, k);
    }
    for (int k = 0; k < n; ++k) x[perm[k]] = y[k];
    return x;
}

struct Dev {
    std::vector<void *> allocs;
    ~Dev() {
        for (void *p : allocs) (void)hipFree(p);
    }
    double *alloc(size_t n) {
        void *p = nullptr;
        if (hipMalloc(&p, n * sizeof(double)) != hipSuccess) return nullptr;
        allocs.push_back(p);
        return (double *)p;
    }
};

#define SCF_HIP(x)                                                                          \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) {                                                             \
            rc = pkh::fail(PK_ERR_HIP, std::string(#x ": ") + hipGetErrorString(e_));       \
            goto done;                                                                      \
        }                                                                                   \
    } while (0)
#define SCF_RB(x)                                                                           \
    do {                                                                                    \
        rocblas_status s_ = (x);                                                            \
        if (s_ != rocblas_status_success) {                                                 \
            rc = pkh::fail(PK_ERR_HIP, std::string(#x " failed: ") + rocblas_status_to_string(s_)); \
            goto done;                                                                      \
        }                                                                                   \
    } while (0)
#define SCF_PK(x)                                                                           \
    do {                                                                                    \
        int r_ = (x);                                                                       \
        if (r_ != PK_OK) {                                                                  \
            rc = pkh::fail(r_, std::string(#x ": ") + pk_last_error());                     \
            goto done;                                                                      \
        }                                                                                   \
    } while (0)


// ---------------------------------------------------------------- SAD guess
// compute_sad_guess_rhf (src/scf/sad.cpp:589-643): atomic UHF per element,
// spherical average per shell, block-diagonal assembly; the projection on an
// RHF density runs on the device in pk_scf_rhf.

using Mat = std::vector<double>;  // row-major n x n

Mat mm(int n, const Mat &A, const Mat &B, bool ta = false, bool tb = false) {
    Mat C((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) {
            const double a = ta ? A[(size_t)k * n + i] : A[(size_t)i * n + k];
            if (a == 0.0) continue;
            for (int j = 0; j < n; ++j) C[(size_t)i * n + j] += a * (tb ? B[(size_t)j * n + k] : B[(size_t)k * n + j]);
        }
    return C;
}

// symmetric f(A) = U diag(f(w)) U^T through the Jacobi eigensolver
template <typename Fn>
Mat sym_fun(int n, const Mat &A, Fn f) {
    std::vector<double> w, V;
    pkh::jacobi_eigh(n, A, w, V);
    Mat R((size_t)n * n, 0.0);
    for (int k = 0; k < n; ++k) {
        const double fk = f(w[k]);
        if (fk == 0.0) continue;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) R[(size_t)i * n + j] += V[(size_t)i * n + k] * fk * V[(size_t)j * n + k];
    }
    return R;
}

// atomic UHF (sad.cpp:284-356 through scf.cpp:467-830): multiplicity 1 for even
// Z, 2 for odd Z; core guess per spin; DIIS per spin (6 vectors, restart when
// the error grows 2x after iteration 2); converged when iteration > 1 and |dE|,
// rms and max change of Pa + Pb are below 1e-10.  J and K from the engine.
int atomic_uhf(const pkh::Basis &ab, int engine, double tol_eri, Mat &Pt, Mat &S) {
    const int n = ab.nb;
    const size_t n2 = (size_t)n * n;
    S.assign(n2, 0.0);
    Mat T(n2), V(n2), H(n2);
    pkh::one_electron(ab, S.data(), T.data(), V.data());
    for (size_t i = 0; i < n2; ++i) H[i] = T[i] + V[i];
    const int Z = ab.Z[0], unp = Z % 2, na = (Z + unp) / 2, nbeta = (Z - unp) / 2;
    bool singular = false;
    const Mat X = sym_fun(n, S, [&](double w) {
        if (w < 1e-8) singular = true;
        return w > 0 ? 1.0 / std::sqrt(w) : 0.0;
    });
    if (singular) return pkh::fail(PK_ERR_ARG, "SAD: atomic overlap matrix is near-singular");
    auto dens = [&](const Mat &F, int nocc) {
        std::vector<double> w, U;
        pkh::jacobi_eigh(n, mm(n, mm(n, X, F), X), w, U);
        const Mat C = mm(n, X, U);
        Mat P(n2, 0.0);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                double s = 0;
                for (int k = 0; k < nocc; ++k) s += C[(size_t)i * n + k] * C[(size_t)j * n + k];
                P[(size_t)i * n + j] = s;
            }
        return P;
    };
    Mat Pa = dens(H, na), Pb = dens(H, nbeta);
    pk_basis_desc d{(int32_t)ab.l.size(), ab.center.data(), ab.l.data(), ab.nprim.data(), ab.exponent.data(),
                    ab.coef.data(), ab.prim_norm.data()};
    pk_ctx *ctx = nullptr;
    if (pk_ctx_create(&d, &ctx) != PK_OK) return pkh::fail(PK_ERR_HIP, std::string("SAD: ") + pk_last_error());
    int rc = PK_OK;
    if (pk_ctx_set(ctx, engine, tol_eri, PK_PREC_FP64, 0.0) != PK_OK) rc = pkh::fail(PK_ERR_HIP, pk_last_error());
    std::deque<std::pair<Mat, Mat>> hist[2];  // (F, e) per spin
    double Eprev = 0.0, err_prev = 1e300;
    bool conv = false;
    Mat Ja(n2), Ka(n2), Jb(n2), Kb(n2);
    for (int it = 1; it <= 200 && rc == PK_OK && !conv; ++it) {
        if (pk_fock_jk(ctx, Pa.data(), Ja.data(), Ka.data(), PK_MEM_HOST) != PK_OK ||
            pk_fock_jk(ctx, Pb.data(), Jb.data(), Kb.data(), PK_MEM_HOST) != PK_OK) {
            rc = pkh::fail(PK_ERR_HIP, pk_last_error());
            break;
        }
        Mat F[2] = {Mat(n2), Mat(n2)};
        double E = 0.0;
        for (size_t i = 0; i < n2; ++i) {
            F[0][i] = H[i] + Ja[i] + Jb[i] - Ka[i];
            F[1][i] = H[i] + Ja[i] + Jb[i] - Kb[i];
            E += 0.5 * (Pa[i] * (H[i] + F[0][i]) + Pb[i] * (H[i] + F[1][i]));
        }
        const Mat *Ps[2] = {&Pa, &Pb};
        Mat e[2];
        double cur = 0.0;
        for (int sp = 0; sp < 2; ++sp) {
            const Mat FPS = mm(n, mm(n, F[sp], *Ps[sp]), S), SPF = mm(n, mm(n, S, *Ps[sp]), F[sp]);
            Mat D(n2);
            for (size_t i = 0; i < n2; ++i) D[i] = FPS[i] - SPF[i];
            e[sp] = mm(n, mm(n, X, D), X);
            double ss = 0;
            for (double v : e[sp]) ss += v * v;
            cur = std::max(cur, std::sqrt(ss / (double)n2));
        }
        if (it > 2 && cur > 2.0 * err_prev) hist[0].clear(), hist[1].clear();
        err_prev = cur;
        Mat Pn[2];
        for (int sp = 0; sp < 2; ++sp) {
            hist[sp].emplace_back(F[sp], e[sp]);
            if (hist[sp].size() > 6) hist[sp].pop_front();
            Mat Fd = F[sp];
            const int m = (int)hist[sp].size();
            if (m >= 2) {
                std::vector<double> A((size_t)(m + 1) * (m + 1), 0.0), rhs(m + 1, 0.0);
                for (int a = 0; a < m; ++a) {
                    for (int b2 = 0; b2 < m; ++b2) {
                        double v = 0;
                        for (size_t k = 0; k < n2; ++k) v += hist[sp][a].second[k] * hist[sp][b2].second[k];
                        A[(size_t)a * (m + 1) + b2] = v;
                    }
                    A[(size_t)a * (m + 1) + m] = A[(size_t)m * (m + 1) + a] = -1.0;
                }
                rhs[m] = -1.0;
                const std::vector<double> c = qr_solve(m + 1, A, rhs);
                std::fill(Fd.begin(), Fd.end(), 0.0);
                for (int a = 0; a < m; ++a)
                    for (size_t k = 0; k < n2; ++k) Fd[k] += c[a] * hist[sp][a].first[k];
            }
            Pn[sp] = dens(Fd, sp == 0 ? na : nbeta);
        }
        double dmax = 0, dss = 0;
        for (size_t i = 0; i < n2; ++i) {
            const double dd = (Pn[0][i] + Pn[1][i]) - (Pa[i] + Pb[i]);
            dmax = std::max(dmax, std::fabs(dd));
            dss += dd * dd;
        }
        Pa.swap(Pn[0]);
        Pb.swap(Pn[1]);
        const double dE = std::fabs(E - Eprev);
        Eprev = E;
        conv = it > 1 && dE < 1e-10 && std::sqrt(dss / (double)n2) < 1e-10 && dmax < 1e-10;
    }
    pk_ctx_destroy(ctx);
    if (rc != PK_OK) return rc;
    if (!conv) return pkh::fail(PK_ERR_ARG, "SAD: atomic UHF for Z=" + std::to_string(Z) + " did not converge");
    Pt.assign(n2, 0.0);
    for (size_t i = 0; i < n2; ++i) Pt[i] = Pa[i] + Pb[i];
    return PK_OK;
}

// spherical average (sad.cpp:367-430): Mulliken shell populations, cross-shell
// blocks zeroed, each shell block (N_l / g) pinv(S_shell, 1e-10)
Mat spherical_average(const pkh::Basis &ab, const Mat &P, const Mat &S) {
    const int n = ab.nb;
    const Mat PS = mm(n, P, S);
    Mat out((size_t)n * n, 0.0);
    int f0 = 0;
    for (size_t s = 0; s < ab.l.size(); ++s) {
        const int g = (ab.l[s] + 1) * (ab.l[s] + 2) / 2;
        double N = 0;
        for (int i = 0; i < g; ++i) N += PS[(size_t)(f0 + i) * n + f0 + i];
        Mat Sb((size_t)g * g);
        for (int i = 0; i < g; ++i)
            for (int j = 0; j < g; ++j) Sb[(size_t)i * g + j] = S[(size_t)(f0 + i) * n + f0 + j];
        const Mat Pinv = sym_fun(g, Sb, [](double w) { return w > 1e-10 ? 1.0 / w : 0.0; });
        for (int i = 0; i < g; ++i)
            for (int j = 0; j < g; ++j) out[(size_t)(f0 + i) * n + f0 + j] = N / g * Pinv[(size_t)i * g + j];
        f0 += g;
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) {
            const double v = 0.5 * (out[(size_t)i * n + j] + out[(size_t)j * n + i]);
            out[(size_t)i * n + j] = out[(size_t)j * n + i] = v;
        }
    return out;
}

// block-diagonal raw SAD density in molecular AO order (sad.cpp:436-470)
int sad_raw(const pkh::Basis &b, int engine, double tol_eri, Mat &P) {
    const int n = b.nb, natoms = (int)b.Z.size();
    std::vector<std::vector<int>> ao(natoms);
    int f = 0;
    for (size_t s = 0; s < b.l.size(); ++s) {
        const int g = (b.l[s] + 1) * (b.l[s] + 2) / 2;
        for (int i = 0; i < g; ++i) ao[b.atom[s]].push_back(f + i);
        f += g;
    }
    std::map<int, Mat> cache;
    P.assign((size_t)n * n, 0.0);
    for (int a = 0; a < natoms; ++a) {
        const int z = b.Z[a];
        if (!cache.count(z)) {
            const pkh::Basis ab = pkh::atom_basis(b, a);
            Mat Pt, Sa;
            const int rc = atomic_uhf(ab, engine, tol_eri, Pt, Sa);
            if (rc != PK_OK) return rc;
            cache[z] = spherical_average(ab, Pt, Sa);
        }
        const Mat &B = cache[z];
        const int g = (int)ao[a].size();
        for (int i = 0; i < g; ++i)
            for (int j = 0; j < g; ++j) P[(size_t)ao[a][i] * n + ao[a][j]] = B[(size_t)i * g + j];
    }
    return PK_OK;
}

}  // namespace

extern "C" {

void pk_scf_default_options(pk_scf_options *o) {
    o->tol_energy = 1e-10;
    o->tol_density = 1e-10;
    o->tol_eri = 1e-10;
    o->diis_dim = 8;
    o->max_cycles = 0;
    o->engine = PK_ENGINE_OS;
    o->charge = 0;
    o->use_guess = 0;
    o->guess_density = nullptr;
    o->precision = PK_PREC_FP64;
    o->fp32_cutoff = 0.0;
    o->guess_fallback = 0;
    o->sad_guess = 0;
    o->guess_only = 0;
}

int pk_scf_rhf(const pk_basis *bas, const pk_scf_options *opt, pk_scf_result *res, double *P_out, double *hist,
               int hist_len) {
    if (!bas || !opt || !res) return pkh::fail(PK_ERR_ARG, "pk_scf_rhf: null argument");
    const auto t_start = std::chrono::steady_clock::now();
    const pkh::Basis &b = pkh_basis_of(bas);
    const int n = b.nb;
    const size_t n2 = (size_t)n * n;
    int nel = -opt->charge;
    for (int z : b.Z) nel += z;
    if (nel % 2 != 0) return pkh::fail(PK_ERR_ARG, "RHF requires an even number of electrons (closed shell)");
    const int nocc = nel / 2;
    const int max_cycles = opt->max_cycles > 0 ? opt->max_cycles : (n > 1000 ? 300 : n > 500 ? 200 : n > 250 ? 100 : 50);
    const int diis_dim = opt->diis_dim > 0 ? opt->diis_dim : 8;
    *res = pk_scf_result{};
    res->nuclear_repulsion = pkh::nuclear_repulsion(b);

    int rc = PK_OK;
    Dev dev;
    pk_ctx *ctx = nullptr;
    rocblas_handle hb = nullptr;
    hipStream_t st = nullptr;
    std::vector<double> w(n);
    double *S, *H, *X, *P, *Pn, *F, *G, *M, *D, *W1, *C, *ev, *ee, *part;
    rocblas_int *info;
    std::deque<int> slots;  // DIIS ring: slot indices, oldest first
    std::vector<double *> Fh, Eh;
    std::vector<std::vector<double>> Bm;  // B over slots
    const double one = 1.0, zero = 0.0, two = 2.0;
    double Eprev = 0.0, Etot = 0.0;
    bool have_G = false;  // G of the current P already built (guess comparison)
    const int g2 = (int)((n2 + 255) / 256);
    const int nparts = 256;
    std::vector<double> hpart(2 * nparts);
    pk_basis_desc desc;
    auto gemm = [&](rocblas_operation ta, rocblas_operation tb, int k, double alpha, const double *A, const double *Bm_,
                    double *Cm) {
        return rocblas_dgemm(hb, ta, tb, n, n, k, &alpha, A, n, Bm_, n, &zero, Cm, n);
    };

    // density from a Fock-like matrix Fm: P_target = 2 C_occ C_occ^T with C = X eig(X Fm X)
    auto density = [&](const double *Fm, double *Pt) -> int {
        if (gemm(rocblas_operation_none, rocblas_operation_none, n, 1.0, X, Fm, M) != rocblas_status_success ||
            gemm(rocblas_operation_none, rocblas_operation_none, n, 1.0, M, X, W1) != rocblas_status_success)
            return pkh::fail(PK_ERR_HIP, "gemm failed");
        if (rocsolver_dsyevd(hb, rocblas_evect_original, rocblas_fill_upper, n, W1, n, ev, ee, info) !=
            rocblas_status_success)
            return pkh::fail(PK_ERR_HIP, "syevd failed");
        if (gemm(rocblas_operation_none, rocblas_operation_none, n, 1.0, X, W1, C) != rocblas_status_success)
            return pkh::fail(PK_ERR_HIP, "gemm failed");
        if (rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_transpose, n, n, nocc, &two, C, n, C, n, &zero,
                          Pt, n) != rocblas_status_success)
            return pkh::fail(PK_ERR_HIP, "gemm failed");
        return PK_OK;
    };

    SCF_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    SCF_RB(rocblas_create_handle(&hb));
    SCF_RB(rocblas_set_stream(hb, st));
    pk_basis_describe(bas, &desc);
    SCF_PK(pk_ctx_create(&desc, &ctx));
    SCF_PK(pk_ctx_set_stream(ctx, st));
    SCF_PK(pk_ctx_set(ctx, opt->engine, opt->tol_eri, opt->precision, opt->fp32_cutoff));

    S = dev.alloc(n2), H = dev.alloc(n2), X = dev.alloc(n2), P = dev.alloc(n2), Pn = dev.alloc(n2);
    F = dev.alloc(n2), G = dev.alloc(n2), M = dev.alloc(n2), D = dev.alloc(n2), W1 = dev.alloc(n2), C = dev.alloc(n2);
    ev = dev.alloc(n), ee = dev.alloc(n), part = dev.alloc(2 * nparts);
    info = (rocblas_int *)dev.alloc(1);
    for (int s = 0; s < diis_dim; ++s) {
        Fh.push_back(dev.alloc(n2));
        Eh.push_back(dev.alloc(n2));
    }
    if (!S || !H || !X || !P || !Pn || !F || !G || !M || !D || !W1 || !C || !ev || !ee || !part || !info ||
        !Fh.back() || !Eh.back()) {
        rc = pkh::fail(PK_ERR_ALLOC, "pk_scf_rhf: device allocation failed");
        goto done;
    }
    Bm.assign(diis_dim, std::vector<double>(diis_dim, 0.0));
    // S, and H = T + V, on the device (T through M, V through G: scratch until the first build)
    if ((rc = pkh::one_electron_device(b, S, M, G, st)) != PK_OK) goto done;
    hipLaunchKernelGGL(k_add, dim3(g2), dim3(256), 0, st, M, G, H, n2);
    // X = U s^{-1/2} U^T
    SCF_HIP(hipMemcpyAsync(W1, S, n2 * 8, hipMemcpyDeviceToDevice, st));
    SCF_RB(rocsolver_dsyevd(hb, rocblas_evect_original, rocblas_fill_upper, n, W1, n, ev, ee, info));
    SCF_HIP(hipMemcpyAsync(w.data(), ev, n * 8, hipMemcpyDeviceToHost, st));
    SCF_HIP(hipStreamSynchronize(st));
    if (w[0] < 1e-8) {
        char msg[128];
        std::snprintf(msg, sizeof msg, "Overlap matrix is near-singular (min eigenvalue = %.3e)", w[0]);
        rc = pkh::fail(PK_ERR_ARG, msg);
        goto done;
    }
    SCF_HIP(hipMemcpyAsync(C, W1, n2 * 8, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_scale_cols, dim3(g2), dim3(256), 0, st, C, ev, n);
    SCF_RB(gemm(rocblas_operation_none, rocblas_operation_transpose, n, 1.0, C, W1, X));

    if (opt->use_guess && opt->guess_density) {
        SCF_HIP(hipMemcpyAsync(P, opt->guess_density, n2 * 8, hipMemcpyHostToDevice, st));
        res->guess_used = 1;
    } else if (opt->sad_guess) {
        // SAD: raw block-diagonal density on the host, then the projection on
        // an RHF density (sad.cpp:476-505): the n_occ eigenvectors of X P X
        // with the largest eigenvalues, P = 2 C C^T with C = X U
        {
            Mat Praw;
            if ((rc = sad_raw(b, opt->engine, opt->tol_eri, Praw)) != PK_OK) goto done;
            SCF_HIP(hipMemcpyAsync(D, Praw.data(), n2 * 8, hipMemcpyHostToDevice, st));
            SCF_HIP(hipStreamSynchronize(st));
        }
        SCF_RB(gemm(rocblas_operation_none, rocblas_operation_none, n, 1.0, X, D, M));
        SCF_RB(gemm(rocblas_operation_none, rocblas_operation_none, n, 1.0, M, X, W1));
        SCF_RB(rocsolver_dsyevd(hb, rocblas_evect_original, rocblas_fill_upper, n, W1, n, ev, ee, info));
        SCF_RB(gemm(rocblas_operation_none, rocblas_operation_none, n, 1.0, X, W1, C));
        SCF_RB(rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_transpose, n, n, nocc, &two,
                             C + (size_t)(n - nocc) * n, n, C + (size_t)(n - nocc) * n, n, &zero, P, n));
    } else if ((rc = density(H, P)) != PK_OK) {
        goto done;
    }
    if (opt->guess_only) {
        if (P_out) SCF_HIP(hipMemcpyAsync(P_out, P, n2 * 8, hipMemcpyDeviceToHost, st));
        SCF_HIP(hipStreamSynchronize(st));
        res->converged = 1;
        goto done;
    }
    if (opt->use_guess && opt->guess_density && opt->guess_fallback) {
        // first DIIS error of each candidate: RMS of X (F P S - S P F) X with
        // F = H + G[P]; the core guess replaces the supplied one when smaller
        auto first_error = [&](const double *Pc, double *Gc, double &err) -> int {
            if (pk_fock_2e(ctx, Pc, Gc, PK_MEM_DEVICE) != PK_OK) return pkh::fail(PK_ERR_HIP, pk_last_error());
            pk_stats s;
            pk_ctx_stats(ctx, &s);
            res->fock_ms += s.last_build_ms;
            hipLaunchKernelGGL(k_add, dim3(g2), dim3(256), 0, st, H, Gc, F, n2);
            if (gemm(rocblas_operation_none, rocblas_operation_none, n, 1.0, F, Pc, W1) != rocblas_status_success ||
                gemm(rocblas_operation_none, rocblas_operation_none, n, 1.0, W1, S, M) != rocblas_status_success)
                return pkh::fail(PK_ERR_HIP, "gemm failed");
            hipLaunchKernelGGL(k_antisym, dim3(g2), dim3(256), 0, st, M, D, n);
            if (gemm(rocblas_operation_none, rocblas_operation_none, n, 1.0, X, D, W1) != rocblas_status_success ||
                gemm(rocblas_operation_none, rocblas_operation_none, n, 1.0, W1, X, M) != rocblas_status_success)
                return pkh::fail(PK_ERR_HIP, "gemm failed");
            double v = 0;
            if (rocblas_set_pointer_mode(hb, rocblas_pointer_mode_host) != rocblas_status_success ||
                rocblas_ddot(hb, (rocblas_int)n2, M, 1, M, 1, &v) != rocblas_status_success)
                return pkh::fail(PK_ERR_HIP, "ddot failed");
            err = std::sqrt(v / (double)n2);
            return PK_OK;
        };
        double eg = 0, ec = 0;
        if ((rc = first_error(P, G, eg)) != PK_OK) goto done;
        if ((rc = density(H, Pn)) != PK_OK) goto done;
        if ((rc = first_error(Pn, C, ec)) != PK_OK) goto done;  // C is free until the first density()
        res->guess_diis_error = eg;
        res->core_diis_error = ec;
        if (ec < eg) {
            std::swap(P, Pn);
            SCF_HIP(hipMemcpyAsync(G, C, n2 * 8, hipMemcpyDeviceToDevice, st));
            res->guess_used = 0;
        }
        have_G = true;
    }

    for (int it = 1; it <= max_cycles; ++it) {
        if (!have_G) {
            SCF_PK(pk_fock_2e(ctx, P, G, PK_MEM_DEVICE));
            pk_stats s;
            pk_ctx_stats(ctx, &s);
            res->fock_ms += s.last_build_ms;
        }
        have_G = false;
        hipLaunchKernelGGL(k_add, dim3(g2), dim3(256), 0, st, H, G, F, n2);
        double eph = 0, epf = 0;
        SCF_RB(rocblas_set_pointer_mode(hb, rocblas_pointer_mode_host));
        SCF_RB(rocblas_ddot(hb, (rocblas_int)n2, P, 1, H, 1, &eph));
        SCF_RB(rocblas_ddot(hb, (rocblas_int)n2, P, 1, F, 1, &epf));
        Etot = 0.5 * (eph + epf) + res->nuclear_repulsion;
        if (hist && it - 1 < hist_len) hist[it - 1] = Etot;
        // e = X (F P S - S P F) X
        SCF_RB(gemm(rocblas_operation_none, rocblas_operation_none, n, 1.0, F, P, W1));
        SCF_RB(gemm(rocblas_operation_none, rocblas_operation_none, n, 1.0, W1, S, M));
        hipLaunchKernelGGL(k_antisym, dim3(g2), dim3(256), 0, st, M, D, n);
        SCF_RB(gemm(rocblas_operation_none, rocblas_operation_none, n, 1.0, X, D, W1));
        int slot;
        if ((int)slots.size() < diis_dim) {
            slot = (int)slots.size();
        } else {
            slot = slots.front();
            slots.pop_front();
        }
        SCF_RB(gemm(rocblas_operation_none, rocblas_operation_none, n, 1.0, W1, X, Eh[slot]));
        SCF_HIP(hipMemcpyAsync(Fh[slot], F, n2 * 8, hipMemcpyDeviceToDevice, st));
        slots.push_back(slot);
        for (int s : slots) {
            double v = 0;
            SCF_RB(rocblas_ddot(hb, (rocblas_int)n2, Eh[slot], 1, Eh[s], 1, &v));
            Bm[slot][s] = Bm[s][slot] = v;
        }
        if (it == 1) res->first_diis_error = std::sqrt(Bm[slot][slot] / (double)n2);
        const double *Fd = F;
        if (slots.size() >= 2) {
            const int m = (int)slots.size();
            std::vector<double> A((size_t)(m + 1) * (m + 1), 0.0), rhs(m + 1, 0.0);
            for (int i = 0; i < m; ++i) {
                for (int j = 0; j < m; ++j) A[(size_t)i * (m + 1) + j] = Bm[slots[i]][slots[j]];
                A[(size_t)i * (m + 1) + m] = A[(size_t)m * (m + 1) + i] = -1.0;
            }
            rhs[m] = -1.0;
            const std::vector<double> cf = qr_solve(m + 1, A, rhs);
            SCF_HIP(hipMemsetAsync(M, 0, n2 * 8, st));
            for (int i = 0; i < m; ++i) SCF_RB(rocblas_daxpy(hb, (rocblas_int)n2, &cf[i], Fh[slots[i]], 1, M, 1));
            SCF_HIP(hipMemcpyAsync(D, M, n2 * 8, hipMemcpyDeviceToDevice, st));
            Fd = D;
        }
        if ((rc = density(Fd, Pn)) != PK_OK) goto done;
        hipLaunchKernelGGL(k_diff, dim3(nparts), dim3(256), 0, st, Pn, P, n2, part);
        SCF_HIP(hipMemcpyAsync(hpart.data(), part, 2 * nparts * 8, hipMemcpyDeviceToHost, st));
        SCF_HIP(hipStreamSynchronize(st));
        double dmax = 0, dss = 0;
        for (int i = 0; i < nparts; ++i) {
            dmax = std::max(dmax, hpart[2 * i]);
            dss += hpart[2 * i + 1];
        }
        const double drms = std::sqrt(dss / (double)n2), dE = std::fabs(Etot - Eprev);
        std::swap(P, Pn);
        Eprev = Etot;
        res->iterations = it;
        res->energy = Etot;
        if (it > 1 && dE < opt->tol_energy && drms < opt->tol_density && dmax < opt->tol_density) {
            res->converged = 1;
            break;
        }
    }
    if (P_out) {
        SCF_HIP(hipMemcpyAsync(P_out, P, n2 * 8, hipMemcpyDeviceToHost, st));
        SCF_HIP(hipStreamSynchronize(st));
    }
    if (!res->converged) rc = pkh::fail(1, "SCF did not converge in " + std::to_string(max_cycles) + " iterations");
done:
    if (st) (void)hipStreamSynchronize(st);
    res->wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    if (ctx) pk_ctx_destroy(ctx);
    if (hb) rocblas_destroy_handle(hb);
    if (st) (void)hipStreamDestroy(st);
    return rc;
}

// Batched RHF: every molecule runs the iteration of pk_scf_rhf, but the Fock
// builds of one iteration are one batched build of all molecules (one context,
// merged per-class quartet lists, hipGraph replay); the per-molecule algebra
// (energy, DIIS, diagonalisation) follows on the same stream.  Converged
// molecules keep their density and stop updating.
int pk_scf_rhf_batch(const pk_basis *const *bases, int nmol, const pk_scf_options *opt, pk_scf_result *res,
                     double *const *P_out, const double *const *guess) {
    if (!bases || nmol <= 0 || !opt || !res) return pkh::fail(PK_ERR_ARG, "pk_scf_rhf_batch: null argument");
    const auto t_start = std::chrono::steady_clock::now();
    struct Mol {
        const pkh::Basis *b;
        int n = 0, nocc = 0;
        size_t n2 = 0;
        double *S, *H, *X, *Pn, *F, *M, *D, *W1, *C, *ev, *ee, *part, *P = nullptr, *G = nullptr;
        rocblas_int *info;
        std::vector<double *> Fh, Eh;
        std::deque<int> slots;
        std::vector<std::vector<double>> Bm;
        double Eprev = 0.0, Etot = 0.0;
        const double *Fd = nullptr;  // Fock matrix (DIIS-extrapolated) of this iteration
        bool done = false;
    };
    std::vector<Mol> mols(nmol);
    // molecules of one size and occupation: their density steps run as one
    // batched call sequence (rocBLAS / rocSOLVER batched over pointer arrays)
    struct Group {
        int n = 0, nocc = 0;
        std::vector<int> members;
        double *D = nullptr, *E = nullptr;
        rocblas_int *info = nullptr;
    };
    std::vector<Group> groups;
    Dev dev;
    hipStream_t st = nullptr;
    rocblas_handle hb = nullptr;
    pk_ctx *ctx = nullptr;
    int rc = PK_OK;
    const double one = 1.0, zero = 0.0, two = 2.0;
    const int nparts = 256;
    const int diis_dim = opt->diis_dim > 0 ? opt->diis_dim : 8;
    int max_cycles = opt->max_cycles, auto_cycles = 0;
    std::vector<pk_basis_desc> descs(nmol);
    std::vector<int32_t> lay_nb(nmol);
    std::vector<int64_t> lay_off(nmol);
    size_t ntot = 0;
    double *Pp = nullptr, *Gp = nullptr;
    bool have_G = false;  // Gp already holds G[Pp] (guess fallback)
    (void)one;
    auto gemm_h = [&](rocblas_handle h, Mol &m, rocblas_operation ta, rocblas_operation tb, int k, const double *A,
                      const double *B, double *Cm) {
        return rocblas_dgemm(h, ta, tb, m.n, m.n, k, &one, A, m.n, B, m.n, &zero, Cm, m.n);
    };
    auto gemm = [&](Mol &m, rocblas_operation ta, rocblas_operation tb, int k, const double *A, const double *B,
                    double *Cm) { return gemm_h(hb, m, ta, tb, k, A, B, Cm); };
    auto density_h = [&](rocblas_handle h, Mol &m, const double *Fm, double *Pt) -> int {
        const int n = m.n;
        if (gemm_h(h, m, rocblas_operation_none, rocblas_operation_none, n, m.X, Fm, m.M) != rocblas_status_success ||
            gemm_h(h, m, rocblas_operation_none, rocblas_operation_none, n, m.M, m.X, m.W1) != rocblas_status_success ||
            rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, n, m.W1, n, m.ev, m.ee, m.info) !=
                rocblas_status_success ||
            gemm_h(h, m, rocblas_operation_none, rocblas_operation_none, n, m.X, m.W1, m.C) != rocblas_status_success ||
            rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, n, n, m.nocc, &two, m.C, n, m.C, n,
                          &zero, Pt, n) != rocblas_status_success)
            return pkh::fail(PK_ERR_HIP, "pk_scf_rhf_batch: dense algebra failed");
        return PK_OK;
    };
    auto density = [&](Mol &m, const double *Fm, double *Pt) -> int { return density_h(hb, m, Fm, Pt); };


    SCF_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    SCF_RB(rocblas_create_handle(&hb));
    SCF_RB(rocblas_set_stream(hb, st));
    for (int i = 0; i < nmol; ++i) {
        Mol &m = mols[i];
        if (!bases[i]) {
            rc = pkh::fail(PK_ERR_ARG, "pk_scf_rhf_batch: null basis");
            goto done;
        }
        m.b = &pkh_basis_of(bases[i]);
        m.n = m.b->nb;
        m.n2 = (size_t)m.n * m.n;
        int nel = -opt->charge;
        for (int z : m.b->Z) nel += z;
        if (nel % 2 != 0) {
            rc = pkh::fail(PK_ERR_ARG, "RHF requires an even number of electrons (closed shell)");
            goto done;
        }
        m.nocc = nel / 2;
        pk_basis_describe(bases[i], &descs[i]);
        res[i] = pk_scf_result{};
        res[i].nuclear_repulsion = pkh::nuclear_repulsion(*m.b);
        const size_t n2 = m.n2;
        for (double **b : {&m.S, &m.H, &m.X, &m.Pn, &m.F, &m.M, &m.D, &m.W1, &m.C}) *b = dev.alloc(n2);
        m.ev = dev.alloc(m.n), m.ee = dev.alloc(m.n), m.part = dev.alloc(2 * nparts);
        m.info = (rocblas_int *)dev.alloc(1);
        for (int s = 0; s < diis_dim; ++s) {
            m.Fh.push_back(dev.alloc(n2));
            m.Eh.push_back(dev.alloc(n2));
        }
        if (!m.S || !m.H || !m.X || !m.Pn || !m.F || !m.M || !m.D || !m.W1 || !m.C || !m.ev || !m.ee || !m.part ||
            !m.info || !m.Fh.back() || !m.Eh.back()) {
            rc = pkh::fail(PK_ERR_ALLOC, "pk_scf_rhf_batch: device allocation failed");
            goto done;
        }
        m.Bm.assign(diis_dim, std::vector<double>(diis_dim, 0.0));
        auto_cycles = std::max(auto_cycles, m.n > 1000 ? 300 : m.n > 500 ? 200 : m.n > 250 ? 100 : 50);
    }
    if (max_cycles <= 0) max_cycles = auto_cycles;
    // the batch context: one build per iteration for all molecules
    SCF_PK(pk_batch_create(descs.data(), nmol, &ctx));
    SCF_PK(pk_ctx_set_stream(ctx, st));
    SCF_PK(pk_ctx_set(ctx, opt->engine, opt->tol_eri, opt->precision, opt->fp32_cutoff));
    SCF_PK(pk_ctx_set_graph(ctx, 1));
    SCF_PK(pk_ctx_mol_layout(ctx, lay_nb.data(), lay_off.data()));
    ntot = (size_t)(lay_off[nmol - 1] + (int64_t)lay_nb[nmol - 1] * lay_nb[nmol - 1]);
    Pp = dev.alloc(ntot), Gp = dev.alloc(ntot);
    if (!Pp || !Gp) {
        rc = pkh::fail(PK_ERR_ALLOC, "pk_scf_rhf_batch: device allocation failed");
        goto done;
    }
    for (int i = 0; i < nmol; ++i) {
        Mol &m = mols[i];
        const size_t n2 = m.n2;
        const int n = m.n, g2 = (int)((n2 + 255) / 256);
        m.P = Pp + lay_off[i];
        m.G = Gp + lay_off[i];
        {
            std::vector<double> w(n);
            if ((rc = pkh::one_electron_device(*m.b, m.S, m.M, m.F, st)) != PK_OK) goto done;
            hipLaunchKernelGGL(k_add, dim3(g2), dim3(256), 0, st, m.M, m.F, m.H, n2);
            SCF_HIP(hipMemcpyAsync(m.W1, m.S, n2 * 8, hipMemcpyDeviceToDevice, st));
            SCF_RB(rocsolver_dsyevd(hb, rocblas_evect_original, rocblas_fill_upper, n, m.W1, n, m.ev, m.ee, m.info));
            SCF_HIP(hipMemcpyAsync(w.data(), m.ev, n * 8, hipMemcpyDeviceToHost, st));
            SCF_HIP(hipStreamSynchronize(st));
            if (w[0] < 1e-8) {
                rc = pkh::fail(PK_ERR_ARG, "Overlap matrix is near-singular in molecule " + std::to_string(i));
                goto done;
            }
        }
        SCF_HIP(hipMemcpyAsync(m.C, m.W1, n2 * 8, hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(k_scale_cols, dim3(g2), dim3(256), 0, st, m.C, m.ev, n);
        SCF_RB(gemm(m, rocblas_operation_none, rocblas_operation_transpose, n, m.C, m.W1, m.X));
        if (guess && guess[i]) {
            SCF_HIP(hipMemcpyAsync(m.P, guess[i], n2 * 8, hipMemcpyHostToDevice, st));
            res[i].guess_used = 1;
        } else if ((rc = density(m, m.H, m.P)) != PK_OK) {
            goto done;
        }
    }
    // size groups for the batched density step
    {
        std::map<std::pair<int, int>, int> gi;
        for (int i = 0; i < nmol; ++i) {
            const auto key = std::make_pair(mols[i].n, mols[i].nocc);
            auto f = gi.find(key);
            if (f == gi.end()) {
                f = gi.emplace(key, (int)groups.size()).first;
                groups.emplace_back();
                groups.back().n = mols[i].n;
                groups.back().nocc = mols[i].nocc;
            }
            groups[f->second].members.push_back(i);
        }
        for (Group &g : groups) {
            const size_t cnt = g.members.size();
            g.D = dev.alloc((size_t)g.n * cnt);
            g.E = dev.alloc((size_t)g.n * cnt);
            g.info = (rocblas_int *)dev.alloc(cnt);
            if (!g.D || !g.E || !g.info) {
                rc = pkh::fail(PK_ERR_ALLOC, "pk_scf_rhf_batch: device allocation failed");
                goto done;
            }
        }
    }
    // Guess fallback (the rule of pk_scf_rhf): for every molecule given a
    // guess density, the first DIIS error of that density and of the core
    // guess, each from one batched build of all molecules; the smaller one
    // starts the SCF.  The builds of the chosen densities are kept, so the
    // first iteration does not build again.
    if (opt->guess_fallback && guess) {
        std::vector<int> gm;
        for (int i = 0; i < nmol; ++i)
            if (guess[i]) gm.push_back(i);
        if (!gm.empty()) {
            double *Pc = dev.alloc(ntot), *Gc = dev.alloc(ntot);
            if (!Pc || !Gc) {
                rc = pkh::fail(PK_ERR_ALLOC, "pk_scf_rhf_batch: device allocation failed");
                goto done;
            }
            // RMS of X (F P S - S P F) X with F = H + G for molecule m
            auto first_error = [&](Mol &m, const double *Pm, const double *Gm, double &err) -> int {
                const int n = m.n, g2 = (int)((m.n2 + 255) / 256);
                hipLaunchKernelGGL(k_add, dim3(g2), dim3(256), 0, st, m.H, Gm, m.F, m.n2);
                if (gemm(m, rocblas_operation_none, rocblas_operation_none, n, m.F, Pm, m.W1) != rocblas_status_success ||
                    gemm(m, rocblas_operation_none, rocblas_operation_none, n, m.W1, m.S, m.M) != rocblas_status_success)
                    return pkh::fail(PK_ERR_HIP, "pk_scf_rhf_batch: gemm failed");
                hipLaunchKernelGGL(k_antisym, dim3(g2), dim3(256), 0, st, m.M, m.D, n);
                if (gemm(m, rocblas_operation_none, rocblas_operation_none, n, m.X, m.D, m.W1) != rocblas_status_success ||
                    gemm(m, rocblas_operation_none, rocblas_operation_none, n, m.W1, m.X, m.M) != rocblas_status_success)
                    return pkh::fail(PK_ERR_HIP, "pk_scf_rhf_batch: gemm failed");
                double v = 0;
                if (rocblas_set_pointer_mode(hb, rocblas_pointer_mode_host) != rocblas_status_success ||
                    rocblas_ddot(hb, (rocblas_int)m.n2, m.M, 1, m.M, 1, &v) != rocblas_status_success)
                    return pkh::fail(PK_ERR_HIP, "pk_scf_rhf_batch: ddot failed");
                err = std::sqrt(v / (double)m.n2);
                return PK_OK;
            };
            SCF_HIP(hipMemcpyAsync(Pc, Pp, ntot * 8, hipMemcpyDeviceToDevice, st));
            for (int i : gm)
                if ((rc = density(mols[i], mols[i].H, Pc + lay_off[i])) != PK_OK) goto done;
            std::vector<double> eg(nmol, 0.0), ec(nmol, 0.0);
            for (int pass = 0; pass < 2; ++pass) {
                double *Pb = pass == 0 ? Pp : Pc, *Gb = pass == 0 ? Gp : Gc;
                SCF_PK(pk_fock_2e(ctx, Pb, Gb, PK_MEM_DEVICE));
                pk_stats s2;
                pk_ctx_stats(ctx, &s2);
                for (int i = 0; i < nmol; ++i) res[i].fock_ms += s2.last_build_ms;
                for (int i : gm)
                    if ((rc = first_error(mols[i], Pb + lay_off[i], Gb + lay_off[i], pass == 0 ? eg[i] : ec[i])) !=
                        PK_OK)
                        goto done;
            }
            for (int i : gm) {
                res[i].guess_diis_error = eg[i];
                res[i].core_diis_error = ec[i];
                if (ec[i] < eg[i]) {
                    SCF_HIP(hipMemcpyAsync(Pp + lay_off[i], Pc + lay_off[i], mols[i].n2 * 8, hipMemcpyDeviceToDevice, st));
                    SCF_HIP(hipMemcpyAsync(Gp + lay_off[i], Gc + lay_off[i], mols[i].n2 * 8, hipMemcpyDeviceToDevice, st));
                    res[i].guess_used = 0;
                }
            }
            have_G = true;
        }
    }
    // Iterations.  Every step of every unconverged molecule runs as a few
    // launches over molecule tables (pointer arrays and sizes staged in one
    // host buffer and copied in one transfer): F = H + G with the energy
    // partial sums, the DIIS error Xt (F P S - S P F) X (batched GEMMs per
    // size group), the new row of each molecule's DIIS matrix, then, after
    // the host's small DIIS solves, the extrapolated Fock matrices, the
    // batched density step and the density change (two host synchronisations
    // per iteration).
    {
        const int nblk = 32;
        // per molecule: phase A 17 slots + 3 per DIIS vector, phase B 12 + 2 per DIIS vector
        const size_t stage_bytes = (size_t)nmol * 8 * (32 + 4 * (size_t)diis_dim) + 4096;
        char *dstage = (char *)dev.alloc(stage_bytes / 8);
        // partials: 2 per block for each molecule and each DIIS matrix entry
        const size_t npart = (size_t)nmol * 2 * nblk * (2 + (size_t)diis_dim);
        double *dpart = dev.alloc(npart);
        if (!dstage || !dpart) {
            rc = pkh::fail(PK_ERR_ALLOC, "pk_scf_rhf_batch: device allocation failed");
            goto done;
        }
        std::vector<char> hstage(stage_bytes);
        std::vector<double> hpart(npart);
        size_t used = 0;
        bool overflow = false;
        // place `bytes` at the next 8-byte slot of the staging buffer; returns
        // the device address it will have
        auto put = [&](const void *src, size_t bytes) -> void * {
            const size_t at = used;
            if (at + bytes > hstage.size()) {
                overflow = true;
                return dstage;
            }
            std::memcpy(hstage.data() + at, src, bytes);
            used += (bytes + 7) / 8 * 8;
            return dstage + at;
        };
        const dim3 blk(256);
        for (int it = 1; it <= max_cycles; ++it) {
            double build_ms = 0;
            if (!have_G) {
                SCF_PK(pk_fock_2e(ctx, Pp, Gp, PK_MEM_DEVICE));
                pk_stats s2;
                pk_ctx_stats(ctx, &s2);
                build_ms = s2.last_build_ms;
            }
            have_G = false;
            std::vector<int> act;
            for (int i = 0; i < nmol; ++i)
                if (!mols[i].done) act.push_back(i);
            const int cnt = (int)act.size();
            if (cnt == 0) break;
            const rocblas_operation N = rocblas_operation_none;
            // molecule tables of one size group (device pointer arrays)
            struct GT {
                int n, cnt;
                double *const *t[8];
                const Group *grp;
            };
            // ---- phase A (its device tables live in the staging buffer until
            // phase B rewrites it: they are scoped to the block below)
            used = 0;
            std::vector<int> slot(nmol, -1);
            for (int i : act) {
                Mol &m = mols[i];
                if ((int)m.slots.size() < diis_dim) {
                    slot[i] = (int)m.slots.size();
                } else {
                    slot[i] = m.slots.front();
                    m.slots.pop_front();
                }
                m.slots.push_back(slot[i]);
            }
            std::vector<std::pair<int, int>> pairs;  // DIIS matrix entries: (molecule, slot)
            int np2 = 0;
            {
            std::vector<const double *> vH, vG, vP;
            std::vector<double *> vF;
            std::vector<long long> vn2;
            std::vector<int> vn;
            for (int i : act) {
                vH.push_back(mols[i].H);
                vG.push_back(mols[i].G);
                vF.push_back(mols[i].F);
                vP.push_back(mols[i].P);
                vn2.push_back((long long)mols[i].n2);
                vn.push_back(mols[i].n);
            }
            const double *const *tH = (const double *const *)put(vH.data(), cnt * 8);
            const double *const *tG = (const double *const *)put(vG.data(), cnt * 8);
            double *const *tF = (double *const *)put(vF.data(), cnt * 8);
            const double *const *tP = (const double *const *)put(vP.data(), cnt * 8);
            const long long *tn2 = (const long long *)put(vn2.data(), cnt * 8);
            const int *tn = (const int *)put(vn.data(), cnt * 4);
            // error GEMM tables per group: F, P, S, X, W1, M, D, E_slot
            std::vector<GT> gts;
            std::vector<const double *> vM, vD;
            for (Group &g : groups) {
                std::vector<double *> cols[8];
                for (int i : g.members) {
                    if (mols[i].done) continue;
                    Mol &m = mols[i];
                    double *v[8] = {m.F, m.P, m.S, m.X, m.W1, m.M, m.D, m.Eh[slot[i]]};
                    for (int k = 0; k < 8; ++k) cols[k].push_back(v[k]);
                }
                if (cols[0].empty()) continue;
                GT gt{g.n, (int)cols[0].size(), {}, &g};
                for (int k = 0; k < 8; ++k) gt.t[k] = (double *const *)put(cols[k].data(), cols[k].size() * 8);
                gts.push_back(gt);
            }
            for (int i : act) {
                vM.push_back(mols[i].M);
                vD.push_back(mols[i].D);
            }
            const double *const *tM = (const double *const *)put(vM.data(), cnt * 8);
            double *const *tD = (double *const *)put((const void *)vD.data(), cnt * 8);
            // DIIS matrix row: (E_new . E_s) for every slot s of the molecule
            std::vector<const double *> va, vb;
            std::vector<long long> vpn;
            for (int i : act)
                for (int s3 : mols[i].slots) {
                    va.push_back(mols[i].Eh[slot[i]]);
                    vb.push_back(mols[i].Eh[s3]);
                    vpn.push_back((long long)mols[i].n2);
                    pairs.emplace_back(i, s3);
                }
            np2 = (int)pairs.size();
            const double *const *tA = (const double *const *)put(va.data(), np2 * 8);
            const double *const *tB = (const double *const *)put(vb.data(), np2 * 8);
            const long long *tpn = (const long long *)put(vpn.data(), np2 * 8);
            if (overflow) {
                rc = pkh::fail(PK_ERR_STATE, "pk_scf_rhf_batch: molecule tables exceed their buffer");
                goto done;
            }
            SCF_HIP(hipMemcpyAsync(dstage, hstage.data(), used, hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(kb_fock, dim3(nblk, cnt), blk, 0, st, tH, tG, tF, tP, tn2, dpart);
            for (const GT &g : gts)  // W1 = F P, M = W1 S
                if (rocblas_dgemm_batched(hb, N, N, g.n, g.n, g.n, &one, g.t[0], g.n, g.t[1], g.n, &zero, g.t[4], g.n,
                                          g.cnt) != rocblas_status_success ||
                    rocblas_dgemm_batched(hb, N, N, g.n, g.n, g.n, &one, g.t[4], g.n, g.t[2], g.n, &zero, g.t[5], g.n,
                                          g.cnt) != rocblas_status_success) {
                    rc = pkh::fail(PK_ERR_HIP, "pk_scf_rhf_batch: batched GEMM failed");
                    goto done;
                }
            hipLaunchKernelGGL(kb_antisym, dim3(nblk, cnt), blk, 0, st, tM, tD, tn);
            for (const GT &g : gts)  // W1 = X D, E = W1 X
                if (rocblas_dgemm_batched(hb, N, N, g.n, g.n, g.n, &one, g.t[3], g.n, g.t[6], g.n, &zero, g.t[4], g.n,
                                          g.cnt) != rocblas_status_success ||
                    rocblas_dgemm_batched(hb, N, N, g.n, g.n, g.n, &one, g.t[4], g.n, g.t[3], g.n, &zero, g.t[7], g.n,
                                          g.cnt) != rocblas_status_success) {
                    rc = pkh::fail(PK_ERR_HIP, "pk_scf_rhf_batch: batched GEMM failed");
                    goto done;
                }
            for (int i : act)
                SCF_HIP(hipMemcpyAsync(mols[i].Fh[slot[i]], mols[i].F, mols[i].n2 * 8, hipMemcpyDeviceToDevice, st));
            double *dpart2 = dpart + (size_t)cnt * nblk * 2;
            hipLaunchKernelGGL(kb_dots, dim3(nblk, np2), blk, 0, st, tA, tB, tpn, dpart2);
            SCF_HIP(hipGetLastError());
            SCF_HIP(hipMemcpyAsync(hpart.data(), dpart, ((size_t)cnt + np2) * nblk * 2 * 8, hipMemcpyDeviceToHost, st));
            SCF_HIP(hipStreamSynchronize(st));
            }
            // ---- host: energies, DIIS matrices and coefficients
            for (int k = 0; k < cnt; ++k) {
                Mol &m = mols[act[k]];
                double eph = 0, epf = 0;
                for (int b2 = 0; b2 < nblk; ++b2) {
                    eph += hpart[2 * ((size_t)k * nblk + b2)];
                    epf += hpart[2 * ((size_t)k * nblk + b2) + 1];
                }
                m.Etot = 0.5 * (eph + epf) + res[act[k]].nuclear_repulsion;
                res[act[k]].fock_ms += build_ms;
            }
            for (int q = 0; q < np2; ++q) {
                double v = 0;
                for (int b2 = 0; b2 < nblk; ++b2) v += hpart[2 * (((size_t)cnt + q) * nblk + b2)];
                const int i = pairs[q].first, s3 = pairs[q].second;
                mols[i].Bm[slot[i]][s3] = mols[i].Bm[s3][slot[i]] = v;
            }
            // ---- phase B
            used = 0;
            std::vector<const double *> vsrc;
            std::vector<double> vcoef;
            std::vector<int> vns;
            std::vector<double *> vout;
            std::vector<long long> vcn;
            for (int i : act) {
                Mol &m = mols[i];
                if (it == 1) res[i].first_diis_error = std::sqrt(m.Bm[slot[i]][slot[i]] / (double)m.n2);
                m.Fd = m.F;
                if (m.slots.size() < 2) continue;
                const int k = (int)m.slots.size();
                std::vector<double> A((size_t)(k + 1) * (k + 1), 0.0), rhs(k + 1, 0.0);
                for (int a2 = 0; a2 < k; ++a2) {
                    for (int b2 = 0; b2 < k; ++b2) A[(size_t)a2 * (k + 1) + b2] = m.Bm[m.slots[a2]][m.slots[b2]];
                    A[(size_t)a2 * (k + 1) + k] = A[(size_t)k * (k + 1) + a2] = -1.0;
                }
                rhs[k] = -1.0;
                const std::vector<double> cf = qr_solve(k + 1, A, rhs);
                for (int a2 = 0; a2 < diis_dim; ++a2) {
                    vsrc.push_back(a2 < k ? m.Fh[m.slots[a2]] : m.F);
                    vcoef.push_back(a2 < k ? cf[a2] : 0.0);
                }
                vns.push_back(k);
                vout.push_back(m.D);
                vcn.push_back((long long)m.n2);
                m.Fd = m.D;
            }
            const int ncomb = (int)vout.size();
            const double *const *tsrc = (const double *const *)put(vsrc.data(), vsrc.size() * 8);
            const double *tcoef = (const double *)put(vcoef.data(), vcoef.size() * 8);
            const int *tns = (const int *)put(vns.data(), vns.size() * 4);
            double *const *tout = (double *const *)put(vout.data(), vout.size() * 8);
            const long long *tcn = (const long long *)put(vcn.data(), vcn.size() * 8);
            std::vector<GT> dts;  // density tables per group: X, Fd, M, W1, C, Pn
            for (Group &g : groups) {
                std::vector<double *> cols[6];
                for (int i : g.members) {
                    if (mols[i].done) continue;
                    Mol &m = mols[i];
                    double *v[6] = {m.X, const_cast<double *>(m.Fd), m.M, m.W1, m.C, m.Pn};
                    for (int k = 0; k < 6; ++k) cols[k].push_back(v[k]);
                }
                if (cols[0].empty()) continue;
                GT gt{g.n, (int)cols[0].size(), {}, &g};
                for (int k = 0; k < 6; ++k) gt.t[k] = (double *const *)put(cols[k].data(), cols[k].size() * 8);
                dts.push_back(gt);
            }
            std::vector<const double *> vPn;
            std::vector<double *> vPo;
            for (int i : act) {
                vPn.push_back(mols[i].Pn);
                vPo.push_back(mols[i].P);
            }
            std::vector<long long> vn2b;
            for (int i : act) vn2b.push_back((long long)mols[i].n2);
            const double *const *tPn = (const double *const *)put(vPn.data(), cnt * 8);
            double *const *tPo = (double *const *)put(vPo.data(), cnt * 8);
            const long long *tn2b = (const long long *)put(vn2b.data(), cnt * 8);
            if (overflow) {
                rc = pkh::fail(PK_ERR_STATE, "pk_scf_rhf_batch: molecule tables exceed their buffer");
                goto done;
            }
            SCF_HIP(hipMemcpyAsync(dstage, hstage.data(), used, hipMemcpyHostToDevice, st));
            if (ncomb > 0)
                hipLaunchKernelGGL(kb_combine, dim3(nblk, ncomb), blk, 0, st, tsrc, tcoef, tns, diis_dim, tout, tcn);
            for (const GT &g : dts) {
                const Group &gr = *g.grp;
                const int n = g.n, c2 = g.cnt;
                const rocblas_operation T = rocblas_operation_transpose;
                if (rocblas_dgemm_batched(hb, N, N, n, n, n, &one, g.t[0], n, g.t[1], n, &zero, g.t[2], n, c2) !=
                        rocblas_status_success ||
                    rocblas_dgemm_batched(hb, N, N, n, n, n, &one, g.t[2], n, g.t[0], n, &zero, g.t[3], n, c2) !=
                        rocblas_status_success ||
                    rocsolver_dsyevd_batched(hb, rocblas_evect_original, rocblas_fill_upper, n, g.t[3], n, gr.D, n,
                                             gr.E, n, gr.info, c2) != rocblas_status_success ||
                    rocblas_dgemm_batched(hb, N, N, n, n, n, &one, g.t[0], n, g.t[3], n, &zero, g.t[4], n, c2) !=
                        rocblas_status_success ||
                    rocblas_dgemm_batched(hb, N, T, n, n, gr.nocc, &two, g.t[4], n, g.t[4], n, &zero, g.t[5], n, c2) !=
                        rocblas_status_success) {
                    rc = pkh::fail(PK_ERR_HIP, "pk_scf_rhf_batch: batched density step failed");
                    goto done;
                }
            }
            hipLaunchKernelGGL(kb_diff, dim3(nblk, cnt), blk, 0, st, tPn, tPo, tn2b, dpart);
            SCF_HIP(hipGetLastError());
            SCF_HIP(hipMemcpyAsync(hpart.data(), dpart, (size_t)cnt * nblk * 2 * 8, hipMemcpyDeviceToHost, st));
            SCF_HIP(hipStreamSynchronize(st));
            for (int k = 0; k < cnt; ++k) {
                Mol &m = mols[act[k]];
                pk_scf_result &r = res[act[k]];
                double dmax = 0, dss = 0;
                for (int b2 = 0; b2 < nblk; ++b2) {
                    dmax = std::max(dmax, hpart[2 * ((size_t)k * nblk + b2)]);
                    dss += hpart[2 * ((size_t)k * nblk + b2) + 1];
                }
                const double drms = std::sqrt(dss / (double)m.n2), dE = std::fabs(m.Etot - m.Eprev);
                m.Eprev = m.Etot;
                r.iterations = it;
                r.energy = m.Etot;
                if (it > 1 && dE < opt->tol_energy && drms < opt->tol_density && dmax < opt->tol_density) {
                    r.converged = 1;
                    m.done = true;
                }
            }
        }
    }
    for (int i = 0; i < nmol; ++i)
        if (P_out && P_out[i]) SCF_HIP(hipMemcpyAsync(P_out[i], mols[i].P, mols[i].n2 * 8, hipMemcpyDeviceToHost, st));
    SCF_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < nmol; ++i)
        if (!res[i].converged) rc = pkh::fail(1, "SCF did not converge in " + std::to_string(max_cycles) +
                                                     " iterations for molecule " + std::to_string(i));
done:
    if (st) (void)hipStreamSynchronize(st);
    {
        const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
        for (int i = 0; i < nmol; ++i) res[i].wall_s = wall;
    }
    if (ctx) pk_ctx_destroy(ctx);
    if (hb) rocblas_destroy_handle(hb);
    if (st) (void)hipStreamDestroy(st);
    return rc;
}

}  // extern "C"
