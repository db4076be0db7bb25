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

#include <chrono>
#include <cmath>
#include <deque>
#include <string>
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
    std::vector<double> hS(n2), hT(n2), hV(n2), hH(n2), w(n);
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

    pkh::one_electron(b, hS.data(), hT.data(), hV.data());
    for (size_t i = 0; i < n2; ++i) hH[i] = hT[i] + hV[i];

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
    SCF_HIP(hipMemcpyAsync(S, hS.data(), n2 * 8, hipMemcpyHostToDevice, st));
    SCF_HIP(hipMemcpyAsync(H, hH.data(), n2 * 8, hipMemcpyHostToDevice, st));
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
    } else if ((rc = density(H, P)) != PK_OK) {
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

}  // extern "C"
