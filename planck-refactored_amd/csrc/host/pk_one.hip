// pk_one.hip — one-electron integrals S, T, V on the device.
//
// Reference semantics restated:
//   overlap / kinetic ............ src/integrals/os.cpp:248-395 (_compute_1e, base.h:10-18)
//   nuclear attraction ........... src/integrals/os.cpp:462-602 (_compute_nuclear_attraction, base.h:20-27)
//   contraction and norms ........ src/base/types.h:606-675 (c_A c_B N_A N_B norm_A norm_B)
//
// One thread per basis-function pair i >= j; the value is written to (i, j)
// and (j, i), so the matrices are exactly symmetric.  Per primitive pair the
// 1-D factors are the moments of a Gaussian: with mean offsets c (from A) and
// d (from B) and variance v,
//     I(i, j) = sum_{k<=i, l<=j} C(i,k) C(j,l) c^(i-k) d^(j-l) M_{k+l}(v),
//     M_n(v) = (n-1)!! v^(n/2) for even n, 0 for odd n,
// which is the closed form of the Obara-Saika 1-D recurrence.  Overlap:
// c = PA, d = PB, v = 1/(2 zeta).  Kinetic per axis:
//     T_x = beta (2 j + 1) S(i, j) - 2 beta^2 S(i, j+2) - j (j-1)/2 S(i, j-2).
// Nuclear attraction per nucleus C: the Rys rule of e^{-zeta |PC|^2 u} with
// n = (la+lb)/2 + 1 roots (the engine's own table); per root u the 1-D factor
// has c = PA - u PC, d = c + AB, v = (1 - u)/(2 zeta).  The work is O(nb^2 K^2
// natoms), microseconds to a few milliseconds per molecule, and is launched
// per molecule on the caller's stream.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "pk_host_internal.h"

#define PK_RYS_COEF_QUALIFIER static __device__ const
#include "rys_table.inc"

namespace pk1 {

struct OneArgs {
    const int32_t *fshell;  // shell of each function
    const int32_t *fxyz;    // lx | ly << 4 | lz << 8
    const double *fnorm;    // component norm of each function
    const double *center;   // [3 * nshells]
    const int32_t *sl, *snp, *sp0;
    const double *pexp, *pcn;  // per primitive: exponent, coef * prim_norm
    const double *atom;        // [4 * natoms]: x, y, z, Z
    int natoms, nb;
    int64_t npairs;
    double *S, *T, *V;  // row-major nb x nb (null = not wanted)
};

// sum_{k,l} C(i,k) C(j,l) c^(i-k) d^(j-l) M_{k+l}(v)
__device__ double moment1(double c, double d, double v, int i, int j) {
    double s = 0.0, ck = 1.0, bi = 1.0;
    for (int k = i; k >= 0; --k) {  // ck = c^(i-k), bi = C(i,k)
        double inner = 0.0, dl = 1.0, bj = 1.0;
        for (int l = j; l >= 0; --l) {  // dl = d^(j-l), bj = C(j,l)
            const int n = k + l;
            if ((n & 1) == 0) {
                double m = 1.0;
                for (int t = 1; t < n; t += 2) m *= t * v;
                inner = fma(bj * dl, m, inner);
            }
            dl *= d;
            bj = bj * l / (j - l + 1);
        }
        s = fma(bi * ck, inner, s);
        ck *= c;
        bi = bi * k / (i - k + 1);
    }
    return s;
}

__device__ __forceinline__ double cheb(const double *cf, double x) {
    const double x2 = 2.0 * x;
    double b1 = 0.0, b2 = 0.0;
#pragma unroll
    for (int j = PK_RYS_DEG; j >= 1; --j) {
        const double t = fma(x2, b1, cf[j] - b2);
        b2 = b1;
        b1 = t;
    }
    return fma(x, b1, cf[0] - b2);
}

// n-point rule (roots u = t^2, weights) of e^{-T u}: Chebyshev fits below
// TMAX(n), the Gauss-Hermite limit above
template <int N>
__device__ __forceinline__ void rys_rule(double T, double (&u)[N], double (&w)[N]) {
    if (T >= pk_rys_tmax[N]) {
        const double it = 1.0 / T, sit = sqrt(it);
#pragma unroll
        for (int r = 0; r < N; ++r) {
            u[r] = pk_rys_ax2[pk_rys_abase[N] + r] * it;
            w[r] = pk_rys_aw[pk_rys_abase[N] + r] * sit;
        }
    } else {
        const int i = (int)T;
        const double x = 2.0 * (T - i) - 1.0;
        const double *c = pk_rys_coef + pk_rys_base[N] + i * (2 * N * (PK_RYS_DEG + 1));
#pragma unroll
        for (int r = 0; r < N; ++r) {
            u[r] = cheb(c + r * (PK_RYS_DEG + 1), x);
            w[r] = cheb(c + (N + r) * (PK_RYS_DEG + 1), x);
        }
    }
}

// sum over nuclei of -Z sum_r w_r Ix Iy Iz (without the 2 pi / zeta E factor)
template <int N>
__device__ double nuclear_sum(const OneArgs &a, const double (&P)[3], const double (&PA)[3], const double (&AB)[3],
                              double ze, const int (&ea)[3], const int (&eb)[3]) {
    const double h = 0.5 / ze;
    double acc = 0.0;
    for (int c = 0; c < a.natoms; ++c) {
        const double *C = a.atom + 4 * c;
        const double PC[3] = {P[0] - C[0], P[1] - C[1], P[2] - C[2]};
        const double T = ze * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
        double u[N], w[N];
        rys_rule<N>(T, u, w);
        double s = 0.0;
#pragma unroll
        for (int r = 0; r < N; ++r) {
            const double v = (1.0 - u[r]) * h;
            double prod = w[r];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const double c0 = PA[q] - u[r] * PC[q];
                prod *= moment1(c0, c0 + AB[q], v, ea[q], eb[q]);
            }
            s += prod;
        }
        acc -= C[3] * s;
    }
    return acc;
}

__global__ void __launch_bounds__(256) k_one_electron(OneArgs a) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.npairs) return;
    // k -> (i, j), i >= j, k = i (i + 1) / 2 + j
    int64_t i = (int64_t)((sqrt(8.0 * (double)k + 1.0) - 1.0) * 0.5);
    while (i * (i + 1) / 2 > k) --i;
    while ((i + 1) * (i + 2) / 2 <= k) ++i;
    const int64_t j = k - i * (i + 1) / 2;
    const int sa = a.fshell[i], sb = a.fshell[j];
    const int pa = a.fxyz[i], pb = a.fxyz[j];
    const int ea[3] = {pa & 15, (pa >> 4) & 15, (pa >> 8) & 15};
    const int eb[3] = {pb & 15, (pb >> 4) & 15, (pb >> 8) & 15};
    const int nr = (a.sl[sa] + a.sl[sb]) / 2 + 1;
    const double *A = a.center + 3 * sa, *B = a.center + 3 * sb;
    const double AB[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]};
    const double R2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
    double s = 0.0, t = 0.0, vn = 0.0;
    const int p0a = a.sp0[sa], p0b = a.sp0[sb], npa = a.snp[sa], npb = a.snp[sb];
    for (int p = 0; p < npa; ++p)
        for (int q = 0; q < npb; ++q) {
            const double al = a.pexp[p0a + p], be = a.pexp[p0b + q], ze = al + be, iz = 1.0 / ze;
            const double E = exp(-al * be * iz * R2) * a.pcn[p0a + p] * a.pcn[p0b + q];
            double P[3], PA[3], PB[3];
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                P[x] = (al * A[x] + be * B[x]) * iz;
                PA[x] = P[x] - A[x];
                PB[x] = P[x] - B[x];
            }
            const double h = 0.5 * iz;
            if (a.S || a.T) {
                double o[3], kx[3];
#pragma unroll
                for (int x = 0; x < 3; ++x) {
                    const int y = eb[x];
                    o[x] = moment1(PA[x], PB[x], h, ea[x], y);
                    double kv = be * (2 * y + 1) * o[x] - 2.0 * be * be * moment1(PA[x], PB[x], h, ea[x], y + 2);
                    if (y >= 2) kv -= 0.5 * y * (y - 1) * moment1(PA[x], PB[x], h, ea[x], y - 2);
                    kx[x] = kv;
                }
                const double sf = M_PI * iz * sqrt(M_PI * iz) * E;
                s += sf * o[0] * o[1] * o[2];
                t += sf * (kx[0] * o[1] * o[2] + o[0] * kx[1] * o[2] + o[0] * o[1] * kx[2]);
            }
            if (a.V) {
                double r;
                switch (nr) {
                    case 1: r = nuclear_sum<1>(a, P, PA, AB, ze, ea, eb); break;
                    case 2: r = nuclear_sum<2>(a, P, PA, AB, ze, ea, eb); break;
                    case 3: r = nuclear_sum<3>(a, P, PA, AB, ze, ea, eb); break;
                    default: r = nuclear_sum<4>(a, P, PA, AB, ze, ea, eb); break;
                }
                vn += 2.0 * M_PI * iz * E * r;
            }
        }
    const double nn = a.fnorm[i] * a.fnorm[j];
    const size_t x = (size_t)i * a.nb + j, y = (size_t)j * a.nb + i;
    if (a.S) a.S[x] = a.S[y] = nn * s;
    if (a.T) a.T[x] = a.T[y] = nn * t;
    if (a.V) a.V[x] = a.V[y] = nn * vn;
}

}  // namespace pk1

namespace pkh {

// Device S, T, V of basis b into device buffers (any may be null), on stream
// st; returns after the kernel has completed.
int one_electron_device(const Basis &b, double *dS, double *dT, double *dV, hipStream_t st) {
    const int ns = (int)b.l.size(), nb = b.nb, nat = (int)b.Z.size();
    if (nb <= 0) return fail(PK_ERR_ARG, "one-electron integrals: empty basis");
    for (int s = 0; s < ns; ++s)
        if (b.l[s] < 0 || b.l[s] > 3) return fail(PK_ERR_UNSUPPORTED, "one-electron integrals: shells up to f only");
    std::vector<int32_t> fshell(nb), fxyz(nb), sp0(ns);
    for (int s = 0, f = 0, p = 0; s < ns; ++s) {
        sp0[s] = p;
        p += b.nprim[s];
        const int l = b.l[s];
        for (int lx = l; lx >= 0; --lx)
            for (int ly = l - lx; ly >= 0; --ly, ++f) {
                fshell[f] = s;
                fxyz[f] = lx | ly << 4 | (l - lx - ly) << 8;
            }
    }
    const size_t np = b.exponent.size();
    std::vector<double> pcn(np), atom(4 * (size_t)nat);
    for (size_t p = 0; p < np; ++p) pcn[p] = b.coef[p] * b.prim_norm[p];
    for (int c = 0; c < nat; ++c) {
        for (int q = 0; q < 3; ++q) atom[4 * c + q] = b.xyz[3 * c + q];
        atom[4 * c + 3] = (double)b.Z[c];
    }
    // one device block holding every table
    const size_t i4 = 4 * ((size_t)2 * nb + 3 * ns), d8 = 8 * ((size_t)nb + 3 * ns + 2 * np + 4 * nat);
    char *blk = nullptr;
    hipError_t e = hipMalloc(&blk, d8 + i4);
    if (e != hipSuccess) return fail(PK_ERR_ALLOC, std::string("one-electron integrals: ") + hipGetErrorString(e));
    std::vector<char> h(d8 + i4);
    char *w = h.data();
    auto put = [&](const void *src, size_t bytes) {
        char *at = w;
        std::memcpy(w, src, bytes);
        w += bytes;
        return blk + (at - h.data());
    };
    pk1::OneArgs a{};
    a.fnorm = (const double *)put(b.fnorm.data(), 8 * (size_t)nb);
    a.center = (const double *)put(b.center.data(), 8 * 3 * (size_t)ns);
    a.pexp = (const double *)put(b.exponent.data(), 8 * np);
    a.pcn = (const double *)put(pcn.data(), 8 * np);
    a.atom = (const double *)put(atom.data(), 8 * 4 * (size_t)nat);
    a.fshell = (const int32_t *)put(fshell.data(), 4 * (size_t)nb);
    a.fxyz = (const int32_t *)put(fxyz.data(), 4 * (size_t)nb);
    a.sl = (const int32_t *)put(b.l.data(), 4 * (size_t)ns);
    a.snp = (const int32_t *)put(b.nprim.data(), 4 * (size_t)ns);
    a.sp0 = (const int32_t *)put(sp0.data(), 4 * (size_t)ns);
    a.natoms = nat;
    a.nb = nb;
    a.npairs = (int64_t)nb * (nb + 1) / 2;
    a.S = dS;
    a.T = dT;
    a.V = dV;
    e = hipMemcpyAsync(blk, h.data(), h.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        const int64_t grid = (a.npairs + 255) / 256;
        hipLaunchKernelGGL(pk1::k_one_electron, dim3((unsigned)grid), dim3(256), 0, st, a);
        e = hipGetLastError();
    }
    const hipError_t e2 = hipStreamSynchronize(st);
    if (e == hipSuccess) e = e2;
    (void)hipFree(blk);
    if (e != hipSuccess) return fail(PK_ERR_HIP, std::string("one-electron integrals: ") + hipGetErrorString(e));
    return PK_OK;
}

// host pointers in and out (a temporary stream and device buffers)
int one_electron(const Basis &b, double *S, double *T, double *V) {
    const size_t n2 = (size_t)b.nb * b.nb;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(PK_ERR_HIP, "one-electron integrals: no HIP device visible");
    hipStream_t st = nullptr;
    double *d = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&d, 3 * n2 * 8);
    int rc = PK_OK;
    if (e != hipSuccess) {
        rc = fail(PK_ERR_HIP, std::string("one-electron integrals: ") + hipGetErrorString(e));
    } else {
        rc = one_electron_device(b, S ? d : nullptr, T ? d + n2 : nullptr, V ? d + 2 * n2 : nullptr, st);
        for (int m = 0; m < 3 && rc == PK_OK; ++m) {
            double *dst = m == 0 ? S : (m == 1 ? T : V);
            if (!dst) continue;
            e = hipMemcpy(dst, d + m * n2, n2 * 8, hipMemcpyDeviceToHost);
            if (e != hipSuccess) rc = fail(PK_ERR_HIP, std::string("one-electron integrals: ") + hipGetErrorString(e));
        }
    }
    if (d) (void)hipFree(d);
    if (st) (void)hipStreamDestroy(st);
    return rc;
}

}  // namespace pkh
