// eri_grad.h — two-electron part of the nuclear gradient on the device.
//
// First derivatives of the contracted Cartesian ERIs (the shift rule of
// src/integrals/os.cpp:1296-1420) contracted on the fly with the two-particle
// density of the closed- or open-shell determinant,
//   Gamma_ijkl = 2 Pj_ij Pj_kl - kx (Pk1_ik Pk1_jl + Pk2_ik Pk2_jl),
// into dE2/dR for the centre of every shell, E2 = 1/4 sum Gamma_ijkl (ij|kl)
// (the ERI loop of src/gradient/gradient.cpp:209-232 with its permutation
// rule, gradient.cpp:70-118).
//
// Rys quadrature with one more unit of angular momentum on the centres A, B and
// C, per root and axis:
//   2-D integrals I[e][f], e <= la+lb+1, f <= lc+ld+1 (same recurrences as the
//   energy kernels);
//   transferred factors G[a][b][c][d], a <= la+1, b <= lb+1, c <= lc+1, d <= ld,
//   (a|b) = sum_k C(b,k) AB^(b-k) (a+k|0) and the same on the ket;
//   derivative factors dA = 2 alpha G[a+1] - a G[a-1], dB = 2 beta G[b+1] -
//   b G[b-1], dC = 2 gamma G[c+1] - c G[c-1], with alpha, beta, gamma the
//   primitive exponents on A, B, C (the 2 alpha weight of the reference's
//   wceri_A and its lowering term);
// d(ab|cd)/dA_x = sum_r dA_x Gy Gz and likewise, so each component adds
// Gamma x (three products per centre and axis).  The centre D follows from
// translation invariance: dD = -(dA + dB + dC), summed per quartet.
//
// Every unique shell quartet of the energy work lists (same Schwarz screening,
// per-function strict test) is visited once with its degeneracy factor d; the
// eight index permutations have the same derivative integrals, so with the
// symmetrised density
//   Gs = 2 Pj_ab Pj_cd - kx/2 (Pk_ac Pk_bd + Pk_ad Pk_bc)   (summed over Pk1, Pk2)
// the quartet contributes 1/4 * 8 d sum_abcd Gs d(ab|cd)/dR = 2 d sum Gs d(ab|cd)/dR.
//
// One 64-lane workgroup owns one shell quartet at a time; the class is a runtime
// value, so one compiled kernel serves every class up to (ff|ff).  Per
// primitive quartet: n lanes evaluate the roots, 3n lanes the 2-D integrals of
// all roots into LDS; then per root the transferred and derivative factors are
// spread over all lanes, and each lane adds its share of the components.  LDS:
// 2-D integrals of all roots, the factors of one root, and Gs of the quartet
// (at most 127 KB, (ff|ff)).
#pragma once

#include "eri_generic.h"  // kCartTab, sel3, rys_root_rt
#include "eri_kernels.h"  // wave_sum

namespace pk {

struct GradArgs {
    const int *__restrict__ wave_bra;
    const int *__restrict__ wave_start;
    const int *__restrict__ wave_count;
    const int *__restrict__ kets;
    const PairDev *__restrict__ pairs;
    const PrimDev *__restrict__ prims;
    const double *__restrict__ palpha;  // exponent on the pair's first shell, per primitive pair
    const double *__restrict__ hz;      // 1/(2 zeta) per primitive pair
    const int *__restrict__ pair_sh;    // [2 * npairs]: first and second shell of each pair
    const double *__restrict__ Pj;      // densities scaled by component norms, internal order
    const double *__restrict__ Pk1;
    const double *__restrict__ Pk2;  // may be null
    const double *__restrict__ Q;
    double *g;  // [3 * nshells] dE2/dR per shell centre
    double tol, kx;
    int nb;
    int nitems;
    int la, lb, lc, ld;
};

constexpr int GRAD_NR = 7;  // roots for L + 1 <= 13
constexpr int GRAD_NE = 8;  // 2-D extent e, f <= 7

// C(n, k) for 0 <= k <= n <= 4
__device__ __forceinline__ double binom4(int n, int k) {
    if (k == 0 || k == n) return 1.0;
    if (n == 4 && k == 2) return 6.0;
    return (double)n;  // C(n, 1) = C(n, n-1) = n for n <= 4
}

struct GradTabs {
    unsigned char ab_a[100], ab_b[100], cd_c[100], cd_d[100];  // ab -> (a, b), cd -> (c, d)
    unsigned short ab_g[100][3];                               // compact factor offset of (a, b) per axis
    unsigned char cd_g[100][3];                                // compact factor offset of (c, d) per axis
    unsigned char g_a[256], g_b[256], g_c[256], g_d[256];      // compact index -> exponents
    unsigned char x_a[512], x_b[512], x_c[512], x_d[512];      // extended index -> exponents
};

__global__ __launch_bounds__(64) void k_grad(GradArgs a) {
    extern __shared__ double smem[];
    __shared__ double su[GRAD_NR], sw[GRAD_NR];
    __shared__ GradTabs tb;
    const int lane = threadIdx.x;
    const int item = blockIdx.x;
    if (item >= a.nitems) return;
    const int la = a.la, lb = a.lb, lc = a.lc, ld = a.ld;
    const int NA = (la + 1) * (la + 2) / 2, NB = (lb + 1) * (lb + 2) / 2;
    const int NC = (lc + 1) * (lc + 2) / 2, ND = (ld + 1) * (ld + 2) / 2;
    const int NAB = NA * NB, NCD = NC * ND, NABCD = NAB * NCD;
    const int E1 = la + lb + 1, F1 = lc + ld + 1, n = (E1 + F1 - 1) / 2 + 1;
    const int FS = F1 + 1, NI = (E1 + 1) * FS;
    // extended factors: a <= la+1, b <= lb+1, c <= lc+1, d <= ld
    const int XB = lb + 2, XC = lc + 2, XD = ld + 1, XCD = XC * XD, XBCD = XB * XCD;
    const int NGX = (la + 2) * XBCD;
    // compact factors: a <= la, b <= lb, c <= lc, d <= ld
    const int D2 = lb + 1, D3 = lc + 1, D4 = ld + 1, D34 = D3 * D4;
    const int NG = (la + 1) * D2 * D34;
    double *I2 = smem;            // [3n][E1+1][F1+1]
    double *GX = I2 + 3 * n * NI;  // [3 axes][NGX], one root
    double *FK = GX + 3 * NGX;     // [3 axes][4 kinds: G, dA, dB, dC][NG], one root
    double *gam = FK + 12 * NG;    // [NABCD]

    for (int i = lane; i < NAB; i += 64) {
        const int ia = i / NB, ib = i % NB;
        tb.ab_a[i] = ia;
        tb.ab_b[i] = ib;
        for (int q = 0; q < 3; ++q) tb.ab_g[i][q] = (kCartTab[la][ia][q] * D2 + kCartTab[lb][ib][q]) * D34;
    }
    for (int i = lane; i < NCD; i += 64) {
        const int ic = i / ND, id = i % ND;
        tb.cd_c[i] = ic;
        tb.cd_d[i] = id;
        for (int q = 0; q < 3; ++q) tb.cd_g[i][q] = kCartTab[lc][ic][q] * D4 + kCartTab[ld][id][q];
    }
    for (int g = lane; g < NG; g += 64) {
        int t = g;
        tb.g_d[g] = t % D4;
        t /= D4;
        tb.g_c[g] = t % D3;
        t /= D3;
        tb.g_b[g] = t % D2;
        tb.g_a[g] = t / D2;
    }
    for (int g = lane; g < NGX; g += 64) {
        int t = g;
        tb.x_d[g] = t % XD;
        t /= XD;
        tb.x_c[g] = t % XC;
        t /= XC;
        tb.x_b[g] = t % XB;
        tb.x_a[g] = t / XB;
    }
    const int ab0 = lane / NCD, cd0 = lane % NCD, sab = 64 / NCD, scd = 64 % NCD;
    __syncthreads();

    const int nb = a.nb;
    const int bra = a.wave_bra[item];
    const int kbeg = a.wave_start[item];
    const int kcnt = a.wave_count[item];
    const PairDev pb = a.pairs[bra];
    const int sA = a.pair_sh[2 * bra], sB = a.pair_sh[2 * bra + 1];
    auto at = [nb](const double *M, int r, int c) { return M[(size_t)r * nb + c]; };

    for (int kk = 0; kk < kcnt; ++kk) {
        const int ket = a.kets[kbeg + kk];
        const PairDev pk = a.pairs[ket];
        const int sC = a.pair_sh[2 * ket], sD = a.pair_sh[2 * ket + 1];
        // symmetrised two-particle density of the quartet, with 2 d and the
        // per-function Schwarz test folded in
        double deg = 2.0;
        if (pb.diag) deg *= 0.5;
        if (pk.diag) deg *= 0.5;
        if (bra == ket) deg *= 0.5;
        const bool all_pass = pb.qmin * pk.qmin >= a.tol;
        {
            int ab = ab0, cd = cd0;
            for (int i = lane; i < NABCD; i += 64) {
                const int af = pb.f0 + tb.ab_a[ab], bf = pb.f1 + tb.ab_b[ab];
                const int cf = pk.f0 + tb.cd_c[cd], df = pk.f1 + tb.cd_d[cd];
                double m = deg;
                if (!all_pass && at(a.Q, af, bf) * at(a.Q, cf, df) < a.tol) m = 0.0;
                double x = at(a.Pk1, af, cf) * at(a.Pk1, bf, df) + at(a.Pk1, af, df) * at(a.Pk1, bf, cf);
                if (a.Pk2) x += at(a.Pk2, af, cf) * at(a.Pk2, bf, df) + at(a.Pk2, af, df) * at(a.Pk2, bf, cf);
                gam[i] = m * fma(2.0 * at(a.Pj, af, bf), at(a.Pj, cf, df), -0.5 * a.kx * x);
                cd += scd;
                ab += sab;
                if (cd >= NCD) cd -= NCD, ++ab;
            }
        }
        double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0}, gC[3] = {0.0, 0.0, 0.0};
        __syncthreads();
        for (int kq = 0; kq < pk.nprim; ++kq) {
            const PrimDev q = a.prims[pk.prim_off + kq];
            const double g2 = 2.0 * a.palpha[pk.prim_off + kq];  // 2 gamma (exponent on C)
            const double hq = a.hz[pk.prim_off + kq];
            for (int kp = 0; kp < pb.nprim; ++kp) {
                const PrimDev p = a.prims[pb.prim_off + kp];
                const double a2 = 2.0 * a.palpha[pb.prim_off + kp];  // 2 alpha (on A)
                const double hp = a.hz[pb.prim_off + kp];
                const double b2 = 2.0 * p.zeta - a2;                 // 2 beta (on B)
                const double rs = rsqrt_nr(p.zeta + q.zeta);
                const double inv_de = rs * rs;
                const double rho = p.zeta * q.zeta * inv_de;
                const double PQ[3] = {p.p[0] - q.p[0], p.p[1] - q.p[1], p.p[2] - q.p[2]};
                const double T = rho * fma(PQ[0], PQ[0], fma(PQ[1], PQ[1], PQ[2] * PQ[2]));
                const double pref = p.kc * q.kc * rs;
                if (lane < n) {
                    double u, w;
                    rys_root_rt(n, T, lane, u, w);
                    su[lane] = u;
                    sw[lane] = w;
                }
                __syncthreads();
                if (lane < 3 * n) {
                    const int r = lane / 3, ax = lane - 3 * r;
                    const double ur = su[r];
                    const double B00 = 0.5 * inv_de * ur;
                    const double B10 = hp * (1.0 - q.zeta * inv_de * ur);  // rho / zeta = eta / (zeta + eta)
                    const double B01 = hq * (1.0 - p.zeta * inv_de * ur);
                    const double PQa = sel3(ax, PQ[0], PQ[1], PQ[2]);
                    const double C00 = fma(ur * (-q.zeta * inv_de), PQa, sel3(ax, p.pa[0], p.pa[1], p.pa[2]));
                    const double D00 = fma(ur * (p.zeta * inv_de), PQa, sel3(ax, q.pa[0], q.pa[1], q.pa[2]));
                    double *I = I2 + lane * NI;
                    double c2[GRAD_NE], c1[GRAD_NE], c0[GRAD_NE];
                    c1[0] = ax == 0 ? pref * sw[r] : 1.0;
#pragma unroll
                    for (int e = 1; e < GRAD_NE; ++e) {
                        double t = C00 * c1[e - 1];
                        if (e >= 2) t = fma((double)(e - 1) * B10, c1[e - 2], t);
                        c1[e] = e <= E1 ? t : 0.0;
                    }
#pragma unroll
                    for (int e = 0; e < GRAD_NE; ++e) {
                        if (e <= E1) I[e * FS] = c1[e];
                        c2[e] = 0.0;
                    }
                    for (int f = 1; f <= F1; ++f) {
                        const double bf = (double)(f - 1) * B01;
#pragma unroll
                        for (int e = 0; e < GRAD_NE; ++e) {
                            double t = fma(D00, c1[e], bf * c2[e]);
                            if (e >= 1) t = fma((double)e * B00, c1[e - 1], t);
                            c0[e] = e <= E1 ? t : 0.0;
                        }
#pragma unroll
                        for (int e = 0; e < GRAD_NE; ++e) {
                            if (e <= E1) I[e * FS + f] = c0[e];
                            c2[e] = c1[e];
                            c1[e] = c0[e];
                        }
                    }
                }
                __syncthreads();
                for (int r = 0; r < n; ++r) {
                    // extended transferred factors of root r, three axes
                    for (int t = lane; t < 3 * NGX; t += 64) {
                        const int ax = t >= NGX ? (t >= 2 * NGX ? 2 : 1) : 0;
                        const int g = t - ax * NGX;
                        const int ia = tb.x_a[g], ib = tb.x_b[g], ic = tb.x_c[g], id = tb.x_d[g];
                        double s = 0.0;
                        if (ia + ib <= E1) {
                            const double ABq = sel3(ax, pb.ab[0], pb.ab[1], pb.ab[2]);
                            const double CDq = sel3(ax, pk.ab[0], pk.ab[1], pk.ab[2]);
                            const double *I = I2 + (3 * r + ax) * NI;
                            double pab = 1.0;
                            for (int k = ib; k >= 0; --k) {
                                double scd = 0.0, pcd = 1.0;
                                for (int l = id; l >= 0; --l) {
                                    scd = fma(binom4(id, l) * pcd, I[(ia + k) * FS + ic + l], scd);
                                    pcd *= CDq;
                                }
                                s = fma(binom4(ib, k) * pab, scd, s);
                                pab *= ABq;
                            }
                        }
                        GX[t] = s;
                    }
                    __syncthreads();
                    // plain and derivative factors of root r
                    for (int t = lane; t < 3 * NG; t += 64) {
                        const int ax = t >= NG ? (t >= 2 * NG ? 2 : 1) : 0;
                        const int g = t - ax * NG;
                        const int ia = tb.g_a[g], ib = tb.g_b[g], ic = tb.g_c[g], id = tb.g_d[g];
                        const double *X = GX + ax * NGX;
                        const int x = ia * XBCD + ib * XCD + ic * XD + id;
                        double *F = FK + ax * 4 * NG + g;
                        F[0] = X[x];
                        F[NG] = fma(a2, X[x + XBCD], ia ? -(double)ia * X[x - XBCD] : 0.0);
                        F[2 * NG] = fma(b2, X[x + XCD], ib ? -(double)ib * X[x - XCD] : 0.0);
                        F[3 * NG] = fma(g2, X[x + XD], ic ? -(double)ic * X[x - XD] : 0.0);
                    }
                    __syncthreads();
                    int ab = ab0, cd = cd0;
                    for (int i = lane; i < NABCD; i += 64) {
                        const int o0 = tb.ab_g[ab][0] + tb.cd_g[cd][0];
                        const int o1 = tb.ab_g[ab][1] + tb.cd_g[cd][1] + 4 * NG;
                        const int o2 = tb.ab_g[ab][2] + tb.cd_g[cd][2] + 8 * NG;
                        const double G = gam[i];
                        const double gx = FK[o0], gy = FK[o1], gz = FK[o2];
                        const double yz = G * gy * gz, xz = G * gx * gz, xy = G * gx * gy;
                        gA[0] = fma(FK[o0 + NG], yz, gA[0]);
                        gA[1] = fma(FK[o1 + NG], xz, gA[1]);
                        gA[2] = fma(FK[o2 + NG], xy, gA[2]);
                        gB[0] = fma(FK[o0 + 2 * NG], yz, gB[0]);
                        gB[1] = fma(FK[o1 + 2 * NG], xz, gB[1]);
                        gB[2] = fma(FK[o2 + 2 * NG], xy, gB[2]);
                        gC[0] = fma(FK[o0 + 3 * NG], yz, gC[0]);
                        gC[1] = fma(FK[o1 + 3 * NG], xz, gC[1]);
                        gC[2] = fma(FK[o2 + 3 * NG], xy, gC[2]);
                        cd += scd;
                        ab += sab;
                        if (cd >= NCD) cd -= NCD, ++ab;
                    }
                    __syncthreads();
                }
            }
        }
        // quartet sums over the workgroup; centre D by translation invariance
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const double x = wave_sum(gA[q]), y = wave_sum(gB[q]), z = wave_sum(gC[q]);
            if (lane == 0) {
                unsafeAtomicAdd(&a.g[3 * sA + q], x);
                unsafeAtomicAdd(&a.g[3 * sB + q], y);
                unsafeAtomicAdd(&a.g[3 * sC + q], z);
                unsafeAtomicAdd(&a.g[3 * sD + q], -(x + y + z));
            }
        }
        __syncthreads();
    }
}

// dynamic LDS of the gradient kernel for a class
inline size_t grad_lds_bytes(int la, int lb, int lc, int ld) {
    const size_t na = (la + 1) * (la + 2) / 2, nbb = (lb + 1) * (lb + 2) / 2, nc = (lc + 1) * (lc + 2) / 2,
                 nd = (ld + 1) * (ld + 2) / 2;
    const size_t e1 = la + lb + 1, f1 = lc + ld + 1, n = (e1 + f1 - 1) / 2 + 1;
    const size_t ni = (e1 + 1) * (f1 + 1);
    const size_t ngx = (size_t)(la + 2) * (lb + 2) * (lc + 2) * (ld + 1);
    const size_t ng = (size_t)(la + 1) * (lb + 1) * (lc + 1) * (ld + 1);
    return (3 * n * ni + 3 * ngx + 12 * ng + na * nbb * nc * nd) * sizeof(double);
}

}  // namespace pk
