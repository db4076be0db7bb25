// eri_rys.h — Rys-quadrature shell-quartet kernels for gfx950 (FP64).
//
// One lane owns one contracted shell quartet (ab|cd) of a fixed angular class
// (LA >= LB, LC >= LD, bra class >= ket class).  For every primitive quartet the
// lane evaluates the n = L/2 + 1 point Rys rule, the 2-D integrals I_q(e, f)
// per Cartesian axis (e <= LA+LB, f <= LC+LD), transfers angular momentum
// C -> D on the 2-D integrals (1-D HRR), and accumulates the products into the
// contracted intermediates [e0|cd) for e in levels LA..LA+LB.  After the
// primitive loops the A -> B HRR runs once on the contracted intermediates.
// Large classes are split over the ket components (c, d) into chunks [C0, C1)
// so the accumulators stay in registers.
//
// Reference semantics restated: src/integrals/rys.cpp:250-280 (2-D recurrence),
// rys.cpp:367-485 (primitive assembly), os.cpp:837-879 (A->B transfer).  The
// reference applies both transfers per primitive on a 6-D buffer; here the
// C->D transfer acts on the 2-D integrals and the A->B transfer after
// contraction, which is the same linear map applied in another order.
#pragma once

#include "eri_common.h"

#define PK_RYS_COEF_QUALIFIER static __device__ const
#include "rys_table.inc"

namespace pk {

// n-point Rys rule for the measure e^{-T u}/(2 sqrt u) du on [0,1]:
// Chebyshev fits below TMAX(n), the Gauss-Hermite limit above.
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
        const double x = 2.0 * (T - (double)i) - 1.0, x2 = 2.0 * x;
        const double *c = pk_rys_coef + pk_rys_base[N] + i * (2 * N * (PK_RYS_DEG + 1));
#pragma unroll
        for (int f = 0; f < 2 * N; ++f) {
            const double *cf = c + f * (PK_RYS_DEG + 1);
            double b1 = 0.0, b2 = 0.0;
#pragma unroll
            for (int j = PK_RYS_DEG; j >= 1; --j) {
                const double t = fma(x2, b1, cf[j] - b2);
                b2 = b1;
                b1 = t;
            }
            const double v = fma(x, b1, cf[0] - b2);
            if (f < N)
                u[f] = v;
            else
                w[f - N] = v;
        }
    }
}

// 2-D integrals of one axis with the C->D transfer applied:
// G[e][c][d], e <= E, c <= LC, d <= LD.
template <int E, int LC, int LD>
__device__ __forceinline__ void axis_2d(double seed, double C00, double D00, double B00, double B10, double B01,
                                        double CDq, double (&G)[E + 1][LC + 1][LD + 1]) {
    constexpr int F = LC + LD;
    double I[E + 1][F + 1];
    I[0][0] = seed;
#pragma unroll
    for (int e = 1; e <= E; ++e) {
        I[e][0] = C00 * I[e - 1][0];
        if (e >= 2) I[e][0] = fma((double)(e - 1) * B10, I[e - 2][0], I[e][0]);
    }
#pragma unroll
    for (int f = 1; f <= F; ++f) {
        I[0][f] = D00 * I[0][f - 1];
        if (f >= 2) I[0][f] = fma((double)(f - 1) * B01, I[0][f - 2], I[0][f]);
#pragma unroll
        for (int e = 1; e <= E; ++e) {
            double t = fma(D00, I[e][f - 1], (double)e * B00 * I[e - 1][f - 1]);
            if (f >= 2) t = fma((double)(f - 1) * B01, I[e][f - 2], t);
            I[e][f] = t;
        }
    }
    if constexpr (LD == 0) {
#pragma unroll
        for (int e = 0; e <= E; ++e)
#pragma unroll
            for (int c = 0; c <= LC; ++c) G[e][c][0] = I[e][c];
    } else {
        double H[E + 1][F + 1][LD + 1];
#pragma unroll
        for (int e = 0; e <= E; ++e)
#pragma unroll
            for (int c = 0; c <= F; ++c) H[e][c][0] = I[e][c];
#pragma unroll
        for (int d = 1; d <= LD; ++d)
#pragma unroll
            for (int c = 0; c <= F - d; ++c)
#pragma unroll
                for (int e = 0; e <= E; ++e) H[e][c][d] = fma(CDq, H[e][c][d - 1], H[e][c + 1][d - 1]);
#pragma unroll
        for (int e = 0; e <= E; ++e)
#pragma unroll
            for (int c = 0; c <= LC; ++c)
#pragma unroll
                for (int d = 0; d <= LD; ++d) G[e][c][d] = H[e][c][d];
    }
}

// Class constants.
template <int LA, int LB, int LC, int LD>
struct Cls {
    static constexpr int E = LA + LB, F = LC + LD, L = E + F, NR = L / 2 + 1;
    static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    static constexpr int NE = nlev(LA, E);
    static constexpr int NCD = NC * ND;
    static constexpr int NABCD = NA * NB * NC * ND;
};

// One primitive quartet: accumulate [e0|cd) for the chunk [C0, C1).
template <int LA, int LB, int LC, int LD, int C0, int C1>
__device__ __forceinline__ void prim_quartet(const PrimDev &p, const PrimDev &q, const double cd[3],
                                             double (&acc)[Cls<LA, LB, LC, LD>::NE][C1 - C0]) {
    using K = Cls<LA, LB, LC, LD>;
    constexpr int NR = K::NR, E = K::E;
    const double zeta = p.zeta, eta = q.zeta;
    const double inv_de = 1.0 / (zeta + eta);
    const double rho = zeta * eta * inv_de;
    const double PQx = p.p[0] - q.p[0], PQy = p.p[1] - q.p[1], PQz = p.p[2] - q.p[2];
    const double T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);
    // W - P = -eta/(zeta+eta) PQ,  W - Q = zeta/(zeta+eta) PQ
    const double fp = -eta * inv_de, fq = zeta * inv_de;
    const double pref = p.kc * q.kc * 2.0 * sqrt(rho * 0.31830988618379067154);  // 2 sqrt(rho/pi)
    double u[NR], w[NR];
    rys_rule<NR>(T, u, w);
    const double i2z = 0.5 / zeta, i2e = 0.5 / eta, roz = rho / zeta, roe = rho / eta;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const double ur = u[r];
        const double B00 = 0.5 * inv_de * ur;
        const double B10 = i2z * (1.0 - roz * ur);
        const double B01 = i2e * (1.0 - roe * ur);
        const double uwp = ur * fp, uwq = ur * fq;
        double Gx[E + 1][LC + 1][LD + 1], Gy[E + 1][LC + 1][LD + 1], Gz[E + 1][LC + 1][LD + 1];
        axis_2d<E, LC, LD>(pref * w[r], fma(uwp, PQx, p.pa[0]), fma(uwq, PQx, q.pa[0]), B00, B10, B01, cd[0], Gx);
        axis_2d<E, LC, LD>(1.0, fma(uwp, PQy, p.pa[1]), fma(uwq, PQy, q.pa[1]), B00, B10, B01, cd[1], Gy);
        axis_2d<E, LC, LD>(1.0, fma(uwp, PQz, p.pa[2]), fma(uwq, PQz, q.pa[2]), B00, B10, B01, cd[2], Gz);
#pragma unroll
        for (int ie = 0; ie < K::NE; ++ie) {
            constexpr int dummy = 0;
            (void)dummy;
            const int le = flat_level(LA, ie), ke = flat_k(LA, ie);
            const int ex = cart_x(le, ke), ey = cart_y(le, ke), ez = cart_z(le, ke);
#pragma unroll
            for (int j = 0; j < C1 - C0; ++j) {
                const int ic = (C0 + j) / K::ND, id = (C0 + j) % K::ND;
                const int cx = cart_x(LC, ic), cy = cart_y(LC, ic), cz = cart_z(LC, ic);
                const int dx = cart_x(LD, id), dy = cart_y(LD, id), dz = cart_z(LD, id);
                acc[ie][j] = fma(Gx[ex][cx][dx] * Gy[ey][cy][dy], Gz[ez][cz][dz], acc[ie][j]);
            }
        }
    }
}

// A -> B transfer on contracted intermediates of one ket component:
// in: col[NE] over levels LA..LA+LB;  out: v[NA*NB] in (a, b) order.
template <int LA, int LB, int K>
__device__ __forceinline__ void bra_hrr_step(const double *cur, const double ab[3], double *out) {
    if constexpr (K == LB) {
#pragma unroll
        for (int i = 0; i < ncart(LA) * ncart(LB); ++i) out[i] = cur[i];
    } else {
        constexpr int NK = ncart(K), NK1 = ncart(K + 1);
        constexpr int NAL = nlev(LA, LA + LB - K - 1);
        double nxt[NAL * NK1];
#pragma unroll
        for (int bp = 0; bp < NK1; ++bp) {
            const int bx = cart_x(K + 1, bp), by = cart_y(K + 1, bp), bz = cart_z(K + 1, bp);
            const int ax = bx > 0 ? 0 : (by > 0 ? 1 : 2);
            const int b = cart_index(bx - (ax == 0), by - (ax == 1), bz - (ax == 2));
#pragma unroll
            for (int ia = 0; ia < NAL; ++ia) {
                const int la = flat_level(LA, ia), ka = flat_k(LA, ia);
                const int x = cart_x(la, ka), y = cart_y(la, ka), z = cart_z(la, ka);
                const int iap = flat(LA, x + (ax == 0), y + (ax == 1), z + (ax == 2));
                nxt[ia * NK1 + bp] = fma(ab[ax], cur[ia * NK + b], cur[iap * NK + b]);
            }
        }
        bra_hrr_step<LA, LB, K + 1>(nxt, ab, out);
    }
}

struct JKArgs {
    const int *__restrict__ wave_bra;
    const int *__restrict__ wave_start;
    const int *__restrict__ wave_count;
    const int *__restrict__ kets;
    const PairDev *__restrict__ pairs;
    const PrimDev *__restrict__ prims;
    const double *__restrict__ P;  // density scaled by component norms
    const double *__restrict__ Q;  // Schwarz per function pair (normalised functions)
    double *A;                     // J accumulator, one orientation
    double *B;                     // K accumulator, one orientation
    double tol;
    int nb;
    int nwaves;
};

struct EriArgs {
    const int *__restrict__ bra;
    const int *__restrict__ ket;
    const PairDev *__restrict__ pairs;
    const PrimDev *__restrict__ prims;
    double *out;  // [nquart][NA NB NC ND] raw contracted Cartesian ERIs
    int nquart;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// J/K digestion kernel: one wave = one bra pair and up to 64 ket pairs.
template <int LA, int LB, int LC, int LD, int C0, int C1>
__global__ __launch_bounds__(256) void k_jk(JKArgs a) {
    using K = Cls<LA, LB, LC, LD>;
    constexpr int NA = K::NA, NB = K::NB, NC = K::NC, ND = K::ND, NE = K::NE, NCH = C1 - C0;
    const int wave = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (wave >= a.nwaves) return;
    const int lane = threadIdx.x & 63;
    const int bra = a.wave_bra[wave];
    const int cnt = a.wave_count[wave];
    const bool active = lane < cnt;
    const int ket = a.kets[a.wave_start[wave] + (active ? lane : 0)];
    const PairDev pb = a.pairs[bra];
    const PairDev pk = a.pairs[ket];
    const double cdv[3] = {pk.ab[0], pk.ab[1], pk.ab[2]};

    double acc[NE][NCH];
#pragma unroll
    for (int i = 0; i < NE; ++i)
#pragma unroll
        for (int j = 0; j < NCH; ++j) acc[i][j] = 0.0;

    const int nq = active ? pk.nprim : 0;
    for (int kq = 0; kq < nq; ++kq) {
        const PrimDev q = a.prims[pk.prim_off + kq];
        for (int kp = 0; kp < pb.nprim; ++kp) {
            const PrimDev p = a.prims[pb.prim_off + kp];
            prim_quartet<LA, LB, LC, LD, C0, C1>(p, q, cdv, acc);
        }
    }

    // degeneracy of the unique quartet
    double deg = 1.0;
    if (pb.diag) deg *= 0.5;
    if (pk.diag) deg *= 0.5;
    if (bra == ket) deg *= 0.5;
    if (!active) deg = 0.0;

    const int nb = a.nb;
    double jab[NA * NB];
#pragma unroll
    for (int i = 0; i < NA * NB; ++i) jab[i] = 0.0;
    double kac[NA][NC], kbc[NB][NC], kad[NA][ND], kbd[NB][ND];
#pragma unroll
    for (int i = 0; i < NA; ++i) {
#pragma unroll
        for (int c = 0; c < NC; ++c) kac[i][c] = 0.0;
#pragma unroll
        for (int d = 0; d < ND; ++d) kad[i][d] = 0.0;
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) {
#pragma unroll
        for (int c = 0; c < NC; ++c) kbc[i][c] = 0.0;
#pragma unroll
        for (int d = 0; d < ND; ++d) kbd[i][d] = 0.0;
    }
    double Pab[NA * NB], Qab[NA * NB];
#pragma unroll
    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
        for (int ib = 0; ib < NB; ++ib) {
            Pab[ia * NB + ib] = a.P[(size_t)(pb.f0 + ia) * nb + pb.f1 + ib];
            Qab[ia * NB + ib] = a.Q[(size_t)(pb.f0 + ia) * nb + pb.f1 + ib];
        }

#pragma unroll
    for (int j = 0; j < NCH; ++j) {
        const int ic = (C0 + j) / ND, id = (C0 + j) % ND;
        const int cf = pk.f0 + ic, df = pk.f1 + id;
        double col[NE];
#pragma unroll
        for (int i = 0; i < NE; ++i) col[i] = acc[i][j];
        double v[NA * NB];
        bra_hrr_step<LA, LB, 0>(col, pb.ab, v);
        const double Qcd = a.Q[(size_t)cf * nb + df];
        const double Pcd = a.P[(size_t)cf * nb + df];
        double jcd = 0.0;
#pragma unroll
        for (int ia = 0; ia < NA; ++ia) {
            const int af = pb.f0 + ia;
            const double Pac = a.P[(size_t)af * nb + cf], Pad = a.P[(size_t)af * nb + df];
#pragma unroll
            for (int ib = 0; ib < NB; ++ib) {
                const int bf = pb.f1 + ib;
                const double m = (Qab[ia * NB + ib] * Qcd >= a.tol) ? deg : 0.0;
                const double dv = m * v[ia * NB + ib];
                jcd = fma(dv, Pab[ia * NB + ib], jcd);
                jab[ia * NB + ib] = fma(dv, Pcd, jab[ia * NB + ib]);
                const double Pbd = a.P[(size_t)bf * nb + df], Pbc = a.P[(size_t)bf * nb + cf];
                kac[ia][ic] = fma(dv, Pbd, kac[ia][ic]);
                kad[ia][id] = fma(dv, Pbc, kad[ia][id]);
                kbc[ib][ic] = fma(dv, Pad, kbc[ib][ic]);
                kbd[ib][id] = fma(dv, Pac, kbd[ib][id]);
            }
        }
        if (active) unsafeAtomicAdd(&a.A[(size_t)cf * nb + df], 2.0 * jcd);
    }
    // ket components present in this chunk
    constexpr int c_lo = C0 / ND, c_hi = (C1 - 1) / ND;
    constexpr bool all_d = (C1 - C0) >= ND || (C0 % ND == 0 && C1 % ND == 0);
    if (active) {
#pragma unroll
        for (int ic = c_lo; ic <= c_hi; ++ic) {
            const int cf = pk.f0 + ic;
#pragma unroll
            for (int ia = 0; ia < NA; ++ia) unsafeAtomicAdd(&a.B[(size_t)(pb.f0 + ia) * nb + cf], kac[ia][ic]);
#pragma unroll
            for (int ib = 0; ib < NB; ++ib) unsafeAtomicAdd(&a.B[(size_t)(pb.f1 + ib) * nb + cf], kbc[ib][ic]);
        }
#pragma unroll
        for (int id = 0; id < ND; ++id) {
            bool present = all_d;
#pragma unroll
            for (int j = 0; j < NCH; ++j) present = present || ((C0 + j) % ND == id);
            if (!present) continue;
            const int df = pk.f1 + id;
#pragma unroll
            for (int ia = 0; ia < NA; ++ia) unsafeAtomicAdd(&a.B[(size_t)(pb.f0 + ia) * nb + df], kad[ia][id]);
#pragma unroll
            for (int ib = 0; ib < NB; ++ib) unsafeAtomicAdd(&a.B[(size_t)(pb.f1 + ib) * nb + df], kbd[ib][id]);
        }
    }
    // J_ab: every lane of the wave shares the bra pair
#pragma unroll
    for (int i = 0; i < NA * NB; ++i) {
        const double s = wave_sum(jab[i]);
        if (lane == 0) unsafeAtomicAdd(&a.A[(size_t)(pb.f0 + i / NB) * nb + pb.f1 + i % NB], 2.0 * s);
    }
}

// Raw contracted ERIs of arbitrary quartets of one class (Schwarz table, tests).
template <int LA, int LB, int LC, int LD, int C0, int C1>
__global__ __launch_bounds__(256) void k_eri(EriArgs a) {
    using K = Cls<LA, LB, LC, LD>;
    constexpr int NA = K::NA, NB = K::NB, NC = K::NC, ND = K::ND, NE = K::NE, NCH = C1 - C0;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.nquart) return;
    const PairDev pb = a.pairs[a.bra[q]];
    const PairDev pk = a.pairs[a.ket[q]];
    const double cdv[3] = {pk.ab[0], pk.ab[1], pk.ab[2]};
    double acc[NE][NCH];
#pragma unroll
    for (int i = 0; i < NE; ++i)
#pragma unroll
        for (int j = 0; j < NCH; ++j) acc[i][j] = 0.0;
    for (int kq = 0; kq < pk.nprim; ++kq) {
        const PrimDev qq = a.prims[pk.prim_off + kq];
        for (int kp = 0; kp < pb.nprim; ++kp) {
            const PrimDev p = a.prims[pb.prim_off + kp];
            prim_quartet<LA, LB, LC, LD, C0, C1>(p, qq, cdv, acc);
        }
    }
    double *o = a.out + (size_t)q * K::NABCD;
#pragma unroll
    for (int j = 0; j < NCH; ++j) {
        double col[NE];
#pragma unroll
        for (int i = 0; i < NE; ++i) col[i] = acc[i][j];
        double v[NA * NB];
        bra_hrr_step<LA, LB, 0>(col, pb.ab, v);
#pragma unroll
        for (int i = 0; i < NA * NB; ++i) o[(size_t)i * NC * ND + C0 + j] = v[i];
    }
}

}  // namespace pk
