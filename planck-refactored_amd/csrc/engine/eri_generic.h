// eri_generic.h — cooperative shell-quartet kernel for any class up to (ff|ff).
//
// One 64-lane workgroup owns one contracted shell quartet at a time; the class
// (la, lb, lc, ld) is a runtime value, so one compiled kernel serves every class
// whose contracted intermediates do not fit a single lane (large d/f classes),
// and every ERI-output request (Schwarz table, ERI tensor).
//
// Per primitive quartet: lanes r < n evaluate the Rys roots; lanes t < 3n build
// the 2-D integrals of root t/3 and axis t%3 in registers and write the fully
// transferred 1-D factors G[a][b][c][d] (closed-form transfers
//   (a|b) = sum_k C(b,k) (A-B)^(b-k) (a+k|0),  (c|d) likewise with C-D)
// to LDS; then every lane adds sum_r Gx Gy Gz for its share of the (ab|cd)
// components into LDS accumulators.  J/K digestion reduces each output element
// in one lane (no atomic contention inside the quartet), then FP64 atomics.
// Reference semantics as in eri_prim.h (rys.cpp:250-485, os.cpp:837-879).
#pragma once

#include "eri_prim.h"

namespace pk {

constexpr int GEN_MAX_L = 3;                       // per shell
constexpr int GEN_NR = 7;                          // roots for L <= 12
constexpr int GEN_D = GEN_MAX_L + 1;               // per-axis a, b, c, d extent
constexpr int GEN_G = GEN_D * GEN_D * GEN_D * GEN_D;  // G entries per (root, axis)
constexpr int GEN_MAXC = 10 * 10 * 10 * 10;        // (ff|ff) components

struct GenArgs {
    // JK mode: wave lists of one class (as JKArgs); ERI mode: explicit quartets
    const int *__restrict__ wave_bra;
    const int *__restrict__ wave_start;
    const int *__restrict__ wave_count;
    const int *__restrict__ wave_mol;
    const int *__restrict__ kets;
    const long long *__restrict__ mol_off;  // JK mode: packed matrices per molecule
    const int *__restrict__ mol_nb;
    const int *__restrict__ qbra;  // ERI mode
    const int *__restrict__ qket;
    const PairDev *__restrict__ pairs;
    const PrimDev *__restrict__ prims;
    const double *__restrict__ P;
    const double *__restrict__ Q;
    double *A;
    double *B;
    double *out;  // ERI mode: [nitems][NA NB NC ND]
    double tol;
    int nitems;  // waves (JK) or quartets (ERI)
    int la, lb, lc, ld;
    int eri_mode;
};

__device__ __forceinline__ int comp_axis(int l, int k, int q) {
    int lx = l;
    while (k > l - lx) {
        k -= l - lx + 1;
        --lx;
    }
    return q == 0 ? lx : (q == 1 ? l - lx - k : k);
}

// C(n, k) for 0 <= k <= n <= 3
__device__ __forceinline__ double binom(int n, int k) {
    return (n == 3 && (k == 1 || k == 2)) ? 3.0 : ((n == 2 && k == 1) ? 2.0 : 1.0);
}

__device__ __forceinline__ double sel3(int ax, double x, double y, double z) {
    return ax == 0 ? x : (ax == 1 ? y : z);
}

// Rys rule with a runtime root count (the kernels' tables)
__device__ __forceinline__ void rys_root_rt(int n, double T, int r, double &u, double &w) {
    if (T >= pk_rys_tmax[n]) {
        const double s = rsqrt_nr(T);
        u = pk_rys_ax2[pk_rys_abase[n] + r] * s * s;
        w = pk_rys_aw[pk_rys_abase[n] + r] * s;
    } else {
        const int i = (int)T;
        const double x = 2.0 * (T - (double)i) - 1.0;
        const double *c = pk_rys_coef + pk_rys_base[n] + i * (2 * n * (PK_RYS_DEG + 1));
        u = cheb(c + r * (PK_RYS_DEG + 1), x);
        w = cheb(c + (n + r) * (PK_RYS_DEG + 1), x);
    }
}

__global__ __launch_bounds__(64) void k_generic(GenArgs a) {
    extern __shared__ double smem[];
    __shared__ double su[GEN_NR], sw[GEN_NR];
    const int lane = threadIdx.x;
    const int item = blockIdx.x;
    if (item >= a.nitems) return;
    const int la = a.la, lb = a.lb, lc = a.lc, ld = a.ld;
    const int NA = (la + 1) * (la + 2) / 2, NB = (lb + 1) * (lb + 2) / 2;
    const int NC = (lc + 1) * (lc + 2) / 2, ND = (ld + 1) * (ld + 2) / 2;
    const int NCD = NC * ND, NABCD = NA * NB * NCD;
    const int E = la + lb, F = lc + ld, n = (E + F) / 2 + 1;
    double *G = smem;               // [n][3][GEN_G]
    double *acc = G + n * 3 * GEN_G;  // [NABCD]

    int bra, kbeg, kcnt, nb = 0;
    const double *P = nullptr, *Qm = nullptr;
    double *Am = nullptr, *Bm = nullptr;
    if (a.eri_mode) {
        bra = a.qbra[item];
        kbeg = item;
        kcnt = 1;
    } else {
        bra = a.wave_bra[item];
        kbeg = a.wave_start[item];
        kcnt = a.wave_count[item];
        const int mol = a.wave_mol[item];
        const long long moff = a.mol_off[mol];
        nb = a.mol_nb[mol];
        P = a.P + moff;
        Qm = a.Q + moff;
        Am = a.A + moff;
        Bm = a.B + moff;
    }
    const PairDev pb = a.pairs[bra];

    for (int kk = 0; kk < kcnt; ++kk) {
        const int ket = a.eri_mode ? a.qket[item] : a.kets[kbeg + kk];
        const PairDev pk = a.pairs[ket];
        for (int i = lane; i < NABCD; i += 64) acc[i] = 0.0;
        __syncthreads();
        for (int kq = 0; kq < pk.nprim; ++kq) {
            const PrimDev q = a.prims[pk.prim_off + kq];
            for (int kp = 0; kp < pb.nprim; ++kp) {
                const PrimDev p = a.prims[pb.prim_off + kp];
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
                    const int r = lane / 3, ax = lane % 3;
                    const double ur = su[r];
                    const double B00 = 0.5 * inv_de * ur;
                    const double B10 = (0.5 / p.zeta) * (1.0 - rho / p.zeta * ur);
                    const double B01 = (0.5 / q.zeta) * (1.0 - rho / q.zeta * ur);
                    const double PQa = sel3(ax, PQ[0], PQ[1], PQ[2]);
                    const double C00 = fma(ur * (-q.zeta * inv_de), PQa, sel3(ax, p.pa[0], p.pa[1], p.pa[2]));
                    const double D00 = fma(ur * (p.zeta * inv_de), PQa, sel3(ax, q.pa[0], q.pa[1], q.pa[2]));
                    double I[7][7];
#pragma unroll
                    for (int e = 0; e < 7; ++e)
#pragma unroll
                        for (int f = 0; f < 7; ++f) I[e][f] = 0.0;
                    I[0][0] = ax == 0 ? pref * sw[r] : 1.0;
#pragma unroll
                    for (int e = 1; e < 7; ++e) {
                        double t = C00 * I[e - 1][0];
                        if (e >= 2) t = fma((double)(e - 1) * B10, I[e - 2][0], t);
                        I[e][0] = e <= E ? t : 0.0;
                    }
#pragma unroll
                    for (int f = 1; f < 7; ++f) {
                        if (f <= F) {
                            double t0 = D00 * I[0][f - 1];
                            if (f >= 2) t0 = fma((double)(f - 1) * B01, I[0][f - 2], t0);
                            I[0][f] = t0;
#pragma unroll
                            for (int e = 1; e < 7; ++e) {
                                double t = fma(D00, I[e][f - 1], (double)e * B00 * I[e - 1][f - 1]);
                                if (f >= 2) t = fma((double)(f - 1) * B01, I[e][f - 2], t);
                                I[e][f] = e <= E ? t : 0.0;
                            }
                        }
                    }
                    const double ABq = sel3(ax, pb.ab[0], pb.ab[1], pb.ab[2]);
                    const double CDq = sel3(ax, pk.ab[0], pk.ab[1], pk.ab[2]);
                    double *g = G + (r * 3 + ax) * GEN_G;
#pragma unroll
                    for (int ia = 0; ia < GEN_D; ++ia)
#pragma unroll
                        for (int ib = 0; ib < GEN_D; ++ib)
#pragma unroll
                            for (int ic = 0; ic < GEN_D; ++ic)
#pragma unroll
                                for (int id = 0; id < GEN_D; ++id) {
                                    if (ia > la || ib > lb || ic > lc || id > ld) continue;
                                    double s = 0.0, pab = 1.0;
#pragma unroll
                                    for (int k = ib; k >= 0; --k) {  // (A-B)^(ib-k) C(ib,k)
                                        double scd = 0.0, pcd = 1.0;
#pragma unroll
                                        for (int l = id; l >= 0; --l) {
                                            scd = fma(binom(id, l) * pcd, I[ia + k][ic + l], scd);
                                            pcd *= CDq;
                                        }
                                        s = fma(binom(ib, k) * pab, scd, s);
                                        pab *= ABq;
                                    }
                                    g[((ia * GEN_D + ib) * GEN_D + ic) * GEN_D + id] = s;
                                }
                }
                __syncthreads();
                for (int i = lane; i < NABCD; i += 64) {
                    int t = i;
                    const int id = t % ND;
                    t /= ND;
                    const int ic = t % NC;
                    t /= NC;
                    const int ib = t % NB;
                    const int ia = t / NB;
                    int o[3];
#pragma unroll
                    for (int q3 = 0; q3 < 3; ++q3)
                        o[q3] = ((comp_axis(la, ia, q3) * GEN_D + comp_axis(lb, ib, q3)) * GEN_D +
                                 comp_axis(lc, ic, q3)) * GEN_D + comp_axis(ld, id, q3);
                    double s = 0.0;
                    for (int r = 0; r < n; ++r)
                        s = fma(G[(r * 3) * GEN_G + o[0]] * G[(r * 3 + 1) * GEN_G + o[1]], G[(r * 3 + 2) * GEN_G + o[2]], s);
                    acc[i] += s;
                }
                __syncthreads();
            }
        }
        if (a.eri_mode) {
            double *o = a.out + (size_t)item * NABCD;
            for (int i = lane; i < NABCD; i += 64) o[i] = acc[i];
            __syncthreads();
            continue;
        }
        // ---- J/K digestion of the quartet (bra pb, ket pk)
        double deg = 1.0;
        if (pb.diag) deg *= 0.5;
        if (pk.diag) deg *= 0.5;
        if (bra == ket) deg *= 0.5;
        const bool all_pass = pb.qmin * pk.qmin >= a.tol;
        // apply degeneracy and the function-level Schwarz test in place
        for (int i = lane; i < NABCD; i += 64) {
            double m = deg;
            if (!all_pass) {
                const int cd = i % NCD, ab = i / NCD;
                const int af = pb.f0 + ab / NB, bf = pb.f1 + ab % NB, cf = pk.f0 + cd / ND, df = pk.f1 + cd % ND;
                if (Qm[(size_t)af * nb + bf] * Qm[(size_t)cf * nb + df] < a.tol) m = 0.0;
            }
            acc[i] *= m;
        }
        __syncthreads();
        auto Pv = [&](int r, int c) { return P[(size_t)r * nb + c]; };
        // A_ab += 2 sum_cd v P_cd
        for (int ab = lane; ab < NA * NB; ab += 64) {
            double s = 0.0;
            for (int cd = 0; cd < NCD; ++cd) s = fma(acc[ab * NCD + cd], Pv(pk.f0 + cd / ND, pk.f1 + cd % ND), s);
            unsafeAtomicAdd(&Am[(size_t)(pb.f0 + ab / NB) * nb + pb.f1 + ab % NB], 2.0 * s);
        }
        // A_cd += 2 sum_ab v P_ab
        for (int cd = lane; cd < NCD; cd += 64) {
            double s = 0.0;
            for (int ab = 0; ab < NA * NB; ++ab) s = fma(acc[ab * NCD + cd], Pv(pb.f0 + ab / NB, pb.f1 + ab % NB), s);
            unsafeAtomicAdd(&Am[(size_t)(pk.f0 + cd / ND) * nb + pk.f1 + cd % ND], 2.0 * s);
        }
        // B_ac += sum_bd v P_bd ; B_bc += sum_ad v P_ad ; B_ad += sum_bc v P_bc ; B_bd += sum_ac v P_ac
        for (int t = lane; t < NA * NC; t += 64) {
            const int ia = t / NC, ic = t % NC;
            double s = 0.0;
            for (int ib = 0; ib < NB; ++ib)
                for (int id = 0; id < ND; ++id)
                    s = fma(acc[(ia * NB + ib) * NCD + ic * ND + id], Pv(pb.f1 + ib, pk.f1 + id), s);
            unsafeAtomicAdd(&Bm[(size_t)(pb.f0 + ia) * nb + pk.f0 + ic], s);
        }
        for (int t = lane; t < NB * NC; t += 64) {
            const int ib = t / NC, ic = t % NC;
            double s = 0.0;
            for (int ia = 0; ia < NA; ++ia)
                for (int id = 0; id < ND; ++id)
                    s = fma(acc[(ia * NB + ib) * NCD + ic * ND + id], Pv(pb.f0 + ia, pk.f1 + id), s);
            unsafeAtomicAdd(&Bm[(size_t)(pb.f1 + ib) * nb + pk.f0 + ic], s);
        }
        for (int t = lane; t < NA * ND; t += 64) {
            const int ia = t / ND, id = t % ND;
            double s = 0.0;
            for (int ib = 0; ib < NB; ++ib)
                for (int ic = 0; ic < NC; ++ic)
                    s = fma(acc[(ia * NB + ib) * NCD + ic * ND + id], Pv(pb.f1 + ib, pk.f0 + ic), s);
            unsafeAtomicAdd(&Bm[(size_t)(pb.f0 + ia) * nb + pk.f1 + id], s);
        }
        for (int t = lane; t < NB * ND; t += 64) {
            const int ib = t / ND, id = t % ND;
            double s = 0.0;
            for (int ia = 0; ia < NA; ++ia)
                for (int ic = 0; ic < NC; ++ic)
                    s = fma(acc[(ia * NB + ib) * NCD + ic * ND + id], Pv(pb.f0 + ia, pk.f0 + ic), s);
            unsafeAtomicAdd(&Bm[(size_t)(pb.f1 + ib) * nb + pk.f1 + id], s);
        }
        __syncthreads();
    }
}

// dynamic LDS of the generic kernel for a class: G for n roots + accumulators
inline size_t generic_lds_bytes(int la, int lb, int lc, int ld) {
    const size_t nabcd = (size_t)((la + 1) * (la + 2) / 2) * ((lb + 1) * (lb + 2) / 2) * ((lc + 1) * (lc + 2) / 2) *
                         ((ld + 1) * (ld + 2) / 2);
    const int n = (la + lb + lc + ld) / 2 + 1;
    return ((size_t)n * 3 * GEN_G + nabcd) * sizeof(double);
}

}  // namespace pk
