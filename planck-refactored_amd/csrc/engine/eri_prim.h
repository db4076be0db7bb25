// eri_prim.h — per-primitive-quartet building blocks of the ERI kernels: the
// Boys function, the Rys rule, the Obara-Saika / Head-Gordon-Pople vertical
// recurrence and the Rys 2-D integrals, unrolled per angular class at compile
// time (sfor), so every intermediate lives in registers.  Each block is a
// template over the arithmetic type R: double for the FP64 path, float for the
// quartets of the mixed-precision path whose Schwarz bound lies below the FP32
// cutoff (the Boys large-T branch of those stays FP64).
//
// Reference semantics restated (the contracted Cartesian ERIs are the same
// numbers, computed in another order):
//   Boys F_m(T) ............ src/lookup/boys.cpp:1592-1618 (here: accurate to ~1e-15 relative)
//   OS VRR ................. src/integrals/os.cpp:618-828
//   Rys rule, 2-D VRR ...... src/integrals/rys_roots.cpp:223-235, rys.cpp:250-280, 367-485
//   A->B / C->D transfer ... src/integrals/os.cpp:837-879, 415-459
#pragma once

#include "eri_common.h"

#define PK_RYS_COEF_QUALIFIER static __device__ const
#include "rys_table.inc"
#define PK_BOYS_QUALIFIER static __device__ const
#include "boys_table.inc"

// PrimDev::kc carries K_ab sqrt(zeta) (4/pi)^(1/4), so that the product of a bra
// and a ket record times 1/sqrt(zeta+eta) is K_ab K_cd 2 sqrt(rho/pi).

#define PK_CI(v) decltype(v)::value

namespace pk {

// ---------------------------------------------------------------- helpers
// 1/sqrt(x) to full FP64 precision: hardware estimate and two Newton steps
__device__ __forceinline__ double rsqrt_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    y = fma(y, fma(-h * y, y, 0.5), y);
    y = fma(y, fma(-h * y, y, 0.5), y);
    return y;
}
__device__ __forceinline__ double rsq(double x) { return rsqrt_nr(x); }
__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }

// scalars of one primitive quartet (p on the bra, q on the ket).  hz = 1/(2 zeta)
// and he = 1/(2 eta) come from the pairs' stored reciprocals; rho/zeta =
// eta/(zeta+eta) and rho/eta = zeta/(zeta+eta) are products, so a primitive
// quartet divides nothing (an FP64 division is a ~10-instruction sequence).
template <typename R>
struct PrimQ {
    R zeta, eta, rs, inv_de, rho, PQ[3], T, pref;
    R hz, he, roz, roe;
};

template <typename R>
__device__ __forceinline__ PrimQ<R> prim_quartet(const PrimDev &p, const PrimDev &q, double hzp, double hzq) {
    PrimQ<R> s;
    s.zeta = (R)p.zeta;
    s.eta = (R)q.zeta;
    s.rs = rsq(s.zeta + s.eta);
    s.inv_de = s.rs * s.rs;
    s.rho = s.zeta * s.eta * s.inv_de;
    s.hz = (R)hzp;
    s.he = (R)hzq;
    s.roz = s.eta * s.inv_de;
    s.roe = s.zeta * s.inv_de;
    s.PQ[0] = (R)(p.p[0] - q.p[0]);
    s.PQ[1] = (R)(p.p[1] - q.p[1]);
    s.PQ[2] = (R)(p.p[2] - q.p[2]);
    s.T = s.rho * fma(s.PQ[0], s.PQ[0], fma(s.PQ[1], s.PQ[1], s.PQ[2] * s.PQ[2]));
    s.pref = (R)(p.kc * q.kc) * s.rs;
    return s;
}

// ---------------------------------------------------------------- Boys
// F_0..F_M(T): 8-term Taylor series of F_M about the nearest grid point
// (step 1/16) and downward recursion below 50, with e^{-T} = e^{-x_i} e^{-d};
// F_0 = sqrt(pi/T)/2 and the upward recursion at and above 50.  The upward
// recursion's e^{-T} term is dropped for M <= 4 (the orders of the
// Obara-Saika kernels): there it is below 2e-17 of F_M, and the large-T
// branch then costs a reciprocal square root and M products instead of an
// exponential.
template <int M>
__device__ __forceinline__ void boys(double T, double (&F)[M + 1]) {
    static_assert(M + 8 <= PK_BOYS_COLS, "Boys table too narrow for this class");
    if (T < (double)PK_BOYS_XMAX) {
        const int i = (int)fma(T, (double)PK_BOYS_STEP_INV, 0.5);
        const double d = fma((double)i, -1.0 / PK_BOYS_STEP_INV, T);
        // 24-bit multiply (full rate; i < 2^24)
        const double *row = pk_boys_table + __umul24((unsigned)i, PK_BOYS_COLS) + M;
        double s = row[7];
        s = fma(s, -d * (1.0 / 7.0), row[6]);
        s = fma(s, -d * (1.0 / 6.0), row[5]);
        s = fma(s, -d * (1.0 / 5.0), row[4]);
        s = fma(s, -d * (1.0 / 4.0), row[3]);
        s = fma(s, -d * (1.0 / 3.0), row[2]);
        s = fma(s, -d * (1.0 / 2.0), row[1]);
        s = fma(s, -d, row[0]);
        F[M] = s;
        if constexpr (M > 0) {
            // e^{-d}, |d| <= 1/32: 8 Taylor terms (next term < 1e-17)
            double e = fma(-d, 1.0 / 40320.0, 1.0 / 5040.0);
            e = fma(e, -d, 1.0 / 720.0);
            e = fma(e, -d, 1.0 / 120.0);
            e = fma(e, -d, 1.0 / 24.0);
            e = fma(e, -d, 1.0 / 6.0);
            e = fma(e, -d, 0.5);
            e = fma(e, -d, 1.0);
            e = fma(e, -d, 1.0);
            const double eT = pk_boys_exp[i] * e;
            const double t2 = 2.0 * T;
            sfor_down<M, 0>([&](auto m_) {
                constexpr int m = PK_CI(m_);
                F[m - 1] = fma(t2, F[m], eT) * (1.0 / (2 * m - 1));
            });
        }
    } else {
        const double r = rsqrt_nr(T);
        F[0] = 0.88622692545275801365 * r;  // sqrt(pi/T)/2
        if constexpr (M > 0) {
            const double it2 = 0.5 * r * r;
            if constexpr (M <= 4) {
                sfor<1, M + 1>([&](auto m_) {
                    constexpr int m = PK_CI(m_);
                    F[m] = (2 * m - 1) * F[m - 1] * it2;
                });
            } else {
                const double eT = exp(-T);
                sfor<1, M + 1>([&](auto m_) {
                    constexpr int m = PK_CI(m_);
                    F[m] = ((2 * m - 1) * F[m - 1] - eT) * it2;
                });
            }
        }
    }
}

// FP32 path: 5-term series from the float tables below 50 (|d| <= 1/32, next
// term below 3e-10 relative); the large-T branch in FP64 without the e^{-T}
// term (below 2e-17 of F_M for M <= 4, the orders of the FP32 kernels).
template <int M>
__device__ __forceinline__ void boys(float T, float (&F)[M + 1]) {
    static_assert(M + 5 <= PK_BOYS_COLS, "Boys table too narrow for this class");
    if (T < (float)PK_BOYS_XMAX) {
        const int i = (int)fmaf(T, (float)PK_BOYS_STEP_INV, 0.5f);
        const float d = fmaf((float)i, -1.0f / PK_BOYS_STEP_INV, T);
        const float *row = pk_boys_table_f + __umul24((unsigned)i, PK_BOYS_COLS) + M;
        float s = row[4];
        s = fmaf(s, -d * (1.0f / 4.0f), row[3]);
        s = fmaf(s, -d * (1.0f / 3.0f), row[2]);
        s = fmaf(s, -d * (1.0f / 2.0f), row[1]);
        s = fmaf(s, -d, row[0]);
        F[M] = s;
        if constexpr (M > 0) {
            float e = fmaf(-d, 1.0f / 120.0f, 1.0f / 24.0f);
            e = fmaf(e, -d, 1.0f / 6.0f);
            e = fmaf(e, -d, 0.5f);
            e = fmaf(e, -d, 1.0f);
            e = fmaf(e, -d, 1.0f);
            const float eT = pk_boys_exp_f[i] * e;
            const float t2 = 2.0f * T;
            sfor_down<M, 0>([&](auto m_) {
                constexpr int m = PK_CI(m_);
                F[m - 1] = fmaf(t2, F[m], eT) * (1.0f / (2 * m - 1));
            });
        }
    } else {
        const double Td = (double)T;
        const double r = rsqrt_nr(Td);
        double f = 0.88622692545275801365 * r;
        F[0] = (float)f;
        if constexpr (M > 0) {
            static_assert(M <= 4, "FP32 large-T Boys branch drops e^{-T}: orders up to 4 only");
            const double it2 = 0.5 * r * r;
            sfor<1, M + 1>([&](auto m_) {
                constexpr int m = PK_CI(m_);
                f = (2 * m - 1) * f * it2;
                F[m] = (float)f;
            });
        }
    }
}

// ---------------------------------------------------------------- Rys rule
// Root r of the n-point rule of e^{-T u}/(2 sqrt u) du on [0,1]: Chebyshev fits
// below TMAX(n), the Gauss-Hermite limit above.
template <typename R>
struct RysCtx {
    const double *c;  // coefficients of T's interval (null in the large-T limit)
    R x, it, sit;
};

template <typename R, int N>
__device__ __forceinline__ RysCtx<R> rys_ctx(R T) {
    RysCtx<R> k;
    if (T >= (R)pk_rys_tmax[N]) {
        k.c = nullptr;
        k.sit = rsq(T);
        k.it = k.sit * k.sit;
        k.x = R(0);
    } else {
        const int i = (int)T;
        k.c = pk_rys_coef + pk_rys_base[N] + i * (2 * N * (PK_RYS_DEG + 1));
        k.x = R(2) * (T - (R)i) - R(1);
        k.it = k.sit = R(0);
    }
    return k;
}

// Clenshaw sum of a Chebyshev series (coefficients FP64, arithmetic in R)
template <typename R>
__device__ __forceinline__ R cheb(const double *cf, R x) {
    const R x2 = R(2) * x;
    R b1 = R(0), b2 = R(0);
#pragma unroll
    for (int j = PK_RYS_DEG; j >= 1; --j) {
        const R t = fma(x2, b1, (R)cf[j] - b2);
        b2 = b1;
        b1 = t;
    }
    return fma(x, b1, (R)cf[0] - b2);
}

template <typename R, int N>
__device__ __forceinline__ void rys_root(const RysCtx<R> &k, int r, R &u, R &w) {
    if (k.c == nullptr) {
        u = (R)pk_rys_ax2[pk_rys_abase[N] + r] * k.it;
        w = (R)pk_rys_aw[pk_rys_abase[N] + r] * k.sit;
    } else {
        u = cheb<R>(k.c + r * (PK_RYS_DEG + 1), k.x);
        w = cheb<R>(k.c + (N + r) * (PK_RYS_DEG + 1), k.x);
    }
}

// ---------------------------------------------------------------- class constants
template <int LA, int LB, int LC, int LD>
struct Cls {
    static constexpr int E = LA + LB, F = LC + LD, L = E + F, NR = L / 2 + 1;
    static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    static constexpr int NAB = NA * NB, NCD = NC * ND;
    static constexpr int NE = nlev(LA, E);  // [e0| for levels LA..LA+LB
    static constexpr int NF = nlev(LC, F);  // |f0] for levels LC..LC+LD
    static constexpr int NABCD = NAB * NCD;
};

// axis used to build component (x,y,z) from a lower one: the first non-zero
__host__ __device__ constexpr int build_axis(int x, int y, int z) { return x > 0 ? 0 : (y > 0 ? 1 : 2); }
__host__ __device__ constexpr int comp_q(int l, int k, int q) {
    return q == 0 ? cart_x(l, k) : (q == 1 ? cart_y(l, k) : cart_z(l, k));
}
// flat index (levels from 0) of component k of level l shifted by s along axis q
__host__ __device__ constexpr int flat0_shift(int l, int k, int q, int s) {
    return flat(0, cart_x(l, k) + (q == 0 ? s : 0), cart_y(l, k) + (q == 1 ? s : 0), cart_z(l, k) + (q == 2 ? s : 0));
}

// ---------------------------------------------------------------- OS / HGP VRR
// Accumulate [e0|f0]^(0) for |e| in [LA, E], |f| in [LC, F] of one primitive
// quartet into acc[NE][NF].  V[e][f][m] holds [e0|f0]^(m) for all |e| <= E,
// |f| <= F (flat level-0 indices): bra side first, then the ket side with the
// bra-ket coupling term (Obara & Saika 1986, Head-Gordon & Pople 1988).
template <typename R, int LA, int LB, int LC, int LD>
__device__ __forceinline__ void os_prim(const PrimDev &p, const PrimDev &q, double hzp, double hzq,
                                        R (&acc)[Cls<LA, LB, LC, LD>::NE][Cls<LA, LB, LC, LD>::NF]) {
    using K = Cls<LA, LB, LC, LD>;
    constexpr int E = K::E, F = K::F, L = K::L;
    constexpr int NLE = nlev(0, E), NLF = nlev(0, F);
    const PrimQ<R> s = prim_quartet<R>(p, q, hzp, hzq);
    const R zeta = s.zeta, eta = s.eta, inv_de = s.inv_de, rho = s.rho;
    const R PQ[3] = {s.PQ[0], s.PQ[1], s.PQ[2]};
    R Fm[L + 1];
    boys<L>(s.T, Fm);
    R V[NLE][NLF][L + 1];
    sfor<0, L + 1>([&](auto m_) { V[0][0][PK_CI(m_)] = s.pref * Fm[PK_CI(m_)]; });
    const R fp = -eta * inv_de, fq = zeta * inv_de;  // W - P = fp PQ, W - Q = fq PQ
    const R i2z = s.hz, roz = s.roz, i2e = s.he, roe = s.roe, i2d = R(0.5) * inv_de;
    sfor<1, E + 1>([&](auto le_) {
        constexpr int le = PK_CI(le_);
        sfor<0, ncart(le)>([&](auto k_) {
            constexpr int k = PK_CI(k_);
            constexpr int ax = build_axis(cart_x(le, k), cart_y(le, k), cart_z(le, k));
            constexpr int ei = comp_q(le, k, ax);
            constexpr int e = flat(0, cart_x(le, k), cart_y(le, k), cart_z(le, k));
            constexpr int e1 = flat0_shift(le, k, ax, -1);
            const R pa = (R)p.pa[ax], wp = fp * PQ[ax];
            sfor<0, L - le + 1>([&](auto m_) {
                constexpr int m = PK_CI(m_);
                R t = fma(pa, V[e1][0][m], wp * V[e1][0][m + 1]);
                if constexpr (ei >= 2) {
                    constexpr int e2 = flat0_shift(le, k, ax, -2);
                    t = fma(R(ei - 1) * i2z, fma(-roz, V[e2][0][m + 1], V[e2][0][m]), t);
                }
                V[e][0][m] = t;
            });
        });
    });
    sfor<1, F + 1>([&](auto lf_) {
        constexpr int lf = PK_CI(lf_);
        sfor<0, ncart(lf)>([&](auto kf_) {
            constexpr int kf = PK_CI(kf_);
            constexpr int ax = build_axis(cart_x(lf, kf), cart_y(lf, kf), cart_z(lf, kf));
            constexpr int fi = comp_q(lf, kf, ax);
            constexpr int f = flat(0, cart_x(lf, kf), cart_y(lf, kf), cart_z(lf, kf));
            constexpr int f1 = flat0_shift(lf, kf, ax, -1);
            const R qc = (R)q.pa[ax], wq = fq * PQ[ax];
            sfor<0, E + 1>([&](auto le_) {
                constexpr int le = PK_CI(le_);
                sfor<0, ncart(le)>([&](auto ke_) {
                    constexpr int ke = PK_CI(ke_);
                    constexpr int e = nlev(0, le - 1) + ke;
                    constexpr int ei = comp_q(le, ke, ax);
                    sfor<0, L - le - lf + 1>([&](auto m_) {
                        constexpr int m = PK_CI(m_);
                        R t = fma(qc, V[e][f1][m], wq * V[e][f1][m + 1]);
                        if constexpr (fi >= 2) {
                            constexpr int f2 = flat0_shift(lf, kf, ax, -2);
                            t = fma(R(fi - 1) * i2e, fma(-roe, V[e][f2][m + 1], V[e][f2][m]), t);
                        }
                        if constexpr (ei >= 1) {
                            constexpr int em = flat0_shift(le, ke, ax, -1);
                            t = fma(R(ei) * i2d, V[em][f1][m + 1], t);
                        }
                        V[e][f][m] = t;
                    });
                });
            });
        });
    });
    sfor<0, K::NE>([&](auto ie_) {
        sfor<0, K::NF>([&](auto jf_) {
            constexpr int ie = PK_CI(ie_), jf = PK_CI(jf_);
            acc[ie][jf] += V[nlev(0, LA - 1) + ie][nlev(0, LC - 1) + jf][0];
        });
    });
}

// ---------------------------------------------------------------- Rys 2-D integrals
// G[e][c][d] of one axis: 2-D recurrence (e <= E, f <= LC+LD) then the C->D
// transfer on the ket index.
template <typename R, int E, int LC, int LD>
__device__ __forceinline__ void axis_2d(R seed, R C00, R D00, R B00, R B10, R B01, R CDq,
                                        R (&G)[E + 1][LC + 1][LD + 1]) {
    constexpr int F = LC + LD;
    R I[E + 1][F + 1];
    I[0][0] = seed;
    sfor<1, E + 1>([&](auto e_) {
        constexpr int e = PK_CI(e_);
        I[e][0] = C00 * I[e - 1][0];
        if constexpr (e >= 2) I[e][0] = fma(R(e - 1) * B10, I[e - 2][0], I[e][0]);
    });
    sfor<1, F + 1>([&](auto f_) {
        constexpr int f = PK_CI(f_);
        I[0][f] = D00 * I[0][f - 1];
        if constexpr (f >= 2) I[0][f] = fma(R(f - 1) * B01, I[0][f - 2], I[0][f]);
        sfor<1, E + 1>([&](auto e_) {
            constexpr int e = PK_CI(e_);
            R t = fma(D00, I[e][f - 1], R(e) * B00 * I[e - 1][f - 1]);
            if constexpr (f >= 2) t = fma(R(f - 1) * B01, I[e][f - 2], t);
            I[e][f] = t;
        });
    });
    R H[E + 1][F + 1][LD + 1];
    sfor<0, E + 1>([&](auto e_) { sfor<0, F + 1>([&](auto c_) { H[PK_CI(e_)][PK_CI(c_)][0] = I[PK_CI(e_)][PK_CI(c_)]; }); });
    sfor<1, LD + 1>([&](auto d_) {
        constexpr int d = PK_CI(d_);
        sfor<0, F - d + 1>([&](auto c_) {
            constexpr int c = PK_CI(c_);
            sfor<0, E + 1>([&](auto e_) {
                constexpr int e = PK_CI(e_);
                H[e][c][d] = fma(CDq, H[e][c][d - 1], H[e][c + 1][d - 1]);
            });
        });
    });
    sfor<0, E + 1>([&](auto e_) {
        sfor<0, LC + 1>([&](auto c_) {
            sfor<0, LD + 1>([&](auto d_) {
                G[PK_CI(e_)][PK_CI(c_)][PK_CI(d_)] = H[PK_CI(e_)][PK_CI(c_)][PK_CI(d_)];
            });
        });
    });
}

// One primitive quartet by Rys quadrature: accumulate [e0|cd) of the ket
// components [C0, C1) into acc[NE][C1 - C0]; one root per loop trip.
template <typename R, int LA, int LB, int LC, int LD, int C0, int C1>
__device__ __forceinline__ void rys_prim(const PrimDev &p, const PrimDev &q, double hzp, double hzq, const R cd[3],
                                         R (&acc)[Cls<LA, LB, LC, LD>::NE][C1 - C0]) {
    using K = Cls<LA, LB, LC, LD>;
    constexpr int NR = K::NR, E = K::E;
    const PrimQ<R> s = prim_quartet<R>(p, q, hzp, hzq);
    const R zeta = s.zeta, eta = s.eta, inv_de = s.inv_de, rho = s.rho;
    const R PQx = s.PQ[0], PQy = s.PQ[1], PQz = s.PQ[2];
    const R fp = -eta * inv_de, fq = zeta * inv_de;
    const R i2z = s.hz, i2e = s.he, roz = s.roz, roe = s.roe;
    const RysCtx<R> rk = rys_ctx<R, NR>(s.T);
#pragma unroll 1
    for (int r = 0; r < NR; ++r) {
        R ur, wr;
        rys_root<R, NR>(rk, r, ur, wr);
        const R B00 = R(0.5) * inv_de * ur;
        const R B10 = i2z * (R(1) - roz * ur);
        const R B01 = i2e * (R(1) - roe * ur);
        const R uwp = ur * fp, uwq = ur * fq;
        R Gx[E + 1][LC + 1][LD + 1], Gy[E + 1][LC + 1][LD + 1], Gz[E + 1][LC + 1][LD + 1];
        axis_2d<R, E, LC, LD>(s.pref * wr, fma(uwp, PQx, (R)p.pa[0]), fma(uwq, PQx, (R)q.pa[0]), B00, B10, B01, cd[0],
                              Gx);
        axis_2d<R, E, LC, LD>(R(1), fma(uwp, PQy, (R)p.pa[1]), fma(uwq, PQy, (R)q.pa[1]), B00, B10, B01, cd[1], Gy);
        axis_2d<R, E, LC, LD>(R(1), fma(uwp, PQz, (R)p.pa[2]), fma(uwq, PQz, (R)q.pa[2]), B00, B10, B01, cd[2], Gz);
        sfor<0, K::NE>([&](auto ie_) {
            constexpr int ie = PK_CI(ie_);
            constexpr int le = flat_level(LA, ie), ke = flat_k(LA, ie);
            constexpr int ex = cart_x(le, ke), ey = cart_y(le, ke), ez = cart_z(le, ke);
            sfor<0, C1 - C0>([&](auto j_) {
                constexpr int j = PK_CI(j_);
                constexpr int ic = (C0 + j) / K::ND, id = (C0 + j) % K::ND;
                constexpr int cx = cart_x(LC, ic), cy = cart_y(LC, ic), cz = cart_z(LC, ic);
                constexpr int dx = cart_x(LD, id), dy = cart_y(LD, id), dz = cart_z(LD, id);
                acc[ie][j] = fma(Gx[ex][cx][dx] * Gy[ey][cy][dy], Gz[ez][cz][dz], acc[ie][j]);
            });
        });
    }
}

// ---------------------------------------------------------------- transfers
// A -> B transfer: cur over (a levels LA..LA+LB-K) x (b level K) -> out (a, b)
// of levels (LA, LB).  Used for the bra (A -> B) and, with the ket's angular
// momenta, for the ket (C -> D).
template <typename R, int LA, int LB, int K>
__device__ __forceinline__ void hrr_step(const R (&cur)[nlev(LA, LA + LB - K) * ncart(K)], const double ab[3],
                                         R (&out)[ncart(LA) * ncart(LB)]) {
    if constexpr (K == LB) {
        sfor<0, ncart(LA) * ncart(LB)>([&](auto i_) { out[PK_CI(i_)] = cur[PK_CI(i_)]; });
    } else {
        constexpr int NK = ncart(K), NK1 = ncart(K + 1);
        constexpr int NAL = nlev(LA, LA + LB - K - 1);
        R nxt[NAL * NK1];
        sfor<0, NK1>([&](auto bp_) {
            constexpr int bp = PK_CI(bp_);
            constexpr int bx = cart_x(K + 1, bp), by = cart_y(K + 1, bp), bz = cart_z(K + 1, bp);
            constexpr int ax = build_axis(bx, by, bz);
            constexpr int b = cart_index(bx - (ax == 0), by - (ax == 1), bz - (ax == 2));
            const R abq = (R)ab[ax];
            sfor<0, NAL>([&](auto ia_) {
                constexpr int ia = PK_CI(ia_);
                constexpr int la = flat_level(LA, ia), ka = flat_k(LA, ia);
                constexpr int iap = flat(LA, cart_x(la, ka) + (ax == 0), cart_y(la, ka) + (ax == 1),
                                         cart_z(la, ka) + (ax == 2));
                nxt[ia * NK1 + bp] = fma(abq, cur[ia * NK + b], cur[iap * NK + b]);
            });
        });
        hrr_step<R, LA, LB, K + 1>(nxt, ab, out);
    }
}

}  // namespace pk
