// eri_generic.h — cooperative shell-quartet kernel for any class up to (ff|ff).
//
// One 64-lane workgroup owns one contracted shell quartet at a time; the class
// (la, lb, lc, ld) is a runtime value, so one compiled kernel serves every class
// whose contracted intermediates do not fit a single lane (large d/f classes),
// and every ERI-output request (Schwarz table, ERI tensor).
//
// Per primitive quartet: lanes r < n evaluate the Rys roots; lanes t < 3n build
// the 2-D integrals of root t/3 and axis t%3 into LDS; all lanes then form the
// transferred 1-D factors G[a][b][c][d] (closed-form transfers
//   (a|b) = sum_k C(b,k) (A-B)^(b-k) (a+k|0),  (c|d) likewise with C-D)
// and add sum_r Gx Gy Gz for their share of the (ab|cd) components into LDS
// accumulators.  J/K digestion reduces each output element
// in one lane (no atomic contention inside the quartet), then FP64 atomics.
// Reference semantics as in eri_prim.h (rys.cpp:250-485, os.cpp:837-879).
#pragma once

#include "eri_prim.h"

namespace pk {

constexpr int GEN_MAX_L = 3;  // per shell
constexpr int GEN_NR = 7;     // roots for L <= 12; also the 2-D extent E + 1 <= 7

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
    const double *__restrict__ hz;   // 1/(2 zeta) per primitive pair
    const double *__restrict__ P;    // density of J, scaled
    const double *__restrict__ PK;   // exchange density of B (RHF: P)
    const double *__restrict__ PK2;  // exchange density of B2 (k_generic<2>, one-pass UHF)
    const double *__restrict__ Q;
    double *A;
    double *B;
    double *B2;
    double *out;  // ERI mode: [nitems][NA NB NC ND]
    double tol;
    int nitems;  // waves (JK) or quartets (ERI)
    int la, lb, lc, ld;
    int eri_mode;
};

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

// Sums of nterm terms for nout outputs over the 64 lanes of the workgroup:
// lpo lanes (a power of two) per output each add a strided share, then a
// butterfly inside the aligned lane group; the group's first lane emits.
// Every lane calls it (the shuffles need the whole wave).
template <typename Term, typename Emit>
__device__ __forceinline__ void seg_reduce(int lane, int nout, int nterm, Term term, Emit emit) {
    int sh = 0;  // lpo = 2^sh lanes per output
    while (sh < 6 && (2 << sh) * nout <= 64) ++sh;
    const int lpo = 1 << sh, opp = 64 >> sh;
    const int sl = lane & (lpo - 1);
    for (int o0 = 0; o0 < nout; o0 += opp) {
        const int o = o0 + (lane >> sh);
        double s = 0.0;
        if (o < nout)
            for (int t = sl; t < nterm; t += lpo) s += term(o, t);
        for (int off = lpo >> 1; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
        if (o < nout && sl == 0) emit(o, s);
    }
}

// out[o] = sum_t acc[a_of(o) + b_of(t)] S[t] for o in [0, NOUT), one lane per
// output and the NTERM terms unrolled: b_of(t) of a compile-time t folds to an
// immediate LDS offset, so a term costs two LDS reads and one FMA (the class
// of the kernel fixed at compile time).  emit(o, value) takes the sum.  NL
// lanes share the outputs.
// A scheduling barrier after every PK_GEN_CT_GROUP terms keeps the compiler
// from hoisting all NTERM LDS reads ahead of their FMAs (measured: the two-wave
// (32|31) kernel 256 VGPRs + 94 AGPRs -> 237 VGPRs, 5.2 -> 2.7 ms; (33|21)
// 3.2 -> 1.8 ms on n-C10H22 / cc-pVTZ).
#ifndef PK_GEN_CT_GROUP
#define PK_GEN_CT_GROUP 16
#endif
template <int NOUT, int NTERM, int NL, typename AO, typename BO, typename Emit>
__device__ __forceinline__ void contract_ct(int lane, const double *acc, const double *S, AO a_of, BO b_of, Emit emit) {
    for (int o = lane; o < NOUT; o += NL) {
        const double *base = acc + a_of(o);
        double s[4] = {0.0, 0.0, 0.0, 0.0};  // four independent FMA chains
        sfor<0, NTERM>([&](auto t_) {
            constexpr int t = PK_CI(t_);
            s[t & 3] = fma(base[b_of(t_)], S[t], s[t & 3]);
#if PK_GEN_CT_GROUP > 0
            if constexpr ((t + 1) % PK_GEN_CT_GROUP == 0) __builtin_amdgcn_sched_barrier(0);
#endif
        });
        emit(o, (s[0] + s[1]) + (s[2] + s[3]));
    }
}

// Cartesian exponents (x, y, z) of component k of a shell of angular momentum
// l <= 3, in the order lx descending, then ly descending
static __device__ const unsigned char kCartTab[4][10][3] = {
    {{0, 0, 0}},
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},
    {{2, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 2, 0}, {0, 1, 1}, {0, 0, 2}},
    {{3, 0, 0}, {2, 1, 0}, {2, 0, 1}, {1, 2, 0}, {1, 1, 1}, {1, 0, 2}, {0, 3, 0}, {0, 2, 1}, {0, 1, 2}, {0, 0, 3}}};

// Index tables of one class, built once per workgroup so that the hot loops
// decode their indices by lookups instead of integer divisions.  NPR bounds
// the component pairs of every shell pair (ab, cd, ac, ad, bc, bd), NGT the
// transferred 1-D factors per (root, axis): the largest class's 100 and 256,
// or the sizes of the class a kernel is compiled for (less static LDS, more
// workgroups per CU).
template <int NPR, int NGT>
struct GenTabs {
    unsigned char ab_a[NPR], ab_b[NPR], cd_c[NPR], cd_d[NPR];  // ab -> (a, b), cd -> (c, d)
    unsigned char ac_a[NPR], ac_c[NPR], bd_b[NPR], bd_d[NPR];  // pairs of the exchange sums
    unsigned char bc_b[NPR], bc_c[NPR], ad_a[NPR], ad_d[NPR];
    unsigned short ab_g[NPR][3];                               // G offset of (a, b) per axis
    unsigned char cd_g[NPR][3];                                // G offset of (c, d) per axis
    unsigned char g_a[NGT], g_b[NGT], g_c[NGT], g_d[NGT];      // G index -> exponents
};

// Component pairs bound of a class: the largest of NA NB, NC ND, NA NC,
// NA ND, NB NC, NB ND.
constexpr int gen_pair_bound(int la, int lb, int lc, int ld) {
    const int na = ncart(la), nb = ncart(lb), nc = ncart(lc), nd = ncart(ld);
    int m = na * nb;
    m = m > nc * nd ? m : nc * nd;
    m = m > na * nc ? m : na * nc;
    m = m > na * nd ? m : na * nd;
    m = m > nb * nc ? m : nb * nc;
    return m > nb * nd ? m : nb * nd;
}

// Register accumulation (class fixed at compile time, at most GEN_RACC_MAX
// components per lane of the NL lanes of a quartet, GEN_RACC1_MAX with one
// wave per quartet).  The quartet's integrals then reach LDS only for the
// digestion, over the (by then dead) 2-D and transferred factors when those
// are at least as large.
#ifndef PK_GEN_RACC_MAX
#define PK_GEN_RACC_MAX 15
#endif
#ifndef PK_GEN_RACC1_MAX
#define PK_GEN_RACC1_MAX 17
#endif
constexpr int GEN_RACC_MAX = PK_GEN_RACC_MAX, GEN_RACC1_MAX = PK_GEN_RACC1_MAX;
constexpr int gen_regs_per_lane(int la, int lb, int lc, int ld, int nl) {
    return (ncart(la) * ncart(lb) * ncart(lc) * ncart(ld) + nl - 1) / nl;
}
constexpr bool gen_racc(int la, int lb, int lc, int ld, int nl) {
    return gen_regs_per_lane(la, lb, lc, ld, nl) <= (nl == 64 ? GEN_RACC1_MAX : GEN_RACC_MAX);
}
// Waves per SIMD the compiler must leave room for: two (256 registers) for the
// one-wave kernels that accumulate more than GEN_RACC_MAX components per lane
// in registers (without the bound they take 256 VGPRs and a few AGPRs, one
// wave per SIMD; held to three they spill 88 VGPRs to scratch); one elsewhere.
#ifndef PK_GEN_RACC1_WAVES
#define PK_GEN_RACC1_WAVES 2
#endif
constexpr int gen_min_waves(int la, int lb, int lc, int ld, int nl) {
    return la >= 0 && nl == 64 && gen_racc(la, lb, lc, ld, nl) && gen_regs_per_lane(la, lb, lc, ld, nl) > GEN_RACC_MAX
               ? PK_GEN_RACC1_WAVES
               : 1;
}

// NK = 2: one-pass UHF digestion (J of P into A, exchange of PK into B and of
// PK2 into B2).  TLA..TLD >= 0 fix the class at compile time (every size,
// loop bound and index division below is then a constant); -1 takes it from
// the arguments (one kernel for every class: ERI output, Schwarz table).
// Waves per quartet of the compile-time-class kernel: two for the classes of at
// least 1800 integrals per quartet (measured on n-C10H22 / cc-pVTZ: (dd|dp)
// 5.3 -> 3.5 ms, (fd|fd) 3.2 -> 1.9 ms, (ff|dd) 2.2 -> 1.3 ms), one below
// ((fd|dp) 14.0 -> 15.0 ms, (fp|fp) 3.8 -> 4.6 ms with two).
#ifndef PK_GEN_NW2_MIN
#define PK_GEN_NW2_MIN 1800
#endif
#ifndef PK_GEN_NW4_MIN
#define PK_GEN_NW4_MIN 1000000
#endif
constexpr int gen_waves(int la, int lb, int lc, int ld) {
    const int n = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld);
    return n >= PK_GEN_NW4_MIN ? 4 : (n >= PK_GEN_NW2_MIN ? 2 : 1);
}

// NW waves per workgroup (NL = 64 NW lanes) share each quartet; NW > 1 needs
// the class at compile time.
template <int NK, int TLA = -1, int TLB = -1, int TLC = -1, int TLD = -1, int NW = 1>
__global__ __launch_bounds__(64 * NW, gen_min_waves(TLA, TLB, TLC, TLD, 64 * NW)) void k_generic(GenArgs a) {
    constexpr int NL = 64 * NW;
    static_assert(NW == 1 || (TLA >= 0 && TLB >= 0 && TLC >= 0 && TLD >= 0),
                  "several waves per quartet need the class at compile time (shuffle-free digestion)");
    constexpr bool TMPL = TLA >= 0 && TLB >= 0 && TLC >= 0 && TLD >= 0;
    constexpr int NPR = TMPL ? gen_pair_bound(TLA, TLB, TLC, TLD) : 100;
    constexpr int NGT = TMPL ? (TLA + 1) * (TLB + 1) * (TLC + 1) * (TLD + 1) : 256;
    extern __shared__ double smem[];
    __shared__ double su[GEN_NR], sw[GEN_NR];
    __shared__ double hab[3][10], hcd[3][10];  // C(b,k) AB^(b-k) per axis at [b(b+1)/2 + k]; ket likewise
    // density blocks of the quartet: ab, cd (J density), ac, ad, bc, bd (exchange
    // density of B), then ac, ad, bc, bd of the exchange density of B2
    __shared__ double sP[NK == 2 ? 10 : 6][NPR];
    __shared__ double sQ[2][NPR];              // Schwarz blocks ab, cd
    __shared__ GenTabs<NPR, NGT> tb;
    const int lane = threadIdx.x;
    const int item = blockIdx.x;
    if (item >= a.nitems) return;
    const int la = TLA >= 0 ? TLA : a.la, lb = TLB >= 0 ? TLB : a.lb;
    const int lc = TLC >= 0 ? TLC : a.lc, ld = TLD >= 0 ? TLD : a.ld;
    const int NA = (la + 1) * (la + 2) / 2, NB = (lb + 1) * (lb + 2) / 2;
    const int NC = (lc + 1) * (lc + 2) / 2, ND = (ld + 1) * (ld + 2) / 2;
    const int NAB = NA * NB, NCD = NC * ND, NABCD = NAB * NCD;
    const int E = la + lb, F = lc + ld, n = (E + F) / 2 + 1;
    const int F1 = F + 1, NI = (E + 1) * F1;
    const int D2 = lb + 1, D3 = lc + 1, D4 = ld + 1, D34 = D3 * D4;
    const int NG = (la + 1) * D2 * D34;  // transferred 1-D factors per (root, axis)
    // Class fixed at compile time and at most GEN_RACC_MAX (one wave:
    // GEN_RACC1_MAX) components per lane (measured on n-C10H22 / cc-pVTZ:
    // (31|21) 14.2 -> 11.1 ms, (31|31) 3.8 -> 3.0 ms; with 17 components per
    // lane, held to two waves per SIMD and the integrals over the dead factors
    // in LDS, (32|21) 11.2 -> 8.0 ms and (31|22) 5.3 -> 3.9 ms, from LDS-bound
    // occupancy of 6 to 8 workgroups per CU):
    // the lane's components i = lane + 64 k are accumulated in registers, with
    // their three G offsets (each < 256) packed in one word per component,
    // computed once; the lane's transfer entries t = lane + 64 j likewise carry
    // (root-axis, G index, exponents) packed in one word each.
    constexpr int T_NABCD = TMPL ? ncart(TLA) * ncart(TLB) * ncart(TLC) * ncart(TLD) : 1;
    constexpr int T_N = TMPL ? (TLA + TLB + TLC + TLD) / 2 + 1 : 1;
    constexpr int T_NG = TMPL ? (TLA + 1) * (TLB + 1) * (TLC + 1) * (TLD + 1) : 1;
    constexpr int KREG = (T_NABCD + NL - 1) / NL, KTR = (3 * T_N * T_NG + NL - 1) / NL;
    constexpr bool RACC = TMPL && gen_racc(TLA, TLB, TLC, TLD, NL);
    // register accumulation: the integrals take the place of the 2-D and
    // transferred factors for the digestion when they fit there
    constexpr bool ALIAS = RACC && T_NABCD <= 3 * T_N * ((TLA + TLB + 1) * (TLC + TLD + 1) + T_NG);
    double *I2 = smem;                   // [n * 3][E + 1][F + 1]
    double *G = I2 + n * 3 * NI;         // [n * 3][la + 1][lb + 1][lc + 1][ld + 1]
    double *acc = ALIAS ? smem : G + n * 3 * NG;  // [NABCD]

    for (int i = lane; i < NAB; i += NL) {
        const int ia = i / NB, ib = i % NB;
        tb.ab_a[i] = ia;
        tb.ab_b[i] = ib;
        for (int q = 0; q < 3; ++q) tb.ab_g[i][q] = (kCartTab[la][ia][q] * D2 + kCartTab[lb][ib][q]) * D34;
    }
    for (int i = lane; i < NCD; i += NL) {
        const int ic = i / ND, id = i % ND;
        tb.cd_c[i] = ic;
        tb.cd_d[i] = id;
        for (int q = 0; q < 3; ++q) tb.cd_g[i][q] = kCartTab[lc][ic][q] * D4 + kCartTab[ld][id][q];
    }
    for (int i = lane; i < NA * NC; i += NL) tb.ac_a[i] = i / NC, tb.ac_c[i] = i % NC;
    for (int i = lane; i < NB * ND; i += NL) tb.bd_b[i] = i / ND, tb.bd_d[i] = i % ND;
    for (int i = lane; i < NB * NC; i += NL) tb.bc_b[i] = i / NC, tb.bc_c[i] = i % NC;
    for (int i = lane; i < NA * ND; i += NL) tb.ad_a[i] = i / ND, tb.ad_d[i] = i % ND;
    for (int g = lane; g < NG; g += NL) {
        int t = g;
        tb.g_d[g] = t % D4;
        t /= D4;
        tb.g_c[g] = t % D3;
        t /= D3;
        tb.g_b[g] = t % D2;
        tb.g_a[g] = t / D2;
    }
    // this lane's first (ab, cd) in the flattened component loops, and the
    // step of 64 components in (ab, cd); likewise (root-axis, g) of the
    // transfer loop
    const int ab0 = lane / NCD, cd0 = lane % NCD, sab = NL / NCD, scd = NL % NCD;
    const int ra0 = lane / NG, g0 = lane % NG, sra = NL / NG, sg = NL % NG;
    __syncthreads();

    unsigned goff[RACC ? KREG : 1], gtr[RACC ? KTR : 1];
    double racc[RACC ? KREG : 1];
    if constexpr (RACC) {
        int ab = ab0, cd = cd0;
        sfor<0, KREG>([&](auto k_) {
            constexpr int k = PK_CI(k_);
            unsigned v = 0;
            if (lane + NL * k < T_NABCD)
                v = (unsigned)(tb.ab_g[ab][0] + tb.cd_g[cd][0]) | ((unsigned)(tb.ab_g[ab][1] + tb.cd_g[cd][1]) << 8) |
                    ((unsigned)(tb.ab_g[ab][2] + tb.cd_g[cd][2]) << 16);
            goff[k] = v;
            cd += scd;
            ab += sab;
            if (cd >= NCD) cd -= NCD, ++ab;
        });
        int ra = ra0, g = g0;
        sfor<0, KTR>([&](auto j_) {
            constexpr int j = PK_CI(j_);
            unsigned v = 0xffffffffu;
            if (lane + NL * j < 3 * T_N * T_NG)
                v = (unsigned)g | ((unsigned)ra << 8) | ((unsigned)tb.g_a[g] << 16) | ((unsigned)tb.g_b[g] << 18) |
                    ((unsigned)tb.g_c[g] << 20) | ((unsigned)tb.g_d[g] << 22);
            gtr[j] = v;
            g += sg;
            ra += sra;
            if (g >= NG) g -= NG, ++ra;
        });
    }

    int bra, kbeg, kcnt, nb = 0;
    const double *P = nullptr, *PKm = nullptr, *PK2m = nullptr, *Qm = nullptr;
    double *Am = nullptr, *Bm = nullptr, *B2m = nullptr;
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
        PKm = a.PK + moff;
        Qm = a.Q + moff;
        Am = a.A + moff;
        Bm = a.B + moff;
        if (NK == 2) {
            PK2m = a.PK2 + moff;
            B2m = a.B2 + moff;
        }
    }
    const PairDev pb = a.pairs[bra];
    if (lane < 30) {
        // bra transfer coefficients C(b,k) AB^(b-k), b <= 3
        const int ax = lane / 10, t = lane % 10;
        const int b = t >= 6 ? 3 : (t >= 3 ? 2 : (t >= 1 ? 1 : 0)), k = t - b * (b + 1) / 2;
        const double x = sel3(ax, pb.ab[0], pb.ab[1], pb.ab[2]);
        double p = binom(b, k);
        for (int e = 0; e < b - k; ++e) p *= x;
        hab[ax][t] = p;
    }

    for (int kk = 0; kk < kcnt; ++kk) {
        const int ket = a.eri_mode ? a.qket[item] : a.kets[kbeg + kk];
        const PairDev pk = a.pairs[ket];
        if constexpr (RACC)
            sfor<0, KREG>([&](auto k_) { racc[PK_CI(k_)] = 0.0; });
        else
            for (int i = lane; i < NABCD; i += NL) acc[i] = 0.0;
        if (lane < 30) {
            const int ax = lane / 10, t = lane % 10;
            const int b = t >= 6 ? 3 : (t >= 3 ? 2 : (t >= 1 ? 1 : 0)), k = t - b * (b + 1) / 2;
            const double x = sel3(ax, pk.ab[0], pk.ab[1], pk.ab[2]);
            double p = binom(b, k);
            for (int e = 0; e < b - k; ++e) p *= x;
            hcd[ax][t] = p;
        }
        __syncthreads();
        for (int kq = 0; kq < pk.nprim; ++kq) {
            const PrimDev q = a.prims[pk.prim_off + kq];
            const double hq = a.hz[pk.prim_off + kq];
            for (int kp = 0; kp < pb.nprim; ++kp) {
                const PrimDev p = a.prims[pb.prim_off + kp];
                const double hp = a.hz[pb.prim_off + kp];
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
                // 2-D integrals I[e][f] of root r, axis ax (lane 3r + ax), one
                // column f at a time with the previous two columns in registers
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
                    double c2[GEN_NR], c1[GEN_NR], c0[GEN_NR];
                    c1[0] = ax == 0 ? pref * sw[r] : 1.0;
#pragma unroll
                    for (int e = 1; e < GEN_NR; ++e) {
                        double t = C00 * c1[e - 1];
                        if (e >= 2) t = fma((double)(e - 1) * B10, c1[e - 2], t);
                        c1[e] = e <= E ? t : 0.0;
                    }
#pragma unroll
                    for (int e = 0; e < GEN_NR; ++e) {
                        if (e <= E) I[e * F1] = c1[e];
                        c2[e] = 0.0;
                    }
                    for (int f = 1; f <= F; ++f) {
                        const double bf = (double)(f - 1) * B01;
#pragma unroll
                        for (int e = 0; e < GEN_NR; ++e) {
                            double t = fma(D00, c1[e], bf * c2[e]);
                            if (e >= 1) t = fma((double)e * B00, c1[e - 1], t);
                            c0[e] = e <= E ? t : 0.0;
                        }
#pragma unroll
                        for (int e = 0; e < GEN_NR; ++e) {
                            if (e <= E) I[e * F1 + f] = c0[e];
                            c2[e] = c1[e];
                            c1[e] = c0[e];
                        }
                    }
                }
                __syncthreads();
                // transferred factors G[a][b][c][d] = sum_k sum_l C(b,k) AB^(b-k)
                // C(d,l) CD^(d-l) I[a+k][c+l] for every (root, axis) and G index,
                // spread over the lanes
                if constexpr (RACC) {
                    sfor<0, KTR>([&](auto j_) {
                        const unsigned v = gtr[PK_CI(j_)];
                        if (v != 0xffffffffu) {
                            const int g = v & 255, ra = (v >> 8) & 255;
                            const int ia = (v >> 16) & 3, ib = (v >> 18) & 3, ic = (v >> 20) & 3, id = (v >> 22) & 3;
                            const int ax = ra % 3;
                            const double *I = I2 + ra * NI + ia * F1 + ic;
                            const double *hb = hab[ax] + ib * (ib + 1) / 2, *hd = hcd[ax] + id * (id + 1) / 2;
                            double s = 0.0;
                            sfor<0, TLB + 1>([&](auto k_) {
                                constexpr int k = PK_CI(k_);
                                if (k <= ib) {
                                    double scd = 0.0;
                                    sfor<0, TLD + 1>([&](auto l_) {
                                        constexpr int l = PK_CI(l_);
                                        if (l <= id) scd = fma(hd[l], I[k * F1 + l], scd);
                                    });
                                    s = fma(hb[k], scd, s);
                                }
                            });
                            G[ra * NG + g] = s;
                        }
                    });
                } else {
                    int ra = ra0, g = g0;
                    for (int t = lane; t < 3 * n * NG; t += NL) {
                        const int ax = ra % 3;
                        const int ia = tb.g_a[g], ib = tb.g_b[g], ic = tb.g_c[g], id = tb.g_d[g];
                        const double *I = I2 + ra * NI + ia * F1 + ic;
                        const double *hb = hab[ax] + ib * (ib + 1) / 2, *hd = hcd[ax] + id * (id + 1) / 2;
                        double s = 0.0;
                        for (int k = 0; k <= ib; ++k) {
                            double scd = 0.0;
                            for (int l = 0; l <= id; ++l) scd = fma(hd[l], I[k * F1 + l], scd);
                            s = fma(hb[k], scd, s);
                        }
                        G[ra * NG + g] = s;
                        g += sg;
                        ra += sra;
                        if (g >= NG) g -= NG, ++ra;
                    }
                }
                __syncthreads();
                if constexpr (RACC) {
                    sfor<0, KREG>([&](auto k_) {
                        constexpr int k = PK_CI(k_);
                        const unsigned v = goff[k];
                        const int o0 = v & 255, o1 = (v >> 8) & 255, o2 = v >> 16;
                        double s = 0.0;
                        sfor<0, T_N>([&](auto r_) {
                            constexpr int r = PK_CI(r_);
                            s = fma(G[(r * 3) * T_NG + o0] * G[(r * 3 + 1) * T_NG + o1], G[(r * 3 + 2) * T_NG + o2], s);
                        });
                        racc[k] += s;
                    });
                } else {
                    int ab = ab0, cd = cd0;
                    for (int i = lane; i < NABCD; i += NL) {
                        const int o0 = tb.ab_g[ab][0] + tb.cd_g[cd][0], o1 = tb.ab_g[ab][1] + tb.cd_g[cd][1],
                                  o2 = tb.ab_g[ab][2] + tb.cd_g[cd][2];
                        double s = 0.0;
                        for (int r = 0; r < n; ++r)
                            s = fma(G[(r * 3) * NG + o0] * G[(r * 3 + 1) * NG + o1], G[(r * 3 + 2) * NG + o2], s);
                        acc[i] += s;
                        cd += scd;
                        ab += sab;
                        if (cd >= NCD) cd -= NCD, ++ab;
                    }
                }
                __syncthreads();
            }
        }
        if constexpr (RACC) {
            // the quartet's integrals into LDS for the digestion
            sfor<0, KREG>([&](auto k_) {
                constexpr int k = PK_CI(k_);
                if (lane + NL * k < T_NABCD) acc[lane + NL * k] = racc[k];
            });
            __syncthreads();
        }
        if (a.eri_mode) {
            double *o = a.out + (size_t)item * NABCD;
            for (int i = lane; i < NABCD; i += NL) o[i] = acc[i];
            __syncthreads();
            continue;
        }
        // ---- J/K digestion of the quartet (bra pb, ket pk)
        double deg = 1.0;
        if (pb.diag) deg *= 0.5;
        if (pk.diag) deg *= 0.5;
        if (bra == ket) deg *= 0.5;
        const bool all_pass = pb.qmin * pk.qmin >= a.tol;
        // density (and Schwarz) blocks of the quartet into LDS: independent
        // loads issued together, read many times by the reductions below
        auto Pv = [&](int r, int c) { return P[(size_t)r * nb + c]; };
        for (int i = lane; i < NAB; i += NL) {
            const int af = pb.f0 + tb.ab_a[i], bf = pb.f1 + tb.ab_b[i];
            sP[0][i] = Pv(af, bf);
            if (!all_pass) sQ[0][i] = Qm[(size_t)af * nb + bf];
        }
        for (int i = lane; i < NCD; i += NL) {
            const int cf = pk.f0 + tb.cd_c[i], df = pk.f1 + tb.cd_d[i];
            sP[1][i] = Pv(cf, df);
            if (!all_pass) sQ[1][i] = Qm[(size_t)cf * nb + df];
        }
#pragma unroll
        for (int sp = 0; sp < NK; ++sp) {
            const double *X = sp == 0 ? PKm : PK2m;
            double(*S)[NPR] = sP + 2 + 4 * sp;
            for (int i = lane; i < NA * NC; i += NL) S[0][i] = X[(size_t)(pb.f0 + tb.ac_a[i]) * nb + pk.f0 + tb.ac_c[i]];
            for (int i = lane; i < NA * ND; i += NL) S[1][i] = X[(size_t)(pb.f0 + tb.ad_a[i]) * nb + pk.f1 + tb.ad_d[i]];
            for (int i = lane; i < NB * NC; i += NL) S[2][i] = X[(size_t)(pb.f1 + tb.bc_b[i]) * nb + pk.f0 + tb.bc_c[i]];
            for (int i = lane; i < NB * ND; i += NL) S[3][i] = X[(size_t)(pb.f1 + tb.bd_b[i]) * nb + pk.f1 + tb.bd_d[i]];
        }
        __syncthreads();
        // apply degeneracy and the function-level Schwarz test in place
        {
            int ab = ab0, cd = cd0;
            for (int i = lane; i < NABCD; i += NL) {
                double m = deg;
                if (!all_pass && sQ[0][ab] * sQ[1][cd] < a.tol) m = 0.0;
                acc[i] *= m;
                cd += scd;
                ab += sab;
                if (cd >= NCD) cd -= NCD, ++ab;
            }
        }
        __syncthreads();
        // compile-time contractions for the classes above 1080 integrals (and
        // with several waves, which the shuffle reductions below cannot span);
        // measured: (fd|dp) 13.2 -> 14.0 ms with them, (dd|dp) 5.8 -> 5.3 ms
        if constexpr (TMPL && (NW > 1 || T_NABCD > 1080)) {
            constexpr int cNA = ncart(TLA), cNB = ncart(TLB), cNC = ncart(TLC), cND = ncart(TLD);
            constexpr int cNCD = cNC * cND, cNBCD = cNB * cNCD;
            // A_ab += 2 sum_cd v P_cd ; A_cd += 2 sum_ab v P_ab
            contract_ct<cNA * cNB, cNCD, NL>(
                lane, acc, sP[1], [&](int o) { return o * cNCD; }, [](auto t_) { return (int)PK_CI(t_); },
                [&](int o, double v) { unsafeAtomicAdd(&Am[(size_t)(pb.f0 + o / cNB) * nb + pb.f1 + o % cNB], 2.0 * v); });
            contract_ct<cNCD, cNA * cNB, NL>(
                lane, acc, sP[0], [&](int o) { return o; }, [](auto t_) { return (int)PK_CI(t_) * cNCD; },
                [&](int o, double v) { unsafeAtomicAdd(&Am[(size_t)(pk.f0 + o / cND) * nb + pk.f1 + o % cND], 2.0 * v); });
#pragma unroll
            for (int sp = 0; sp < NK; ++sp) {
                double *Bs = sp == 0 ? Bm : B2m;
                const double(*S)[NPR] = sP + 2 + 4 * sp;  // ac, ad, bc, bd
                // B_ac += sum_bd v P_bd
                contract_ct<cNA * cNC, cNB * cND, NL>(
                    lane, acc, S[3], [&](int o) { return (o / cNC) * cNBCD + (o % cNC) * cND; },
                    [](auto t_) {
                        constexpr int t = PK_CI(t_);
                        return (t / cND) * cNCD + t % cND;
                    },
                    [&](int o, double v) { unsafeAtomicAdd(&Bs[(size_t)(pb.f0 + o / cNC) * nb + pk.f0 + o % cNC], v); });
                // B_bc += sum_ad v P_ad
                contract_ct<cNB * cNC, cNA * cND, NL>(
                    lane, acc, S[1], [&](int o) { return (o / cNC) * cNCD + (o % cNC) * cND; },
                    [](auto t_) {
                        constexpr int t = PK_CI(t_);
                        return (t / cND) * cNBCD + t % cND;
                    },
                    [&](int o, double v) { unsafeAtomicAdd(&Bs[(size_t)(pb.f1 + o / cNC) * nb + pk.f0 + o % cNC], v); });
                // B_ad += sum_bc v P_bc
                contract_ct<cNA * cND, cNB * cNC, NL>(
                    lane, acc, S[2], [&](int o) { return (o / cND) * cNBCD + o % cND; },
                    [](auto t_) {
                        constexpr int t = PK_CI(t_);
                        return (t / cNC) * cNCD + (t % cNC) * cND;
                    },
                    [&](int o, double v) { unsafeAtomicAdd(&Bs[(size_t)(pb.f0 + o / cND) * nb + pk.f1 + o % cND], v); });
                // B_bd += sum_ac v P_ac
                contract_ct<cNB * cND, cNA * cNC, NL>(
                    lane, acc, S[0], [&](int o) { return (o / cND) * cNCD + o % cND; },
                    [](auto t_) {
                        constexpr int t = PK_CI(t_);
                        return (t / cNC) * cNBCD + (t % cNC) * cND;
                    },
                    [&](int o, double v) { unsafeAtomicAdd(&Bs[(size_t)(pb.f1 + o / cND) * nb + pk.f1 + o % cND], v); });
            }
        } else {
            auto at = [&](int ia, int ib, int ic, int id) { return acc[(ia * NB + ib) * NCD + ic * ND + id]; };
            // A_ab += 2 sum_cd v P_cd ; A_cd += 2 sum_ab v P_ab
            seg_reduce(lane, NAB, NCD, [&](int ab, int cd) { return acc[ab * NCD + cd] * sP[1][cd]; },
                       [&](int ab, double v) {
                           unsafeAtomicAdd(&Am[(size_t)(pb.f0 + tb.ab_a[ab]) * nb + pb.f1 + tb.ab_b[ab]], 2.0 * v);
                       });
            seg_reduce(lane, NCD, NAB, [&](int cd, int ab) { return acc[ab * NCD + cd] * sP[0][ab]; },
                       [&](int cd, double v) {
                           unsafeAtomicAdd(&Am[(size_t)(pk.f0 + tb.cd_c[cd]) * nb + pk.f1 + tb.cd_d[cd]], 2.0 * v);
                       });
            // B_ac += sum_bd v P_bd ; B_bc += sum_ad v P_ad ; B_ad += sum_bc v P_bc ; B_bd += sum_ac v P_ac
            // (and the same into B2 with the second exchange density)
    #pragma unroll
            for (int sp = 0; sp < NK; ++sp) {
                double *Bs = sp == 0 ? Bm : B2m;
                const double(*S)[NPR] = sP + 2 + 4 * sp;  // ac, ad, bc, bd
                seg_reduce(lane, NA * NC, NB * ND, [&](int o, int t) {
                    return at(tb.ac_a[o], tb.bd_b[t], tb.ac_c[o], tb.bd_d[t]) * S[3][t];
                }, [&](int o, double v) { unsafeAtomicAdd(&Bs[(size_t)(pb.f0 + tb.ac_a[o]) * nb + pk.f0 + tb.ac_c[o]], v); });
                seg_reduce(lane, NB * NC, NA * ND, [&](int o, int t) {
                    return at(tb.ad_a[t], tb.bc_b[o], tb.bc_c[o], tb.ad_d[t]) * S[1][t];
                }, [&](int o, double v) { unsafeAtomicAdd(&Bs[(size_t)(pb.f1 + tb.bc_b[o]) * nb + pk.f0 + tb.bc_c[o]], v); });
                seg_reduce(lane, NA * ND, NB * NC, [&](int o, int t) {
                    return at(tb.ad_a[o], tb.bc_b[t], tb.bc_c[t], tb.ad_d[o]) * S[2][t];
                }, [&](int o, double v) { unsafeAtomicAdd(&Bs[(size_t)(pb.f0 + tb.ad_a[o]) * nb + pk.f1 + tb.ad_d[o]], v); });
                seg_reduce(lane, NB * ND, NA * NC, [&](int o, int t) {
                    return at(tb.ac_a[t], tb.bd_b[o], tb.ac_c[t], tb.bd_d[o]) * S[0][t];
                }, [&](int o, double v) { unsafeAtomicAdd(&Bs[(size_t)(pb.f1 + tb.bd_b[o]) * nb + pk.f1 + tb.bd_d[o]], v); });
            }
        }
        __syncthreads();
    }
}

// dynamic LDS of the generic kernel for a class: 2-D integrals and transferred
// factors for n roots x 3 axes, and the quartet's accumulators (over the
// factors when the kernel accumulates in registers, nl_regacc > 0 lanes per
// quartet, and they fit there)
inline size_t generic_lds_bytes(int la, int lb, int lc, int ld, int nl_regacc = 0) {
    const size_t nabcd = (size_t)ncart(la) * ncart(lb) * ncart(lc) * ncart(ld);
    const size_t n = (la + lb + lc + ld) / 2 + 1;
    const size_t ni = (size_t)(la + lb + 1) * (lc + ld + 1), ng = (size_t)(la + 1) * (lb + 1) * (lc + 1) * (ld + 1);
    const size_t fac = n * 3 * (ni + ng);
    const bool alias = nl_regacc > 0 && gen_racc(la, lb, lc, ld, nl_regacc) && nabcd <= fac;
    return (alias ? fac : fac + nabcd) * sizeof(double);
}

}  // namespace pk
