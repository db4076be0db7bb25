// eri_kernels.h — the per-class shell-quartet kernels (gfx950, FP64).
//
// One lane owns one contracted shell quartet (ab|cd) of a fixed angular class
// (LA >= LB, LC >= LD, bra class >= ket class).  The primitive loops run ket
// primitives outer (per lane) and bra primitives inner (the bra pair is the
// same for the whole wave, so its data comes through scalar loads).  Two
// algorithms share the digestion:
//   ALG_OS  : Obara-Saika / Head-Gordon-Pople VRR to the contracted [e0|f0],
//             then the A->B and C->D transfers once per quartet;
//   ALG_RYS : Rys quadrature, C->D transfer on the 2-D integrals, contracted
//             [e0|cd) for a chunk [C0, C1) of ket components, then A->B.
// J/K digestion: for the unique quartet with degeneracy factor d and value v,
//   A_ab += 2 d v P_cd, A_cd += 2 d v P_ab           (J = A + A^T)
//   B_ac += d v P_bd, B_bc += d v P_ad, B_ad += d v P_bc, B_bd += d v P_ac   (K = B + B^T)
// with the reference's screening of every AO quartet whose Q_ab Q_cd is below
// tol_eri (src/integrals/os.cpp:1475); A_ab is reduced over the wave first,
// everything else is added with FP64 atomics.
#pragma once

#include "eri_prim.h"

namespace pk {

enum { ALG_OS = 0, ALG_RYS = 1 };

// Work of one class: waves (bra pair, ket range, molecule).  The matrices of
// molecule m start at mol_off[m] of the packed arrays, leading dimension
// mol_nb[m]; a single molecule is the batch of one.
struct JKArgs {
    const int *__restrict__ wave_bra;
    const int *__restrict__ wave_start;
    const int *__restrict__ wave_count;
    const int *__restrict__ wave_mol;
    const int *__restrict__ kets;
    const PairDev *__restrict__ pairs;
    const PrimDev *__restrict__ prims;
    const double *__restrict__ hz;  // 1/(2 zeta) per primitive pair (beside prims)
    const long long *__restrict__ mol_off;
    const int *__restrict__ mol_nb;
    const double *__restrict__ P;    // density of J (RHF: also of K), scaled by component norms
    const float *__restrict__ P32;   // the same density in FP32 (read by the FP32 kernels)
    const double *__restrict__ PK;   // one-pass UHF: exchange density of B (Pa), scaled
    const double *__restrict__ PK2;  // one-pass UHF: exchange density of B2 (Pb), scaled
    const double *__restrict__ Q;    // Schwarz per function pair (normalised functions)
    double *A;                       // J accumulator, one orientation
    double *B;                       // K accumulator, one orientation
    double *B2;                      // second K accumulator (one-pass UHF)
    double tol;
    int nwaves;
    int remap;  // 1: each XCD takes a contiguous range of waves (grid a multiple of 8)
    // segment mode (see k_jk): segments of waves of one bra, block starts of
    // the function types per molecule, LDS columns per exchange row
    const int *__restrict__ seg;   // [nseg + 1] first wave of each segment
    const int *__restrict__ tblk;  // [5 * nmol]: first function of type s, p, d, f, and nb
    int nseg;
    int kcols;
    int segmode;
};

// Wave handled by this wave slot.  With remap, workgroup b (dispatched to XCD
// b % 8) takes logical block (b % 8) * (grid / 8) + b / 8, a bijection on the
// grid when the grid is a multiple of 8, so consecutive waves (same bra pair,
// neighbouring kets) share one XCD's L2.
__device__ __forceinline__ int wave_index(int remap) {
    int blk = blockIdx.x;
    if (remap) {
        const int per = gridDim.x >> 3;
        blk = (blk & 7) * per + (blk >> 3);
    }
    return __builtin_amdgcn_readfirstlane(blk * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Contracted Rys intermediates [e0|cd) of the chunk [C0, C1): acc[NE][NCH].
template <typename R, int LA, int LB, int LC, int LD, int C0, int C1>
__device__ __forceinline__ void rys_contract(const PairDev &pb, const PairDev &pk, int nq, int np,
                                             const PrimDev *__restrict__ prims, const double *__restrict__ hz,
                                             R (&acc)[Cls<LA, LB, LC, LD>::NE][C1 - C0]) {
    using K = Cls<LA, LB, LC, LD>;
    sfor<0, K::NE>([&](auto i_) { sfor<0, C1 - C0>([&](auto j_) { acc[PK_CI(i_)][PK_CI(j_)] = R(0); }); });
    const R cdv[3] = {(R)pk.ab[0], (R)pk.ab[1], (R)pk.ab[2]};
    for (int kq = 0; kq < nq; ++kq) {
        const PrimDev q = prims[pk.prim_off + kq];
        const double hq = hz[pk.prim_off + kq];
        for (int kp = 0; kp < np; ++kp) {
            const PrimDev p = prims[pb.prim_off + kp];
            rys_prim<R, LA, LB, LC, LD, C0, C1>(p, q, hz[pb.prim_off + kp], hq, cdv, acc);
        }
    }
}

// Contracted (ab|cd) of all ket components by OS/HGP: V[NAB][NCD].
template <typename R, int LA, int LB, int LC, int LD>
__device__ __forceinline__ void os_values(const PairDev &pb, const PairDev &pk, int nq, int np,
                                          const PrimDev *__restrict__ prims, const double *__restrict__ hz,
                                          R (&V)[Cls<LA, LB, LC, LD>::NAB][Cls<LA, LB, LC, LD>::NCD]) {
    using K = Cls<LA, LB, LC, LD>;
    constexpr int NE = K::NE, NF = K::NF, NAB = K::NAB;
    R acc[NE][NF];
    sfor<0, NE>([&](auto i_) { sfor<0, NF>([&](auto j_) { acc[PK_CI(i_)][PK_CI(j_)] = R(0); }); });
    for (int kq = 0; kq < nq; ++kq) {
        const PrimDev q = prims[pk.prim_off + kq];
        const double hq = hz[pk.prim_off + kq];
        for (int kp = 0; kp < np; ++kp) {
            const PrimDev p = prims[pb.prim_off + kp];
            os_prim<R, LA, LB, LC, LD>(p, q, hz[pb.prim_off + kp], hq, acc);
        }
    }
    // A->B per ket column f, then C->D per bra component ab
    R Tb[NAB][NF];
    sfor<0, NF>([&](auto jf_) {
        constexpr int jf = PK_CI(jf_);
        R col[NE], out[NAB];
        sfor<0, NE>([&](auto i_) { col[PK_CI(i_)] = acc[PK_CI(i_)][jf]; });
        hrr_step<R, LA, LB, 0>(col, pb.ab, out);
        sfor<0, NAB>([&](auto i_) { Tb[PK_CI(i_)][jf] = out[PK_CI(i_)]; });
    });
    sfor<0, NAB>([&](auto i_) { hrr_step<R, LC, LD, 0>(Tb[PK_CI(i_)], pk.ab, V[PK_CI(i_)]); });
}

// Classes whose per-lane kernel runs in segment mode (LDS exchange rows, see
// k_jk).  Measured per class on n-C10H22 / cc-pVTZ (segment mode / plain):
// kets with a p or d second shell 0.26-0.83 (the Rys classes (30|22) 0.26,
// (30|21) 0.31, (21|11) 0.43); kets (ss) under a d or f bra 0.75-0.83; but
// (00|00) 1.61, kets (ds) 1.16-1.30 except (20|20) 0.86, kets (fs) 1.05-2.02,
// and the (ff| bras (20 exchange rows, few kets per bra) 1.10-2.02.
constexpr bool seg_class(int la, int lb, int lc, int ld) {
    if (la == 3 && lb == 3) return false;
    if (ld >= 1) return true;
    if (lc == 1) return la >= 1 && !(la == 1 && lb == 0);
    if (lc == 0) return la >= 2;
    return la == 2 && lb == 0 && lc == 2;
}

// Classes whose plain kernel is also compiled in segment mode, chosen at run
// time for lists of few primitive quartets per shell quartet: (ss|ss), (ps|ss),
// (ps|ps), (pp|ss).  In 6-31G* water clusters (about 11-13 primitive quartets
// per shell quartet) they do ~5 exchange atomics per quartet and are bound by
// atomic throughput; in cc-pVTZ alkanes (40-80) by the primitive loop, where
// the plain kernel is faster.
constexpr bool seg_alt_class(int la, int lb, int lc, int ld) {
    return !seg_class(la, lb, lc, ld) && la + lb + lc + ld <= 2;
}

// J/K kernel: one wave = one bra pair and up to 64 ket pairs.  The integrals
// are computed in R; the J and K accumulators are FP64.
// NK = 2: one-pass UHF digestion, J of P (= Pa + Pb) into A and the exchange of
// PK (= Pa) into B and of PK2 (= Pb) into B2 from the same integrals.
//
// Segment mode (a.segmode): the waves of a workgroup take the waves
// [seg[s], seg[s + 1]) of one bra pair, in strides of the waves per workgroup.
// Every exchange contribution then lands in a row of the bra's two shells
// (B_ac, B_ad: rows a; B_bc, B_bd: rows b) and in a column of the kets' two
// shell types (internal order: a molecule's functions grouped by type,
// tblk[5 m + l] the first function of type l), so they are added into LDS rows
// krow[NK][NA + NB][kcols] (LDS atomics), the wave sums of J_ab into NAB more
// LDS doubles, and every element goes to the global accumulator once, when
// the segment is done.  The global FP64 atomics are
// executed beyond the L2 of the eight XCDs (about 1000 cycles each, and every
// later load of the wave waits for them); this replaces ~NA ND + NB ND of them
// per quartet by a share of one flush per segment.  J_cd stays a global atomic
// per ket component.
// SEGV: 1 / 0 compiles the segment-mode / plain kernel, -1 takes seg_class.
template <typename R, int LA, int LB, int LC, int LD, int ALG, int C0, int C1, int NK = 1, int SEGV = -1>
__global__ __launch_bounds__(256) void k_jk(JKArgs a) {
    using K = Cls<LA, LB, LC, LD>;
    constexpr int NA = K::NA, NB = K::NB, ND = K::ND, NAB = K::NAB, NE = K::NE, NCH = C1 - C0;
    constexpr int NROW = NA + NB;
    extern __shared__ double krow[];
    constexpr bool SEG = SEGV >= 0 ? SEGV != 0 : seg_class(LA, LB, LC, LD);
    const bool segmode = SEG && a.segmode != 0;
    int w_first, w_end, w_step, w_lead;
    if (segmode) {
        int blk = blockIdx.x;
        if (a.remap) {
            const int per = gridDim.x >> 3;
            blk = (blk & 7) * per + (blk >> 3);
        }
        if (blk >= a.nseg) return;  // the whole workgroup
        w_lead = a.seg[blk];
        w_end = a.seg[blk + 1];
        w_step = (int)(blockDim.x >> 6);
        w_first = w_lead + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    } else {
        const int wave = wave_index(a.remap);
        if (wave >= a.nwaves) return;
        w_lead = w_first = wave;
        w_end = wave + 1;
        w_step = 1;
    }
    const int mol = a.wave_mol[w_lead];
    const long long moff = a.mol_off[mol];
    const int nb = a.mol_nb[mol];
    // per-quartet digestion sums in D: FP64, or FP32 for the FP32 kernels (the
    // quartets below the FP32 cutoff), which then read an FP32 copy of the
    // density; the J and K accumulators themselves are FP64 (atomics)
    using D = std::conditional_t<std::is_same_v<R, float>, float, double>;
    const D *__restrict__ P = std::is_same_v<D, float> ? (const D *)(const void *)(a.P32 + moff)
                                                       : (const D *)(const void *)(a.P + moff);
    static_assert(NK == 1 || std::is_same_v<D, double>, "two-density digestion is FP64");
    const D *__restrict__ PK = NK == 2 ? (const D *)(const void *)(a.PK + moff) : P;   // exchange density of B
    const D *__restrict__ P2 = NK == 2 ? (const D *)(const void *)(a.PK2 + moff) : P;  // exchange density of B2
    const double *__restrict__ Qm = a.Q + moff;
    double *Am = a.A + moff;
    double *Bm = a.B + moff;
    double *B2m = NK == 2 ? a.B2 + moff : Bm;
    const int lane = threadIdx.x & 63;

    // LDS rows of the segment: column of a c function f: f - cbase; of a d
    // function: dcol0 + f - dbase
    int cbase = 0, dbase = 0, dcol0 = 0, ncols = 0;
    if (segmode) {
        const int *tb = a.tblk + 5 * mol;
        cbase = tb[LC];
        dbase = tb[LD];
        const int nc = tb[LC + 1] - tb[LC];
        dcol0 = LC == LD ? 0 : nc;
        ncols = LC == LD ? nc : nc + tb[LD + 1] - tb[LD];
        const int nz = NK * NROW * a.kcols + NAB;
        for (int i = threadIdx.x; i < nz; i += blockDim.x) krow[i] = 0.0;
        __syncthreads();
    }
    // exchange contribution v to row r (0..NA-1: a functions, NA..: b
    // functions) and column function f of the ket's first (c) or second (d) shell
    auto kadd = [&](const PairDev &pb, int sp, int r, int f, bool dcol, double v) {
        if (segmode) {
            const int col = dcol ? dcol0 + f - dbase : f - cbase;
            __hip_atomic_fetch_add(&krow[(sp * NROW + r) * a.kcols + col], v, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        } else {
            double *Bs = sp == 0 ? Bm : B2m;
            const int rf = r < NA ? pb.f0 + r : pb.f1 + (r - NA);
            unsafeAtomicAdd(&Bs[(size_t)rf * nb + f], v);
        }
    };

    double *jrow = krow + NK * NROW * a.kcols;  // segment mode: J_ab of the segment

    for (int wave = w_first; wave < w_end; wave += w_step) {
        // the bra of every wave of a segment is the segment's; it is read per
        // wave all the same, so that no bra-only value is held across waves
        // (registers)
        const int bra = a.wave_bra[wave];
        const PairDev pb = a.pairs[bra];
        const int cnt = a.wave_count[wave];
        const bool active = lane < cnt;
        const int ket = a.kets[a.wave_start[wave] + (active ? lane : 0)];
        const PairDev pk = a.pairs[ket];
        const int nq = active ? pk.nprim : 0;

        D deg = D(1);
        if (pb.diag) deg *= D(0.5);
        if (pk.diag) deg *= D(0.5);
        if (bra == ket) deg *= D(0.5);
        if (!active) deg = D(0);
        // every function quartet passes the Schwarz test when the smallest
        // function-pair bounds do; otherwise each one is tested
        const bool all_pass = pb.qmin * pk.qmin >= a.tol;

        D jab[NAB], kac[NK][NA], kbc[NK][NB], kad[NK][NA][ND], kbd[NK][NB][ND];
        sfor<0, NAB>([&](auto i_) { jab[PK_CI(i_)] = D(0); });
        sfor<0, NK>([&](auto s_) {
            constexpr int sp = PK_CI(s_);
            sfor<0, NA>([&](auto i_) {
                kac[sp][PK_CI(i_)] = D(0);
                sfor<0, ND>([&](auto d_) { kad[sp][PK_CI(i_)][PK_CI(d_)] = D(0); });
            });
            sfor<0, NB>([&](auto i_) {
                kbc[sp][PK_CI(i_)] = D(0);
                sfor<0, ND>([&](auto d_) { kbd[sp][PK_CI(i_)][PK_CI(d_)] = D(0); });
            });
        });

        // The work lists order each bra's kets by ordering group (a bucket of
        // primitive counts), then first shell, then second shell, so lanes with
        // equal (group, first ket shell) are contiguous; (a|c) and (b|c) exchange
        // sums are added over such a segment (segmented scan over the wave) and
        // stored once, by its last lane.
        const int ckey = active ? (pk.grp << 20) + pk.f0 : -1 - lane;
        bool seg_same[6];
        sfor<0, 6>([&](auto k_) {
            constexpr int off = 1 << PK_CI(k_);
            // the shuffle runs on every lane (a source lane outside the exec mask reads as 0)
            const int up = __shfl_up(ckey, off, 64);
            seg_same[PK_CI(k_)] = lane >= off && up == ckey;
        });
        const int down = __shfl_down(ckey, 1, 64);
        const bool seg_tail = lane == 63 || down != ckey;
        auto seg_sum = [&](D v) {
            sfor<0, 6>([&](auto k_) {
                const D o = __shfl_up(v, 1 << PK_CI(k_), 64);
                if (seg_same[PK_CI(k_)]) v += o;
            });
            return v;
        };

        // digestion of one ket component (ic, id) with its bra block v[NAB]
        auto digest = [&](auto jj_, const R (&v)[NAB]) {
            constexpr int jj = PK_CI(jj_);
            constexpr int ic = (C0 + jj) / ND, id = (C0 + jj) % ND;
            const int cf = pk.f0 + ic, df = pk.f1 + id;
            const D Pcd = P[(size_t)cf * nb + df];
            const double Qcd = Qm[(size_t)cf * nb + df];
            D jcd = D(0);
            sfor<0, NA>([&](auto ia_) {
                constexpr int ia = PK_CI(ia_);
                const int af = pb.f0 + ia;
                const D *KD[2] = {PK, P2};
                D Pac[NK], Pad[NK];
                sfor<0, NK>([&](auto s_) {
                    constexpr int sp = PK_CI(s_);
                    Pac[sp] = KD[sp][(size_t)af * nb + cf];
                    Pad[sp] = KD[sp][(size_t)af * nb + df];
                });
                sfor<0, NB>([&](auto ib_) {
                    constexpr int ib = PK_CI(ib_);
                    const int bf = pb.f1 + ib;
                    const size_t abi = (size_t)af * nb + bf;
                    const D m = (all_pass || Qm[abi] * Qcd >= a.tol) ? deg : D(0);
                    const D dv = m * (D)v[ia * NB + ib];
                    jcd = fma(dv, P[abi], jcd);
                    jab[ia * NB + ib] = fma(dv, Pcd, jab[ia * NB + ib]);
                    sfor<0, NK>([&](auto s_) {
                        constexpr int sp = PK_CI(s_);
                        kac[sp][ia] = fma(dv, KD[sp][(size_t)bf * nb + df], kac[sp][ia]);
                        kad[sp][ia][id] = fma(dv, KD[sp][(size_t)bf * nb + cf], kad[sp][ia][id]);
                        kbc[sp][ib] = fma(dv, Pad[sp], kbc[sp][ib]);
                        kbd[sp][ib][id] = fma(dv, Pac[sp], kbd[sp][ib][id]);
                    });
                });
            });
            if (active) unsafeAtomicAdd(&Am[(size_t)cf * nb + df], 2.0 * (double)jcd);
            // last ket component of this c in the chunk: flush the (a|c), (b|c)
            // sums, added first over the lanes that share the ket's first shell
            if constexpr (id == ND - 1 || jj == NCH - 1) {
                sfor<0, NK>([&](auto s_) {
                    constexpr int sp = PK_CI(s_);
                    sfor<0, NA>([&](auto ia_) {
                        const D v = seg_sum(kac[sp][PK_CI(ia_)]);
                        if (active && seg_tail) kadd(pb, sp, PK_CI(ia_), cf, false, (double)v);
                    });
                    sfor<0, NB>([&](auto ib_) {
                        const D v = seg_sum(kbc[sp][PK_CI(ib_)]);
                        if (active && seg_tail) kadd(pb, sp, NA + PK_CI(ib_), cf, false, (double)v);
                    });
                    sfor<0, NA>([&](auto i_) { kac[sp][PK_CI(i_)] = D(0); });
                    sfor<0, NB>([&](auto i_) { kbc[sp][PK_CI(i_)] = D(0); });
                });
            }
        };

        if constexpr (ALG == ALG_OS) {
            static_assert(C0 == 0 && C1 == K::NCD, "the OS kernels take all ket components");
            R V[NAB][K::NCD];
            os_values<R, LA, LB, LC, LD>(pb, pk, nq, pb.nprim, a.prims, a.hz, V);
            sfor<0, NCH>([&](auto jj_) {
                R v[NAB];
                sfor<0, NAB>([&](auto i_) { v[PK_CI(i_)] = V[PK_CI(i_)][PK_CI(jj_)]; });
                digest(jj_, v);
            });
        } else {
            R acc[NE][NCH];
            rys_contract<R, LA, LB, LC, LD, C0, C1>(pb, pk, nq, pb.nprim, a.prims, a.hz, acc);
            sfor<0, NCH>([&](auto jj_) {
                R col[NE], v[NAB];
                sfor<0, NE>([&](auto i_) { col[PK_CI(i_)] = acc[PK_CI(i_)][PK_CI(jj_)]; });
                hrr_step<R, LA, LB, 0>(col, pb.ab, v);
                digest(jj_, v);
            });
        }
        if (active) {
            sfor<0, ND>([&](auto id_) {
                constexpr int id = PK_CI(id_);
                constexpr bool present = [] {
                    for (int j = 0; j < NCH; ++j)
                        if ((C0 + j) % ND == id) return true;
                    return false;
                }();
                if constexpr (present) {
                    const int df = pk.f1 + id;
                    sfor<0, NK>([&](auto s_) {
                        constexpr int sp = PK_CI(s_);
                        sfor<0, NA>([&](auto ia_) { kadd(pb, sp, PK_CI(ia_), df, true, (double)kad[sp][PK_CI(ia_)][id]); });
                        sfor<0, NB>([&](auto ib_) { kadd(pb, sp, NA + PK_CI(ib_), df, true, (double)kbd[sp][PK_CI(ib_)][id]); });
                    });
                }
            });
        }
        // J_ab: sum over the wave, one add per element
        sfor<0, NAB>([&](auto i_) {
            constexpr int i = PK_CI(i_);
            const double sj = wave_sum((double)jab[i]);
            if (lane == 0) {
                if (segmode)
                    __hip_atomic_fetch_add(&jrow[i], sj, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                else
                    unsafeAtomicAdd(&Am[(size_t)(pb.f0 + i / NB) * nb + pb.f1 + i % NB], 2.0 * sj);
            }
        });
    }
    if (segmode) {
        // every wave of the workgroup reaches this barrier (no early exit above)
        __syncthreads();
        const PairDev pb = a.pairs[a.wave_bra[w_lead]];
        const int per = NROW * ncols;
        for (int i = threadIdx.x; i < NK * per; i += blockDim.x) {
            const int sp = i / per, rc = i - sp * per;
            const int r = rc / ncols, col = rc - r * ncols;
            const double v = krow[(sp * NROW + r) * a.kcols + col];
            if (v != 0.0) {
                double *Bs = sp == 0 ? Bm : B2m;
                const int rf = r < NA ? pb.f0 + r : pb.f1 + (r - NA);
                const int f = (LC == LD || col < dcol0) ? cbase + col : dbase + (col - dcol0);
                unsafeAtomicAdd(&Bs[(size_t)rf * nb + f], v);
            }
        }
        for (int i = threadIdx.x; i < NAB; i += blockDim.x)
            unsafeAtomicAdd(&Am[(size_t)(pb.f0 + i / NB) * nb + pb.f1 + i % NB], 2.0 * jrow[i]);
    }
}

}  // namespace pk
