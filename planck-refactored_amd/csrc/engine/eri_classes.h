// eri_classes.h — the list of quartet classes compiled into the engine and the
// per-class launchers.  The kernels are instantiated in pk_kernels.hip, one
// translation unit per shard so the classes compile in parallel.
#pragma once

#include "eri_kernels.h"

#define PK_MAX_L 3  // s, p, d, f shells

// X(la, lb, lc, ld, shard) for every class with bra class >= ket class, in
// the order cb (cb + 1) / 2 + ck of the pair classes (0,0) (1,0) (1,1) (2,0)
// (2,1) (2,2) (3,0) (3,1) (3,2) (3,3); shards balance the compile work.
#define PK_FOR_EACH_CLASS(X) \
    X(0, 0, 0, 0, 8) \
    X(1, 0, 0, 0, 13) \
    X(1, 0, 1, 0, 12) \
    X(1, 1, 0, 0, 15) \
    X(1, 1, 1, 0, 13) \
    X(1, 1, 1, 1, 7) \
    X(2, 0, 0, 0, 11) \
    X(2, 0, 1, 0, 6) \
    X(2, 0, 1, 1, 11) \
    X(2, 0, 2, 0, 12) \
    X(2, 1, 0, 0, 14) \
    X(2, 1, 1, 0, 14) \
    X(2, 1, 1, 1, 12) \
    X(2, 1, 2, 0, 8) \
    X(2, 1, 2, 1, 10) \
    X(2, 2, 0, 0, 11) \
    X(2, 2, 1, 0, 9) \
    X(2, 2, 1, 1, 8) \
    X(2, 2, 2, 0, 13) \
    X(2, 2, 2, 1, 3) \
    X(2, 2, 2, 2, 0) \
    X(3, 0, 0, 0, 10) \
    X(3, 0, 1, 0, 8) \
    X(3, 0, 1, 1, 10) \
    X(3, 0, 2, 0, 15) \
    X(3, 0, 2, 1, 14) \
    X(3, 0, 2, 2, 7) \
    X(3, 0, 3, 0, 13) \
    X(3, 1, 0, 0, 9) \
    X(3, 1, 1, 0, 13) \
    X(3, 1, 1, 1, 11) \
    X(3, 1, 2, 0, 15) \
    X(3, 1, 2, 1, 9) \
    X(3, 1, 2, 2, 9) \
    X(3, 1, 3, 0, 9) \
    X(3, 1, 3, 1, 9) \
    X(3, 2, 0, 0, 7) \
    X(3, 2, 1, 0, 14) \
    X(3, 2, 1, 1, 6) \
    X(3, 2, 2, 0, 12) \
    X(3, 2, 2, 1, 9) \
    X(3, 2, 2, 2, 9) \
    X(3, 2, 3, 0, 4) \
    X(3, 2, 3, 1, 9) \
    X(3, 2, 3, 2, 9) \
    X(3, 3, 0, 0, 10) \
    X(3, 3, 1, 0, 15) \
    X(3, 3, 1, 1, 2) \
    X(3, 3, 2, 0, 5) \
    X(3, 3, 2, 1, 9) \
    X(3, 3, 2, 2, 6) \
    X(3, 3, 3, 0, 1) \
    X(3, 3, 3, 1, 9) \
    X(3, 3, 3, 2, 6) \
    X(3, 3, 3, 3, 9)
#define PK_NSHARDS 16

namespace pk {

// Contracted Rys accumulators per lane: a class whose NE x NCD intermediates
// exceed the budget runs in chunks of ket components, each chunk recomputing
// the roots and 2-D integrals.  Measured per class on n-C10H22 / cc-pVTZ,
// budget 64 against 48: 0.56-0.88x the time for the chunked d/f classes (fewer
// chunks), and (fp|dp) gains a per-lane kernel (9 chunks; 0.63x the
// cooperative kernel's time); budget 80 against 64: 0.83-0.94x for most chunked
// classes ((fp|ps) 4.9 -> 3.7 ms, (fp|pp) 5.1 -> 4.2 ms, (fp|dp) 7.0 -> 6.2 ms).
#ifndef PK_ACC_BUDGET
#define PK_ACC_BUDGET 80
#endif
constexpr int kAccBudget = PK_ACC_BUDGET;
constexpr int kMaxChunks = 12;  // larger classes run in the cooperative kernel
// Which f classes get per-lane kernels is decided at this budget: at 80,
// (fp|dd) and (fp|fp) would fit 12 and 10 chunks, but (fp|fp) is slower so
// (2.0 -> 2.7 ms) and the two triple the kernel shards' build time.
constexpr int kTmplBudget = 64;

#ifndef PK_ACC_BUDGET32
#define PK_ACC_BUDGET32 128
#endif
// FP32 kernels (the quartets below the FP32 cutoff) hold their accumulators and
// 2-D integrals in half the registers, so they take twice the ket components
// per chunk: fewer chunks, each recomputing the roots and 2-D integrals
constexpr int kAccBudget32 = PK_ACC_BUDGET32;

constexpr int chunk_size(int la, int lb, int lc, int ld, int budget = kAccBudget) {
    const int ne = nlev(la, la + lb), ncd = ncart(lc) * ncart(ld);
    int ch = budget / ne;
    if (ch < 1) ch = 1;
    if (ch >= ncd) return ncd;
    // the same number of chunks, of balanced sizes
    const int nch = (ncd + ch - 1) / ch;
    return (ncd + nch - 1) / nch;
}

// classes with a per-lane Rys kernel: every s/p/d class, and f classes of at
// most kMaxChunks ket-component chunks (the rest run in the cooperative kernel)
constexpr bool rys_tmpl_available(int la, int lb, int lc, int ld) {
    const int ncd = ncart(lc) * ncart(ld), ch = chunk_size(la, lb, lc, ld, kTmplBudget);
    return la <= 2 || (ncd + ch - 1) / ch <= kMaxChunks;
}

// classes with an Obara-Saika kernel: L <= 4 and [e0|f0] accumulators <= 54
constexpr bool os_available(int la, int lb, int lc, int ld) {
    return la + lb + lc + ld <= 4 && nlev(la, la + lb) * nlev(lc, lc + ld) <= 54;
}

// dynamic LDS of a k_jk launch: the exchange rows of segment mode
inline size_t jk_lds_bytes(const JKArgs &a, int la, int lb, int nk) {
    return a.segmode ? ((size_t)nk * (ncart(la) + ncart(lb)) * a.kcols + ncart(la) * ncart(lb)) * sizeof(double) : 0;
}

// grid x block (block 64, 128 or 256 lanes = 1, 2 or 4 waves)
template <typename R, int LA, int LB, int LC, int LD, int C0>
void launch_jk_rys(const JKArgs &a, int grid, int block, hipStream_t s, int &nlaunch) {
    constexpr int NCD = ncart(LC) * ncart(LD);
    if constexpr (rys_tmpl_available(LA, LB, LC, LD) && C0 < NCD) {
        constexpr int CH = chunk_size(LA, LB, LC, LD, std::is_same_v<R, float> ? kAccBudget32 : kAccBudget);
        constexpr int C1 = C0 + CH < NCD ? C0 + CH : NCD;
        hipLaunchKernelGGL((k_jk<R, LA, LB, LC, LD, ALG_RYS, C0, C1>), dim3(grid), dim3(block),
                           jk_lds_bytes(a, LA, LB, 1), s, a);
        ++nlaunch;
        launch_jk_rys<R, LA, LB, LC, LD, C1>(a, grid, block, s, nlaunch);
    }
}

template <typename R, int LA, int LB, int LC, int LD, int NK = 1>
void launch_jk_os(const JKArgs &a, int grid, int block, hipStream_t s, int &nlaunch) {
    if constexpr (os_available(LA, LB, LC, LD)) {
        constexpr int NCD = ncart(LC) * ncart(LD);
        hipLaunchKernelGGL((k_jk<R, LA, LB, LC, LD, ALG_OS, 0, NCD, NK>), dim3(grid), dim3(block),
                           jk_lds_bytes(a, LA, LB, NK), s, a);
        ++nlaunch;
    }
}

// the segment-mode variant of a seg_alt_class (no launch elsewhere)
template <typename R, int LA, int LB, int LC, int LD, int NK = 1>
void launch_jk_os_seg(const JKArgs &a, int grid, int block, hipStream_t s, int &nlaunch) {
    if constexpr (os_available(LA, LB, LC, LD) && seg_alt_class(LA, LB, LC, LD)) {
        constexpr int NCD = ncart(LC) * ncart(LD);
        hipLaunchKernelGGL((k_jk<R, LA, LB, LC, LD, ALG_OS, 0, NCD, NK, 1>), dim3(grid), dim3(block),
                           jk_lds_bytes(a, LA, LB, NK), s, a);
        ++nlaunch;
    }
}

}  // namespace pk
