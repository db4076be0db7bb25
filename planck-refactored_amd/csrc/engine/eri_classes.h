// eri_classes.h — the list of quartet classes compiled into the engine and the
// per-class launchers.  The kernels are instantiated in pk_kernels.hip, one
// translation unit per shard so the classes compile in parallel.
#pragma once

#include "eri_kernels.h"

#define PK_MAX_L 2  // s, p, d shells

// X(la, lb, lc, ld, shard) for every class with bra class >= ket class, in
// the order cb (cb + 1) / 2 + ck of the pair classes (0,0) (1,0) (1,1) (2,0) (2,1) (2,2)
#define PK_FOR_EACH_CLASS(X)                                                                  \
    X(0, 0, 0, 0, 0)                                                                          \
    X(1, 0, 0, 0, 0) X(1, 0, 1, 0, 1)                                                         \
    X(1, 1, 0, 0, 1) X(1, 1, 1, 0, 2) X(1, 1, 1, 1, 3)                                        \
    X(2, 0, 0, 0, 0) X(2, 0, 1, 0, 1) X(2, 0, 1, 1, 2) X(2, 0, 2, 0, 3)                       \
    X(2, 1, 0, 0, 4) X(2, 1, 1, 0, 5) X(2, 1, 1, 1, 6) X(2, 1, 2, 0, 7) X(2, 1, 2, 1, 4)      \
    X(2, 2, 0, 0, 5) X(2, 2, 1, 0, 6) X(2, 2, 1, 1, 7) X(2, 2, 2, 0, 2) X(2, 2, 2, 1, 3)      \
    X(2, 2, 2, 2, 0)
#define PK_NSHARDS 8

namespace pk {

constexpr int kAccBudget = 48;  // contracted Rys accumulators per lane

constexpr int chunk_size(int la, int lb, int lc, int ld) {
    const int ne = nlev(la, la + lb), ncd = ncart(lc) * ncart(ld);
    int ch = kAccBudget / ne;
    if (ch < 1) ch = 1;
    return ch < ncd ? ch : ncd;
}

// classes with an Obara-Saika kernel: L <= 4 and [e0|f0] accumulators <= 54
constexpr bool os_available(int la, int lb, int lc, int ld) {
    return la + lb + lc + ld <= 4 && nlev(la, la + lb) * nlev(lc, lc + ld) <= 54;
}

template <int LA, int LB, int LC, int LD, int C0>
void launch_jk_rys(const JKArgs &a, int grid, hipStream_t s, int &nlaunch) {
    constexpr int NCD = ncart(LC) * ncart(LD);
    if constexpr (C0 < NCD) {
        constexpr int CH = chunk_size(LA, LB, LC, LD);
        constexpr int C1 = C0 + CH < NCD ? C0 + CH : NCD;
        hipLaunchKernelGGL((k_jk<LA, LB, LC, LD, ALG_RYS, C0, C1>), dim3(grid), dim3(256), 0, s, a);
        ++nlaunch;
        launch_jk_rys<LA, LB, LC, LD, C1>(a, grid, s, nlaunch);
    }
}

template <int LA, int LB, int LC, int LD, int C0>
void launch_eri_rys(const EriArgs &a, int grid, hipStream_t s) {
    constexpr int NCD = ncart(LC) * ncart(LD);
    if constexpr (C0 < NCD) {
        constexpr int CH = chunk_size(LA, LB, LC, LD);
        constexpr int C1 = C0 + CH < NCD ? C0 + CH : NCD;
        hipLaunchKernelGGL((k_eri<LA, LB, LC, LD, ALG_RYS, C0, C1>), dim3(grid), dim3(256), 0, s, a);
        launch_eri_rys<LA, LB, LC, LD, C1>(a, grid, s);
    }
}

template <int LA, int LB, int LC, int LD>
void launch_jk_os(const JKArgs &a, int grid, hipStream_t s, int &nlaunch) {
    if constexpr (os_available(LA, LB, LC, LD)) {
        constexpr int NCD = ncart(LC) * ncart(LD);
        hipLaunchKernelGGL((k_jk<LA, LB, LC, LD, ALG_OS, 0, NCD>), dim3(grid), dim3(256), 0, s, a);
        ++nlaunch;
    }
}

template <int LA, int LB, int LC, int LD>
void launch_eri_os(const EriArgs &a, int grid, hipStream_t s) {
    if constexpr (os_available(LA, LB, LC, LD)) {
        constexpr int NCD = ncart(LC) * ncart(LD);
        hipLaunchKernelGGL((k_eri<LA, LB, LC, LD, ALG_OS, 0, NCD>), dim3(grid), dim3(256), 0, s, a);
    }
}

}  // namespace pk
