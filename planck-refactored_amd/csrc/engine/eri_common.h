// eri_common.h — shared layout and compile-time Cartesian index helpers of the
// ERI engine.  Included by the HIP translation units only.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace pk {

// Compile-time loop: f(std::integral_constant<int, i>) for i in [B, E).  Every
// index built from the loop variable is a constant expression, so the arrays
// of the kernels stay in registers without relying on loop unrolling.
template <int B, int E, typename F>
__host__ __device__ __forceinline__ void sfor(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        sfor<B + 1, E>(f);
    }
}
// descending: f(i) for i = B, B-1, ..., E+1
template <int B, int E, typename F>
__host__ __device__ __forceinline__ void sfor_down(F &&f) {
    if constexpr (B > E) {
        f(std::integral_constant<int, B>{});
        sfor_down<B - 1, E>(f);
    }
}

// ---------------------------------------------------------------- device records
// One shell pair (a, b) with L_a >= L_b ("first" shell carries the larger L).
// f0/f1 are the first basis-function indices of the two shells, ab = A - B.
struct PairDev {
    int f0, f1;
    int prim_off, nprim;
    int diag;  // 1 when both shells are the same shell
    int grp;   // ordering group (primitive-count bucket) of the pair as a ket
    double ab[3];
    double qmin;  // smallest Schwarz bound Q_ij over the pair's functions
};

// One primitive pair: zeta = alpha + beta, p = Gaussian product centre,
// pa = P - A (A = first shell), kc = K_ab sqrt(zeta) (4/pi)^(1/4) with
// K_ab = (pi/zeta)^1.5 exp(-alpha beta/zeta |AB|^2) c_a N_a c_b N_b
// (contraction coefficient x Nc, primitive norm).
struct PrimDev {
    double zeta, kc;
    double p[3];
    double pa[3];
};

// ---------------------------------------------------------------- Cartesian order
// Components of a shell of angular momentum l in the reference order
// (src/basis/basis.cpp:20-37): lx descending, then ly descending.
__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }

__host__ __device__ constexpr int cart_x(int l, int k) {
    int lx = l;
    while (k > l - lx) {
        k -= l - lx + 1;
        --lx;
    }
    return lx;
}
__host__ __device__ constexpr int cart_z(int l, int k) {
    int lx = l;
    while (k > l - lx) {
        k -= l - lx + 1;
        --lx;
    }
    return k;
}
__host__ __device__ constexpr int cart_y(int l, int k) { return l - cart_x(l, k) - cart_z(l, k); }
__host__ __device__ constexpr int cart_index(int x, int y, int z) {
    return ((y + z) * (y + z + 1)) / 2 + z;  // (l - x)(l - x + 1)/2 + z
}
// number of components in the levels lo..hi
__host__ __device__ constexpr int nlev(int lo, int hi) {
    int s = 0;
    for (int l = lo; l <= hi; ++l) s += ncart(l);
    return s;
}
// flat index of (x, y, z) in the concatenated level list starting at level lo
__host__ __device__ constexpr int flat(int lo, int x, int y, int z) { return nlev(lo, x + y + z - 1) + cart_index(x, y, z); }
// component of the flat list starting at lo: level and index within level
__host__ __device__ constexpr int flat_level(int lo, int i) {
    int l = lo;
    while (i >= ncart(l)) {
        i -= ncart(l);
        ++l;
    }
    return l;
}
__host__ __device__ constexpr int flat_k(int lo, int i) {
    int l = lo;
    while (i >= ncart(l)) {
        i -= ncart(l);
        ++l;
    }
    return i;
}

// Pair class index of (la, lb), la >= lb: (0,0)=0 (1,0)=1 (1,1)=2 (2,0)=3 ...
__host__ __device__ constexpr int pair_class(int la, int lb) { return la * (la + 1) / 2 + lb; }

// Flop model v1 of one surviving shell quartet with K primitive quartets:
//   K [n (12 (la+lb+1)(lc+ld+1) + 3 Nc) + 20 n] + 12 Nc
// (2-D Rys recurrences, assembly, roots/weights, J/K digestion).
__host__ __device__ constexpr double model_flops(int la, int lb, int lc, int ld, double K) {
    const int L = la + lb + lc + ld, n = L / 2 + 1;
    const double Nc = (double)ncart(la) * ncart(lb) * ncart(lc) * ncart(ld);
    return K * (n * (12.0 * (la + lb + 1) * (lc + ld + 1) + 3.0 * Nc) + 20.0 * n) + 12.0 * Nc;
}

}  // namespace pk
