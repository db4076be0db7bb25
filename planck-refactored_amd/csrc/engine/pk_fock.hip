// pk_fock.hip — host side of the MI355X Fock-build engine and its C ABI
// (include/pk_fock.h).  Per-lane class kernels live in eri_kernels.h (instantiated
// in pk_kernels.hip), the cooperative kernel for large f classes and for ERI
// output in eri_generic.h; this file builds the shell pairs, the Schwarz table
// and the screened quartet work lists, and launches the kernels on the
// context's stream.
//
// Reference path replaced: src/integrals/base.h:57-93 (_compute_2e_fock[_uhf]),
// src/integrals/os.cpp:1435-1625 and rys.cpp:701-824 (the Fock builds),
// os.cpp:1703-1752 (Schwarz table), shellpair.cpp:7-22 (pairs).
#include "pk_fock.h"
#include "pk_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "eri_classes.h"
#include "eri_generic.h"
#include "eri_grad.h"

// per-class launchers, defined in pk_kernels.hip
#define PK_DECL(a, b, c, d, shard)                                                             \
    void pk_launch_jk_rys_##a##b##c##d(const pk::JKArgs &, int, int, hipStream_t, int &);           \
    void pk_launch_jk_os_##a##b##c##d(const pk::JKArgs &, int, int, hipStream_t, int &);            \
    void pk_launch_jk_rys32_##a##b##c##d(const pk::JKArgs &, int, int, hipStream_t, int &);         \
    void pk_launch_jk_os32_##a##b##c##d(const pk::JKArgs &, int, int, hipStream_t, int &);         \
    void pk_launch_jk_os_uhf_##a##b##c##d(const pk::JKArgs &, int, int, hipStream_t, int &);          \
    void pk_launch_jk_os_seg_##a##b##c##d(const pk::JKArgs &, int, int, hipStream_t, int &);          \
    void pk_launch_jk_os32_seg_##a##b##c##d(const pk::JKArgs &, int, int, hipStream_t, int &);        \
    void pk_launch_jk_os_uhf_seg_##a##b##c##d(const pk::JKArgs &, int, int, hipStream_t, int &);
PK_FOR_EACH_CLASS(PK_DECL)
#undef PK_DECL

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define PK_HIP(expr)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail(PK_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_));           \
    } while (0)

constexpr int kMaxL = PK_MAX_L;
constexpr double kPrimCut = 1e-17;  // primitive-pair screening bound (see add_molecule)
// Class kernels of one build run on the context stream and up to kSideQueues
// more streams when the build holds at least kSideMinQuartets shell quartets
// (GPU_MAX_HW_QUEUES is 4: one queue per stream).
constexpr int kSideQueues = 3;
constexpr int64_t kSideMinQuartets = 1 << 20;
// Segment mode of the per-lane kernels (see k_jk): at most kSegWaves waves of
// one bra per segment, and at most kSegLdsMax bytes of LDS exchange rows.
constexpr int kSegWaves = 16;
// primitive quartets per shell quartet below which a list of a pk::seg_alt_class
// runs in segment mode (6-31G* water clusters: 11-13; cc-pVTZ alkanes: 40-80)
constexpr double kSegAltPrim = 24.0;
constexpr size_t kSegLdsMax = 64 * 1024;
// lanes per workgroup of a class list until pk_ctx_autotune picks its own;
// one wave per workgroup for the segment-mode lists of a batch with at most
// kSegOneWaveLds bytes of exchange rows (see launch_list)
constexpr int kDefaultBlock = 256;
constexpr size_t kSegOneWaveLds = 24 * 1024;
constexpr int kPairClasses = (kMaxL + 1) * (kMaxL + 2) / 2;
constexpr int kQuartClasses = kPairClasses * (kPairClasses + 1) / 2;

struct ClassFns {
    int la, lb, lc, ld;
    bool has_os;    // Obara-Saika per-lane kernel
    bool has_tmpl;  // Rys per-lane kernel (else the cooperative kernel)
    bool seg;       // the per-lane kernels run in segment mode (pk::seg_class)
    bool seg_alt;   // the OS kernel also exists in segment mode (pk::seg_alt_class)
    void (*jk_os)(const pk::JKArgs &, int, int, hipStream_t, int &);
    void (*jk_rys)(const pk::JKArgs &, int, int, hipStream_t, int &);
    void (*jk_os32)(const pk::JKArgs &, int, int, hipStream_t, int &);  // FP32 integrals
    void (*jk_rys32)(const pk::JKArgs &, int, int, hipStream_t, int &);
    void (*jk_os_uhf)(const pk::JKArgs &, int, int, hipStream_t, int &);  // FP64, two exchange densities
    void (*jk_os_seg)(const pk::JKArgs &, int, int, hipStream_t, int &);  // segment-mode variants
    void (*jk_os32_seg)(const pk::JKArgs &, int, int, hipStream_t, int &);
    void (*jk_os_uhf_seg)(const pk::JKArgs &, int, int, hipStream_t, int &);
};

#define PK_ENTRY(a, b, c, d, shard)                                                             \
    ClassFns{a, b, c, d, pk::os_available(a, b, c, d), pk::rys_tmpl_available(a, b, c, d),      \
             pk::seg_class(a, b, c, d), pk::os_available(a, b, c, d) && pk::seg_alt_class(a, b, c, d), \
             &pk_launch_jk_os_##a##b##c##d, &pk_launch_jk_rys_##a##b##c##d,                     \
             &pk_launch_jk_os32_##a##b##c##d, &pk_launch_jk_rys32_##a##b##c##d,                 \
             &pk_launch_jk_os_uhf_##a##b##c##d, &pk_launch_jk_os_seg_##a##b##c##d,                   \
             &pk_launch_jk_os32_seg_##a##b##c##d, &pk_launch_jk_os_uhf_seg_##a##b##c##d},
// Quartet classes (bra class >= ket class), indexed cb (cb + 1) / 2 + ck with
// pair classes (0,0) (1,0) (1,1) (2,0) (2,1) (2,2) (3,0) (3,1) (3,2) (3,3).
const ClassFns kClasses[kQuartClasses] = {PK_FOR_EACH_CLASS(PK_ENTRY)};
#undef PK_ENTRY

int quart_class(int cb, int ck) { return cb * (cb + 1) / 2 + ck; }
int class_la(int c) {
    int l = 0;
    while ((l + 1) * (l + 2) / 2 <= c) ++l;
    return l;
}
int class_lb(int c) { return c - class_la(c) * (class_la(c) + 1) / 2; }

// ---------------------------------------------------------------- small kernels
// Packed matrices: molecule m (blockIdx.y) has nb = mol_nb[m] functions, its
// nb x nb block at mol_off[m].  Inside the engine a molecule's functions are
// in the internal order: shells grouped by angular momentum (s, p, d, f), in
// basis order within a group, components contiguous.  Kets of one bra that
// share their first shell then have second shells of one type at neighbouring
// positions, so a wave's atomic adds into a row of J or K land in a few cache
// lines (the L2 executes a 64-lane FP64 atomic per cache line touched: ~24 G
// lines/s, against ~180 G lanes/s when the lanes share lines).  pos[foff[m] +
// f] is the internal position of the caller's function f; the component
// norms nrm[foff[m] + x] are stored by internal position x.
struct MolTab {
    const long long *off;
    const int *nb;
    const int *foff;
    const int *pos;
};

// Pn (internal order) = n n^T o P (caller's order)
__global__ void k_scale_density(const double *__restrict__ P, const double *__restrict__ nrm,
                                double *__restrict__ Pn, MolTab t) {
    const int m = blockIdx.y, nb = t.nb[m];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nb * nb) return;
    const int r = (int)(i / nb), c = (int)(i % nb);
    const int *ps = t.pos + t.foff[m];
    const int pr = ps[r], pc = ps[c];
    const double *n = nrm + t.foff[m];
    const size_t o = (size_t)t.off[m];
    Pn[o + (size_t)pr * nb + pc] = n[pr] * n[pc] * P[o + i];
}

// out (caller's order) = in (internal order)
__global__ void k_unpermute(const double *__restrict__ in, double *__restrict__ out, MolTab t) {
    const int m = blockIdx.y, nb = t.nb[m];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nb * nb) return;
    const int r = (int)(i / nb), c = (int)(i % nb);
    const int *ps = t.pos + t.foff[m];
    const size_t o = (size_t)t.off[m];
    out[o + i] = in[o + (size_t)ps[r] * nb + ps[c]];
}

// y = (float) x
__global__ void k_to_f32(const double *__restrict__ x, float *__restrict__ y, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (float)x[i];
}

// z = x + y
__global__ void k_sum2(const double *__restrict__ x, const double *__restrict__ y, double *__restrict__ z, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = x[i] + y[i];
}

__global__ void k_fill(double *__restrict__ x, size_t n, double v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

// J = n n^T o (A + A^T), K = n n^T o (B + B^T) (A, B internal order; J, K, G
// caller's order); optional G = J - K/2
__global__ void k_finalize(const double *__restrict__ A, const double *__restrict__ B,
                           const double *__restrict__ nrm, double *J, double *K, double *G, MolTab t) {
    const int m = blockIdx.y, nb = t.nb[m];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nb * nb) return;
    const int r = (int)(i / nb), c = (int)(i % nb);
    const int *ps = t.pos + t.foff[m];
    const int pr = ps[r], pc = ps[c];
    const double *n = nrm + t.foff[m];
    const size_t o = (size_t)t.off[m];
    const double s = n[pr] * n[pc];
    const size_t x = o + (size_t)pr * nb + pc, y = o + (size_t)pc * nb + pr;
    const double j = s * (A[x] + A[y]);
    const double k = s * (B[x] + B[y]);
    if (J) J[o + i] = j;
    if (K) K[o + i] = k;
    if (G) G[o + i] = j - 0.5 * k;
}

struct TuneEntry {
    int cls, prec, block, remap, alt;
};

struct ClassWork {
    int cls = 0;
    int prec = PK_PREC_FP64;  // PK_PREC_MIXED: the FP32 list of the class
    int block = 0;            // lanes per workgroup (autotuned or PK_BLOCK; 0: the default of the launch)
    int block_run = 0;        // lanes per workgroup of the last launch
    int remap = 0;            // XCD-contiguous wave ranges (autotuned)
    int alt = 0;              // alternative kernel (autotuned): FP64 integrals for an FP32 list,
                              // the cooperative kernel for a per-lane Rys class
    int nwaves = 0;
    int64_t nquart = 0;
    int64_t nprimq = 0;
    double flops = 0.0;
    int *d_wave = nullptr;  // [4][nwaves]: bra, start, count, molecule
    int *d_kets = nullptr;
    int nseg = 0;          // segments: runs of at most kSegWaves waves of one bra
    int *d_seg = nullptr;  // [nseg + 1] first wave of each segment
    int kcols = 0;         // LDS columns of an exchange row (largest molecule)
    int kernels = 0;
    double ms_total = 0.0;
    int64_t builds = 0;
    double cost_ms = 0.0;  // measured time of the list in the last timed build (0: not yet)
    hipEvent_t e0 = nullptr, e1 = nullptr;
};

template <typename T>
int dev_upload(T **dst, const T *src, size_t n, hipStream_t s) {
    *dst = nullptr;
    if (n == 0) return PK_OK;
    PK_HIP(hipMalloc((void **)dst, n * sizeof(T)));
    PK_HIP(hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
    return PK_OK;
}

template <typename T>
int dev_upload(T **dst, const std::vector<T> &src, hipStream_t s) {
    return dev_upload(dst, src.data(), src.size(), s);
}

}  // namespace

struct pk_ctx {
    int nb = 0, ns = 0;  // functions and shells over all molecules
    int nmol = 1;
    std::vector<int> mol_nb, mol_foff;  // functions per molecule, first function
    std::vector<int> tblk;              // [5 * nmol]: first function of type s, p, d, f (internal order), nb
    int *d_tblk = nullptr;
    bool seg_on = true;                 // segment mode of the per-lane kernels (PK_EXCHANGE_ROWS=0: off)
    std::vector<long long> mol_off;     // offset of the molecule's nb x nb block
    size_t n2 = 0;                      // packed matrix length
    int nbmax = 0;
    std::vector<int> pair_mol;
    long long *d_mol_off = nullptr;
    int *d_mol_nb = nullptr, *d_mol_foff = nullptr;
    bool graph_on = false, graph_ok = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    double *d_G = nullptr;
    std::vector<TuneEntry> tuned;  // autotuned launch shape per (class, precision)
    std::vector<int> sh_l, sh_np, sh_poff, sh_f0;
    std::vector<double> sh_c, ex, co, pn;
    std::vector<double> fnorm;  // component norms by internal position
    std::vector<int> fpos;      // internal position of each caller's function (per molecule)
    std::vector<int> fnat;      // caller's function at each internal position
    int *d_fpos = nullptr;
    std::vector<pk::PairDev> pairs;
    std::vector<int> pair_cls;
    std::vector<double> pair_q;
    std::vector<pk::PrimDev> prims;
    std::vector<double> palpha;  // exponent on the pair's first shell, per primitive pair
    std::vector<double> hz;      // 1/(2 zeta) per primitive pair
    std::vector<int> pair_sh;    // first and second shell of each pair
    std::vector<double> Q;  // packed nb x nb blocks, normalised functions, internal order
    pk::PairDev *d_pairs = nullptr;
    pk::PrimDev *d_prims = nullptr;
    double *d_Q = nullptr, *d_fnorm = nullptr, *d_P = nullptr, *d_Pn = nullptr, *d_A = nullptr, *d_B = nullptr;
    double *d_J = nullptr, *d_K = nullptr;
    double *d_J2 = nullptr;  // UHF: J of the second density
    double *d_hz = nullptr;
    double *d_palpha = nullptr, *d_Pk1n = nullptr, *d_Pk2n = nullptr, *d_gsh = nullptr;  // gradient, UHF
    double *d_B2 = nullptr;  // one-pass UHF: second exchange accumulator
    float *d_P32 = nullptr;  // mixed precision: FP32 copy of the scaled density
    int *d_pair_sh = nullptr;
    std::vector<ClassWork> work;
    bool work_ok = false;
    int engine = PK_ENGINE_OS;
    double tol = 1e-10;
    int precision = PK_PREC_FP64;
    double cutoff = 0.0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // side queues of a build: the class kernels are spread over the context
    // stream and these (default priority), forked and joined by events
    hipStream_t side[kSideQueues] = {};
    hipEvent_t fork_ev = nullptr, join_ev[kSideQueues] = {};
    bool side_ok = false;
    bool side_off = false;  // PK_SIDE_QUEUES=0
    double seg_alt_prim = kSegAltPrim;  // PK_SEG_ALT_PRIM
    int block0 = kDefaultBlock;         // PK_BLOCK: lanes per workgroup of every list before autotuning
    bool block_forced = false;          // PK_BLOCK given: no per-list default
    bool cost_learned = false;  // the lists' times have been measured (first large build)
    bool learning = false;      // the current build records per-list events for that
    bool profile = false;
    pk_stats st{};
};

namespace {

// OS kernels where they exist unless the Rys engine is forced
bool use_os(const pk_ctx *c, const ClassFns &f) { return f.has_os && c->engine != PK_ENGINE_RYS; }

// Kernels that run a class list (pk_class_stats.kind): 0 / 3 per-lane OS in
// FP64 / FP32, 1 / 4 per-lane Rys in FP64 / FP32, 2 cooperative (FP64).
int list_kind(const pk_ctx *c, const ClassWork &w) {
    const ClassFns &f = kClasses[w.cls];
    const bool fp32 = w.prec == PK_PREC_MIXED && !w.alt;
    if (use_os(c, f)) return fp32 ? 3 : 0;
    if (f.has_tmpl && !(w.alt && w.prec == PK_PREC_FP64)) return fp32 ? 4 : 1;
    return 2;
}

void drop_graph(pk_ctx *c) {
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
    c->graph_ok = false;
}

void free_work(pk_ctx *c) {
    for (ClassWork &w : c->work) {
        if (w.d_wave) (void)hipFree(w.d_wave);
        if (w.d_kets) (void)hipFree(w.d_kets);
        if (w.d_seg) (void)hipFree(w.d_seg);
        if (w.e0) (void)hipEventDestroy(w.e0);
        if (w.e1) (void)hipEventDestroy(w.e1);
    }
    c->work.clear();
    c->work_ok = false;
    c->cost_learned = false;
    drop_graph(c);
}

// Raw contracted ERIs (no component norms) of the quartets (bra[i] | ket[i]),
// all of class (cb, ck) with cb >= ck.  out gets NA NB NC ND values per quartet.
// One 64-lane workgroup per item; the dynamic LDS holds the transferred 1-D
// factors and the quartet's accumulators (at most 123 KB for (ff|ff)).
// Cooperative kernels with the class fixed at compile time, for the J/K lists
// of the classes without a per-lane kernel (null elsewhere), with
// pk::gen_waves(class) waves per quartet.
struct GenKernels {
    const void *k1, *k2;  // one and two exchange densities
    int waves;
};
template <int A, int B, int C, int D>
constexpr GenKernels gen_kernels() {
    if constexpr (!pk::rys_tmpl_available(A, B, C, D)) {
        constexpr int W = pk::gen_waves(A, B, C, D);
        return GenKernels{(const void *)pk::k_generic<1, A, B, C, D, W>, (const void *)pk::k_generic<2, A, B, C, D, W>, W};
    } else {
        return GenKernels{nullptr, nullptr, 1};
    }
}
#define PK_GEN_ENTRY(a, b, c, d, shard) gen_kernels<a, b, c, d>(),
const GenKernels kGenKernels[kQuartClasses] = {PK_FOR_EACH_CLASS(PK_GEN_ENTRY)};

int launch_generic(const pk::GenArgs &a, hipStream_t s) {
    if (a.nitems <= 0) return PK_OK;
    if (!a.eri_mode) {
        const int cb = pk::pair_class(a.la, a.lb), ck = pk::pair_class(a.lc, a.ld);
        const GenKernels &g = kGenKernels[quart_class(cb, ck)];
        const void *k = a.B2 ? g.k2 : g.k1;
        if (k) {
            static std::atomic<int> attr_done[2 * kQuartClasses];  // zero-initialised
            const size_t bytes = pk::generic_lds_bytes(a.la, a.lb, a.lc, a.ld, 64 * g.waves);
            std::atomic<int> &done = attr_done[2 * quart_class(cb, ck) + (a.B2 ? 1 : 0)];
            if (!done.load()) {
                PK_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)pk::generic_lds_bytes(kMaxL, kMaxL, kMaxL, kMaxL)));
                done.store(1);
            }
            pk::GenArgs x = a;
            void *args[] = {&x};
            PK_HIP(hipLaunchKernel(k, dim3(a.nitems), dim3(64 * g.waves), args, bytes, s));
            return PK_OK;
        }
    }
    static const size_t max_bytes = pk::generic_lds_bytes(kMaxL, kMaxL, kMaxL, kMaxL);
    static const hipError_t attr1 =
        hipFuncSetAttribute((const void *)pk::k_generic<1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_bytes);
    static const hipError_t attr2 =
        hipFuncSetAttribute((const void *)pk::k_generic<2>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_bytes);
    PK_HIP(attr1);
    PK_HIP(attr2);
    const size_t bytes = pk::generic_lds_bytes(a.la, a.lb, a.lc, a.ld);
    if (a.B2)
        hipLaunchKernelGGL(pk::k_generic<2>, dim3(a.nitems), dim3(64), bytes, s, a);
    else
        hipLaunchKernelGGL(pk::k_generic<1>, dim3(a.nitems), dim3(64), bytes, s, a);
    PK_HIP(hipGetLastError());
    return PK_OK;
}

int run_eri_class(pk_ctx *c, int cb, int ck, const std::vector<int> &bra, const std::vector<int> &ket,
                  std::vector<double> &out) {
    const ClassFns &f = kClasses[quart_class(cb, ck)];
    const int nabcd = pk::ncart(f.la) * pk::ncart(f.lb) * pk::ncart(f.lc) * pk::ncart(f.ld);
    const int nq = (int)bra.size();
    out.assign((size_t)nq * nabcd, 0.0);
    if (nq == 0) return PK_OK;
    int *d_b = nullptr, *d_k = nullptr;
    double *d_o = nullptr;
    int rc;
    if ((rc = dev_upload(&d_b, bra, c->stream)) != PK_OK) return rc;
    if ((rc = dev_upload(&d_k, ket, c->stream)) != PK_OK) return rc;
    PK_HIP(hipMalloc((void **)&d_o, out.size() * sizeof(double)));
    pk::GenArgs a{};
    a.qbra = d_b;
    a.qket = d_k;
    a.pairs = c->d_pairs;
    a.prims = c->d_prims;
    a.hz = c->d_hz;
    a.out = d_o;
    a.nitems = nq;
    a.la = f.la;
    a.lb = f.lb;
    a.lc = f.lc;
    a.ld = f.ld;
    a.eri_mode = 1;
    if ((rc = launch_generic(a, c->stream)) != PK_OK) return rc;
    PK_HIP(hipMemcpyAsync(out.data(), d_o, out.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PK_HIP(hipStreamSynchronize(c->stream));
    (void)hipFree(d_b);
    (void)hipFree(d_k);
    (void)hipFree(d_o);
    return PK_OK;
}

int build_schwarz(pk_ctx *c) {
    c->Q.assign(c->n2, 0.0);
    for (int cls = 0; cls < kPairClasses; ++cls) {
        std::vector<int> idx;
        for (int p = 0; p < (int)c->pairs.size(); ++p)
            if (c->pair_cls[p] == cls) idx.push_back(p);
        if (idx.empty()) continue;
        std::vector<double> out;
        int rc = run_eri_class(c, cls, cls, idx, idx, out);
        if (rc != PK_OK) return rc;
        const int la = class_la(cls), lb = class_lb(cls);
        const int na = pk::ncart(la), nbb = pk::ncart(lb), nab = na * nbb;
        for (size_t q = 0; q < idx.size(); ++q) {
            const pk::PairDev &p = c->pairs[idx[q]];
            const int m = c->pair_mol[idx[q]], nb = c->mol_nb[m];
            const double *fn = c->fnorm.data() + c->mol_foff[m];
            double *Qm = c->Q.data() + c->mol_off[m];
            double qmax = 0.0, qmin = 1e300;
            for (int ia = 0; ia < na; ++ia)
                for (int ib = 0; ib < nbb; ++ib) {
                    const int i = p.f0 + ia, j = p.f1 + ib;
                    const double raw = out[q * (size_t)nab * nab + (size_t)(ia * nbb + ib) * nab + ia * nbb + ib];
                    const double v = fn[i] * fn[j] * std::sqrt(std::fabs(raw));
                    Qm[(size_t)i * nb + j] = Qm[(size_t)j * nb + i] = v;
                    qmax = std::max(qmax, v);
                    qmin = std::min(qmin, v);
                }
            c->pair_q[idx[q]] = qmax;
            c->pairs[idx[q]].qmin = qmin;
        }
    }
    return PK_OK;
}

// Runs f(i) for i in [0, n) on up to 16 host threads (dynamic chunks of 64).
template <typename F>
void parallel_for(size_t n, F f) {
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const unsigned nt = (unsigned)std::min<size_t>(hw, (n + 63) / 64);
    if (nt <= 1) {
        for (size_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<size_t> next{0};
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&] {
            for (size_t i0; (i0 = next.fetch_add(64)) < n;)
                for (size_t i = i0; i < std::min(n, i0 + 64); ++i) f(i);
        });
    for (std::thread &t : th) t.join();
}

// Screened work lists: for every quartet class (cb >= ck), every molecule and
// every bra pair, the ket pairs of the same molecule with
// qmax_bra * qmax_ket >= tol (and ket <= bra when cb == ck), sorted by
// primitive count and shells, cut into waves of at most 64 kets; the waves of all
// molecules form one list per class.  In mixed precision the kets with
// qmax_bra * qmax_ket < cutoff form a second, FP32 list of the class (classes
// of the cooperative kernel stay FP64).
int ensure_side(pk_ctx *c);

int build_work(pk_ctx *c) {
    free_work(c);
    const int np = (int)c->pairs.size();
    // by_cls[m][class]: pairs of molecule m, q descending
    std::vector<std::vector<std::vector<int>>> by_cls(c->nmol, std::vector<std::vector<int>>(kPairClasses));
    for (int p = 0; p < np; ++p) by_cls[c->pair_mol[p]][c->pair_cls[p]].push_back(p);
    for (auto &mv : by_cls)
        for (auto &v : mv)
            std::stable_sort(v.begin(), v.end(), [&](int x, int y) { return c->pair_q[x] > c->pair_q[y]; });
    // ket order inside a wave: ordering group descending, then first and
    // second function (one 64-bit key per pair)
    std::vector<uint64_t> okey(np);
    for (int p = 0; p < np; ++p)
        okey[p] = ((uint64_t)(2047 - c->pairs[p].grp) << 40) | ((uint64_t)c->pairs[p].f0 << 20) | (uint64_t)c->pairs[p].f1;
    c->st.quartets = 0;
    c->st.primitive_quartets = 0;
    c->st.model_flops = 0.0;
    const bool mixed = c->precision == PK_PREC_MIXED;
    // ord[m][class]: positions in by_cls[m][class] in ket order (okey
    // ascending), so that a bra's kets -- a prefix range of the q order -- come
    // out in ket order by one pass over it instead of a sort per bra
    std::vector<std::vector<std::vector<int>>> ord(c->nmol, std::vector<std::vector<int>>(kPairClasses));
    parallel_for((size_t)c->nmol, [&](size_t m) {
        for (int cl = 0; cl < kPairClasses; ++cl) {
            const std::vector<int> &v = by_cls[m][cl];
            std::vector<int> &o = ord[m][cl];
            o.resize(v.size());
            for (size_t i = 0; i < v.size(); ++i) o[i] = (int)i;
            std::sort(o.begin(), o.end(), [&](int x, int y) { return okey[v[x]] < okey[v[y]]; });
        }
    });
    struct BraRange {
        size_t k0 = 0, k1 = 0;  // positions [k0, k1) of the q-ordered kets
        int64_t nprimq = 0;
        double flops = 0.0;
    };
    for (int cb = 0; cb < kPairClasses; ++cb)
        for (int ck = 0; ck <= cb; ++ck) {
            const int cls = quart_class(cb, ck);
            const ClassFns &f = kClasses[cls];
            const bool split = mixed && f.has_tmpl;
            // kets per work item: a wave of 64 lanes for the per-lane kernels,
            // a few quartets per workgroup for the cooperative kernel
            const size_t group = f.has_tmpl ? 64 : 4;
            // bra tasks of every molecule, in order
            std::vector<std::pair<int, int>> tasks;
            for (int m = 0; m < c->nmol; ++m)
                if (!by_cls[m][cb].empty() && !by_cls[m][ck].empty())
                    for (int ib = 0; ib < (int)by_cls[m][cb].size(); ++ib) tasks.emplace_back(m, ib);
            if (tasks.empty()) continue;
            for (int prec : {PK_PREC_FP64, PK_PREC_MIXED}) {
                if (prec == PK_PREC_MIXED && !split) continue;
                // pass 1: each bra's ket range (binary searches only)
                std::vector<BraRange> outs(tasks.size());
                parallel_for(tasks.size(), [&](size_t t) {
                    const int m = tasks[t].first, ib = tasks[t].second;
                    const std::vector<int> &B = by_cls[m][cb], &Kl = by_cls[m][ck];
                    const double qb = c->pair_q[B[ib]];
                    // number of kets (q descending) with qb * qk >= thr
                    auto count = [&](double thr) {
                        size_t lo = 0, hi = Kl.size();
                        while (lo < hi) {
                            const size_t mid = (lo + hi) / 2;
                            if (qb * c->pair_q[Kl[mid]] >= thr)
                                lo = mid + 1;
                            else
                                hi = mid;
                        }
                        return cb == ck ? std::min(lo, (size_t)ib + 1) : lo;
                    };
                    const size_t n_tol = count(c->tol);
                    const size_t n_hi = split ? std::min(count(c->cutoff), n_tol) : n_tol;
                    BraRange &o = outs[t];
                    o.k0 = prec == PK_PREC_FP64 ? 0 : n_hi;
                    o.k1 = std::max(o.k0, prec == PK_PREC_FP64 ? n_hi : n_tol);
                });
                ClassWork w;
                w.cls = cls;
                w.prec = prec;
                w.block = c->block_forced ? c->block0 : 0;
                // offsets of every bra's kets and waves (prefix sums), then the
                // flat arrays filled in parallel
                std::vector<size_t> koff(tasks.size() + 1, 0), woff(tasks.size() + 1, 0);
                for (size_t t = 0; t < tasks.size(); ++t) {
                    const size_t n = outs[t].k1 - outs[t].k0;
                    koff[t + 1] = koff[t] + n;
                    woff[t + 1] = woff[t] + (n + group - 1) / group;
                }
                if (koff.back() > (size_t)INT32_MAX) return fail(PK_ERR_ARG, "build_work: more than 2^31 quartets in one class");
                const size_t nw = woff.back();
                // the ket array is written entirely below: no value-initialisation
                // (it holds one index per quartet, up to GBs for large batches)
                std::unique_ptr<int[]> kets(new int[std::max<size_t>(1, koff.back())]);
                std::vector<int> packed(4 * nw);
                // pass 2: every bra's kets in ket order (ordering group
                // descending, then first and second function) straight into the
                // flat array, and its waves
                parallel_for(tasks.size(), [&](size_t t) {
                    BraRange &o = outs[t];
                    const size_t n = o.k1 - o.k0;
                    if (n == 0) return;
                    const int m = tasks[t].first;
                    const std::vector<int> &Kl = by_cls[m][ck];
                    int *dst = kets.get() + koff[t];
                    if (Kl.size() <= 16 * n + 64) {
                        size_t j = 0;
                        for (const int i : ord[m][ck])
                            if ((size_t)i >= o.k0 && (size_t)i < o.k1) {
                                dst[j++] = Kl[i];
                                if (j == n) break;
                            }
                    } else {
                        std::vector<std::pair<uint64_t, int>> kk(n);
                        for (size_t i = o.k0; i < o.k1; ++i) kk[i - o.k0] = {okey[Kl[i]], Kl[i]};
                        std::sort(kk.begin(), kk.end());
                        for (size_t i = 0; i < n; ++i) dst[i] = kk[i].second;
                    }
                    const int b = by_cls[m][cb][tasks[t].second];
                    const int kb = c->pairs[b].nprim;
                    for (size_t i = 0; i < n; ++i) {
                        const double K = (double)kb * c->pairs[dst[i]].nprim;
                        o.nprimq += (int64_t)K;
                        o.flops += pk::model_flops(f.la, f.lb, f.lc, f.ld, K);
                    }
                    size_t wi = woff[t];
                    for (size_t s0 = 0; s0 < n; s0 += group, ++wi) {
                        packed[wi] = b;
                        packed[nw + wi] = (int)(koff[t] + s0);
                        packed[2 * nw + wi] = (int)std::min<size_t>(group, n - s0);
                        packed[3 * nw + wi] = m;
                    }
                });
                for (size_t t = 0; t < tasks.size(); ++t) {
                    w.nquart += (int64_t)(outs[t].k1 - outs[t].k0);
                    w.nprimq += outs[t].nprimq;
                    w.flops += outs[t].flops;
                }
                if (w.nquart == 0) continue;
                w.nwaves = (int)nw;
                int rc;
                if ((rc = dev_upload(&w.d_wave, packed, c->stream)) != PK_OK) return rc;
                if ((rc = dev_upload(&w.d_kets, kets.get(), koff.back(), c->stream)) != PK_OK) return rc;
                if (f.has_tmpl) {
                    // segments of the per-lane kernels: each bra's waves in runs
                    // of at most kSegWaves; LDS columns: the two ket shell types
                    // of the largest molecule
                    std::vector<int> seg;
                    for (size_t t = 0; t < tasks.size(); ++t) {
                        for (size_t w0 = woff[t]; w0 < woff[t + 1]; w0 += kSegWaves) seg.push_back((int)w0);
                        if (woff[t + 1] > woff[t]) {
                            const int *tb = &c->tblk[5 * tasks[t].first];
                            const int ncol = tb[f.lc + 1] - tb[f.lc] + (f.lc == f.ld ? 0 : tb[f.ld + 1] - tb[f.ld]);
                            w.kcols = std::max(w.kcols, ncol);
                        }
                    }
                    seg.push_back((int)nw);
                    w.nseg = (int)seg.size() - 1;
                    if ((rc = dev_upload(&w.d_seg, seg, c->stream)) != PK_OK) return rc;
                }
                c->st.quartets += w.nquart;
                c->st.primitive_quartets += w.nprimq;
                c->st.model_flops += w.flops;
                c->work.push_back(w);
            }
        }
    // largest classes first
    std::sort(c->work.begin(), c->work.end(), [](const ClassWork &x, const ClassWork &y) { return x.flops > y.flops; });
    // launch shapes chosen by pk_ctx_autotune for this context, if any
    for (ClassWork &w : c->work)
        for (const TuneEntry &t : c->tuned)
            if (t.cls == w.cls && t.prec == w.prec) {
                w.block = t.block;
                w.remap = t.remap;
                w.alt = t.alt;
            }
    PK_HIP(hipStreamSynchronize(c->stream));
    if (c->st.quartets >= kSideMinQuartets) {
        int rc;
        if ((rc = ensure_side(c)) != PK_OK) return rc;
    }
    c->work_ok = true;
    return PK_OK;
}

int ensure_buffers(pk_ctx *c) {
    const size_t n2 = c->n2 * sizeof(double);
    if (!c->d_P) PK_HIP(hipMalloc((void **)&c->d_P, n2));
    if (!c->d_Pn) PK_HIP(hipMalloc((void **)&c->d_Pn, n2));
    if (!c->d_A) PK_HIP(hipMalloc((void **)&c->d_A, n2));
    if (!c->d_B) PK_HIP(hipMalloc((void **)&c->d_B, n2));
    if (!c->d_J) PK_HIP(hipMalloc((void **)&c->d_J, n2));
    if (!c->d_K) PK_HIP(hipMalloc((void **)&c->d_K, n2));
    if (!c->d_G) PK_HIP(hipMalloc((void **)&c->d_G, n2));
    if (!c->d_P32 && c->precision == PK_PREC_MIXED) PK_HIP(hipMalloc((void **)&c->d_P32, c->n2 * sizeof(float)));
    return PK_OK;
}

// Enqueue the kernels of one build on the context stream: P (device) -> A, B
// accumulators -> J, K, G (any may be null).  Class events only when profiling
// (never while the stream is being captured).
// Per-build scratch: the scaled density and the J, K accumulators (packed,
// c->n2 doubles each).
// One-pass UHF adds the scaled exchange densities PKa, PKb and the second K
// accumulator B2 (all null for RHF; Pn then holds Pa + Pb).
// P32 is an FP32 copy of Pn for the FP32 kernels of the mixed precision.
struct BuildBufs {
    double *Pn, *A, *B;
    float *P32 = nullptr;
    double *PKa = nullptr, *PKb = nullptr, *B2 = nullptr;
};

BuildBufs ctx_bufs(pk_ctx *c) { return BuildBufs{c->d_Pn, c->d_A, c->d_B, c->d_P32}; }

// Enqueue the class kernels of one build (the density already scaled into
// sb.Pn, the accumulators cleared); nl counts the launches.  With sb.B2 set
// (one-pass UHF) every list runs FP64 two-density digestion: the per-lane OS
// kernels where the class has them, else the cooperative kernel.
int launch_list(pk_ctx *c, ClassWork &w, const BuildBufs &sb, bool events, hipStream_t s, int &nl);

// Side streams and events of a context, created on first use.
int ensure_side(pk_ctx *c) {
    if (c->side_ok) return PK_OK;
    if (!c->fork_ev) PK_HIP(hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming));
    for (int i = 0; i < kSideQueues; ++i) {
        if (!c->side[i]) PK_HIP(hipStreamCreateWithFlags(&c->side[i], hipStreamNonBlocking));
        if (!c->join_ev[i]) PK_HIP(hipEventCreateWithFlags(&c->join_ev[i], hipEventDisableTiming));
    }
    c->side_ok = true;
    return PK_OK;
}

int launch_classes(pk_ctx *c, const BuildBufs &sb, bool events, int &nl) {
    int rc;
    hipStream_t s = c->stream;
    // Concurrent queues (not while profiling: the class events time one list
    // at a time): lists in descending model cost, each to the queue with the
    // least cost so far, so that the small classes fill the tails of the
    // large ones.  All lists add into the same accumulators with atomics.
    if (!events && c->side_ok && !c->side_off && c->st.quartets >= kSideMinQuartets && c->work.size() > 1) {
        PK_HIP(hipEventRecord(c->fork_ev, s));
        hipStream_t q[kSideQueues + 1] = {s};
        for (int i = 0; i < kSideQueues; ++i) {
            q[i + 1] = c->side[i];
            PK_HIP(hipStreamWaitEvent(c->side[i], c->fork_ev, 0));
        }
        std::vector<size_t> order(c->work.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = i;
        // cost of a list: its measured time when the context has one (the first
        // large build of a context times every list, one at a time), else the
        // model flops over a typical rate of its kernel (TFLOP/s: per-lane OS
        // ~10, per-lane Rys ~3, cooperative ~1.5)
        std::vector<double> cost(c->work.size());
        for (size_t i = 0; i < cost.size(); ++i) {
            const ClassWork &w = c->work[i];
            const int kind = list_kind(c, w) % 3;  // 0 OS, 1 Rys, 2 cooperative
            cost[i] = w.cost_ms > 0.0 ? w.cost_ms : w.flops * 1e-9 / (kind == 0 ? 10.0 : kind == 1 ? 3.0 : 1.5);
        }
        std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return cost[x] > cost[y]; });
        double load[kSideQueues + 1] = {};
        for (size_t i : order) {
            int k = 0;
            for (int j = 1; j <= kSideQueues; ++j)
                if (load[j] < load[k]) k = j;
            load[k] += cost[i];
            if ((rc = launch_list(c, c->work[i], sb, false, q[k], nl)) != PK_OK) return rc;
        }
        for (int i = 0; i < kSideQueues; ++i) {
            PK_HIP(hipEventRecord(c->join_ev[i], c->side[i]));
            PK_HIP(hipStreamWaitEvent(s, c->join_ev[i], 0));
        }
        return PK_OK;
    }
    for (ClassWork &w : c->work)
        if ((rc = launch_list(c, w, sb, events, s, nl)) != PK_OK) return rc;
    return PK_OK;
}

// Enqueue the kernels of one class list on stream s.
int launch_list(pk_ctx *c, ClassWork &w, const BuildBufs &sb, bool events, hipStream_t s, int &nl) {
    int rc;
    {
        const int *wv = w.d_wave;
        pk::JKArgs a{};
        a.wave_bra = wv;
        a.wave_start = wv + w.nwaves;
        a.wave_count = wv + 2 * w.nwaves;
        a.wave_mol = wv + 3 * w.nwaves;
        a.kets = w.d_kets;
        a.pairs = c->d_pairs;
        a.prims = c->d_prims;
        a.hz = c->d_hz;
        a.mol_off = c->d_mol_off;
        a.mol_nb = c->d_mol_nb;
        a.P = sb.Pn;
        a.P32 = sb.P32;
        a.PK = sb.PKa;
        a.PK2 = sb.PKb;
        a.Q = c->d_Q;
        a.A = sb.A;
        a.B = sb.B;
        a.B2 = sb.B2;
        a.tol = c->tol;
        a.nwaves = w.nwaves;
        a.remap = w.remap;
        a.seg = w.d_seg;
        a.tblk = c->d_tblk;
        a.nseg = w.nseg;
        a.kcols = w.kcols;
        const bool uhf = sb.B2 != nullptr;
        if (events) {
            if (!w.e0) PK_HIP(hipEventCreate(&w.e0));
            if (!w.e1) PK_HIP(hipEventCreate(&w.e1));
            PK_HIP(hipEventRecord(w.e0, s));
        }
        const int before = nl;
        const ClassFns &f = kClasses[w.cls];
        const bool fp32 = w.prec == PK_PREC_MIXED && !w.alt;
        // segment mode when the exchange rows of the call fit the LDS budget
        // (two rows sets for the one-pass UHF OS kernels); the low-L OS classes
        // compiled both ways take it for lists of few primitive quartets per
        // shell quartet (atomic-bound), the plain kernel otherwise
        const bool seg_alt = f.seg_alt && use_os(c, f) && (double)w.nprimq <= c->seg_alt_prim * (double)w.nquart;
        const int nk = uhf && use_os(c, f) ? 2 : 1;
        {
            const size_t rows = (size_t)nk * (pk::ncart(f.la) + pk::ncart(f.lb)) * w.kcols +
                                (size_t)pk::ncart(f.la) * pk::ncart(f.lb);
            a.segmode = c->seg_on && (f.seg || seg_alt) && w.nseg > 0 && rows * sizeof(double) <= kSegLdsMax;
        }
        const bool alt = seg_alt && a.segmode;  // the segment-mode variant of a plain OS class
        // lanes per workgroup: the autotuned (or PK_BLOCK) shape, else one wave
        // per workgroup for the segment-mode lists of a batch of molecules whose
        // exchange rows leave room for at least 6 such workgroups per CU, and
        // kDefaultBlock otherwise.  Measured per class list: on the 256 water
        // clusters (rows of 2-19 KB) every list is faster with one wave (batched
        // build 356 -> 256 ms); on n-C10H22 / cc-pVTZ (one molecule, segments of
        // many kets) most lists are faster with four ((pp|ss) 7.4 -> 3.5 ms,
        // (dd|dp) 11.4 -> 7.4 ms; a few lose up to 0.8 ms)
        const bool one_wave =
            a.segmode && c->nmol > 1 && pk::jk_lds_bytes(a, f.la, f.lb, nk) <= kSegOneWaveLds;
        const int block = w.block > 0 ? w.block : (one_wave ? 64 : kDefaultBlock);
        w.block_run = block;
        const int wpb = block / 64;
        int grid = a.segmode ? w.nseg : (w.nwaves + wpb - 1) / wpb;
        if (w.remap) grid = (grid + 7) / 8 * 8;
        if (uhf && use_os(c, f)) {
            (alt ? f.jk_os_uhf_seg : f.jk_os_uhf)(a, grid, block, s, nl);
        } else if (uhf && f.has_tmpl) {
            // per-lane Rys kernels once per spin density: J of each into A
            // (summing to J of Pa + Pb), exchange into B and B2
            pk::JKArgs x = a;
            x.P = sb.PKa;
            f.jk_rys(x, grid, block, s, nl);
            x.P = sb.PKb;
            x.B = sb.B2;
            f.jk_rys(x, grid, block, s, nl);
        } else if (uhf) {
            // handled by the cooperative kernel below
        } else if (use_os(c, f)) {
            if (alt)
                (fp32 ? f.jk_os32_seg : f.jk_os_seg)(a, grid, block, s, nl);
            else
                (fp32 ? f.jk_os32 : f.jk_os)(a, grid, block, s, nl);
        } else if (f.has_tmpl && !(w.alt && w.prec == PK_PREC_FP64)) {
            (fp32 ? f.jk_rys32 : f.jk_rys)(a, grid, block, s, nl);
        }
        if ((uhf && !use_os(c, f) && !f.has_tmpl) ||
            (!uhf && !use_os(c, f) && !(f.has_tmpl && !(w.alt && w.prec == PK_PREC_FP64)))) {
            pk::GenArgs g{};
            g.wave_bra = a.wave_bra;
            g.wave_start = a.wave_start;
            g.wave_count = a.wave_count;
            g.wave_mol = a.wave_mol;
            g.kets = a.kets;
            g.mol_off = a.mol_off;
            g.mol_nb = a.mol_nb;
            g.pairs = a.pairs;
            g.prims = a.prims;
            g.hz = a.hz;
            g.P = a.P;
            g.PK = uhf ? a.PK : a.P;
            g.PK2 = a.PK2;
            g.Q = a.Q;
            g.A = a.A;
            g.B = a.B;
            g.B2 = a.B2;
            g.tol = a.tol;
            g.nitems = w.nwaves;
            g.la = f.la;
            g.lb = f.lb;
            g.lc = f.lc;
            g.ld = f.ld;
            if ((rc = launch_generic(g, s)) != PK_OK) return rc;
            ++nl;
        }
        w.kernels = nl - before;
        if (events) PK_HIP(hipEventRecord(w.e1, s));
    }
    return PK_OK;
}

int launch_build(pk_ctx *c, const double *dP, double *dJ, double *dK, double *dG, const BuildBufs &sb,
                 bool events) {
    int rc;
    hipStream_t s = c->stream;
    const MolTab t{c->d_mol_off, c->d_mol_nb, c->d_mol_foff, c->d_fpos};
    const dim3 g2((unsigned)(((size_t)c->nbmax * c->nbmax + 255) / 256), (unsigned)c->nmol);
    const int gf = (int)((c->n2 + 255) / 256);
    hipLaunchKernelGGL(k_scale_density, g2, dim3(256), 0, s, dP, c->d_fnorm, sb.Pn, t);
    hipLaunchKernelGGL(k_fill, dim3(gf), dim3(256), 0, s, sb.A, c->n2, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(gf), dim3(256), 0, s, sb.B, c->n2, 0.0);
    int nl = 3;
    if (c->precision == PK_PREC_MIXED) {
        if (!sb.P32) return fail(PK_ERR_STATE, "mixed precision build without an FP32 density buffer");
        hipLaunchKernelGGL(k_to_f32, dim3(gf), dim3(256), 0, s, sb.Pn, sb.P32, c->n2);
        ++nl;
    }
    if ((rc = launch_classes(c, sb, events, nl)) != PK_OK) return rc;
    hipLaunchKernelGGL(k_finalize, g2, dim3(256), 0, s, sb.A, sb.B, c->d_fnorm, dJ, dK, dG, t);
    ++nl;
    PK_HIP(hipGetLastError());
    c->st.kernel_launches = nl;
    return PK_OK;
}

// One-pass UHF build: Pn = n n^T o (Pa + Pb) for J, PKa, PKb the scaled spin
// densities for the two exchange matrices; each integral computed once.
int launch_build_uhf(pk_ctx *c, const double *dPa, const double *dPb, double *dJt, double *dKa, double *dKb,
                     const BuildBufs &sb, bool events) {
    int rc;
    hipStream_t s = c->stream;
    const MolTab t{c->d_mol_off, c->d_mol_nb, c->d_mol_foff, c->d_fpos};
    const dim3 g2((unsigned)(((size_t)c->nbmax * c->nbmax + 255) / 256), (unsigned)c->nmol);
    const int gf = (int)((c->n2 + 255) / 256);
    hipLaunchKernelGGL(k_scale_density, g2, dim3(256), 0, s, dPa, c->d_fnorm, sb.PKa, t);
    hipLaunchKernelGGL(k_scale_density, g2, dim3(256), 0, s, dPb, c->d_fnorm, sb.PKb, t);
    hipLaunchKernelGGL(k_sum2, dim3(gf), dim3(256), 0, s, sb.PKa, sb.PKb, sb.Pn, c->n2);
    hipLaunchKernelGGL(k_fill, dim3(gf), dim3(256), 0, s, sb.A, c->n2, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(gf), dim3(256), 0, s, sb.B, c->n2, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(gf), dim3(256), 0, s, sb.B2, c->n2, 0.0);
    int nl = 6;
    if ((rc = launch_classes(c, sb, events, nl)) != PK_OK) return rc;
    hipLaunchKernelGGL(k_finalize, g2, dim3(256), 0, s, sb.A, sb.B, c->d_fnorm, dJt, dKa, (double *)nullptr, t);
    hipLaunchKernelGGL(k_finalize, g2, dim3(256), 0, s, sb.A, sb.B2, c->d_fnorm, (double *)nullptr, dKb,
                       (double *)nullptr, t);
    nl += 2;
    PK_HIP(hipGetLastError());
    c->st.kernel_launches = nl;
    return PK_OK;
}

// One build.  With graphs enabled (and profiling off) the kernels of a build
// are captured once into a hipGraph that reads the context's own density
// buffer and writes its own J, K, G buffers, and every later build replays
// it; the caller's buffers are copied in and out around the replay.
int build_device(pk_ctx *c, const double *dP, double *dJ, double *dK, double *dG) {
    int rc;
    if (!c->work_ok && (rc = build_work(c)) != PK_OK) return rc;
    if ((rc = ensure_buffers(c)) != PK_OK) return rc;
    hipStream_t s = c->stream;
    const size_t bytes = c->n2 * sizeof(double);
    // the first build that would use the side queues times every list (one
    // at a time), so that later builds balance the queues by measured cost
    c->learning = !c->cost_learned && !c->profile && c->side_ok && !c->side_off &&
                  c->st.quartets >= kSideMinQuartets && c->work.size() > 1;
    if (!c->graph_on || c->profile || c->learning) {
        PK_HIP(hipEventRecord(c->ev0, s));
        if ((rc = launch_build(c, dP, dJ, dK, dG, ctx_bufs(c), c->profile || c->learning)) != PK_OK) return rc;
        PK_HIP(hipEventRecord(c->ev1, s));
        return PK_OK;
    }
    if (!c->graph_ok) {
        drop_graph(c);
        PK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int rc_cap = launch_build(c, c->d_P, c->d_J, c->d_K, c->d_G, ctx_bufs(c), false);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (rc_cap != PK_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc_cap;
        }
        PK_HIP(e);
        c->graph = g;
        PK_HIP(hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0));
        c->graph_ok = true;
    }
    PK_HIP(hipEventRecord(c->ev0, s));
    if (dP != c->d_P) PK_HIP(hipMemcpyAsync(c->d_P, dP, bytes, hipMemcpyDeviceToDevice, s));
    PK_HIP(hipGraphLaunch(c->gexec, s));
    if (dJ && dJ != c->d_J) PK_HIP(hipMemcpyAsync(dJ, c->d_J, bytes, hipMemcpyDeviceToDevice, s));
    if (dK && dK != c->d_K) PK_HIP(hipMemcpyAsync(dK, c->d_K, bytes, hipMemcpyDeviceToDevice, s));
    if (dG && dG != c->d_G) PK_HIP(hipMemcpyAsync(dG, c->d_G, bytes, hipMemcpyDeviceToDevice, s));
    PK_HIP(hipEventRecord(c->ev1, s));
    return PK_OK;
}

int finish_timing(pk_ctx *c) {
    PK_HIP(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    PK_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->st.last_build_ms = ms;
    c->st.last_eri_kernel_ms = ms;
    if (c->profile || c->learning) {
        double eri = 0.0;
        for (ClassWork &w : c->work) {
            float m = 0.f;
            PK_HIP(hipEventElapsedTime(&m, w.e0, w.e1));
            if (c->profile) {
                w.ms_total += m;
                w.builds += 1;
            }
            w.cost_ms = m;
            eri += m;
        }
        c->st.last_eri_kernel_ms = eri;
        c->cost_learned = true;
        c->learning = false;
    }
    return PK_OK;
}

}  // namespace

int pk_build_with_scratch(pk_ctx *c, const double *P, double *J, double *K, double *scratch, size_t each,
                          const char *caller) {
    const std::string who = caller ? caller : "pk_build_with_scratch";
    if (!c || !P || !scratch) return fail(PK_ERR_ARG, who + ": null context, density or scratch");
    if (c->n2 > each)
        return fail(PK_ERR_ARG, who + ": packed matrices of " + std::to_string(c->n2) +
                                    " elements exceed the scratch limit of " + std::to_string(each) + " elements");
    if (!c->work_ok)
        return fail(PK_ERR_STATE, who + ": the work lists of the current settings are not built "
                                        "(call pk_fock_jk once after pk_ctx_create / pk_ctx_set)");
    BuildBufs sb{scratch, scratch + each, scratch + 2 * each};
    sb.P32 = reinterpret_cast<float *>(scratch + 3 * each);
    return launch_build(c, P, J, K, nullptr, sb, false);
}

int pk_set_error(int code, const char *msg) { return fail(code, msg ? msg : ""); }

// ====================================================================== C ABI
extern "C" {

const char *pk_last_error(void) { return g_err.c_str(); }

}  // extern "C"

namespace {

// Ordering group of a pair with n primitive pairs: floor(log_1.5 n).  Kets of
// one group run the primitive loop for up to 1.5 times as many iterations as
// the group's smallest member (idle lanes), in exchange for longer runs of
// neighbouring second shells in a wave (fewer cache lines per atomic add).
// Measured on n-C20H42 / cc-pVDZ (ms per build): exact primitive counts 60.9,
// ratio 1.5 58.3; an earlier sweep gave 62.2 / 61.1 / 61.0 / 61.3 / 63.4 for
// exact / 1.25 / 1.5 / 2 / 3.
constexpr double kGroupRatio = 1.5;
int prim_group(int n) { return (int)std::floor(std::log((double)n) / std::log(kGroupRatio) + 1e-9); }

// Shells, component norms and shell pairs of molecule m (pairs within the
// molecule only), appended to the context's arrays.
int add_molecule(pk_ctx *c, const pk_basis_desc *d, int m) {
    if (!d || d->nshells <= 0 || !d->center || !d->l || !d->nprim || !d->exponent || !d->coef || !d->prim_norm)
        return fail(PK_ERR_ARG, "pk_ctx_create: null or empty basis description (molecule " + std::to_string(m) + ")");
    const int s0 = c->ns;
    int off = (int)c->ex.size(), f0 = 0;
    const int poff0 = off;
    for (int s = 0; s < d->nshells; ++s) {
        const int l = d->l[s];
        if (l < 0 || l > kMaxL)
            return fail(PK_ERR_UNSUPPORTED, "pk_ctx_create: shell " + std::to_string(s) + " has l=" + std::to_string(l) +
                                                "; this build supports l <= " + std::to_string(kMaxL));
        if (d->nprim[s] <= 0) return fail(PK_ERR_ARG, "pk_ctx_create: shell with no primitives");
    }
    // internal order: shells grouped by l, basis order within a group
    std::vector<int> spos(d->nshells), snat(d->nshells);
    {
        int pos = 0;
        for (int l = 0; l <= kMaxL; ++l) {
            c->tblk.push_back(pos);
            for (int s = 0; s < d->nshells; ++s)
                if (d->l[s] == l) {
                    spos[s] = pos;
                    pos += pk::ncart(l);
                }
        }
        c->tblk.push_back(pos);
        for (int s = 0, f = 0; s < d->nshells; ++s) {
            snat[s] = f;
            f += pk::ncart(d->l[s]);
        }
        const size_t fb = c->fnorm.size();
        c->fnorm.resize(fb + pos);
        c->fnat.resize(fb + pos);
        c->fpos.resize(fb + pos);
        for (int s = 0; s < d->nshells; ++s)
            for (int k = 0; k < pk::ncart(d->l[s]); ++k) {
                c->fpos[fb + snat[s] + k] = spos[s] + k;
                c->fnat[fb + spos[s] + k] = snat[s] + k;
            }
    }
    const size_t fbase = (size_t)c->nb;  // this molecule's first entry in fnorm / fpos / fnat
    for (int s = 0; s < d->nshells; ++s) {
        const int l = d->l[s], np = d->nprim[s];
        c->sh_l.push_back(l);
        c->sh_np.push_back(np);
        c->sh_poff.push_back(off);
        c->sh_f0.push_back(spos[s]);
        for (int q = 0; q < 3; ++q) c->sh_c.push_back(d->center[3 * s + q]);
        for (int k = 0; k < np; ++k) {
            c->ex.push_back(d->exponent[off - poff0 + k]);
            c->co.push_back(d->coef[off - poff0 + k]);
            c->pn.push_back(d->prim_norm[off - poff0 + k]);
        }
        // component norms 1/sqrt((2lx-1)!!(2ly-1)!!(2lz-1)!!)  (gaussian.cpp:286-298)
        for (int k = 0; k < pk::ncart(l); ++k) {
            auto df = [](int n) {
                int r = 1;
                while (n > 0) {
                    r *= n;
                    n -= 2;
                }
                return r;
            };
            const int x = pk::cart_x(l, k), y = pk::cart_y(l, k), z = pk::cart_z(l, k);
            c->fnorm[fbase + spos[s] + k] = 1.0 / std::sqrt((double)(df(2 * x - 1) * df(2 * y - 1) * df(2 * z - 1)));
        }
        off += np;
        f0 += pk::ncart(l);
    }
    // the kernels pack (pair primitive count, first function) into one int
    if (f0 >= (1 << 20)) return fail(PK_ERR_ARG, "pk_ctx_create: more than 2^20 functions in one molecule");
    for (int s = s0; s < s0 + d->nshells; ++s)
        if (c->sh_np[s] * c->sh_np[s] >= 2048) return fail(PK_ERR_ARG, "pk_ctx_create: shell with more than 45 primitives");
    c->ns += d->nshells;
    c->mol_foff.push_back(c->nb);
    c->mol_nb.push_back(f0);
    c->mol_off.push_back((long long)c->n2);
    c->nb += f0;
    c->n2 += (size_t)f0 * f0;
    c->nbmax = std::max(c->nbmax, f0);
    // shell pairs a >= b, first shell carries the larger l (types.h:606-675 pair data)
    const size_t pair0 = c->pairs.size();
    std::vector<std::vector<pk::PrimDev>> pp_all;
    std::vector<std::vector<double>> pa_all;
    double kmax = 0.0;  // max over primitive pairs of |kc| / sqrt(zeta)
    for (int A = s0; A < c->ns; ++A)
        for (int Bs = s0; Bs <= A; ++Bs) {
            const int sa = c->sh_l[A] >= c->sh_l[Bs] ? A : Bs, sb = sa == A ? Bs : A;
            pk::PairDev p{};
            p.f0 = c->sh_f0[sa];
            p.f1 = c->sh_f0[sb];
            p.diag = A == Bs;
            const double *Ac = &c->sh_c[3 * sa], *Bc = &c->sh_c[3 * sb];
            for (int q = 0; q < 3; ++q) p.ab[q] = Ac[q] - Bc[q];
            const double R2 = p.ab[0] * p.ab[0] + p.ab[1] * p.ab[1] + p.ab[2] * p.ab[2];
            std::vector<pk::PrimDev> pps;
            std::vector<double> pal;
            for (int i = 0; i < c->sh_np[sa]; ++i)
                for (int j = 0; j < c->sh_np[sb]; ++j) {
                    const double al = c->ex[c->sh_poff[sa] + i], be = c->ex[c->sh_poff[sb] + j];
                    const double ca = c->co[c->sh_poff[sa] + i] * c->pn[c->sh_poff[sa] + i];
                    const double cb = c->co[c->sh_poff[sb] + j] * c->pn[c->sh_poff[sb] + j];
                    pk::PrimDev pp{};
                    pp.zeta = al + be;
                    const double iz = 1.0 / pp.zeta;
                    // K_ab sqrt(zeta) (4/pi)^(1/4): two records and 1/sqrt(zeta+eta)
                    // give K_ab K_cd 2 sqrt(rho/pi)
                    pp.kc = std::pow(M_PI * iz, 1.5) * std::exp(-(al * be * iz) * R2) * ca * cb *
                            std::sqrt(pp.zeta) * 1.0622519320271969145;
                    for (int q = 0; q < 3; ++q) {
                        pp.p[q] = (al * Ac[q] + be * Bc[q]) * iz;
                        pp.pa[q] = pp.p[q] - Ac[q];
                    }
                    kmax = std::max(kmax, std::fabs(pp.kc) / std::sqrt(pp.zeta));
                    pps.push_back(pp);
                    pal.push_back(al);
                }
            pp_all.push_back(std::move(pps));
            pa_all.push_back(std::move(pal));
            c->pair_sh.push_back(sa);
            c->pair_sh.push_back(sb);
            c->pairs.push_back(p);
            c->pair_cls.push_back(pk::pair_class(c->sh_l[sa], c->sh_l[sb]));
            c->pair_q.push_back(0.0);
            c->pair_mol.push_back(m);
        }
    // Primitive screening: a primitive pair whose prefactor bound
    // |kc_p| max_q |kc_q| / sqrt(eta_q) (>= |kc_p kc_q| / sqrt(zeta_p + eta_q))
    // is below kPrimCut cannot move any integral by more than that times its
    // angular factors; it is dropped (distant tight primitives: e^{-ab R^2/z}).
    for (size_t k = 0; k < pp_all.size(); ++k) {
        pk::PairDev &p = c->pairs[pair0 + k];
        p.prim_off = (int)c->prims.size();
        for (size_t i = 0; i < pp_all[k].size(); ++i)
            if (std::fabs(pp_all[k][i].kc) * kmax >= kPrimCut) {
                c->prims.push_back(pp_all[k][i]);
                c->palpha.push_back(pa_all[k][i]);
                c->hz.push_back(0.5 / pp_all[k][i].zeta);
            }
        p.nprim = (int)c->prims.size() - p.prim_off;
        p.grp = prim_group(p.nprim);
    }
    return PK_OK;
}

int ctx_build(const pk_basis_desc *descs, int nmol, pk_ctx **out) {
    if (!descs || !out || nmol <= 0) return fail(PK_ERR_ARG, "pk_ctx_create: null or empty basis description");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(PK_ERR_HIP, "pk_ctx_create: no HIP device visible");
    pk_ctx *c = new pk_ctx;
    c->nmol = nmol;
    for (int m = 0; m < nmol; ++m) {
        const int rc = add_molecule(c, &descs[m], m);
        if (rc != PK_OK) {
            delete c;
            return rc;
        }
    }
    if (c->prims.size() > (size_t)INT32_MAX || c->n2 > (size_t)1 << 40) {
        delete c;
        return fail(PK_ERR_ARG, "pk_ctx_create: batch too large");
    }
    c->st.nbasis = c->nb;
    c->st.nshells = c->ns;
    // diagnostic switch: PK_EXCHANGE_ROWS=0 runs the per-lane kernels without
    // the LDS exchange rows (one global atomic per exchange contribution)
    if (const char *e = std::getenv("PK_EXCHANGE_ROWS")) c->seg_on = std::atoi(e) != 0;
    // diagnostic switch: PK_SIDE_QUEUES=0 keeps every class kernel of a build
    // on the context stream (one kernel at a time, as a kernel trace times them)
    if (const char *e = std::getenv("PK_SIDE_QUEUES")) c->side_off = std::atoi(e) == 0;
    // diagnostic knob: PK_SEG_ALT_PRIM overrides the primitive quartets per
    // shell quartet below which the low-L OS lists run in segment mode
    if (const char *e = std::getenv("PK_SEG_ALT_PRIM")) c->seg_alt_prim = std::atof(e);
    // diagnostic knob: PK_BLOCK = 64, 128 or 256 lanes per workgroup of every
    // class list until pk_ctx_autotune chooses per list
    if (const char *e = std::getenv("PK_BLOCK")) {
        const int b = std::atoi(e);
        if (b == 64 || b == 128 || b == 256) c->block0 = b, c->block_forced = true;
    }
    c->st.shell_pairs = (int64_t)c->pairs.size();
    int rc = PK_OK;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        pk_ctx_destroy(c);
        return fail(PK_ERR_HIP, "pk_ctx_create: stream/event creation failed");
    }
    c->own_stream = true;
    if ((rc = dev_upload(&c->d_mol_off, c->mol_off, c->stream)) != PK_OK ||
        (rc = dev_upload(&c->d_mol_nb, c->mol_nb, c->stream)) != PK_OK ||
        (rc = dev_upload(&c->d_mol_foff, c->mol_foff, c->stream)) != PK_OK ||
        (rc = dev_upload(&c->d_tblk, c->tblk, c->stream)) != PK_OK ||
        (rc = dev_upload(&c->d_pairs, c->pairs, c->stream)) != PK_OK ||
        (rc = dev_upload(&c->d_prims, c->prims, c->stream)) != PK_OK ||
        (rc = dev_upload(&c->d_palpha, c->palpha, c->stream)) != PK_OK ||
        (rc = dev_upload(&c->d_hz, c->hz, c->stream)) != PK_OK ||
        (rc = dev_upload(&c->d_pair_sh, c->pair_sh, c->stream)) != PK_OK ||
        (rc = dev_upload(&c->d_fnorm, c->fnorm, c->stream)) != PK_OK ||
        (rc = dev_upload(&c->d_fpos, c->fpos, c->stream)) != PK_OK || (rc = build_schwarz(c)) != PK_OK ||
        (rc = dev_upload(&c->d_Q, c->Q, c->stream)) != PK_OK ||
        hipMemcpyAsync(c->d_pairs, c->pairs.data(), c->pairs.size() * sizeof(pk::PairDev), hipMemcpyHostToDevice,
                       c->stream) != hipSuccess) {
        if (rc == PK_OK) rc = fail(PK_ERR_HIP, "pk_ctx_create: pair upload failed");
        const std::string msg = g_err;
        pk_ctx_destroy(c);
        return fail(rc, msg);
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        pk_ctx_destroy(c);
        return fail(PK_ERR_HIP, "pk_ctx_create: device error during setup");
    }
    *out = c;
    return PK_OK;
}

}  // namespace

extern "C" {

int pk_ctx_create(const pk_basis_desc *d, pk_ctx **out) { return ctx_build(d, 1, out); }

int pk_batch_create(const pk_basis_desc *descs, int nmol, pk_ctx **out) { return ctx_build(descs, nmol, out); }

int pk_ctx_nmol(const pk_ctx *c) { return c ? c->nmol : -1; }

int pk_ctx_mol_layout(const pk_ctx *c, int32_t *nb, int64_t *offset) {
    if (!c) return fail(PK_ERR_ARG, "pk_ctx_mol_layout: null context");
    for (int m = 0; m < c->nmol; ++m) {
        if (nb) nb[m] = c->mol_nb[m];
        if (offset) offset[m] = c->mol_off[m];
    }
    return PK_OK;
}

// Launch-shape autotuner: every class list is timed (HIP events, `reps` builds
// per candidate after one untimed build) under each candidate shape --
// workgroups of 1, 2 or 4 waves, with and without XCD-contiguous wave ranges
// -- and keeps the fastest; the choice persists for the context and is
// re-applied when the work lists are rebuilt.
int pk_ctx_autotune(pk_ctx *c, int reps) {
    if (!c) return fail(PK_ERR_ARG, "pk_ctx_autotune: null context");
    if (reps < 1) reps = 1;
    int rc;
    if (!c->work_ok && (rc = build_work(c)) != PK_OK) return rc;
    if ((rc = ensure_buffers(c)) != PK_OK) return rc;
    hipStream_t s = c->stream;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((c->n2 + 255) / 256)), dim3(256), 0, s, c->d_P, c->n2, 0.01);
    const bool prof = c->profile;
    c->profile = true;
    // launch shapes (lanes per workgroup, XCD remap) of the class kernels, and
    // the alternative kernel of the list (see ClassWork::alt) in its own shape
    constexpr int kCand = 7;
    const int cand[kCand][3] = {{64, 0, 0}, {128, 0, 0}, {256, 0, 0}, {64, 1, 0}, {128, 1, 0}, {256, 1, 0}, {256, 0, 1}};
    const size_t nw = c->work.size();
    std::vector<double> best(nw, 1e300);
    std::vector<int> pick(nw, 2);
    auto run = [&](int k, int nrep, bool keep) -> int {
        for (ClassWork &w : c->work) {
            w.block = cand[k][0];
            w.remap = cand[k][1];
            w.alt = cand[k][2];
        }
        std::vector<double> t(nw, 0.0);
        for (int r = 0; r < nrep; ++r) {
            int rc2;
            if ((rc2 = launch_build(c, c->d_P, c->d_J, c->d_K, nullptr, ctx_bufs(c), true)) != PK_OK) return rc2;
            PK_HIP(hipStreamSynchronize(s));
            for (size_t i = 0; i < nw; ++i) {
                float ms = 0.f;
                PK_HIP(hipEventElapsedTime(&ms, c->work[i].e0, c->work[i].e1));
                t[i] += ms;
            }
        }
        if (keep)
            for (size_t i = 0; i < nw; ++i)
                if (t[i] < best[i]) {
                    best[i] = t[i];
                    pick[i] = k;
                }
        return PK_OK;
    };
    rc = run(2, 1, false);
    for (int k = 0; k < kCand && rc == PK_OK; ++k) rc = run(k, reps, true);
    c->profile = prof;
    if (rc != PK_OK) return rc;
    c->tuned.clear();
    for (size_t i = 0; i < nw; ++i) {
        ClassWork &w = c->work[i];
        w.block = cand[pick[i]][0];
        w.remap = cand[pick[i]][1];
        w.alt = cand[pick[i]][2];
        w.ms_total = 0.0;
        w.builds = 0;
        c->tuned.push_back(TuneEntry{w.cls, w.prec, w.block, w.remap, w.alt});
    }
    drop_graph(c);
    return PK_OK;
}

int pk_ctx_set_graph(pk_ctx *c, int enable) {
    if (!c) return fail(PK_ERR_ARG, "pk_ctx_set_graph: null context");
    c->graph_on = enable != 0;
    if (!c->graph_on) drop_graph(c);
    return PK_OK;
}

int pk_ctx_destroy(pk_ctx *c) {
    if (!c) return PK_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_work(c);
    for (void *p : {(void *)c->d_pairs, (void *)c->d_prims, (void *)c->d_Q, (void *)c->d_fnorm, (void *)c->d_P,
                    (void *)c->d_Pn, (void *)c->d_A, (void *)c->d_B, (void *)c->d_J, (void *)c->d_K, (void *)c->d_G, (void *)c->d_J2,
                    (void *)c->d_palpha, (void *)c->d_hz, (void *)c->d_B2, (void *)c->d_P32, (void *)c->d_Pk1n, (void *)c->d_Pk2n, (void *)c->d_gsh, (void *)c->d_pair_sh,
                    (void *)c->d_mol_off, (void *)c->d_mol_nb, (void *)c->d_mol_foff, (void *)c->d_tblk, (void *)c->d_fpos})
        if (p) (void)hipFree(p);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
    for (int i = 0; i < kSideQueues; ++i) {
        if (c->join_ev[i]) (void)hipEventDestroy(c->join_ev[i]);
        if (c->side[i]) (void)hipStreamDestroy(c->side[i]);
    }
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return PK_OK;
}

int pk_ctx_set(pk_ctx *c, int engine, double tol, int precision, double cutoff) {
    if (!c) return fail(PK_ERR_ARG, "pk_ctx_set: null context");
    if (engine < PK_ENGINE_OS || engine > PK_ENGINE_AUTO) return fail(PK_ERR_ARG, "pk_ctx_set: unknown engine");
    if (!(tol >= 0.0)) return fail(PK_ERR_ARG, "pk_ctx_set: tol_eri must be >= 0");
    if (precision != PK_PREC_FP64 && precision != PK_PREC_MIXED)
        return fail(PK_ERR_ARG, "pk_ctx_set: unknown precision");
    if (precision == PK_PREC_MIXED && !(cutoff >= 0.0))
        return fail(PK_ERR_ARG, "pk_ctx_set: the FP32 cutoff must be >= 0");
    if (precision == PK_PREC_FP64) cutoff = 0.0;
    // the kernel of each class (OS or Rys) is chosen at launch time, so a
    // captured graph holds the previous engine's kernels
    if (engine != c->engine) drop_graph(c);
    c->engine = engine;
    if (tol != c->tol || precision != c->precision || cutoff != c->cutoff) {
        c->tol = tol;
        c->precision = precision;
        c->cutoff = cutoff;
        c->work_ok = false;
    }
    return PK_OK;
}

int pk_ctx_set_stream(pk_ctx *c, void *stream) {
    if (!c) return fail(PK_ERR_ARG, "pk_ctx_set_stream: null context");
    (void)hipStreamSynchronize(c->stream);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    if (stream) {
        c->stream = (hipStream_t)stream;
        c->own_stream = false;
    } else {
        PK_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->own_stream = true;
    }
    return PK_OK;
}

int pk_ctx_nbasis(const pk_ctx *c) { return c ? c->nb : -1; }

int pk_ctx_profile(pk_ctx *c, int enable) {
    if (!c) return fail(PK_ERR_ARG, "pk_ctx_profile: null context");
    c->profile = enable != 0;
    if (c->profile)
        for (ClassWork &w : c->work) {
            w.ms_total = 0.0;
            w.builds = 0;
        }
    return PK_OK;
}

int pk_ctx_class_count(const pk_ctx *c) { return c ? (int)c->work.size() : -1; }

int pk_ctx_class_stats(const pk_ctx *c, int i, pk_class_stats *o) {
    if (!c || !o || i < 0 || i >= (int)c->work.size()) return fail(PK_ERR_ARG, "pk_ctx_class_stats: bad index");
    const ClassWork &w = c->work[i];
    const ClassFns &f = kClasses[w.cls];
    o->la = f.la;
    o->lb = f.lb;
    o->lc = f.lc;
    o->ld = f.ld;
    o->kernels = w.kernels;
    o->quartets = w.nquart;
    o->primitive_quartets = w.nprimq;
    o->model_flops = w.flops;
    o->ms_total = w.ms_total;
    o->builds = w.builds;
    o->precision = w.prec;
    o->block = w.block > 0 ? w.block : w.block_run;
    o->remap = w.remap;
    o->alt = w.alt;
    o->kind = list_kind(c, w);
    o->waves = w.nwaves;
    return PK_OK;
}

int pk_ctx_stats(const pk_ctx *c, pk_stats *out) {
    if (!c || !out) return fail(PK_ERR_ARG, "pk_ctx_stats: null argument");
    *out = c->st;
    return PK_OK;
}

static int fock_common(pk_ctx *c, const double *P, double *J, double *K, double *G, int memkind) {
    if (!c || !P) return fail(PK_ERR_ARG, "pk_fock: null context or density");
    if (memkind != PK_MEM_HOST && memkind != PK_MEM_DEVICE) return fail(PK_ERR_ARG, "pk_fock: bad memkind");
    int rc;
    if ((rc = ensure_buffers(c)) != PK_OK) return rc;
    const size_t n2 = c->n2 * sizeof(double);
    if (memkind == PK_MEM_DEVICE) {
        if ((rc = build_device(c, P, J, K, G)) != PK_OK) return rc;
        return finish_timing(c);
    }
    PK_HIP(hipMemcpyAsync(c->d_P, P, n2, hipMemcpyHostToDevice, c->stream));
    double *dJ = J ? c->d_J : nullptr, *dK = K ? c->d_K : nullptr, *dG = G ? c->d_G : nullptr;
    if ((rc = build_device(c, c->d_P, dJ, dK, dG)) != PK_OK) return rc;
    if (J) PK_HIP(hipMemcpyAsync(J, c->d_J, n2, hipMemcpyDeviceToHost, c->stream));
    if (K) PK_HIP(hipMemcpyAsync(K, c->d_K, n2, hipMemcpyDeviceToHost, c->stream));
    if (G) PK_HIP(hipMemcpyAsync(G, c->d_G, n2, hipMemcpyDeviceToHost, c->stream));
    PK_HIP(hipStreamSynchronize(c->stream));
    return finish_timing(c);
}

int pk_fock_jk(pk_ctx *c, const double *P, double *J, double *K, int memkind) {
    return fock_common(c, P, J, K, nullptr, memkind);
}

int pk_fock_2e(pk_ctx *c, const double *P, double *G, int memkind) {
    if (!G) return fail(PK_ERR_ARG, "pk_fock_2e: null G");
    return fock_common(c, P, nullptr, nullptr, G, memkind);
}

int pk_fock_jk_uhf(pk_ctx *c, const double *Pa, const double *Pb, double *Jt, double *Ka, double *Kb, int memkind) {
    if (!c || !Pa || !Pb) return fail(PK_ERR_ARG, "pk_fock_jk_uhf: null argument");
    if (memkind != PK_MEM_HOST && memkind != PK_MEM_DEVICE) return fail(PK_ERR_ARG, "pk_fock_jk_uhf: bad memkind");
    int rc;
    if (!c->work_ok && (rc = build_work(c)) != PK_OK) return rc;
    if ((rc = ensure_buffers(c)) != PK_OK) return rc;
    const size_t n2 = c->n2 * sizeof(double);
    if (!c->d_J2) PK_HIP(hipMalloc((void **)&c->d_J2, n2));
    if (!c->d_Pk1n) PK_HIP(hipMalloc((void **)&c->d_Pk1n, n2));
    if (!c->d_Pk2n) PK_HIP(hipMalloc((void **)&c->d_Pk2n, n2));
    if (!c->d_B2) PK_HIP(hipMalloc((void **)&c->d_B2, n2));
    hipStream_t s = c->stream;
    // one pass over the integrals: J of Pa + Pb, the exchange of each spin
    // density from the same integrals (two exchange accumulators)
    BuildBufs sb = ctx_bufs(c);
    sb.PKa = c->d_Pk1n;
    sb.PKb = c->d_Pk2n;
    sb.B2 = c->d_B2;
    PK_HIP(hipEventRecord(c->ev0, s));
    if (memkind == PK_MEM_DEVICE) {
        if ((rc = launch_build_uhf(c, Pa, Pb, Jt, Ka, Kb, sb, c->profile)) != PK_OK) return rc;
        PK_HIP(hipEventRecord(c->ev1, s));
        PK_HIP(hipStreamSynchronize(s));
        return finish_timing(c);
    }
    // host memory: staged through the context's buffers
    PK_HIP(hipMemcpyAsync(c->d_P, Pa, n2, hipMemcpyHostToDevice, s));
    PK_HIP(hipMemcpyAsync(c->d_J2, Pb, n2, hipMemcpyHostToDevice, s));
    if ((rc = launch_build_uhf(c, c->d_P, c->d_J2, c->d_J, c->d_K, c->d_G, sb, c->profile)) != PK_OK) return rc;
    PK_HIP(hipEventRecord(c->ev1, s));
    if (Jt) PK_HIP(hipMemcpyAsync(Jt, c->d_J, n2, hipMemcpyDeviceToHost, s));
    if (Ka) PK_HIP(hipMemcpyAsync(Ka, c->d_K, n2, hipMemcpyDeviceToHost, s));
    if (Kb) PK_HIP(hipMemcpyAsync(Kb, c->d_G, n2, hipMemcpyDeviceToHost, s));
    PK_HIP(hipStreamSynchronize(s));
    return finish_timing(c);
}

// Two-electron gradient: the densities are scaled into the internal order,
// then every class list of the energy build (same screening) runs through the
// derivative kernel; the per-shell sums leave through the context's buffer.
int pk_eri_gradient(pk_ctx *c, const double *Pj, const double *Pk1, const double *Pk2, double kx, double *grad,
                    int memkind) {
    if (!c || !Pj || !Pk1 || !grad) return fail(PK_ERR_ARG, "pk_eri_gradient: null argument");
    if (memkind != PK_MEM_HOST && memkind != PK_MEM_DEVICE) return fail(PK_ERR_ARG, "pk_eri_gradient: bad memkind");
    if (c->nmol != 1) return fail(PK_ERR_ARG, "pk_eri_gradient: one molecule only");
    int rc;
    if (!c->work_ok && (rc = build_work(c)) != PK_OK) return rc;
    if ((rc = ensure_buffers(c)) != PK_OK) return rc;
    const size_t n2 = c->n2 * sizeof(double);
    if (!c->d_Pk1n) PK_HIP(hipMalloc((void **)&c->d_Pk1n, n2));
    if (!c->d_Pk2n) PK_HIP(hipMalloc((void **)&c->d_Pk2n, n2));
    if (!c->d_gsh) PK_HIP(hipMalloc((void **)&c->d_gsh, 3 * (size_t)c->ns * sizeof(double)));
    static const hipError_t attr = hipFuncSetAttribute(
        (const void *)pk::k_grad, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pk::grad_lds_bytes(kMaxL, kMaxL, kMaxL, kMaxL));
    PK_HIP(attr);
    hipStream_t s = c->stream;
    const double *src[3] = {Pj, Pk1, Pk2};
    if (memkind == PK_MEM_HOST) {
        // staging through the context's P, J and K buffers
        double *stage[3] = {c->d_P, c->d_J, c->d_K};
        for (int i = 0; i < 3; ++i)
            if (src[i]) {
                PK_HIP(hipMemcpyAsync(stage[i], src[i], n2, hipMemcpyHostToDevice, s));
                src[i] = stage[i];
            }
    }
    const MolTab t{c->d_mol_off, c->d_mol_nb, c->d_mol_foff, c->d_fpos};
    const dim3 g2((unsigned)(((size_t)c->nbmax * c->nbmax + 255) / 256), 1);
    double *dst[3] = {c->d_Pn, c->d_Pk1n, c->d_Pk2n};
    for (int i = 0; i < 3; ++i)
        if (src[i]) hipLaunchKernelGGL(k_scale_density, g2, dim3(256), 0, s, src[i], c->d_fnorm, dst[i], t);
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((3 * c->ns + 255) / 256)), dim3(256), 0, s, c->d_gsh,
                       (size_t)3 * c->ns, 0.0);
    for (const ClassWork &w : c->work) {
        const ClassFns &f = kClasses[w.cls];
        const int *wv = w.d_wave;
        pk::GradArgs a{};
        a.wave_bra = wv;
        a.wave_start = wv + w.nwaves;
        a.wave_count = wv + 2 * w.nwaves;
        a.kets = w.d_kets;
        a.pairs = c->d_pairs;
        a.prims = c->d_prims;
        a.palpha = c->d_palpha;
        a.hz = c->d_hz;
        a.pair_sh = c->d_pair_sh;
        a.Pj = c->d_Pn;
        a.Pk1 = c->d_Pk1n;
        a.Pk2 = Pk2 ? c->d_Pk2n : nullptr;
        a.Q = c->d_Q;
        a.g = c->d_gsh;
        a.tol = c->tol;
        a.kx = kx;
        a.nb = c->nb;
        a.nitems = w.nwaves;
        a.la = f.la;
        a.lb = f.lb;
        a.lc = f.lc;
        a.ld = f.ld;
        hipLaunchKernelGGL(pk::k_grad, dim3(w.nwaves), dim3(64), pk::grad_lds_bytes(f.la, f.lb, f.lc, f.ld), s, a);
    }
    PK_HIP(hipGetLastError());
    const size_t gb = 3 * (size_t)c->ns * sizeof(double);
    PK_HIP(hipMemcpyAsync(grad, c->d_gsh, gb, memkind == PK_MEM_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, s));
    PK_HIP(hipStreamSynchronize(s));
    return PK_OK;
}

int pk_schwarz(pk_ctx *c, double *Q, int memkind) {
    if (!c || !Q) return fail(PK_ERR_ARG, "pk_schwarz: null argument");
    if (memkind == PK_MEM_DEVICE) {
        const MolTab t{c->d_mol_off, c->d_mol_nb, c->d_mol_foff, c->d_fpos};
        const dim3 g2((unsigned)(((size_t)c->nbmax * c->nbmax + 255) / 256), (unsigned)c->nmol);
        hipLaunchKernelGGL(k_unpermute, g2, dim3(256), 0, c->stream, c->d_Q, Q, t);
        PK_HIP(hipGetLastError());
        PK_HIP(hipStreamSynchronize(c->stream));
    } else {
        for (int m = 0; m < c->nmol; ++m) {
            const int nb = c->mol_nb[m];
            const int *ps = c->fpos.data() + c->mol_foff[m];
            const double *in = c->Q.data() + c->mol_off[m];
            double *out = Q + c->mol_off[m];
            for (int r = 0; r < nb; ++r)
                for (int q = 0; q < nb; ++q) out[(size_t)r * nb + q] = in[(size_t)ps[r] * nb + ps[q]];
        }
    }
    return PK_OK;
}

int pk_eri_tensor(pk_ctx *c, double *out) {
    if (!c || !out) return fail(PK_ERR_ARG, "pk_eri_tensor: null argument");
    if (c->nmol != 1) return fail(PK_ERR_ARG, "pk_eri_tensor: one molecule only");
    const size_t nb = c->nb, nb2 = nb * nb, nb3 = nb2 * nb;
    if ((double)nb3 * nb * 8.0 > 2147483648.0) return fail(PK_ERR_ARG, "pk_eri_tensor: nb^4 tensor exceeds 2 GiB");
    std::fill(out, out + nb3 * nb, 0.0);
    const int np = (int)c->pairs.size();
    std::vector<std::vector<int>> qb(kQuartClasses), qk(kQuartClasses);
    for (int p = 0; p < np; ++p)
        for (int q = 0; q <= p; ++q) {
            int b = p, k = q;
            if (c->pair_cls[b] < c->pair_cls[k]) std::swap(b, k);
            const int cls = quart_class(c->pair_cls[b], c->pair_cls[k]);
            qb[cls].push_back(b);
            qk[cls].push_back(k);
        }
    for (int cls = 0; cls < kQuartClasses; ++cls) {
        if (qb[cls].empty()) continue;
        const ClassFns &f = kClasses[cls];
        std::vector<double> v;
        int rc = run_eri_class(c, pk::pair_class(f.la, f.lb), pk::pair_class(f.lc, f.ld), qb[cls], qk[cls], v);
        if (rc != PK_OK) return rc;
        const int na = pk::ncart(f.la), nbs = pk::ncart(f.lb), nc = pk::ncart(f.lc), nd = pk::ncart(f.ld);
        for (size_t q = 0; q < qb[cls].size(); ++q) {
            const pk::PairDev &B = c->pairs[qb[cls][q]], &K = c->pairs[qk[cls][q]];
            for (int ia = 0; ia < na; ++ia)
                for (int ib = 0; ib < nbs; ++ib)
                    for (int ic = 0; ic < nc; ++ic)
                        for (int id = 0; id < nd; ++id) {
                            // internal positions (screening, norms), then the caller's order
                            size_t i = B.f0 + ia, j = B.f1 + ib, k = K.f0 + ic, l = K.f1 + id;
                            if (c->Q[i * nb + j] * c->Q[k * nb + l] < c->tol) continue;
                            const double val = c->fnorm[i] * c->fnorm[j] * c->fnorm[k] * c->fnorm[l] *
                                               v[q * (size_t)na * nbs * nc * nd + ((ia * nbs + ib) * nc + ic) * nd + id];
                            i = c->fnat[i];
                            j = c->fnat[j];
                            k = c->fnat[k];
                            l = c->fnat[l];
                            const size_t idx[8] = {i * nb3 + j * nb2 + k * nb + l, j * nb3 + i * nb2 + k * nb + l,
                                                   i * nb3 + j * nb2 + l * nb + k, j * nb3 + i * nb2 + l * nb + k,
                                                   k * nb3 + l * nb2 + i * nb + j, l * nb3 + k * nb2 + i * nb + j,
                                                   k * nb3 + l * nb2 + j * nb + i, l * nb3 + k * nb2 + j * nb + i};
                            for (size_t x : idx) out[x] = val;
                        }
        }
    }
    return PK_OK;
}

}  // extern "C"
