/*
 * pk_fock.h — C ABI of the MI355X direct-SCF two-electron Fock-build engine.
 *
 * Drop-in for the two-electron path of HemanthHaridas/planck-refactored:
 *   Eigen::MatrixXd _compute_2e_fock(shell_pairs, density, nbasis, engine, tol_eri, sym_ops)
 *       src/integrals/base.h:57-73  ->  pk_fock_2e / pk_fock_jk
 *   std::pair<MatrixXd,MatrixXd> _compute_2e_fock_uhf(shell_pairs, Pa, Pb, nbasis, engine, tol_eri, sym_ops)
 *       src/integrals/base.h:75-93  ->  pk_fock_jk_uhf
 *   std::vector<double> _compute_2e(shell_pairs, nbasis, engine, tol_eri, sym_ops)
 *       src/integrals/base.h:39-55  ->  pk_eri_tensor (small systems only)
 *   static MatrixXd _compute_schwarz_table(shell_pairs, nbasis, sym_ops)
 *       src/integrals/os.cpp:1703-1752, rys.cpp:509-556  ->  pk_schwarz
 *   build_shellpairs(basis)   src/integrals/shellpair.cpp:7-22   ->  pk_ctx_create
 *   OptionsIntegral{_tol_eri, _engine}   src/base/types.h:407-411  ->  pk_ctx_set
 *
 * Conventions (identical to the reference):
 *   - Cartesian Gaussian basis; shells in basis order, the functions of a shell
 *     contiguous in the order lx descending, then ly descending
 *     (src/basis/basis.cpp:20-37); component norm 1/sqrt((2lx-1)!!(2ly-1)!!(2lz-1)!!).
 *   - Matrices are nb x nb, symmetric, column-major (so row-major is the same).
 *   - Schwarz screening drops an AO quartet (ij|kl) when Q_ij * Q_kl < tol_eri
 *     with Q_ij = sqrt|(ij|ij)|  (src/integrals/os.cpp:1475).
 *   - RHF: G = J[P] - K[P] / 2 with P the total density (os.cpp:1509-1522).
 *   - UHF: G_s = J[Pa + Pb] - K[P_s] (os.cpp:1600-1620).
 *
 * Errors: every call returns PK_OK (0) or a negative PK_ERR_* code; the text of
 * the last error on the calling thread is pk_last_error().  No exception
 * crosses this boundary.  Threading: calls on one context are serialized by
 * the caller; each context owns one HIP stream (settable).
 */
#ifndef PK_FOCK_H
#define PK_FOCK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PK_OK 0
#define PK_ERR_ARG (-1)
#define PK_ERR_HIP (-2)
#define PK_ERR_UNSUPPORTED (-3)
#define PK_ERR_ALLOC (-4)
#define PK_ERR_STATE (-5)

/* IntegralMethod, src/base/types.h:95-100 */
enum { PK_ENGINE_OS = 0, PK_ENGINE_RYS = 1, PK_ENGINE_AUTO = 2 };
/* arithmetic of the ERI kernels */
enum { PK_PREC_FP64 = 0, PK_PREC_MIXED = 1 };
/* where the P / J / K / G pointers live */
enum { PK_MEM_HOST = 0, PK_MEM_DEVICE = 1 };

/* One contracted Cartesian shell per entry (src/base/types.h:263-276).
 * coef[i] is the file coefficient times the contracted normalisation Nc
 * (src/basis/gaussian.cpp:274-281); prim_norm[i] is the primitive
 * normalisation N_i (src/basis/basis.cpp:39-48).  exponent/coef/prim_norm are
 * the concatenation over shells in order, sum(nprim) entries each. */
typedef struct pk_basis_desc {
    int32_t nshells;
    const double *center;    /* [3 * nshells], Bohr */
    const int32_t *l;        /* [nshells], 0..3 */
    const int32_t *nprim;    /* [nshells] */
    const double *exponent;  /* [sum nprim] */
    const double *coef;      /* [sum nprim] */
    const double *prim_norm; /* [sum nprim] */
} pk_basis_desc;

typedef struct pk_ctx pk_ctx;

/* Counters of the last Fock build and of the screened work list. */
typedef struct pk_stats {
    int64_t nbasis;
    int64_t nshells;
    int64_t shell_pairs;          /* unique shell pairs a >= b */
    int64_t quartets;             /* shell quartets surviving the Schwarz test */
    int64_t primitive_quartets;   /* sum over surviving quartets of K_ab * K_cd */
    double model_flops;           /* algorithmic flop model v1 of the surviving work */
    double last_build_ms;         /* device time of the last build (HIP events) */
    double last_eri_kernel_ms;    /* device time spent in ERI/digestion kernels */
    int64_t kernel_launches;      /* kernels launched by the last build */
} pk_stats;

/* Build the context: upload shells, build primitive shell-pair data, compute
 * the Schwarz table on the device.  Once per geometry (shellpair.cpp:7-22,
 * os.cpp:1703-1752).  Defaults: engine os, tol_eri 1e-10, FP64. */
int pk_ctx_create(const pk_basis_desc *basis, pk_ctx **out);
int pk_ctx_destroy(pk_ctx *ctx);

/* Batch of nmol independent molecules in one context: the batched form of
 * the reference's per-molecule _compute_2e_fock (base.h:57-73), one call
 * building every molecule's J and K.  Matrices are packed: molecule m's
 * nb_m x nb_m block starts at element offset[m] (pk_ctx_mol_layout); the
 * quartet lists of all molecules are merged per angular class, so one build
 * is one kernel per class for the whole batch.  pk_ctx_create is the batch of
 * one.  pk_fock_jk / pk_fock_2e / pk_fock_jk_uhf / pk_schwarz / pk_ctx_set
 * work on batches unchanged (packed arrays); pk_eri_tensor takes one molecule. */
int pk_batch_create(const pk_basis_desc *basis, int nmol, pk_ctx **out);
int pk_ctx_nmol(const pk_ctx *ctx);
int pk_ctx_mol_layout(const pk_ctx *ctx, int32_t *nb, int64_t *offset);

/* hipGraph mode: the kernels of a build are captured once and replayed on
 * every later build (inputs and outputs copied through the context's own
 * buffers).  Off by default; recaptured after any change of the work lists. */
int pk_ctx_set_graph(pk_ctx *ctx, int enable);

/* Launch-shape autotuner: times every class list under each candidate shape
 * (1, 2 or 4 waves per workgroup; with or without XCD-contiguous wave ranges;
 * the list's alternative kernel: FP64 integrals for an FP32 list of the mixed
 * precision, the cooperative kernel for a class with per-lane Rys kernels)
 * over `reps` builds and keeps the fastest per class for this context.  The
 * density buffer is overwritten. */
int pk_ctx_autotune(pk_ctx *ctx, int reps);

/* Engine and threshold (types.h:407-411).  precision PK_PREC_MIXED computes
 * quartets whose Schwarz bound is below fp32_cutoff in FP32 (accumulation in
 * FP64); PK_PREC_FP64 ignores fp32_cutoff. */
int pk_ctx_set(pk_ctx *ctx, int engine, double tol_eri, int precision, double fp32_cutoff);
/* HIP stream (hipStream_t passed as void*; NULL = the context's own stream). */
int pk_ctx_set_stream(pk_ctx *ctx, void *hip_stream);
int pk_ctx_nbasis(const pk_ctx *ctx);
int pk_ctx_stats(const pk_ctx *ctx, pk_stats *out);

/* J[P] and K[P] for a symmetric density P (nb x nb).  J or K may be NULL. */
int pk_fock_jk(pk_ctx *ctx, const double *P, double *J, double *K, int memkind);
/* G = J[P] - K[P]/2, the return value of _compute_2e_fock (base.h:57-73). */
int pk_fock_2e(pk_ctx *ctx, const double *P, double *G, int memkind);
/* UHF: Jt = J[Pa + Pb], Ka = K[Pa], Kb = K[Pb]; G_a = Jt - Ka, G_b = Jt - Kb
 * (_compute_2e_fock_uhf, base.h:75-93).  One build (FP64): the Obara-Saika
 * and cooperative classes compute each quartet once and digest it into J of
 * Pa + Pb and both exchange matrices; the per-lane Rys classes run once per
 * spin density.  Host or device pointers; Jt, Ka or Kb may be NULL.  Packed
 * arrays for batches, as pk_fock_jk.  Runs eagerly (no graph). */
int pk_fock_jk_uhf(pk_ctx *ctx, const double *Pa, const double *Pb, double *Jt, double *Ka, double *Kb,
                   int memkind);

/* pk_fock_jk without any device allocation in the call.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  P, J, K are device
 * pointers; the build is enqueued on the context's stream and the call returns
 * without synchronising.  The per-build scratch is a static device array of
 * the library (28 MiB per GPU, held until the process ends): the packed
 * matrices may hold at most 1048576 elements (one molecule of up to 1024
 * functions; PK_ERR_ARG beyond), calls that overlap on one GPU race, and the
 * work lists of the current settings must already exist (pk_fock_jk once after
 * pk_ctx_create / pk_ctx_set; PK_ERR_STATE otherwise). */
int pk_fock_jk_noalloc(pk_ctx *ctx, const double *P, double *J, double *K);

/* Q_ij = sqrt|(ij|ij)|, nb x nb (the reference's Schwarz table). */
int pk_schwarz(pk_ctx *ctx, double *Q, int memkind);
/* Full screened nb^4 ERI tensor, index ((i*nb + j)*nb + k)*nb + l, as the
 * reference's _compute_2e returns it (base.h:39-55).  Host memory; refused
 * when nb^4 * 8 bytes exceeds 2 GiB. */
int pk_eri_tensor(pk_ctx *ctx, double *out);

/* Two-electron part of the nuclear gradient: the contraction of the ERI first
 * derivatives (ObaraSaika::_compute_eri_deriv_elem, src/integrals/os.cpp:1296-1420)
 * with the two-particle density in the ERI loop of the reference's gradients
 * (src/gradient/gradient.cpp:209-232 closed shell, :436-463 UHF):
 *   grad[3 s + q] = dE2 / d(centre of shell s)_q,  E2 = 1/4 sum_ijkl Gamma_ijkl (ij|kl),
 *   Gamma_ijkl = 2 Pj_ij Pj_kl - kx (Pk1_ik Pk1_jl + Pk2_ik Pk2_jl).
 * RHF: Pj = Pk1 = P (total density), Pk2 = NULL, kx = 1; RKS: kx = the exact-
 * exchange fraction (gradient.cpp:307-310); UHF: Pj = Pa + Pb, Pk1 = Pa,
 * Pk2 = Pb, kx = 2 (gradient.cpp:428-434).  grad has 3 * nshells entries in the
 * caller's shell order; the caller adds the shells of each atom (the
 * reference's build_shell_atom_map, gradient.cpp:20-46).  FP64, the context's
 * tol_eri screening; one molecule; host or device pointers. */
int pk_eri_gradient(pk_ctx *ctx, const double *Pj, const double *Pk1, const double *Pk2, double kx, double *grad,
                    int memkind);

/* Per-class device timing (HIP events around each class's kernels on the
 * context stream).  Disabled by default; enabling resets the counters. */
typedef struct pk_class_stats {
    int32_t la, lb, lc, ld;       /* angular class of the kernels */
    int32_t kernels;              /* kernels per build (ket-component chunks) */
    int64_t quartets;
    int64_t primitive_quartets;
    double model_flops;           /* flop model v1 of one build of this class */
    double ms_total;              /* summed device time over profiled builds */
    int64_t builds;               /* profiled builds */
    int32_t precision;            /* PK_PREC_FP64, or PK_PREC_MIXED for the FP32 list */
    int32_t block;                /* lanes per workgroup of the class kernels */
    int32_t remap;                /* 1: XCD-contiguous wave ranges */
    int64_t waves;                /* work items (one bra, up to 64 kets per wave) */
    int32_t alt;                  /* 1: the list runs its alternative kernel (autotuned): FP64
                                     integrals for an FP32 list, or the cooperative kernel for a
                                     class with per-lane Rys kernels */
    int32_t kind;                 /* kernels that run the list: 0 per-lane Obara-Saika (FP64),
                                     1 per-lane Rys (FP64), 2 cooperative (FP64), 3 per-lane
                                     Obara-Saika (FP32), 4 per-lane Rys (FP32) */
} pk_class_stats;
int pk_ctx_profile(pk_ctx *ctx, int enable);
int pk_ctx_class_count(const pk_ctx *ctx);
int pk_ctx_class_stats(const pk_ctx *ctx, int index, pk_class_stats *out);

const char *pk_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* PK_FOCK_H */
