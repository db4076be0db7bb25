/*
 * pk_ref — CPU oracle for the direct-SCF Fock build (TEST INFRASTRUCTURE ONLY).
 *
 * This library restates, in plain C++ on the CPU, the FP64 algorithm of Planck's
 * two-electron Fock build so that the HIP engine can be checked against it.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product never links, imports or calls it.
 *
 * Every entry point mirrors one reference function (file:line in pk_ref.cpp).
 * All matrices are nb x nb, column-major (symmetric, so row/col-major coincide).
 */
#ifndef PK_REF_H
#define PK_REF_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pkref_ctx pkref_ctx;

/* engine: 0 = Obara-Saika (engine os), 1 = Rys (engine rys), 2 = auto */
pkref_ctx *pkref_create(const char *gbs_path, int natoms, const int *Z,
                        const double *xyz_bohr, int engine, double tol_eri);
void pkref_destroy(pkref_ctx *ctx);
const char *pkref_last_error(void);

int pkref_nbasis(const pkref_ctx *ctx);
int pkref_nshells(const pkref_ctx *ctx);
/* shell table: per shell L, nprim, atom, center[3]; primitives exps/coef(xNc)/norm */
int pkref_shell_info(const pkref_ctx *ctx, int s, int *L, int *nprim, int *atom,
                     double *center, double *exps, double *coefs, double *norms);

double pkref_nuclear_repulsion(const pkref_ctx *ctx);
void pkref_overlap(const pkref_ctx *ctx, double *S);
void pkref_kinetic(const pkref_ctx *ctx, double *T);
void pkref_nuclear(const pkref_ctx *ctx, double *V);
void pkref_schwarz(pkref_ctx *ctx, double *Q);

/* one contracted Cartesian ERI (ij|kl) by the ctx's engine routing */
double pkref_eri(pkref_ctx *ctx, int i, int j, int k, int l);
/* one contracted Cartesian ERI forced through engine e (0 os, 1 rys) */
double pkref_eri_engine(pkref_ctx *ctx, int e, int i, int j, int k, int l);
/* full nb^4 tensor with Schwarz screening applied (small nb only) */
int pkref_eri_full(pkref_ctx *ctx, double *out);

/* true-direct J and K for density P (no nb^4 buffer); nthreads<=0 -> all */
int pkref_fock_jk(pkref_ctx *ctx, const double *P, double *J, double *K, int nthreads);
/* same restricted to bra AO pairs p with p % stride == offset (timing samples) */
int pkref_fock_jk_sample(pkref_ctx *ctx, const double *P, double *J, double *K,
                         int nthreads, long stride, long offset, double *n_quartets);

/* Elements (ei[e], ej[e]) of J and K of density P, e < nel, each summed on
 * its own over every AO quartet the full build keeps (same screening): equal
 * to those elements of pkref_fock_jk at O(nb^2) ERIs per element, so large
 * molecules can be checked on a sample of elements. */
int pkref_jk_elements(pkref_ctx *ctx, const double *P, int nel, const int *ei, const int *ej, double *J,
                      double *K, int nthreads);

/* 12 first derivatives d(ij|kl)/dX_q of one contracted ERI, i <= j, k <= l,
 * out[cen*3 + q] with cen 0..3 = centres of i, j, k, l (os.cpp:1296-1420). */
int pkref_eri_deriv(pkref_ctx *ctx, int i, int j, int k, int l, double *out12);
/* Two-electron gradient, natoms x 3 row-major (gradient.cpp:209-232, 70-118):
 * 1/4 sum Gamma_ijkl d(ij|kl)/dR with
 * Gamma = 2 Pj_ij Pj_kl - kx (Pk1_ik Pk1_jl + Pk2_ik Pk2_jl); Pk2 may be NULL. */
int pkref_eri_gradient(pkref_ctx *ctx, const double *Pj, const double *Pk1, const double *Pk2, double kx,
                       double *grad, int nthreads);

/* Primitive-pair screening of the engine (not the reference's behaviour),
 * for measuring its effect: rule 0 the prefactor bound, rule 1 the primitive
 * Schwarz bound (pk_ref.cpp).  Returns the number of shell primitive pairs
 * dropped; the Schwarz table is recomputed afterwards. */
int pkref_screen_primitives(pkref_ctx *ctx, double cut, int rule);

/* RHF SCF exactly as run_rhf (core guess).  hist gets E_total per iteration. */
int pkref_rhf(pkref_ctx *ctx, int charge, double tol_energy, double tol_density,
              int diis_dim, int max_cycles, double *energy, int *iterations,
              double *P_out, double *hist, int hist_len);

/* Boys F_n(x) by the table + 6-term Taylor lookup */
double pkref_boys(int n, double x);
/* 1: every later Boys evaluation (pkref_boys, the ERIs, the nuclear
 * attraction) uses the accurate series instead of the reference's lookup;
 * 0: the lookup (default).  Process-wide; not the reference's behaviour. */
void pkref_set_boys_exact(int on);
/* n-point Rys rule by the moment/Stieltjes/Jacobi construction, 1 <= n <= 5
 * (returns -1 and NaNs beyond: the method is inaccurate there; the oracle's
 * Rys and auto engines compute components needing more roots by OS) */
int pkref_rys_rule(int n, double T, double *roots, double *weights);
/* auto-engine routing decision for one Cartesian component quartet */
int pkref_auto_prefers_rys(const int *lA, const int *lB, const int *lC, const int *lD);

/* dense symmetric eigensolver used by the SCF (ascending eigenvalues) */
int pkref_syev(int n, const double *A, double *w, double *V);

#ifdef __cplusplus
}
#endif
#endif
