/*
 * pk_scf.h — C ABI of the host-side layer above the Fock-build engine:
 * basis construction, one-electron integrals and the RHF SCF driver.
 *
 * Reference interfaces replaced:
 *   read_gbs_basis(file, molecule, basis_type)   src/basis/gaussian.cpp:183-302  -> pk_basis_load
 *   _compute_1e(shell_pairs, nbasis, ...)        src/integrals/base.h:10-18      -> pk_one_electron (S, T)
 *   _compute_nuclear_attraction(...)             src/integrals/base.h:20-27      -> pk_one_electron (V)
 *   Calculator::_compute_nuclear_repulsion()     src/base/types.h:934-964        -> pk_nuclear_repulsion
 *   SCF::run_rhf(calculator, shell_pairs, pcm)   src/scf/scf.cpp:175-441         -> pk_scf_rhf
 *
 * The SCF driver keeps every nb x nb matrix on the device (rocBLAS GEMMs,
 * rocSOLVER syevd) and obtains G = J - K/2 from pk_fock_2e each iteration.
 * Return codes are those of pk_fock.h; the message of the last error on the
 * calling thread is pk_host_last_error().
 */
#ifndef PK_SCF_H
#define PK_SCF_H

#include <stdint.h>

#include "pk_fock.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pk_basis pk_basis;

/* Gaussian94 basis file for the atoms Z[i] at xyz_bohr[3 i..3 i + 2] (Bohr).
 * Shells per atom in file order, SP split into S then P, coefficients scaled
 * by the contracted normalisation (gaussian.cpp:54-163, 183-302). */
int pk_basis_load(const char *gbs_path, int natoms, const int32_t *Z, const double *xyz_bohr, pk_basis **out);
void pk_basis_free(pk_basis *b);
int pk_basis_nbasis(const pk_basis *b);
/* pointers into b, valid until pk_basis_free */
int pk_basis_describe(const pk_basis *b, pk_basis_desc *desc);

/* S, T, V (nb x nb, host memory; any may be NULL), computed on the device by
 * a HIP kernel; PK_ERR_HIP when no HIP device is visible (no CPU path). */
int pk_one_electron(const pk_basis *b, double *S, double *T, double *V);
double pk_nuclear_repulsion(const pk_basis *b);

typedef struct pk_scf_options {
    double tol_energy;   /* default 1e-10 (types.h:337) */
    double tol_density;  /* default 1e-10 */
    double tol_eri;      /* default 1e-10 (types.h:410) */
    int32_t diis_dim;    /* default 8 */
    int32_t max_cycles;  /* 0 = by basis size (types.h:358-363) */
    int32_t engine;      /* PK_ENGINE_* */
    int32_t charge;
    int32_t use_guess;   /* 1: start from guess_density instead of the core guess */
    const double *guess_density;
    int32_t precision;   /* PK_PREC_FP64 (default) or PK_PREC_MIXED */
    double fp32_cutoff;  /* mixed: quartets with Schwarz bound below it run in FP32 */
    int32_t guess_fallback; /* with use_guess: start from the core guess instead when
                               its first DIIS error is smaller (one extra Fock build) */
    int32_t sad_guess;   /* 1: SAD initial density (sad.cpp:589-643) instead of the
                            core Hamiltonian (ignored with use_guess) */
    int32_t guess_only;  /* 1: return the initial density in P_out, no iterations */
} pk_scf_options;

typedef struct pk_scf_result {
    double energy;           /* E_total of the last iteration */
    double nuclear_repulsion;
    int32_t iterations;
    int32_t converged;
    double first_diis_error; /* RMS of the first DIIS error matrix */
    double wall_s;           /* guess to convergence */
    double fock_ms;          /* summed device time of the Fock builds */
    int32_t guess_used;      /* 1: the supplied guess density, 0: the core guess */
    double guess_diis_error; /* first DIIS errors of the supplied and the core */
    double core_diis_error;  /* guess (guess_fallback only; else 0) */
} pk_scf_result;

void pk_scf_default_options(pk_scf_options *o);
/* RHF direct SCF (scf.cpp:175-441).  hist (may be NULL) receives E_total per
 * iteration; P_out (may be NULL) the final density. */
int pk_scf_rhf(const pk_basis *b, const pk_scf_options *opt, pk_scf_result *res, double *P_out, double *hist,
               int hist_len);

/* Batched RHF over nmol independent molecules (each the iteration of
 * pk_scf_rhf): the Fock builds of one iteration are one batched, hipGraph-
 * replayed build of every molecule (pk_batch_create); the per-molecule algebra
 * follows on the same stream.  res[nmol]; P_out[i] and guess[i] (host nb_i^2
 * densities) may be NULL, as may the arrays.  guess_fallback applies per
 * molecule as in pk_scf_rhf (the first DIIS errors of the supplied and the
 * core densities come from two batched builds).  wall_s is the batch's wall time;
 * fock_ms the summed batched-build time over the molecule's iterations.  Not
 * in the reference (its SCF handles one molecule). */
int pk_scf_rhf_batch(const pk_basis *const *bases, int nmol, const pk_scf_options *opt, pk_scf_result *res,
                     double *const *P_out, const double *const *guess);

const char *pk_host_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* PK_SCF_H */
