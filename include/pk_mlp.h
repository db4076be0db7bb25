/*
 * pk_mlp.h — C ABI of the atom-pair-block network of the learned SCF initial
 * guess: a fully connected network n_in -> hidden -> hidden -> n_out with tanh
 * activations, FP32, trained and evaluated on the device by hand-written HIP
 * kernels (tiled FP32 GEMMs with fused bias / activation, their backward
 * passes, a mean-squared-error loss and Adam).
 *
 * No counterpart in the reference: its initial guesses are the core
 * Hamiltonian (src/scf/scf.cpp:33-45) and SAD (src/scf/sad.cpp:589-643); this
 * network predicts density blocks that replace them, with those guesses as the
 * fallback (see pkfock/guess.py).
 *
 * Buffers: X, T, Y are device pointers, row-major [batch][n]; parameters are
 * W1 [n_in][hidden], b1, W2 [hidden][hidden], b2, W3 [hidden][n_out], b3 in
 * that order (pk_mlp_nparams floats).  Errors: PK_OK / PK_ERR_* from pk_fock.h,
 * message in pk_mlp_last_error().  One context per network, one stream.
 */
#ifndef PK_MLP_H
#define PK_MLP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pk_mlp pk_mlp;

/* Network with Glorot-uniform weights (seeded), zero biases, Adam state zero.
 * max_batch bounds the batch of every later call (workspace is sized for it). */
int pk_mlp_create(int n_in, int hidden, int n_out, int max_batch, uint64_t seed, pk_mlp **out);
int pk_mlp_destroy(pk_mlp *m);
int pk_mlp_set_stream(pk_mlp *m, void *stream);
size_t pk_mlp_nparams(const pk_mlp *m);
/* host copies of the parameters / gradients (pk_mlp_nparams floats) */
int pk_mlp_get_params(pk_mlp *m, float *host);
int pk_mlp_set_params(pk_mlp *m, const float *host);
int pk_mlp_get_grads(pk_mlp *m, float *host);

/* Y = net(X) for batch rows (inference). */
int pk_mlp_forward(pk_mlp *m, const float *X, int batch, float *Y);
/* Forward and backward of L = mean over rows and outputs of (net(X) - T)^2;
 * the gradients are left in the context; *loss (host) receives L. */
int pk_mlp_loss_grad(pk_mlp *m, const float *X, const float *T, int batch, float *loss);
/* One Adam step with the stored gradients (bias-corrected, step count kept). */
int pk_mlp_adam_step(pk_mlp *m, float lr, float beta1, float beta2, float eps);

const char *pk_mlp_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
