// pk_internal.h — engine functions shared between the engine's translation
// units; not part of the C ABI.
#pragma once

#include <stddef.h>

#include "pk_fock.h"

// One J/K build from the device density P into the device matrices J and K on
// the context's stream, with caller-provided per-build scratch: three arrays
// of `each` doubles (scaled density, J and K accumulators) followed by `each`
// floats (FP32 copy of the density for mixed precision), 3.5 `each` doubles
// in all.  Allocates nothing
// and returns without synchronising.  Fails with PK_ERR_ARG when the packed
// matrices of the context exceed `each` elements and with PK_ERR_STATE when
// the work lists of the current settings have not been built (pk_fock_jk
// builds them).  The error text names `caller`.
int pk_build_with_scratch(pk_ctx *ctx, const double *P, double *J, double *K, double *scratch, size_t each,
                          const char *caller);
// Sets the thread's error text (pk_last_error) and returns code.
int pk_set_error(int code, const char *msg);
