// pk_host_internal.h — host-side basis record shared by pk_host.cpp and pk_scf.hip.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "pk_scf.h"

typedef struct ihipStream_t *hipStream_t;

namespace pkh {

struct Basis {
    int nb = 0;
    std::vector<int32_t> Z;
    std::vector<double> xyz;  // Bohr
    std::vector<double> center;
    std::vector<int32_t> l, nprim, atom;
    std::vector<double> exponent, coef, prim_norm;
    std::vector<double> fnorm;  // component norm per basis function
};

extern thread_local std::string g_err;
int fail(int code, const std::string &m);
int load_basis(const char *path, int natoms, const int32_t *Z, const double *xyz, Basis &b);
double nuclear_repulsion(const Basis &b);
// S, T, V (row-major nb x nb, symmetric) computed on the device: host
// pointers (one_electron) or device pointers on the caller's stream
// (one_electron_device, returns after the kernel completed); null = skip
int one_electron(const Basis &b, double *S, double *T, double *V);
int one_electron_device(const Basis &b, double *dS, double *dT, double *dV, hipStream_t st);
// the shells of atom a moved to the origin (a one-atom basis of the same element)
Basis atom_basis(const Basis &b, int a);
// symmetric eigenproblem by cyclic Jacobi rotations (small matrices):
// A (n x n, row-major, symmetric) -> w ascending, V columns (row-major V[i*n+k])
void jacobi_eigh(int n, std::vector<double> A, std::vector<double> &w, std::vector<double> &V);

}  // namespace pkh

const pkh::Basis &pkh_basis_of(const pk_basis *b);
