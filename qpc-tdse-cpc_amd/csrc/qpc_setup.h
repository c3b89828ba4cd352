// qpc_setup.h -- internal header of the host-side set-up (public API in include/qpc_gpu.h).
#pragma once
#include "../../include/qpc_gpu.h"
#include "qpc_common.h"

namespace qpc {
// symmetric tridiagonal eigen-decomposition (diagonal d[n], off-diagonal
// e[n-1]): eigenvalues w ascending, eigenvector k in z[k*n .. k*n+n)
int tridiag_eigen(int n, const double* d, const double* e, double* w, double* z);
}  // namespace qpc
