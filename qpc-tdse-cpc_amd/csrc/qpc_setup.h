// qpc_setup.h -- internal header of the host-side set-up (public API in include/qpc_gpu.h).
#pragma once
#include "../../include/qpc_gpu.h"
#include "qpc_common.h"

#include <array>

namespace qpc {
// symmetric tridiagonal eigen-decomposition (diagonal d[n], off-diagonal
// e[n-1]): eigenvalues w ascending, eigenvector k in z[k*n .. k*n+n)
int tridiag_eigen(int n, const double* d, const double* e, double* w, double* z);
// angular eigen-data of the elliptically polarised steps (qpc_setup_angular_ep,
// and with x1_2_y1 / tran_y1 those of qpc_setup_angular_lgwr); null merged
// outputs are skipped
int angular_ep(int M, int L, int m0, std::array<double*, 4> eig, double* tran_x1, double* tran_y2,
               std::array<double*, 3> merged, double* x1_2_y1 = nullptr, double* tran_y1 = nullptr);
}  // namespace qpc
