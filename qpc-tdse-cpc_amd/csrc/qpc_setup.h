// qpc_setup.h -- internal header of the host-side set-up (public API in include/qpc_gpu.h).
#pragma once
#include "../../include/qpc_gpu.h"
#include "qpc_common.h"

#include <array>
#include <vector>

namespace qpc {
// symmetric tridiagonal eigen-decomposition (diagonal d[n], off-diagonal
// e[n-1]): eigenvalues w ascending, eigenvector k in z[k*n .. k*n+n)
int tridiag_eigen(int n, const double* d, const double* e, double* w, double* z);
// The angular operators of the elliptically polarised steps, x1 = Y11,
// x2 = -i Q11, y1 = -i Y11, y2 = Q11: the two with multiplier -i become real
// symmetric after the phase change diag(i^im) (eigenvector components of
// channel j times i^(im_j)).  angular_blocks builds each operator's dense
// parity blocks B[op][pb] (n_pb x n_pb, il of parity pb in the order idx[pb];
// par and loc give each channel's block and position); angular_rows orders
// one operator's eigenpairs as converter::catagorize does (even, zero for odd
// M, odd; ascending within each), from its per-parity eigenvalues.
constexpr bool kAngImagMult[4] = {false, true, true, false};
struct AngBlocks {
    std::vector<int> idx[2], par, loc;
    std::vector<double> B[4][2];
};
struct AngRow {
    int pb, k;  // parity block and local eigenvector
};
void angular_blocks(int M, int L, int m0, AngBlocks& ab);
// The four angular operators couple m to m +- 1 only, so each parity block's
// spectrum is symmetric about 0: in catagorize's order the even (odd)
// category's values are w and -w from either end.  pair_spectrum makes the
// computed pairs exact negatives (their mean magnitude) and the zero category
// exactly 0, so that sweeps of S + i e X at e and -e can share one table
// (its complex conjugate).
void pair_spectrum(double* w, int M, int L);
int angular_rows(int M, int L, const std::vector<double> w[2], std::vector<AngRow>& rows);
// angular eigen-data of the elliptically polarised steps (qpc_setup_angular_ep,
// and with x1_2_y1 / tran_y1 those of qpc_setup_angular_lgwr); null merged
// outputs are skipped
int angular_ep(int M, int L, int m0, std::array<double*, 4> eig, double* tran_x1, double* tran_y2,
               std::array<double*, 3> merged, double* x1_2_y1 = nullptr, double* tran_y1 = nullptr);
}  // namespace qpc
