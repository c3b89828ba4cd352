// qpc_kernels.h -- launchers of the HIP kernels of the propagator.
#pragma once
#include "qpc_common.h"

namespace qpc {

constexpr int kMaxRank = 8;  // band half-width supported by the register-resident kernels

// dst[m][k][x] = sum_i P[m][k][i] src[m][i][x]   (trans=0)
// dst[m][k][x] = sum_i P[m][i][k] src[m][i][x]   (trans=1), x in [0,X), real
void launch_rot_real(hipStream_t s, const double* P, const double* src, double* dst, int64_t M, int64_t L,
                     int64_t X, int trans);
// complex: trans=1 uses conj(P[m][i][k])
void launch_rot_cplx(hipStream_t s, const cplx* P, const cplx* src, cplx* dst, int64_t M, int64_t L, int64_t N,
                     int trans);
// cn_propagate_kernel_avx1 per rotated channel (reference propagate.hpp:3-98)
void launch_cn_onthefly(hipStream_t s, const double* smat, const double* rmat, const double* eig, int symflag,
                        double fzdt, cplx* z, cplx* ytmp, cplx* utmp, int64_t nch, int64_t N, int64_t p);
// cn_hc_propagate0 (reference propagate.hpp:311-331)
void launch_hc_free(hipStream_t s, const cplx* upper, const cplx* lower, cplx* coef, cplx* w, int64_t M, int64_t L,
                    int64_t m0, int64_t N, int64_t p);
// absorber_mask::operator() (reference absorber.hpp:52-64)
void launch_mask(hipStream_t s, const double* mask, const double* chol, cplx* coef, cplx* t, int64_t nch, int64_t N,
                 int64_t p);
// per-channel partial sums of <c|S|c> into red[nch]
void launch_norm(hipStream_t s, const double* S, const cplx* coef, double* red, int64_t nch, int64_t N, int64_t p);
// per-channel partial sums of Re observable_z into red[M*L]
void launch_dipole_z(hipStream_t s, const double* rband, const double* y10, const cplx* coef, double* red, int64_t M,
                     int64_t L, int64_t N, int64_t p);

}  // namespace qpc
