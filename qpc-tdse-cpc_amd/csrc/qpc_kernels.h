// qpc_kernels.h -- launchers of the HIP kernels of the propagator.
#pragma once
#include "qpc_common.h"

namespace qpc {

// rotations (FP64 matrix cores): largest l-dimension and row tiles
constexpr int kRotMaxL = 96;
constexpr int kRotMaxRT = kRotMaxL / 16;
// elements per m of a prepared operator
int64_t rot_prep_elems(int64_t L);
// prepared operator of P' = P (trans = 0) or P^H (trans = 1; P^T for real P)
void launch_rot_prep_real(hipStream_t s, const double* P, double* Pp, int64_t M, int64_t L, int trans);
void launch_rot_prep_cplx(hipStream_t s, const cplx* P, cplx* Pp, int64_t M, int64_t L, int trans);
// dst[m][k][x] = sum_i P'[m][k][i] src[m][i][x], x < N, with Pp from launch_rot_prep_*
void launch_rot_real(hipStream_t s, const double* Pp, const cplx* src, cplx* dst, int64_t M, int64_t L, int64_t N);
void launch_rot_cplx(hipStream_t s, const cplx* Pp, const cplx* src, cplx* dst, int64_t M, int64_t L, int64_t N);
// per-channel partial sums of <c|S|c> into red[nch]
void launch_norm(hipStream_t s, const double* S, const cplx* coef, double* red, int64_t nch, int64_t N, int64_t p);
// per-channel partial sums of Re observable_z into red[M*L]
void launch_dipole_z(hipStream_t s, const double* rband, const double* y10, const cplx* coef, double* red, int64_t M,
                     int64_t L, int64_t N, int64_t p);

}  // namespace qpc
