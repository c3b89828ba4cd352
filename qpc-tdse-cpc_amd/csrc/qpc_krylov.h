// qpc_krylov.h -- device pieces of the Krylov (S-orthogonal Lanczos) propagator,
// propagator_arnoldi with propagator_hc_vgwz::multiply (la_1e.hpp:90-335, 480-497).
#pragma once
#include "qpc_common.h"

namespace qpc {

constexpr int kDotBlocks = 128;  // partial sums per inner product (summed on the host in order)

// u1 = H(az) u0 over every channel (la_1e.hpp:480-497): H_l per channel and the
// l -> l+-1 couplings  +i az q(l-1) F - i az y(l-1) G  (from l-1) and
// -i az y(l) G - i az q(l) F  (from l+1); F = <Bi|1/r|Bj> symmetric, G =
// <Bi|d/dr Bj> antisymmetric; all bands diagonal-major (D[j*N+i] = A(i,i+j)),
// H per l at Hd + l*(p+1)*N; p is n_rank (7 or 8).
void launch_hmul(hipStream_t s, const double* Hd, const double* Fd, const double* Gd, const double* y10,
                 const double* q10, double az, int M, int L, int m0, int N, int p, const cplx* u0, cplx* u1);
// part[i*kDotBlocks + b] = partial sum over block b of conj(U_i) . v, i < k
void launch_dots(hipStream_t s, const cplx* U, int64_t stride, int k, const cplx* v, int64_t n, cplx* part);
// v -= sum_{i<k} c[i] U_i
void launch_axpy_multi(hipStream_t s, const cplx* U, int64_t stride, int k, const cplx* c, cplx* v, int64_t n);
// v *= a
void launch_scale(hipStream_t s, cplx* v, double a, int64_t n);
// w = sum_{i<k} c[i] U_i
void launch_combine(hipStream_t s, const cplx* U, int64_t stride, int k, const cplx* c, cplx* w, int64_t n);

}  // namespace qpc
