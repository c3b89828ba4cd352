// qpc_krylov.hip -- device pieces of the Krylov propagator (gfx950).
#include "qpc_krylov.h"

namespace qpc {
namespace {

// (A x)_i of a band with kNe entries per row (n_rank + 1), symmetric (SIGN =
// 1) or antisymmetric (SIGN = -1), diagonal-major
template <int SIGN, int kNe>
__device__ inline cplx band_row(const double* __restrict__ D, const cplx* __restrict__ x, int i, int N) {
    cplx t = {0.0, 0.0};
#pragma unroll
    for (int j = 1; j < kNe; ++j)
        if (i - j >= 0) {
            const double a = D[(int64_t)j * N + i - j];
            t += (SIGN > 0 ? a : -a) * x[i - j];
        }
#pragma unroll
    for (int k = 0; k < kNe; ++k)
        if (i + k < N) t += D[(int64_t)k * N + i] * x[i + k];
    return t;
}

__device__ inline cplx times_i(cplx z, double a) { return {-a * z.im, a * z.re}; }  // i a z

template <int kNe>
__global__ __launch_bounds__(256) void k_hmul(const double* __restrict__ Hd, const double* __restrict__ Fd,
                                              const double* __restrict__ Gd, const double* __restrict__ y10,
                                              const double* __restrict__ q10, double az, int L, int m0, int N,
                                              const cplx* __restrict__ u0, cplx* __restrict__ u1) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int ch = blockIdx.y;
    if (i >= N) return;
    const int im = ch / L, il = ch % L;
    const int m = m0 + im;
    const int l = il + (m < 0 ? -m : m);
    const cplx* x = u0 + (int64_t)ch * N;
    cplx h = band_row<1, kNe>(Hd + (int64_t)l * kNe * N, x, i, N);
    if (il > 0) {
        const cplx f = band_row<1, kNe>(Fd, x - N, i, N), g = band_row<-1, kNe>(Gd, x - N, i, N);
        h += times_i(f, az * q10[ch - 1]);
        h -= times_i(g, az * y10[ch - 1]);
    }
    if (il < L - 1) {
        const cplx f = band_row<1, kNe>(Fd, x + N, i, N), g = band_row<-1, kNe>(Gd, x + N, i, N);
        h -= times_i(g, az * y10[ch]);
        h -= times_i(f, az * q10[ch]);
    }
    u1[(int64_t)ch * N + i] = h;
}

__global__ __launch_bounds__(256) void k_dots(const cplx* __restrict__ U, int64_t stride, const cplx* __restrict__ v,
                                              int64_t n, cplx* __restrict__ part) {
    __shared__ double sre[256], sim[256];
    const int b = blockIdx.x, vi = blockIdx.y;
    const cplx* u = U + (int64_t)vi * stride;
    const int64_t chunk = (n + kDotBlocks - 1) / kDotBlocks;
    const int64_t e0 = (int64_t)b * chunk, e1 = e0 + chunk < n ? e0 + chunk : n;
    double re = 0.0, im = 0.0;
    for (int64_t e = e0 + threadIdx.x; e < e1; e += 256) {
        const cplx a = u[e], c = v[e];
        re += a.re * c.re + a.im * c.im;
        im += a.re * c.im - a.im * c.re;
    }
    sre[threadIdx.x] = re;
    sim[threadIdx.x] = im;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            sre[threadIdx.x] += sre[threadIdx.x + st];
            sim[threadIdx.x] += sim[threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(int64_t)vi * kDotBlocks + b] = cplx{sre[0], sim[0]};
}

__global__ __launch_bounds__(256) void k_axpy_multi(const cplx* __restrict__ U, int64_t stride, int k,
                                                    const cplx* __restrict__ c, cplx* __restrict__ v, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    cplx t = v[e];
    for (int i = 0; i < k; ++i) {
        const cplx ci = c[i];
        if (ci.re != 0.0 || ci.im != 0.0) t -= U[(int64_t)i * stride + e] * ci;
    }
    v[e] = t;
}

__global__ __launch_bounds__(256) void k_scale(cplx* __restrict__ v, double a, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e < n) v[e] = v[e] * a;
}

__global__ __launch_bounds__(256) void k_combine(const cplx* __restrict__ U, int64_t stride, int k,
                                                 const cplx* __restrict__ c, cplx* __restrict__ w, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    cplx t = U[e] * c[0];
    for (int i = 1; i < k; ++i) t += U[(int64_t)i * stride + e] * c[i];
    w[e] = t;
}

}  // namespace

void launch_hmul(hipStream_t s, const double* Hd, const double* Fd, const double* Gd, const double* y10,
                 const double* q10, double az, int M, int L, int m0, int N, int p, const cplx* u0, cplx* u1) {
    if (p == 8)
        hipLaunchKernelGGL(k_hmul<9>, dim3((N + 255) / 256, M * L), dim3(256), 0, s, Hd, Fd, Gd, y10, q10, az, L, m0,
                           N, u0, u1);
    else
        hipLaunchKernelGGL(k_hmul<8>, dim3((N + 255) / 256, M * L), dim3(256), 0, s, Hd, Fd, Gd, y10, q10, az, L, m0,
                           N, u0, u1);
}
void launch_dots(hipStream_t s, const cplx* U, int64_t stride, int k, const cplx* v, int64_t n, cplx* part) {
    hipLaunchKernelGGL(k_dots, dim3(kDotBlocks, k), dim3(256), 0, s, U, stride, v, n, part);
}
void launch_axpy_multi(hipStream_t s, const cplx* U, int64_t stride, int k, const cplx* c, cplx* v, int64_t n) {
    hipLaunchKernelGGL(k_axpy_multi, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, U, stride, k, c, v, n);
}
void launch_scale(hipStream_t s, cplx* v, double a, int64_t n) {
    hipLaunchKernelGGL(k_scale, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, v, a, n);
}
void launch_combine(hipStream_t s, const cplx* U, int64_t stride, int k, const cplx* c, cplx* w, int64_t n) {
    hipLaunchKernelGGL(k_combine, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, U, stride, k, c, w, n);
}

}  // namespace qpc
