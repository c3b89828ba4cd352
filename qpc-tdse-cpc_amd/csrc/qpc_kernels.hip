// qpc_kernels.hip -- HIP kernels (gfx950) of the Crank-Nicolson propagator.
//
// Reference operation order is kept inside each recurrence so that results
// agree with the reference's FP64 CPU run to rounding.  Citations are
// file:line in zzh535897/QPC-TDSE-CPC.
#include "qpc_kernels.h"

namespace qpc {
namespace {

// ---------------------------------------------------------------------------
// angular rotations: cn_transform_hl::construct / eliminate (cn_1e.hpp:101-147)
// One thread owns one column x and KT output rows; the KT x L slice of P
// sits in LDS.
constexpr int kRotThreads = 256;
constexpr int kRotKT = 8;

__global__ __launch_bounds__(kRotThreads) void k_rot_real(const double* __restrict__ P, const double* __restrict__ src,
                                                          double* __restrict__ dst, int L, int64_t X, int trans) {
    extern __shared__ double sP[];  // [kRotKT][L]
    const int m = blockIdx.z;
    const int k0 = blockIdx.y * kRotKT;
    const double* Pm = P + (int64_t)m * L * L;
    for (int t = threadIdx.x; t < kRotKT * L; t += blockDim.x) {
        int kk = t / L, i = t % L;
        int k = k0 + kk;
        double v = 0.0;
        if (k < L) v = trans ? Pm[(int64_t)i * L + k] : Pm[(int64_t)k * L + i];
        sP[t] = v;
    }
    __syncthreads();
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= X) return;
    const double* s = src + (int64_t)m * L * X + x;
    double acc[kRotKT];
#pragma unroll
    for (int kk = 0; kk < kRotKT; ++kk) acc[kk] = 0.0;
    for (int i = 0; i < L; ++i) {
        double v = s[(int64_t)i * X];
#pragma unroll
        for (int kk = 0; kk < kRotKT; ++kk) acc[kk] += sP[kk * L + i] * v;
    }
    double* d = dst + (int64_t)m * L * X + x;
#pragma unroll
    for (int kk = 0; kk < kRotKT; ++kk)
        if (k0 + kk < L) d[(int64_t)(k0 + kk) * X] = acc[kk];
}

constexpr int kRotKTc = 4;
__global__ __launch_bounds__(kRotThreads) void k_rot_cplx(const cplx* __restrict__ P, const cplx* __restrict__ src,
                                                          cplx* __restrict__ dst, int L, int64_t N, int trans) {
    extern __shared__ double sPd[];  // [kRotKTc][L] complex
    cplx* sP = reinterpret_cast<cplx*>(sPd);
    const int m = blockIdx.z;
    const int k0 = blockIdx.y * kRotKTc;
    const cplx* Pm = P + (int64_t)m * L * L;
    for (int t = threadIdx.x; t < kRotKTc * L; t += blockDim.x) {
        int kk = t / L, i = t % L;
        int k = k0 + kk;
        cplx v = {0.0, 0.0};
        if (k < L) v = trans ? conjc(Pm[(int64_t)i * L + k]) : Pm[(int64_t)k * L + i];
        sP[t] = v;
    }
    __syncthreads();
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= N) return;
    const cplx* s = src + (int64_t)m * L * N + x;
    cplx acc[kRotKTc];
#pragma unroll
    for (int kk = 0; kk < kRotKTc; ++kk) acc[kk] = {0.0, 0.0};
    for (int i = 0; i < L; ++i) {
        cplx v = s[(int64_t)i * N];
#pragma unroll
        for (int kk = 0; kk < kRotKTc; ++kk) acc[kk] += sP[kk * L + i] * v;
    }
    cplx* d = dst + (int64_t)m * L * N + x;
#pragma unroll
    for (int kk = 0; kk < kRotKTc; ++kk)
        if (k0 + kk < L) d[(int64_t)(k0 + kk) * N] = acc[kk];
}

// ---------------------------------------------------------------------------
// cn_propagate_kernel_avx1<symflag> (propagate.hpp:3-98), one lane per
// rotated channel, ifet built as in cn_hl_propagate1/2 (propagate.hpp:343,400).
template <int P, int SYM>
__global__ __launch_bounds__(64) void k_cn_onthefly(const double* __restrict__ smat, const double* __restrict__ rmat,
                                                    const double* __restrict__ eig, double fzdt, cplx* __restrict__ zall,
                                                    cplx* __restrict__ yall, cplx* __restrict__ uall, int nch, int N) {
    const int iml = blockIdx.x * blockDim.x + threadIdx.x;
    if (iml >= nch) return;
    constexpr int nd = P, ne = P + 1;
    cplx ifet;
    if (SYM == 0) {
        cplx ifdt = {0.0, fzdt / 2};
        ifet = ifdt * eig[iml];
    } else {
        double ifdt = -fzdt / 2;
        ifet = {eig[iml] * ifdt, 0.0};
    }
    cplx* z = zall + (int64_t)iml * N;
    cplx* y = yall + (int64_t)iml * N;
    cplx* u = uall + (int64_t)iml * N * P;
    cplx a[2 * nd + 1], b[2 * nd + 1];
    for (int i = 0; i < N; ++i) {
#pragma unroll
        for (int j = 1; j < ne; ++j) {  // lower panel
            const int k = i * ne - j * nd;
            const int jj = nd - j;
            if (k < 0) {
                a[jj] = {0.0, 0.0};
                b[jj] = {0.0, 0.0};
            } else {
                const double s = smat[k], r = rmat[k];
                const cplx fr = ifet * r;
                if (SYM == 1) {
                    a[jj] = cplx{s + fr.re, fr.im};
                    b[jj] = cplx{s - fr.re, -fr.im};
                } else {
                    a[jj] = cplx{s - fr.re, -fr.im};
                    b[jj] = cplx{s + fr.re, fr.im};
                }
            }
        }
#pragma unroll
        for (int j = 0; j < ne; ++j) {  // upper
            const int k = i * ne + j;
            const double s = smat[k], r = rmat[k];
            const cplx fr = ifet * r;
            a[nd + j] = cplx{s - fr.re, -fr.im};
            b[nd + j] = cplx{s + fr.re, fr.im};
        }
        const int l = nd > i ? nd - i : 0;
        const int h = N - 1 - nd > i ? 2 * nd + 1 : nd + N - i;
#pragma unroll
        for (int j = 0; j < nd; ++j) {
            if (j >= l) {
#pragma unroll
                for (int k = 0; k < nd; ++k) b[j + k + 1] -= b[j] * u[(int64_t)(i + j - nd) * P + k];
            }
        }
#pragma unroll
        for (int j = 0; j < nd; ++j) u[(int64_t)i * P + j] = b[j + nd + 1] / b[nd];
        cplx v = {0.0, 0.0};
#pragma unroll
        for (int j = 0; j < 2 * nd + 1; ++j)
            if (j >= l && j < h) v += a[j] * z[i + j - nd];
#pragma unroll
        for (int j = 0; j < nd; ++j)
            if (j >= l) v -= b[j] * y[i + j - nd];
        y[i] = v / b[nd];
    }
    for (int i = N; i-- > 0;) {
        const int jmax = ne < N - i ? ne : N - i;
        cplx v = {0.0, 0.0};
#pragma unroll
        for (int j = 1; j < ne; ++j)
            if (j < jmax) v += y[i + j] * u[(int64_t)i * P + j - 1];
        const cplx yi = y[i] - v;
        y[i] = yi;
        z[i] = yi;
    }
}

// ---------------------------------------------------------------------------
// cn_hc_propagate0 (propagate.hpp:311-331): w = upper_l c (symb_mul_vecd
// loop path, blas2.hpp:487-507), w = L^-1 w (ltrb_ldl_vecd, blas2.hpp:1058-1069),
// c = (D L^T)^-1 w (utrb_ldl_vecd, blas2.hpp:1082-1095).  One lane per channel.
template <int P>
__global__ __launch_bounds__(64) void k_hc_free(const cplx* __restrict__ upper, const cplx* __restrict__ lower,
                                                cplx* __restrict__ coef, cplx* __restrict__ wall, int M, int L, int m0,
                                                int N) {
    const int iml = blockIdx.x * blockDim.x + threadIdx.x;
    if (iml >= M * L) return;
    constexpr int ne = P + 1;
    const int im = iml / L, il = iml % L;
    const int mabs = (im + m0) < 0 ? -(im + m0) : (im + m0);
    const int l = il + mabs;
    const cplx* U = upper + (int64_t)l * N * ne;
    const cplx* F = lower + (int64_t)l * N * ne;
    cplx* c = coef + (int64_t)iml * N;
    cplx* w = wall + (int64_t)iml * N;
    for (int i = 0; i < N; ++i) {
        cplx t = {0.0, 0.0};
#pragma unroll
        for (int j = 1; j < ne; ++j)
            if (j <= i) t += U[(int64_t)(i - j) * ne + j] * c[i - j];
#pragma unroll
        for (int k = 0; k < ne; ++k)
            if (i + k < N) t += U[(int64_t)i * ne + k] * c[i + k];
        w[i] = t;
    }
    for (int i = 0; i < N; ++i) {
        cplx x = w[i];
#pragma unroll
        for (int j = ne - 1; j >= 1; --j)
            if (j <= i) x -= w[i - j] * F[(int64_t)(i - j) * ne + j];
        w[i] = x;
    }
    for (int i = N; i-- > 0;) {
        cplx x = w[i] / F[(int64_t)i * ne];
#pragma unroll
        for (int k = 1; k < ne; ++k)
            if (i + k < N) x -= c[i + k] * F[(int64_t)i * ne + k];
        c[i] = x;
    }
}

// ---------------------------------------------------------------------------
// absorber_mask::operator() (absorber.hpp:52-64): t = M w; t = L^-1 t
// (ltrb_inv_vecd); w = L^-T t (utrb_inv_vecd), L = chol(S).  One lane per channel.
template <int P>
__global__ __launch_bounds__(64) void k_mask(const double* __restrict__ mask, const double* __restrict__ chol,
                                             cplx* __restrict__ coef, cplx* __restrict__ tall, int nch, int N) {
    const int iml = blockIdx.x * blockDim.x + threadIdx.x;
    if (iml >= nch) return;
    constexpr int ne = P + 1;
    cplx* w = coef + (int64_t)iml * N;
    cplx* t = tall + (int64_t)iml * N;
    for (int i = 0; i < N; ++i) {
        cplx s = {0.0, 0.0};
#pragma unroll
        for (int j = 1; j < ne; ++j)
            if (j <= i) s += mask[(int64_t)(i - j) * ne + j] * w[i - j];
#pragma unroll
        for (int k = 0; k < ne; ++k)
            if (i + k < N) s += mask[(int64_t)i * ne + k] * w[i + k];
        t[i] = s;
    }
    for (int i = 0; i < N; ++i) {
        cplx x = t[i];
#pragma unroll
        for (int j = 1; j < ne; ++j)
            if (j <= i) x -= t[i - j] * chol[(int64_t)(i - j) * ne + j];
        t[i] = x / chol[(int64_t)i * ne];
    }
    for (int i = N; i-- > 0;) {
        cplx x = t[i];
#pragma unroll
        for (int k = 1; k < ne; ++k)
            if (i + k < N) x -= w[i + k] * chol[(int64_t)i * ne + k];
        w[i] = x / chol[(int64_t)i * ne];
    }
}

// ---------------------------------------------------------------------------
// bilinear forms x^H A y with a real symmetric band (symb_prj_vecd,
// blas2.hpp:1110-1127); one block per channel, fixed-shape tree reduction.
constexpr int kRedThreads = 256;
__device__ inline cplx band_row(const double* A, const cplx* y, int i, int N, int ne) {
    cplx t = {0.0, 0.0};
    for (int j = 1; j < ne; ++j)
        if (j <= i) t += A[(int64_t)(i - j) * ne + j] * y[i - j];
    for (int k = 0; k < ne; ++k)
        if (i + k < N) t += A[(int64_t)i * ne + k] * y[i + k];
    return t;
}
__device__ inline double block_sum(double v) {
    __shared__ double sh[kRedThreads];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = kRedThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    return sh[0];
}

__global__ __launch_bounds__(kRedThreads) void k_norm(const double* __restrict__ S, const cplx* __restrict__ coef,
                                                      double* __restrict__ red, int N, int ne) {
    const int iml = blockIdx.x;
    const cplx* x = coef + (int64_t)iml * N;
    double acc = 0.0;
    for (int i = threadIdx.x; i < N; i += blockDim.x) acc += mulc(band_row(S, x, i, N, ne), x[i]).re;
    double s = block_sum(acc);
    if (threadIdx.x == 0) red[iml] = s;
}

// observable_z::observe (observable/1e.hpp:26-49), real part per channel
__global__ __launch_bounds__(kRedThreads) void k_dipole_z(const double* __restrict__ R, const double* __restrict__ y10,
                                                          const cplx* __restrict__ coef, double* __restrict__ red,
                                                          int L, int N, int ne) {
    const int iml = blockIdx.x;
    const int im = iml / L, il = iml % L;
    const cplx* lhs = coef + (int64_t)iml * N;
    const cplx* base = coef + (int64_t)im * L * N;
    double acc = 0.0;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        if (il > 0) acc += mulc(band_row(R, base + (int64_t)(il - 1) * N, i, N, ne), lhs[i]).re * y10[im * L + il - 1];
        if (il + 1 < L) acc += mulc(band_row(R, base + (int64_t)(il + 1) * N, i, N, ne), lhs[i]).re * y10[im * L + il];
    }
    double s = block_sum(acc);
    if (threadIdx.x == 0) red[iml] = s;
}

template <template <int> class F, class... Args>
void dispatch_rank(int64_t p, Args&&... args) {
    switch (p) {
        case 1: F<1>::run(args...); break;
        case 2: F<2>::run(args...); break;
        case 3: F<3>::run(args...); break;
        case 4: F<4>::run(args...); break;
        case 5: F<5>::run(args...); break;
        case 6: F<6>::run(args...); break;
        case 7: F<7>::run(args...); break;
        case 8: F<8>::run(args...); break;
        default: break;
    }
}

template <int P>
struct RunCN {
    static void run(hipStream_t s, const double* smat, const double* rmat, const double* eig, int symflag, double fzdt,
                    cplx* z, cplx* y, cplx* u, int nch, int N) {
        dim3 grid((nch + 63) / 64), block(64);
        if (symflag == 0)
            hipLaunchKernelGGL((k_cn_onthefly<P, 0>), grid, block, 0, s, smat, rmat, eig, fzdt, z, y, u, nch, N);
        else
            hipLaunchKernelGGL((k_cn_onthefly<P, 1>), grid, block, 0, s, smat, rmat, eig, fzdt, z, y, u, nch, N);
    }
};
template <int P>
struct RunHC {
    static void run(hipStream_t s, const cplx* up, const cplx* lo, cplx* coef, cplx* w, int M, int L, int m0, int N) {
        dim3 grid((M * L + 63) / 64), block(64);
        hipLaunchKernelGGL((k_hc_free<P>), grid, block, 0, s, up, lo, coef, w, M, L, m0, N);
    }
};
template <int P>
struct RunMask {
    static void run(hipStream_t s, const double* mask, const double* chol, cplx* coef, cplx* t, int nch, int N) {
        dim3 grid((nch + 63) / 64), block(64);
        hipLaunchKernelGGL((k_mask<P>), grid, block, 0, s, mask, chol, coef, t, nch, N);
    }
};

}  // namespace

void launch_rot_real(hipStream_t s, const double* P, const double* src, double* dst, int64_t M, int64_t L, int64_t X,
                     int trans) {
    dim3 grid((unsigned)((X + kRotThreads - 1) / kRotThreads), (unsigned)((L + kRotKT - 1) / kRotKT), (unsigned)M);
    size_t lds = sizeof(double) * kRotKT * L;
    hipLaunchKernelGGL(k_rot_real, grid, dim3(kRotThreads), lds, s, P, src, dst, (int)L, X, trans);
}

void launch_rot_cplx(hipStream_t s, const cplx* P, const cplx* src, cplx* dst, int64_t M, int64_t L, int64_t N,
                     int trans) {
    dim3 grid((unsigned)((N + kRotThreads - 1) / kRotThreads), (unsigned)((L + kRotKTc - 1) / kRotKTc), (unsigned)M);
    size_t lds = sizeof(cplx) * kRotKTc * L;
    hipLaunchKernelGGL(k_rot_cplx, grid, dim3(kRotThreads), lds, s, P, src, dst, (int)L, N, trans);
}

void launch_cn_onthefly(hipStream_t s, const double* smat, const double* rmat, const double* eig, int symflag,
                        double fzdt, cplx* z, cplx* ytmp, cplx* utmp, int64_t nch, int64_t N, int64_t p) {
    dispatch_rank<RunCN>(p, s, smat, rmat, eig, symflag, fzdt, z, ytmp, utmp, (int)nch, (int)N);
}

void launch_hc_free(hipStream_t s, const cplx* upper, const cplx* lower, cplx* coef, cplx* w, int64_t M, int64_t L,
                    int64_t m0, int64_t N, int64_t p) {
    dispatch_rank<RunHC>(p, s, upper, lower, coef, w, (int)M, (int)L, (int)m0, (int)N);
}

void launch_mask(hipStream_t s, const double* mask, const double* chol, cplx* coef, cplx* t, int64_t nch, int64_t N,
                 int64_t p) {
    dispatch_rank<RunMask>(p, s, mask, chol, coef, t, (int)nch, (int)N);
}

void launch_norm(hipStream_t s, const double* S, const cplx* coef, double* red, int64_t nch, int64_t N, int64_t p) {
    hipLaunchKernelGGL(k_norm, dim3((unsigned)nch), dim3(kRedThreads), 0, s, S, coef, red, (int)N, (int)(p + 1));
}

void launch_dipole_z(hipStream_t s, const double* rband, const double* y10, const cplx* coef, double* red, int64_t M,
                     int64_t L, int64_t N, int64_t p) {
    hipLaunchKernelGGL(k_dipole_z, dim3((unsigned)(M * L)), dim3(kRedThreads), 0, s, rband, y10, coef, red, (int)L,
                       (int)N, (int)(p + 1));
}

}  // namespace qpc
