// qpc_common.h -- shared device/host helpers of the MI355X propagator.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace qpc {

// complex FP64 stored as interleaved {re,im} (16 B), the reference's layout
// (libraries/support_avx.h:59,146).  Quotients follow the reference's
// avx_div: (x*conj(y))/|y|^2 (libraries/avx/avx128_complex.h:75-85).
struct cplx {
    double re, im;
};

__host__ __device__ inline cplx cmake(double r, double i) { return cplx{r, i}; }
__host__ __device__ inline cplx operator+(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
__host__ __device__ inline cplx operator-(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
__host__ __device__ inline cplx operator*(cplx a, cplx b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ inline cplx operator*(cplx a, double b) { return {a.re * b, a.im * b}; }
__host__ __device__ inline cplx operator*(double a, cplx b) { return {a * b.re, a * b.im}; }
__host__ __device__ inline cplx operator/(cplx x, cplx y) {
    double d = y.re * y.re + y.im * y.im;
    return {(y.re * x.re + y.im * x.im) / d, (y.re * x.im - y.im * x.re) / d};
}
__host__ __device__ inline cplx operator/(cplx x, double y) { return {x.re / y, x.im / y}; }
__host__ __device__ inline cplx crecip(cplx y) {
    double d = y.re * y.re + y.im * y.im;
    return {y.re / d, -y.im / d};
}
__host__ __device__ inline cplx& operator+=(cplx& a, cplx b) { a = a + b; return a; }
__host__ __device__ inline cplx& operator-=(cplx& a, cplx b) { a = a - b; return a; }
__host__ __device__ inline cplx conjc(cplx a) { return {a.re, -a.im}; }
// x * conj(y)
__host__ __device__ inline cplx mulc(cplx x, cplx y) {
    return {y.re * x.re + y.im * x.im, y.re * x.im - y.im * x.re};
}

}  // namespace qpc

#define QPC_HIP_CHECK(ctx, call)                                                        \
    do {                                                                                \
        hipError_t _e = (call);                                                         \
        if (_e != hipSuccess) {                                                         \
            (ctx)->set_error(std::string(#call) + ": " + hipGetErrorString(_e));        \
            return QPC_E_HIP;                                                           \
        }                                                                               \
    } while (0)
