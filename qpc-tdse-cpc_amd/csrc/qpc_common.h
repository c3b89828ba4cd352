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
// The arithmetic of the field-free set-up (cn_component_hc::initialize): the
// H_l assembly of integrate_hc (basisfunc.hpp:92-106) and one row of lu_zsb's
// complex symmetric LDL^T (algorithm/ed.h:208-289), in real operations that
// are not contracted into FMAs, so that the host and the device set-up give
// the same bits.
__host__ __device__ inline double hc_entry(double v, double cent, double k, double c2m, double t) {
#pragma clang fp contract(off)
    return v + cent * k + c2m * t;
}
// f = l d (the multiplier of row j-k)
__host__ __device__ inline cplx ldlt_mul(cplx l, cplx d) {
#pragma clang fp contract(off)
    return {l.re * d.re - l.im * d.im, l.re * d.im + l.im * d.re};
}
// a -= f x
__host__ __device__ inline void ldlt_sub(cplx& a, cplx f, cplx x) {
#pragma clang fp contract(off)
    const double tr = f.re * x.re - f.im * x.im, ti = f.re * x.im + f.im * x.re;
    a.re = a.re - tr;
    a.im = a.im - ti;
}
// 1 / d
__host__ __device__ inline cplx ldlt_recip(cplx y) {
#pragma clang fp contract(off)
    const double d = y.re * y.re + y.im * y.im;
    return {y.re / d, -y.im / d};
}
__host__ __device__ inline double ldlt_norm2(cplx y) {
#pragma clang fp contract(off)
    return y.re * y.re + y.im * y.im;
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
