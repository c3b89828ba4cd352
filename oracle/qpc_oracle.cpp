// qpc_oracle.cpp -- CPU oracle for the QPC-TDSE Crank-Nicolson propagator.
//
// TEST INFRASTRUCTURE ONLY: a C++17/OpenMP restatement of the reference
// algorithm (zzh535897/QPC-TDSE-CPC), used as the parity checker and as the
// timed CPU baseline.  The product library never links or calls it.
// Citations are file:line in the reference repository.
//
// Pins (see tests/): Gauss-Legendre vs numpy fixtures; B-spline partition of
// unity and polynomial exactness; S/H/R/D bands vs dense numpy integrals;
// hydrogen spectrum -1/(2n^2); lu_zsb reconstruction; CN kernels vs dense
// numpy solves; Y10 eigenvalues = Gauss-Legendre nodes (Golub-Welsch);
// P^H diag(eig) P = angular operator; Cayley phase of eigenstates; unitarity;
// time reversal; dense brute-force Strang product at small sizes.
#include "qpc_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------
// complex FP64 with the reference's arithmetic (avx128_complex.h:31-98):
// product a*b, quotient (x*conj(y))/|y|^2, interleaved {re,im}.
struct cplx {
    double re, im;
};
static inline cplx operator+(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
static inline cplx operator-(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
static inline cplx operator*(cplx a, cplx b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
static inline cplx operator*(cplx a, double b) { return {a.re * b, a.im * b}; }
static inline cplx operator*(double a, cplx b) { return {a * b.re, a * b.im}; }
static inline cplx operator/(cplx x, cplx y) {
    double d = y.re * y.re + y.im * y.im;
    return {(y.re * x.re + y.im * x.im) / d, (y.re * x.im - y.im * x.re) / d};
}
static inline cplx operator/(cplx x, double y) { return {x.re / y, x.im / y}; }
static inline cplx operator/(double x, cplx y) {
    double d = y.re * y.re + y.im * y.im;
    return {x * y.re / d, -x * y.im / d};
}
static inline cplx& operator+=(cplx& a, cplx b) { a = a + b; return a; }
static inline cplx& operator-=(cplx& a, cplx b) { a = a - b; return a; }
static inline cplx& operator*=(cplx& a, cplx b) { a = a * b; return a; }
static inline cplx& operator/=(cplx& a, cplx b) { a = a / b; return a; }
static inline double cnorm(cplx a) { return a.re * a.re + a.im * a.im; }
static inline cplx conjc(cplx a) { return {a.re, -a.im}; }
// x * conj(y)  (avx_mulc, avx128_complex.h:55-62)
static inline cplx mulc(cplx x, cplx y) {
    return {y.re * x.re + y.im * x.im, y.re * x.im - y.im * x.re};
}
static const cplx czero = {0.0, 0.0};

static constexpr double kPi = 3.14159265358979323846;  // utilities/const.h:5

// ---------------------------------------------------------------------------
// Gauss-Legendre nodes: textbook Newton iteration on P_n in long double
// (stands in for gsl_integration_glfixed, basis_bsplines.h:377-413).
// xpos/wpos receive the n/2 (or (n+1)/2) non-negative nodes ascending.
static void gl_half(int n, std::vector<long double>& xpos, std::vector<long double>& wpos) {
    int h = (n + 1) / 2;
    xpos.assign(h, 0.0L);
    wpos.assign(h, 0.0L);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int k = 0; k < h; ++k) {
        long double z = cosl(pi * (k + 0.75L) / (n + 0.5L));
        long double pp = 1.0L;
        for (int it = 0; it < 200; ++it) {
            long double p0 = 1.0L, p1 = z;
            for (int j = 2; j <= n; ++j) {
                long double p2 = ((2 * j - 1) * z * p1 - (j - 1) * p0) / j;
                p0 = p1;
                p1 = p2;
            }
            pp = n * (z * p1 - p0) / (z * z - 1.0L);
            long double dz = p1 / pp;
            z -= dz;
            if (fabsl(dz) < 1e-20L) break;
        }
        {
            long double p0 = 1.0L, p1 = z;
            for (int j = 2; j <= n; ++j) {
                long double p2 = ((2 * j - 1) * z * p1 - (j - 1) * p0) / j;
                p0 = p1;
                p1 = p2;
            }
            pp = n * (z * p1 - p0) / (z * z - 1.0L);
        }
        // k-th root counted from the top; store ascending among non-negatives
        xpos[h - 1 - k] = z;
        wpos[h - 1 - k] = 2.0L / ((1.0L - z * z) * pp * pp);
    }
    if (n % 2 == 1) xpos[0] = 0.0L;
}

// point i (ascending) on [a,b], the unpacking of gsl_integration_glfixed_point
static void gl_point(int n, const std::vector<long double>& xp, const std::vector<long double>& wp,
                     double a, double b, int i, double* xi, double* wi) {
    const double A = (b - a) / 2;
    const double B = (a + b) / 2;
    if (n % 2 == 1) {
        int k = i - n / 2;
        int s = k < 0 ? -1 : 1;
        *xi = B + s * A * (double)xp[s * k];
        *wi = A * (double)wp[s * k];
    } else if (i < n / 2) {
        int j = n / 2 - 1 - i;
        *xi = B - A * (double)xp[j];
        *wi = A * (double)wp[j];
    } else {
        int j = i - n / 2;
        *xi = B + A * (double)xp[j];
        *wi = A * (double)wp[j];
    }
}

// ---------------------------------------------------------------------------
// B-spline axis: bsplines<N,p,0,0>::axis (basis_bsplines.h:107-527).
struct Axis {
    long N = 0, p = 0, q = 0;           // q = n_glin = 2p+2 (:110)
    std::vector<double> kn;              // kn[i+1] = knot(i), i = -1..N+p+1
    std::vector<double> gx, gw;          // [s*q+k], active intervals s = 0..N+p-1
    std::vector<double> v0, v1;          // [(s*(p+1)+t)*q+k], spline s-p+t

    double knot(long i) const { return kn[i + 1]; }

    // cal<p,m> (basis_bsplines.h:234-260), the same recursion
    double cal(long pp, long m, long i, double x) const {
        if (m == 0) {
            if (pp == 0) return (x >= knot(i) && x < knot(i + 1)) ? 1.0 : 0.0;
            double wl = knot(i + pp) - knot(i);
            double wr = knot(i + pp + 1) - knot(i + 1);
            double xl = wl == 0. ? 0. : (x - knot(i)) / wl * cal(pp - 1, 0, i, x);
            double xr = wr == 0. ? 0. : (x - knot(i + 1 + pp)) / wr * cal(pp - 1, 0, i + 1, x);
            return xl - xr;
        }
        if (pp == 0) return 0.0;
        double wl = knot(i + pp) - knot(i);
        double wr = knot(i + pp + 1) - knot(i + 1);
        double xl = wl == 0. ? 0. : cal(pp - 1, m - 1, i, x) / wl;
        double xr = wr == 0. ? 0. : cal(pp - 1, m - 1, i + 1, x) / wr;
        return (xl - xr) * pp;
    }

    // initialize_axis_linear_step (basis_bsplines.h:450-459) + initialize (:375-414)
    void init_linear(long N_, long p_, double xmin, double xmax) {
        N = N_;
        p = p_;
        q = 2 * p + 2;
        const long n_knot = N + p + 3;  // n_lfbd = n_rtbd = 1 (:111-113)
        kn.assign(n_knot, 0.0);
        double step = (xmax - xmin) / (double)(N + 2 - p);
        for (long i = 0; i < p; ++i) kn[i] = xmin;                                  // heads
        for (long i = p; i < N + 2; ++i) kn[i] = xmin + (double)(size_t)(i - p) * step;  // elses
        for (long i = N + 2; i < n_knot; ++i) kn[i] = xmax;                         // tails
        std::vector<long double> xp, wp;
        gl_half((int)q, xp, wp);
        const long ns = N + p;
        gx.assign(ns * q, 0.0);
        gw.assign(ns * q, 0.0);
        v0.assign(ns * (p + 1) * q, 0.0);
        v1.assign(ns * (p + 1) * q, 0.0);
        for (long s = 0; s < ns; ++s) {
            double xl = knot(s), xr = knot(s + 1);
            if (xl == xr) {
                for (long k = 0; k < q; ++k) gx[s * q + k] = xl;
                continue;  // zero weight and zero values
            }
            for (long k = 0; k < q; ++k) {
                double xi, wi;
                gl_point((int)q, xp, wp, xl, xr, (int)k, &xi, &wi);
                gx[s * q + k] = xi;
                gw[s * q + k] = wi;
                for (long t = 0; t <= p; ++t) {
                    long idx = s - t;  // spline index held at slot p-t
                    if (idx < 0 || idx >= N) continue;  // dropped end splines
                    v0[(s * (p + 1) + (p - t)) * q + k] = cal(p, 0, idx, xi);
                    v1[(s * (p + 1) + (p - t)) * q + k] = cal(p, 1, idx, xi);
                }
            }
        }
    }

    // integrate<mi,mj> (basis_bsplines.h:262-289)
    template <class F>
    double integrate(int mi, int mj, const F& f, long i, long j) const {
        double retn = 0.;
        if (i >= N || j >= N) return retn;
        long imin = std::max(i, j);
        long imax = std::min(i, j) + p;
        const std::vector<double>& lhs = mi ? v1 : v0;
        const std::vector<double>& rhs = mj ? v1 : v0;
        for (long s = imin; s <= imax; ++s) {
            if (gw[s * q] == 0.0) continue;  // zero-length interval
            const double* l = &lhs[(s * (p + 1) + (p + i - s)) * q];
            const double* r = &rhs[(s * (p + 1) + (p + j - s)) * q];
            for (long k = 0; k < q; ++k)
                retn += gw[s * q + k] * f(gx[s * q + k]) * l[k] * r[k];
        }
        return retn;
    }
};

// integrate_0r0 / integrate_0r1 (basisfunc.hpp:63-81): band[ir*ne+ic]
template <class F>
static void band_0r0(const Axis& ax, std::vector<double>& out, const F& f) {
    long N = ax.N, ne = ax.p + 1;
    out.assign(N * ne, 0.0);
    for (long ir = 0; ir < N; ++ir)
        for (long ic = 0; ic < ne; ++ic) out[ir * ne + ic] = ax.integrate(0, 0, f, ir, ir + ic);
}
template <class F>
static void band_0r1(const Axis& ax, std::vector<double>& out, const F& f) {
    long N = ax.N, ne = ax.p + 1;
    out.assign(N * ne, 0.0);
    for (long ir = 0; ir < N; ++ir)
        for (long ic = 0; ic < ne; ++ic) out[ir * ne + ic] = ax.integrate(0, 1, f, ir, ir + ic);
}

// ---------------------------------------------------------------------------
// Banded kernels restated from blas2.hpp (generic loop paths, :370-447, :452-508, :1006-1109)

// symb_mul_vecd_loop (blas2.hpp:370-377) with m = ne, q = m-1
template <class A, class X>
static inline cplx symb_loop(const A* a, const X* x, long q, long jmin, long jmax, long kmin, long kmax) {
    cplx t = czero;
    for (long j = jmin; j < jmax; ++j) t += a[-j * q] * x[-j];
    for (long k = kmin; k < kmax; ++k) t += a[k] * x[k];
    return t;
}
// symb_mul_vecd<n,m,0> loop branch (blas2.hpp:487-507); requires n >= 2m
template <class A>
static void symb_mul(long n, long m, const A* a, const cplx* x, cplx* y) {
    long q = m - 1;
    for (long i = 0; i < m; ++i) y[i] = symb_loop(a + i * m, x + i, q, 1, i + 1, 0, m);
    for (long i = m; i < n - m; ++i) y[i] = symb_loop(a + i * m, x + i, q, 1, m, 0, m);
    for (long i = n - m; i < n; ++i) y[i] = symb_loop(a + i * m, x + i, q, 1, m, 0, n - i);
}
// ltrb_ldl_vecd (blas2.hpp:1058-1069; kernel :434-440)
static void ltrb_ldl(long n, long m, const cplx* a, const cplx* x, cplx* y) {
    long q = m - 1;
    auto loop = [&](const cplx* ai, cplx xv, cplx* yi, long jmin, long jmax) {
        for (long j = jmax; j-- > jmin;) xv -= yi[-j] * ai[-j * q];
        *yi = xv;
    };
    for (long i = 0; i < m; ++i) loop(a + i * m, x[i], y + i, 1, i + 1);
    for (long i = m; i < n; ++i) loop(a + i * m, x[i], y + i, 1, m);
}
// utrb_ldl_vecd (blas2.hpp:1082-1095; kernel :441-447)
static void utrb_ldl(long n, long m, const cplx* a, const cplx* x, cplx* y) {
    auto loop = [&](const cplx* ai, cplx xv, cplx* yi, long kmin, long kmax) {
        xv /= *ai;
        for (long k = kmin; k < kmax; ++k) xv -= yi[k] * ai[k];
        *yi = xv;
    };
    for (long i = 0; i < m; ++i) {
        long s = n - 1 - i;
        loop(a + s * m, x[s], y + s, 1, i + 1);
    }
    for (long i = m; i < n; ++i) {
        long s = n - 1 - i;
        loop(a + s * m, x[s], y + s, 1, m);
    }
}
// ltrb_inv_vecd (blas2.hpp:1006-1017; kernel :421-427) with a real band
static void ltrb_inv(long n, long m, const double* a, const cplx* x, cplx* y) {
    long q = m - 1;
    auto loop = [&](const double* ai, cplx xv, cplx* yi, long jmin, long jmax) {
        for (long j = jmin; j < jmax; ++j) xv -= yi[-j] * ai[-j * q];
        *yi = xv / *ai;
    };
    for (long i = 0; i < m; ++i) loop(a + i * m, x[i], y + i, 1, i + 1);
    for (long i = m; i < n; ++i) loop(a + i * m, x[i], y + i, 1, m);
}
// utrb_inv_vecd (blas2.hpp:1030-1043; kernel :428-433) with a real band
static void utrb_inv(long n, long m, const double* a, const cplx* x, cplx* y) {
    auto loop = [&](const double* ai, cplx xv, cplx* yi, long kmin, long kmax) {
        for (long k = kmin; k < kmax; ++k) xv -= yi[k] * ai[k];
        *yi = xv / *ai;
    };
    for (long i = 0; i < m; ++i) {
        long s = n - 1 - i;
        loop(a + s * m, x[s], y + s, 1, i + 1);
    }
    for (long i = m; i < n; ++i) {
        long s = n - 1 - i;
        loop(a + s * m, x[s], y + s, 1, m);
    }
}
// symb_prj_vecd_naive (blas2.hpp:1110-1127): x^H A y, A real symmetric band
static cplx symb_prj(long n, long m, const double* a, const cplx* x, const cplx* y) {
    long q = m - 1;
    cplx tmp = czero;
    for (long i = 0; i < m; ++i) tmp += mulc(symb_loop(a + i * m, y + i, q, 1, i + 1, 0, m), x[i]);
    for (long i = m; i < n - m; ++i) tmp += mulc(symb_loop(a + i * m, y + i, q, 1, m, 0, m), x[i]);
    for (long i = n - m; i < n; ++i) tmp += mulc(symb_loop(a + i * m, y + i, q, 1, m, 0, n - i), x[i]);
    return tmp;
}

// impl_ed_simd::lu_zsb (algorithm/ed.h:208-289), generic loop_kernel (:249-267)
static int lu_zsb(long n, long p, cplx* a) {
    int info = 0;
    const long ne = p + 1;
    auto kernel = [&](cplx* ai, long kmax) {
        for (long k = 1; k < kmax; ++k) {
            cplx* as = ai - p * k;
            cplx aa = as[0] * as[-k];
            for (long i = 0; i < ne - k; ++i) ai[i] -= aa * as[i];
        }
        if (cnorm(*ai) < 1e-31) info++;
        cplx a0 = 1. / (*ai);
        for (long i = 1; i < ne; ++i) ai[i] *= a0;
    };
    for (long j = 0; j < ne; ++j) kernel(a + j * ne, 1 + j);
    for (long j = ne; j < n - p; ++j) kernel(a + j * ne, ne);
    for (long j = n - p; j < n; ++j) {
        for (long i = n - j; i < ne; ++i) a[j * ne + i] = czero;
        kernel(a + j * ne, ne);
    }
    return info;
}

// banded Cholesky A = L L^T (LAPACKE_dpbtrf 'L' col-major on the reference's
// row-major upper band, ed.h:182-194), then the trailing clear (:186-192)
static int chol_band(long n, long p, double* a) {
    const long ne = p + 1;
    for (long j = 0; j < n; ++j) {
        double ajj = a[j * ne];
        if (!(ajj > 0.0)) return (int)(j + 1);
        ajj = std::sqrt(ajj);
        a[j * ne] = ajj;
        long kn = std::min(p, n - 1 - j);
        for (long k = 1; k <= kn; ++k) a[j * ne + k] /= ajj;
        for (long b = 1; b <= kn; ++b)
            for (long c = b; c <= kn; ++c) a[(j + b) * ne + (c - b)] -= a[j * ne + c] * a[j * ne + b];
    }
    for (long j = 0; j < n; ++j)
        for (long k = 0; k < ne; ++k)
            if (j + k >= n) a[j * ne + k] = 0.0;
    return 0;
}

// cn_propagate_kernel_avx1<symflag> (propagate.hpp:3-98).  ifet is real for
// symflag 1 and complex for symflag 0; both carried as cplx here.
static void cn_kernel(long n, long p, int symflag, const double* smat, const double* rmat, cplx ifet,
                      cplx* zsrc, cplx* ytmp, cplx* utmp) {
    const long nd = p, ne = p + 1;
    std::vector<cplx> av(2 * nd + 1), bv(2 * nd + 1);
    cplx* a = av.data();
    cplx* b = bv.data();
    auto u = [&](long i, long j) -> cplx& { return utmp[i * nd + j]; };
    for (long i = 0; i < n; ++i) {
        for (long j = 1; j < ne; ++j) {  // lower panel
            long k = i * ne - j * nd;
            long jj = nd - j;
            if (k < 0) {
                a[jj] = czero;
                b[jj] = czero;
            } else {
                double s = smat[k], r = rmat[k];
                if (symflag == 1) {
                    a[jj] = cplx{s, 0} + ifet * r;
                    b[jj] = cplx{s, 0} - ifet * r;
                } else {
                    a[jj] = cplx{s, 0} - ifet * r;
                    b[jj] = cplx{s, 0} + ifet * r;
                }
            }
        }
        for (long j = 0; j < ne; ++j) {  // upper
            long k = i * ne + j;
            double s = smat[k], r = rmat[k];
            a[nd + j] = cplx{s, 0} - ifet * r;
            b[nd + j] = cplx{s, 0} + ifet * r;
        }
        long l = nd > i ? nd - i : 0;
        long h = n - 1 - nd > i ? 2 * nd + 1 : nd + n - i;
        for (long j = l; j < nd; ++j)
            for (long k = 0; k < nd; ++k) b[j + k + 1] -= b[j] * u(i + j - nd, k);
        for (long j = 0; j < nd; ++j) u(i, j) = b[j + nd + 1] / b[nd];
        cplx v = czero;
        for (long j = l; j < h; ++j) v += a[j] * zsrc[i + j - nd];
        for (long j = l; j < nd; ++j) v -= b[j] * ytmp[i + j - nd];
        ytmp[i] = v / b[nd];
    }
    for (long i = n; i-- > 0;) {
        long jmax = std::min(ne, n - i);
        cplx v = czero;
        for (long j = 1; j < jmax; ++j) v += ytmp[i + j] * u(i, j - 1);
        ytmp[i] -= v;
        zsrc[i] = ytmp[i];
    }
}

// ---------------------------------------------------------------------------
// real symmetric tridiagonal eigen-decomposition: implicit QL with Wilkinson
// shifts (textbook tql2), eigenvectors accumulated; sorted ascending.
// z row-major z[k*n+i] = component i of eigenvector k.
static int tridiag_eig(long n, const double* diag, const double* off, double* eig, double* z) {
    std::vector<double> d(diag, diag + n), e(n, 0.0);
    for (long i = 0; i + 1 < n; ++i) e[i] = off[i];
    std::vector<double> Z(n * n, 0.0);  // column-major accumulator Z[i + k*n]
    for (long i = 0; i < n; ++i) Z[i + i * n] = 1.0;
    for (long l = 0; l < n; ++l) {
        int iter = 0;
        long m;
        do {
            for (m = l; m < n - 1; ++m) {
                double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= 1e-300 + 2.2e-16 * 0.5 * dd) break;
            }
            if (m != l) {
                if (iter++ == 200) return -1;
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = std::hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
                double s = 1.0, c = 1.0, pp = 0.0;
                long i;
                for (i = m - 1; i >= l; --i) {
                    double f = s * e[i], b = c * e[i];
                    r = std::hypot(f, g);
                    e[i + 1] = r;
                    if (r == 0.0) {
                        d[i + 1] -= pp;
                        e[m] = 0.0;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - pp;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    pp = s * r;
                    d[i + 1] = g + pp;
                    g = c * r - b;
                    for (long k = 0; k < n; ++k) {
                        double fz = Z[k + (i + 1) * n];
                        Z[k + (i + 1) * n] = s * Z[k + i * n] + c * fz;
                        Z[k + i * n] = c * Z[k + i * n] - s * fz;
                    }
                }
                if (r == 0.0 && i >= l) continue;
                d[l] -= pp;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
    std::vector<long> ord(n);
    for (long i = 0; i < n; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](long a, long b) { return d[a] < d[b]; });
    for (long k = 0; k < n; ++k) {
        eig[k] = d[ord[k]];
        for (long i = 0; i < n; ++i) z[k * n + i] = Z[i + ord[k] * n];
    }
    return 0;
}

// ---------------------------------------------------------------------------
// symmetric banded LDL^T of a real band (no pivoting) used for Sylvester
// inertia counts; returns the number of negative pivots.
static long inertia_count(long n, long p, const double* h, const double* s, double sigma,
                          std::vector<double>& work) {
    const long ne = p + 1;
    work.assign(n * ne, 0.0);
    for (long i = 0; i < n * ne; ++i) work[i] = h[i] - sigma * s[i];
    long neg = 0;
    for (long j = 0; j < n; ++j) {
        double* aj = &work[j * ne];
        long kmax = std::min(j, p);
        for (long k = 1; k <= kmax; ++k) {
            const double* as = aj - p * k;
            double aa = as[0] * as[-k];
            for (long i = 0; i < ne - k; ++i) aj[i] -= aa * as[i];
        }
        if (aj[0] < 0) ++neg;
        double a0 = aj[0] == 0.0 ? 1e300 : 1.0 / aj[0];
        for (long i = 1; i < ne; ++i) aj[i] = (j + i < n) ? aj[i] * a0 : 0.0;
    }
    return neg;
}
// solve (LDL^T) x = b with the factor from inertia_count (real band, real vector)
static void ldl_solve_real(long n, long p, const double* f, std::vector<double>& x) {
    const long ne = p + 1;
    for (long i = 0; i < n; ++i) {
        double v = x[i];
        for (long j = 1; j <= std::min(i, p); ++j) v -= f[(i - j) * ne + j] * x[i - j];
        x[i] = v;
    }
    for (long i = n; i-- > 0;) {
        double v = x[i] / f[i * ne];
        for (long k = 1; k < ne && i + k < n; ++k) v -= f[i * ne + k] * x[i + k];
        x[i] = v;
    }
}
static void band_mul_real(long n, long p, const double* a, const double* x, double* y) {
    const long ne = p + 1;
    for (long i = 0; i < n; ++i) {
        double t = 0.0;
        for (long j = 1; j <= std::min(i, p); ++j) t += a[(i - j) * ne + j] * x[i - j];
        for (long k = 0; k < ne && i + k < n; ++k) t += a[i * ne + k] * x[i + k];
        y[i] = t;
    }
}
// eigenpair `in` (ascending) of H v = E S v: inertia bisection, then inverse
// iteration; v^T S v = 1 (the normalisation of LAPACKE_dsbgv, ed.h:105-112).
static int gen_eig_band(long n, long p, const double* h, const double* s, long in, double* energy,
                        std::vector<double>& v) {
    std::vector<double> work;
    double lo = -1.0, hi = 1.0;
    while (inertia_count(n, p, h, s, lo, work) > in) lo *= 2.0;
    while (inertia_count(n, p, h, s, hi, work) <= in) hi *= 2.0;
    for (int it = 0; it < 200; ++it) {
        double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (inertia_count(n, p, h, s, mid, work) > in) hi = mid;
        else lo = mid;
    }
    double lam = 0.5 * (lo + hi);
    double sigma = lam - 1e-9 * std::max(1.0, std::fabs(lam));
    inertia_count(n, p, h, s, sigma, work);
    v.assign(n, 1.0);
    std::vector<double> t(n);
    for (int it = 0; it < 6; ++it) {
        band_mul_real(n, p, s, v.data(), t.data());
        ldl_solve_real(n, p, work.data(), t);
        band_mul_real(n, p, s, t.data(), v.data());
        double nn = 0.0;
        for (long i = 0; i < n; ++i) nn += t[i] * v[i];
        nn = 1.0 / std::sqrt(nn);
        for (long i = 0; i < n; ++i) v[i] = t[i] * nn;
    }
    // Rayleigh quotient energy
    band_mul_real(n, p, h, v.data(), t.data());
    double e = 0.0;
    for (long i = 0; i < n; ++i) e += v[i] * t[i];
    *energy = e;
    long imax = 0;
    for (long i = 0; i < n; ++i)
        if (std::fabs(v[i]) > std::fabs(v[imax])) imax = i;
    if (v[imax] < 0)
        for (long i = 0; i < n; ++i) v[i] = -v[i];
    return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
struct qor_ctx {
    qor_dims d{};
    long N = 0, p = 0, ne = 0, L = 0, M = 0, m0 = 0, lmax = 0;
    Axis ax;
    std::vector<double> S, chol, H, rinv, rdif, rpow1, mask, dvdr;
    std::vector<cplx> upper, lower;
    std::vector<double> yeig, yprj, qeig, y10, q10;
    std::vector<cplx> qprj, merged;
    std::vector<cplx> tmpc;  // cn_workspace_so::tmpc (workspace.hpp:25)
    // propagator_cn_hc_vgwr angular data (cn_1e.hpp:818-832)
    int ep_isub = -1;
    bool have_ep = false;
    std::vector<double> ep_eig[4], ep_tx1, ep_ty2;  // zeig of x1, x2, y1, y2
    // propagator_cn_hc_lgwr angular data (cn_1e.hpp:745-765): zeig of x1 and
    // y1, tran_x1 (real), x1_2_y1 = U[y1] U[x1]^H and tran_y1 (complex)
    int lr_isub = -1;
    bool have_lgwr = false;
    std::vector<double> lr_ex1, lr_ey1, lr_tx1;
    std::vector<cplx> lr_x1y1, lr_ty1;
    std::vector<cplx> ep_x1x2, ep_x2y1, ep_y1y2;
    int lu_info = 0;

    long in_l(long im, long il) const { return il + std::labs(im + m0); }  // dimension/1e.hpp:109-112
};

static int64_t arr_size(const qor_ctx* c, int what) {
    const long N = c->N, ne = c->ne, L = c->L, M = c->M;
    switch (what) {
        case QOR_S: case QOR_CHOL: case QOR_RINV: case QOR_RDIF: case QOR_RPOW1: case QOR_MASK:
        case QOR_DVDR:
            return N * ne;
        case QOR_H: return (c->lmax + 1) * N * ne;
        case QOR_UPPER: case QOR_LOWER: return 2 * (c->lmax + 1) * N * ne;
        case QOR_YEIG: case QOR_QEIG: case QOR_Y10: case QOR_Q10: return M * L;
        case QOR_YPRJ: return M * L * L;
        case QOR_QPRJ: case QOR_MERGED: return 2 * M * L * L;
        case QOR_KNOTS: return N + c->p + 3;
        case QOR_GSX: case QOR_GSW: return (N + c->p) * c->ax.q;
        default: return -1;
    }
}

static double* arr_ptr(qor_ctx* c, int what) {
    switch (what) {
        case QOR_S: return c->S.data();
        case QOR_CHOL: return c->chol.data();
        case QOR_H: return c->H.data();
        case QOR_RINV: return c->rinv.data();
        case QOR_RDIF: return c->rdif.data();
        case QOR_RPOW1: return c->rpow1.data();
        case QOR_MASK: return c->mask.data();
        case QOR_DVDR: return c->dvdr.data();
        case QOR_UPPER: return (double*)c->upper.data();
        case QOR_LOWER: return (double*)c->lower.data();
        case QOR_YEIG: return c->yeig.data();
        case QOR_YPRJ: return c->yprj.data();
        case QOR_QEIG: return c->qeig.data();
        case QOR_QPRJ: return (double*)c->qprj.data();
        case QOR_MERGED: return (double*)c->merged.data();
        case QOR_KNOTS: return c->ax.kn.data();
        case QOR_GSX: return c->ax.gx.data();
        case QOR_GSW: return c->ax.gw.data();
        case QOR_Y10: return c->y10.data();
        case QOR_Q10: return c->q10.data();
        default: return nullptr;
    }
}

// cn_component_hc::initialize (cn_1e.hpp:342-360): upper = S - i dt/2 H_l,
// lower = LDL^T(S + i dt/2 H_l)
static void build_factors(qor_ctx* c) {
    const long N = c->N, ne = c->ne, nl = c->lmax + 1;
    c->upper.assign(nl * N * ne, czero);
    c->lower.assign(nl * N * ne, czero);
    int err = 0;
#pragma omp parallel for reduction(+ : err) schedule(static)
    for (long l = 0; l < nl; ++l) {
        for (long i = 0; i < N * ne; ++i) {
            double si = c->S[i];
            cplx hi = cplx{0.0, 1.0} * c->H[l * N * ne + i] * c->d.para_dt * 0.5;
            c->upper[l * N * ne + i] = cplx{si, 0.0} - hi;
            c->lower[l * N * ne + i] = cplx{si, 0.0} + hi;
        }
        err += lu_zsb(N, c->p, &c->lower[l * N * ne]);
    }
    c->lu_info = err;
}

// operator_y10/q10::solve_eigen (operators/1e.hpp:304-333, 421-448) and the
// projector set-up of propagator_cn_hc_vgwz::initialize (cn_1e.hpp:673-679).
static int build_angular(qor_ctx* c) {
    const long L = c->L, M = c->M;
    c->y10.assign(M * L, 0.0);
    c->q10.assign(M * L, 0.0);
    for (long im = 0; im < M; ++im)
        for (long il = 0; il < L; ++il) {
            long m = im + c->m0;
            long l = il + std::labs(m);
            // integrate_cos, integrate_sin2dcos (basisfunc.hpp:222-239)
            double y = std::sqrt(double((l + 1) * (l + 1) - m * m) / double((2 * l + 1) * (2 * l + 3)));
            c->y10[im * L + il] = y;
            c->q10[im * L + il] = y * (double)(l + 1);
        }
    c->yeig.assign(M * L, 0.0);
    c->qeig.assign(M * L, 0.0);
    c->yprj.assign(M * L * L, 0.0);
    c->qprj.assign(M * L * L, czero);
    c->merged.assign(M * L * L, czero);
    std::vector<double> diag(L, 0.0), off(L, 0.0), z(L * L);
    for (long im = 0; im < M; ++im) {
        // Y10 with multiplier 1: real symmetric tridiagonal, off-diagonal y(im,il)
        for (long il = 0; il < L; ++il) off[il] = c->y10[im * L + il];
        if (tridiag_eig(L, diag.data(), off.data(), &c->yeig[im * L], z.data())) return -1;
        for (long k = 0; k < L; ++k)
            for (long i = 0; i < L; ++i) c->yprj[(im * L + k) * L + i] = z[k * L + i];
        // i*Q10 with multiplier -i: Hermitian tridiagonal H(l+1,l) = -i q_l.
        // D = diag(i^l) maps it to the real tridiagonal with off-diagonal -q_l:
        // H = D T D^H, eigenvectors v = D v'.  After conjugate() (cn_1e.hpp:677)
        // P_Q[k][l] = conj(v_k(l)) = (-i)^l v'_k(l).
        for (long il = 0; il < L; ++il) off[il] = -c->q10[im * L + il];
        if (tridiag_eig(L, diag.data(), off.data(), &c->qeig[im * L], z.data())) return -1;
        static const cplx mi_pow[4] = {{1, 0}, {0, -1}, {-1, 0}, {0, 1}};
        for (long k = 0; k < L; ++k)
            for (long i = 0; i < L; ++i) c->qprj[(im * L + k) * L + i] = mi_pow[i % 4] * z[k * L + i];
        // merge_two: P_Q * P_Y^H (cn_1e.hpp:76-99)
        for (long k = 0; k < L; ++k)
            for (long j = 0; j < L; ++j) {
                cplx acc = czero;
                for (long i = 0; i < L; ++i) acc += c->qprj[(im * L + k) * L + i] * c->yprj[(im * L + j) * L + i];
                c->merged[(im * L + k) * L + j] = acc;
            }
    }
    return 0;
}

extern "C" {

int qor_create(const qor_dims* d, qor_ctx** out) {
    if (!d || !out) return -1;
    if (d->n_rank < 1 || d->n_ndim < 2 * (d->n_rank + 1) || d->n_ldim < 1 || d->n_mdim < 1) return -2;
    qor_ctx* c = new qor_ctx();
    c->d = *d;
    c->N = d->n_ndim;
    c->p = d->n_rank;
    c->ne = c->p + 1;
    c->L = d->n_ldim;
    c->M = d->n_mdim;
    c->m0 = d->n_mmin;
    long m1 = c->m0 + c->M - 1;
    c->lmax = std::max(std::labs(c->m0), std::labs(m1)) + c->L - 1;  // dimension/1e.hpp:49-50
    c->ax.init_linear(c->N, c->p, d->para_rmin, d->para_rmax);
    const Axis& ax = c->ax;
    // operator_sc (operators/1e.hpp:1232-1245)
    band_0r0(ax, c->S, [](double) { return 1.0; });
    c->chol = c->S;
    if (chol_band(c->N, c->p, c->chol.data())) {
        delete c;
        return -3;
    }
    // operator_rinv1, operator_rdif1, operator_rpow1 (operators/1e.hpp:138-193)
    band_0r0(ax, c->rinv, [](double r) { return 1. / (r + 1e-50); });
    band_0r1(ax, c->rdif, [](double) { return 1.0; });
    band_0r0(ax, c->rpow1, [](double r) { return r; });
    const double zc = d->para_zc;
    band_0r0(ax, c->dvdr, [zc](double r) { return zc / (r * r + 1e-50); });
    // operator_hc (operators/1e.hpp:639-647) -> integrate_hc (basisfunc.hpp:92-106)
    {
        const long N = c->N, ne = c->ne, nl = c->lmax + 1;
        c->H.assign(nl * N * ne, 0.0);
        const double coef = 0.5 / d->para_me;
        auto potn = [zc](double r) { return -zc / (r + 1e-50); };  // potential.hpp:13-16
#pragma omp parallel for schedule(dynamic)
        for (long l = 0; l < nl; ++l) {
            size_t lu = (size_t)l;
            auto part1 = [lu](double r) { return lu * (lu + 1) / (r * r + 1e-50); };
            auto part2 = [](double) { return 1.; };
            for (long ir = 0; ir < N; ++ir)
                for (long ic = 0; ic < ne; ++ic)
                    c->H[(l * N + ir) * ne + ic] = ax.integrate(0, 0, potn, ir, ir + ic) +
                                                   ax.integrate(0, 0, part1, ir, ir + ic) * coef +
                                                   ax.integrate(1, 1, part2, ir, ir + ic) * coef;
        }
    }
    // absorber_mask::initialize (absorber.hpp:42-49)
    {
        const double rmin = d->mask_rmin, rmax = d->mask_rmax, idx = d->mask_fact;
        band_0r0(ax, c->mask, [=](double r) {
            return r < rmin ? 1.0 : r > rmax ? 0.0 : std::pow(std::fabs(std::cos(0.5 * kPi * (r - rmin) / (rmax - rmin))), idx);
        });
    }
    build_factors(c);
    if (build_angular(c)) {
        delete c;
        return -4;
    }
    c->tmpc.assign(c->M * c->L * c->N, czero);
    *out = c;
    return 0;
}

void qor_destroy(qor_ctx* c) { delete c; }

int64_t qor_size(const qor_ctx* c, int what) { return c ? arr_size(c, what) : -1; }

int qor_get(const qor_ctx* c, int what, double* dst) {
    if (!c || !dst) return -1;
    int64_t n = arr_size(c, what);
    double* src = arr_ptr(const_cast<qor_ctx*>(c), what);
    if (n < 0 || !src) return -1;
    std::memcpy(dst, src, sizeof(double) * n);
    return 0;
}

int qor_set(qor_ctx* c, int what, const double* src) {
    if (!c || !src) return -1;
    int64_t n = arr_size(c, what);
    double* dst = arr_ptr(c, what);
    if (n < 0 || !dst) return -1;
    std::memcpy(dst, src, sizeof(double) * n);
    return 0;
}

int qor_lu_info(const qor_ctx* c) { return c ? c->lu_info : -1; }
int qor_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}

int qor_eigenstate(const qor_ctx* c, int64_t im, int64_t il, int64_t in, double* energy, double* coef) {
    if (!c || !coef) return -1;
    if (im < 0 || im >= c->M || il < 0 || il >= c->L || in < 0 || in >= c->N) return -2;
    const long N = c->N;
    std::vector<double> v;
    double e = 0.0;
    long l = c->in_l(im, il);
    gen_eig_band(N, c->p, &c->H[l * N * c->ne], c->S.data(), in, &e, v);
    std::memset(coef, 0, sizeof(double) * 2 * c->M * c->L * N);
    cplx* psi = (cplx*)coef + (im * c->L + il) * N;
    for (long i = 0; i < N; ++i) psi[i] = cplx{v[i], 0.0};
    if (energy) *energy = e;
    return 0;
}

// eigenstate_sph::initialize (population_1e.hpp:57-118) with the criterion
// "the nsel lowest states of every l": the left-hand basis S v (:108-113)
int qor_projection_basis(const qor_ctx* c, int64_t nsel, double* W, double* E) {
    if (!c || !W || nsel < 1 || nsel > c->N) return -1;
    const long N = c->N, nl = c->lmax + 1;
    int bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : bad)
    for (long l = 0; l < nl; ++l) {
        std::vector<double> v;
        for (long n = 0; n < nsel; ++n) {
            double e = 0.0;
            if (gen_eig_band(N, c->p, &c->H[l * N * c->ne], c->S.data(), n, &e, v)) {
                ++bad;
                continue;
            }
            band_mul_real(N, c->p, c->S.data(), v.data(), W + (l * nsel + n) * N);
            if (E) E[l * nsel + n] = e;
        }
    }
    return bad ? -2 : 0;
}

// eigenstate_sph::projection (population_1e.hpp:29-55): per channel (im,il)
// with l = il + |m0+im|, proj[(im*L+il)*stride + in] = sum_ir W[l][in][ir] * c[im][il][ir]
int qor_projection(const qor_ctx* c, const double* coef, int64_t nsel, const double* W, double* proj,
                   int64_t stride) {
    if (!c || !coef || !W || !proj || nsel < 1 || stride < nsel) return -1;
    const long N = c->N, L = c->L, M = c->M;
    const cplx* cf = (const cplx*)coef;
    cplx* pr = (cplx*)proj;
#pragma omp parallel for collapse(2) schedule(static)
    for (long im = 0; im < M; ++im)
        for (long il = 0; il < L; ++il) {
            const long l = c->in_l(im, il);
            const cplx* x = cf + (im * L + il) * N;
            for (long in = 0; in < nsel; ++in) {
                const double* w = W + (l * nsel + in) * N;
                cplx t = czero;
                for (long ir = 0; ir < N; ++ir) t += w[ir] * x[ir];
                pr[(im * L + il) * stride + in] = t;
            }
        }
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// the propagator pieces of propagator_cn_hc_vgwz (cn_1e.hpp:621-715)

// cn_transform_hl::construct / eliminate (cn_1e.hpp:101-147) as loops
// blocked over r (a block of kRotRB rows of every source channel stays in
// cache while all destination channels are formed; the reference calls a
// BLAS GEMM here).  Each sum runs over i in order, so the blocking does not
// change a bit.
// dst[im][k][:] = sum_i P[im][k][i] * src[im][i][:]       (construct, trans=0)
// dst[im][k][:] = sum_i conj(P[im][i][k]) * src[im][i][:] (eliminate, trans=1)
static const long kRotRB = 64;
static void rotate_real(const qor_ctx* c, const double* P, const cplx* src, cplx* dst, int trans) {
    const long L = c->L, N = c->N, M = c->M, nb = (N + kRotRB - 1) / kRotRB;
#pragma omp parallel for collapse(2) schedule(static)
    for (long im = 0; im < M; ++im)
        for (long b = 0; b < nb; ++b) {
            const long r0 = 2 * b * kRotRB, r1 = 2 * std::min(N, (b + 1) * kRotRB);
            for (long k = 0; k < L; ++k) {
                double* yd = (double*)(dst + (im * L + k) * N);
                for (long r = r0; r < r1; ++r) yd[r] = 0.0;
                for (long i = 0; i < L; ++i) {
                    double pk = trans ? P[(im * L + i) * L + k] : P[(im * L + k) * L + i];
                    const double* xd = (const double*)(src + (im * L + i) * N);
                    for (long r = r0; r < r1; ++r) yd[r] += pk * xd[r];
                }
            }
        }
}
static void rotate_cplx(const qor_ctx* c, const cplx* P, const cplx* src, cplx* dst, int trans) {
    const long L = c->L, N = c->N, M = c->M, nb = (N + kRotRB - 1) / kRotRB;
#pragma omp parallel for collapse(2) schedule(static)
    for (long im = 0; im < M; ++im)
        for (long b = 0; b < nb; ++b) {
            const long r0 = b * kRotRB, r1 = std::min(N, (b + 1) * kRotRB);
            for (long k = 0; k < L; ++k) {
                cplx* y = dst + (im * L + k) * N;
                for (long r = r0; r < r1; ++r) y[r] = czero;
                for (long i = 0; i < L; ++i) {
                    cplx pk = trans ? conjc(P[(im * L + i) * L + k]) : P[(im * L + k) * L + i];
                    const cplx* x = src + (im * L + i) * N;
                    for (long r = r0; r < r1; ++r) y[r] += pk * x[r];
                }
            }
        }
}

// cn_hl_propagate1/2 (propagate.hpp:333-441) over all rotated channels of tmpc
static void hl_propagate(const qor_ctx* c, cplx* buf, const double* eig, const double* rmat, int symflag,
                         double fzdt) {
    const long N = c->N, p = c->p, ML = c->M * c->L;
#pragma omp parallel
    {
        std::vector<cplx> ytmp(N), utmp(N * p);
#pragma omp for schedule(static)
        for (long iml = 0; iml < ML; ++iml) {
            cplx ifet;
            if (symflag == 0) {
                cplx ifdt = cplx{0.0, 1.0} * (fzdt / 2);  // propagate.hpp:343
                ifet = ifdt * eig[iml];
            } else {
                double ifdt = -fzdt / 2;  // propagate.hpp:400
                ifet = cplx{eig[iml] * ifdt, 0.0};
            }
            cn_kernel(N, p, symflag, c->S.data(), rmat, ifet, buf + iml * N, ytmp.data(), utmp.data());
        }
    }
}

// cn_hc_propagate0 (propagate.hpp:311-331)
static void hc_propagate0(const qor_ctx* c, cplx* coef) {
    const long N = c->N, ne = c->ne, L = c->L, M = c->M;
#pragma omp parallel
    {
        std::vector<cplx> w(N);
#pragma omp for collapse(2) schedule(static)
        for (long im = 0; im < M; ++im)
            for (long il = 0; il < L; ++il) {
                long l = c->in_l(im, il);
                cplx* cc = coef + (im * L + il) * N;
                symb_mul(N, ne, &c->upper[l * N * ne], cc, w.data());
                ltrb_ldl(N, ne, &c->lower[l * N * ne], w.data(), w.data());
                utrb_ldl(N, ne, &c->lower[l * N * ne], w.data(), cc);
            }
    }
}

// ---------------------------------------------------------------------------
// the propagator pieces of propagator_cn_hc_vgwr (cn_1e.hpp:812-951)

// sub-block ranges of dimension_1e (dimension/1e.hpp:52-80): msub (rotations)
// and mpro (sweeps) for the selector isub
struct EpRange {
    long s0, s1, p0, p1;
};
static EpRange ep_range(const qor_ctx* c, int isub) {
    const long M = c->M, L = c->L, ML = M * L;
    const long n_zero = M % 2 == 0 ? 0 : L;
    const long n_even = (M / 2 * 2) * ((L + 1) / 2);
    EpRange r{0, ML, 0, ML};
    if (isub == 2) {
        r.s1 = n_even + n_zero;
        r.p1 = n_even;
    } else if (isub == 3) {
        r.s0 = n_even;
        r.p0 = n_even + n_zero;
    }
    return r;
}

// cn_transform_hm::construct<isrc,idst> (cn_1e.hpp:222-247): rows [d0,d1) of
// temp from rows [s0,s1) of coef, C' = P C with the cblas operand offsets of
// the reference (complex: P + s0 + ML*d0; real: P + d0*ML, n = 2N)
static void hm_construct_cplx(const qor_ctx* c, const cplx* P, long s0, long s1, long d0, long d1, const cplx* src,
                              cplx* dst) {
    const long N = c->N, ML = c->M * c->L;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < d1 - d0; ++i) {
        cplx* y = dst + (d0 + i) * N;
        for (long r = 0; r < N; ++r) y[r] = czero;
        for (long k = 0; k < s1 - s0; ++k) {
            const cplx a = P[(d0 + i) * ML + s0 + k];
            const cplx* x = src + (s0 + k) * N;
            for (long r = 0; r < N; ++r) y[r] += a * x[r];
        }
    }
}
static void hm_construct_real(const qor_ctx* c, const double* P, long s0, long s1, long d0, long d1,
                              const cplx* src, cplx* dst) {
    const long N = c->N, ML = c->M * c->L;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < d1 - d0; ++i) {
        double* y = (double*)(dst + (d0 + i) * N);
        for (long r = 0; r < 2 * N; ++r) y[r] = 0.0;
        for (long k = 0; k < s1 - s0; ++k) {
            const double a = P[(d0 + i) * ML + k];
            const double* x = (const double*)(src + (s0 + k) * N);
            for (long r = 0; r < 2 * N; ++r) y[r] += a * x[r];
        }
    }
}
// cn_transform_hm::eliminate<isrc,idst> (cn_1e.hpp:249-273): C = P^H C'
// (complex) or P^T C' (real), operand P + d0 + s0*ML
static void hm_eliminate_cplx(const qor_ctx* c, const cplx* P, long s0, long s1, long d0, long d1, const cplx* src,
                              cplx* dst) {
    const long N = c->N, ML = c->M * c->L;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < d1 - d0; ++i) {
        cplx* y = dst + (d0 + i) * N;
        for (long r = 0; r < N; ++r) y[r] = czero;
        for (long k = 0; k < s1 - s0; ++k) {
            const cplx a = conjc(P[(s0 + k) * ML + d0 + i]);
            const cplx* x = src + (s0 + k) * N;
            for (long r = 0; r < N; ++r) y[r] += a * x[r];
        }
    }
}
static void hm_eliminate_real(const qor_ctx* c, const double* P, long s0, long s1, long d0, long d1,
                              const cplx* src, cplx* dst) {
    const long N = c->N, ML = c->M * c->L;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < d1 - d0; ++i) {
        double* y = (double*)(dst + (d0 + i) * N);
        for (long r = 0; r < 2 * N; ++r) y[r] = 0.0;
        for (long k = 0; k < s1 - s0; ++k) {
            const double a = P[(s0 + k) * ML + d0 + i];
            const double* x = (const double*)(src + (s0 + k) * N);
            for (long r = 0; r < 2 * N; ++r) y[r] += a * x[r];
        }
    }
}

// cn_hm_propagate1/2<isub> (propagate.hpp:443-568) on rotated channels [p0,p1)
// of buf: symflag 0 (R symmetric, ifdt = i frdt/2) or 1 (R = d/dr, ifdt = -frdt/2)
static void hm_propagate(const qor_ctx* c, cplx* buf, const double* zeig, const double* rmat, int symflag, double frdt,
                         long p0, long p1) {
    const long N = c->N, p = c->p;
#pragma omp parallel
    {
        std::vector<cplx> ytmp(N), utmp(N * p);
#pragma omp for schedule(static)
        for (long iml = p0; iml < p1; ++iml) {
            cplx ifet;
            if (symflag == 0) {
                cplx ifdt = cplx{0.0, 1.0} * (frdt / 2);  // propagate.hpp:463
                ifet = zeig[iml] * ifdt;
            } else {
                double ifdt = -frdt / 2;  // propagate.hpp:525
                ifet = cplx{zeig[iml] * ifdt, 0.0};
            }
            cn_kernel(N, p, symflag, c->S.data(), rmat, ifet, buf + iml * N, ytmp.data(), utmp.data());
        }
    }
}

extern "C" {

// ---------------------------------------------------------------------------
// the controller network, FP64 (layout and loss as include/qpc_mlp.h)
static constexpr int kNI = 7, kNH = 64, kNO = 2;
static constexpr int kW1 = 0, kB1 = kW1 + kNH * kNI, kW2 = kB1 + kNH, kB2 = kW2 + kNH * kNH, kW3 = kB2 + kNH,
                     kB3 = kW3 + kNO * kNH, kNP = kB3 + kNO;
static double gelu(double x) { return 0.5 * x * (1.0 + std::erf(x / std::sqrt(2.0))); }
static double gelu_d(double x) {
    return 0.5 * (1.0 + std::erf(x / std::sqrt(2.0))) + x * std::exp(-0.5 * x * x) / std::sqrt(2.0 * kPi);
}
struct MlpAct {
    double z1[kNH], h1[kNH], z2[kNH], h2[kNH], y[kNO];
};
static void mlp_fwd(const double* P, const double* x, MlpAct& a) {
    for (int j = 0; j < kNH; ++j) {
        double z = P[kB1 + j];
        for (int i = 0; i < kNI; ++i) z += P[kW1 + j * kNI + i] * x[i];
        a.z1[j] = z;
        a.h1[j] = gelu(z);
    }
    for (int j = 0; j < kNH; ++j) {
        double z = P[kB2 + j];
        for (int k = 0; k < kNH; ++k) z += P[kW2 + j * kNH + k] * a.h1[k];
        a.z2[j] = z;
        a.h2[j] = gelu(z);
    }
    for (int o = 0; o < kNO; ++o) {
        double y = P[kB3 + o];
        for (int k = 0; k < kNH; ++k) y += P[kW3 + o * kNH + k] * a.h2[k];
        a.y[o] = y;
    }
}

int qor_mlp_forward(const double* P, const double* X, int64_t n, double* Y) {
    if (!P || !X || !Y || n < 1) return -1;
    MlpAct a;
    for (int64_t s = 0; s < n; ++s) {
        mlp_fwd(P, X + s * kNI, a);
        for (int o = 0; o < kNO; ++o) Y[s * kNO + o] = a.y[o];
    }
    return 0;
}

int qor_mlp_loss_grad(const double* P, const double* X, const double* T, int64_t n, double* loss, double* G) {
    if (!P || !X || !T || !loss || !G || n < 1) return -1;
    for (int p = 0; p < kNP; ++p) G[p] = 0.0;
    double L = 0.0;
    MlpAct a;
    for (int64_t s = 0; s < n; ++s) {
        const double* x = X + s * kNI;
        mlp_fwd(P, x, a);
        double dy[kNO], d2[kNH], d1[kNH];
        for (int o = 0; o < kNO; ++o) {
            const double r = a.y[o] - T[s * kNO + o];
            L += r * r;
            dy[o] = r / (double)n;
            G[kB3 + o] += dy[o];
            for (int k = 0; k < kNH; ++k) G[kW3 + o * kNH + k] += dy[o] * a.h2[k];
        }
        for (int k = 0; k < kNH; ++k) {
            double g = 0.0;
            for (int o = 0; o < kNO; ++o) g += P[kW3 + o * kNH + k] * dy[o];
            d2[k] = g * gelu_d(a.z2[k]);
            G[kB2 + k] += d2[k];
            for (int i = 0; i < kNH; ++i) G[kW2 + k * kNH + i] += d2[k] * a.h1[i];
        }
        for (int j = 0; j < kNH; ++j) {
            double g = 0.0;
            for (int k = 0; k < kNH; ++k) g += P[kW2 + k * kNH + j] * d2[k];
            d1[j] = g * gelu_d(a.z1[j]);
            G[kB1 + j] += d1[j];
            for (int i = 0; i < kNI; ++i) G[kW1 + j * kNI + i] += d1[j] * x[i];
        }
    }
    *loss = L / (2.0 * (double)n);
    return 0;
}

int qor_set_ep(qor_ctx* c, int isub, const double* ex1, const double* ex2, const double* ey1, const double* ey2,
               const double* tx1, const double* ty2, const double* x1x2, const double* x2y1, const double* y1y2) {
    if (!c || !ex1 || !ex2 || !ey1 || !ey2 || !tx1 || !ty2 || !x1x2 || !x2y1 || !y1y2) return -1;
    if (isub < -1 || isub > 3) return -2;
    const long ML = c->M * c->L;
    const double* e[4] = {ex1, ex2, ey1, ey2};
    for (int k = 0; k < 4; ++k) c->ep_eig[k].assign(e[k], e[k] + ML);
    c->ep_tx1.assign(tx1, tx1 + ML * ML);
    c->ep_ty2.assign(ty2, ty2 + ML * ML);
    c->ep_x1x2.assign((const cplx*)x1x2, (const cplx*)x1x2 + ML * ML);
    c->ep_x2y1.assign((const cplx*)x2y1, (const cplx*)x2y1 + ML * ML);
    c->ep_y1y2.assign((const cplx*)y1y2, (const cplx*)y1y2 + ML * ML);
    c->ep_isub = isub;
    c->have_ep = true;
    return 0;
}

int qor_step_vgwr(qor_ctx* c, double* coef_d, double ax_ti, double ax_tf, double ay_ti, double ay_tf) {
    if (!c || !coef_d || !c->have_ep) return -1;
    const long ML = c->M * c->L;
    const EpRange s = ep_range(c, c->ep_isub), all{0, ML, 0, ML};
    const double step = c->d.para_dt * 0.5;  // prop_x1..prop_y2 step (cn_1e.hpp:873-876)
    const double *ex1 = c->ep_eig[0].data(), *ex2 = c->ep_eig[1].data(), *ey1 = c->ep_eig[2].data(),
                 *ey2 = c->ep_eig[3].data();
    const double *rd = c->rdif.data(), *rr = c->rinv.data();
    cplx* coef = (cplx*)coef_d;  // "coef.data"
    cplx* work = c->tmpc.data();  // "work.tmpc"
    ax_ti /= 2.0;
    ay_ti /= 2.0;
    ax_tf /= 2.0;
    ay_tf /= 2.0;
    hm_construct_real(c, c->ep_tx1.data(), all.s0, all.s1, s.s0, s.s1, coef, work);     // tran_x1.construct<-1,isub>
    hm_propagate(c, work, ex1, rd, 1, (-ax_ti) * step, s.p0, s.p1);                      // prop_x1.propagate2(-ax_ti)
    std::swap(work, coef);
    hm_construct_cplx(c, c->ep_x1x2.data(), s.s0, s.s1, s.s0, s.s1, coef, work);        // x1_2_x2.construct
    hm_propagate(c, work, ex2, rr, 0, (ax_ti) * step, s.p0, s.p1);                       // prop_x2.propagate1(ax_ti)
    std::swap(work, coef);
    hm_construct_cplx(c, c->ep_x2y1.data(), s.s0, s.s1, s.s0, s.s1, coef, work);        // x2_2_y1.construct
    hm_propagate(c, work, ey1, rd, 1, (-ay_ti) * step, s.p0, s.p1);                      // prop_y1.propagate2(-ay_ti)
    std::swap(work, coef);
    hm_construct_cplx(c, c->ep_y1y2.data(), s.s0, s.s1, s.s0, s.s1, coef, work);        // y1_2_y2.construct
    hm_propagate(c, work, ey2, rr, 0, (-ay_ti) * step, s.p0, s.p1);                      // prop_y2.propagate1(-ay_ti)
    hm_eliminate_real(c, c->ep_ty2.data(), s.s0, s.s1, all.s0, all.s1, work, coef);     // tran_y2.eliminate<isub,-1>
    hc_propagate0(c, coef);                                                               // prop_hc.propagate0
    hm_construct_real(c, c->ep_ty2.data(), all.s0, all.s1, s.s0, s.s1, coef, work);     // tran_y2.construct<-1,isub>
    hm_propagate(c, work, ey2, rr, 0, (-ay_tf) * step, s.p0, s.p1);                      // prop_y2.propagate1(-ay_tf)
    hm_eliminate_cplx(c, c->ep_y1y2.data(), s.s0, s.s1, s.s0, s.s1, work, coef);        // y1_2_y2.eliminate
    std::swap(work, coef);
    hm_propagate(c, work, ey1, rd, 1, (-ay_tf) * step, s.p0, s.p1);                      // prop_y1.propagate2(-ay_tf)
    hm_eliminate_cplx(c, c->ep_x2y1.data(), s.s0, s.s1, s.s0, s.s1, work, coef);        // x2_2_y1.eliminate
    std::swap(work, coef);
    hm_propagate(c, work, ex2, rr, 0, (ax_tf) * step, s.p0, s.p1);                       // prop_x2.propagate1(ax_tf)
    hm_eliminate_cplx(c, c->ep_x1x2.data(), s.s0, s.s1, s.s0, s.s1, work, coef);        // x1_2_x2.eliminate
    std::swap(work, coef);
    hm_propagate(c, work, ex1, rd, 1, (-ax_tf) * step, s.p0, s.p1);                      // prop_x1.propagate2(-ax_tf)
    hm_eliminate_real(c, c->ep_tx1.data(), s.s0, s.s1, all.s0, all.s1, work, coef);     // tran_x1.eliminate<isub,-1>
    // six swaps: coef is the caller's array again
    return 0;
}

int qor_run_ep(qor_ctx* c, double* coef, const double* axtab, const double* aytab, int64_t nt_field, int64_t it0,
               int64_t it1, int use_mask) {
    if (!c || !coef || !c->have_ep) return -1;
    for (int64_t it = it0; it < it1; ++it) {
        if (it < nt_field) qor_step_vgwr(c, coef, axtab[it], axtab[it + 1], aytab[it], aytab[it + 1]);
        else qor_step_free(c, coef);
        if (use_mask) qor_mask(c, coef);
    }
    return 0;
}

int qor_step_vgwz(qor_ctx* c, double* coef_d, double az_ti, double az_tf) {
    if (!c || !coef_d) return -1;
    cplx* coef = (cplx*)coef_d;
    const double step = c->d.para_dt * 0.5;  // prop_z1/prop_z2 step (cn_1e.hpp:670-671)
    cplx* bufA = coef;          // "coef.data"
    cplx* bufB = c->tmpc.data();  // "work.tmpc"
    az_ti = -az_ti;
    az_tf = -az_tf;
    rotate_real(c, c->yprj.data(), bufA, bufB, 0);                   // tran_z2.construct
    hl_propagate(c, bufB, c->yeig.data(), c->rdif.data(), 1, az_ti * step);  // prop_z2.propagate2
    std::swap(bufA, bufB);                                          // swap(work.tmpc, coef.data)
    rotate_cplx(c, c->merged.data(), bufA, bufB, 0);                 // z2_2_z1.construct
    hl_propagate(c, bufB, c->qeig.data(), c->rinv.data(), 0, az_ti * step);  // prop_z1.propagate1
    rotate_cplx(c, c->qprj.data(), bufB, bufA, 1);                   // tran_z1.eliminate
    hc_propagate0(c, bufA);                                         // prop_hc.propagate0
    rotate_cplx(c, c->qprj.data(), bufA, bufB, 0);                   // tran_z1.construct
    hl_propagate(c, bufB, c->qeig.data(), c->rinv.data(), 0, az_tf * step);  // prop_z1.propagate1
    rotate_cplx(c, c->merged.data(), bufB, bufA, 1);                 // z2_2_z1.eliminate
    std::swap(bufA, bufB);                                          // swap
    hl_propagate(c, bufB, c->yeig.data(), c->rdif.data(), 1, az_tf * step);  // prop_z2.propagate2
    rotate_real(c, c->yprj.data(), bufB, bufA, 1);                   // tran_z2.eliminate
    // two swaps: bufA is the caller's array again
    return 0;
}

// propagator_cn_hc_lgwz::propagate(coef, ez_ti, ez_tf) (cn_1e.hpp:598-613):
// the length-gauge step.  tran_z1 is the real P_Y of Y10 (solve_eigen with
// idty, :590), prop_z1 = cn_component_hl(S, r, dt/2) (:587) with the Y10
// eigenvalues; propagate1 acts on work.tmpc between construct and eliminate,
// with no pointer swaps.
int qor_step_lgwz(qor_ctx* c, double* coef_d, double ez_ti, double ez_tf) {
    if (!c || !coef_d) return -1;
    cplx* coef = (cplx*)coef_d;
    cplx* tmp = c->tmpc.data();
    const double step = c->d.para_dt * 0.5;
    rotate_real(c, c->yprj.data(), coef, tmp, 0);                           // tran_z1.construct
    hl_propagate(c, tmp, c->yeig.data(), c->rpow1.data(), 0, ez_ti * step);  // prop_z1.propagate1(ez_ti)
    rotate_real(c, c->yprj.data(), tmp, coef, 1);                           // tran_z1.eliminate
    hc_propagate0(c, coef);                                                 // prop_hc.propagate0
    rotate_real(c, c->yprj.data(), coef, tmp, 0);                           // tran_z1.construct
    hl_propagate(c, tmp, c->yeig.data(), c->rpow1.data(), 0, ez_tf * step);  // prop_z1.propagate1(ez_tf)
    rotate_real(c, c->yprj.data(), tmp, coef, 1);                           // tran_z1.eliminate
    return 0;
}

// ---------------------------------------------------------------------------
// propagator_cn_hc_lgwr (cn_1e.hpp:717-810): the elliptically polarised step
// in the length gauge, x1 (Y11, mult 1) and y1 (Y11, mult -i) sweeps of
// S + i e r on the (m,l) grid around the field-free solve
int qor_set_lgwr(qor_ctx* c, int isub, const double* ex1, const double* ey1, const double* tx1, const double* x1y1,
                 const double* ty1) {
    if (!c || !ex1 || !ey1 || !tx1 || !x1y1 || !ty1) return -1;
    if (isub < -1 || isub > 3) return -2;
    const long ML = c->M * c->L;
    c->lr_ex1.assign(ex1, ex1 + ML);
    c->lr_ey1.assign(ey1, ey1 + ML);
    c->lr_tx1.assign(tx1, tx1 + ML * ML);
    c->lr_x1y1.assign((const cplx*)x1y1, (const cplx*)x1y1 + ML * ML);
    c->lr_ty1.assign((const cplx*)ty1, (const cplx*)ty1 + ML * ML);
    c->lr_isub = isub;
    c->have_lgwr = true;
    return 0;
}

// propagator_cn_hc_lgwr::propagate<isub> (cn_1e.hpp:767-803)
int qor_step_lgwr(qor_ctx* c, double* coef_d, double ex_ti, double ex_tf, double ey_ti, double ey_tf) {
    if (!c || !coef_d || !c->have_lgwr) return -1;
    const long ML = c->M * c->L;
    const EpRange s = ep_range(c, c->lr_isub), all{0, ML, 0, ML};
    const double step = c->d.para_dt * 0.5;  // prop_x1, prop_y1 step (cn_1e.hpp:758-759)
    const double *ex1 = c->lr_ex1.data(), *ey1 = c->lr_ey1.data(), *rr = c->rpow1.data();
    cplx* coef = (cplx*)coef_d;  // "coef.data"
    cplx* work = c->tmpc.data();  // "work.tmpc"
    ex_ti /= 2.0;
    ey_ti /= 2.0;
    ex_tf /= 2.0;
    ey_tf /= 2.0;
    hm_construct_real(c, c->lr_tx1.data(), all.s0, all.s1, s.s0, s.s1, coef, work);     // tran_x1.construct<-1,isub>
    hm_propagate(c, work, ex1, rr, 0, ex_ti * step, s.p0, s.p1);                         // prop_x1.propagate1(ex_ti)
    std::swap(work, coef);
    hm_construct_cplx(c, c->lr_x1y1.data(), s.s0, s.s1, s.s0, s.s1, coef, work);        // x1_2_y1.construct
    hm_propagate(c, work, ey1, rr, 0, ey_ti * step, s.p0, s.p1);                         // prop_y1.propagate1(ey_ti)
    hm_eliminate_cplx(c, c->lr_ty1.data(), s.s0, s.s1, all.s0, all.s1, work, coef);     // tran_y1.eliminate<isub,-1>
    hc_propagate0(c, coef);                                                               // prop_hc.propagate0
    hm_construct_cplx(c, c->lr_ty1.data(), all.s0, all.s1, s.s0, s.s1, coef, work);     // tran_y1.construct<-1,isub>
    hm_propagate(c, work, ey1, rr, 0, ey_tf * step, s.p0, s.p1);                         // prop_y1.propagate1(ey_tf)
    hm_eliminate_cplx(c, c->lr_x1y1.data(), s.s0, s.s1, s.s0, s.s1, work, coef);        // x1_2_y1.eliminate
    std::swap(work, coef);
    hm_propagate(c, work, ex1, rr, 0, ex_tf * step, s.p0, s.p1);                         // prop_x1.propagate1(ex_tf)
    hm_eliminate_real(c, c->lr_tx1.data(), s.s0, s.s1, all.s0, all.s1, work, coef);     // tran_x1.eliminate<isub,-1>
    // two swaps: coef is the caller's array again
    return 0;
}

int qor_run_lgwr(qor_ctx* c, double* coef, const double* extab, const double* eytab, int64_t nt_field, int64_t it0,
                 int64_t it1, int use_mask) {
    if (!c || !coef || !c->have_lgwr) return -1;
    for (int64_t it = it0; it < it1; ++it) {
        if (it < nt_field) qor_step_lgwr(c, coef, extab[it], extab[it + 1], eytab[it], eytab[it + 1]);
        else qor_step_free(c, coef);
        if (use_mask) qor_mask(c, coef);
    }
    return 0;
}

int qor_run_lg(qor_ctx* c, double* coef, const double* etab, int64_t nt_field, int64_t it0, int64_t it1,
               int use_mask) {
    if (!c || !coef) return -1;
    for (int64_t it = it0; it < it1; ++it) {
        if (it < nt_field) qor_step_lgwz(c, coef, etab[it], etab[it + 1]);
        else qor_step_free(c, coef);
        if (use_mask) qor_mask(c, coef);
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Krylov propagator: propagator_arnoldi (la_1e.hpp:90-335) with
// propagator_hc_vgwz::multiply (la_1e.hpp:480-497).  operator_sc::inverse_all
// (operators/1e.hpp:1253-1260) offsets channel i by i*N*N, which leaves the
// array for every channel but the first; it is restated here as intended,
// one Cholesky solve per channel.

// asyb_mul_vecd (blas2.hpp:389-397, 838-): y = A x, A antisymmetric band
// stored upper (A(i,i+k) = a[i*m+k], A(i+j,i) = -a[i*m+j])
static void asyb_mul(long n, long m, const double* a, const cplx* x, cplx* y) {
    for (long i = 0; i < n; ++i) {
        cplx t = czero;
        for (long j = 1; j < m && j <= i; ++j) t -= a[(i - j) * m + j] * x[i - j];
        for (long k = 0; k < m && i + k < n; ++k) t += a[i * m + k] * x[i + k];
        y[i] = t;
    }
}
// symb_mul_vecd with general n (no n >= 2m requirement)
static void sym_mul(long n, long m, const double* a, const cplx* x, cplx* y) {
    for (long i = 0; i < n; ++i) {
        cplx t = czero;
        for (long j = 1; j < m && j <= i; ++j) t += a[(i - j) * m + j] * x[i - j];
        for (long k = 0; k < m && i + k < n; ++k) t += a[i * m + k] * x[i + k];
        y[i] = t;
    }
}

// u1 = H(az) u0 (la_1e.hpp:480-497)
static void krylov_multiply(const qor_ctx* c, double az, const cplx* u0, cplx* u1) {
    const long N = c->N, L = c->L, M = c->M, ne = c->ne;
    const cplx iaz = {0.0, az};
#pragma omp parallel for collapse(2) schedule(static)
    for (long im = 0; im < M; ++im)
        for (long il = 0; il < L; ++il) {
            const cplx* x = u0 + (im * L + il) * N;
            cplx* y = u1 + (im * L + il) * N;
            std::vector<cplx> t(N);
            sym_mul(N, ne, &c->H[c->in_l(im, il) * N * ne], x, y);
            if (il > 0) {
                const cplx cq = iaz * c->q10[im * L + il - 1], cy = (czero - iaz) * c->y10[im * L + il - 1];
                sym_mul(N, ne, c->rinv.data(), x - N, t.data());
                for (long i = 0; i < N; ++i) y[i] += t[i] * cq;
                asyb_mul(N, ne, c->rdif.data(), x - N, t.data());
                for (long i = 0; i < N; ++i) y[i] += t[i] * cy;
            }
            if (il < L - 1) {
                const cplx cy = (czero - iaz) * c->y10[im * L + il], cq = (czero - iaz) * c->q10[im * L + il];
                asyb_mul(N, ne, c->rdif.data(), x + N, t.data());
                for (long i = 0; i < N; ++i) y[i] += t[i] * cy;
                sym_mul(N, ne, c->rinv.data(), x + N, t.data());
                for (long i = 0; i < N; ++i) y[i] += t[i] * cq;
            }
        }
}

// <u|v> summed over the whole state (vecd_proj_vecd, blas1: sum conj(u) v)
static cplx dot_all(long n, const cplx* u, const cplx* v) {
    double re = 0.0, im = 0.0;
    for (long i = 0; i < n; ++i) {
        re += u[i].re * v[i].re + u[i].im * v[i].im;
        im += u[i].re * v[i].im - u[i].im * v[i].re;
    }
    return {re, im};
}

static double snorm_all(const qor_ctx* c, const cplx* u) {
    double s = 0.0;
    for (long ch = 0; ch < c->M * c->L; ++ch) s += symb_prj(c->N, c->ne, c->S.data(), u + ch * c->N, u + ch * c->N).re;
    return s;
}

int qor_krylov_step(qor_ctx* c, double* coef_d, double az, double dt, int64_t kmax, double emin, double emax,
                    int64_t* m_used, double* eps_out, double* ab_out) {
    if (!c || !coef_d || kmax < 1) return -1;
    const long N = c->N, nl = c->M * c->L * N;
    cplx* w = (cplx*)coef_d;
    std::vector<cplx> U((kmax + 2) * nl);
    std::vector<double> a(kmax + 2, 0.0), b(kmax + 2, 0.0);
    std::vector<cplx> x(kmax + 2);
    auto u = [&](long i) { return U.data() + i * nl; };
    // construct_sub (la_1e.hpp:134-154)
    for (long e = 0; e < nl; ++e) u(0)[e] = w[e];
    double norm = std::sqrt(snorm_all(c, u(0)));
    for (long e = 0; e < nl; ++e) u(0)[e] = u(0)[e] * (1.0 / norm);
    b[0] = norm;
    // Krylov loop (la_1e.hpp:318-328) with orthogonalize_ip (:195-240)
    bool more = true;
    long m = 0;
    while (more && m < kmax) {
        krylov_multiply(c, az, u(m), u(m + 1));
        const long k = ++m;
        cplx* uk = u(k);
        for (long i = 0; i < k; ++i) x[i] = dot_all(nl, u(i), uk);
        for (long ch = 0; ch < c->M * c->L; ++ch) {  // S \ (S u_k), channel by channel
            ltrb_inv(N, c->ne, c->chol.data(), uk + ch * N, uk + ch * N);
            utrb_inv(N, c->ne, c->chol.data(), uk + ch * N, uk + ch * N);
        }
        for (long i = 0; i < k; ++i)
            if (i == k - 1 || cnorm(x[i]) > 4e-27)
                for (long e = 0; e < nl; ++e) uk[e] -= u(i)[e] * x[i];
        double nk = snorm_all(c, uk);
        bool stop = false;
        if (nk < 1e-13) {
            nk = 0.0;
            stop = true;
        } else {
            nk = std::sqrt(nk);
            for (long e = 0; e < nl; ++e) uk[e] = uk[e] * (1.0 / nk);
        }
        a[k - 1] = x[k - 1].re;
        b[k] = nk;
        more = !stop;
    }
    if (ab_out)
        for (long i = 0; i < m; ++i) {
            ab_out[2 * i] = a[i];
            ab_out[2 * i + 1] = b[i + 1];
        }
    // propagate_sub (la_1e.hpp:157-182): eigen-decomposition of the tridiagonal
    std::vector<double> lam(m), z(m * m), off(m > 1 ? m - 1 : 1, 0.0);
    for (long i = 0; i + 1 < m; ++i) off[i] = b[i + 1];
    if (tridiag_eig(m, a.data(), off.data(), lam.data(), z.data())) return -2;
    std::vector<cplx> wk(m);
    for (long i = 0; i < m; ++i) {
        const double ph = -dt * lam[i];
        wk[i] = cplx{std::cos(ph), std::sin(ph)} * z[i * m];
    }
    for (long j = 0; j < m; ++j) {
        cplx t = czero;
        for (long i = 0; i < m; ++i) t += wk[i] * z[i * m + j];
        x[j] = t * b[0];
    }
    // eliminate_sub (la_1e.hpp:185-193)
    for (long e = 0; e < nl; ++e) {
        cplx t = czero;
        for (long i = 0; i < m; ++i) t += u(i)[e] * x[i];
        w[e] = t;
    }
    // criterion (la_1e.hpp:296-310)
    const double eps = m > 1 ? cnorm(x[m - 1]) : 0.0;
    if (m_used) *m_used = m;
    if (eps_out) *eps_out = eps;
    return eps > emax ? 0 : (eps > emin ? 1 : 2);
}

int qor_step_free(qor_ctx* c, double* coef) {
    if (!c || !coef) return -1;
    hc_propagate0(c, (cplx*)coef);
    return 0;
}

int qor_mask(qor_ctx* c, double* coef_d) {
    if (!c || !coef_d) return -1;
    cplx* coef = (cplx*)coef_d;
    const long N = c->N, ne = c->ne, ML = c->M * c->L;
#pragma omp parallel
    {
        std::vector<cplx> t(N);
#pragma omp for schedule(static)
        for (long iml = 0; iml < ML; ++iml) {
            cplx* w = coef + iml * N;
            symb_mul(N, ne, c->mask.data(), w, t.data());
            ltrb_inv(N, ne, c->chol.data(), t.data(), t.data());
            utrb_inv(N, ne, c->chol.data(), t.data(), w);
        }
    }
    return 0;
}

double qor_norm(const qor_ctx* c, const double* coef) {
    const long N = c->N, ne = c->ne, ML = c->M * c->L;
    double nrm = 0.0;
    for (long iml = 0; iml < ML; ++iml) {
        const cplx* w = (const cplx*)coef + iml * N;
        nrm += symb_prj(N, ne, c->S.data(), w, w).re;
    }
    return nrm;
}

double qor_outer_norm(const qor_ctx* c, const double* coef, double r0) {
    std::vector<double> sw;
    band_0r0(c->ax, sw, [r0](double r) { return r >= r0 ? 1.0 : 0.0; });
    const long N = c->N, ne = c->ne, ML = c->M * c->L;
    double nrm = 0.0;
    for (long iml = 0; iml < ML; ++iml) {
        const cplx* w = (const cplx*)coef + iml * N;
        nrm += symb_prj(N, ne, sw.data(), w, w).re;
    }
    return nrm;
}

// observable_z::observe (observable/1e.hpp:26-49)
double qor_dipole_z(const qor_ctx* c, const double* coef, int which) {
    const long N = c->N, ne = c->ne, L = c->L, M = c->M;
    const double* band = which ? c->dvdr.data() : c->rpow1.data();
    const cplx* w = (const cplx*)coef;
    cplx obsv = czero;
    for (long im = 0; im < M; ++im)
        for (long il = 0; il < L; ++il) {
            const cplx* lhs = w + (im * L + il) * N;
            const cplx* rhs = w + im * L * N;
            if (il > 0) obsv += symb_prj(N, ne, band, lhs, rhs + N * (il - 1)) * c->y10[im * L + il - 1];
            if (il + 1 < L) obsv += symb_prj(N, ne, band, lhs, rhs + N * (il + 1)) * c->y10[im * L + il];
        }
    return obsv.re;
}

// observable_r::observe_rm / observe_rp (observable/1e.hpp:76-126) with
// operator_y11<-1> (outer sinexp1, inner sinexp2) and operator_y11<+1> (outer
// sinexp3, inner sinexp4) from the closed forms of integrator_angu
// (utilities/basisfunc.hpp:243-279), and the index map in_check<dm, dl>
// (dimension/1e.hpp:97-100); <x> = Re(rm + rp)/2, <y> = Im(rp - rm)/2
// (observable/1e.hpp:72-73, sphTDSE.cpp:524-527).
static double sinexp(int which, long m, long l) {
    auto q = [](double a, double b) { return b != 0.0 ? a / b : 0.0; };
    switch (which) {
        case 1: return -std::sqrt(q((double)((l + m + 2) * (l + m + 1)), (double)((2 * l + 1) * (2 * l + 3))));
        case 2: return +std::sqrt(q((double)((l - m) * (l - m - 1)), (double)((2 * l + 1) * (2 * l - 1))));
        case 3: return -std::sqrt(q((double)((l + m) * (l + m - 1)), (double)((2 * l + 1) * (2 * l - 1))));
        default: return +std::sqrt(q((double)((l - m + 2) * (l - m + 1)), (double)((2 * l + 1) * (2 * l + 3))));
    }
}
// in_check<msh, lsh>: im2 = im + msh in [0, M) and il2 = l(im, il) + lsh - |m(im2)| in [0, L)
static bool in_check(const qor_ctx* c, long im, long il, long msh, long lsh, long* im2, long* il2) {
    *im2 = im + msh;
    if (*im2 < 0 || *im2 >= c->M) return false;
    *il2 = c->in_l(im, il) + lsh - std::labs(*im2 + c->m0);
    return *il2 >= 0 && *il2 < c->L;
}
int qor_dipole_xy(const qor_ctx* c, const double* coef, int which, double* x, double* y) {
    if (!c || !coef || !x || !y) return -1;
    const long N = c->N, ne = c->ne, L = c->L, M = c->M;
    const double* band = which ? c->dvdr.data() : c->rpow1.data();
    const cplx* w = (const cplx*)coef;
    cplx rm = czero, rp = czero;
    long im2, il2;
    for (long im = 0; im + 1 < M; ++im)  // observe_rm (:76-100)
        for (long il = 0; il < L; ++il) {
            const long m = im + c->m0, l = c->in_l(im, il);
            const cplx* lhs = w + (im * L + il) * N;
            if (in_check(c, im, il, +1, +1, &im2, &il2))
                rm += symb_prj(N, ne, band, lhs, w + (im2 * L + il2) * N) * sinexp(1, m, l);
            if (in_check(c, im, il, +1, -1, &im2, &il2))
                rm += symb_prj(N, ne, band, lhs, w + (im2 * L + il2) * N) * sinexp(2, m, l);
        }
    for (long im = 1; im < M; ++im)  // observe_rp (:102-126)
        for (long il = 0; il < L; ++il) {
            const long m = im + c->m0, l = c->in_l(im, il);
            const cplx* lhs = w + (im * L + il) * N;
            if (in_check(c, im, il, -1, +1, &im2, &il2))
                rp += symb_prj(N, ne, band, lhs, w + (im2 * L + il2) * N) * sinexp(4, m, l);
            if (in_check(c, im, il, -1, -1, &im2, &il2))
                rp += symb_prj(N, ne, band, lhs, w + (im2 * L + il2) * N) * sinexp(3, m, l);
        }
    *x = (rm + rp).re / 2.0;
    *y = (rp - rm).im / 2.0;
    return 0;
}

// spectrum_c::measure_lmd (spectrum/momentum_1e.hpp:384-401):
// spec[(ikr*M + im)*L + il] = sqrt(2/pi) cps(ikr,l) sum_j c(im,il)_j ijl(ikr,l,j),
// l = in_l(im, il); ijl[(ikr*(lmax+1) + l)*N + j], cps complex [ikr*(lmax+1) + l]
int qor_measure_lmd(const qor_ctx* c, const double* coef, int64_t nkr, const double* ijl, const double* cps,
                    double* out) {
    if (!c || !coef || !ijl || !cps || !out) return -1;
    const long N = c->N, L = c->L, M = c->M, nl = c->lmax + 1;
    const cplx* w = (const cplx*)coef;
    const cplx* cp = (const cplx*)cps;
    cplx* o = (cplx*)out;
#pragma omp parallel for schedule(static)
    for (long ikr = 0; ikr < nkr; ++ikr)
        for (long im = 0; im < M; ++im)
            for (long il = 0; il < L; ++il) {
                const long l = c->in_l(im, il);
                const cplx* cj = w + (im * L + il) * N;
                const double* f = ijl + ((size_t)ikr * nl + l) * N;
                cplx tmq = czero;
                for (long j = 0; j < N; ++j) tmq += cj[j] * f[j];
                tmq = tmq * cp[ikr * nl + l] * 0.797884560802865;  // sqrt(2/pi)
                o[(ikr * M + im) * L + il] = tmq;
            }
    return 0;
}

// spectrum_c_tsf::prepare + accumulate, z-polarised (momentum_1e.hpp:609-627,
// 675-719); a field-free step is accumulate(dt, coef) (:631-673), which with
// disx = disy = 0 is the same with az0 = az1 = 0.  temp/temq = sum_i coef(i)
// b0/br over the non-zero B-splines at r0; then time += dt,
// disz += (az0 + az1) dt / 2, and for every (kr, kt, kp)
//   sum = sum_(im,il) y0 e0 (-i)^(l+1) (jdb - i az1 (j0 temp(l-1) pm(l-1) + j0 temp(l+1) pm(l))),
//   jdb = -(j0 temq - jr temp) / 2,
//   work += sum exp(i (k^2 time / 2 + k cos(kt) disz)) dt sqrt(2/pi).
// state = {time, disz}; work complex [nkr][nkt][nkp].
int qor_tsurff_accumulate(const qor_ctx* c, const double* coef, double az0, double az1, double dt, int64_t nkr,
                          int64_t nkt, int64_t nkp, const double* kr, const double* kt, int64_t nb, const int64_t* ib,
                          const double* b0, const double* br, const double* j0, const double* jr, const double* y0,
                          const double* e0, const double* pmat0, double* state, double* work) {
    if (!c || !coef || !kr || !kt || !ib || !b0 || !br || !j0 || !jr || !y0 || !e0 || !pmat0 || !state || !work)
        return -1;
    const long N = c->N, L = c->L, M = c->M, nl = c->lmax + 1;
    const cplx* w = (const cplx*)coef;
    const cplx* ep = (const cplx*)e0;
    cplx* out = (cplx*)work;
    std::vector<cplx> temp((size_t)M * L), temq((size_t)M * L);
    for (long im = 0; im < M; ++im)
        for (long il = 0; il < L; ++il) {
            const cplx* cj = w + (im * L + il) * N;
            cplx s0 = czero, s1 = czero;
            for (long i = 0; i < nb; ++i) s0 += cj[ib[i]] * b0[i];
            for (long i = 0; i < nb; ++i) s1 += cj[ib[i]] * br[i];
            temp[im * L + il] = s0;
            temq[im * L + il] = s1;
        }
    state[0] += dt;
    state[1] += 0.5 * (az0 + az1) * dt;
    const double time = state[0], disz = state[1];
    const cplx iaz = {0.0, az1};
    static const cplx pmi[4] = {{0.0, -1.0}, {-1.0, 0.0}, {0.0, 1.0}, {1.0, 0.0}};  // (-i)^(l+1)
#pragma omp parallel for schedule(static)
    for (long ikr = 0; ikr < nkr; ++ikr)
        for (long ikt = 0; ikt < nkt; ++ikt)
            for (long ikp = 0; ikp < nkp; ++ikp) {
                cplx sum = czero;
                for (long im = 0; im < M; ++im)
                    for (long il = 0; il < L; ++il) {
                        const long l = c->in_l(im, il);
                        const cplx y0val = (y0[(ikt * M + im) * L + il] * ep[ikp * M + im]) * pmi[l % 4];
                        const double j0val = j0[ikr * nl + l], jrval = jr[ikr * nl + l];
                        const cplx b0_prev = il > 0 ? temp[im * L + il - 1] : czero;
                        const double pm_prev = il > 0 ? pmat0[im * L + il - 1] : 0.0;
                        const cplx b0_this = temp[im * L + il], br_this = temq[im * L + il];
                        const cplx b0_next = il + 1 < L ? temp[im * L + il + 1] : czero;
                        const double pm_next = il + 1 < L ? pmat0[im * L + il] : 0.0;
                        const cplx jdb = -0.5 * (j0val * br_this - jrval * b0_this);
                        const cplx jbp = j0val * b0_prev, jbn = j0val * b0_next;
                        sum += y0val * (jdb - iaz * jbp * pm_prev - iaz * jbn * pm_next);
                    }
                const double phase = 0.5 * kr[ikr] * kr[ikr] * time + kr[ikr] * std::cos(kt[ikt]) * disz;
                out[(ikr * nkt + ikt) * nkp + ikp] += sum * cplx{std::cos(phase), std::sin(phase)} * dt *
                                                      0.797884560802865;
            }
    return 0;
}

// spectrum_c_tsf::accumulate (xy) (momentum_1e.hpp:716-776) and, with
// ax = ay = 0, the field-free form of the elliptically polarised run (:631-673,
// disz = 0).  time += dt, disx/disy += the mean of the step's end values
// times dt; per (kr, kt, kp) the channel term is jdb minus the A.r couplings
// to the (m-+1, l-+1) neighbours found by dimension_1e::in_check
// (dimension/1e.hpp:97-100), with iap = i (ax1 + i ay1) / 2 for m+1 and
// ias = i (ax1 - i ay1) / 2 for m-1 and the sin(th) e^(-i ph) elements
// pmat1 (outer, l+1) and pmat2 (inner, l-1) of integrate_sinexp1/2
// (basisfunc.hpp:241-259); the phase is k^2 time/2 + k sin(kt) (cos(kp) disx +
// sin(kp) disy).  state = {time, disx, disy}.
int qor_tsurff_accumulate_xy(const qor_ctx* c, const double* coef, double ax0, double ax1, double ay0, double ay1,
                             double dt, int64_t nkr, int64_t nkt, int64_t nkp, const double* kr, const double* kt,
                             const double* kp, int64_t nb, const int64_t* ib, const double* b0, const double* br,
                             const double* j0, const double* jr, const double* y0, const double* e0,
                             const double* pmat1, const double* pmat2, double* state, double* work) {
    if (!c || !coef || !kr || !kt || !kp || !ib || !b0 || !br || !j0 || !jr || !y0 || !e0 || !pmat1 || !pmat2 ||
        !state || !work)
        return -1;
    const long N = c->N, L = c->L, M = c->M, nl = c->lmax + 1;
    const cplx* w = (const cplx*)coef;
    const cplx* ep = (const cplx*)e0;
    cplx* out = (cplx*)work;
    std::vector<cplx> temp((size_t)M * L), temq((size_t)M * L);
    for (long im = 0; im < M; ++im)
        for (long il = 0; il < L; ++il) {
            const cplx* cj = w + (im * L + il) * N;
            cplx s0 = czero, s1 = czero;
            for (long i = 0; i < nb; ++i) s0 += cj[ib[i]] * b0[i];
            for (long i = 0; i < nb; ++i) s1 += cj[ib[i]] * br[i];
            temp[im * L + il] = s0;
            temq[im * L + il] = s1;
        }
    const cplx I = {0.0, 1.0};
    const cplx iap = (cplx{ax1, 0.0} + cplx{0.0, ay1}) * I * 0.5;
    const cplx ias = (cplx{ax1, 0.0} - cplx{0.0, ay1}) * I * 0.5;
    state[0] += dt;
    state[1] += 0.5 * (ax0 + ax1) * dt;
    state[2] += 0.5 * (ay0 + ay1) * dt;
    const double time = state[0], disx = state[1], disy = state[2];
    // neighbour (jm, jl) of (im, il) at m + dm, l + dl, if inside the grid
    auto nb_of = [&](long im, long il, long dm, long dl, long& jm, long& jl) {
        jm = im + dm;
        if (jm < 0 || jm >= M) return false;
        jl = c->in_l(im, il) + dl - std::labs(c->m0 + jm);
        return jl >= 0 && jl < L;
    };
    static const cplx pmi[4] = {{0.0, -1.0}, {-1.0, 0.0}, {0.0, 1.0}, {1.0, 0.0}};  // (-i)^(l+1)
#pragma omp parallel for schedule(static)
    for (long ikr = 0; ikr < nkr; ++ikr)
        for (long ikt = 0; ikt < nkt; ++ikt)
            for (long ikp = 0; ikp < nkp; ++ikp) {
                cplx sum = czero;
                for (long im = 0; im < M; ++im)
                    for (long il = 0; il < L; ++il) {
                        const long l = c->in_l(im, il);
                        const cplx y0val = (y0[(ikt * M + im) * L + il] * ep[ikp * M + im]) * pmi[l % 4];
                        const double j0val = j0[ikr * nl + l], jrval = jr[ikr * nl + l];
                        const cplx b0 = temp[im * L + il], brv = temq[im * L + il];
                        cplx t = -0.5 * (j0val * brv - jrval * b0);
                        long jm, jl;
                        if (nb_of(im, il, -1, -1, jm, jl)) t -= ias * j0val * temp[jm * L + jl] * pmat1[jm * L + jl];
                        if (nb_of(im, il, -1, +1, jm, jl)) t -= ias * j0val * temp[jm * L + jl] * pmat2[jm * L + jl];
                        if (nb_of(im, il, +1, -1, jm, jl)) t -= iap * j0val * temp[jm * L + jl] * pmat2[im * L + il];
                        if (nb_of(im, il, +1, +1, jm, jl)) t -= iap * j0val * temp[jm * L + jl] * pmat1[im * L + il];
                        sum += y0val * t;
                    }
                const double k = kr[ikr], st = std::sin(kt[ikt]);
                const double phase = 0.5 * k * k * time + k * st * std::cos(kp[ikp]) * disx +
                                     k * st * std::sin(kp[ikp]) * disy;
                out[(ikr * nkt + ikt) * nkp + ikp] += sum * cplx{std::cos(phase), std::sin(phase)} * dt *
                                                      0.797884560802865;
            }
    return 0;
}

int qor_run(qor_ctx* c, double* coef, const double* atab, int64_t nt_field, int64_t it0, int64_t it1,
            int use_mask, double* obs) {
    if (!c || !coef) return -1;
    for (int64_t it = it0; it < it1; ++it) {
        if (obs) {
            obs[2 * (it - it0) + 0] = qor_norm(c, coef);
            obs[2 * (it - it0) + 1] = qor_dipole_z(c, coef, 0);
        }
        if (it < nt_field) qor_step_vgwz(c, coef, atab[it], atab[it + 1]);
        else qor_step_free(c, coef);
        if (use_mask) qor_mask(c, coef);
    }
    return 0;
}

int qor_run_record(qor_ctx* c, double* coef, const double* atab, int64_t nt_field, int64_t it0, int64_t it1,
                   int use_mask, double r0, double* obs8) {
    if (!c || !coef) return -1;
    for (int64_t it = it0; it < it1; ++it) {
        if (obs8) {
            double* o = obs8 + 8 * (it - it0);
            o[0] = qor_norm(c, coef);
            o[1] = qor_dipole_z(c, coef, 0);
            o[2] = qor_dipole_z(c, coef, 1);
            o[3] = qor_outer_norm(c, coef, r0);
            qor_dipole_xy(c, coef, 0, &o[4], &o[5]);
            qor_dipole_xy(c, coef, 1, &o[6], &o[7]);
        }
        if (it < nt_field) qor_step_vgwz(c, coef, atab[it], atab[it + 1]);
        else qor_step_free(c, coef);
        if (use_mask) qor_mask(c, coef);
    }
    return 0;
}

double qor_sinen2(double e0, double w0, double nc, double n0, double c0, double t) {
    // field_component_sinen<2>::initialize / operator() (field.h:99-123)
    double para_pd = 6.283185307179586 / w0;
    double para_cep = 3.141592653589793 * c0;
    double para_tau = nc * para_pd;
    double para_ini = n0 * para_pd;
    double ts = (t - para_ini) / para_tau;
    double ev = t > para_ini && t < (para_ini + para_tau) ? std::pow(std::sin(3.1415926535897932 * ts), 2) : 0.0;
    double ca = e0 * std::sin(w0 * t + para_cep);
    return ca * ev;
}

int64_t qor_para_nt(double w0, double nc, double n0, double dt) {
    double para_pd = 6.283185307179586 / w0;
    double ti = n0 * para_pd;
    double tf = ti + nc * para_pd;
    return (int64_t)(1ul + (size_t)((tf - ti) / dt));  // field.h:279
}

int qor_lu_zsb(int64_t n, int64_t p, double* band) { return lu_zsb(n, p, (cplx*)band); }

void qor_cn_kernel(int64_t n, int64_t p, int symflag, const double* smat, const double* rmat, double ifet_re,
                   double ifet_im, double* z) {
    std::vector<cplx> ytmp(n), utmp(n * p);
    cn_kernel(n, p, symflag, smat, rmat, cplx{ifet_re, ifet_im}, (cplx*)z, ytmp.data(), utmp.data());
}

void qor_symb_mul(int64_t n, int64_t p, const double* cband, const double* x, double* y) {
    symb_mul(n, p + 1, (const cplx*)cband, (const cplx*)x, (cplx*)y);
}

void qor_ldl_solve(int64_t n, int64_t p, const double* factor, double* x) {
    ltrb_ldl(n, p + 1, (const cplx*)factor, (const cplx*)x, (cplx*)x);
    utrb_ldl(n, p + 1, (const cplx*)factor, (const cplx*)x, (cplx*)x);
}

int qor_chol(int64_t n, int64_t p, double* band) { return chol_band(n, p, band); }

int qor_tridiag_eig(int64_t n, const double* diag, const double* off, double* eig, double* z) {
    return tridiag_eig(n, diag, off, eig, z);
}

void qor_gauss_legendre(int64_t n, double* x, double* w) {
    std::vector<long double> xp, wp;
    gl_half((int)n, xp, wp);
    for (int64_t i = 0; i < n; ++i) gl_point((int)n, xp, wp, -1.0, 1.0, (int)i, &x[i], &w[i]);
}

}  // extern "C"
