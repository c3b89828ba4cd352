// qpc_setup.cpp -- host-side operator set-up of the propagator (runs once).
//
// Builds, on the host and in the reference's layouts, everything the device
// time loop consumes: the B-spline radial basis and its Gauss-Legendre
// quadrature, the radial bands S, chol(S), 1/r, d/dr, r, <dV/dr>, the
// absorber band, H_l for l = 0..lmax, the cn_component_hc bands
// (S -/+ i dt/2 H_l, the latter LDL^T-factorised), the angular eigen-data of
// the velocity-gauge propagator and field-free eigenstates.  This is the
// host half of propagator_cn_hc_vgwz::initialize (propagator/cn_1e.hpp:642-680),
// absorber_mask::initialize (absorber.hpp:42-49) and
// coefficient::initialize (coefficient/1e.hpp:89-143) of the reference.
#include "qpc_setup.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#include <thread>
#include <array>
#include <atomic>
#include <complex>

namespace qpc {
namespace {

constexpr double kPi = 3.14159265358979323846;

// QL with implicit shifts on the symmetric tridiagonal (d, e), e[i] coupling
// i and i+1, accumulating the plane rotations into the rows of V (V[i*n+k]:
// component k of vector i, so a rotation streams two contiguous rows).  V
// enters as the transformation that produced the tridiagonal (the identity
// for a tridiagonal problem) and leaves as the eigenvectors; d leaves as the
// eigenvalues (unordered).
int ql_implicit(int n, std::vector<double>& d, std::vector<double>& e, double* V) {
    e.resize(n, 0.0);
    for (int l = 0; l < n; ++l) {
        for (int iter = 0;; ++iter) {
            int m = l;
            for (; m < n - 1; ++m) {
                double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= 1e-300 || std::fabs(e[m]) <= 1.1102230246251565e-16 * dd) break;
            }
            if (m == l) break;
            if (iter > 300) return -1;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
            double s = 1.0, c = 1.0, p = 0.0;
            bool underflow = false;
            for (int i = m - 1; i >= l; --i) {
                double f = s * e[i], b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    underflow = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                double* vi = V + (size_t)i * n;
                double* vj = V + (size_t)(i + 1) * n;
                for (int k = 0; k < n; ++k) {
                    const double t = vj[k];
                    vj[k] = s * vi[k] + c * t;
                    vi[k] = c * vi[k] - s * t;
                }
            }
            if (underflow) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    return 0;
}

// eigenvalues ascending (stable) and eigenvectors as rows z[k*n+i]
void sort_eigen(int n, const std::vector<double>& d, const std::vector<double>& V, std::vector<double>& w,
                std::vector<double>& z) {
    std::vector<int> ord(n);
    for (int i = 0; i < n; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return d[a] < d[b]; });
    w.assign(n, 0.0);
    z.assign((size_t)n * n, 0.0);
    for (int k = 0; k < n; ++k) {
        w[k] = d[ord[k]];
        std::copy(V.begin() + (size_t)ord[k] * n, V.begin() + (size_t)(ord[k] + 1) * n, z.begin() + (size_t)k * n);
    }
}

// symmetric tridiagonal eigen-decomposition; eigenvalues ascending and
// eigenvectors as rows z[k*n+i].  Used for Gauss-Legendre (Golub-Welsch) and
// for the angular operators.
int tridiag_ql(int n, std::vector<double> d, std::vector<double> e, std::vector<double>& w, std::vector<double>& z) {
    std::vector<double> V((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
    if (ql_implicit(n, d, e, V.data())) return -1;
    sort_eigen(n, d, V, w, z);
    return 0;
}

// dense real symmetric eigen-decomposition (A row-major, full storage):
// Householder tridiagonalisation A = Q T Q^T (Golub & Van Loan, Matrix
// Computations, Alg. 8.3.1), Q accumulated backwards, then QL on T with the
// rotations applied to the rows of Q^T.  Eigenvalues ascending, eigenvectors
// as rows z[k*n+i].
int sym_eigen(int n, std::vector<double> A, std::vector<double>& w, std::vector<double>& z) {
    std::vector<double> d(n, 0.0), e(n, 0.0), beta(n, 0.0), vs((size_t)n * n, 0.0), p(n), q(n);
    for (int k = 0; k + 2 < n; ++k) {
        const int m = n - k - 1;  // trailing block rows/cols k+1 .. n-1
        double* v = &vs[(size_t)k * n];
        double alpha = 0.0;
        for (int i = 0; i < m; ++i) {
            v[i] = A[(size_t)(k + 1 + i) * n + k];
            alpha += v[i] * v[i];
        }
        alpha = std::sqrt(alpha);
        if (alpha == 0.0) {
            e[k] = 0.0;
            continue;
        }
        const double x0 = v[0];
        const double v0 = x0 + std::copysign(alpha, x0);
        for (int i = 1; i < m; ++i) v[i] /= v0;
        v[0] = 1.0;
        double vv = 0.0;
        for (int i = 0; i < m; ++i) vv += v[i] * v[i];
        const double b = 2.0 / vv;
        beta[k] = b;
        e[k] = -std::copysign(alpha, x0);
        // p = b A22 v, q = p - (b/2)(p.v) v, A22 -= v q^T + q v^T
        for (int i = 0; i < m; ++i) {
            const double* a = &A[(size_t)(k + 1 + i) * n + k + 1];
            double t = 0.0;
            for (int j = 0; j < m; ++j) t += a[j] * v[j];
            p[i] = b * t;
        }
        double pv = 0.0;
        for (int i = 0; i < m; ++i) pv += p[i] * v[i];
        for (int i = 0; i < m; ++i) q[i] = p[i] - 0.5 * b * pv * v[i];
        for (int i = 0; i < m; ++i) {
            double* a = &A[(size_t)(k + 1 + i) * n + k + 1];
            const double vi = v[i], qi = q[i];
            for (int j = 0; j < m; ++j) a[j] -= vi * q[j] + qi * v[j];
        }
    }
    for (int i = 0; i < n; ++i) d[i] = A[(size_t)i * n + i];
    if (n >= 2) e[n - 2] = A[(size_t)(n - 1) * n + n - 2];
    // Q = H_0 H_1 ... H_{n-3}, built backwards on the trailing blocks
    std::vector<double> Q((size_t)n * n, 0.0), s(n);
    for (int i = 0; i < n; ++i) Q[(size_t)i * n + i] = 1.0;
    for (int k = n - 3; k >= 0; --k) {
        if (beta[k] == 0.0) continue;
        const int m = n - k - 1;
        const double* v = &vs[(size_t)k * n];
        std::fill(s.begin(), s.end(), 0.0);
        for (int r = 0; r < m; ++r) {
            const double* qr = &Q[(size_t)(k + 1 + r) * n + k + 1];
            for (int c = 0; c < m; ++c) s[c] += v[r] * qr[c];
        }
        for (int r = 0; r < m; ++r) {
            double* qr = &Q[(size_t)(k + 1 + r) * n + k + 1];
            const double f = beta[k] * v[r];
            for (int c = 0; c < m; ++c) qr[c] -= f * s[c];
        }
    }
    // rows of V = columns of Q
    std::vector<double> V((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) V[(size_t)k * n + i] = Q[(size_t)i * n + k];
    e.resize(n - 1 > 0 ? n - 1 : 0);
    if (ql_implicit(n, d, e, V.data())) return -1;
    sort_eigen(n, d, V, w, z);
    return 0;
}

// run f(i) for i in [0, n) on up to hw threads
template <class F>
void parallel_for(int n, F f) {
    const int hw = std::max(1, std::min(n, (int)std::thread::hardware_concurrency()));
    std::atomic<int> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < hw; ++t)
        th.emplace_back([&] {
            for (int i; (i = next.fetch_add(1)) < n;) f(i);
        });
    for (auto& x : th) x.join();
}

// Gauss-Legendre on [-1,1] by Golub-Welsch, polished by one Newton step on
// P_n; nodes ascending.
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
    std::vector<double> d(n, 0.0), e(n, 0.0), ev, z;
    for (int k = 1; k < n; ++k) e[k - 1] = k / std::sqrt(4.0 * k * k - 1.0);
    tridiag_ql(n, d, e, ev, z);
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int k = 0; k < n; ++k) {
        long double t = ev[k];
        long double p0 = 1.0L, p1 = t, dp = 1.0L;
        for (int it = 0; it < 3; ++it) {
            p0 = 1.0L;
            p1 = t;
            for (int j = 2; j <= n; ++j) {
                long double p2 = ((2 * j - 1) * t * p1 - (j - 1) * p0) / j;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (t * p1 - p0) / (t * t - 1.0L);
            t -= p1 / dp;
        }
        p0 = 1.0L;
        p1 = t;
        for (int j = 2; j <= n; ++j) {
            long double p2 = ((2 * j - 1) * t * p1 - (j - 1) * p0) / j;
            p0 = p1;
            p1 = p2;
        }
        dp = n * (t * p1 - p0) / (t * t - 1.0L);
        x[k] = (double)t;
        w[k] = (double)(2.0L / ((1.0L - t * t) * dp * dp));
    }
    // enforce the exact symmetry of the rule
    for (int k = 0; k < n / 2; ++k) {
        double a = 0.5 * (x[n - 1 - k] - x[k]);
        double b = 0.5 * (w[n - 1 - k] + w[k]);
        x[k] = -a;
        x[n - 1 - k] = a;
        w[k] = w[n - 1 - k] = b;
    }
    if (n % 2) x[n / 2] = 0.0;
}

// B-spline basis of degree p on the reference's linear knot sequence
// (bss::bsplines<N,p,0,0>::axis::initialize_axis_linear_step,
// basis_bsplines.h:417-459): knots t_{-1..p-1} = rmin, t_k = rmin + (k-p+1)h,
// t_{N+1..N+p+1} = rmax, h = (rmax-rmin)/(N+2-p); B-splines -1 and N dropped.
struct Basis {
    int N = 0, p = 0, q = 0;
    std::vector<double> t;          // t[k+1] = knot k
    std::vector<double> gx, gw;     // per active interval s = 0..N+p-1, q points
    std::vector<double> b0, b1;     // [(s*q+k)*(p+1)+r]: spline s-p+r, value / derivative

    double knot(int k) const { return t[k + 1]; }

    // all p+1 non-zero B-splines of degree p and their first derivatives at x
    // in knot span [knot(s), knot(s+1)) (de Boor's triangular scheme)
    void eval(int s, double x, double* val, double* der) const {
        std::vector<double> left(p + 1), right(p + 1);
        std::vector<std::vector<double>> ndu(p + 1, std::vector<double>(p + 1, 0.0));
        ndu[0][0] = 1.0;
        for (int j = 1; j <= p; ++j) {
            left[j] = x - knot(s + 1 - j);
            right[j] = knot(s + j) - x;
            double saved = 0.0;
            for (int r = 0; r < j; ++r) {
                ndu[j][r] = right[r + 1] + left[j - r];
                double tmp = ndu[j][r] == 0.0 ? 0.0 : ndu[r][j - 1] / ndu[j][r];
                ndu[r][j] = saved + right[r + 1] * tmp;
                saved = left[j - r] * tmp;
            }
            ndu[j][j] = saved;
        }
        for (int r = 0; r <= p; ++r) val[r] = ndu[r][p];
        // first derivative: p * (N_{r-1,p-1}/(t_{r+p}-t_r) - N_{r,p-1}/(t_{r+p+1}-t_{r+1}))
        for (int r = 0; r <= p; ++r) {
            double d = 0.0;
            int ir = s - p + r;  // spline index
            if (r >= 1) {
                double den = knot(ir + p) - knot(ir);
                if (den != 0.0) d += ndu[r - 1][p - 1] / den;
            }
            if (r <= p - 1) {
                double den = knot(ir + p + 1) - knot(ir + 1);
                if (den != 0.0) d -= ndu[r][p - 1] / den;
            }
            der[r] = p * d;
        }
    }

    void build(int N_, int p_, double rmin, double rmax) {
        N = N_;
        p = p_;
        q = 2 * p + 2;  // Gauss points per interval (basis_bsplines.h:110)
        const int nk = N + p + 3;
        t.assign(nk, 0.0);
        const double h = (rmax - rmin) / (double)(N + 2 - p);
        for (int k = 0; k < nk; ++k) {
            if (k < p) t[k] = rmin;
            else if (k < N + 2) t[k] = rmin + (double)(k - p) * h;
            else t[k] = rmax;
        }
        std::vector<double> xg, wg;
        gauss_legendre(q, xg, wg);
        const int ns = N + p;
        gx.assign((size_t)ns * q, 0.0);
        gw.assign((size_t)ns * q, 0.0);
        b0.assign((size_t)ns * q * (p + 1), 0.0);
        b1.assign((size_t)ns * q * (p + 1), 0.0);
        std::vector<double> v(p + 1), dv(p + 1);
        for (int s = 0; s < ns; ++s) {
            const double a = knot(s), b = knot(s + 1);
            if (!(b > a)) continue;
            const double half = 0.5 * (b - a), mid = 0.5 * (a + b);
            for (int k = 0; k < q; ++k) {
                const double x = mid + half * xg[k];
                gx[(size_t)s * q + k] = x;
                gw[(size_t)s * q + k] = half * wg[k];
                eval(s, x, v.data(), dv.data());
                for (int r = 0; r <= p; ++r) {
                    const int idx = s - p + r;
                    if (idx < 0 || idx >= N) continue;  // the two dropped end splines
                    b0[((size_t)s * q + k) * (p + 1) + r] = v[r];
                    b1[((size_t)s * q + k) * (p + 1) + r] = dv[r];
                }
            }
        }
    }

    // band[i*ne + c] = int B_i^(mi) f B_{i+c}^(mj) dr, zero for i+c >= N
    template <class F>
    void band(std::vector<double>& out, int mi, int mj, const F& f) const {
        const int ne = p + 1;
        out.assign((size_t)N * ne, 0.0);
        const std::vector<double>& L = mi ? b1 : b0;
        const std::vector<double>& R = mj ? b1 : b0;
        for (int s = 0; s < N + p; ++s) {
            if (gw[(size_t)s * q] == 0.0) continue;
            for (int k = 0; k < q; ++k) {
                const double x = gx[(size_t)s * q + k];
                const double wf = gw[(size_t)s * q + k] * f(x);
                const double* lv = &L[((size_t)s * q + k) * (p + 1)];
                const double* rv = &R[((size_t)s * q + k) * (p + 1)];
                for (int r = 0; r <= p; ++r) {
                    const int i = s - p + r;
                    if (i < 0 || i >= N) continue;
                    for (int c = 0; c <= p - r; ++c) {
                        const int j = i + c;
                        if (j >= N) break;
                        out[(size_t)i * ne + c] += wf * lv[r] * rv[r + c];
                    }
                }
            }
        }
    }
};

// complex-symmetric banded LDL^T without pivoting, in the lu_zsb format
// (algorithm/ed.h:208-289): f[i*ne] = D_i, f[i*ne+j] = L(i+j,i); returns the
// number of pivots with |D|^2 < 1e-31.
int ldlt_csym(int n, int p, cplx* a) {
    const int ne = p + 1;
    int tiny = 0;
    for (int j = n - p; j < n; ++j)
        for (int i = n - j; i < ne; ++i) a[(size_t)j * ne + i] = {0.0, 0.0};
    for (int j = 0; j < n; ++j) {
        cplx* aj = a + (size_t)j * ne;
        const int kmax = std::min(j, p);
        for (int k = 1; k <= kmax; ++k) {
            const cplx* ak = a + (size_t)(j - k) * ne;  // row j-k: D at [0], L(j-k+c, j-k) at [c]
            const cplx f = ldlt_mul(ak[k], ak[0]);       // L(j, j-k) D_(j-k)
            for (int i = 0; i + k < ne; ++i) ldlt_sub(aj[i], f, ak[k + i]);
        }
        if (ldlt_norm2(aj[0]) < 1e-31) ++tiny;
        const cplx r = ldlt_recip(aj[0]);
        for (int i = 1; i < ne; ++i) aj[i] = ldlt_mul(aj[i], r);
    }
    return tiny;
}

// real symmetric banded Cholesky S = L L^T in the dpbtrf 'L' layout
int chol_real(int n, int p, double* a) {
    const int ne = p + 1;
    for (int j = 0; j < n; ++j) {
        double* aj = a + (size_t)j * ne;
        const int kmax = std::min(j, p);
        for (int k = 1; k <= kmax; ++k) {
            const double* ak = a + (size_t)(j - k) * ne;
            const double lk = ak[k];
            for (int i = 0; i + k < ne && j + i < n; ++i) aj[i] -= lk * ak[k + i];
        }
        if (!(aj[0] > 0.0)) return j + 1;
        aj[0] = std::sqrt(aj[0]);
        for (int i = 1; i < ne; ++i) aj[i] = (j + i < n) ? aj[i] / aj[0] : 0.0;
    }
    return 0;
}

// real symmetric banded LDL^T of H - sigma S; returns the negative-pivot count
int inertia(int n, int p, const double* h, const double* s, double sigma, std::vector<double>& f) {
    const int ne = p + 1;
    f.assign((size_t)n * ne, 0.0);
    for (size_t i = 0; i < f.size(); ++i) f[i] = h[i] - sigma * s[i];
    int neg = 0;
    for (int j = 0; j < n; ++j) {
        double* aj = &f[(size_t)j * ne];
        const int kmax = std::min(j, p);
        for (int k = 1; k <= kmax; ++k) {
            const double* ak = &f[(size_t)(j - k) * ne];
            const double l = ak[k];
            const double g = l * ak[0];
            for (int i = 0; i + k < ne; ++i) aj[i] -= g * ak[k + i];
        }
        if (aj[0] < 0) ++neg;
        const double r = aj[0] != 0.0 ? 1.0 / aj[0] : 1e300;
        for (int i = 1; i < ne; ++i) aj[i] = (j + i < n) ? aj[i] * r : 0.0;
    }
    return neg;
}

void band_apply(int n, int p, const double* a, const double* x, double* y) {
    const int ne = p + 1;
    for (int i = 0; i < n; ++i) {
        double t = 0.0;
        for (int j = 1; j <= std::min(i, p); ++j) t += a[(size_t)(i - j) * ne + j] * x[i - j];
        for (int c = 0; c < ne && i + c < n; ++c) t += a[(size_t)i * ne + c] * x[i + c];
        y[i] = t;
    }
}

void ldl_solve(int n, int p, const double* f, double* x) {
    const int ne = p + 1;
    for (int i = 0; i < n; ++i)
        for (int j = 1; j <= std::min(i, p); ++j) x[i] -= f[(size_t)(i - j) * ne + j] * x[i - j];
    for (int i = n - 1; i >= 0; --i) {
        double v = x[i] / f[(size_t)i * ne];
        for (int c = 1; c < ne && i + c < n; ++c) v -= f[(size_t)i * ne + c] * x[i + c];
        x[i] = v;
    }
}

// eigenpair `in` (ascending) of the banded pencil H v = E S v, v^T S v = 1:
// Sylvester-inertia bisection for E, then shift-inverted iteration.
int pencil_eigen(int n, int p, const double* h, const double* s, int in, double& E, std::vector<double>& v) {
    std::vector<double> f;
    double lo = -1.0, hi = 1.0;
    for (int g = 0; g < 200 && inertia(n, p, h, s, lo, f) > in; ++g) lo *= 2;
    for (int g = 0; g < 200 && inertia(n, p, h, s, hi, f) <= in; ++g) hi *= 2;
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        (inertia(n, p, h, s, mid, f) > in ? hi : lo) = mid;
    }
    const double lam = 0.5 * (lo + hi);
    inertia(n, p, h, s, lam - 1e-9 * std::max(1.0, std::fabs(lam)), f);
    v.assign(n, 1.0);
    std::vector<double> t(n), u(n);
    for (int it = 0; it < 6; ++it) {
        band_apply(n, p, s, v.data(), t.data());
        ldl_solve(n, p, f.data(), t.data());
        band_apply(n, p, s, t.data(), u.data());
        double nn = 0.0;
        for (int i = 0; i < n; ++i) nn += t[i] * u[i];
        if (!(nn > 0.0)) return -1;
        nn = 1.0 / std::sqrt(nn);
        for (int i = 0; i < n; ++i) v[i] = t[i] * nn;
    }
    band_apply(n, p, h, v.data(), t.data());
    E = 0.0;
    for (int i = 0; i < n; ++i) E += v[i] * t[i];
    int imax = 0;
    for (int i = 0; i < n; ++i)
        if (std::fabs(v[i]) > std::fabs(v[imax])) imax = i;
    if (v[imax] < 0)
        for (double& x : v) x = -x;
    return 0;
}

}  // namespace
}  // namespace qpc

// ---------------------------------------------------------------------------
// Angular eigen-data of propagator_cn_hc_vgwr (cn_1e.hpp:850-895)

// closed forms of integrator_angu (utilities/basisfunc.hpp:243-302) for the
// couplings (m,l) -> (m+1,l+1) ("outer") and (m+1,l-1) ("inner")
static void ep_forms(long m, long l, double& se1, double& se2, double& cl1, double& cl2) {
    auto ratio = [](long a, long b) { return b != 0 ? double(a) / double(b) : 0.0; };
    se1 = -std::sqrt(ratio((l + m + 2) * (l + m + 1), (2 * l + 1) * (2 * l + 3)));
    se2 = +std::sqrt(ratio((l - m) * (l - m - 1), (2 * l + 1) * (2 * l - 1)));
    cl1 = -std::sqrt(ratio((l + m + 1) * (l + m + 2), (2 * l + 1) * (2 * l + 3))) * (double)(l + 1);
    cl2 = -std::sqrt(ratio((l - m) * (l - m - 1), (2 * l + 1) * (2 * l - 1))) * (double)l;
}

void qpc::angular_blocks(int M, int L, int m0, AngBlocks& ab) {
    const int ML = M * L;
    // operators x1 = Y11 (mult 1), x2 = Q11 (mult -i), y1 = Y11 (mult -i), y2 = Q11 (mult 1)
    // (cn_1e.hpp:878-881); Q11<-1> stores minus the closed form (operators/1e.hpp:500-513).
    // The lower panel A[(m+1,l+-1),(m,l)] = value*mult (:365-385); with mult = -i
    // the diagonal change D = diag(i^im) gives the real symmetric B = D^H A D,
    // B[(m+1,.),(m,.)] = -value, and eigenvectors D v_B.
    const bool is_q[4] = {false, true, false, true};
    ab.par.assign(ML, 0);
    ab.loc.assign(ML, 0);
    for (auto& v : ab.idx) v.clear();
    for (int im = 0; im < M; ++im)
        for (int il = 0; il < L; ++il) {
            const int i = im * L + il, pb = il % 2;
            ab.par[i] = pb;
            ab.loc[i] = (int)ab.idx[pb].size();
            ab.idx[pb].push_back(i);
        }
    // dense real symmetric matrix of each (operator, parity) block
    for (int op = 0; op < 4; ++op)
        for (int pb = 0; pb < 2; ++pb) ab.B[op][pb].assign(ab.idx[pb].size() * ab.idx[pb].size(), 0.0);
    for (int im = 0; im + 1 < M; ++im)
        for (int il = 0; il < L; ++il) {
            const long m = m0 + im, l = il + std::labs(m), m2 = m + 1;
            double se1, se2, cl1, cl2;
            ep_forms(m, l, se1, se2, cl1, cl2);
            const double vy[2] = {se1, se2}, vq[2] = {cl1, cl2};
            for (int t = 0; t < 2; ++t) {
                const long jl = l + (t == 0 ? 1 : -1) - std::labs(m2);  // dimension_1e::in_check (dimension/1e.hpp:93-97)
                if (jl < 0 || jl >= L) continue;
                const int i = im * L + il, j = (im + 1) * L + (int)jl;
                const int pb = ab.par[i];
                const size_t n = ab.idx[pb].size();
                for (int op = 0; op < 4; ++op) {
                    double v = is_q[op] ? -vq[t] : vy[t];
                    if (kAngImagMult[op]) v = -v;
                    ab.B[op][pb][(size_t)ab.loc[j] * n + ab.loc[i]] = v;
                    ab.B[op][pb][(size_t)ab.loc[i] * n + ab.loc[j]] = v;
                }
            }
        }
}

void qpc::pair_spectrum(double* w, int M, int L) {
    const int n_zero = M % 2 == 1 ? L : 0;
    const int n_even = (M / 2 * 2) * ((L + 1) / 2), ML = M * L;
    const int blk[2][2] = {{0, n_even}, {n_even + n_zero, ML}};
    for (const auto& b : blk) {
        const int n = b[1] - b[0];
        double* v = w + b[0];
        for (int k = 0; k < n / 2; ++k) {
            const double a = 0.5 * (v[n - 1 - k] - v[k]);
            v[k] = -a;
            v[n - 1 - k] = a;
        }
        if (n % 2) v[n / 2] = 0.0;
    }
    for (int k = n_even; k < n_even + n_zero; ++k) w[k] = 0.0;
}

int qpc::angular_rows(int M, int L, const std::vector<double> w[2], std::vector<AngRow>& rows) {
    // converter::catagorize (utilities/converter.hpp:27-107): even, zero, odd,
    // each in ascending order; zero only for odd M (|eig| < 1e-8)
    const int n_zero = M % 2 == 1 ? L : 0;
    const int n_even = (M / 2 * 2) * ((L + 1) / 2), n_odds = (M / 2 * 2) * (L / 2);
    std::vector<AngRow> cat[3];
    for (int pb = 0; pb < 2; ++pb)
        for (int k = 0; k < (int)w[pb].size(); ++k) {
            const double e = w[pb][k];
            const int c = (M % 2 == 1 && std::fabs(e) < 1e-8) ? 1 : (pb == 0 ? 0 : 2);
            cat[c].push_back({pb, k});
        }
    std::stable_sort(cat[1].begin(), cat[1].end(),
                     [&](const AngRow& a, const AngRow& b) { return w[a.pb][a.k] < w[b.pb][b.k]; });
    if ((int)cat[0].size() != n_even || (int)cat[1].size() != n_zero || (int)cat[2].size() != n_odds)
        return QPC_E_SETUP;
    rows.clear();
    for (auto& c : cat) rows.insert(rows.end(), c.begin(), c.end());
    return QPC_OK;
}

int qpc::angular_ep(int M, int L, int m0, std::array<double*, 4> eig, double* tran_x1, double* tran_y2,
                    std::array<double*, 3> merged, double* x1_2_y1, double* tran_y1) {
    const int ML = M * L;
    const bool* imag_mult = kAngImagMult;
    AngBlocks ab;
    angular_blocks(M, L, m0, ab);
    const auto& idx = ab.idx;
    auto& B = ab.B;
    std::vector<double> w[4][2], z[4][2];
    std::atomic<int> bad{0};
    parallel_for(8, [&](int t) {
        const int op = t / 2, pb = t % 2;
        if (sym_eigen((int)idx[pb].size(), B[op][pb], w[op][pb], z[op][pb])) ++bad;
    });
    if (bad) return QPC_E_SETUP;
    using Row = AngRow;
    std::vector<Row> rows[4];
    for (int op = 0; op < 4; ++op) {
        if (angular_rows(M, L, w[op], rows[op])) return QPC_E_SETUP;
        for (int r = 0; r < ML; ++r) eig[op][r] = w[op][rows[op][r].pb][rows[op][r].k];
        pair_spectrum(eig[op], M, L);
    }
    // component j of row r of the unconjugated projector (zhbev's eigenvector)
    auto phase = [&](int op, int j) -> cplx {
        if (!imag_mult[op]) return {1.0, 0.0};
        static const cplx ipow[4] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
        return ipow[(j / L) % 4];
    };
    // tran_x1, tran_y2: real projectors (copy_real, cn_1e.hpp:183-193)
    double* tr[2] = {tran_x1, tran_y2};
    const int trop[2] = {0, 3};
    for (int t = 0; t < 2; ++t) {
        const int op = trop[t];
        std::fill(tr[t], tr[t] + (size_t)ML * ML, 0.0);
        for (int r = 0; r < ML; ++r) {
            const Row& R = rows[op][r];
            const size_t n = idx[R.pb].size();
            for (size_t q = 0; q < n; ++q) tr[t][(size_t)r * ML + idx[R.pb][q]] = z[op][R.pb][(size_t)R.k * n + q];
        }
    }
    // tran_y1 of the length gauge: the conjugated complex projector of y1
    // (conjugate(), cn_1e.hpp:762), rows as catagorize orders them
    if (tran_y1) {
        cplx* out = reinterpret_cast<cplx*>(tran_y1);
        std::fill(out, out + (size_t)ML * ML, cplx{0.0, 0.0});
        for (int r = 0; r < ML; ++r) {
            const Row& R = rows[2][r];
            const size_t n = idx[R.pb].size();
            for (size_t q = 0; q < n; ++q) {
                const cplx v = phase(2, idx[R.pb][q]) * z[2][R.pb][(size_t)R.k * n + q];
                out[(size_t)r * ML + idx[R.pb][q]] = cplx{v.re, -v.im};
            }
        }
    }
    // merge_two (cn_1e.hpp:161-181): P_b P_a^H with P = conj(projector):
    // out[r][r'] = sum_j conj(prj_b[r][j]) prj_a[r'][j], nonzero only within a parity
    const int mg[4][2] = {{1, 0}, {2, 1}, {3, 2}, {2, 0}};  // (b, a): x1_2_x2, x2_2_y1, y1_2_y2, x1_2_y1
    double* const mout[4] = {merged[0], merged[1], merged[2], x1_2_y1};
    for (int t = 0; t < 4; ++t) {
        if (!mout[t]) continue;
        const int ob = mg[t][0], oa = mg[t][1];
        cplx* out = reinterpret_cast<cplx*>(mout[t]);
        std::fill(out, out + (size_t)ML * ML, cplx{0.0, 0.0});
        for (int pb = 0; pb < 2; ++pb) {
            const int n = (int)idx[pb].size();
            std::vector<int> rb, ra;  // global rows of each operator in this parity
            for (int r = 0; r < ML; ++r) {
                if (rows[ob][r].pb == pb) rb.push_back(r);
                if (rows[oa][r].pb == pb) ra.push_back(r);
            }
            // Yt[q][c] = prj_a[ra[c]][idx[q]] (split re/im), X rows conj(prj_b)
            std::vector<double> ytr((size_t)n * n), yti((size_t)n * n);
            for (int c = 0; c < n; ++c) {
                const Row& R = rows[oa][ra[c]];
                for (int q = 0; q < n; ++q) {
                    const cplx v = phase(oa, idx[pb][q]) * z[oa][pb][(size_t)R.k * n + q];
                    ytr[(size_t)q * n + c] = v.re;
                    yti[(size_t)q * n + c] = v.im;
                }
            }
            parallel_for(n, [&](int ir) {
                const Row& R = rows[ob][rb[ir]];
                std::vector<double> cr(n, 0.0), ci(n, 0.0);
                for (int q = 0; q < n; ++q) {
                    const cplx pv = phase(ob, idx[pb][q]) * z[ob][pb][(size_t)R.k * n + q];
                    const double xr = pv.re, xi = -pv.im;  // conj
                    if (xr == 0.0 && xi == 0.0) continue;
                    const double* yr = &ytr[(size_t)q * n];
                    const double* yi = &yti[(size_t)q * n];
                    for (int c = 0; c < n; ++c) {
                        cr[c] += xr * yr[c] - xi * yi[c];
                        ci[c] += xr * yi[c] + xi * yr[c];
                    }
                }
                cplx* o = out + (size_t)rb[ir] * ML;
                for (int c = 0; c < n; ++c) o[ra[c]] = cplx{cr[c], ci[c]};
            });
        }
    }
    return QPC_OK;
}

int qpc::tridiag_eigen(int n, const double* d, const double* e, double* w, double* z) {
    std::vector<double> dv(d, d + n), ev(e, e + (n > 1 ? n - 1 : 0)), wv, zv;
    if (tridiag_ql(n, dv, ev, wv, zv)) return -1;
    std::copy(wv.begin(), wv.end(), w);
    std::copy(zv.begin(), zv.end(), z);
    return 0;
}

namespace qpc {
// ---------------------------------------------------------------------------
// Regular Coulomb functions F_l(eta, rho), l = 0..lmax, by Steed's method
// (the continued fractions of Barnett, Feng, Steed & Goldfarb, Comput. Phys.
// Commun. 8 (1974) 377): CF1 gives F'/F at lmax (modified Lentz; the sign of
// F_lmax follows the sign changes of the denominators), the recurrences
//   F_{l-1} = (S_l F_l + F'_l) / R_l,   F'_{l-1} = S_l F_{l-1} - R_l F_l
// (S_l = l/rho + eta/l, R_l = sqrt(1 + eta^2/l^2)) carry F down to l = 0,
// CF2 gives p + iq = (G' + iF')/(G + iF) at l = 0,
//   p + iq = i(1 - eta/rho) + (i/rho) a b / (2(rho - eta + i) +) (a+1)(b+1) / (2(rho - eta + 2i) +) ...
// with a = i eta, b = 1 + i eta, and the Wronskian F'G - FG' = 1 fixes
// |F_0| = (q + (F'_0/F_0 - p)^2 / q)^(-1/2).  CF2 converges ever more slowly
// as rho -> 0; there (rho <= 0.05 and 2|eta| rho <= 2) F_0 comes from its
// power series instead (Abramowitz & Stegun 14.1.3-14.1.7):
//   F_0 = C_0 rho sum_n a_n rho^n,  a_0 = 1, a_1 = eta,
//   n(n+1) a_n = 2 eta a_{n-1} - a_{n-2},  C_0^2 = 2 pi eta / (exp(2 pi eta) - 1).
// Returns 0, or -1 when a continued fraction does not converge.
int coulomb_f(int lmax, double eta, double rho, double* F, double* Fd = nullptr);

static double coulomb_f0_series(double eta, double rho) {
    const double x = 2.0 * kPi * eta;
    const double C0 = x != 0.0 ? std::sqrt(x / std::expm1(x)) : 1.0;
    double a2 = 1.0, a1 = eta, s = 1.0 + eta * rho, pw = rho;
    int small = 0;
    for (int n = 2; n < 1000 && small < 2; ++n) {
        const double an = (2.0 * eta * a1 - a2) / (double)(n * (n + 1));
        pw *= rho;
        const double t = an * pw;
        s += t;
        a2 = a1;
        a1 = an;
        small = std::fabs(t) < 1e-17 * std::fabs(s) ? small + 1 : 0;
    }
    return C0 * rho * s;
}

int coulomb_f(int lmax, double eta, double rho, double* F, double* Fd) {
    const double tiny = 1e-300, acc = 1e-16;
    auto S = [&](double k) { return k / rho + eta / k; };
    auto R2 = [&](double k) { return 1.0 + eta * eta / (k * k); };
    const int L = lmax;
    double f = S(L + 1);
    if (f == 0.0) f = tiny;
    double C = f, D = 0.0, sgn = 1.0;
    bool ok = false;
    for (int j = 1; j < 4000000; ++j) {
        const double a = -R2(L + j), b = S(L + j) + S(L + j + 1);
        D = b + a * D;
        if (D == 0.0) D = tiny;
        C = b + a / C;
        if (C == 0.0) C = tiny;
        D = 1.0 / D;
        if (D < 0.0) sgn = -sgn;
        const double dl = C * D;
        f *= dl;
        if (std::fabs(dl - 1.0) < acc) {
            ok = true;
            break;
        }
    }
    if (!ok) return -1;
    std::vector<double> Fp(L + 1);
    F[L] = sgn * 1e-30;
    Fp[L] = f * F[L];
    for (int k = L; k > 0; --k) {
        const double Rk = std::sqrt(R2(k)), Sk = S(k);
        F[k - 1] = (Sk * F[k] + Fp[k]) / Rk;
        Fp[k - 1] = Sk * F[k - 1] - Rk * F[k];
        if (std::fabs(F[k - 1]) > 1e250)
            for (int q = k - 1; q <= L; ++q) {
                F[q] *= 1e-250;
                Fp[q] *= 1e-250;
            }
    }
    if (rho <= 0.05 && 2.0 * std::fabs(eta) * rho <= 2.0) {
        const double scale = coulomb_f0_series(eta, rho) / F[0];
        for (int k = 0; k <= L; ++k) F[k] *= scale;
        if (Fd)
            for (int k = 0; k <= L; ++k) Fd[k] = Fp[k] * scale;
        return 0;
    }
    const double f0 = Fp[0] / F[0];
    using cd = std::complex<double>;
    const cd a(0.0, eta), b(1.0, eta);
    cd fr(tiny, 0.0), Cc = fr, Dc(0.0, 0.0);
    ok = false;
    for (int k = 0; k < 4000000; ++k) {
        const cd ak = (a + (double)k) * (b + (double)k);
        const cd bk = 2.0 * cd(rho - eta, (double)(k + 1));
        Dc = bk + ak * Dc;
        if (Dc == 0.0) Dc = tiny;
        Cc = bk + ak / Cc;
        if (Cc == 0.0) Cc = tiny;
        Dc = 1.0 / Dc;
        const cd dl = Cc * Dc;
        fr *= dl;
        if (std::abs(dl - 1.0) < acc) {
            ok = true;
            break;
        }
    }
    if (!ok) return -1;
    const cd pq = cd(0.0, 1.0 - eta / rho) + cd(0.0, 1.0 / rho) * fr;
    const double p = pq.real(), q = pq.imag();
    const double F0 = std::copysign(1.0 / std::sqrt((f0 - p) * (f0 - p) / q + q), F[0]);
    const double scale = F0 / F[0];
    for (int k = 0; k <= L; ++k) F[k] *= scale;
    if (Fd)
        for (int k = 0; k <= L; ++k) Fd[k] = Fp[k] * scale;
    return 0;
}

// Coulomb phase sigma_l = arg Gamma(l + 1 + i eta), l = 0..lmax: sigma_0 from
// Stirling's series of ln Gamma(z) at z = 1 + n + i eta (n = 20) shifted down
// by ln Gamma(z) = ln Gamma(z + n) - sum ln(z + k), then
// sigma_l = sigma_{l-1} + atan2(eta, l).
void coulomb_sigma(int lmax, double eta, double* sig) {
    using cd = std::complex<double>;
    const int n = 20;
    const cd z(1.0 + n, eta);
    // ln Gamma(z) ~ (z - 1/2) ln z - z + ln(2 pi)/2 + sum B_2k / (2k(2k-1) z^(2k-1))
    static const double B[] = {1.0 / 6, -1.0 / 30, 1.0 / 42, -1.0 / 30, 5.0 / 66, -691.0 / 2730, 7.0 / 6};
    cd lg = (z - 0.5) * std::log(z) - z + 0.5 * std::log(2.0 * kPi);
    cd zp = z;
    for (int k = 1; k <= 7; ++k) {
        lg += B[k - 1] / ((2.0 * k) * (2.0 * k - 1.0) * zp);
        zp *= z * z;
    }
    double s = lg.imag();
    for (int k = 0; k < n; ++k) s -= std::atan2(eta, 1.0 + k);
    sig[0] = s;
    for (int l = 1; l <= lmax; ++l) sig[l] = sig[l - 1] + std::atan2(eta, (double)l);
}
}  // namespace qpc

using qpc::cplx;

struct qpc_setup {
    qpc_config cfg{};
    int N = 0, p = 0, ne = 0, L = 0, M = 0, m0 = 0, lmax = 0;
    qpc::Basis basis;
    std::vector<double> S, chol, rinv, rdif, rpow1, dvdr, mask, souter, H, yeig, qeig, yprj, y10, q10;
    std::vector<double> hV, hK, hT;  // the l-independent parts of integrate_hc
    std::vector<cplx> upper, lower, qprj, merged;
    int lu_tiny = 0;
};

extern "C" {

int qpc_setup_build(const qpc_config* cfg, qpc_setup** out) {
    if (!cfg || !out) return QPC_E_ARG;
    *out = nullptr;
    const qpc_dims& d = cfg->dims;
    if (d.n_rank < 1 || d.n_rank > 8 || d.n_ndim < 2 * (d.n_rank + 1) || d.n_ldim < 1 || d.n_mdim < 1 ||
        !(cfg->para_rmax > cfg->para_rmin) || !(cfg->para_me > 0) || !(d.para_dt > 0))
        return QPC_E_SHAPE;
    qpc_setup* st = new qpc_setup();
    st->cfg = *cfg;
    st->N = (int)d.n_ndim;
    st->p = (int)d.n_rank;
    st->ne = st->p + 1;
    st->L = (int)d.n_ldim;
    st->M = (int)d.n_mdim;
    st->m0 = (int)d.n_mmin;
    const int m1 = st->m0 + st->M - 1;
    st->lmax = std::max(std::abs(st->m0), std::abs(m1)) + st->L - 1;
    const int N = st->N, p = st->p, ne = st->ne, L = st->L, M = st->M;
    st->basis.build(N, p, cfg->para_rmin, cfg->para_rmax);
    const qpc::Basis& B = st->basis;
    const double zc = cfg->para_zc;
    B.band(st->S, 0, 0, [](double) { return 1.0; });
    B.band(st->rinv, 0, 0, [](double r) { return 1.0 / (r + 1e-50); });
    B.band(st->rdif, 0, 1, [](double) { return 1.0; });
    B.band(st->rpow1, 0, 0, [](double r) { return r; });
    B.band(st->dvdr, 0, 0, [zc](double r) { return zc / (r * r + 1e-50); });
    {
        const double a = cfg->mask_rmin, b = cfg->mask_rmax, e = cfg->mask_fact;
        B.band(st->mask, 0, 0, [=](double r) {
            return r < a ? 1.0 : r > b ? 0.0 : std::pow(std::fabs(std::cos(0.5 * qpc::kPi * (r - a) / (b - a))), e);
        });
    }
    {
        // overlap restricted to the absorbing region r >= mask_rmin (the "outer" norm)
        const double a = cfg->mask_rmin;
        B.band(st->souter, 0, 0, [=](double r) { return r >= a ? 1.0 : 0.0; });
    }
    st->chol = st->S;
    if (qpc::chol_real(N, p, st->chol.data())) {
        delete st;
        return QPC_E_SETUP;
    }
    // H_l = <V> + l(l+1)/(2m) <1/r^2> + 1/(2m) <B'|B'>  (basisfunc.hpp:92-106)
    std::vector<double>&V = st->hV, &K = st->hK, &T = st->hT;
    B.band(V, 0, 0, [zc](double r) { return -zc / (r + 1e-50); });
    B.band(K, 0, 0, [](double r) { return 1.0 / (r * r + 1e-50); });
    B.band(T, 1, 1, [](double) { return 1.0; });
    const double c2m = 0.5 / cfg->para_me;
    const size_t band = (size_t)N * ne;
    st->H.assign((size_t)(st->lmax + 1) * band, 0.0);
    st->upper.assign((size_t)(st->lmax + 1) * band, cplx{0, 0});
    st->lower.assign((size_t)(st->lmax + 1) * band, cplx{0, 0});
    const double hdt = 0.5 * d.para_dt;
    for (int l = 0; l <= st->lmax; ++l) {
        const double cent = (double)l * (l + 1) * c2m;
        double* h = &st->H[(size_t)l * band];
        cplx* up = &st->upper[(size_t)l * band];
        cplx* lo = &st->lower[(size_t)l * band];
        for (size_t i = 0; i < band; ++i) {
            h[i] = qpc::hc_entry(V[i], cent, K[i], c2m, T[i]);
            up[i] = cplx{st->S[i], -hdt * h[i]};
            lo[i] = cplx{st->S[i], hdt * h[i]};
        }
        st->lu_tiny += qpc::ldlt_csym(N, p, lo);
    }
    // angular: Y10 (real symmetric tridiagonal) and the Hermitian i*Q10
    // matrix with lower entries -i q_l (operators/1e.hpp:304-333, 421-448);
    // i*Q10 = D T D^H with D = diag(i^l) and T real with off-diagonal -q_l.
    st->y10.assign((size_t)M * L, 0.0);
    st->q10.assign((size_t)M * L, 0.0);
    st->yeig.assign((size_t)M * L, 0.0);
    st->qeig.assign((size_t)M * L, 0.0);
    st->yprj.assign((size_t)M * L * L, 0.0);
    st->qprj.assign((size_t)M * L * L, cplx{0, 0});
    st->merged.assign((size_t)M * L * L, cplx{0, 0});
    for (int im = 0; im < M; ++im) {
        const long m = im + st->m0;
        std::vector<double> dz(L, 0.0), ey(L, 0.0), eq(L, 0.0), w, z;
        for (int il = 0; il < L; ++il) {
            const long l = il + std::labs(m);
            const double y = std::sqrt(double((l + 1) * (l + 1) - m * m) / double((2 * l + 1) * (2 * l + 3)));
            st->y10[(size_t)im * L + il] = y;
            st->q10[(size_t)im * L + il] = y * (double)(l + 1);
            ey[il] = y;
            eq[il] = -y * (double)(l + 1);
        }
        if (qpc::tridiag_ql(L, dz, ey, w, z)) {
            delete st;
            return QPC_E_SETUP;
        }
        for (int k = 0; k < L; ++k) {
            st->yeig[(size_t)im * L + k] = w[k];
            for (int i = 0; i < L; ++i) st->yprj[((size_t)im * L + k) * L + i] = z[(size_t)k * L + i];
        }
        if (qpc::tridiag_ql(L, dz, eq, w, z)) {
            delete st;
            return QPC_E_SETUP;
        }
        // T has a zero diagonal, so its spectrum is symmetric about 0 (the
        // similarity diag((-1)^i) maps T to -T); the ascending eigenvalues
        // are made exact pairs (w[L-1-k] = -w[k], the middle one 0 for odd L),
        // a change at the rounding level that lets the R-part sweeps of
        // opposite eigenvalues share their conjugate tables
        for (int k = 0; k < L / 2; ++k) {
            const double v = 0.5 * (w[L - 1 - k] - w[k]);
            w[k] = -v;
            w[L - 1 - k] = v;
        }
        if (L % 2) w[L / 2] = 0.0;
        static const cplx kMinusIPow[4] = {{1, 0}, {0, -1}, {-1, 0}, {0, 1}};
        for (int k = 0; k < L; ++k) {
            st->qeig[(size_t)im * L + k] = w[k];
            for (int i = 0; i < L; ++i)
                st->qprj[((size_t)im * L + k) * L + i] = kMinusIPow[i % 4] * z[(size_t)k * L + i];
        }
        for (int k = 0; k < L; ++k)
            for (int j = 0; j < L; ++j) {
                cplx acc{0, 0};
                for (int i = 0; i < L; ++i)
                    acc += st->qprj[((size_t)im * L + k) * L + i] * st->yprj[((size_t)im * L + j) * L + i];
                st->merged[((size_t)im * L + k) * L + j] = acc;
            }
    }
    *out = st;
    return QPC_OK;
}

void qpc_setup_free(qpc_setup* st) { delete st; }

int64_t qpc_setup_size(const qpc_setup* st, int what) {
    if (!st) return -1;
    const int64_t band = (int64_t)st->N * st->ne, nl = st->lmax + 1, ml = (int64_t)st->M * st->L;
    switch (what) {
        case QPC_SETUP_S: case QPC_SETUP_CHOL: case QPC_SETUP_RINV: case QPC_SETUP_RDIF: case QPC_SETUP_RPOW1:
        case QPC_SETUP_DVDR: case QPC_SETUP_MASK: case QPC_SETUP_SOUTER: case QPC_SETUP_HV: case QPC_SETUP_HK:
        case QPC_SETUP_HT:
            return band;
        case QPC_SETUP_H: return nl * band;
        case QPC_SETUP_UPPER: case QPC_SETUP_LOWER: return 2 * nl * band;
        case QPC_SETUP_YEIG: case QPC_SETUP_QEIG: case QPC_SETUP_Y10: case QPC_SETUP_Q10: return ml;
        case QPC_SETUP_YPRJ: return ml * st->L;
        case QPC_SETUP_QPRJ: case QPC_SETUP_MERGED: return 2 * ml * st->L;
        case QPC_SETUP_KNOTS: return (int64_t)st->N + st->p + 1;
        default: return -1;
    }
}

int qpc_setup_get(const qpc_setup* st, int what, double* dst) {
    if (!st || !dst) return QPC_E_ARG;
    const int64_t n = qpc_setup_size(st, what);
    if (n < 0) return QPC_E_ARG;
    const double* src = nullptr;
    switch (what) {
        case QPC_SETUP_KNOTS:
            // bsplines::knot() from the first active knot, size() + 1 of them (basis_bsplines.h:149-177)
            for (int64_t k = 0; k < n; ++k) dst[k] = st->basis.knot((int)k);
            return QPC_OK;
        case QPC_SETUP_S: src = st->S.data(); break;
        case QPC_SETUP_CHOL: src = st->chol.data(); break;
        case QPC_SETUP_RINV: src = st->rinv.data(); break;
        case QPC_SETUP_RDIF: src = st->rdif.data(); break;
        case QPC_SETUP_RPOW1: src = st->rpow1.data(); break;
        case QPC_SETUP_DVDR: src = st->dvdr.data(); break;
        case QPC_SETUP_MASK: src = st->mask.data(); break;
        case QPC_SETUP_SOUTER: src = st->souter.data(); break;
        case QPC_SETUP_HV: src = st->hV.data(); break;
        case QPC_SETUP_HK: src = st->hK.data(); break;
        case QPC_SETUP_HT: src = st->hT.data(); break;
        case QPC_SETUP_H: src = st->H.data(); break;
        case QPC_SETUP_UPPER: src = (const double*)st->upper.data(); break;
        case QPC_SETUP_LOWER: src = (const double*)st->lower.data(); break;
        case QPC_SETUP_YEIG: src = st->yeig.data(); break;
        case QPC_SETUP_QEIG: src = st->qeig.data(); break;
        case QPC_SETUP_Y10: src = st->y10.data(); break;
        case QPC_SETUP_Q10: src = st->q10.data(); break;
        case QPC_SETUP_YPRJ: src = st->yprj.data(); break;
        case QPC_SETUP_QPRJ: src = (const double*)st->qprj.data(); break;
        case QPC_SETUP_MERGED: src = (const double*)st->merged.data(); break;
        default: return QPC_E_ARG;
    }
    std::memcpy(dst, src, sizeof(double) * (size_t)n);
    return QPC_OK;
}

int qpc_setup_eigenstate(const qpc_setup* st, int64_t im, int64_t il, int64_t in, double* energy, double* coef) {
    if (!st || !coef) return QPC_E_ARG;
    if (im < 0 || im >= st->M || il < 0 || il >= st->L || in < 0 || in >= st->N) return QPC_E_ARG;
    const int N = st->N;
    const long l = il + std::labs(im + st->m0);
    std::vector<double> v;
    double E = 0.0;
    if (qpc::pencil_eigen(N, st->p, &st->H[(size_t)l * N * st->ne], st->S.data(), (int)in, E, v))
        return QPC_E_SETUP;
    std::memset(coef, 0, sizeof(double) * 2 * (size_t)st->M * st->L * N);
    double* psi = coef + 2 * ((size_t)im * st->L + il) * N;
    for (int i = 0; i < N; ++i) psi[2 * i] = v[i];
    if (energy) *energy = E;
    return QPC_OK;
}

int qpc_setup_lu_info(const qpc_setup* st) { return st ? st->lu_tiny : QPC_E_ARG; }

int qpc_setup_projection_basis(const qpc_setup* st, int64_t nsel, double* W, double* E) {
    if (!st || !W || nsel < 1 || nsel > st->N) return QPC_E_ARG;
    const int N = st->N, nl = st->lmax + 1;
    std::atomic<int> bad{0}, next{0};
    auto work = [&]() {
        std::vector<double> v;
        for (int l = next++; l < nl; l = next++) {
            const double* h = &st->H[(size_t)l * N * st->ne];
            for (int64_t n = 0; n < nsel; ++n) {
                double e = 0.0;
                if (qpc::pencil_eigen(N, st->p, h, st->S.data(), (int)n, e, v)) {
                    ++bad;
                    continue;
                }
                qpc::band_apply(N, st->p, st->S.data(), v.data(), W + ((size_t)l * nsel + n) * N);
                if (E) E[(size_t)l * nsel + n] = e;
            }
        }
    };
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned k = 0; k + 1 < nt; ++k) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
    return bad ? QPC_E_SETUP : QPC_OK;
}

int qpc_coulomb_f(int64_t lmax, double eta, double rho, double* F) {
    if (!F || lmax < 0 || !(rho > 0.0) || !std::isfinite(eta)) return QPC_E_ARG;
    return qpc::coulomb_f((int)lmax, eta, rho, F) ? QPC_E_NUMERIC : QPC_OK;
}

int qpc_coulomb_sigma(int64_t lmax, double eta, double* sig) {
    if (!sig || lmax < 0 || !std::isfinite(eta)) return QPC_E_ARG;
    qpc::coulomb_sigma((int)lmax, eta, sig);
    return QPC_OK;
}

int qpc_setup_pcs(const qpc_setup* st, int64_t nkr, double krmin, double krmax, double zc, double rb, double* kr,
                  double* ijl, double* cps) {
    if (!st || !kr || !ijl || !cps || nkr < 1 || !(krmin > 0.0) || !(krmax >= krmin)) return QPC_E_ARG;
    const int N = st->N, p = st->p, nl = st->lmax + 1;
    const qpc::Basis& B = st->basis;
    // set_sequence_lin (spectrum/inventory.h:15-26)
    const double dk = (krmax - krmin) / (double)(nkr > 1 ? nkr - 1 : 1);
    for (int64_t i = 0; i < nkr; ++i) kr[i] = (double)i * dk + krmin;
    if (nkr > 1) kr[nkr - 1] = krmax;
    const int ns = N + p, q = B.q;
    std::atomic<int> bad{0};
    std::atomic<int64_t> next{0};
    auto work = [&]() {
        std::vector<double> F(nl), sig(nl), fg((size_t)nl * ns * q);
        for (int64_t ik = next++; ik < nkr; ik = next++) {
            const double k = kr[ik], eta = -zc / k;
            // F_l(eta, k r) on the Gauss points (zero where the weight is zero,
            // the zero-length end intervals), times the plane-wave window when rb > 0
            std::fill(fg.begin(), fg.end(), 0.0);
            for (int s = 0; s < ns; ++s) {
                if (B.gw[(size_t)s * q] == 0.0) continue;
                for (int g = 0; g < q; ++g) {
                    const double r = B.gx[(size_t)s * q + g];
                    if (!(r > 0.0)) continue;
                    if (qpc::coulomb_f(nl - 1, eta, k * r, F.data())) {
                        ++bad;
                        continue;
                    }
                    const double win = rb > 0.0 ? 1.0 / (1.0 + std::exp(-5.0 * (r - rb))) : 1.0;
                    for (int l = 0; l < nl; ++l) fg[((size_t)l * ns + s) * q + g] = F[l] * win;
                }
            }
            // ijl(ikr, l, j) = int B_j(r) F_l(eta, k r) dr / k (momentum_1e.hpp:136-139)
            double* out = ijl + (size_t)ik * nl * N;
            std::fill(out, out + (size_t)nl * N, 0.0);
            for (int l = 0; l < nl; ++l)
                for (int s = 0; s < ns; ++s) {
                    if (B.gw[(size_t)s * q] == 0.0) continue;
                    for (int g = 0; g < q; ++g) {
                        const double wf = B.gw[(size_t)s * q + g] * fg[((size_t)l * ns + s) * q + g] / k;
                        const double* bv = &B.b0[((size_t)s * q + g) * (p + 1)];
                        for (int r = 0; r <= p; ++r) {
                            const int j = s - p + r;
                            if (j >= 0 && j < N) out[(size_t)l * N + j] += wf * bv[r];
                        }
                    }
                }
            // cps(ikr, l) = exp(i (sigma_l(eta) - l pi/2)) (momentum_1e.hpp:140-144)
            qpc::coulomb_sigma(nl - 1, eta, sig.data());
            for (int l = 0; l < nl; ++l) {
                const double ph = sig[l] - 0.5 * qpc::kPi * l;
                cps[2 * ((size_t)ik * nl + l)] = std::cos(ph);
                cps[2 * ((size_t)ik * nl + l) + 1] = std::sin(ph);
            }
        }
    };
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t + 1 < nt; ++t) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
    return bad ? QPC_E_NUMERIC : QPC_OK;
}

// t-SURFF tables (spectrum_c_tsf::initialize, momentum_1e.hpp:779-880).
// Axes: set_sequence_lin, or for flag 1 k = sqrt(2 E) on a linear energy
// axis (set_sequence_sqr, spectrum/inventory.h:15-39).  The non-zero
// B-splines and their derivatives at r0; j0 = F_l(eta, k r0) / k and
// jr = F_l'(eta, k r0) with eta = -zc / k (r j_l and its r-derivative when
// zc = 0); y0 = Y_lm(theta_k, 0) (spherical-harmonic normalisation,
// Condon-Shortley phase, Y_l,-m = (-1)^m Y_lm on the axis); e0 = exp(i m phi_k);
// pmat0 = <l m|cos|l+1 m> (angular integrate_cos).
int qpc_setup_tsurff(const qpc_setup* st, const qpc_tsurff_cfg* tc, double* kr, double* kt, double* kp, int64_t* nb,
                     int64_t* ib, double* b0, double* br, double* j0, double* jr, double* y0, double* e0,
                     double* pmat0) {
    if (!st || !tc || !kr || !kt || !kp || !nb || !ib || !b0 || !br || !j0 || !jr || !y0 || !e0 || !pmat0)
        return QPC_E_ARG;
    if (tc->nkr < 1 || tc->nkt < 1 || tc->nkp < 1 || !(tc->krmin > 0.0) || !(tc->krmax >= tc->krmin) ||
        !(tc->r0 > st->cfg.para_rmin) || !(tc->r0 < st->cfg.para_rmax))
        return QPC_E_ARG;
    const int N = st->N, p = st->p, M = st->M, L = st->L, m0 = st->m0, nl = st->lmax + 1;
    auto lin = [](double* a, double lo, double hi, int64_t n) {
        const double d = (hi - lo) / (double)(n > 1 ? n - 1 : 1);
        for (int64_t i = 0; i < n; ++i) a[i] = (double)i * d + lo;
        if (n > 1) a[n - 1] = hi;
    };
    lin(kp, tc->kpmin, tc->kpmax, tc->nkp);
    lin(kt, tc->ktmin, tc->ktmax, tc->nkt);
    if (tc->flag == 0) {
        lin(kr, tc->krmin, tc->krmax, tc->nkr);
    } else {
        lin(kr, 0.5 * tc->krmin * tc->krmin, 0.5 * tc->krmax * tc->krmax, tc->nkr);
        for (int64_t i = 0; i < tc->nkr; ++i) kr[i] = std::sqrt(2.0 * kr[i]);
    }
    // B-splines at r0 (the knot span holding r0; basis index = spline index)
    const qpc::Basis& B = st->basis;
    int s = 0;
    while (s + 1 < N + p && !(B.knot(s + 1) > tc->r0)) ++s;
    std::vector<double> v(p + 1), dv(p + 1);
    B.eval(s, tc->r0, v.data(), dv.data());
    int64_t n = 0;
    for (int r = 0; r <= p; ++r) {
        const int i = s - p + r;
        if (i < 0 || i >= N || !(v[r] > 0.0)) continue;
        ib[n] = i;
        b0[n] = v[r];
        br[n] = dv[r];
        ++n;
    }
    *nb = n;
    // Coulomb functions at k r0
    std::vector<double> F(nl), Fd(nl);
    for (int64_t ik = 0; ik < tc->nkr; ++ik) {
        const double k = kr[ik];
        if (qpc::coulomb_f(nl - 1, -tc->zc / k, k * tc->r0, F.data(), Fd.data())) return QPC_E_NUMERIC;
        for (int l = 0; l < nl; ++l) {
            j0[ik * nl + l] = F[l] / k;
            jr[ik * nl + l] = Fd[l];
        }
    }
    // normalised associated Legendre functions, m >= 0, by the standard
    // recurrences: P_mm = -sqrt((2m+1)/(2m)) sin(t) P_(m-1)(m-1), P_00 = 1/sqrt(4 pi),
    // P_(m+1)m = sqrt(2m+3) x P_mm, P_lm = a (x P_(l-1)m - b P_(l-2)m),
    // a = sqrt((4l^2-1)/(l^2-m^2)), b = sqrt(((l-1)^2-m^2)/(4(l-1)^2-1))
    const int lmax = nl - 1;
    std::vector<double> P((size_t)nl * nl);
    for (int64_t it = 0; it < tc->nkt; ++it) {
        const double x = std::cos(kt[it]), sn = std::sqrt(std::max(0.0, 1.0 - x * x));
        std::fill(P.begin(), P.end(), 0.0);
        double pmm = 1.0 / std::sqrt(4.0 * qpc::kPi);
        for (int m = 0; m <= lmax; ++m) {
            if (m > 0) pmm *= -std::sqrt((2.0 * m + 1.0) / (2.0 * m)) * sn;
            P[(size_t)m * nl + m] = pmm;
            if (m + 1 <= lmax) P[(size_t)(m + 1) * nl + m] = std::sqrt(2.0 * m + 3.0) * x * pmm;
            for (int l = m + 2; l <= lmax; ++l) {
                const double a = std::sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
                const double b = std::sqrt(((l - 1.0) * (l - 1.0) - (double)m * m) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
                P[(size_t)l * nl + m] = a * (x * P[(size_t)(l - 1) * nl + m] - b * P[(size_t)(l - 2) * nl + m]);
            }
        }
        for (int im = 0; im < M; ++im)
            for (int il = 0; il < L; ++il) {
                const int m = m0 + im, am = m < 0 ? -m : m, l = il + am;
                const double y = l <= lmax ? P[(size_t)l * nl + am] : 0.0;
                y0[(it * M + im) * L + il] = (m < 0 && (am & 1)) ? -y : y;
            }
    }
    for (int64_t ip = 0; ip < tc->nkp; ++ip)
        for (int im = 0; im < M; ++im) {
            const double a = (double)(m0 + im) * kp[ip];
            e0[2 * (ip * M + im)] = std::cos(a);
            e0[2 * (ip * M + im) + 1] = std::sin(a);
        }
    for (int im = 0; im < M; ++im)
        for (int il = 0; il < L; ++il) {
            const double m = m0 + im, l = il + std::fabs(m);
            pmat0[im * L + il] = std::sqrt(((l + 1.0) * (l + 1.0) - m * m) / ((2.0 * l + 1.0) * (2.0 * l + 3.0)));
        }
    return QPC_OK;
}

int qpc_setup_tsurff_xy(const qpc_setup* st, double* pmat1, double* pmat2) {
    if (!st || !pmat1 || !pmat2) return QPC_E_ARG;
    // integrate_sinexp1/2 (basisfunc.hpp:241-259) per channel (m, l)
    for (int im = 0; im < st->M; ++im)
        for (int il = 0; il < st->L; ++il) {
            const long m = st->m0 + im, l = il + std::labs(m);
            double se1, se2, cl1, cl2;
            ep_forms(m, l, se1, se2, cl1, cl2);
            pmat1[im * st->L + il] = se1;
            pmat2[im * st->L + il] = se2;
        }
    return QPC_OK;
}

int qpc_setup_angular_ep(int64_t M, int64_t L, int64_t m0, double* eig_x1, double* eig_x2, double* eig_y1,
                         double* eig_y2, double* tran_x1, double* tran_y2, double* x1_2_x2, double* x2_2_y1,
                         double* y1_2_y2) {
    if (M < 1 || L < 1 || M * L > 65535 || !eig_x1 || !eig_x2 || !eig_y1 || !eig_y2 || !tran_x1 || !tran_y2 ||
        !x1_2_x2 || !x2_2_y1 || !y1_2_y2)
        return QPC_E_ARG;
    return qpc::angular_ep((int)M, (int)L, (int)m0, {eig_x1, eig_x2, eig_y1, eig_y2}, tran_x1, tran_y2,
                           {x1_2_x2, x2_2_y1, y1_2_y2});
}

int qpc_setup_angular_lgwr(int64_t M, int64_t L, int64_t m0, double* eig_x1, double* eig_y1, double* tran_x1,
                           double* x1_2_y1, double* tran_y1) {
    if (M < 1 || L < 1 || M * L > 65535 || !eig_x1 || !eig_y1 || !tran_x1 || !x1_2_y1 || !tran_y1) return QPC_E_ARG;
    // the Q11 operators are not needed: their eigen-data go to scratch
    const size_t ML = (size_t)(M * L);
    std::vector<double> e2(ML), e4(ML), ty2(ML * ML);
    return qpc::angular_ep((int)M, (int)L, (int)m0, {eig_x1, e2.data(), eig_y1, e4.data()}, tran_x1, ty2.data(),
                           {nullptr, nullptr, nullptr}, x1_2_y1, tran_y1);
}

}  // extern "C"
