// qpc_solve.hip -- partitioned banded CN solves, their tables and the
// time-batched factorisation of the coupling sweeps (gfx950).
#include "qpc_solve.h"

namespace qpc {
namespace {

__device__ inline cplx tmul(double a, cplx b) { return {a * b.re, a * b.im}; }
__device__ inline cplx tmul(cplx a, cplx b) { return a * b; }
__device__ inline double tzero(double) { return 0.0; }
__device__ inline cplx tzero(cplx) { return {0.0, 0.0}; }
__device__ inline double tone(double) { return 1.0; }
__device__ inline cplx tone(cplx) { return {1.0, 0.0}; }

constexpr int kMaxExt = kCP + 2 * kPK;  // partitions touched by one chunk
constexpr int kMaxRows = kMaxExt * kPM; // their rows

// ---------------------------------------------------------------------------
// right-hand side v_i = (B z)_i for a row i < N
template <int RHS>
__device__ inline cplx rhs_row(const RhsArgs& ra, int ch, int tabid, const cplx* __restrict__ z, int i, int N) {
    constexpr int ne = kP + 1;
    cplx v = {0.0, 0.0};
    if (RHS == RHS_CBAND) {
        const cplx* U = ra.cband + (int64_t)tabid * N * ne;
#pragma unroll
        for (int j = 1; j <= kP; ++j)
            if (i - j >= 0) v += U[(int64_t)(i - j) * ne + j] * z[i - j];
#pragma unroll
        for (int k = 0; k < ne; ++k)
            if (i + k < N) v += U[(int64_t)i * ne + k] * z[i + k];
    } else if (RHS == RHS_RBAND) {
        const double* B = ra.rband;
#pragma unroll
        for (int j = 1; j <= kP; ++j)
            if (i - j >= 0) v += B[(int64_t)(i - j) * ne + j] * z[i - j];
#pragma unroll
        for (int k = 0; k < ne; ++k)
            if (i + k < N) v += B[(int64_t)i * ne + k] * z[i + k];
    } else {
        const double* S = ra.rband;
        const double* X = ra.xband;
        const double e = ra.eig[ch] * ra.ifdt;
        cplx sz = {0.0, 0.0}, xl = {0.0, 0.0}, xu = {0.0, 0.0};
#pragma unroll
        for (int j = 1; j <= kP; ++j)
            if (i - j >= 0) {
                const int64_t k = (int64_t)(i - j) * ne + j;
                sz += S[k] * z[i - j];
                xl += X[k] * z[i - j];
            }
#pragma unroll
        for (int k = 0; k < ne; ++k)
            if (i + k < N) {
                const int64_t kk = (int64_t)i * ne + k;
                sz += S[kk] * z[i + k];
                xu += X[kk] * z[i + k];
            }
        if (RHS == RHS_SX_SYM) {
            // a = S - ifet X with ifet = i*e:  v = Sz - i e (Xz)
            const cplx xz = xl + xu;
            v = {sz.re + e * xz.im, sz.im - e * xz.re};
        } else {
            // a = S + e X (lower) and S - e X (upper), e real
            v = {sz.re + e * (xl.re - xu.re), sz.im + e * (xl.im - xu.im)};
        }
    }
    return v;
}

// forward recurrence over the kPM rows of one partition from state h
// (h[0] = y_{r0-1}, ...); the partition's table rows are loaded up front
template <class TF, bool STORE>
__device__ inline void fwd_part(const TF* __restrict__ T, const cplx* __restrict__ v, cplx* __restrict__ ys, int r0,
                                cplx (&h)[kP]) {
    TF c[kPM][kP + 1];
#pragma unroll
    for (int s = 0; s < kPM; ++s)
#pragma unroll
        for (int q = 0; q <= kP; ++q) c[s][q] = T[(int64_t)(r0 + s) * kTab + q];
#pragma unroll
    for (int s = 0; s < kPM; ++s) {
        cplx acc = v[s];
#pragma unroll
        for (int j = 0; j < kP; ++j) acc -= tmul(c[s][j], h[j]);
        const cplx y = tmul(c[s][kP], acc);
#pragma unroll
        for (int j = kP - 1; j > 0; --j) h[j] = h[j - 1];
        h[0] = y;
        if (STORE) ys[s] = y;
    }
}

// backward recurrence over the rows of one partition from state h (h[0] = x_{r0+kPM} ...)
template <class TF>
__device__ inline void bwd_part(const TF* __restrict__ T, const cplx* __restrict__ ys, cplx* __restrict__ xs, int r0,
                                cplx (&h)[kP]) {
    TF c[kPM][kP + 1];
#pragma unroll
    for (int s = 0; s < kPM; ++s)
#pragma unroll
        for (int q = 0; q <= kP; ++q) c[s][q] = T[(int64_t)(r0 + s) * kTab + kP + 1 + q];
#pragma unroll
    for (int s = kPM - 1; s >= 0; --s) {
        cplx x = tmul(c[s][kP], ys[s]);
#pragma unroll
        for (int k = 0; k < kP; ++k) x -= tmul(c[s][k], h[k]);
#pragma unroll
        for (int k = kP - 1; k > 0; --k) h[k] = h[k - 1];
        h[0] = x;
        if (xs) xs[s] = x;
    }
}

// Horner evaluation of the truncated carry chain, one lane per component:
// acc <- E[gp] + T[gp] acc for gp = g -/+ k, k = kPK..1.  The transfer rows do
// not depend on acc, so the fully unrolled loop lets their loads run ahead.
template <class TF>
__device__ inline cplx carry(const TF* __restrict__ TR, const cplx (*sE)[kP], int gA, int g, int dir, int P, int r,
                             int lane_base) {
    cplx acc = {0.0, 0.0};
    const int rr = r < kP ? r : 0;
#pragma unroll
    for (int k = kPK; k >= 1; --k) {
        const int gp = dir == 0 ? g - k : g + k;
        const bool ok = gp >= 0 && gp < P;  // exact start: no state beyond the channel ends
        TF tr[kP];
        const TF* Tm = TR + ((int64_t)(ok ? gp : 0) * 2 + dir) * kP * kP + rr * kP;
#pragma unroll
        for (int j = 0; j < kP; ++j) tr[j] = Tm[j];
        cplx full[kP];
#pragma unroll
        for (int j = 0; j < kP; ++j) {
            full[j].re = __shfl(acc.re, lane_base + j, 64);
            full[j].im = __shfl(acc.im, lane_base + j, 64);
        }
        if (ok) {
            cplx nv = sE[gp - gA][rr];
#pragma unroll
            for (int j = 0; j < kP; ++j) nv += tmul(tr[j], full[j]);
            acc = nv;
        }
    }
    return acc;
}

template <class TF, int RHS>
__global__ __launch_bounds__(kSolveThreads) void k_bsolve(RhsArgs ra, const TF* __restrict__ tab,
                                                          const TF* __restrict__ trf, TabMap tm,
                                                          const cplx* __restrict__ zin, cplx* __restrict__ zout,
                                                          Geom geo) {
    __shared__ cplx sv[kMaxRows];
    __shared__ cplx sy[kMaxRows];
    __shared__ cplx sE[kMaxExt][kP];
    __shared__ cplx sC[kMaxExt][kP];

    const int ch = blockIdx.y;
    const int N = geo.N, P = geo.P;
    const int g0 = blockIdx.x * kCP;
    const int g1 = min(P, g0 + kCP);
    const int gA = max(0, g0 - kPK), gB = min(P, g1 + kPK);
    const int R0 = gA * kPM, R1 = gB * kPM;
    const int tid = threadIdx.x;

    int tabid = ch;
    if (tm.mode == 1) tabid = 0;
    else if (tm.mode == 2) {
        const int im = ch / tm.L, il = ch % tm.L;
        const int m = im + tm.m0;
        tabid = il + (m < 0 ? -m : m);
    }
    const TF* T = tab + (int64_t)tabid * geo.Np * kTab;
    const TF* TR = trf + (int64_t)tabid * P * 2 * kP * kP;
    const cplx* z = zin + (int64_t)ch * N;
    cplx* xo = zout + (int64_t)ch * N;

    // phase 1: right-hand side for every touched row (padding rows: 0)
    for (int i = R0 + tid; i < R1; i += kSolveThreads)
        sv[i - R0] = i < N ? rhs_row<RHS>(ra, ch, tabid, z, i, N) : cplx{0.0, 0.0};
    __syncthreads();

    // phase 2: forward pass A (zero incoming state), end states -> sE
    if (tid < gB - gA) {
        const int g = gA + tid;
        cplx h[kP];
#pragma unroll
        for (int j = 0; j < kP; ++j) h[j] = {0.0, 0.0};
        fwd_part<TF, false>(T, sv + (g - gA) * kPM, nullptr, g * kPM, h);
#pragma unroll
        for (int j = 0; j < kP; ++j) sE[tid][j] = h[j];
    }
    __syncthreads();

    // phase 3: forward carry for partitions [g0, gB), 8 lanes per partition
    for (int base = 0; base < gB - g0; base += kSolveThreads / 8) {
        const int pp = base + (tid >> 3), r = tid & 7;
        const int g = g0 + pp;
        if (g < gB) {
            const cplx c = carry<TF>(TR, sE, gA, g, 0, P, r, (tid & 63) & ~7);
            if (r < kP) sC[g - gA][r] = c;
        }
    }
    __syncthreads();

    // phase 4: forward pass B for partitions [g0, gB): y -> sy
    if (tid < gB - g0) {
        const int g = g0 + tid;
        cplx h[kP];
#pragma unroll
        for (int j = 0; j < kP; ++j) h[j] = g == 0 ? cplx{0.0, 0.0} : sC[g - gA][j];
        fwd_part<TF, true>(T, sv + (g - gA) * kPM, sy + (g - gA) * kPM, g * kPM, h);
    }
    __syncthreads();

    // phase 5: backward pass A (zero incoming state) for [g0, gB), end states -> sE
    if (tid < gB - g0) {
        const int g = g0 + tid;
        cplx h[kP];
#pragma unroll
        for (int j = 0; j < kP; ++j) h[j] = {0.0, 0.0};
        bwd_part<TF>(T, sy + (g - gA) * kPM, nullptr, g * kPM, h);
#pragma unroll
        for (int j = 0; j < kP; ++j) sE[g - gA][j] = h[j];
    }
    __syncthreads();

    // phase 6: backward carry for partitions [g0, g1)
    for (int base = 0; base < g1 - g0; base += kSolveThreads / 8) {
        const int pp = base + (tid >> 3), r = tid & 7;
        const int g = g0 + pp;
        if (g < g1) {
            const cplx c = carry<TF>(TR, sE, gA, g, 1, P, r, (tid & 63) & ~7);
            if (r < kP) sC[g - gA][r] = c;
        }
    }
    __syncthreads();

    // phase 7: backward pass B for [g0, g1): x -> sv (free now), then to global
    if (tid < g1 - g0) {
        const int g = g0 + tid;
        cplx h[kP];
#pragma unroll
        for (int j = 0; j < kP; ++j) h[j] = g == P - 1 ? cplx{0.0, 0.0} : sC[g - gA][j];
        bwd_part<TF>(T, sy + (g - gA) * kPM, sv + (g - gA) * kPM, g * kPM, h);
    }
    __syncthreads();
    for (int i = g0 * kPM + tid; i < min(N, g1 * kPM); i += kSolveThreads) xo[i] = sv[i - R0];
}

// ---------------------------------------------------------------------------
// transfer matrices: one lane per (table, partition, direction), all seven
// unit responses propagated together so each table row is read once
template <class TF>
__global__ __launch_bounds__(256) void k_transfer(const TF* __restrict__ tab, TF* __restrict__ trf, int ntab, Geom geo) {
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int P = geo.P;
    const int64_t total = (int64_t)ntab * P * 2;
    if (id >= total) return;
    const int dir = (int)(id % 2);
    const int g = (int)((id / 2) % P);
    const int t = (int)(id / (2 * (int64_t)P));
    const TF* T = tab + (int64_t)t * geo.Np * kTab;
    const int r0 = g * kPM;
    TF H[kP][kP];  // H[q][col]: state entry q of the response to unit vector col
#pragma unroll
    for (int q = 0; q < kP; ++q)
#pragma unroll
        for (int col = 0; col < kP; ++col) H[q][col] = q == col ? tone(TF{}) : tzero(TF{});
    if (dir == 0) {
#pragma unroll
        for (int s = 0; s < kPM; ++s) {
            const TF* c = T + (int64_t)(r0 + s) * kTab;
            TF f[kP + 1];
#pragma unroll
            for (int q = 0; q <= kP; ++q) f[q] = c[q];
            TF y[kP];
#pragma unroll
            for (int col = 0; col < kP; ++col) {
                TF acc = tzero(TF{});
#pragma unroll
                for (int q = 0; q < kP; ++q) acc = acc - f[q] * H[q][col];
                y[col] = f[kP] * acc;
            }
#pragma unroll
            for (int q = kP - 1; q > 0; --q)
#pragma unroll
                for (int col = 0; col < kP; ++col) H[q][col] = H[q - 1][col];
#pragma unroll
            for (int col = 0; col < kP; ++col) H[0][col] = y[col];
        }
    } else {
#pragma unroll
        for (int s = kPM - 1; s >= 0; --s) {
            const TF* c = T + (int64_t)(r0 + s) * kTab + kP + 1;
            TF f[kP];
#pragma unroll
            for (int q = 0; q < kP; ++q) f[q] = c[q];
            TF x[kP];
#pragma unroll
            for (int col = 0; col < kP; ++col) {
                TF acc = tzero(TF{});
#pragma unroll
                for (int q = 0; q < kP; ++q) acc = acc - f[q] * H[q][col];
                x[col] = acc;
            }
#pragma unroll
            for (int q = kP - 1; q > 0; --q)
#pragma unroll
                for (int col = 0; col < kP; ++col) H[q][col] = H[q - 1][col];
#pragma unroll
            for (int col = 0; col < kP; ++col) H[0][col] = x[col];
        }
    }
    TF* out = trf + (((int64_t)t * P + g) * 2 + dir) * kP * kP;
#pragma unroll
    for (int q = 0; q < kP; ++q)
#pragma unroll
        for (int col = 0; col < kP; ++col) out[q * kP + col] = H[q][col];
}

// ---------------------------------------------------------------------------
// tables from the lu_zsb factor (ed.h:208-289): f[i*ne] = D_i, f[i*ne+j] = L(i+j,i)
__global__ __launch_bounds__(256) void k_tab_ldlt(const cplx* __restrict__ lower, cplx* __restrict__ tab, int nl,
                                                  Geom geo) {
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int N = geo.N, Np = geo.Np;
    if (id >= (int64_t)nl * Np) return;
    const int l = (int)(id / Np), i = (int)(id % Np);
    constexpr int ne = kP + 1;
    const cplx* f = lower + (int64_t)l * N * ne;
    cplx* o = tab + ((int64_t)l * Np + i) * kTab;
    if (i >= N) {  // identity padding row
#pragma unroll
        for (int q = 0; q < kTab; ++q) o[q] = {0.0, 0.0};
        o[kP] = {1.0, 0.0};
        o[kTab - 1] = {1.0, 0.0};
        return;
    }
#pragma unroll
    for (int j = 1; j <= kP; ++j) o[j - 1] = i - j >= 0 ? f[(int64_t)(i - j) * ne + j] : cplx{0.0, 0.0};
    o[kP] = {1.0, 0.0};
#pragma unroll
    for (int k = 1; k <= kP; ++k) o[kP + k] = i + k < N ? f[(int64_t)i * ne + k] : cplx{0.0, 0.0};
    o[kTab - 1] = crecip(f[(int64_t)i * ne]);
}

// tables from chol(S) (ed.h:182-194): c[i*ne+j] = L(i+j,i)
__global__ __launch_bounds__(256) void k_tab_chol(const double* __restrict__ c, double* __restrict__ tab, Geom geo) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int N = geo.N;
    if (i >= geo.Np) return;
    constexpr int ne = kP + 1;
    double* o = tab + (int64_t)i * kTab;
    if (i >= N) {
#pragma unroll
        for (int q = 0; q < kTab; ++q) o[q] = 0.0;
        o[kP] = 1.0;
        o[kTab - 1] = 1.0;
        return;
    }
    const double d = 1.0 / c[(int64_t)i * ne];
#pragma unroll
    for (int j = 1; j <= kP; ++j) o[j - 1] = i - j >= 0 ? c[(int64_t)(i - j) * ne + j] : 0.0;
    o[kP] = d;
#pragma unroll
    for (int k = 1; k <= kP; ++k) o[kP + k] = i + k < N ? c[(int64_t)i * ne + k] * d : 0.0;
    o[kTab - 1] = d;
}

// ---------------------------------------------------------------------------
// reference-order LU of b = S + ifet X (propagate.hpp:33-86), one lane per
// (time point, channel), tables written row by row.  The last seven U rows
// live in a register ring (row i in slot i mod 8).
template <int SYM>
struct FactorTypes;
template <>
struct FactorTypes<0> {
    using T = cplx;
};
template <>
struct FactorTypes<1> {
    using T = double;
};

__device__ inline cplx fdiv(cplx a, cplx b) { return a / b; }
__device__ inline double fdiv(double a, double b) { return a / b; }
__device__ inline cplx frcp(cplx b) { return crecip(b); }
__device__ inline double frcp(double b) { return 1.0 / b; }

template <int SYM>
__global__ __launch_bounds__(64) void k_factor(const double* __restrict__ S, const double* __restrict__ X,
                                               const double* __restrict__ eig, const double* __restrict__ ifdt,
                                               int nt, int nch, Geom geo, typename FactorTypes<SYM>::T* __restrict__ tab) {
    using T = typename FactorTypes<SYM>::T;
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nt * nch) return;
    const int N = geo.N;
    const int t = id / nch, ch = id % nch;
    constexpr int nd = kP, ne = kP + 1;
    const double e = eig[ch] * ifdt[t];
    T* out = tab + (int64_t)id * geo.Np * kTab;
    T u[8][nd];
#pragma unroll
    for (int s = 0; s < 8; ++s)
#pragma unroll
        for (int k = 0; k < nd; ++k) u[s][k] = tzero(T{});
    for (int i0 = 0; i0 < N; i0 += 8) {
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int i = i0 + s;
            if (i < N) {
                T b[2 * nd + 1];
                // lower panel: column i-j (propagate.hpp:35-53)
#pragma unroll
                for (int j = 1; j < ne; ++j) {
                    const int k = i * ne - j * nd;
                    const int jj = nd - j;
                    if (k < 0) {
                        b[jj] = tzero(T{});
                    } else {
                        const double sv = S[k], rv = X[k];
                        if constexpr (SYM == 0) b[jj] = cplx{sv, e * rv};  // s + ifet r, ifet = i e
                        else b[jj] = sv - e * rv;                            // s - ifet r (antisymmetric)
                    }
                }
                // upper (propagate.hpp:55-62)
#pragma unroll
                for (int j = 0; j < ne; ++j) {
                    const int k = i * ne + j;
                    const double sv = S[k], rv = X[k];
                    if constexpr (SYM == 0) b[nd + j] = cplx{sv, e * rv};
                    else b[nd + j] = sv + e * rv;
                }
                const int l = nd > i ? nd - i : 0;
                // elimination with the previous U rows (propagate.hpp:67-71)
#pragma unroll
                for (int j = 0; j < nd; ++j) {
                    if (j >= l) {
                        const int slot = (s + j + 1) & 7;  // row i+j-7
#pragma unroll
                        for (int k = 0; k < nd; ++k) b[j + k + 1] = b[j + k + 1] - b[j] * u[slot][k];
                    }
                }
                // u(i,:) (propagate.hpp:72-75)
#pragma unroll
                for (int j = 0; j < nd; ++j) u[s][j] = fdiv(b[j + nd + 1], b[nd]);
                T* o = out + (int64_t)i * kTab;
#pragma unroll
                for (int j = 1; j <= nd; ++j) o[j - 1] = b[nd - j];  // multiplier of column i-j
                o[nd] = frcp(b[nd]);
#pragma unroll
                for (int k = 1; k <= nd; ++k) o[nd + k] = i + k < N ? u[s][k - 1] : tzero(T{});
                o[kTab - 1] = tone(T{});
            }
        }
    }
    // identity padding rows
    for (int i = N; i < geo.Np; ++i) {
        T* o = out + (int64_t)i * kTab;
#pragma unroll
        for (int q = 0; q < kTab; ++q) o[q] = tzero(T{});
        o[nd] = tone(T{});
        o[kTab - 1] = tone(T{});
    }
}

}  // namespace

// ---------------------------------------------------------------------------
template <class TF>
static void bsolve(hipStream_t s, int rhs, const RhsArgs& ra, const TF* tab, const TF* trf, TabMap tm, const cplx* zin,
                   cplx* zout, int nch, Geom geo) {
    dim3 grid(geo.nchunk, nch), block(kSolveThreads);
    switch (rhs) {
        case RHS_CBAND: hipLaunchKernelGGL((k_bsolve<TF, RHS_CBAND>), grid, block, 0, s, ra, tab, trf, tm, zin, zout, geo); break;
        case RHS_RBAND: hipLaunchKernelGGL((k_bsolve<TF, RHS_RBAND>), grid, block, 0, s, ra, tab, trf, tm, zin, zout, geo); break;
        case RHS_SX_SYM: hipLaunchKernelGGL((k_bsolve<TF, RHS_SX_SYM>), grid, block, 0, s, ra, tab, trf, tm, zin, zout, geo); break;
        case RHS_SX_ASY: hipLaunchKernelGGL((k_bsolve<TF, RHS_SX_ASY>), grid, block, 0, s, ra, tab, trf, tm, zin, zout, geo); break;
        default: break;
    }
}

void launch_bsolve_c(hipStream_t s, int rhs, const RhsArgs& ra, const cplx* tab, const cplx* trf, TabMap tm,
                     const cplx* zin, cplx* zout, int nch, Geom geo) {
    bsolve<cplx>(s, rhs, ra, tab, trf, tm, zin, zout, nch, geo);
}
void launch_bsolve_r(hipStream_t s, int rhs, const RhsArgs& ra, const double* tab, const double* trf, TabMap tm,
                     const cplx* zin, cplx* zout, int nch, Geom geo) {
    bsolve<double>(s, rhs, ra, tab, trf, tm, zin, zout, nch, geo);
}

void launch_transfer_c(hipStream_t s, const cplx* tab, cplx* trf, int ntab, Geom geo) {
    const int64_t total = (int64_t)ntab * geo.P * 2;
    hipLaunchKernelGGL(k_transfer<cplx>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, tab, trf, ntab, geo);
}
void launch_transfer_r(hipStream_t s, const double* tab, double* trf, int ntab, Geom geo) {
    const int64_t total = (int64_t)ntab * geo.P * 2;
    hipLaunchKernelGGL(k_transfer<double>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, tab, trf, ntab, geo);
}

void launch_tab_ldlt(hipStream_t s, const cplx* lower, cplx* tab, int nl, Geom geo) {
    const int64_t total = (int64_t)nl * geo.Np;
    hipLaunchKernelGGL(k_tab_ldlt, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, lower, tab, nl, geo);
}
void launch_tab_chol(hipStream_t s, const double* chol, double* tab, Geom geo) {
    hipLaunchKernelGGL(k_tab_chol, dim3((unsigned)((geo.Np + 255) / 256)), dim3(256), 0, s, chol, tab, geo);
}

void launch_factor_sym(hipStream_t s, const double* S, const double* X, const double* eig, const double* ifdt, int nt,
                       int nch, Geom geo, cplx* tab) {
    const int total = nt * nch;
    hipLaunchKernelGGL(k_factor<0>, dim3((unsigned)((total + 63) / 64)), dim3(64), 0, s, S, X, eig, ifdt, nt, nch, geo,
                       tab);
}
void launch_factor_asy(hipStream_t s, const double* S, const double* X, const double* eig, const double* ifdt, int nt,
                       int nch, Geom geo, double* tab) {
    const int total = nt * nch;
    hipLaunchKernelGGL(k_factor<1>, dim3((unsigned)((total + 63) / 64)), dim3(64), 0, s, S, X, eig, ifdt, nt, nch, geo,
                       tab);
}

}  // namespace qpc
