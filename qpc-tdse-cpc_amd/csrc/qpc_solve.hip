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

constexpr int kMaxExt = kCP + 2 * kPK;            // partitions touched by one chunk
constexpr int kMaxRows = kMaxExt * kPM + kPM;     // their rows (last partition may be longer)

// ---------------------------------------------------------------------------
// right-hand side v_i = (B z)_i
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

template <class TF>
__device__ inline void load_row(const TF* __restrict__ T, int i, TF (&c)[kTab]) {
    const TF* r = T + (int64_t)i * kTab;
#pragma unroll
    for (int q = 0; q < kTab; ++q) c[q] = r[q];
}

// forward recurrence over rows [r0,r1) from state h (h[0] = y_{r0-1} ...);
// y written to ys (rows relative to R0) when STORE
template <class TF, bool STORE>
__device__ inline void fwd_rows(const TF* __restrict__ T, const cplx* __restrict__ v, cplx* __restrict__ ys, int R0,
                                int r0, int r1, cplx (&h)[kP]) {
    for (int i = r0; i < r1; ++i) {
        TF c[kTab];
        load_row(T, i, c);
        cplx acc = v[i - R0];
#pragma unroll
        for (int j = 0; j < kP; ++j) acc -= tmul(c[j], h[j]);
        const cplx y = tmul(c[kP], acc);
#pragma unroll
        for (int j = kP - 1; j > 0; --j) h[j] = h[j - 1];
        h[0] = y;
        if (STORE) ys[i - R0] = y;
    }
}

// backward recurrence over rows (r1 > i >= r0) from state h (h[0] = x_{r1} ...)
template <class TF, bool STORE>
__device__ inline void bwd_rows(const TF* __restrict__ T, const cplx* __restrict__ ys, cplx* __restrict__ xo, int R0,
                                int r0, int r1, cplx (&h)[kP]) {
    for (int i = r1 - 1; i >= r0; --i) {
        TF c[kTab];
        load_row(T, i, c);
        cplx x = tmul(c[kTab - 1], ys[i - R0]);
#pragma unroll
        for (int k = 0; k < kP; ++k) x -= tmul(c[kP + 1 + k], h[k]);
#pragma unroll
        for (int k = kP - 1; k > 0; --k) h[k] = h[k - 1];
        h[0] = x;
        if (STORE) xo[i] = x;
    }
}

// Horner evaluation of the truncated carry chain, one lane per component:
// acc <- E[gp] + T[gp] acc for gp = g-dir*k, k = kPK..1
template <class TF>
__device__ inline cplx carry(const TF* __restrict__ TR, const cplx (*sE)[kP], int gA, int g, int dir, int P, int r,
                             int lane_base) {
    cplx acc = {0.0, 0.0};
    for (int k = kPK; k >= 1; --k) {
        const int gp = dir == 0 ? g - k : g + k;
        if (gp < 0 || gp >= P) continue;  // exact start: no state beyond the channel ends
        const TF* Tm = TR + ((int64_t)gp * 2 + dir) * kP * kP + r * kP;
        cplx nv = sE[gp - gA][r < kP ? r : 0];
        cplx full[kP];
#pragma unroll
        for (int j = 0; j < kP; ++j) {
            full[j].re = __shfl(acc.re, lane_base + j, 64);
            full[j].im = __shfl(acc.im, lane_base + j, 64);
        }
        if (r < kP) {
#pragma unroll
            for (int j = 0; j < kP; ++j) nv += tmul(Tm[j], full[j]);
        }
        acc = nv;
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
    const int R0 = part_r0(gA), R1 = part_r1(gB - 1, P, N);
    const int tid = threadIdx.x;

    int tabid = ch;
    if (tm.mode == 1) tabid = 0;
    else if (tm.mode == 2) {
        const int im = ch / tm.L, il = ch % tm.L;
        const int m = im + tm.m0;
        tabid = il + (m < 0 ? -m : m);
    }
    const TF* T = tab + (int64_t)tabid * N * kTab;
    const TF* TR = trf + (int64_t)tabid * P * 2 * kP * kP;
    const cplx* z = zin + (int64_t)ch * N;
    cplx* xo = zout + (int64_t)ch * N;

    // phase 1: right-hand side for every touched row
    for (int i = R0 + tid; i < R1; i += blockDim.x) sv[i - R0] = rhs_row<RHS>(ra, ch, tabid, z, i, N);
    __syncthreads();

    // phase 2: forward pass A (zero incoming state), end states -> sE
    if (tid < gB - gA) {
        const int g = gA + tid;
        cplx h[kP];
#pragma unroll
        for (int j = 0; j < kP; ++j) h[j] = {0.0, 0.0};
        fwd_rows<TF, false>(T, sv, sy, R0, part_r0(g), part_r1(g, P, N), h);
#pragma unroll
        for (int j = 0; j < kP; ++j) sE[tid][j] = h[j];
    }
    __syncthreads();

    // phase 3: forward carry for partitions [g0, gB)
    {
        const int pp = tid >> 3, r = tid & 7;
        const int g = g0 + pp;
        const bool act = g < gB;
        const cplx c = act ? carry<TF>(TR, sE, gA, g, 0, P, r, (tid & 63) & ~7) : cplx{0.0, 0.0};
        if (act && r < kP) sC[g - gA][r] = c;
    }
    __syncthreads();

    // phase 4: forward pass B for partitions [g0, gB): y -> sy
    if (tid < gB - g0) {
        const int g = g0 + tid;
        cplx h[kP];
#pragma unroll
        for (int j = 0; j < kP; ++j) h[j] = g == 0 ? cplx{0.0, 0.0} : sC[g - gA][j];
        fwd_rows<TF, true>(T, sv, sy, R0, part_r0(g), part_r1(g, P, N), h);
    }
    __syncthreads();

    // phase 5: backward pass A (zero incoming state) for [g0, gB), end states -> sE
    if (tid < gB - g0) {
        const int g = g0 + tid;
        cplx h[kP];
#pragma unroll
        for (int j = 0; j < kP; ++j) h[j] = {0.0, 0.0};
        bwd_rows<TF, false>(T, sy, xo, R0, part_r0(g), part_r1(g, P, N), h);
#pragma unroll
        for (int j = 0; j < kP; ++j) sE[g - gA][j] = h[j];
    }
    __syncthreads();

    // phase 6: backward carry for partitions [g0, g1)
    {
        const int pp = tid >> 3, r = tid & 7;
        const int g = g0 + pp;
        const bool act = g < g1;
        const cplx c = act ? carry<TF>(TR, sE, gA, g, 1, P, r, (tid & 63) & ~7) : cplx{0.0, 0.0};
        if (act && r < kP) sC[g - gA][r] = c;
    }
    __syncthreads();

    // phase 7: backward pass B for [g0, g1): x -> global
    if (tid < g1 - g0) {
        const int g = g0 + tid;
        cplx h[kP];
#pragma unroll
        for (int j = 0; j < kP; ++j) h[j] = g == P - 1 ? cplx{0.0, 0.0} : sC[g - gA][j];
        bwd_rows<TF, true>(T, sy, xo, R0, part_r0(g), part_r1(g, P, N), h);
    }
}

// ---------------------------------------------------------------------------
// transfer matrices: one lane per (table, partition, direction, unit vector)
template <class TF>
__global__ __launch_bounds__(256) void k_transfer(const TF* __restrict__ tab, TF* __restrict__ trf, int ntab, Geom geo) {
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int N = geo.N, P = geo.P;
    const int64_t total = (int64_t)ntab * P * 2 * kP;
    if (id >= total) return;
    const int j = (int)(id % kP);
    const int dir = (int)((id / kP) % 2);
    const int g = (int)((id / (2 * kP)) % P);
    const int t = (int)(id / (2 * kP * (int64_t)P));
    const TF* T = tab + (int64_t)t * N * kTab;
    const int r0 = part_r0(g), r1 = part_r1(g, P, N);
    TF h[kP];
#pragma unroll
    for (int q = 0; q < kP; ++q) h[q] = tzero(TF{});
    h[j] = tone(TF{});
    if (dir == 0) {
        for (int i = r0; i < r1; ++i) {
            const TF* c = T + (int64_t)i * kTab;
            TF acc = tzero(TF{});
#pragma unroll
            for (int q = 0; q < kP; ++q) acc = acc - c[q] * h[q];
            const TF y = c[kP] * acc;
#pragma unroll
            for (int q = kP - 1; q > 0; --q) h[q] = h[q - 1];
            h[0] = y;
        }
    } else {
        for (int i = r1 - 1; i >= r0; --i) {
            const TF* c = T + (int64_t)i * kTab;
            TF x = tzero(TF{});
#pragma unroll
            for (int q = 0; q < kP; ++q) x = x - c[kP + 1 + q] * h[q];
#pragma unroll
            for (int q = kP - 1; q > 0; --q) h[q] = h[q - 1];
            h[0] = x;
        }
    }
    TF* out = trf + (((int64_t)t * P + g) * 2 + dir) * kP * kP;
#pragma unroll
    for (int q = 0; q < kP; ++q) out[q * kP + j] = h[q];
}

// ---------------------------------------------------------------------------
// tables from the lu_zsb factor (ed.h:208-289): f[i*ne] = D_i, f[i*ne+j] = L(i+j,i)
__global__ __launch_bounds__(256) void k_tab_ldlt(const cplx* __restrict__ lower, cplx* __restrict__ tab, int nl, int N) {
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (int64_t)nl * N) return;
    const int l = (int)(id / N), i = (int)(id % N);
    constexpr int ne = kP + 1;
    const cplx* f = lower + (int64_t)l * N * ne;
    cplx* o = tab + ((int64_t)l * N + i) * kTab;
#pragma unroll
    for (int j = 1; j <= kP; ++j) o[j - 1] = i - j >= 0 ? f[(int64_t)(i - j) * ne + j] : cplx{0.0, 0.0};
    o[kP] = {1.0, 0.0};
#pragma unroll
    for (int k = 1; k <= kP; ++k) o[kP + k] = i + k < N ? f[(int64_t)i * ne + k] : cplx{0.0, 0.0};
    o[kTab - 1] = crecip(f[(int64_t)i * ne]);
}

// tables from chol(S) (ed.h:182-194): c[i*ne+j] = L(i+j,i)
__global__ __launch_bounds__(256) void k_tab_chol(const double* __restrict__ c, double* __restrict__ tab, int N) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    constexpr int ne = kP + 1;
    double* o = tab + (int64_t)i * kTab;
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
                                               int nt, int nch, int N, typename FactorTypes<SYM>::T* __restrict__ tab) {
    using T = typename FactorTypes<SYM>::T;
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nt * nch) return;
    const int t = id / nch, ch = id % nch;
    constexpr int nd = kP, ne = kP + 1;
    const double e = eig[ch] * ifdt[t];
    T* out = tab + (int64_t)id * N * kTab;
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
    const int64_t total = (int64_t)ntab * geo.P * 2 * kP;
    hipLaunchKernelGGL(k_transfer<cplx>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, tab, trf, ntab, geo);
}
void launch_transfer_r(hipStream_t s, const double* tab, double* trf, int ntab, Geom geo) {
    const int64_t total = (int64_t)ntab * geo.P * 2 * kP;
    hipLaunchKernelGGL(k_transfer<double>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, tab, trf, ntab, geo);
}

void launch_tab_ldlt(hipStream_t s, const cplx* lower, cplx* tab, int nl, int N) {
    const int64_t total = (int64_t)nl * N;
    hipLaunchKernelGGL(k_tab_ldlt, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, lower, tab, nl, N);
}
void launch_tab_chol(hipStream_t s, const double* chol, double* tab, int N) {
    hipLaunchKernelGGL(k_tab_chol, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, chol, tab, N);
}

void launch_factor_sym(hipStream_t s, const double* S, const double* X, const double* eig, const double* ifdt, int nt,
                       int nch, int N, cplx* tab) {
    const int total = nt * nch;
    hipLaunchKernelGGL(k_factor<0>, dim3((unsigned)((total + 63) / 64)), dim3(64), 0, s, S, X, eig, ifdt, nt, nch, N, tab);
}
void launch_factor_asy(hipStream_t s, const double* S, const double* X, const double* eig, const double* ifdt, int nt,
                       int nch, int N, double* tab) {
    const int total = nt * nch;
    hipLaunchKernelGGL(k_factor<1>, dim3((unsigned)((total + 63) / 64)), dim3(64), 0, s, S, X, eig, ifdt, nt, nch, N,
                       tab);
}

}  // namespace qpc
