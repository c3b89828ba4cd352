// qpc_solve.hip -- partitioned banded CN solves, their tables and the
// time-batched factorisation of the coupling sweeps (gfx950).
#include "qpc_solve.h"

namespace qpc {
namespace {

__device__ inline cplx tmul(double a, cplx b) { return {a * b.re, a * b.im}; }
// a + t x with fused multiply-adds (4 for complex t, 2 for real t)
__device__ inline cplx tmac(cplx a, cplx t, cplx x) {
    return {fma(-t.im, x.im, fma(t.re, x.re, a.re)), fma(t.im, x.re, fma(t.re, x.im, a.im))};
}
__device__ inline cplx tmac(cplx a, double t, cplx x) { return {fma(t, x.re, a.re), fma(t, x.im, a.im)}; }
// a - t x with fused multiply-adds
__device__ inline cplx tmsub(cplx a, cplx t, cplx x) {
    return {fma(t.im, x.im, fma(-t.re, x.re, a.re)), fma(-t.im, x.re, fma(-t.re, x.im, a.im))};
}
__device__ inline double tzero(double) { return 0.0; }
__device__ inline cplx tzero(cplx) { return {0.0, 0.0}; }
__device__ inline double tone(double) { return 1.0; }
__device__ inline cplx tone(cplx) { return {1.0, 0.0}; }

// row length and table stride of the two table formats (qpc_solve.h)
template <class TF>
struct TabFmt;
template <>
struct TabFmt<double> {
    static constexpr int row = kTab;
    __host__ __device__ static int64_t elems(const Geom& g) { return g.tab_elems; }
};
template <>
struct TabFmt<cplx> {
    static constexpr int row = kTabC;
    __host__ __device__ static int64_t elems(const Geom& g) { return g.ctab_elems; }
};
// first element of partition g's column in a table
template <class TF>
__device__ inline int64_t part_col(int g) {
    return (int64_t)(g >> 6) * kPM * TabFmt<TF>::row * kLane + (g & 63);
}

__device__ inline int table_of(const TabMap& tm, int ch) {
    if (tm.mode == 1) return 0;
    if (tm.mode == 3) return ch % tm.L;
    if (tm.mode == 2) {
        const int im = ch / tm.L, il = ch % tm.L;
        const int m = im + tm.m0;
        return il + (m < 0 ? -m : m);
    }
    return ch;
}

// ---------------------------------------------------------------------------
// Right-hand sides.  Their loads are unconditional, from indices clamped
// into [0, N), and band entries outside the matrix are zero (stored zero, or
// zeroed after loading), so every load of a row is in flight at once: a load
// under a branch compiles to a load-and-wait pair.  The sums keep their
// order; the extra terms add exact zeros.

// Row i < N of a per-table complex band (diagonal-major, B(i, i+k) =
// U[k*N + i], zero where i + k >= N): its entries, then (B z)_i from
// zl[j-1] = z_{i-j} (j = 1..kP) and zu[k] = z_{i+k} (k = 0..kP)
struct CBandRow {
    cplx ul[kP], uu[kP + 1];
    __device__ inline void load(const cplx* __restrict__ U, int i, int N) {
#pragma unroll
        for (int j = 1; j <= kP; ++j) ul[j - 1] = U[(int64_t)j * N + max(i - j, 0)];
#pragma unroll
        for (int k = 0; k <= kP; ++k) uu[k] = U[(int64_t)k * N + i];
    }
    // explicit fused multiply-adds, so that every kernel forming the row
    // rounds it alike (a complex product leaves the contraction open)
    __device__ inline cplx apply(const cplx (&zl)[kP], const cplx (&zu)[kP + 1], int i) const {
        cplx v = {0.0, 0.0};
#pragma unroll
        for (int j = 1; j <= kP; ++j) {
            const cplx u = i - j >= 0 ? ul[j - 1] : cplx{0.0, 0.0};
            v = tmac(v, u, zl[j - 1]);
        }
#pragma unroll
        for (int k = 0; k <= kP; ++k) v = tmac(v, uu[k], zu[k]);
        return v;
    }
};

// Row i < N of the bands shared by every channel: S (or the mask) and, for
// the RHS_SX_* kinds, X.  Lower entries B(i, i-j) = D[j][i-j] (zero above
// the first row), upper B(i, i+k) = D[k][i] (stored zero where i + k >= N).
template <int RHS>
struct SBandRow {
    static constexpr bool SX = RHS != RHS_RBAND;
    double sl[kP], su[kP + 1], xl[kP], xu[kP + 1];
    __device__ inline void load(const RhsArgs& ra, int i, int N) {
#pragma unroll
        for (int j = 1; j <= kP; ++j) {
            const int64_t k = (int64_t)j * N + max(i - j, 0);
            sl[j - 1] = ra.rband[k];
            xl[j - 1] = SX ? ra.xband[k] : 0.0;
        }
#pragma unroll
        for (int k = 0; k <= kP; ++k) {
            su[k] = ra.rband[(int64_t)k * N + i];
            xu[k] = SX ? ra.xband[(int64_t)k * N + i] : 0.0;
        }
#pragma unroll
        for (int j = 1; j <= kP; ++j)
            if (i - j < 0) sl[j - 1] = xl[j - 1] = 0.0;
    }
    // (B z)_i from zlo[j-1] = z_{i-j}, zup[k] = z_{i+k}; e = eig * ifdt of the channel
    __device__ inline cplx apply(const cplx (&zlo)[kP], const cplx (&zup)[kP + 1], double e) const {
        cplx sz = {0.0, 0.0}, zl = {0.0, 0.0}, zu = {0.0, 0.0};
#pragma unroll
        for (int j = 1; j <= kP; ++j) {
            sz += sl[j - 1] * zlo[j - 1];
            if (SX) zl += xl[j - 1] * zlo[j - 1];
        }
#pragma unroll
        for (int k = 0; k <= kP; ++k) {
            sz += su[k] * zup[k];
            if (SX) zu += xu[k] * zup[k];
        }
        if (RHS == RHS_SX_SYM) {
            // a = S - ifet X with ifet = i*e:  v = Sz - i e (Xz)
            const cplx xz = zl + zu;
            return {sz.re + e * xz.im, sz.im - e * xz.re};
        } else if (RHS == RHS_SX_ASY) {
            // a = S + e X (lower) and S - e X (upper), e real
            return {sz.re + e * (zl.re - zu.re), sz.im + e * (zl.im - zu.im)};
        }
        return sz;
    }
};

// LDS slot of row s of partition g in the row-staging buffer: the partition
// index is XOR-swizzled with s so that row-ordered accesses (consecutive s of
// one g) and partition-ordered accesses (consecutive g of one s) both spread
// over the banks without padding
__device__ inline int sv_slot(int s, int g, int NT) { return s * NT + (g ^ s); }

// forward recurrence over the kPM rows of partition g of a real LU table
// from state h (h[j] = y_{r0-1-j}); v/y rows live in the row-staging buffer.
// Table rows are fetched in blocks of RB rows, every load of a block issued
// before its first use.
template <class TF, bool STORE>
__device__ inline void fwd_pass(const TF* __restrict__ Tg, cplx* sv, int NT, int g, cplx (&h)[kP]) {
    constexpr int RB = 192 / (2 * (int)sizeof(TF));  // rows per block: 192 VGPRs of coefficients
    // the block loop stays rolled: the whole solve kernel must fit the
    // instruction cache (a fully unrolled solve is ~100 KB of code)
#pragma unroll 1
    for (int s0 = 0; s0 < kPM; s0 += RB) {
        TF c[RB][kP + 1];
#pragma unroll
        for (int b = 0; b < RB; ++b)
#pragma unroll
            for (int q = 0; q <= kP; ++q)
                if (s0 + b < kPM) c[b][q] = Tg[(int64_t)((s0 + b) * kTab + q) * kLane];
        cplx vr[RB];
#pragma unroll
        for (int b = 0; b < RB; ++b)
            if (s0 + b < kPM) vr[b] = sv[sv_slot(s0 + b, g, NT)];
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            const int s = s0 + b;
            if (s < kPM) {
                cplx acc = vr[b];
#pragma unroll
                for (int j = 0; j < kP; ++j) acc -= tmul(c[b][j], h[j]);
                const cplx y = tmul(c[b][kP], acc);
#pragma unroll
                for (int j = kP - 1; j > 0; --j) h[j] = h[j - 1];
                h[0] = y;
                if (STORE) sv[sv_slot(s, g, NT)] = y;
            }
        }
    }
}

// backward recurrence over the rows of partition g of a real LU table from
// state h (h[k] = x_{r0+kPM+k})
template <class TF, bool STORE>
__device__ inline void bwd_pass(const TF* __restrict__ Tg, cplx* sv, int NT, int g, cplx (&h)[kP]) {
    constexpr int RB = 192 / (2 * (int)sizeof(TF));
#pragma unroll 1
    for (int s0 = kPM - 1; s0 >= 0; s0 -= RB) {
        TF c[RB][kP + 1];
#pragma unroll
        for (int b = 0; b < RB; ++b)
#pragma unroll
            for (int q = 0; q <= kP; ++q)
                if (s0 - b >= 0) c[b][q] = Tg[(int64_t)((s0 - b) * kTab + kP + 1 + q) * kLane];
        cplx yr[RB];
#pragma unroll
        for (int b = 0; b < RB; ++b)
            if (s0 - b >= 0) yr[b] = sv[sv_slot(s0 - b, g, NT)];
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            const int s = s0 - b;
            if (s >= 0) {
                cplx x = tmul(c[b][kP], yr[b]);
#pragma unroll
                for (int k = 0; k < kP; ++k) x -= tmul(c[b][k], h[k]);
#pragma unroll
                for (int k = kP - 1; k > 0; --k) h[k] = h[k - 1];
                h[0] = x;
                if (STORE) sv[sv_slot(s, g, NT)] = x;
            }
        }
    }
}

// Complex symmetric tables.  The forward coefficient of z_{i-1-j} in row i
// is E_{j+1} of row i-1-j; a register ring keeps, for each of the seven rows
// above, the entries later rows still need: ring[m][k] = E_{k+1}(i-1-m) for
// k >= m.  At a partition's first row the ring holds the last seven rows of
// the partition before (pass B; pass A starts from a zero state, where they
// multiply zeros and are left zero).
struct SymRing {
    cplx r[kP][kP];
    __device__ inline void clear() {
#pragma unroll
        for (int m = 0; m < kP; ++m)
#pragma unroll
            for (int k = m; k < kP; ++k) r[m][k] = {0.0, 0.0};
    }
    // the last kP rows of the partition whose table column starts at Tq
    __device__ inline void load_tail(const cplx* __restrict__ Tq) {
#pragma unroll
        for (int m = 0; m < kP; ++m)
#pragma unroll
            for (int k = m; k < kP; ++k) r[m][k] = Tq[(int64_t)((kPM - 1 - m) * kTabC + k) * kLane];
    }
    // row i done: its entries e[0..kP-1] become ring row 0
    __device__ inline void push(const cplx (&e)[kTabC]) {
#pragma unroll
        for (int m = kP - 1; m > 0; --m)
#pragma unroll
            for (int k = m; k < kP; ++k) r[m][k] = r[m - 1][k];
#pragma unroll
        for (int k = 0; k < kP; ++k) r[0][k] = e[k];
    }
};

// forward recurrence over the kPM rows of partition g of a complex symmetric
// table from state h (h[j] = z_{r0-1-j}); pass B (STORE) writes y = rD z
template <bool STORE>
__device__ inline void fwd_pass_sym(const cplx* __restrict__ Tg, SymRing& ring, cplx* sv, int NT, int g,
                                    cplx (&h)[kP]) {
    constexpr int RB = 6;  // rows per block: 48 loads in flight (a wave counts at most 63)
    static_assert(kPM % RB == 0, "a pass is a whole number of blocks");
#pragma unroll 1
    for (int s0 = 0; s0 < kPM; s0 += RB) {
        cplx c[RB][kTabC];
#pragma unroll
        for (int b = 0; b < RB; ++b)
#pragma unroll
            for (int q = 0; q < kTabC; ++q) c[b][q] = Tg[(int64_t)((s0 + b) * kTabC + q) * kLane];
        cplx vr[RB];  // the block's right-hand sides, read before the recurrence needs them
#pragma unroll
        for (int b = 0; b < RB; ++b) vr[b] = sv[sv_slot(s0 + b, g, NT)];
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            const int s = s0 + b;
            cplx acc = vr[b];
#pragma unroll
            for (int j = 0; j < kP; ++j) acc = tmsub(acc, ring.r[j][j], h[j]);
#pragma unroll
            for (int j = kP - 1; j > 0; --j) h[j] = h[j - 1];
            h[0] = acc;
            if (STORE) sv[sv_slot(s, g, NT)] = c[b][kP] * acc;
            ring.push(c[b]);
        }
    }
}

// backward recurrence over the rows of partition g of a complex symmetric
// table from state h (h[k] = x_{r0+kPM+k}); the rows hold y = rD z
template <bool STORE>
__device__ inline void bwd_pass_sym(const cplx* __restrict__ Tg, cplx* sv, int NT, int g, cplx (&h)[kP]) {
    constexpr int RB = 8;  // rows per block: 56 loads in flight (a wave counts at most 63)
    static_assert(kPM % RB == 0, "a pass is a whole number of blocks");
#pragma unroll 1
    for (int s0 = kPM - 1; s0 >= 0; s0 -= RB) {
        cplx c[RB][kP];
#pragma unroll
        for (int b = 0; b < RB; ++b)
#pragma unroll
            for (int q = 0; q < kP; ++q) c[b][q] = Tg[(int64_t)((s0 - b) * kTabC + q) * kLane];
        cplx yr[RB];
#pragma unroll
        for (int b = 0; b < RB; ++b) yr[b] = sv[sv_slot(s0 - b, g, NT)];
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            const int s = s0 - b;
            cplx x = yr[b];
#pragma unroll
            for (int k = 0; k < kP; ++k) x = tmsub(x, c[b][k], h[k]);
#pragma unroll
            for (int k = kP - 1; k > 0; --k) h[k] = h[k - 1];
            h[0] = x;
            if (STORE) sv[sv_slot(s, g, NT)] = x;
        }
    }
}

// truncated carry by Horner: acc <- E[gp] + T[gp] acc for gp = g -/+ k,
// k = PK..1 (partitions beyond the channel ends do not exist: exact start).
// The first existing term starts from the zero state, so its transfer
// product vanishes and acc = E[gp].  End state and transfer entry e of
// partition gp are column gp - base of sE and of sT (row e).
// Transfer entry e of partition gp: staged in LDS (column gp - base of row e)
// or read from the transfer set in global memory.
template <class TF, bool LDS_T>
struct TrfAt {
    const TF* p;  // LDS: the direction's staged rows; global: the direction's set
    int NT, base;
    __device__ inline TF operator()(int gp, int e) const {
        if constexpr (LDS_T) return p[e * NT + gp - base];
        else return p[(int64_t)(gp >> 6) * 2 * kTrf * kLane + (int64_t)e * kLane + (gp & 63)];
    }
};

template <class TF, int PK, bool LDS_T>
__device__ inline void carry(TrfAt<TF, LDS_T> T, const cplx* sE, int NT, int g, int dir, int P, int base,
                             cplx (&acc)[kP]) {
#pragma unroll
    for (int r = 0; r < kP; ++r) acc[r] = {0.0, 0.0};
    bool started = false;
#pragma unroll 1
    for (int k = PK; k >= 1; --k) {
        const int gp = dir == 0 ? g - k : g + k;
        if (gp >= 0 && gp < P) {
            const int lg = gp - base;
            if (!started) {
#pragma unroll
                for (int r = 0; r < kP; ++r) acc[r] = sE[r * NT + lg];
                started = true;
                continue;
            }
            // the whole transfer matrix and end state are read before the
            // first product (reads in flight together, not row by row)
            TF t[kP][kP];
            cplx e[kP];
#pragma unroll
            for (int r = 0; r < kP; ++r) {
#pragma unroll
                for (int c = 0; c < kP; ++c) t[r][c] = T(gp, r * kP + c);
                e[r] = sE[r * NT + lg];
            }
            cplx nv[kP];
#pragma unroll
            for (int r = 0; r < kP; ++r) {
                cplx a = e[r];
#pragma unroll
                for (int c = 0; c < kP; ++c) a = tmac(a, t[r][c], acc[c]);
                nv[r] = a;
            }
#pragma unroll
            for (int r = 0; r < kP; ++r) acc[r] = nv[r];
        }
    }
}

// v = B z for per-table bands (RHS_CBAND), one lane per (row, channel)
__global__ __launch_bounds__(256) void k_rhs(RhsArgs ra, TabMap tm, const cplx* __restrict__ zin,
                                             cplx* __restrict__ vout, int N) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int ch = blockIdx.y;
    if (i >= N) return;
    CBandRow b;
    b.load(ra.cband + (int64_t)table_of(tm, ch) * (kP + 1) * N, i, N);
    const cplx* z = zin + (int64_t)ch * N;
    cplx zl[kP], zu[kP + 1];
#pragma unroll
    for (int j = 1; j <= kP; ++j) zl[j - 1] = z[max(i - j, 0)];
#pragma unroll
    for (int k = 0; k <= kP; ++k) zu[k] = z[min(i + k, N - 1)];
    vout[(int64_t)ch * N + i] = b.apply(zl, zu, i);
}

// v = B z for bands shared by every channel (RHS_RBAND, RHS_SX_*): a lane
// loads the band entries of its row once and applies them to kRhsCG channels
// (2 measured fastest: 4 keeps too many loads live, 1 repeats the band loads)
constexpr int kRhsCG = 2;
template <int RHS>
__global__ __launch_bounds__(256) void k_rhs_shared(RhsArgs ra, const cplx* __restrict__ zin,
                                                    cplx* __restrict__ vout, int N, int nch) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    SBandRow<RHS> b;
    b.load(ra, i, N);
#pragma unroll
    for (int c = 0; c < kRhsCG; ++c) {
        const int ch = blockIdx.y * kRhsCG + c;
        if (ch < nch) {
            const cplx* z = zin + (int64_t)ch * N;
            cplx zlo[kP], zup[kP + 1];
#pragma unroll
            for (int j = 1; j <= kP; ++j) zlo[j - 1] = z[max(i - j, 0)];
#pragma unroll
            for (int k = 0; k <= kP; ++k) zup[k] = z[min(i + k, N - 1)];
            const double e = SBandRow<RHS>::SX ? ra.eig[ch] * ra.ifdt : 0.0;
            vout[(int64_t)ch * N + i] = b.apply(zlo, zup, e);
        }
    }
}

// x = U^-1 L^-1 v, v = B z.  A channel's partitions are cut
// into segments of up to SN partitions, one 64-lane work-group per (segment,
// channel), one lane per partition, so that a channel spreads over several
// compute units.  A work-group also runs the PK partitions on either side of
// its segment (halo): the forward carries of its own partitions need the
// pass-A end states of the PK partitions before them, and the backward
// carries need the PK partitions after the segment forward-solved.  Every
// value computed for an owned partition depends on these partitions only, so
// the segmentation does not change a single bit of the result.  The staged
// partitions' transfer sets are fetched into LDS together with the
// right-hand side, so the carries make no dependent memory round trips.
// Requires SN + 2 PK <= 64.  ROLE is the RhsKind of the operator; it sets the
// carry depth PK (qpc_solve.h).  On the small grids that stage the transfer
// sets (LDS_T: one or two work-groups per CU) a work-group has kStageWaves
// waves: all of them fetch the right-hand side and the transfer sets and
// write the result, on SIMDs that would otherwise idle, and wave 0 runs the
// passes and carries.
//
// FUSE (LDS_T grids only): src is the state z and the work-group forms
// v = B z itself while staging (the RhsKind ROLE with the bands of ra, the
// same arithmetic as k_rhs / k_rhs_shared): the staged rows and kP rows on
// either side of z go to LDS, and each staged row's band entries are read
// from L2.  Other work-groups read z rows this one owns, so src and zout may
// not overlap.  Without FUSE, src is v.
constexpr int kStageWaves = 4;
template <class TF, int ROLE, bool LDS_T, bool FUSE>
__global__ __launch_bounds__(kLane * (LDS_T ? kStageWaves : 1)) void k_bsolve(
    const TF* __restrict__ tab, const TF* __restrict__ trf, TabMap tm, const cplx* __restrict__ src,
    cplx* __restrict__ zout, Geom geo, int SN, RhsArgs ra) {
    static_assert(LDS_T || !FUSE, "the fused right-hand side is built for the staged grids");
    constexpr int NT = kLane;
    constexpr int NTH = kLane * (LDS_T ? kStageWaves : 1);  // threads of the work-group
    constexpr int KV = kPM * NT / NTH;                       // staged rows per thread
    static_assert(kPM * NT % NTH == 0, "whole staged rows per thread");
    constexpr int PK = ROLE == RHS_SX_ASY ? kPKD : kPKS;
    constexpr int NZ = kPM * NT + 2 * kP;  // z rows of a fused right-hand side
    __shared__ cplx sv[kPM * NT];
    __shared__ cplx sE[kP * NT];
    // transfer sets of the staged partitions, both directions (LDS_T only)
    __shared__ TF sT[LDS_T ? 2 : 1][LDS_T ? kTrf * NT : 1];
    // z rows [r0 - kP, r0 + NZ - kP) (FUSE only; zero outside [0, N))
    __shared__ cplx sz[FUSE ? NZ : 1];

    const int ch = blockIdx.y;
    const int N = geo.N, P = geo.P;
    const int k0 = blockIdx.x * SN, k1 = min(P, k0 + SN);  // owned partitions [k0, k1)
    const int ga = max(0, k0 - PK), gb = min(P, k1 + PK);  // staged partitions [ga, gb)
    const int tid = threadIdx.x;
    const int g = ga + tid;  // this lane's partition (wave 0)
    const bool run = tid < kLane && g < gb;
    const int tabid = table_of(tm, ch);
    const int64_t r0 = (int64_t)ga * kPM;  // first staged row
    const int nrow = (gb - ga) * kPM;
    const int nv = min(nrow, N - (int)r0);  // staged rows below N
    cplx* xo = zout + (int64_t)ch * N + r0;
    const TF* T0 = tab + (int64_t)tabid * TabFmt<TF>::elems(geo);
    const TF* Tg = T0 + part_col<TF>(g);
    constexpr bool SYM = TabFmt<TF>::row == kTabC;

    // right-hand side rows (padding rows: 0), or the z rows it is formed
    // from (the transfer sets follow during pass A).  The row loads are
    // unconditional (indices clamped into the channel) so that all of them
    // are in flight at once.
    if constexpr (FUSE) {
        constexpr int KZ = (NZ + NTH - 1) / NTH;
        const cplx* zc = src + (int64_t)ch * N;
        const int zlo = (int)r0 - kP;
        cplx zr[KZ];
#pragma unroll
        for (int k = 0; k < KZ; ++k) zr[k] = zc[min(max(zlo + tid + k * NTH, 0), N - 1)];
#pragma unroll
        for (int k = 0; k < KZ; ++k) {
            const int j = tid + k * NTH, row = zlo + j;
            if (j < NZ) sz[j] = row >= 0 && row < N ? zr[k] : cplx{0.0, 0.0};
        }
    } else {
        const cplx* v = src + (int64_t)ch * N + r0;
        cplx vr[KV];
#pragma unroll
        for (int k = 0; k < KV; ++k) vr[k] = v[min(tid + k * NTH, nv - 1)];
#pragma unroll
        for (int k = 0; k < KV; ++k) {
            const int i = tid + k * NTH;
            if (i < nrow) sv[sv_slot(i % kPM, i / kPM, NT)] = i < nv ? vr[k] : cplx{0.0, 0.0};
        }
    }
    const TF* TR = trf + (int64_t)tabid * geo.trf_elems;
    __syncthreads();
    if constexpr (FUSE) {
        // v rows of the staged partitions from z in LDS: staged row i is
        // channel row r0 + i, its z_{r0+i+d} sits at sz[i + kP + d]
        const double e = ROLE == RHS_SX_SYM || ROLE == RHS_SX_ASY ? ra.eig[ch] * ra.ifdt : 0.0;
#pragma unroll
        for (int k = 0; k < KV; ++k) {
            const int i = tid + k * NTH;
            const int row = min((int)r0 + i, N - 1);  // rows from N on are padding: v = 0
            cplx zl[kP], zu[kP + 1];
#pragma unroll
            for (int j = 1; j <= kP; ++j) zl[j - 1] = sz[i + kP - j];
#pragma unroll
            for (int q = 0; q <= kP; ++q) zu[q] = sz[i + kP + q];
            cplx vv;
            if constexpr (ROLE == RHS_CBAND) {
                CBandRow b;
                b.load(ra.cband + (int64_t)tabid * (kP + 1) * N, row, N);
                vv = b.apply(zl, zu, row);
            } else {
                SBandRow<ROLE> b;
                b.load(ra, row, N);
                vv = b.apply(zl, zu, e);
            }
            if (i < nrow) sv[sv_slot(i % kPM, i / kPM, NT)] = i < nv ? vv : cplx{0.0, 0.0};
        }
        __syncthreads();
    }
    const TrfAt<TF, LDS_T> Tf{LDS_T ? &sT[0][0] : TR, NT, ga};
    const TrfAt<TF, LDS_T> Tb{LDS_T ? &sT[LDS_T ? 1 : 0][0] : TR + (int64_t)kTrf * kLane, NT, ga};

    // forward pass A on every staged partition (wave 0).  Meanwhile the
    // other waves fetch the staged partitions' transfer sets, which the
    // carries read only after the next barrier: entry q = d * kTrf + e of
    // partition gl goes to row q of the staging buffer, wave w (>= 1)
    // fetching the entries q = w - 1 mod (kStageWaves - 1).
    if constexpr (LDS_T) {
        constexpr int NW = kStageWaves - 1, NQ = 2 * kTrf, QW = (NQ + NW - 1) / NW;
        const int ln = tid & (kLane - 1), w = tid / kLane - 1;
        const int gl = ga + ln;
        if (w >= 0 && gl < gb) {
            const TF* Tr = TR + (int64_t)(gl >> 6) * 2 * kTrf * kLane + (gl & 63);
            TF tv[QW];
#pragma unroll
            for (int j = 0; j < QW; ++j) tv[j] = Tr[(int64_t)min(w + j * NW, NQ - 1) * kLane];
            TF* st = &sT[0][0];
#pragma unroll
            for (int j = 0; j < QW; ++j)
                if (w + j * NW < NQ) st[(w + j * NW) * NT + ln] = tv[j];
        }
    }
    cplx h[kP];
    if (run) {
#pragma unroll
        for (int j = 0; j < kP; ++j) h[j] = {0.0, 0.0};
        if constexpr (SYM) {
            SymRing ring;
            ring.clear();
            fwd_pass_sym<false>(Tg, ring, sv, NT, tid, h);
        } else {
            fwd_pass<TF, false>(Tg, sv, NT, tid, h);
        }
#pragma unroll
        for (int j = 0; j < kP; ++j) sE[j * NT + tid] = h[j];
    }
    __syncthreads();

    // forward carry and pass B on the owned partitions and the trailing halo
    const bool fw = run && g >= k0;
    if (fw) {
        carry<TF, PK, LDS_T>(Tf, sE, NT, g, 0, P, ga, h);
        if constexpr (SYM) {
            SymRing ring;
            if (g > 0) ring.load_tail(T0 + part_col<TF>(g - 1));
            else ring.clear();
            fwd_pass_sym<true>(Tg, ring, sv, NT, tid, h);
        } else {
            fwd_pass<TF, true>(Tg, sv, NT, tid, h);
        }
    }
    __syncthreads();

    // backward pass A on the same partitions
    if (fw) {
#pragma unroll
        for (int j = 0; j < kP; ++j) h[j] = {0.0, 0.0};
        if constexpr (SYM) bwd_pass_sym<false>(Tg, sv, NT, tid, h);
        else bwd_pass<TF, false>(Tg, sv, NT, tid, h);
#pragma unroll
        for (int j = 0; j < kP; ++j) sE[j * NT + tid] = h[j];
    }
    __syncthreads();

    // backward carry and pass B on the owned partitions
    if (fw && g < k1) {
        carry<TF, PK, LDS_T>(Tb, sE, NT, g, 1, P, ga, h);
        if constexpr (SYM) bwd_pass_sym<true>(Tg, sv, NT, tid, h);
        else bwd_pass<TF, true>(Tg, sv, NT, tid, h);
    }
    __syncthreads();
    const int i0 = (k0 - ga) * kPM, i1 = min((k1 - ga) * kPM, nv);
#pragma unroll
    for (int k = 0; k < KV; ++k) {
        const int i = tid + k * NTH;
        if (i >= i0 && i < i1) xo[i] = sv[sv_slot(i % kPM, i / kPM, NT)];
    }
}

// ---------------------------------------------------------------------------
// transfer sets of real LU tables: one lane per (table, partition); the
// seven unit responses of each direction are propagated together so every
// table row is read once
template <class TF>
__global__ __launch_bounds__(kLane) void k_transfer(const TF* __restrict__ tab, TF* __restrict__ trf, Geom geo) {
    const int g = blockIdx.y * kLane + threadIdx.x;
    const int t = blockIdx.x;
    if (g >= geo.P) return;
    const TF* Tg = tab + (int64_t)t * geo.tab_elems + (int64_t)(g >> 6) * kPM * kTab * kLane + (g & 63);
    TF* out = trf + (int64_t)t * geo.trf_elems + (int64_t)((g >> 6) * 2) * kTrf * kLane + (g & 63);
    for (int dir = 0; dir < 2; ++dir) {
        TF H[kP][kP];  // H[q][col]: state entry q of the response to unit vector col
#pragma unroll
        for (int q = 0; q < kP; ++q)
#pragma unroll
            for (int col = 0; col < kP; ++col) H[q][col] = q == col ? tone(TF{}) : tzero(TF{});
        for (int ss = 0; ss < kPM; ++ss) {
            const int s = dir == 0 ? ss : kPM - 1 - ss;
            const TF* c = Tg + (int64_t)(s * kTab + (dir == 0 ? 0 : kP + 1)) * kLane;
            TF f[kP + 1];
#pragma unroll
            for (int q = 0; q <= kP; ++q) f[q] = c[(int64_t)q * kLane];
            TF y[kP];
#pragma unroll
            for (int col = 0; col < kP; ++col) {
                TF acc = tzero(TF{});
#pragma unroll
                for (int q = 0; q < kP; ++q) acc = acc - f[q] * H[q][col];
                y[col] = dir == 0 ? f[kP] * acc : acc;
            }
#pragma unroll
            for (int q = kP - 1; q > 0; --q)
#pragma unroll
                for (int col = 0; col < kP; ++col) H[q][col] = H[q - 1][col];
#pragma unroll
            for (int col = 0; col < kP; ++col) H[0][col] = y[col];
        }
        TF* o = out + (int64_t)dir * kTrf * kLane;
#pragma unroll
        for (int q = 0; q < kP; ++q)
#pragma unroll
            for (int col = 0; col < kP; ++col) o[(int64_t)(q * kP + col) * kLane] = H[q][col];
    }
}

// transfer sets of complex symmetric tables (the z recurrence forward, the x
// recurrence backward); the forward responses start from the coefficients of
// the previous partition's last rows, as the solve's pass B does
__global__ __launch_bounds__(kLane) void k_transfer_sym(const cplx* __restrict__ tab, cplx* __restrict__ trf,
                                                        Geom geo) {
    const int g = blockIdx.y * kLane + threadIdx.x;
    const int t = blockIdx.x;
    if (g >= geo.P) return;
    const cplx* T0 = tab + (int64_t)t * geo.ctab_elems;
    const cplx* Tg = T0 + part_col<cplx>(g);
    cplx* out = trf + (int64_t)t * geo.trf_elems + (int64_t)((g >> 6) * 2) * kTrf * kLane + (g & 63);
    cplx H[kP][kP];  // H[q][col]: state entry q of the response to unit vector col
    // forward
#pragma unroll
    for (int q = 0; q < kP; ++q)
#pragma unroll
        for (int col = 0; col < kP; ++col) H[q][col] = q == col ? cplx{1.0, 0.0} : cplx{0.0, 0.0};
    SymRing ring;
    if (g > 0) ring.load_tail(T0 + part_col<cplx>(g - 1));
    else ring.clear();
    for (int s = 0; s < kPM; ++s) {
        cplx e[kTabC];
#pragma unroll
        for (int q = 0; q < kTabC; ++q) e[q] = Tg[(int64_t)(s * kTabC + q) * kLane];
        cplx y[kP];
#pragma unroll
        for (int col = 0; col < kP; ++col) {
            cplx acc = {0.0, 0.0};
#pragma unroll
            for (int q = 0; q < kP; ++q) acc = acc - ring.r[q][q] * H[q][col];
            y[col] = acc;
        }
#pragma unroll
        for (int q = kP - 1; q > 0; --q)
#pragma unroll
            for (int col = 0; col < kP; ++col) H[q][col] = H[q - 1][col];
#pragma unroll
        for (int col = 0; col < kP; ++col) H[0][col] = y[col];
        ring.push(e);
    }
#pragma unroll
    for (int q = 0; q < kP; ++q)
#pragma unroll
        for (int col = 0; col < kP; ++col) out[(int64_t)(q * kP + col) * kLane] = H[q][col];
    // backward
#pragma unroll
    for (int q = 0; q < kP; ++q)
#pragma unroll
        for (int col = 0; col < kP; ++col) H[q][col] = q == col ? cplx{1.0, 0.0} : cplx{0.0, 0.0};
    for (int ss = 0; ss < kPM; ++ss) {
        const int s = kPM - 1 - ss;
        cplx e[kP];
#pragma unroll
        for (int q = 0; q < kP; ++q) e[q] = Tg[(int64_t)(s * kTabC + q) * kLane];
        cplx y[kP];
#pragma unroll
        for (int col = 0; col < kP; ++col) {
            cplx acc = {0.0, 0.0};
#pragma unroll
            for (int q = 0; q < kP; ++q) acc = acc - e[q] * H[q][col];
            y[col] = acc;
        }
#pragma unroll
        for (int q = kP - 1; q > 0; --q)
#pragma unroll
            for (int col = 0; col < kP; ++col) H[q][col] = H[q - 1][col];
#pragma unroll
        for (int col = 0; col < kP; ++col) H[0][col] = y[col];
    }
    cplx* o = out + (int64_t)kTrf * kLane;
#pragma unroll
    for (int q = 0; q < kP; ++q)
#pragma unroll
        for (int col = 0; col < kP; ++col) o[(int64_t)(q * kP + col) * kLane] = H[q][col];
}

// ---------------------------------------------------------------------------
// complex symmetric tables from the lu_zsb factor (ed.h:208-289):
// f[i*ne] = D_i, f[i*ne+j] = L(i+j,i); one lane per (table, partition)
__global__ __launch_bounds__(kLane) void k_tab_ldlt(const cplx* __restrict__ lower, cplx* __restrict__ tab, Geom geo) {
    const int g = blockIdx.x * kLane + threadIdx.x;
    const int l = blockIdx.y;
    if (g >= geo.P) return;
    const int N = geo.N;
    constexpr int ne = kP + 1;
    const cplx* f = lower + (int64_t)l * N * ne;
    cplx* Tg = tab + (int64_t)l * geo.ctab_elems + part_col<cplx>(g);
    for (int s = 0; s < kPM; ++s) {
        const int i = g * kPM + s;
        cplx o[kTabC];
        if (i >= N) {  // identity padding row
#pragma unroll
            for (int q = 0; q < kP; ++q) o[q] = {0.0, 0.0};
            o[kP] = {1.0, 0.0};
        } else {
#pragma unroll
            for (int k = 1; k <= kP; ++k) o[k - 1] = i + k < N ? f[(int64_t)i * ne + k] : cplx{0.0, 0.0};
            o[kP] = crecip(f[(int64_t)i * ne]);
        }
#pragma unroll
        for (int q = 0; q < kTabC; ++q) Tg[(int64_t)(s * kTabC + q) * kLane] = o[q];
    }
}

// tables from chol(S) (ed.h:182-194): c[i*ne+j] = L(i+j,i)
__global__ __launch_bounds__(kLane) void k_tab_chol(const double* __restrict__ c, double* __restrict__ tab, Geom geo) {
    const int g = blockIdx.x * kLane + threadIdx.x;
    if (g >= geo.P) return;
    const int N = geo.N;
    constexpr int ne = kP + 1;
    double* Tg = tab + (int64_t)(g >> 6) * kPM * kTab * kLane + (g & 63);
    for (int s = 0; s < kPM; ++s) {
        const int i = g * kPM + s;
        double o[kTab];
        if (i >= N) {
#pragma unroll
            for (int q = 0; q < kTab; ++q) o[q] = 0.0;
            o[kP] = 1.0;
            o[kTab - 1] = 1.0;
        } else {
            const double d = 1.0 / c[(int64_t)i * ne];
#pragma unroll
            for (int j = 1; j <= kP; ++j) o[j - 1] = i - j >= 0 ? c[(int64_t)(i - j) * ne + j] : 0.0;
            o[kP] = d;
#pragma unroll
            for (int k = 1; k <= kP; ++k) o[kP + k] = i + k < N ? c[(int64_t)i * ne + k] * d : 0.0;
            o[kTab - 1] = d;
        }
#pragma unroll
        for (int q = 0; q < kTab; ++q) Tg[(int64_t)(s * kTab + q) * kLane] = o[q];
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_band_diag(const T* __restrict__ src, T* __restrict__ dst, int64_t total,
                                                   int N) {
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= total) return;
    constexpr int ne = kP + 1;
    const int64_t b = id / ((int64_t)ne * N);
    const int64_t r = id % ((int64_t)ne * N);
    const int j = (int)(r / N), i = (int)(r % N);
    dst[id] = src[(b * N + i) * ne + j];
}

// ---------------------------------------------------------------------------
// reference-order LU of b = S + ifet X (propagate.hpp:33-86), one lane per
// (time point, channel, partition).  A lane restarts the elimination kRestart
// rows above its partition (as if the matrix began there) and writes the
// table rows of its own partition.  The last seven U rows and the upper-band
// entries of S and X that later rows need as their lower panel live in
// register rings (row i in slot i mod 8).
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

__device__ inline cplx frcp(cplx b) { return crecip(b); }
__device__ inline double frcp(double b) { return 1.0 / b; }

template <int SYM>
__global__ __launch_bounds__(kLane) void k_factor(const double* __restrict__ S, const double* __restrict__ X,
                                                  const double* __restrict__ eig, int64_t eig_tstride,
                                                  const double* __restrict__ ifdt, int nch, Geom geo,
                                                  typename FactorTypes<SYM>::T* __restrict__ tab) {
    using T = typename FactorTypes<SYM>::T;
    const int g = blockIdx.x * kLane + threadIdx.x;
    const int ch = blockIdx.y, t = blockIdx.z;
    if (g >= geo.P) return;
    const int N = geo.N;
    constexpr int nd = kP, ne = kP + 1;
    const double e = eig[(int64_t)t * eig_tstride + ch] * ifdt[t];
    // SYM 0 (S + i e X, X symmetric) writes complex symmetric rows [u(i,0..6), 1/pivot]:
    // for a symmetric matrix the unit upper factor's rows are the columns of
    // L in A = L D L^T and the pivots are D; SYM 1 writes real LU rows
    constexpr int ROW = TabFmt<T>::row;
    T* Tg = tab + ((int64_t)t * nch + ch) * TabFmt<T>::elems(geo) + part_col<T>(g);
    const int r0 = g * kPM;
    const int start = r0 > kRestart ? r0 - kRestart : 0;  // multiple of 8
    const int end = min(N, r0 + kPM);
    T u[8][nd];
    double ws[8][nd], wx[8][nd];  // entries 1..7 of the upper bands of recent rows
#pragma unroll
    for (int s = 0; s < 8; ++s)
#pragma unroll
        for (int k = 0; k < nd; ++k) {
            u[s][k] = tzero(T{});
            ws[s][k] = 0.0;
            wx[s][k] = 0.0;
        }
    for (int i0 = start; i0 < end; i0 += 8) {
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int i = i0 + s;
            if (i < end) {
                double sr[ne], xr[ne];
#pragma unroll
                for (int j = 0; j < ne; ++j) {
                    sr[j] = S[(int64_t)i * ne + j];
                    xr[j] = X[(int64_t)i * ne + j];
                }
                const int avail = i - start;  // rows above i in the restarted matrix
                T b[2 * nd + 1];
                // lower panel: column i-j (propagate.hpp:35-53)
#pragma unroll
                for (int j = 1; j <= nd; ++j) {
                    const int slot = (s - j) & 7;
                    if (j > avail) {
                        b[nd - j] = tzero(T{});
                    } else {
                        const double sv = ws[slot][j - 1], rv = wx[slot][j - 1];
                        if constexpr (SYM == 0) b[nd - j] = cplx{sv, e * rv};  // s + ifet r, ifet = i e
                        else b[nd - j] = sv - e * rv;                            // s - ifet r (antisymmetric)
                    }
                }
                // upper (propagate.hpp:55-62)
#pragma unroll
                for (int j = 0; j < ne; ++j) {
                    if constexpr (SYM == 0) b[nd + j] = cplx{sr[j], e * xr[j]};
                    else b[nd + j] = sr[j] + e * xr[j];
                }
#pragma unroll
                for (int j = 1; j < ne; ++j) {
                    ws[s][j - 1] = sr[j];
                    wx[s][j - 1] = xr[j];
                }
                const int l = avail < nd ? nd - avail : 0;
                // elimination with the previous U rows (propagate.hpp:67-71)
#pragma unroll
                for (int j = 0; j < nd; ++j) {
                    if (j >= l) {
                        const int slot = (s + j + 1) & 7;  // row i+j-7
#pragma unroll
                        for (int k = 0; k < nd; ++k) b[j + k + 1] = b[j + k + 1] - b[j] * u[slot][k];
                    }
                }
                // u(i,:) = b / pivot (propagate.hpp:72-75), via one reciprocal
                const T rp = frcp(b[nd]);
#pragma unroll
                for (int j = 0; j < nd; ++j) u[s][j] = b[j + nd + 1] * rp;
                if (i >= r0) {
                    T* o = Tg + (int64_t)(i - r0) * ROW * kLane;
                    if constexpr (SYM == 0) {
#pragma unroll
                        for (int k = 1; k <= nd; ++k) o[(int64_t)(k - 1) * kLane] = i + k < N ? u[s][k - 1] : tzero(T{});
                        o[(int64_t)nd * kLane] = rp;
                    } else {
#pragma unroll
                        for (int j = 1; j <= nd; ++j) o[(int64_t)(j - 1) * kLane] = b[nd - j];  // multiplier of column i-j
                        o[(int64_t)nd * kLane] = rp;
#pragma unroll
                        for (int k = 1; k <= nd; ++k)
                            o[(int64_t)(nd + k) * kLane] = i + k < N ? u[s][k - 1] : tzero(T{});
                        o[(int64_t)(kTab - 1) * kLane] = tone(T{});
                    }
                }
            }
        }
    }
    // identity padding rows of the last partition
    for (int i = max(N, r0); i < r0 + kPM; ++i) {
        T* o = Tg + (int64_t)(i - r0) * ROW * kLane;
#pragma unroll
        for (int q = 0; q < ROW; ++q) o[(int64_t)q * kLane] = tzero(T{});
        o[(int64_t)nd * kLane] = tone(T{});
        o[(int64_t)(ROW - 1) * kLane] = tone(T{});
    }
}

}  // namespace

// ---------------------------------------------------------------------------
void launch_rhs(hipStream_t s, int rhs, const RhsArgs& ra, TabMap tm, const cplx* zin, cplx* vout, int nch, Geom geo) {
    dim3 rgrid((geo.N + 255) / 256, nch);
    dim3 sgrid((geo.N + 255) / 256, (nch + kRhsCG - 1) / kRhsCG);
    switch (rhs) {
        case RHS_CBAND: hipLaunchKernelGGL(k_rhs, rgrid, dim3(256), 0, s, ra, tm, zin, vout, geo.N); break;
        case RHS_RBAND: hipLaunchKernelGGL(k_rhs_shared<RHS_RBAND>, sgrid, dim3(256), 0, s, ra, zin, vout, geo.N, nch); break;
        case RHS_SX_SYM: hipLaunchKernelGGL(k_rhs_shared<RHS_SX_SYM>, sgrid, dim3(256), 0, s, ra, zin, vout, geo.N, nch); break;
        case RHS_SX_ASY: hipLaunchKernelGGL(k_rhs_shared<RHS_SX_ASY>, sgrid, dim3(256), 0, s, ra, zin, vout, geo.N, nch); break;
        default: break;
    }
}

// Largest grid whose solves stage the transfer sets in LDS, with four waves
// per work-group and the right-hand side formed in the launch.  One such
// work-group fits a CU (107-157 KB of LDS), so up to 512 run in at most two
// rounds over the 256 CUs; larger grids read the transfer sets from L2 and
// keep several single-wave work-groups per CU (64 batched shots in buckets
// of 468-720 channels: 7.59e8 coefficient updates/s with this limit, 7.01e8
// with 1024, 6.81e8 with 4096).
constexpr int64_t kStageGrid = 512;

template <class TF, int ROLE>
static void bsolve(hipStream_t s, const TF* tab, const TF* trf, TabMap tm, const cplx* zin, cplx* vbuf,
                   cplx* zout, int nch, Geom geo, const RhsArgs* ra) {
    constexpr int PK = ROLE == RHS_SX_ASY ? kPKD : kPKS;
    constexpr int smax = seg_max(PK);
    const int nseg = (geo.P + smax - 1) / smax;
    const int sn = (geo.P + nseg - 1) / nseg;  // <= smax, so sn + 2 PK <= 64
    // staged grids (kStageGrid): transfer sets in LDS, four waves, and the
    // right-hand side formed in the launch when z and x do not overlap
    const bool staged = (int64_t)nseg * nch <= kStageGrid;
    const dim3 grid(nseg, nch);
    if (ra) {
        // right-hand side formed in the solve (staged grids, disjoint z and x)
        const uintptr_t a0 = (uintptr_t)zin, b0 = (uintptr_t)zout;
        const uintptr_t len = sizeof(cplx) * (uintptr_t)nch * geo.N;
        if (staged && (a0 + len <= b0 || b0 + len <= a0)) {
            hipLaunchKernelGGL((k_bsolve<TF, ROLE, true, true>), grid, dim3(kLane * kStageWaves), 0, s, tab, trf, tm,
                               zin, zout, geo, sn, *ra);
            return;
        }
        launch_rhs(s, ROLE, *ra, tm, zin, vbuf, nch, geo);
        zin = vbuf;
    }
    if (staged)
        hipLaunchKernelGGL((k_bsolve<TF, ROLE, true, false>), grid, dim3(kLane * kStageWaves), 0, s, tab, trf, tm,
                           zin, zout, geo, sn, RhsArgs{});
    else
        hipLaunchKernelGGL((k_bsolve<TF, ROLE, false, false>), grid, dim3(kLane), 0, s, tab, trf, tm, zin, zout, geo,
                           sn, RhsArgs{});
}

void launch_bsolve_c(hipStream_t s, int role, const cplx* tab, const cplx* trf, TabMap tm, const cplx* vin,
                     cplx* zout, int nch, Geom geo) {
    if (role == RHS_CBAND) bsolve<cplx, RHS_CBAND>(s, tab, trf, tm, vin, nullptr, zout, nch, geo, nullptr);
    else bsolve<cplx, RHS_SX_SYM>(s, tab, trf, tm, vin, nullptr, zout, nch, geo, nullptr);
}
void launch_bsolve_r(hipStream_t s, int role, const double* tab, const double* trf, TabMap tm, const cplx* vin,
                     cplx* zout, int nch, Geom geo) {
    if (role == RHS_RBAND) bsolve<double, RHS_RBAND>(s, tab, trf, tm, vin, nullptr, zout, nch, geo, nullptr);
    else bsolve<double, RHS_SX_ASY>(s, tab, trf, tm, vin, nullptr, zout, nch, geo, nullptr);
}
void launch_rsolve_c(hipStream_t s, int role, const RhsArgs& ra, const cplx* tab, const cplx* trf, TabMap tm,
                     const cplx* zin, cplx* vbuf, cplx* zout, int nch, Geom geo) {
    if (role == RHS_CBAND) bsolve<cplx, RHS_CBAND>(s, tab, trf, tm, zin, vbuf, zout, nch, geo, &ra);
    else bsolve<cplx, RHS_SX_SYM>(s, tab, trf, tm, zin, vbuf, zout, nch, geo, &ra);
}
void launch_rsolve_r(hipStream_t s, int role, const RhsArgs& ra, const double* tab, const double* trf, TabMap tm,
                     const cplx* zin, cplx* vbuf, cplx* zout, int nch, Geom geo) {
    if (role == RHS_RBAND) bsolve<double, RHS_RBAND>(s, tab, trf, tm, zin, vbuf, zout, nch, geo, &ra);
    else bsolve<double, RHS_SX_ASY>(s, tab, trf, tm, zin, vbuf, zout, nch, geo, &ra);
}

void launch_transfer_c(hipStream_t s, const cplx* tab, cplx* trf, int ntab, Geom geo) {
    hipLaunchKernelGGL(k_transfer_sym, dim3(ntab, geo.NG), dim3(kLane), 0, s, tab, trf, geo);
}
void launch_transfer_r(hipStream_t s, const double* tab, double* trf, int ntab, Geom geo) {
    hipLaunchKernelGGL(k_transfer<double>, dim3(ntab, geo.NG), dim3(kLane), 0, s, tab, trf, geo);
}

void launch_tab_ldlt(hipStream_t s, const cplx* lower, cplx* tab, int nl, Geom geo) {
    hipLaunchKernelGGL(k_tab_ldlt, dim3(geo.NG, nl), dim3(kLane), 0, s, lower, tab, geo);
}
void launch_tab_chol(hipStream_t s, const double* chol, double* tab, Geom geo) {
    hipLaunchKernelGGL(k_tab_chol, dim3(geo.NG), dim3(kLane), 0, s, chol, tab, geo);
}

void launch_band_diag_r(hipStream_t s, const double* src, double* dst, int n, int N) {
    const int64_t total = (int64_t)n * (kP + 1) * N;
    hipLaunchKernelGGL(k_band_diag<double>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, src, dst, total, N);
}
void launch_band_diag_c(hipStream_t s, const cplx* src, cplx* dst, int n, int N) {
    const int64_t total = (int64_t)n * (kP + 1) * N;
    hipLaunchKernelGGL(k_band_diag<cplx>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, src, dst, total, N);
}

void launch_factor_sym(hipStream_t s, const double* S, const double* X, const double* eig, const double* ifdt, int nt,
                       int nch, Geom geo, cplx* tab, int64_t eig_tstride) {
    hipLaunchKernelGGL(k_factor<0>, dim3(geo.NG, nch, nt), dim3(kLane), 0, s, S, X, eig, eig_tstride, ifdt, nch, geo,
                       tab);
}
void launch_factor_asy(hipStream_t s, const double* S, const double* X, const double* eig, const double* ifdt, int nt,
                       int nch, Geom geo, double* tab, int64_t eig_tstride) {
    hipLaunchKernelGGL(k_factor<1>, dim3(geo.NG, nch, nt), dim3(kLane), 0, s, S, X, eig, eig_tstride, ifdt, nch, geo,
                       tab);
}

}  // namespace qpc
