// qpc_solve.hip -- partitioned banded CN solves, their tables and the
// time-batched factorisation of the coupling sweeps (gfx950), built for band
// half-widths p = 7 and p = 8 (template parameter P).
#include "qpc_solve.h"

#include <algorithm>
#include <cmath>
#include <vector>

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
template <class TF, int P>
struct TabFmt;
template <int P>
struct TabFmt<double, P> {
    static constexpr int row = 2 * (P + 1);
    __host__ __device__ static int64_t elems(const Geom& g) { return g.tab_elems; }
};
template <int P>
struct TabFmt<cplx, P> {
    static constexpr int row = P + 1;
    __host__ __device__ static int64_t elems(const Geom& g) { return g.ctab_elems; }
};
// first element of partition g's column in a table
template <class TF, int P>
__device__ inline int64_t part_col(int g) {
    return (int64_t)(g >> 6) * kPM * TabFmt<TF, P>::row * kLane + (g & 63);
}
template <int P>
constexpr int trf_n() { return P * P; }  // transfer matrix entries

__device__ inline int table_of(const TabMap& tm, int ch) {
    if (tm.mode == 4) return tm.map[ch] & kMapIdx;
    if (tm.mode == 1) return 0;
    if (tm.mode == 3) return ch % tm.L;
    if (tm.mode == 2) {
        const int im = ch / tm.L, il = ch % tm.L;
        const int m = im + tm.m0;
        return il + (m < 0 ? -m : m);
    }
    return ch;
}

// a channel whose operator is the complex conjugate of its table's (mode 4):
// it is solved as conj(A^-1 B conj(z)), the table's own arithmetic on the
// conjugated input with its scalar, then conjugated
__device__ inline bool conj_of(const TabMap& tm, int ch) { return tm.mode == 4 && (tm.map[ch] & kMapConj); }
__device__ inline cplx cj_if(bool c, cplx v) { return c ? cplx{v.re, -v.im} : v; }

// ---------------------------------------------------------------------------
// Right-hand sides.  Their loads are unconditional, from indices clamped
// into [0, N), and band entries outside the matrix are zero (stored zero, or
// zeroed after loading), so every load of a row is in flight at once: a load
// under a branch compiles to a load-and-wait pair.  The sums keep their
// order; the extra terms add exact zeros.

// Row i < N of a per-table complex band (diagonal-major, B(i, i+k) =
// U[k*N + i], zero where i + k >= N): its entries, then (B z)_i from
// zl[j-1] = z_{i-j} (j = 1..P) and zu[k] = z_{i+k} (k = 0..P)
template <int P>
struct CBandRow {
    cplx ul[P], uu[P + 1];
    __device__ inline void load(const cplx* __restrict__ U, int i, int N) {
#pragma unroll
        for (int j = 1; j <= P; ++j) ul[j - 1] = U[(int64_t)j * N + max(i - j, 0)];
#pragma unroll
        for (int k = 0; k <= P; ++k) uu[k] = U[(int64_t)k * N + i];
    }
    // explicit fused multiply-adds, so that every kernel forming the row
    // rounds it alike (a complex product leaves the contraction open)
    __device__ inline cplx apply(const cplx (&zl)[P], const cplx (&zu)[P + 1], int i) const {
        cplx v = {0.0, 0.0};
#pragma unroll
        for (int j = 1; j <= P; ++j) {
            const cplx u = i - j >= 0 ? ul[j - 1] : cplx{0.0, 0.0};
            v = tmac(v, u, zl[j - 1]);
        }
#pragma unroll
        for (int k = 0; k <= P; ++k) v = tmac(v, uu[k], zu[k]);
        return v;
    }
};

// Row i < N of the bands shared by every channel: the mask, or for the
// RHS_SX_* kinds S and X, whose entries come in {S, X} pairs (one load for
// both).  Lower entries B(i, i-j) = D[j][i-j] (zero above the first row),
// upper B(i, i+k) = D[k][i] (stored zero where i + k >= N).
template <int RHS, int P>
struct SBandRow {
    static constexpr bool SX = RHS != RHS_RBAND;
    double sl[P], su[P + 1], xl[P], xu[P + 1];
    __device__ inline void load(const RhsArgs& ra, int i, int N) {
#pragma unroll
        for (int j = 1; j <= P; ++j) {
            const int64_t k = (int64_t)j * N + max(i - j, 0);
            if constexpr (SX) {
                const cplx q = ra.sx[k];
                sl[j - 1] = q.re;
                xl[j - 1] = q.im;
            } else {
                sl[j - 1] = ra.rband[k];
                xl[j - 1] = 0.0;
            }
        }
#pragma unroll
        for (int k = 0; k <= P; ++k) {
            if constexpr (SX) {
                const cplx q = ra.sx[(int64_t)k * N + i];
                su[k] = q.re;
                xu[k] = q.im;
            } else {
                su[k] = ra.rband[(int64_t)k * N + i];
                xu[k] = 0.0;
            }
        }
#pragma unroll
        for (int j = 1; j <= P; ++j)
            if (i - j < 0) sl[j - 1] = xl[j - 1] = 0.0;
    }
    // (B z)_i from zlo[j-1] = z_{i-j}, zup[k] = z_{i+k}; e = eig * ifdt of the channel
    __device__ inline cplx apply(const cplx (&zlo)[P], const cplx (&zup)[P + 1], double e) const {
        cplx sz = {0.0, 0.0}, zl = {0.0, 0.0}, zu = {0.0, 0.0};
#pragma unroll
        for (int j = 1; j <= P; ++j) {
            sz += sl[j - 1] * zlo[j - 1];
            if (SX) zl += xl[j - 1] * zlo[j - 1];
        }
#pragma unroll
        for (int k = 0; k <= P; ++k) {
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
template <bool STORE, int P>
__device__ inline void fwd_pass(const double* __restrict__ Tg, cplx* sv, int NT, int g, cplx (&h)[P]) {
    constexpr int R = TabFmt<double, P>::row;
    constexpr int RB = 12;  // rows per block
    // the block loop stays rolled: the whole solve kernel must fit the
    // instruction cache (a fully unrolled solve is ~100 KB of code)
#pragma unroll 1
    for (int s0 = 0; s0 < kPM; s0 += RB) {
        double c[RB][P + 1];
#pragma unroll
        for (int b = 0; b < RB; ++b)
#pragma unroll
            for (int q = 0; q <= P; ++q)
                if (s0 + b < kPM) c[b][q] = Tg[(int64_t)((s0 + b) * R + q) * kLane];
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
                for (int j = 0; j < P; ++j) acc -= tmul(c[b][j], h[j]);
                const cplx y = tmul(c[b][P], acc);
#pragma unroll
                for (int j = P - 1; j > 0; --j) h[j] = h[j - 1];
                h[0] = y;
                if (STORE) sv[sv_slot(s, g, NT)] = y;
            }
        }
    }
}

// backward recurrence over the rows of partition g of a real LU table from
// state h (h[k] = x_{r0+kPM+k})
template <bool STORE, int P>
__device__ inline void bwd_pass(const double* __restrict__ Tg, cplx* sv, int NT, int g, cplx (&h)[P]) {
    constexpr int R = TabFmt<double, P>::row;
    constexpr int RB = 12;
#pragma unroll 1
    for (int s0 = kPM - 1; s0 >= 0; s0 -= RB) {
        double c[RB][P + 1];
#pragma unroll
        for (int b = 0; b < RB; ++b)
#pragma unroll
            for (int q = 0; q <= P; ++q)
                if (s0 - b >= 0) c[b][q] = Tg[(int64_t)((s0 - b) * R + P + 1 + q) * kLane];
        cplx yr[RB];
#pragma unroll
        for (int b = 0; b < RB; ++b)
            if (s0 - b >= 0) yr[b] = sv[sv_slot(s0 - b, g, NT)];
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            const int s = s0 - b;
            if (s >= 0) {
                cplx x = tmul(c[b][P], yr[b]);
#pragma unroll
                for (int k = 0; k < P; ++k) x -= tmul(c[b][k], h[k]);
#pragma unroll
                for (int k = P - 1; k > 0; --k) h[k] = h[k - 1];
                h[0] = x;
                if (STORE) sv[sv_slot(s, g, NT)] = x;
            }
        }
    }
}

// Complex symmetric tables.  The forward coefficient of z_{i-1-j} in row i
// is E_{j+1} of row i-1-j.  The sequential solve and the transfer kernels
// keep a register ring of, for each of the P rows above, the entries later
// rows still need: ring[m][k] = E_{k+1}(i-1-m) for k >= m; at a partition's
// first row it holds the last P rows of the partition before.
template <int P>
struct SymRing {
    static constexpr int R = P + 1;
    cplx r[P][P];
    __device__ inline void clear() {
#pragma unroll
        for (int m = 0; m < P; ++m)
#pragma unroll
            for (int k = m; k < P; ++k) r[m][k] = {0.0, 0.0};
    }
    // the last P rows of the partition whose table column starts at Tq
    __device__ inline void load_tail(const cplx* __restrict__ Tq) {
#pragma unroll
        for (int m = 0; m < P; ++m)
#pragma unroll
            for (int k = m; k < P; ++k) r[m][k] = Tq[(int64_t)((kPM - 1 - m) * R + k) * kLane];
    }
    // row i done: its entries e[0..P-1] become ring row 0
    __device__ inline void push(const cplx (&e)[P + 1]) {
#pragma unroll
        for (int m = P - 1; m > 0; --m)
#pragma unroll
            for (int k = m; k < P; ++k) r[m][k] = r[m - 1][k];
#pragma unroll
        for (int k = 0; k < P; ++k) r[0][k] = e[k];
    }
};

// Forward recurrence over the kPM rows of partition g of a complex symmetric
// table from state h (h[j] = z_{r0-1-j}); pass B (STORE) writes y = rD z.
// The coefficient of z_{i-1-j} in row i is E_{j+1} of row i-1-j: the
// coefficients are gathered from the rows above, a row of the partition
// before (column Tp) for the first rows of the partition.  Pass A starts
// from the zero state, where those entries multiply zeros; the first
// partition has none before, and any finite entry (Tp = Tg) does the same.
template <bool STORE, int P>
__device__ inline void fwd_pass_gs(const cplx* __restrict__ Tg, const cplx* __restrict__ Tp, cplx* sv, int NT, int g,
                                   cplx (&h)[P]) {
    constexpr int R = P + 1;
    constexpr int RB = 6;  // rows per block: 6(P+1) loads in flight (a wave counts at most 63)
    static_assert(kPM % RB == 0, "a pass is a whole number of blocks");
#pragma unroll 1
    for (int s0 = 0; s0 < kPM; s0 += RB) {
        cplx c[RB][R];
#pragma unroll
        for (int b = 0; b < RB; ++b) {
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const int rr = s0 + b - 1 - q;
                const cplx* a = rr >= 0 ? Tg + (int64_t)(rr * R + q) * kLane : Tp + (int64_t)((kPM + rr) * R + q) * kLane;
                c[b][q] = *a;
            }
            c[b][P] = Tg[(int64_t)((s0 + b) * R + P) * kLane];
        }
        cplx vr[RB];
#pragma unroll
        for (int b = 0; b < RB; ++b) vr[b] = sv[sv_slot(s0 + b, g, NT)];
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            cplx acc = vr[b];
#pragma unroll
            for (int j = 0; j < P; ++j) acc = tmsub(acc, c[b][j], h[j]);
#pragma unroll
            for (int j = P - 1; j > 0; --j) h[j] = h[j - 1];
            h[0] = acc;
            if (STORE) sv[sv_slot(s0 + b, g, NT)] = c[b][P] * acc;
        }
    }
}

// backward recurrence over the rows of partition g of a complex symmetric
// table from state h (h[k] = x_{r0+kPM+k}); the rows hold y = rD z
template <bool STORE, int P>
__device__ inline void bwd_pass_sym(const cplx* __restrict__ Tg, cplx* sv, int NT, int g, cplx (&h)[P]) {
    constexpr int R = P + 1;
    constexpr int RB = P == 7 ? 8 : 6;  // rows per block: 56 / 48 loads in flight (a wave counts at most 63)
    static_assert(kPM % RB == 0, "a pass is a whole number of blocks");
#pragma unroll 1
    for (int s0 = kPM - 1; s0 >= 0; s0 -= RB) {
        cplx c[RB][P];
#pragma unroll
        for (int b = 0; b < RB; ++b)
#pragma unroll
            for (int q = 0; q < P; ++q) c[b][q] = Tg[(int64_t)((s0 - b) * R + q) * kLane];
        cplx yr[RB];
#pragma unroll
        for (int b = 0; b < RB; ++b) yr[b] = sv[sv_slot(s0 - b, g, NT)];
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            const int s = s0 - b;
            cplx x = yr[b];
#pragma unroll
            for (int k = 0; k < P; ++k) x = tmsub(x, c[b][k], h[k]);
#pragma unroll
            for (int k = P - 1; k > 0; --k) h[k] = h[k - 1];
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
template <class TF, bool LDS_T, int P>
struct TrfAt {
    const TF* p;  // LDS: the direction's staged rows; global: the direction's set
    int NT, base;
    __device__ inline TF operator()(int gp, int e) const {
        if constexpr (LDS_T) return p[e * NT + gp - base];
        else return p[(int64_t)(gp >> 6) * 2 * trf_n<P>() * kLane + (int64_t)e * kLane + (gp & 63)];
    }
};

template <class TF, bool LDS_T, int P>
__device__ inline void carry(TrfAt<TF, LDS_T, P> T, const cplx* sE, int NT, int g, int dir, int NP, int base, int PK,
                             cplx (&acc)[P]) {
#pragma unroll
    for (int r = 0; r < P; ++r) acc[r] = {0.0, 0.0};
    bool started = false;
#pragma unroll 1
    for (int k = PK; k >= 1; --k) {
        const int gp = dir == 0 ? g - k : g + k;
        if (gp >= 0 && gp < NP) {
            const int lg = gp - base;
            if (!started) {
#pragma unroll
                for (int r = 0; r < P; ++r) acc[r] = sE[r * NT + lg];
                started = true;
                continue;
            }
            // the transfer matrix row by row (LDS, or global memory when
            // LDS_T is off), each row's reads in flight together
            cplx nv[P];
#pragma unroll
            for (int r = 0; r < P; ++r) {
                TF t[P];
#pragma unroll
                for (int c = 0; c < P; ++c) t[c] = T(gp, r * P + c);
                cplx a = sE[r * NT + lg];
#pragma unroll
                for (int c = 0; c < P; ++c) a = tmac(a, t[c], acc[c]);
                nv[r] = a;
            }
#pragma unroll
            for (int r = 0; r < P; ++r) acc[r] = nv[r];
        }
    }
}

// waves of a work-group on the grids that stage the transfer sets in LDS
constexpr int kStageWaves = 4;

// The carry of every staged partition spread over the kStageWaves waves of
// a staged work-group: lane ln of each wave serves partition g = base + ln,
// and wave w forms rows w and w + 4 of each Horner step's product; the state
// goes through LDS (two buffers of P x NT) with one barrier per step.  The
// same products in the same order as carry(), so the same bits.  Every wave
// calls it (it holds barriers); lanes with active false only keep their
// state.  Returns the buffer that holds the result.
template <class TF, int P>
__device__ inline int carry4(TrfAt<TF, true, P> T, const cplx* sE, cplx* sA, int NT, int g, int dir, int NP, int base,
                             int PK, int w, int ln, bool active) {
    static_assert(P <= 2 * kStageWaves, "two rows per wave");
    // the partitions on the dir side that exist: the first of them starts the chain
    const int ks = !active ? 0 : dir == 0 ? min(PK, g) : min(PK, NP - 1 - g);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int r = w + kStageWaves * i;
        if (r < P) sA[r * NT + ln] = ks > 0 ? sE[r * NT + (dir == 0 ? g - ks : g + ks) - base] : cplx{0.0, 0.0};
    }
    __syncthreads();
    int cur = 0;
#pragma unroll 1
    for (int k = PK - 1; k >= 1; --k) {
        const cplx* a = sA + cur * P * NT;
        cplx* o = sA + (1 - cur) * P * NT;
        if (k < ks) {
            const int gp = dir == 0 ? g - k : g + k;
            cplx acc[P];
#pragma unroll
            for (int c = 0; c < P; ++c) acc[c] = a[c * NT + ln];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int r = w + kStageWaves * i;
                if (r < P) {
                    TF t[P];
#pragma unroll
                    for (int c = 0; c < P; ++c) t[c] = T(gp, r * P + c);
                    cplx v = sE[r * NT + gp - base];
#pragma unroll
                    for (int c = 0; c < P; ++c) v = tmac(v, t[c], acc[c]);
                    o[r * NT + ln] = v;
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int r = w + kStageWaves * i;
                if (r < P) o[r * NT + ln] = a[r * NT + ln];
            }
        }
        __syncthreads();
        cur ^= 1;
    }
    return cur;
}

// v = B z for per-table bands (RHS_CBAND), one lane per (row, channel)
template <int P>
__global__ __launch_bounds__(256) void k_rhs(RhsArgs ra, TabMap tm, const cplx* __restrict__ zin,
                                             cplx* __restrict__ vout, int N) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int ch = blockIdx.y;
    if (i >= N) return;
    CBandRow<P> b;
    b.load(ra.cband + (int64_t)table_of(tm, ch) * (P + 1) * N, i, N);
    const cplx* z = zin + (int64_t)ch * N;
    cplx zl[P], zu[P + 1];
#pragma unroll
    for (int j = 1; j <= P; ++j) zl[j - 1] = z[max(i - j, 0)];
#pragma unroll
    for (int k = 0; k <= P; ++k) zu[k] = z[min(i + k, N - 1)];
    vout[(int64_t)ch * N + i] = b.apply(zl, zu, i);
}

// v = B z for bands shared by every channel (RHS_RBAND, RHS_SX_*): a lane
// loads the band entries of its row once and applies them to kRhsCG channels
// (2 measured fastest: 4 keeps too many loads live, 1 repeats the band loads)
constexpr int kRhsCG = 2;
template <int RHS, int P>
__global__ __launch_bounds__(256) void k_rhs_shared(RhsArgs ra, const cplx* __restrict__ zin,
                                                    cplx* __restrict__ vout, int N, int nch) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    SBandRow<RHS, P> b;
    b.load(ra, i, N);
#pragma unroll
    for (int c = 0; c < kRhsCG; ++c) {
        const int ch = blockIdx.y * kRhsCG + c;
        if (ch < nch) {
            const cplx* z = zin + (int64_t)ch * N;
            cplx zlo[P], zup[P + 1];
#pragma unroll
            for (int j = 1; j <= P; ++j) zlo[j - 1] = z[max(i - j, 0)];
#pragma unroll
            for (int k = 0; k <= P; ++k) zup[k] = z[min(i + k, N - 1)];
            const double e = SBandRow<RHS, P>::SX ? ra.eig[ch] * ra.ifdt : 0.0;
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
// the segmentation does not change a single bit of the result.  Requires
// SN + 2 PK <= 64.  ROLE is the RhsKind of the operator.  On the small grids
// that stage the transfer sets (LDS_T: one or two work-groups per CU) a
// work-group has kStageWaves waves: all of them fetch the right-hand side and
// the transfer sets and write the result, on SIMDs that would otherwise idle,
// and wave 0 runs the passes and carries.  Both directions' transfer sets are
// staged during forward pass A when they fit the LDS with the other buffers
// (p = 7, and real tables); otherwise (complex tables at p = 8) the forward
// sets are staged then and the backward sets, into the same buffer, during
// backward pass A.
//
// FUSE (LDS_T grids only): src is the state z and the work-group forms
// v = B z itself while staging (the RhsKind ROLE with the bands of ra, the
// same arithmetic as k_rhs / k_rhs_shared): the staged rows and P rows on
// either side of z go to LDS, and each staged row's band entries are read
// from L2.  Other work-groups read z rows this one owns, so src and zout may
// not overlap.  Without FUSE, src is v.
constexpr int kLdsBytes = 160 * 1024;
template <class TF, int P, bool FUSE, int ND>
constexpr int bsolve_lds() {
    return kPM * kLane * 16 + P * kLane * 16 + ND * P * P * kLane * (int)sizeof(TF) +
           (FUSE ? (kPM * kLane + 2 * P) * 16 : 0) + 2 * P * kLane * 16;
}
template <class TF, int P, bool FUSE>
constexpr bool both_trf_fit() { return bsolve_lds<TF, P, FUSE, 2>() <= kLdsBytes; }

// helper waves 1..kStageWaves-1 copy transfer entries q = Q0 .. Q0+NQ-1 of the
// staged partitions [ga, gb) into st (row q - Q0, column gl - ga)
template <class TF, int P, int Q0, int NQ>
__device__ inline void stage_trf(const TF* __restrict__ TR, TF* st, int tid, int ga, int gb) {
    constexpr int NW = kStageWaves - 1, QW = (NQ + NW - 1) / NW;
    const int ln = tid & (kLane - 1), w = tid / kLane - 1;
    const int gl = ga + ln;
    if (w >= 0 && gl < gb) {
        const TF* Tr = TR + (int64_t)(gl >> 6) * 2 * trf_n<P>() * kLane + (gl & 63) + (int64_t)Q0 * kLane;
        TF tv[QW];
#pragma unroll
        for (int j = 0; j < QW; ++j) tv[j] = Tr[(int64_t)min(w + j * NW, NQ - 1) * kLane];
#pragma unroll
        for (int j = 0; j < QW; ++j)
            if (w + j * NW < NQ) st[(w + j * NW) * kLane + ln] = tv[j];
    }
}

template <class TF, int ROLE, bool LDS_T, bool FUSE, int P>
__global__ __launch_bounds__(kLane * (LDS_T ? kStageWaves : 1)) void k_bsolve(
    const TF* __restrict__ tab, const TF* __restrict__ trf, TabMap tm, const cplx* __restrict__ src,
    cplx* __restrict__ zout, Geom geo, int SN, int PK, RhsArgs ra) {
    static_assert(LDS_T || !FUSE, "the fused right-hand side is built for the staged grids");
    constexpr int NT = kLane;
    constexpr int NTH = kLane * (LDS_T ? kStageWaves : 1);  // threads of the work-group
    constexpr int KV = kPM * NT / NTH;                       // staged rows per thread
    static_assert(kPM * NT % NTH == 0, "whole staged rows per thread");
    constexpr int KT = trf_n<P>();
    constexpr bool TWO = both_trf_fit<TF, P, FUSE>();
    static_assert(!LDS_T || bsolve_lds<TF, P, FUSE, TWO ? 2 : 1>() <= kLdsBytes, "LDS budget");
    constexpr int NZ = kPM * NT + 2 * P;  // z rows of a fused right-hand side
    __shared__ cplx sv[kPM * NT];
    __shared__ cplx sE[P * NT];
    // transfer sets of the staged partitions (LDS_T only): both directions,
    // or one direction at a time
    __shared__ TF sT[LDS_T ? (TWO ? 2 : 1) : 1][LDS_T ? KT * NT : 1];
    // z rows [r0 - P, r0 + NZ - P) (FUSE only; zero outside [0, N))
    __shared__ cplx sz[FUSE ? NZ : 1];
    // carry state of the staged partitions (LDS_T: carry4)
    __shared__ cplx sA[LDS_T ? 2 * P * NT : 1];

    int ch = blockIdx.y, seg = blockIdx.x;
    if (tm.order) {
        const int o = tm.order[blockIdx.y * gridDim.x + blockIdx.x];
        ch = o / (int)gridDim.x;
        seg = o - ch * (int)gridDim.x;
    }
    const int N = geo.N, NP = geo.P;
    const int k0 = seg * SN, k1 = min(NP, k0 + SN);  // owned partitions [k0, k1)
    const int ga = max(0, k0 - PK), gb = min(NP, k1 + PK);  // staged partitions [ga, gb)
    const int tid = threadIdx.x;
    const int g = ga + tid;  // this lane's partition (wave 0)
    const bool run = tid < kLane && g < gb;
    const int tabid = table_of(tm, ch);
    const bool cj = conj_of(tm, ch);
    const int64_t r0 = (int64_t)ga * kPM;  // first staged row
    const int nrow = (gb - ga) * kPM;
    const int nv = min(nrow, N - (int)r0);  // staged rows below N
    cplx* xo = zout + (int64_t)ch * N + r0;
    const TF* T0 = tab + (int64_t)tabid * TabFmt<TF, P>::elems(geo);
    const TF* Tg = T0 + part_col<TF, P>(g);
    // the column of the partition before (the partition itself for the first)
    const TF* Tp = T0 + part_col<TF, P>(g > 0 ? g - 1 : 0);
    constexpr bool SYM = TabFmt<TF, P>::row == P + 1;
    const int w = tid / kLane, ln = tid % kLane;  // wave, lane (carry4)

    // right-hand side rows (padding rows: 0), or the z rows it is formed
    // from (the transfer sets follow during pass A).  The row loads are
    // unconditional (indices clamped into the channel) so that all of them
    // are in flight at once.
    if constexpr (FUSE) {
        constexpr int KZ = (NZ + NTH - 1) / NTH;
        const cplx* zc = src + (int64_t)ch * N;
        const int zlo = (int)r0 - P;
        cplx zr[KZ];
#pragma unroll
        for (int k = 0; k < KZ; ++k) zr[k] = cj_if(cj, zc[min(max(zlo + tid + k * NTH, 0), N - 1)]);
#pragma unroll
        for (int k = 0; k < KZ; ++k) {
            const int j = tid + k * NTH, row = zlo + j;
            if (j < NZ) sz[j] = row >= 0 && row < N ? zr[k] : cplx{0.0, 0.0};
        }
    } else {
        const cplx* v = src + (int64_t)ch * N + r0;
        cplx vr[KV];
#pragma unroll
        for (int k = 0; k < KV; ++k) vr[k] = cj_if(cj, v[min(tid + k * NTH, nv - 1)]);
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
        // channel row r0 + i, its z_{r0+i+d} sits at sz[i + P + d]
        // a conjugate channel forms its table's right-hand side (scalar -e)
        const double e0 = ROLE == RHS_SX_SYM || ROLE == RHS_SX_ASY ? ra.eig[ch] * ra.ifdt : 0.0;
        const double e = cj ? -e0 : e0;
#pragma unroll
        for (int k = 0; k < KV; ++k) {
            const int i = tid + k * NTH;
            const int row = min((int)r0 + i, N - 1);  // rows from N on are padding: v = 0
            cplx zl[P], zu[P + 1];
#pragma unroll
            for (int j = 1; j <= P; ++j) zl[j - 1] = sz[i + P - j];
#pragma unroll
            for (int q = 0; q <= P; ++q) zu[q] = sz[i + P + q];
            cplx vv;
            if constexpr (ROLE == RHS_CBAND) {
                CBandRow<P> b;
                b.load(ra.cband + (int64_t)tabid * (P + 1) * N, row, N);
                vv = b.apply(zl, zu, row);
            } else {
                SBandRow<ROLE, P> b;
                b.load(ra, row, N);
                vv = b.apply(zl, zu, e);
            }
            if (i < nrow) sv[sv_slot(i % kPM, i / kPM, NT)] = i < nv ? vv : cplx{0.0, 0.0};
        }
        __syncthreads();
    }
    const TrfAt<TF, LDS_T, P> Tf{LDS_T ? &sT[0][0] : TR, NT, ga};
    const TrfAt<TF, LDS_T, P> Tb{LDS_T ? &sT[TWO ? 1 : 0][0] : TR + (int64_t)KT * kLane, NT, ga};

    // forward pass A on every staged partition (wave 0).  Meanwhile the
    // other waves fetch the staged partitions' transfer sets, which the
    // carries read only after the next barrier.
    if constexpr (LDS_T) stage_trf<TF, P, 0, TWO ? 2 * KT : KT>(TR, &sT[0][0], tid, ga, gb);
    cplx h[P];
    if (run) {
#pragma unroll
        for (int j = 0; j < P; ++j) h[j] = {0.0, 0.0};
        if constexpr (SYM) fwd_pass_gs<false, P>(Tg, Tp, sv, NT, tid, h);
        else fwd_pass<false, P>(Tg, sv, NT, tid, h);
#pragma unroll
        for (int j = 0; j < P; ++j) sE[j * NT + tid] = h[j];
    }
    __syncthreads();

    // forward carry and pass B on the owned partitions and the trailing halo
    const bool fw = run && g >= k0;
    if constexpr (LDS_T) {
        // every wave takes part in the carry; lane ln serves partition ga + ln
        const int gl = ga + ln;
        const int cur = carry4<TF, P>(Tf, sE, sA, NT, gl, 0, NP, ga, PK, w, ln, gl < gb && gl >= k0);
        if (fw) {
#pragma unroll
            for (int j = 0; j < P; ++j) h[j] = sA[(cur * P + j) * NT + tid];
        }
    } else if (fw) {
        carry<TF, LDS_T, P>(Tf, sE, NT, g, 0, NP, ga, PK, h);
    }
    if (fw) {
        if constexpr (SYM) fwd_pass_gs<true, P>(Tg, Tp, sv, NT, tid, h);
        else fwd_pass<true, P>(Tg, sv, NT, tid, h);
    }
    __syncthreads();

    // backward pass A on the same partitions; the backward transfer sets
    // replace the forward ones now when both do not fit at once
    if constexpr (LDS_T && !TWO) stage_trf<TF, P, KT, KT>(TR, &sT[0][0], tid, ga, gb);
    if (fw) {
#pragma unroll
        for (int j = 0; j < P; ++j) h[j] = {0.0, 0.0};
        if constexpr (SYM) bwd_pass_sym<false, P>(Tg, sv, NT, tid, h);
        else bwd_pass<false, P>(Tg, sv, NT, tid, h);
#pragma unroll
        for (int j = 0; j < P; ++j) sE[j * NT + tid] = h[j];
    }
    __syncthreads();

    // backward carry and pass B on the owned partitions
    const bool bw = fw && g < k1;
    if constexpr (LDS_T) {
        const int gl = ga + ln;
        const int cur = carry4<TF, P>(Tb, sE, sA, NT, gl, 1, NP, ga, PK, w, ln, gl < gb && gl >= k0 && gl < k1);
        if (bw) {
#pragma unroll
            for (int j = 0; j < P; ++j) h[j] = sA[(cur * P + j) * NT + tid];
        }
    } else if (bw) {
        carry<TF, LDS_T, P>(Tb, sE, NT, g, 1, NP, ga, PK, h);
    }
    if (bw) {
        if constexpr (SYM) bwd_pass_sym<true, P>(Tg, sv, NT, tid, h);
        else bwd_pass<true, P>(Tg, sv, NT, tid, h);
    }
    __syncthreads();
    const int i0 = (k0 - ga) * kPM, i1 = min((k1 - ga) * kPM, nv);
#pragma unroll
    for (int k = 0; k < KV; ++k) {
        const int i = tid + k * NTH;
        if (i >= i0 && i < i1) xo[i] = cj_if(cj, sv[sv_slot(i % kPM, i / kPM, NT)]);
    }
}

// Sequential solve x = U^-1 L^-1 v over all N rows, one lane per channel:
// the fallback for carries deeper than kPKMax partitions (exact for any
// operator the tables hold).  Table rows are fetched in blocks of RB rows,
// every load of a block issued before its first use; zout holds y between
// the two passes.
template <int P>
__global__ __launch_bounds__(64) void k_seqsolve_c(const cplx* __restrict__ tab, TabMap tm,
                                                   const cplx* __restrict__ vin, cplx* __restrict__ zout, Geom geo,
                                                   int nch) {
    constexpr int R = P + 1, RB = 6;
    const int ch = blockIdx.x * 64 + threadIdx.x;
    if (ch >= nch) return;
    const int N = geo.N;
    const cplx* T0 = tab + (int64_t)table_of(tm, ch) * geo.ctab_elems;
    const bool cj = conj_of(tm, ch);
    const cplx* v = vin + (int64_t)ch * N;
    cplx* x = zout + (int64_t)ch * N;
    SymRing<P> ring;
    ring.clear();
    cplx h[P];
#pragma unroll
    for (int j = 0; j < P; ++j) h[j] = {0.0, 0.0};
    for (int i0 = 0; i0 < N; i0 += RB) {
        cplx c[RB][R], vr[RB];
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            const int i = min(i0 + b, N - 1);
            const cplx* Tq = T0 + part_col<cplx, P>(i / kPM) + (int64_t)(i % kPM) * R * kLane;
#pragma unroll
            for (int q = 0; q < R; ++q) c[b][q] = Tq[(int64_t)q * kLane];
            vr[b] = cj_if(cj, v[i]);
        }
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            if (i0 + b < N) {
                cplx acc = vr[b];
#pragma unroll
                for (int j = 0; j < P; ++j) acc = tmsub(acc, ring.r[j][j], h[j]);
#pragma unroll
                for (int j = P - 1; j > 0; --j) h[j] = h[j - 1];
                h[0] = acc;
                x[i0 + b] = c[b][P] * acc;
                ring.push(c[b]);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < P; ++j) h[j] = {0.0, 0.0};
    for (int i1 = N - 1; i1 >= 0; i1 -= RB) {
        cplx c[RB][P], yr[RB];
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            const int i = max(i1 - b, 0);
            const cplx* Tq = T0 + part_col<cplx, P>(i / kPM) + (int64_t)(i % kPM) * R * kLane;
#pragma unroll
            for (int q = 0; q < P; ++q) c[b][q] = Tq[(int64_t)q * kLane];
            yr[b] = x[i];
        }
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            if (i1 - b >= 0) {
                cplx xv = yr[b];
#pragma unroll
                for (int k = 0; k < P; ++k) xv = tmsub(xv, c[b][k], h[k]);
#pragma unroll
                for (int k = P - 1; k > 0; --k) h[k] = h[k - 1];
                h[0] = xv;
                x[i1 - b] = cj_if(cj, xv);
            }
        }
    }
}

template <int P>
__global__ __launch_bounds__(64) void k_seqsolve_r(const double* __restrict__ tab, TabMap tm,
                                                   const cplx* __restrict__ vin, cplx* __restrict__ zout, Geom geo,
                                                   int nch) {
    constexpr int R = 2 * (P + 1), RB = 6;
    const int ch = blockIdx.x * 64 + threadIdx.x;
    if (ch >= nch) return;
    const int N = geo.N;
    const double* T0 = tab + (int64_t)table_of(tm, ch) * geo.tab_elems;
    const bool cj = conj_of(tm, ch);  // tables shared by conjugation are complex only: false here
    const cplx* v = vin + (int64_t)ch * N;
    cplx* x = zout + (int64_t)ch * N;
    cplx h[P];
#pragma unroll
    for (int j = 0; j < P; ++j) h[j] = {0.0, 0.0};
    for (int i0 = 0; i0 < N; i0 += RB) {
        double c[RB][P + 1];
        cplx vr[RB];
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            const int i = min(i0 + b, N - 1);
            const double* Tq = T0 + part_col<double, P>(i / kPM) + (int64_t)(i % kPM) * R * kLane;
#pragma unroll
            for (int q = 0; q <= P; ++q) c[b][q] = Tq[(int64_t)q * kLane];
            vr[b] = cj_if(cj, v[i]);
        }
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            if (i0 + b < N) {
                cplx acc = vr[b];
#pragma unroll
                for (int j = 0; j < P; ++j) acc -= tmul(c[b][j], h[j]);
                const cplx y = tmul(c[b][P], acc);
#pragma unroll
                for (int j = P - 1; j > 0; --j) h[j] = h[j - 1];
                h[0] = y;
                x[i0 + b] = y;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < P; ++j) h[j] = {0.0, 0.0};
    for (int i1 = N - 1; i1 >= 0; i1 -= RB) {
        double c[RB][P + 1];
        cplx yr[RB];
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            const int i = max(i1 - b, 0);
            const double* Tq = T0 + part_col<double, P>(i / kPM) + (int64_t)(i % kPM) * R * kLane;
#pragma unroll
            for (int q = 0; q <= P; ++q) c[b][q] = Tq[(int64_t)(P + 1 + q) * kLane];
            yr[b] = x[i];
        }
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            if (i1 - b >= 0) {
                cplx xv = tmul(c[b][P], yr[b]);
#pragma unroll
                for (int k = 0; k < P; ++k) xv -= tmul(c[b][k], h[k]);
#pragma unroll
                for (int k = P - 1; k > 0; --k) h[k] = h[k - 1];
                h[0] = xv;
                x[i1 - b] = cj_if(cj, xv);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// transfer sets of real LU tables: one lane per (table, partition); the
// P unit responses of each direction are propagated together so every
// table row is read once
template <int P>
__global__ __launch_bounds__(kLane) void k_transfer(const double* __restrict__ tab, double* __restrict__ trf,
                                                    Geom geo) {
    constexpr int R = TabFmt<double, P>::row, KT = trf_n<P>();
    const int g = blockIdx.y * kLane + threadIdx.x;
    const int t = blockIdx.x;
    if (g >= geo.P) return;
    const double* Tg = tab + (int64_t)t * geo.tab_elems + part_col<double, P>(g);
    double* out = trf + (int64_t)t * geo.trf_elems + (int64_t)((g >> 6) * 2) * KT * kLane + (g & 63);
    for (int dir = 0; dir < 2; ++dir) {
        double H[P][P];  // H[q][col]: state entry q of the response to unit vector col
#pragma unroll
        for (int q = 0; q < P; ++q)
#pragma unroll
            for (int col = 0; col < P; ++col) H[q][col] = q == col ? 1.0 : 0.0;
        for (int ss = 0; ss < kPM; ++ss) {
            const int s = dir == 0 ? ss : kPM - 1 - ss;
            const double* c = Tg + (int64_t)(s * R + (dir == 0 ? 0 : P + 1)) * kLane;
            double f[P + 1];
#pragma unroll
            for (int q = 0; q <= P; ++q) f[q] = c[(int64_t)q * kLane];
            double y[P];
#pragma unroll
            for (int col = 0; col < P; ++col) {
                double acc = 0.0;
#pragma unroll
                for (int q = 0; q < P; ++q) acc = acc - f[q] * H[q][col];
                y[col] = dir == 0 ? f[P] * acc : acc;
            }
#pragma unroll
            for (int q = P - 1; q > 0; --q)
#pragma unroll
                for (int col = 0; col < P; ++col) H[q][col] = H[q - 1][col];
#pragma unroll
            for (int col = 0; col < P; ++col) H[0][col] = y[col];
        }
        double* o = out + (int64_t)dir * KT * kLane;
#pragma unroll
        for (int q = 0; q < P; ++q)
#pragma unroll
            for (int col = 0; col < P; ++col) o[(int64_t)(q * P + col) * kLane] = H[q][col];
    }
}

// transfer sets of complex symmetric tables (the z recurrence forward, the x
// recurrence backward); the forward responses start from the coefficients of
// the previous partition's last rows, as the solve's pass B does.  The P unit
// responses are propagated NC at a time (P = 8: two sweeps of 4, which keeps
// the responses in registers), each sweep reading the partition's rows.
template <int P>
__global__ __launch_bounds__(kLane) void k_transfer_sym(const cplx* __restrict__ tab, cplx* __restrict__ trf,
                                                        Geom geo, int64_t stride) {
    constexpr int R = P + 1, KT = trf_n<P>();
    constexpr int NC = P == 7 ? 7 : 4;
    static_assert(P % NC == 0, "whole column groups");
    const int g = blockIdx.y * kLane + threadIdx.x;
    const int64_t t = (int64_t)blockIdx.z * stride + blockIdx.x;  // table index
    if (g >= geo.P) return;
    const cplx* T0 = tab + t * geo.ctab_elems;
    const cplx* Tg = T0 + part_col<cplx, P>(g);
    cplx* out = trf + (int64_t)t * geo.trf_elems + (int64_t)((g >> 6) * 2) * KT * kLane + (g & 63);
#pragma unroll 1
    for (int c0 = 0; c0 < P; c0 += NC) {
        cplx H[P][NC];  // H[q][c]: state entry q of the response to unit vector c0 + c
        // forward
#pragma unroll
        for (int q = 0; q < P; ++q)
#pragma unroll
            for (int c = 0; c < NC; ++c) H[q][c] = q == c0 + c ? cplx{1.0, 0.0} : cplx{0.0, 0.0};
        SymRing<P> ring;
        if (g > 0) ring.load_tail(T0 + part_col<cplx, P>(g - 1));
        else ring.clear();
        for (int s = 0; s < kPM; ++s) {
            cplx e[R];
#pragma unroll
            for (int q = 0; q < R; ++q) e[q] = Tg[(int64_t)(s * R + q) * kLane];
            cplx y[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                cplx acc = {0.0, 0.0};
#pragma unroll
                for (int q = 0; q < P; ++q) acc = acc - ring.r[q][q] * H[q][c];
                y[c] = acc;
            }
#pragma unroll
            for (int q = P - 1; q > 0; --q)
#pragma unroll
                for (int c = 0; c < NC; ++c) H[q][c] = H[q - 1][c];
#pragma unroll
            for (int c = 0; c < NC; ++c) H[0][c] = y[c];
            ring.push(e);
        }
#pragma unroll
        for (int q = 0; q < P; ++q)
#pragma unroll
            for (int c = 0; c < NC; ++c) out[(int64_t)(q * P + c0 + c) * kLane] = H[q][c];
        // backward
#pragma unroll
        for (int q = 0; q < P; ++q)
#pragma unroll
            for (int c = 0; c < NC; ++c) H[q][c] = q == c0 + c ? cplx{1.0, 0.0} : cplx{0.0, 0.0};
        for (int ss = 0; ss < kPM; ++ss) {
            const int s = kPM - 1 - ss;
            cplx e[P];
#pragma unroll
            for (int q = 0; q < P; ++q) e[q] = Tg[(int64_t)(s * R + q) * kLane];
            cplx y[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                cplx acc = {0.0, 0.0};
#pragma unroll
                for (int q = 0; q < P; ++q) acc = acc - e[q] * H[q][c];
                y[c] = acc;
            }
#pragma unroll
            for (int q = P - 1; q > 0; --q)
#pragma unroll
                for (int c = 0; c < NC; ++c) H[q][c] = H[q - 1][c];
#pragma unroll
            for (int c = 0; c < NC; ++c) H[0][c] = y[c];
        }
        cplx* o = out + (int64_t)KT * kLane;
#pragma unroll
        for (int q = 0; q < P; ++q)
#pragma unroll
            for (int c = 0; c < NC; ++c) o[(int64_t)(q * P + c0 + c) * kLane] = H[q][c];
    }
}

// ---------------------------------------------------------------------------
// complex symmetric tables from the lu_zsb factor (ed.h:208-289):
// f[i*ne] = D_i, f[i*ne+j] = L(i+j,i); one lane per (table, partition)
template <int P>
__global__ __launch_bounds__(kLane) void k_tab_ldlt(const cplx* __restrict__ lower, cplx* __restrict__ tab, Geom geo) {
    constexpr int R = P + 1, ne = P + 1;
    const int g = blockIdx.x * kLane + threadIdx.x;
    const int l = blockIdx.y;
    if (g >= geo.P) return;
    const int N = geo.N;
    const cplx* f = lower + (int64_t)l * N * ne;
    cplx* Tg = tab + (int64_t)l * geo.ctab_elems + part_col<cplx, P>(g);
    for (int s = 0; s < kPM; ++s) {
        const int i = g * kPM + s;
        cplx o[R];
        if (i >= N) {  // identity padding row
#pragma unroll
            for (int q = 0; q < P; ++q) o[q] = {0.0, 0.0};
            o[P] = {1.0, 0.0};
        } else {
#pragma unroll
            for (int k = 1; k <= P; ++k) o[k - 1] = i + k < N ? f[(int64_t)i * ne + k] : cplx{0.0, 0.0};
            o[P] = crecip(f[(int64_t)i * ne]);
        }
#pragma unroll
        for (int q = 0; q < R; ++q) Tg[(int64_t)(s * R + q) * kLane] = o[q];
    }
}

// tables from chol(S) (ed.h:182-194): c[i*ne+j] = L(i+j,i)
template <int P>
__global__ __launch_bounds__(kLane) void k_tab_chol(const double* __restrict__ c, double* __restrict__ tab, Geom geo) {
    constexpr int R = TabFmt<double, P>::row, ne = P + 1;
    const int g = blockIdx.x * kLane + threadIdx.x;
    if (g >= geo.P) return;
    const int N = geo.N;
    double* Tg = tab + part_col<double, P>(g);
    for (int s = 0; s < kPM; ++s) {
        const int i = g * kPM + s;
        double o[R];
        if (i >= N) {
#pragma unroll
            for (int q = 0; q < R; ++q) o[q] = 0.0;
            o[P] = 1.0;
            o[R - 1] = 1.0;
        } else {
            const double d = 1.0 / c[(int64_t)i * ne];
#pragma unroll
            for (int j = 1; j <= P; ++j) o[j - 1] = i - j >= 0 ? c[(int64_t)(i - j) * ne + j] : 0.0;
            o[P] = d;
#pragma unroll
            for (int k = 1; k <= P; ++k) o[P + k] = i + k < N ? c[(int64_t)i * ne + k] * d : 0.0;
            o[R - 1] = d;
        }
#pragma unroll
        for (int q = 0; q < R; ++q) Tg[(int64_t)(s * R + q) * kLane] = o[q];
    }
}

__global__ __launch_bounds__(256) void k_band_pair(const double* __restrict__ S, const double* __restrict__ X,
                                                   cplx* __restrict__ dst, int64_t total) {
    const int64_t id = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (id < total) dst[id] = cplx{S[id], X[id]};
}

template <class T>
__global__ __launch_bounds__(256) void k_band_diag(const T* __restrict__ src, T* __restrict__ dst, int64_t total,
                                                   int N, int ne) {
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= total) return;
    const int64_t b = id / ((int64_t)ne * N);
    const int64_t r = id % ((int64_t)ne * N);
    const int j = (int)(r / N), i = (int)(r % N);
    dst[id] = src[(b * N + i) * ne + j];
}

// ---------------------------------------------------------------------------
// reference-order LU of b = S + ifet X (propagate.hpp:33-86), one lane per
// (time point, channel, partition).  A lane restarts the elimination
// `restart` rows above its partition (as if the matrix began there) and
// writes the table rows of its own partition.  The last P U rows live in a
// register ring of P + 1 slots (row i in slot (i - start) mod (P+1)); the row
// loop is unrolled by the ring length so that every slot index is a
// constant.  The lower panel of row i (S and X of rows i-P..i-1, the rows
// just eliminated) comes from a second register ring at p = 7 and is read
// again from the band arrays, where it is in cache, at p = 8.
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

template <int SYM, int P>
__global__ __launch_bounds__(kLane) void k_factor(const double* __restrict__ S, const double* __restrict__ X,
                                                  const double* __restrict__ eig, int64_t eig_tstride,
                                                  const double* __restrict__ ifdt, int nch, Geom geo, int restart,
                                                  typename FactorTypes<SYM>::T* __restrict__ tab, int64_t tstride) {
    using T = typename FactorTypes<SYM>::T;
    constexpr int nd = P, ne = P + 1, RS = P + 1;
    const int g = blockIdx.x * kLane + threadIdx.x;
    const int ch = blockIdx.y, t = blockIdx.z;
    if (g >= geo.P) return;
    const int N = geo.N;
    const double e = eig[(int64_t)t * eig_tstride + ch] * ifdt[t];
    // SYM 0 (S + i e X, X symmetric) writes complex symmetric rows [u(i,0..P-1), 1/pivot]:
    // for a symmetric matrix the unit upper factor's rows are the columns of
    // L in A = L D L^T and the pivots are D; SYM 1 writes real LU rows
    constexpr int ROW = TabFmt<T, P>::row;
    T* Tg = tab + ((int64_t)t * tstride + ch) * TabFmt<T, P>::elems(geo) + part_col<T, P>(g);
    const int r0 = g * kPM;
    const int start = r0 > restart ? r0 - restart : 0;
    const int end = min(N, r0 + kPM);
    T u[RS][nd];
    // p = 7: entries 1..P of the upper bands of the last RS rows (the lower
    // panel of later rows) stay in a register ring beside u; p = 8 reads them
    // again from the band arrays (the rings would not fit the registers)
    constexpr bool WRING = P == 7;
    double ws[WRING ? RS : 1][nd], wx[WRING ? RS : 1][nd];
#pragma unroll
    for (int s = 0; s < RS; ++s)
#pragma unroll
        for (int k = 0; k < nd; ++k) {
            u[s][k] = tzero(T{});
            if constexpr (WRING) {
                ws[s][k] = 0.0;
                wx[s][k] = 0.0;
            }
        }
    for (int i0 = start; i0 < end; i0 += RS) {
#pragma unroll
        for (int s = 0; s < RS; ++s) {
            const int i = i0 + s;
            if (i < end) {
                const int avail = i - start;  // rows above i in the restarted matrix
                double sr[ne], xr[ne], sl[nd], xl[nd];
#pragma unroll
                for (int j = 0; j < ne; ++j) {
                    sr[j] = S[(int64_t)i * ne + j];
                    xr[j] = X[(int64_t)i * ne + j];
                }
                // entry j of the upper bands of row i-j (clamped index: rows
                // above the restart are zeroed below)
#pragma unroll
                for (int j = 1; j <= nd; ++j) {
                    if constexpr (WRING) {
                        const int slot = (s - j + RS) % RS;  // row i-j
                        sl[j - 1] = ws[slot][j - 1];
                        xl[j - 1] = wx[slot][j - 1];
                    } else {
                        const int64_t k = (int64_t)max(i - j, 0) * ne + j;
                        sl[j - 1] = S[k];
                        xl[j - 1] = X[k];
                    }
                }
                if constexpr (WRING) {
#pragma unroll
                    for (int j = 1; j < ne; ++j) {
                        ws[s][j - 1] = sr[j];
                        wx[s][j - 1] = xr[j];
                    }
                }
                T b[2 * nd + 1];
                // lower panel: column i-j (propagate.hpp:35-53)
#pragma unroll
                for (int j = 1; j <= nd; ++j) {
                    if (j > avail) {
                        b[nd - j] = tzero(T{});
                    } else {
                        const double sv = sl[j - 1], rv = xl[j - 1];
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
                const int l = avail < nd ? nd - avail : 0;
                // elimination with the previous U rows (propagate.hpp:67-71)
#pragma unroll
                for (int j = 0; j < nd; ++j) {
                    if (j >= l) {
                        const int slot = (s + j + 1) % RS;  // row i+j-P
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
                        o[(int64_t)(ROW - 1) * kLane] = tone(T{});
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
template <int P>
static void rhs_launch(hipStream_t s, int rhs, const RhsArgs& ra, TabMap tm, const cplx* zin, cplx* vout, int nch,
                       Geom geo) {
    dim3 rgrid((geo.N + 255) / 256, nch);
    dim3 sgrid((geo.N + 255) / 256, (nch + kRhsCG - 1) / kRhsCG);
    switch (rhs) {
        case RHS_CBAND: hipLaunchKernelGGL(k_rhs<P>, rgrid, dim3(256), 0, s, ra, tm, zin, vout, geo.N); break;
        case RHS_RBAND:
            hipLaunchKernelGGL((k_rhs_shared<RHS_RBAND, P>), sgrid, dim3(256), 0, s, ra, zin, vout, geo.N, nch);
            break;
        case RHS_SX_SYM:
            hipLaunchKernelGGL((k_rhs_shared<RHS_SX_SYM, P>), sgrid, dim3(256), 0, s, ra, zin, vout, geo.N, nch);
            break;
        case RHS_SX_ASY:
            hipLaunchKernelGGL((k_rhs_shared<RHS_SX_ASY, P>), sgrid, dim3(256), 0, s, ra, zin, vout, geo.N, nch);
            break;
        default: break;
    }
}

void launch_rhs(hipStream_t s, int rhs, const RhsArgs& ra, TabMap tm, const cplx* zin, cplx* vout, int nch, Geom geo) {
    if (geo.hw == 8) rhs_launch<8>(s, rhs, ra, tm, zin, vout, nch, geo);
    else rhs_launch<7>(s, rhs, ra, tm, zin, vout, nch, geo);
}

// Largest grid whose solves stage the transfer sets in LDS (geo.stage_grid,
// default kStageGrid), with four waves per work-group and the right-hand side
// formed in the launch.  One such work-group fits a CU (107-157 KB of LDS),
// so up to 512 run in at most two rounds over the 256 CUs; larger grids read
// the transfer sets from L2 and keep several single-wave work-groups per CU
// (64 batched shots in buckets of 468-720 channels: 7.59e8 coefficient
// updates/s with 512, 7.01e8 with 1024, 6.81e8 with 4096).

int solve_segments(const Geom& geo, int pk) {
    if (pk > kPKMax) return 0;  // the sequential solve: no segments
    const int smax = seg_max(pk);
    const int nseg0 = std::min(geo.P, (geo.P + smax - 1) / smax + std::max(0, geo.seg_add));
    const int sn = (geo.P + nseg0 - 1) / nseg0;
    return (geo.P + sn - 1) / sn;
}

void solve_order(const int* tab, int nch, int nseg, int* order) {
    // (table, segment, channel) ascending, dealt to the XCDs in contiguous
    // chunks; XCD x's k-th work-group has linear id x + kXcds k
    std::vector<int> e((size_t)nch * nseg);
    for (int ch = 0; ch < nch; ++ch)
        for (int sg = 0; sg < nseg; ++sg) e[(size_t)ch * nseg + sg] = ch * nseg + sg;
    std::stable_sort(e.begin(), e.end(), [&](int a, int b) {
        const int ta = tab[a / nseg], tb = tab[b / nseg];
        if (ta != tb) return ta < tb;
        return a % nseg < b % nseg;
    });
    const int T = nch * nseg, base = T / kXcds, extra = T % kXcds;
    int j = 0;
    for (int x = 0; x < kXcds; ++x) {
        const int n = base + (x < extra ? 1 : 0);
        for (int k = 0; k < n; ++k) order[x + kXcds * k] = e[j++];
    }
}

template <class TF, int ROLE, int P>
static void bsolve(hipStream_t s, const TF* tab, const TF* trf, TabMap tm, const cplx* zin, cplx* vbuf,
                   cplx* zout, int nch, Geom geo, const RhsArgs* ra, int pk) {
    const int PK = pk > 0 ? pk : (ROLE == RHS_SX_ASY ? kPKD : kPKS);
    if (PK > kPKMax) {
        // the carry would reach further than a work-group holds: sequential solve
        if (ra) {
            rhs_launch<P>(s, ROLE, *ra, tm, zin, vbuf, nch, geo);
            zin = vbuf;
        }
        if constexpr (sizeof(TF) == sizeof(cplx))
            hipLaunchKernelGGL(k_seqsolve_c<P>, dim3((nch + 63) / 64), dim3(64), 0, s, tab, tm, zin, zout, geo, nch);
        else
            hipLaunchKernelGGL(k_seqsolve_r<P>, dim3((nch + 63) / 64), dim3(64), 0, s, tab, tm, zin, zout, geo, nch);
        return;
    }
    const int smax = seg_max(PK);
    // the fewest segments that hold the partitions, plus geo.seg_add, none empty
    const int nseg0 = std::min(geo.P, (geo.P + smax - 1) / smax + std::max(0, geo.seg_add));
    const int sn = (geo.P + nseg0 - 1) / nseg0;  // <= smax, so sn + 2 PK <= 64
    const int nseg = (geo.P + sn - 1) / sn;
    if (tm.order && nseg != tm.order_nseg) tm.order = nullptr;  // built for another launch shape
    // staged grids (geo.stage_grid): transfer sets in LDS, four waves, and the
    // right-hand side formed in the launch when z and x do not overlap
    const bool staged = (int64_t)nseg * nch <= geo.stage_grid;
    const dim3 grid(nseg, nch);
    if (ra) {
        // right-hand side formed in the solve (staged grids, disjoint z and x)
        const uintptr_t a0 = (uintptr_t)zin, b0 = (uintptr_t)zout;
        const uintptr_t len = sizeof(cplx) * (uintptr_t)nch * geo.N;
        if (staged && (a0 + len <= b0 || b0 + len <= a0)) {
            hipLaunchKernelGGL((k_bsolve<TF, ROLE, true, true, P>), grid, dim3(kLane * kStageWaves), 0, s, tab, trf,
                               tm, zin, zout, geo, sn, PK, *ra);
            return;
        }
        rhs_launch<P>(s, ROLE, *ra, tm, zin, vbuf, nch, geo);
        zin = vbuf;
    }
    if (staged)
        hipLaunchKernelGGL((k_bsolve<TF, ROLE, true, false, P>), grid, dim3(kLane * kStageWaves), 0, s, tab, trf, tm,
                           zin, zout, geo, sn, PK, RhsArgs{});
    else
        hipLaunchKernelGGL((k_bsolve<TF, ROLE, false, false, P>), grid, dim3(kLane), 0, s, tab, trf, tm, zin, zout,
                           geo, sn, PK, RhsArgs{});
}

template <class TF, int ROLE>
static void bsolve_w(hipStream_t s, const TF* tab, const TF* trf, TabMap tm, const cplx* zin, cplx* vbuf, cplx* zout,
                     int nch, Geom geo, const RhsArgs* ra, int pk) {
    if (geo.hw == 8) bsolve<TF, ROLE, 8>(s, tab, trf, tm, zin, vbuf, zout, nch, geo, ra, pk);
    else bsolve<TF, ROLE, 7>(s, tab, trf, tm, zin, vbuf, zout, nch, geo, ra, pk);
}

void launch_bsolve_c(hipStream_t s, int role, const cplx* tab, const cplx* trf, TabMap tm, const cplx* vin,
                     cplx* zout, int nch, Geom geo, int pk) {
    if (role == RHS_CBAND) bsolve_w<cplx, RHS_CBAND>(s, tab, trf, tm, vin, nullptr, zout, nch, geo, nullptr, pk);
    else bsolve_w<cplx, RHS_SX_SYM>(s, tab, trf, tm, vin, nullptr, zout, nch, geo, nullptr, pk);
}
void launch_bsolve_r(hipStream_t s, int role, const double* tab, const double* trf, TabMap tm, const cplx* vin,
                     cplx* zout, int nch, Geom geo, int pk) {
    if (role == RHS_RBAND) bsolve_w<double, RHS_RBAND>(s, tab, trf, tm, vin, nullptr, zout, nch, geo, nullptr, pk);
    else bsolve_w<double, RHS_SX_ASY>(s, tab, trf, tm, vin, nullptr, zout, nch, geo, nullptr, pk);
}
void launch_rsolve_c(hipStream_t s, int role, const RhsArgs& ra, const cplx* tab, const cplx* trf, TabMap tm,
                     const cplx* zin, cplx* vbuf, cplx* zout, int nch, Geom geo, int pk) {
    if (role == RHS_CBAND) bsolve_w<cplx, RHS_CBAND>(s, tab, trf, tm, zin, vbuf, zout, nch, geo, &ra, pk);
    else bsolve_w<cplx, RHS_SX_SYM>(s, tab, trf, tm, zin, vbuf, zout, nch, geo, &ra, pk);
}
void launch_rsolve_r(hipStream_t s, int role, const RhsArgs& ra, const double* tab, const double* trf, TabMap tm,
                     const cplx* zin, cplx* vbuf, cplx* zout, int nch, Geom geo, int pk) {
    if (role == RHS_RBAND) bsolve_w<double, RHS_RBAND>(s, tab, trf, tm, zin, vbuf, zout, nch, geo, &ra, pk);
    else bsolve_w<double, RHS_SX_ASY>(s, tab, trf, tm, zin, vbuf, zout, nch, geo, &ra, pk);
}

void launch_transfer_c(hipStream_t s, const cplx* tab, cplx* trf, int ntab, Geom geo) {
    launch_transfer_c(s, tab, trf, ntab, 1, 0, geo);
}
void launch_transfer_c(hipStream_t s, const cplx* tab, cplx* trf, int per_t, int nt, int64_t stride, Geom geo) {
    const dim3 grid(per_t, geo.NG, nt);
    if (geo.hw == 8) hipLaunchKernelGGL(k_transfer_sym<8>, grid, dim3(kLane), 0, s, tab, trf, geo, stride);
    else hipLaunchKernelGGL(k_transfer_sym<7>, grid, dim3(kLane), 0, s, tab, trf, geo, stride);
}
void launch_transfer_r(hipStream_t s, const double* tab, double* trf, int ntab, Geom geo) {
    if (geo.hw == 8) hipLaunchKernelGGL(k_transfer<8>, dim3(ntab, geo.NG), dim3(kLane), 0, s, tab, trf, geo);
    else hipLaunchKernelGGL(k_transfer<7>, dim3(ntab, geo.NG), dim3(kLane), 0, s, tab, trf, geo);
}

void launch_tab_ldlt(hipStream_t s, const cplx* lower, cplx* tab, int nl, Geom geo) {
    if (geo.hw == 8) hipLaunchKernelGGL(k_tab_ldlt<8>, dim3(geo.NG, nl), dim3(kLane), 0, s, lower, tab, geo);
    else hipLaunchKernelGGL(k_tab_ldlt<7>, dim3(geo.NG, nl), dim3(kLane), 0, s, lower, tab, geo);
}
void launch_tab_chol(hipStream_t s, const double* chol, double* tab, Geom geo) {
    if (geo.hw == 8) hipLaunchKernelGGL(k_tab_chol<8>, dim3(geo.NG), dim3(kLane), 0, s, chol, tab, geo);
    else hipLaunchKernelGGL(k_tab_chol<7>, dim3(geo.NG), dim3(kLane), 0, s, chol, tab, geo);
}

void launch_band_diag_r(hipStream_t s, const double* src, double* dst, int n, int N, int p) {
    const int64_t total = (int64_t)n * (p + 1) * N;
    hipLaunchKernelGGL(k_band_diag<double>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, src, dst, total, N,
                       p + 1);
}
void launch_band_pair(hipStream_t s, const double* S, const double* X, cplx* dst, int N, int p) {
    const int64_t total = (int64_t)(p + 1) * N;
    hipLaunchKernelGGL(k_band_pair, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, S, X, dst, total);
}
void launch_band_diag_c(hipStream_t s, const cplx* src, cplx* dst, int n, int N, int p) {
    const int64_t total = (int64_t)n * (p + 1) * N;
    hipLaunchKernelGGL(k_band_diag<cplx>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, src, dst, total, N,
                       p + 1);
}

void launch_factor_sym(hipStream_t s, const double* S, const double* X, const double* eig, const double* ifdt, int nt,
                       int nch, Geom geo, cplx* tab, int64_t eig_tstride, int restart, int64_t tab_tstride) {
    const int rs = restart > 0 ? restart : kRestart;
    const int64_t ts = tab_tstride > 0 ? tab_tstride : nch;
    if (geo.hw == 8)
        hipLaunchKernelGGL((k_factor<0, 8>), dim3(geo.NG, nch, nt), dim3(kLane), 0, s, S, X, eig, eig_tstride, ifdt,
                           nch, geo, rs, tab, ts);
    else
        hipLaunchKernelGGL((k_factor<0, 7>), dim3(geo.NG, nch, nt), dim3(kLane), 0, s, S, X, eig, eig_tstride, ifdt,
                           nch, geo, rs, tab, ts);
}
void launch_factor_asy(hipStream_t s, const double* S, const double* X, const double* eig, const double* ifdt, int nt,
                       int nch, Geom geo, double* tab, int64_t eig_tstride, int restart) {
    const int rs = restart > 0 ? restart : kRestart;
    if (geo.hw == 8)
        hipLaunchKernelGGL((k_factor<1, 8>), dim3(geo.NG, nch, nt), dim3(kLane), 0, s, S, X, eig, eig_tstride, ifdt,
                           nch, geo, rs, tab, (int64_t)nch);
    else
        hipLaunchKernelGGL((k_factor<1, 7>), dim3(geo.NG, nch, nt), dim3(kLane), 0, s, S, X, eig, eig_tstride, ifdt,
                           nch, geo, rs, tab, (int64_t)nch);
}

// ---------------------------------------------------------------------------
// host: reach of the recurrences (qpc_solve.h)
namespace {
struct hc {
    double re, im;
};
inline hc hmul(hc a, hc b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline hc hsub(hc a, hc b) { return {a.re - b.re, a.im - b.im}; }
inline hc hdiv(hc a, hc b) {
    const double d = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
inline double habs(hc a) { return std::sqrt(a.re * a.re + a.im * a.im); }

// rows from a unit source until the response of L^-1 (fwd) or U^-1 (bwd)
// stays below tol times its peak, in the window [w0, w0 + W) of A
Reach window_reach(const double* S, const double* X, int p, double e, int asym, int w0, int W, double tol) {
    const int ne = p + 1;
    // A(i, j) in the window, banded: U rows (unit diagonal after scaling) and L multipliers
    auto upper = [&](int i, int k) -> hc {  // A(i, i+k), k >= 0
        const double sv = S[(size_t)(w0 + i) * ne + k], xv = X[(size_t)(w0 + i) * ne + k];
        return asym ? hc{sv + e * xv, 0.0} : hc{sv, e * xv};
    };
    auto lower = [&](int i, int k) -> hc {  // A(i+k, i), k >= 1
        const double sv = S[(size_t)(w0 + i) * ne + k], xv = X[(size_t)(w0 + i) * ne + k];
        return asym ? hc{sv - e * xv, 0.0} : hc{sv, e * xv};
    };
    // dense-banded LU without pivoting: U[i][k] = U(i, i+k), Lm[i][k] = L(i+k, i)
    std::vector<hc> U((size_t)W * ne), Lm((size_t)W * ne);
    std::vector<hc> row(2 * p + 1);
    for (int i = 0; i < W; ++i) {
        // row i of A from column i-p to i+p, minus the eliminated rows above
        for (int c = -p; c <= p; ++c) {
            const int j = i + c;
            hc v = {0.0, 0.0};
            if (j >= 0 && j < W) v = c >= 0 ? (c <= p ? upper(i, c) : hc{0, 0}) : lower(j, -c);
            row[c + p] = v;
        }
        for (int k = std::max(0, i - p); k < i; ++k) {
            // eliminate column k with U row k: l = row[k] / U(k,k)
            const hc l = hdiv(row[k - i + p], U[(size_t)k * ne]);
            Lm[(size_t)k * ne + (i - k)] = l;
            for (int c = 0; c <= p && k + c < W; ++c) {
                const int col = k + c - i + p;
                if (col > k - i + p && col <= 2 * p) row[col] = hsub(row[col], hmul(l, U[(size_t)k * ne + c]));
            }
            row[k - i + p] = {0.0, 0.0};
        }
        for (int c = 0; c <= p; ++c) U[(size_t)i * ne + c] = (i + c < W) ? row[c + p] : hc{0.0, 0.0};
    }
    const int src = std::min(60, W / 4);
    Reach r{0, 0, 0};
    // restarted elimination: the same LU begun afresh at row src (rows above
    // it dropped, as k_factor restarts above each partition); rst is the
    // distance after which its U rows stay within 1e-14 of the row's largest
    // entry from the full elimination's (the difference falls about three
    // decades per 24 rows to a few ulps, where rounding keeps it until the
    // two eliminations become bit-identical some 110-130 rows on)
    {
        std::vector<hc> Ur((size_t)W * ne);
        int last = src;
        for (int i = src; i < W; ++i) {
            for (int c = -p; c <= p; ++c) {
                const int j = i + c;
                hc v = {0.0, 0.0};
                if (j >= src && j < W) v = c >= 0 ? upper(i, c) : lower(j, -c);
                row[c + p] = v;
            }
            for (int k = std::max(src, i - p); k < i; ++k) {
                const hc l = hdiv(row[k - i + p], Ur[(size_t)k * ne]);
                for (int c = 0; c <= p && k + c < W; ++c) {
                    const int col = k + c - i + p;
                    if (col > k - i + p && col <= 2 * p) row[col] = hsub(row[col], hmul(l, Ur[(size_t)k * ne + c]));
                }
                row[k - i + p] = {0.0, 0.0};
            }
            double big = 0.0, dif = 0.0;
            for (int c = 0; c <= p; ++c) {
                Ur[(size_t)i * ne + c] = (i + c < W) ? row[c + p] : hc{0.0, 0.0};
                big = std::max(big, habs(U[(size_t)i * ne + c]));
                dif = std::max(dif, habs(hsub(Ur[(size_t)i * ne + c], U[(size_t)i * ne + c])));
            }
            if (dif > 1e-14 * big) last = i;
        }
        r.rst = last - src + 1;
    }
    // forward: L y = e_src (unit lower)
    {
        std::vector<hc> y(W, hc{0.0, 0.0});
        y[src] = {1.0, 0.0};
        double peak = 1.0;
        int last = src;
        std::vector<double> mag(W, 0.0);
        for (int i = src + 1; i < W; ++i) {
            hc v = {0.0, 0.0};
            for (int k = 1; k <= p && i - k >= src; ++k) v = hsub(v, hmul(Lm[(size_t)(i - k) * ne + k], y[i - k]));
            y[i] = v;
            mag[i] = habs(v);
            peak = std::max(peak, mag[i]);
        }
        for (int i = src + 1; i < W; ++i)
            if (mag[i] > tol * peak) last = i;
        r.fwd = last - src;
    }
    // backward: (D^-1 U) x = e_src (unit upper)
    {
        const int sb = W - 1 - src;
        std::vector<hc> x(W, hc{0.0, 0.0});
        std::vector<double> mag(W, 0.0);
        x[sb] = {1.0, 0.0};
        double peak = 1.0;
        int first = sb;
        for (int i = sb - 1; i >= 0; --i) {
            hc v = {0.0, 0.0};
            const hc d = U[(size_t)i * ne];
            for (int k = 1; k <= p && i + k <= sb; ++k) v = hsub(v, hmul(hdiv(U[(size_t)i * ne + k], d), x[i + k]));
            x[i] = v;
            mag[i] = habs(v);
            peak = std::max(peak, mag[i]);
        }
        for (int i = sb - 1; i >= 0; --i)
            if (mag[i] > tol * peak) first = i;
        r.bwd = sb - first;
    }
    return r;
}
}  // namespace

Reach solve_reach(const double* S, const double* X, int N, int p, double e, int asym, double tol) {
    const int W = std::min(N, 900);
    Reach a = window_reach(S, X, p, e, asym, 0, W, tol);
    if (N > W)
        for (int w0 : {(N - W) / 2, N - W}) {
            const Reach b = window_reach(S, X, p, e, asym, w0, W, tol);
            a.fwd = std::max(a.fwd, b.fwd);
            a.bwd = std::max(a.bwd, b.bwd);
            a.rst = std::max(a.rst, b.rst);
        }
    return a;
}

DPlan coupling_plan(const double* S, const double* X, int N, int p, double emax, int asym) {
    DPlan d{asym ? kPKD : kPKS, kRestart};
    if (!(emax > 0.0)) return d;
    int reach = 0, rst = 0;
    for (double f : {1.0, 0.5, 0.25, 0.0}) {
        const Reach r = solve_reach(S, X, N, p, emax * f, asym);
        reach = std::max(reach, std::max(r.fwd, r.bwd));
        rst = std::max(rst, r.rst);
    }
    const int W = std::min(N, 900);
    const int NP = (N + kPM - 1) / kPM;
    int pk = (reach + kPM - 1) / kPM + 1;
    if (reach + std::min(60, W / 4) + 1 >= W) {
        // the response had not decayed inside the window: a carry over every
        // partition of the channel is exact when the whole operator was the
        // window, else nothing short of the sequential solve is
        pk = W == N ? NP - 1 : kPKMax + 1;
    }
    if (pk >= NP - 1) pk = std::max(NP - 1, 1);  // the carry reaches the channel start: exact
    if (pk > kPKMax) pk = kPKMax + 1;
    d.pk = std::max(pk, std::min(kPKS, std::max(NP - 1, 1)));
    d.restart = std::max(kRestart, (rst + 15) / 16 * 16);
    return d;
}

}  // namespace qpc
