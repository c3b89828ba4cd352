// qpc_solve.h -- partitioned banded Crank-Nicolson solves on gfx950.
//
// Every banded CN solve of the propagator (field-free, the two rotated
// coupling sweeps and the absorber) is  x = A^-1 (B z)  with B banded and A
// factored into banded triangular factors of half-width p (n_rank; the
// kernels are built for p = 7 and p = 8), held as "solve tables" and applied
// by one kernel.  Two row formats:
//
// Real LU tables (D-part sweeps, S + e X with X antisymmetric; absorber
// chol(S)), 2(p+1) coefficients per row [F_1..F_p, rdf, G_1..G_p, rdb]:
//   forward   y_i = rdf_i * (v_i - sum_{j=1..p} F_ij y_{i-j})
//   backward  x_i = rdb_i * y_i   - sum_{k=1..p} G_ik x_{i+k}
//
// Complex symmetric tables (R-part sweeps S + i e X with X symmetric, and the
// field-free S + i dt/2 H_l), A = L D L^T with L unit lower, p+1 coefficients
// per row [E_1..E_p, rD] with E_k(i) = L(i+k, i) and rD_i = 1 / D_i:
//   forward   z_i = v_i - sum_{j=1..p} E_j(i-j) z_{i-j},   y_i = rD_i z_i
//   backward  x_i = y_i - sum_{k=1..p} E_k(i) x_{i+k}
// The forward coefficients of row i are entries of the p rows above it, so
// one row serves both directions: half the bytes of an LU row.
//
// Rows are cut into partitions of kPM rows (the last one padded with identity
// rows: F = G = E = 0, rdf = rdb = rD = 1, v = 0, which leaves rows < N
// exact).  One lane owns one partition.  Pass A runs each partition from a
// zero state; the true incoming state is recovered from the pass-A end
// states of the PK preceding partitions through the partitions' p x p
// transfer matrices (state independent, built with the tables); pass B
// reruns each partition from that state.  The neglected tail is the response
// of L^-1 and U^-1 across PK*kPM rows, kept below 1e-17 relative: PK = 6 (144
// rows) for the R-part, field-free and absorber solves, whose factors'
// inverses fall below 1e-17 within 111-129 rows on every grid spacing and time
// step measured (tests/test_carry_depth.py); for the D-part, S + e d/dr,
// whose backward reach grows with |e|/h (186 rows at e/h = 0.18, ~600 at
// e/h = 1), the host measures the reach on the uploaded operators at the
// run's largest |e| (solve_reach) and sets PK per run; a reach beyond
// kPKMax partitions runs the sequential solve (one lane per channel over all
// rows), exact for any e.  The result equals the sequential solve to rounding.
//
// Device layout ("partition-interleaved"): lanes of a wave own 64 consecutive
// partitions, so element (partition g, row s of the partition, coefficient q)
// of a table with R coefficients per row sits at
// ((g/64 * kPM + s) * R + q) * 64 + g%64: every table or transfer access of a
// wave is one contiguous 64-element run.
#pragma once
#include "qpc_common.h"

namespace qpc {

constexpr int kPM = 24;        // rows per partition
constexpr int kPKS = 6;        // carry look-back of the R-part, field-free and absorber solves (144 rows)
constexpr int kPKD = 8;        // default carry look-back of the D-part solves (192 rows)
constexpr int kPKMax = 28;     // deepest partitioned carry (seg_max 8); deeper reaches run sequentially
constexpr int kLane = 64;      // partitions per wave group
constexpr int kMaxGroups = 4;  // wave groups per channel: N <= 4 * 64 * 24 = 6144
// solve segments: up to seg_max(PK) partitions per 64-lane work-group, plus
// PK halo partitions on either side
constexpr int seg_max(int pk) { return kLane - 2 * pk; }
// default factorisation restart: a partition's LU rows are rebuilt from
// kRestart rows above it, where the elimination state has contracted to
// rounding (the Schur complement of the banded S + e X converges
// geometrically); the host raises it to the measured forward reach
constexpr int kRestart = 64;
// band half-widths the kernels are built for
inline bool width_supported(int p) { return p == 7 || p == 8; }

// grids of at most this many solve work-groups stage the transfer sets in LDS
// (four waves per work-group, the right-hand side formed in the launch)
constexpr int64_t kStageGrid = 512;

struct Geom {
    int N;      // rows of the operator
    int P;      // partitions
    int NG;     // wave groups (ceil(P / 64)); the solve uses 64 * NG lanes per channel
    int hw;     // band half-width p
    int64_t tab_elems;   // elements per real table  (NG * 64 * kPM * 2(p+1))
    int64_t ctab_elems;  // elements per complex symmetric table (NG * 64 * kPM * (p+1))
    int64_t trf_elems;   // elements per transfer set (NG * 64 * 2 * p^2)
    // launch shape of the partitioned solves (the autotuner's choices; the
    // result does not depend on them): segments per channel beyond the fewest
    // that hold the partitions, and the staged-grid limit
    int seg_add = 0;
    int64_t stage_grid = kStageGrid;
    // the run's table batches: 0 factorised beside the steps (second
    // stream), 1 the steps wait for each batch as soon as it is queued
    int serial = 0;
};
inline Geom make_geom(int N, int p) {
    Geom g;
    g.N = N;
    g.hw = p;
    g.P = (N + kPM - 1) / kPM;
    g.NG = (g.P + kLane - 1) / kLane;
    g.tab_elems = (int64_t)g.NG * kLane * kPM * 2 * (p + 1);
    g.ctab_elems = (int64_t)g.NG * kLane * kPM * (p + 1);
    g.trf_elems = (int64_t)g.NG * kLane * 2 * p * p;
    return g;
}

// right-hand-side operator B of the solve
enum RhsKind {
    RHS_CBAND = 0,  // complex symmetric band (S - i dt/2 H_l, cn_hc_propagate0)
    RHS_RBAND = 1,  // real symmetric band (absorber mask)
    RHS_SX_SYM = 2, // S - ifet*X, X real symmetric (1/r), ifet complex      (propagate.hpp:33-62, symflag 0)
    RHS_SX_ASY = 3, // S + ifet*X lower / S - ifet*X upper, X antisymmetric (symflag 1), ifet real
};

// Bands on device are diagonal-major: D[j * N + i] = B(i, i + j), j = 0..p,
// zero where i + j >= N, so the lanes of a wave (consecutive rows) read
// consecutive words.
struct RhsArgs {
    const cplx* cband;      // RHS_CBAND: diagonal-major bands per table id, stride (p+1)*N
    const double* rband;    // RHS_RBAND: the mask band, diagonal-major
    const cplx* sx;         // RHS_SX_*: {S, X} pairs, diagonal-major (one 16-byte load per entry of both)
    const double* eig;      // RHS_SX_*: eigenvalue per channel
    double ifdt;            // RHS_SX_*: ifet = eig * ifdt (times i for RHS_SX_SYM)
};

// channel -> table id: free solve uses l(ch) = il + |m0 + im|; others use ch or 0
struct TabMap {
    int mode;   // 0: table = ch, 1: table = 0, 2: table = l(ch), 3: table = ch % L (batched shots, m = 0),
                // 4: table = map[ch] & kMapIdx, conjugated when map[ch] & kMapConj (shared tables)
    int L, m0;
    const int* map = nullptr;
    // optional order of the (segment, channel) work-groups of a partitioned
    // solve: work-group i (linear, segment fastest) runs order[i] = ch * nseg
    // + seg.  Work-groups are dealt to the XCDs round-robin by linear id, so
    // an order that gives the channels of one shared table ids of one residue
    // mod 8 keeps them on one XCD and its L2 (solve_order builds it).
    const int* order = nullptr;
    int order_nseg = 0;  // the segments per channel order was built for
};
constexpr int kMapConj = 1 << 30, kMapIdx = kMapConj - 1;
constexpr int kXcds = 8;
// segments per channel of the partitioned solve for this launch shape and
// carry depth (bsolve's choice; 0 when the carry is too deep for it and the
// solve runs sequentially), and the XCD-grouping order of its
// work-groups for channels whose tables are tab[ch] (nch channels)
int solve_segments(const Geom& geo, int pk);
void solve_order(const int* tab, int nch, int nseg, int* order);

// v = B z for nch channels (vout: nch * N)
void launch_rhs(hipStream_t s, int rhs, const RhsArgs& ra, TabMap tm, const cplx* zin, cplx* vout, int nch, Geom geo);
// x = A^-1 v for nch channels; tab/trf hold the tables and transfer sets of
// every table id, strides geo.ctab_elems (complex symmetric) or
// geo.tab_elems (real LU) and geo.trf_elems; role is the RhsKind of the
// operator (kernel name only).  vin and zout may not alias.  pk is the carry
// depth in partitions (0: the operator's default, kPKS or kPKD; > kPKMax:
// the sequential solve).
void launch_bsolve_c(hipStream_t s, int role, const cplx* tab, const cplx* trf, TabMap tm, const cplx* vin,
                     cplx* zout, int nch, Geom geo, int pk = 0);
void launch_bsolve_r(hipStream_t s, int role, const double* tab, const double* trf, TabMap tm, const cplx* vin,
                     cplx* zout, int nch, Geom geo, int pk = 0);

// v = B z (the RhsKind role with the bands of ra) and x = A^-1 v in one call.
// On grids that stage the transfer sets in LDS, and when zin and zout do not
// overlap, the solve forms v itself; otherwise launch_rhs writes v into vbuf
// (nch * N) and the solve reads it.
void launch_rsolve_c(hipStream_t s, int role, const RhsArgs& ra, const cplx* tab, const cplx* trf, TabMap tm,
                     const cplx* zin, cplx* vbuf, cplx* zout, int nch, Geom geo, int pk = 0);
void launch_rsolve_r(hipStream_t s, int role, const RhsArgs& ra, const double* tab, const double* trf, TabMap tm,
                     const cplx* zin, cplx* vbuf, cplx* zout, int nch, Geom geo, int pk = 0);

// transfer sets (both directions) of ntab tables
void launch_transfer_c(hipStream_t s, const cplx* tab, cplx* trf, int ntab, Geom geo);
// transfer sets of tables [t * stride, t * stride + per_t) for t < nt
void launch_transfer_c(hipStream_t s, const cplx* tab, cplx* trf, int per_t, int nt, int64_t stride, Geom geo);
void launch_transfer_r(hipStream_t s, const double* tab, double* trf, int ntab, Geom geo);

// complex symmetric tables from an lu_zsb LDL^T factor (nl bands, row-major N*(p+1))
// and from chol(S) (row-major N*(p+1))
void launch_tab_ldlt(hipStream_t s, const cplx* lower, cplx* tab, int nl, Geom geo);
void launch_tab_chol(hipStream_t s, const double* chol, double* tab, Geom geo);

// row-major band (N*(p+1), B(i,i+j) at i*(p+1)+j) -> diagonal-major, n bands
void launch_band_diag_r(hipStream_t s, const double* src, double* dst, int n, int N, int p);
// two diagonal-major real bands -> {S, X} pairs (RhsArgs::sx), (p+1)*N entries
void launch_band_pair(hipStream_t s, const double* S, const double* X, cplx* dst, int N, int p);
void launch_band_diag_c(hipStream_t s, const cplx* src, cplx* dst, int n, int N, int p);

// reference-order LU of (S + ifet X) for nt time points x nch channels
// (cn_propagate_kernel_avx1 elimination, propagate.hpp:33-86) on row-major
// S, X; ifdt[t] is the per-time-point scalar of cn_hl_propagate1/2
// (propagate.hpp:343,400).  Tables of (t, ch) at (t * nch + ch) * ctab_elems
// (launch_factor_sym: complex symmetric rows) or * tab_elems (launch_factor_asy:
// real LU rows), with e = eig[t*eig_tstride + ch] * ifdt[t] (eig_tstride 0:
// one eigenvalue per channel for every time point).  restart: rows of the
// restarted elimination above each partition (0: kRestart).  tab_tstride:
// tables per time point in the layout (0: nch), so that nch of them can be
// written into room for more.
void launch_factor_sym(hipStream_t s, const double* S, const double* X, const double* eig, const double* ifdt,
                       int nt, int nch, Geom geo, cplx* tab, int64_t eig_tstride = 0, int restart = 0,
                       int64_t tab_tstride = 0);
void launch_factor_asy(hipStream_t s, const double* S, const double* X, const double* eig, const double* ifdt,
                       int nt, int nch, Geom geo, double* tab, int64_t eig_tstride = 0, int restart = 0);

// ---- host ------------------------------------------------------------------
// Reach of the recurrences of A = S + e X (row-major bands N*(p+1)): asym 0,
// A = S + i e X with X symmetric; asym 1, A = S + e X with X antisymmetric
// (lower triangle S - e X).  The reference-order LU without pivoting of
// windows of A (at the start, in the middle and at the end of the grid: the
// operators are uniform in the interior, X = r grows to the end); fwd / bwd
// are the rows a unit source takes, through L^-1 / U^-1, to fall below tol
// times the response's peak; a response still above it at the window's end
// gives the window length.  rst: the rows after which an elimination begun
// afresh (the rows above dropped) has the full elimination's U rows to 1e-14
// of each row: the factorisation restart it needs.
struct Reach {
    int fwd, bwd, rst;
};
Reach solve_reach(const double* S, const double* X, int N, int p, double e, int asym, double tol = 1e-17);
// carry depth (partitions) and factorisation restart (rows) of coupling
// sweeps S + e X (asym as solve_reach) for scalars |e| <= emax: the reach at
// emax, emax/2, emax/4 and 0 plus one partition, and the largest restart
// distance there (at least kRestart); a depth above kPKMax selects the
// sequential solve
struct DPlan {
    int pk, restart;
};
DPlan coupling_plan(const double* S, const double* X, int N, int p, double emax, int asym);
// the D-part sweeps (X = d/dr)
inline DPlan d_part_plan(const double* S, const double* X, int N, int p, double emax) {
    return coupling_plan(S, X, N, p, emax, 1);
}

}  // namespace qpc
