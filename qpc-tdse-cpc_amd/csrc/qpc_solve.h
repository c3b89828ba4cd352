// qpc_solve.h -- partitioned banded Crank-Nicolson solves on gfx950.
//
// Every banded CN solve of the propagator (field-free, the two rotated
// coupling sweeps and the absorber) is  x = U^-1 L^-1 (B z)  with B banded and
// L, U banded triangular of half-width p = 7.  All three pieces are turned
// into "solve tables" (16 coefficients per row) and applied by one kernel:
//
//   forward   y_i = rdf_i * (v_i - sum_{j=1..p} F_ij y_{i-j})
//   backward  x_i = rdb_i * y_i   - sum_{k=1..p} G_ik x_{i+k}
//
// table row layout: [F_1..F_7, rdf, G_1..G_7, rdb] (real or complex).
//
// The rows of a channel are padded to a multiple of kPM with identity rows
// (F = G = 0, rdf = rdb = 1, v = 0), which leaves rows < N exact, and cut
// into partitions of kPM rows.  Pass A runs each
// partition from a zero state; the true incoming state is recovered from the
// pass-A end states of the kPK preceding partitions through the partitions'
// 7x7 transfer matrices (state independent, precomputed with the tables);
// pass B reruns each partition from that state.  The neglected tail is the
// transfer over kPK*kPM = 168 rows, below 1e-20 for every operator of the
// propagator (measured decay of L^-1 and U^-1), so the result equals the
// sequential solve to rounding.
#pragma once
#include "qpc_common.h"

namespace qpc {

constexpr int kP = 7;          // band half-width the solver is built for (n_rank)
constexpr int kTab = 16;       // coefficients per table row
constexpr int kPM = 8;         // rows per partition
constexpr int kPK = 21;        // carry look-back, partitions (168 rows)
constexpr int kCP = 32;        // partitions per work-group chunk
constexpr int kSolveThreads = 256;

struct Geom {
    int N;      // rows of the operator
    int Np;     // rows padded to a multiple of kPM (table rows per table)
    int P;      // partitions = Np / kPM
    int nchunk; // chunks per channel
};
inline Geom make_geom(int N) {
    Geom g;
    g.N = N;
    g.Np = (N + kPM - 1) / kPM * kPM;
    g.P = g.Np / kPM;
    g.nchunk = (g.P + kCP - 1) / kCP;
    return g;
}

// right-hand-side operator B of the solve
enum RhsKind {
    RHS_CBAND = 0,  // complex symmetric band (S - i dt/2 H_l, cn_hc_propagate0)
    RHS_RBAND = 1,  // real symmetric band (absorber mask)
    RHS_SX_SYM = 2, // S - ifet*X, X real symmetric (1/r), ifet complex      (propagate.hpp:33-62, symflag 0)
    RHS_SX_ASY = 3, // S + ifet*X lower / S - ifet*X upper, X antisymmetric (symflag 1), ifet real
};

struct RhsArgs {
    const cplx* cband;      // RHS_CBAND: bands per table id, stride N*(p+1)
    const double* rband;    // RHS_RBAND / RHS_SX_*: S (or mask)
    const double* xband;    // RHS_SX_*: X
    const double* eig;      // RHS_SX_*: eigenvalue per channel
    double ifdt;            // RHS_SX_*: ifet = eig * ifdt (times i for RHS_SX_SYM)
};

// channel -> table id: free solve uses l(ch) = il + |m0 + im|; others use ch or 0
struct TabMap {
    int mode;   // 0: table = ch, 1: table = 0, 2: table = l(ch)
    int L, m0;
};

void launch_bsolve_c(hipStream_t s, int rhs, const RhsArgs& ra, const cplx* tab, const cplx* trf, TabMap tm,
                     const cplx* zin, cplx* zout, int nch, Geom geo);
void launch_bsolve_r(hipStream_t s, int rhs, const RhsArgs& ra, const double* tab, const double* trf, TabMap tm,
                     const cplx* zin, cplx* zout, int nch, Geom geo);

// transfer matrices of every partition of ntab tables: trf[t][g][dir][7][7]
void launch_transfer_c(hipStream_t s, const cplx* tab, cplx* trf, int ntab, Geom geo);
void launch_transfer_r(hipStream_t s, const double* tab, double* trf, int ntab, Geom geo);

// tables from an lu_zsb LDL^T factor (complex, nl bands) and from chol(S)
void launch_tab_ldlt(hipStream_t s, const cplx* lower, cplx* tab, int nl, Geom geo);
void launch_tab_chol(hipStream_t s, const double* chol, double* tab, Geom geo);

// reference-order LU of (S + ifet X) for nt time points x nch channels
// (cn_propagate_kernel_avx1 elimination, propagate.hpp:33-86); ifdt[t] is the
// per-time-point scalar of cn_hl_propagate1/2 (propagate.hpp:343,400)
void launch_factor_sym(hipStream_t s, const double* S, const double* X, const double* eig, const double* ifdt,
                       int nt, int nch, Geom geo, cplx* tab);
void launch_factor_asy(hipStream_t s, const double* S, const double* X, const double* eig, const double* ifdt,
                       int nt, int nch, Geom geo, double* tab);

}  // namespace qpc
