/*
 * qpc_oracle.h -- CPU oracle for the QPC-TDSE Crank-Nicolson propagator.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain C++17/OpenMP restatement
 * of the reference algorithm (zzh535897/QPC-TDSE-CPC), used as the parity
 * checker by tests/, by __graft_entry__.smoke() and as the timed CPU baseline
 * ("port") in bench.py.  Nothing in the product library links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root).  Array layouts are the reference's:
 *   - complex FP64 is interleaved {re,im}, 16 bytes (support_avx.h:59,146);
 *   - coefficients c[(im*L + il)*N + ir]            (spherical/inventory.h:70-71);
 *   - real/complex bands band[i*ne + j] = A(i,i+j)  (operators/1e.hpp:20);
 *   - LDL^T factor f[i*ne] = D_i, f[i*ne+j] = L(i+j,i)  (algorithm/ed.h:208-289);
 *   - Cholesky factor chol[i*ne+j] = L(i+j,i)        (algorithm/ed.h:182-194);
 *   - angular projectors P[im][k][i] row-major      (propagator/cn_1e.hpp:23).
 * Band entries with i+j >= N (outside the matrix) are stored as zero: no
 * in-range result of the reference depends on them (ed.h:280-287,
 * blas2.hpp:501-506, propagate.hpp:64-97).
 */
#ifndef QPC_ORACLE_H
#define QPC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors the compile-time config fields of config/<task>.cpp used on the
 * propagation path (config/example01_a.cpp:20-117). */
typedef struct {
    int64_t n_ndim;   /* N, radial B-splines                      */
    int64_t n_rank;   /* p, B-spline degree = band half-width     */
    int64_t n_ldim;   /* L                                        */
    int64_t n_mdim;   /* M                                        */
    int64_t n_mmin;   /* m0                                       */
    double  para_rmin, para_rmax;
    double  para_zc;  /* hydrogenic charge (potential.hpp:9-17)   */
    double  para_me;  /* mass                                     */
    double  para_dt;
    double  mask_rmin, mask_rmax, mask_fact;
} qor_dims;

typedef struct qor_ctx qor_ctx;

/* Array identifiers for qor_get / qor_size. */
enum {
    QOR_S = 0,       /* real band  N*ne          operator_sc data            */
    QOR_CHOL,        /* real band  N*ne          operator_sc cholesky        */
    QOR_H,           /* real bands (lmax+1)*N*ne operator_hc                 */
    QOR_RINV,        /* real band  N*ne          operator_rinv1              */
    QOR_RDIF,        /* real band  N*ne          operator_rdif1              */
    QOR_RPOW1,       /* real band  N*ne          <Bi|r|Bj> (dipole)          */
    QOR_MASK,        /* real band  N*ne          absorber_mask data          */
    QOR_UPPER,       /* cplx bands (lmax+1)*N*ne S - i dt/2 H_l              */
    QOR_LOWER,       /* cplx bands (lmax+1)*N*ne LDL^T(S + i dt/2 H_l)       */
    QOR_YEIG,        /* real  M*L     eigenvalues of Y10 (D-part)            */
    QOR_YPRJ,        /* real  M*L*L   P_Y (tran_z2)                          */
    QOR_QEIG,        /* real  M*L     eigenvalues of i*Q10 (R-part)          */
    QOR_QPRJ,        /* cplx  M*L*L   P_Q (tran_z1, conjugated)              */
    QOR_MERGED,      /* cplx  M*L*L   P_Q * P_Y^H (z2_2_z1)                  */
    QOR_KNOTS,       /* real  N+p+3   knots, index -1..N+p+1                 */
    QOR_GSX,         /* real  (N+p)*q Gauss points on active intervals       */
    QOR_GSW,         /* real  (N+p)*q Gauss weights                          */
    QOR_Y10,         /* real  M*L     closed-form <l|cos|l+1>                */
    QOR_Q10,         /* real  M*L     closed-form Q10                        */
    QOR_DVDR,        /* real band N*ne <Bi|dV/dr|Bj> (acceleration)          */
    QOR_NARRAYS
};

int     qor_create(const qor_dims* d, qor_ctx** out);
void    qor_destroy(qor_ctx* c);
int64_t qor_size(const qor_ctx* c, int what);             /* number of doubles */
int     qor_get(const qor_ctx* c, int what, double* dst); /* copies qor_size doubles */
int     qor_lu_info(const qor_ctx* c);                    /* lu_zsb tiny-pivot count */
int     qor_set_threads(int n);                           /* OpenMP threads of later calls (n <= 0: query); returns the count */

/* Replace a setup array (used by tests to inject modified operators). */
int     qor_set(qor_ctx* c, int what, const double* src);

/* coefficient::initialize<1> (coefficient/1e.hpp:89-143) with cf=1:
 * eigenvector `in` of (H_l, S) normalised v^T S v = 1.  Sign fixed so that
 * the largest-magnitude component is positive (the reference leaves it to
 * LAPACKE_dsbgv).  coef has M*L*N complex entries (interleaved doubles). */
int     qor_eigenstate(const qor_ctx* c, int64_t im, int64_t il, int64_t in,
                       double* energy, double* coef);

/* eigenstate_sph::initialize (population_1e.hpp:57-118) for the criterion
 * "the nsel lowest states of every l": W[l][n][ir] = (S v_{n,l})_ir (the
 * left-hand basis, :108-113) with v^T S v = 1, sign as qor_eigenstate;
 * E[l][n] (may be NULL) the energies.  W has (lmax+1)*nsel*N doubles. */
int     qor_projection_basis(const qor_ctx* c, int64_t nsel, double* W, double* E);
/* eigenstate_sph::projection (population_1e.hpp:29-55):
 * proj[(im*L+il)*stride + in] = sum_ir W[l][in][ir] * coef[im][il][ir],
 * l = il + |m0+im|, in < nsel (complex, interleaved doubles). */
int     qor_projection(const qor_ctx* c, const double* coef, int64_t nsel, const double* W,
                       double* proj, int64_t stride);

/* One with-field step of propagator_cn_hc_vgwz (cn_1e.hpp:682-709). */
int     qor_step_vgwz(qor_ctx* c, double* coef, double a_i, double a_f);
/* Field-free step propagate(coef) (cn_1e.hpp:711-714). */
int     qor_step_free(qor_ctx* c, double* coef);
/* absorber_mask::operator() (absorber.hpp:52-64). */
int     qor_mask(qor_ctx* c, double* coef);
/* Driver loop of realtime_lp (sphTDSE.cpp:786-831) over steps [it0,it1):
 * step it uses A(t_it)=atab[it], A(t_it+1)=atab[it+1] while it < nt_field,
 * field-free afterwards; then the mask if use_mask.  obs (may be NULL)
 * receives per step {norm, Re z} measured BEFORE the propagation as the
 * reference's obsv::observe does. */
int     qor_run(qor_ctx* c, double* coef, const double* atab, int64_t nt_field,
                int64_t it0, int64_t it1, int use_mask, double* obs);

/* Length gauge, z polarisation: propagator_cn_hc_lgwz::propagate(coef, ez_ti,
 * ez_tf) (cn_1e.hpp:598-613), the P_Y rotations (Y10 eigenvectors) around
 * the sweeps (S + i e r) with e = yeig * ez * dt/4, and the field-free
 * factor; qor_run_lg the driver loop with E(t) at the time points. */
int     qor_step_lgwz(qor_ctx* c, double* coef, double ez_ti, double ez_tf);
int     qor_run_lg(qor_ctx* c, double* coef, const double* etab, int64_t nt_field, int64_t it0, int64_t it1,
                   int use_mask);

/* qor_run with every observable of the library's in-loop recording: obs8
 * (may be NULL) receives per step {norm, Re z (f = r), Re z with f = dV/dr,
 * outer norm over r >= r0, x, y (f = r), x, y (f = dV/dr)}, observed
 * BEFORE the propagation. */
int     qor_run_record(qor_ctx* c, double* coef, const double* atab, int64_t nt_field, int64_t it0,
                       int64_t it1, int use_mask, double r0, double* obs8);

/* Elliptically polarised (x-y) field: propagator_cn_hc_vgwr (cn_1e.hpp:812-951).
 * qor_set_ep takes its angular eigen-data, M*L = ML channels:
 *   isub          sub-block selector of propagate<isub> (dimension/1e.hpp:57-80):
 *                 -1 (or 0, 1) all channels, 2 gathered even, 3 gathered odd;
 *   eig_x1..y2    ML doubles each: prop_x1.zeig (Y11, mult 1), prop_x2.zeig
 *                 (Q11, mult -i), prop_y1.zeig (Y11, mult -i), prop_y2.zeig
 *                 (Q11, mult 1), ordered by converter::catagorize
 *                 (utilities/converter.hpp:27-107);
 *   tran_x1/y2    ML*ML doubles, row-major (real parts of the projectors);
 *   x1_2_x2, x2_2_y1, y1_2_y2   ML*ML complex (merge_two products).
 * qor_step_vgwr is propagate<isub>(coef, ax_ti, ax_tf, ay_ti, ay_tf)
 * (cn_1e.hpp:898-945); qor_run_ep the loop of the EP task
 * (binsrc/sphTDSE.cpp:921-951): with-field steps while it < nt_field,
 * field-free afterwards, then the mask if use_mask. */
int     qor_set_ep(qor_ctx* c, int isub, const double* eig_x1, const double* eig_x2, const double* eig_y1,
                   const double* eig_y2, const double* tran_x1, const double* tran_y2, const double* x1_2_x2,
                   const double* x2_2_y1, const double* y1_2_y2);
int     qor_step_vgwr(qor_ctx* c, double* coef, double ax_ti, double ax_tf, double ay_ti, double ay_tf);
/* propagator_cn_hc_lgwr (cn_1e.hpp:717-810), the length gauge of the same
 * grid: qor_set_lgwr takes the zeig of x1 (Y11) and y1 (Y11 with -i), the
 * real tran_x1, and x1_2_y1 = U[y1] U[x1]^H and tran_y1 = U[y1] (complex,
 * U = conj of the eigenvector rows); qor_step_lgwr is propagate<isub>(coef,
 * ex_ti, ex_tf, ey_ti, ey_tf) (:767-803), qor_run_lgwr its driver loop. */
int     qor_set_lgwr(qor_ctx* c, int isub, const double* eig_x1, const double* eig_y1, const double* tran_x1,
                     const double* x1_2_y1, const double* tran_y1);
int     qor_step_lgwr(qor_ctx* c, double* coef, double ex_ti, double ex_tf, double ey_ti, double ey_tf);
int     qor_run_lgwr(qor_ctx* c, double* coef, const double* extab, const double* eytab, int64_t nt_field,
                     int64_t it0, int64_t it1, int use_mask);
int     qor_run_ep(qor_ctx* c, double* coef, const double* axtab, const double* aytab, int64_t nt_field,
                   int64_t it0, int64_t it1, int use_mask);

/* The learned Krylov controller's network (no counterpart in the reference;
 * the product's is include/qpc_mlp.h): a 7 -> 64 -> 64 -> 2 perceptron with
 * exact-erf GELU, evaluated here in FP64.  params (4802 doubles) in the
 * layout W1[64][7], b1[64], W2[64][64], b2[64], W3[2][64], b3[2];
 * qor_mlp_forward: Y[n][2] = f(X[n][7]); qor_mlp_loss_grad: the loss
 * L = 1/(2n) sum (y - t)^2 and dL/dparams by back-propagation. */
int     qor_mlp_forward(const double* params, const double* X, int64_t n, double* Y);
int     qor_mlp_loss_grad(const double* params, const double* X, const double* T, int64_t n, double* loss,
                          double* grad);

/* One step of the Krylov propagator propagator_arnoldi::propagate1
 * (la_1e.hpp:312-335) for H(az) of propagator_hc_vgwz::multiply
 * (la_1e.hpp:480-497): coef <- exp(-i dt S^-1 H(az)) coef on the Krylov
 * space of at most kmax vectors (S-orthogonal Lanczos, thresholded full
 * re-orthogonalisation, tridiagonal eigen-solve).  The reference's
 * operator_sc::inverse_all (operators/1e.hpp:1253-1260) is restated as
 * intended (one Cholesky solve per channel).  Returns the criterion of
 * la_1e.hpp:296-310 (0: estimate > emax, 1: in [emin, emax], 2: < emin);
 * m_used / eps_out (may be NULL) get the subspace size and the estimate
 * |x_{m-1}|^2; ab_out (may be NULL, 2*kmax doubles) the Lanczos a_i, b_{i+1}. */
int     qor_krylov_step(qor_ctx* c, double* coef, double az, double dt, int64_t kmax, double emin,
                        double emax, int64_t* m_used, double* eps_out, double* ab_out);

/* <c|S|c> summed over channels (operator_rsym::observe, operators/1e.hpp:69-78). */
double  qor_norm(const qor_ctx* c, const double* coef);
/* Windowed norm: sum over channels of <c|S_w|c> with S_w = <Bi|theta(r-r0)|Bj>. */
double  qor_outer_norm(const qor_ctx* c, const double* coef, double r0);
/* observable_z::observe (observable/1e.hpp:26-49) with f(r)=r (which=0) or
 * f(r)=dV/dr (which=1); returns the real part as obsv::observe does. */
double  qor_dipole_z(const qor_ctx* c, const double* coef, int which);

/* observable_r (observable/1e.hpp:53-129) with f(r)=r (which=0) or dV/dr
 * (which=1): *x = Re(rm + rp)/2, *y = Im(rp - rm)/2 as obsv::observe takes
 * them (binsrc/sphTDSE.cpp:524-527). */
int     qor_dipole_xy(const qor_ctx* c, const double* coef, int which, double* x, double* y);

/* spectrum_c::measure_lmd (spectrum/momentum_1e.hpp:384-401), the partial
 * photoelectron momentum amplitudes saved as /lmd1:
 * out[(ikr*M + im)*L + il] = sqrt(2/pi) cps[ikr][l] sum_j c[im][il][j] ijl[ikr][l][j]
 * (complex), with ijl (nkr*(lmax+1)*N doubles) and cps (complex nkr*(lmax+1))
 * of spectrum_c_pcs::initialize. */
int     qor_measure_lmd(const qor_ctx* c, const double* coef, int64_t nkr, const double* ijl, const double* cps,
                        double* out);

/* spectrum_c_tsf::accumulate, z-polarised (momentum_1e.hpp:609-627, 675-719;
 * a field-free step is az0 = az1 = 0, :631-673) on the state after a step:
 * state = {time, disz} is advanced, work (complex [nkr][nkt][nkp]) gains the
 * step's flux term.  Tables as qpc_setup_tsurff makes them (ib, b0, br of
 * nb splines; j0, jr [nkr][lmax+1]; y0 [nkt][M][L]; e0 complex [nkp][M];
 * pmat0 [M][L]); kt are the polar angles. */
int     qor_tsurff_accumulate(const qor_ctx* c, const double* coef, double az0, double az1, double dt,
                              int64_t nkr, int64_t nkt, int64_t nkp, const double* kr, const double* kt,
                              int64_t nb, const int64_t* ib, const double* b0, const double* br,
                              const double* j0, const double* jr, const double* y0, const double* e0,
                              const double* pmat0, double* state, double* work);

/* spectrum_c_tsf::accumulate (xy) (momentum_1e.hpp:716-776), the elliptically
 * polarised form; with ax = ay = 0 the field-free form of that run.
 * state = {time, disx, disy}; kt, kp the polar and azimuthal angles; pmat1,
 * pmat2 [M][L] the sin(th) e^(-i ph) elements of integrate_sinexp1/2
 * (basisfunc.hpp:241-259); other tables as qor_tsurff_accumulate. */
int     qor_tsurff_accumulate_xy(const qor_ctx* c, const double* coef, double ax0, double ax1, double ay0,
                                 double ay1, double dt, int64_t nkr, int64_t nkt, int64_t nkp, const double* kr,
                                 const double* kt, const double* kp, int64_t nb, const int64_t* ib,
                                 const double* b0, const double* br, const double* j0, const double* jr,
                                 const double* y0, const double* e0, const double* pmat1, const double* pmat2,
                                 double* state, double* work);

/* field_component_sinen<2> (field.h:90-133): A(t) for one component. */
double  qor_sinen2(double e0, double w0, double nc, double n0, double c0, double t);
/* field::initialize_t para_nt (field.h:245-281) for one component starting at
 * n0=0: 1 + floor((tf-ti)/dt). */
int64_t qor_para_nt(double w0, double nc, double n0, double dt);

/* ---- kernel-level entry points (used by tests to pin each kernel) ---- */
/* impl_ed_simd::lu_zsb (algorithm/ed.h:208-289), in place on a complex band. */
int     qor_lu_zsb(int64_t n, int64_t p, double* band);
/* cn_propagate_kernel_avx1 (propagate.hpp:3-98): z <- (S+aR)\(S-aR) z with
 * symflag 0 (R symmetric, a complex) or 1 (R antisymmetric, a real). */
void    qor_cn_kernel(int64_t n, int64_t p, int symflag, const double* smat,
                      const double* rmat, double ifet_re, double ifet_im, double* z);
/* symb_mul_vecd<n,ne,0> for a complex band (blas2.hpp:487-507). */
void    qor_symb_mul(int64_t n, int64_t p, const double* cband, const double* x, double* y);
/* ltrb_ldl_vecd then utrb_ldl_vecd (blas2.hpp:1058-1095). */
void    qor_ldl_solve(int64_t n, int64_t p, const double* factor, double* x);
/* banded Cholesky of a real symmetric band (LAPACKE_dpbtrf 'L' + clear,
 * ed.h:182-194); returns the LAPACK-style info. */
int     qor_chol(int64_t n, int64_t p, double* band);
/* real symmetric tridiagonal eigen-decomposition, ascending; z row-major
 * z[k*n+i] = component i of eigenvector k. */
int     qor_tridiag_eig(int64_t n, const double* diag, const double* off, double* eig, double* z);
/* Gauss-Legendre nodes/weights on [-1,1], ascending. */
void    qor_gauss_legendre(int64_t n, double* x, double* w);

#ifdef __cplusplus
}
#endif
#endif
