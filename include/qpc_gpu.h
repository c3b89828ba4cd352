/*
 * qpc_gpu.h -- C ABI of the MI355X Crank-Nicolson propagator for QPC-TDSE.
 *
 * The library replaces the propagation loop of the reference solver
 * (zzh535897/QPC-TDSE-CPC): the propagator object used by the driver's time
 * loop (binsrc/sphTDSE.cpp:786-831) and the absorber applied after every step
 * (sphTDSE.cpp:807-810).  The wavefunction stays resident in device memory
 * between calls; host arrays are only read at upload time and written by
 * qpc_get_state / qpc_observe.
 *
 * Layouts are exactly the reference's (so the reference's own set-up output
 * can be handed over without reshuffling):
 *   complex FP64            interleaved {re,im}, 16 B  (libraries/support_avx.h:59,146)
 *   coefficients            c[(im*L + il)*N + ir]      (spherical/inventory.h:70-71)
 *   radial bands            band[i*ne + j] = A(i,i+j), ne = n_rank+1 (operators/1e.hpp:20)
 *   LDL^T factor (lu_zsb)   f[i*ne] = D_i, f[i*ne+j] = L(i+j,i)      (algorithm/ed.h:208-289)
 *   Cholesky of S           chol[i*ne+j] = L(i+j,i)    (algorithm/ed.h:182-194)
 *   angular projectors      P[im][k][i] row-major      (propagator/cn_1e.hpp:16-147)
 *
 * Every entry point returns 0 on success and a negative status otherwise
 * (QPC_E_*); qpc_last_error() gives a message.  No exceptions cross the ABI.
 * A context is used from one host thread at a time.
 */
#ifndef QPC_GPU_H
#define QPC_GPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QPC_OK            0
#define QPC_E_ARG        -1   /* null pointer or out-of-range argument        */
#define QPC_E_SHAPE      -2   /* dimensions the kernels do not support        */
#define QPC_E_HIP        -3   /* a HIP runtime call failed                    */
#define QPC_E_STATE      -4   /* operands not uploaded yet                    */
#define QPC_E_NUMERIC    -5   /* non-finite value or zero pivot detected      */
#define QPC_E_SETUP      -6   /* host set-up (basis, eigen-solve) failed      */

typedef struct qpc_ctx qpc_ctx;

/* Problem shape: the template parameters of dimension_1e
 * (spherical/dimension/1e.hpp:19-50) plus the time step of the config file
 * (config/example01_a.cpp:86). */
typedef struct {
    int64_t n_ndim;   /* N  radial B-splines                 */
    int64_t n_rank;   /* p  B-spline degree (band half-width) */
    int64_t n_ldim;   /* L  angular channels per m           */
    int64_t n_mdim;   /* M  number of m values               */
    int64_t n_mmin;   /* m0 smallest m                       */
    double  para_dt;  /* time step                           */
} qpc_dims;

/* Creates a context on the current HIP device and allocates device memory
 * for the state and every operand.  Replaces the allocation half of
 * propagator_cn_hc_vgwz::initialize (propagator/cn_1e.hpp:642-668). */
int  qpc_create(const qpc_dims* dims, qpc_ctx** out);
int  qpc_destroy(qpc_ctx* ctx);
const char* qpc_last_error(const qpc_ctx* ctx);

/* Radial operators, reference band layout, N*ne doubles each:
 * S (operator_sc data), chol (operator_sc::cholesky, operators/1e.hpp:1232-1245),
 * rinv (operator_rinv1), rdif (operator_rdif1), rpow1 (<Bi|r|Bj>, may be NULL),
 * mask (absorber_mask data, absorber.hpp:42-49, may be NULL). */
int  qpc_upload_radial(qpc_ctx* ctx, const double* S, const double* chol, const double* rinv,
                       const double* rdif, const double* rpow1, const double* mask);

/* cn_component_hc bands (propagator/cn_1e.hpp:342-360): upper = S - i dt/2 H_l
 * and lower = lu_zsb(S + i dt/2 H_l), (lmax+1)*N*ne complex each. */
int  qpc_upload_hc(qpc_ctx* ctx, const double* upper, const double* lower);
/* The field-free bands built on the device, in place of qpc_upload_hc:
 * cn_component_hc::initialize (cn_1e.hpp:642-680) with
 * H_l = <Bi|V|Bj> + l(l+1)/(2 mass) <Bi|1/r^2|Bj> + 1/(2 mass) <Bi'|Bj'>
 * (integrate_hc, basisfunc.hpp:92-106) for l = 0..lmax, upper = S - i dt/2 H_l
 * and lower = the lu_zsb LDL^T factor of S + i dt/2 H_l (algorithm/ed.h:
 * 208-289), one l per device thread.  V, K, T: the three l-independent bands
 * (N*ne doubles each, the layout of S; QPC_SETUP_HV/HK/HT).  S comes from
 * qpc_upload_radial, which must come first.  QPC_E_NUMERIC when a pivot
 * has |D|^2 < 1e-31 (lu_zsb's info). */
int  qpc_build_hc(qpc_ctx* ctx, const double* V, const double* K, const double* T, double mass);

/* Angular eigen-data of propagator_cn_hc_vgwz (propagator/cn_1e.hpp:673-679):
 * yeig/qeig M*L doubles (prop_z2.yeig, prop_z1.yeig), yprj M*L*L doubles
 * (tran_z2), qprj and merged M*L*L complex (tran_z1 after conjugate(),
 * z2_2_z1 = P_Q P_Y^H).  y10 (M*L doubles, operator_y10 data) is optional
 * (NULL) and only needed by qpc_observe(QPC_OBS_DIPOLE_Z). */
int  qpc_upload_angular(qpc_ctx* ctx, const double* yeig, const double* yprj, const double* qeig,
                        const double* qprj, const double* merged, const double* y10);

/* State transfer, M*L*N complex (interleaved doubles). */
int  qpc_set_state(qpc_ctx* ctx, const double* coef);
int  qpc_get_state(qpc_ctx* ctx, double* coef);

/* propagator_cn_hc_vgwz::propagate(coef, A(t_i), A(t_f)) (propagator/cn_1e.hpp:682-709).
 * Any A: the carry depth and factorisation restart of the D-part sweeps
 * (S + e d/dr, e = yeig A dt/4) follow the reach measured on the uploaded
 * operators at the run's largest |e| (qpc_coupling_plan), with a sequential
 * solve beyond what the partitioned one holds; QPC_E_NUMERIC only for a
 * non-finite A (qpc_run, qpc_run_batch and qpc_run_ep alike).
 * The per-step entry points (qpc_step, qpc_step_lg, qpc_step_free,
 * qpc_mask) return when the work is queued; qpc_get_state, qpc_observe and
 * qpc_sync wait for it.  qpc_step keeps the tables of its A(t_f): the next
 * call with A(t_i) equal to it (the driver's loop) factorises A(t_f) alone;
 * any other call, a qpc_run or an upload starts afresh. */
int  qpc_step(qpc_ctx* ctx, double a_i, double a_f);
/* Length gauge, z polarisation: propagator_cn_hc_lgwz::propagate(coef,
 * E(t_i), E(t_f)) (propagator/cn_1e.hpp:598-613), selected by gauge 1 for
 * the LP task (binsrc/sphTDSE.cpp:574-579): the Y10 rotations around the
 * sweeps (S + i e r), e = yeig E dt/4, and the field-free factor.  Needs
 * rpow1 (<Bi|r|Bj>) from qpc_upload_radial.  qpc_run_lg is qpc_run with
 * etab[it] = E at the time points (the driver's laser values in length
 * gauge); the field-free step and the absorber are the velocity gauge's.
 * qpc_step_lg keeps its E(t_f) tables as qpc_step does. */
int  qpc_step_lg(qpc_ctx* ctx, double e_i, double e_f);
int  qpc_run_lg(qpc_ctx* ctx, const double* etab, int64_t nt_field, int64_t it0, int64_t it1, int use_mask);
/* propagator_cn_hc_vgwz::propagate(coef), field-free (propagator/cn_1e.hpp:711-714). */
int  qpc_step_free(qpc_ctx* ctx);
/* absorber_mask::operator()(coef, s) (absorber.hpp:52-64). */
int  qpc_mask(qpc_ctx* ctx);

/* The driver's time loop realtime_lp (binsrc/sphTDSE.cpp:786-831) on device:
 * for it in [it0,it1): with-field step with A(t_it)=atab[it], A(t_it+1)=atab[it+1]
 * while it < nt_field, field-free step afterwards; then the absorber when
 * use_mask != 0.  atab holds nt_field+1 values (A at the time-grid points).
 * Returns when the work is queued; qpc_sync waits for it. */
int  qpc_run(qpc_ctx* ctx, const double* atab, int64_t nt_field, int64_t it0, int64_t it1, int use_mask);
int  qpc_sync(qpc_ctx* ctx);

/* Observables of the current state (synchronous):
 *   QPC_OBS_NORM      sum_ml <c|S|c>              (operators/1e.hpp:69-78)
 *   QPC_OBS_DIPOLE_Z  Re observable_z with f(r)=r (observable/1e.hpp:26-49)  */
#define QPC_OBS_NORM      0
#define QPC_OBS_DIPOLE_Z  1
#define QPC_OBS_OUTER     2   /* sum_ml <c|S_out|c>, S_out from qpc_upload_outer (the windowed
                                 overlap of observable_molec, observable/1e.hpp:176-200) */
#define QPC_OBS_ACCEL_Z   3   /* Re observable_z with f(r) = dV/dr, the dipole acceleration of the
                                 HHG spectrum (observable/1e.hpp:26-49 with func_obsv_z = dV/dr);
                                 band from qpc_upload_accel */
#define QPC_OBS_DIPOLE_X  4   /* observable_r with f(r) = r: Re(rm + rp)/2 (observable/1e.hpp:53-129) */
#define QPC_OBS_DIPOLE_Y  5   /*                             Im(rp - rm)/2                         */
#define QPC_OBS_ACCEL_X   6   /* the same with f(r) = dV/dr (band from qpc_upload_accel)          */
#define QPC_OBS_ACCEL_Y   7
int  qpc_observe(qpc_ctx* ctx, int what, double* out);
/* Overlap band of the outer region (N*ne doubles, reference band layout). */
int  qpc_upload_outer(qpc_ctx* ctx, const double* band);
/* <Bi|dV/dr|Bj> (N*ne doubles, reference band layout; integrate_0r0 with
 * f = dV/dr, e.g. Z/r^2 for hydrogenic{Z}) for QPC_OBS_ACCEL_Z. */
int  qpc_upload_accel(qpc_ctx* ctx, const double* band);

/* Observables recorded inside the time loops, as the driver's obsv::observe
 * does before every step (binsrc/sphTDSE.cpp:793-795, 516-529): qpc_run,
 * qpc_run_lg and qpc_run_ep observe the state before each step they queue,
 * on the device, into a history of `cap` rows; qpc_record_read copies the
 * rows recorded so far, 8 doubles each: norm, Re <z> (f = r), z acceleration
 * (f = dV/dr), outer norm, <x>, <y> (observable_r with f = r,
 * observable/1e.hpp:53-129, as obsv::observe takes them:
 * x = Re(rm + rp)/2, y = Im(rp - rm)/2), x and y acceleration (f = dV/dr);
 * 0 for an observable not in flags.  qpc_record(ctx, 0, 0) turns the
 * recording off; a new call starts an empty history.  A run that would
 * overflow the history returns QPC_E_ARG before queuing anything.  Not for
 * batch contexts. */
#define QPC_REC_NORM       1
#define QPC_REC_DIPOLE     2    /* <z>                        */
#define QPC_REC_ACCEL      4    /* z acceleration             */
#define QPC_REC_OUTER      8
#define QPC_REC_DIPOLE_XY 16    /* <x>, <y>                   */
#define QPC_REC_ACCEL_XY  32    /* x and y acceleration       */
#define QPC_REC_ALL       63
#define QPC_REC_COLUMNS    8
int  qpc_record(qpc_ctx* ctx, int flags, int64_t cap);
int  qpc_record_read(qpc_ctx* ctx, double* out, int64_t* n);
/* The observables of qpc_observe without any allocation.  No counterpart in
 * the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  Writes the QPC_REC_* observables
 * in `flags` of the current state to out[0..7] (DEVICE memory, 8 doubles:
 * one row of the qpc_record history; 0 where the flag is clear), summed on
 * the device in a fixed order; queued on the context's stream (qpc_sync
 * waits).  The per-channel partial sums live in a static device array of
 * 8 x 65536 doubles (4 MiB per GPU, held until the process ends): at most
 * 65536 channels (QPC_E_SHAPE beyond).  Every context on one GPU shares that
 * array: calls that overlap on one GPU (different streams) race. */
int  qpc_observe_noalloc(qpc_ctx* ctx, int flags, double* out);

/* Projection onto field-free eigenstates, eigenstate_sph
 * (spectrum/population_1e.hpp:7-119).  qpc_upload_projection takes the
 * left-hand basis of initialize (:57-118) for the criterion "the nsel lowest
 * states of every l": W[l][n][ir] = (S v_{n,l})_ir, (lmax+1)*nsel*N doubles,
 * v normalised v^T S v = 1 (qpc_setup_projection_basis builds it).
 * qpc_project is projection (:29-55) on the current state:
 * proj[(im*L+il)*stride + n] = sum_ir W[l][n][ir] * c[im][il][ir] (complex,
 * interleaved), l = il + |m0+im|, n < nsel; entries n in [nsel, stride) are
 * set to 0 (stride = N gives the reference's M*L*N layout). */
int  qpc_upload_projection(qpc_ctx* ctx, int64_t nsel, const double* W);

/* Krylov propagator: propagator_arnoldi (propagator/la_1e.hpp:90-335) with
 * H(t) of propagator_hc_vgwz::multiply (la_1e.hpp:480-497).
 * qpc_upload_krylov takes the field-free bands H_l ((lmax+1)*N*ne, operator_hc
 * data) and the off-diagonals of Y10 and Q10 (M*L each, operator_y10 /
 * operator_q10 data, entry il couples il and il+1); S, its Cholesky factor,
 * 1/r and d/dr come from qpc_upload_radial.
 * qpc_krylov_step is propagate1 (la_1e.hpp:312-335) on the current state:
 * coef <- exp(-i dt S^-1 H(az)) coef in an S-orthogonal Lanczos space of at
 * most kmax vectors, stopping early when the new vector's S-norm falls below
 * 1e-13; *crit receives the criterion of la_1e.hpp:296-310 (0: estimate >
 * emax, 1: in (emin, emax], 2: <= emin), m_used / eps (may be NULL) the
 * subspace size and the estimate |x_{m-1}|^2.  The reference's
 * operator_sc::inverse_all (operators/1e.hpp:1253-1260) is applied as
 * intended, channel by channel. */
int  qpc_upload_krylov(qpc_ctx* ctx, const double* H, const double* y10, const double* q10);
/* spectrum_c::measure_lmd (spectrum/momentum_1e.hpp:384-401) on the
 * current state, the partial photoelectron momentum amplitudes the driver
 * saves as /lmd1 (binsrc/sphTDSE.cpp:367-371): qpc_upload_pcs takes the
 * Coulomb-wave data of spectrum_c_pcs::initialize, ijl[ikr][l][j]
 * (nkr*(lmax+1)*N doubles) and cps[ikr][l] (complex nkr*(lmax+1)), in the
 * reference's layouts (qpc_setup_pcs builds them for a hydrogenic
 * potential); qpc_measure_lmd writes lmd[(ikr*M + im)*L + il] =
 * sqrt(2/pi) cps[ikr][l] sum_j c[im][il][j] ijl[ikr][l][j] (complex), the
 * contraction over j as one FP64 matrix-core GEMM per l (the projection's). */
int  qpc_upload_pcs(qpc_ctx* ctx, int64_t nkr, const double* ijl, const double* cps);
int  qpc_measure_lmd(qpc_ctx* ctx, double* lmd);
/* t-SURFF, spectrum_c_tsf (spectrum/momentum_1e.hpp:531-880), velocity
 * gauge, z polarisation.  qpc_upload_tsurff takes the tables of
 * spectrum_c_tsf::initialize (qpc_setup_tsurff builds them) and clears the
 * accumulated amplitudes and the Volkov clock (clearwork, :580-589).
 * qpc_tsurff(ctx, 1) makes qpc_run call accumulate after every step, before
 * the absorber, as the driver does (binsrc/sphTDSE.cpp:798-804: with
 * A(t_i), A(t_f) during the pulse, the field-free form after it);
 * qpc_tsurff_accumulate is that call on the current state, for callers that
 * step with qpc_step.  One accumulate: time += dt, disz += (az0 + az1) dt / 2,
 * then for every (kr, kt, kp)
 *   work += sqrt(2/pi) dt exp(i (kr^2 time/2 + kr cos(kt) disz))
 *           sum_(m,l) Y_lm(kt) e^(i m kp) (-i)^(l+1) [ -(j0 psi' - jr psi)/2
 *           - i az1 j0 (psi_(l-1) <l-1|cos|l> + psi_(l+1) <l|cos|l+1>) ],
 * psi, psi' = the channel's radial function and derivative at r0 (:609-719).
 * qpc_tsurff_read waits and returns work (complex [nkr][nkt][nkp], the
 * driver's /pmd2), time and disz.  At most 4096 channels (QPC_E_SHAPE);
 * single contexts only; qpc_run_lg refuses a run with t-SURFF on. */
typedef struct {
    double r0, zc;                  /* t-SURFF radius; charge of the Coulomb waves (0: plane waves) */
    double krmin, krmax, ktmin, ktmax, kpmin, kpmax;
    int64_t nkr, nkt, nkp;
    int flag;                       /* spec_axis_type: 0 linear in k, 1 linear in k^2/2 */
} qpc_tsurff_cfg;
int  qpc_upload_tsurff(qpc_ctx* ctx, const qpc_tsurff_cfg* tc, const double* kr, const double* kt,
                       const double* kp, int64_t nb, const int64_t* ib, const double* b0, const double* br,
                       const double* j0, const double* jr, const double* y0, const double* e0,
                       const double* pmat0);
int  qpc_tsurff(qpc_ctx* ctx, int on);
int  qpc_tsurff_accumulate(qpc_ctx* ctx, double az0, double az1);
int  qpc_tsurff_clear(qpc_ctx* ctx);
int  qpc_tsurff_read(qpc_ctx* ctx, double* work, double* time, double* disz);
/* Elliptical polarisation, accumulate (xy) (momentum_1e.hpp:716-776):
 * qpc_upload_tsurff_xy adds pmat1, pmat2 (M*L doubles each, qpc_setup_tsurff_xy
 * builds them) to the tables of qpc_upload_tsurff; with t-SURFF on, qpc_run_ep
 * accumulates after every step (binsrc/sphTDSE.cpp:939-945: A_x, A_y at the
 * step's ends during the pulse, the field-free form after it).
 * qpc_tsurff_accumulate_xy is that call on the current state: time += dt,
 * disx (disy) += (ax0 + ax1) dt / 2 ((ay0 + ay1) dt / 2), and the channel term
 * -(j0 psi' - jr psi)/2 - i j0 (A.r psi) with the A.r couplings of the
 * (m -+ 1, l -+ 1) neighbours (sin e^(+-i phi) elements, A_x(t_f), A_y(t_f)),
 * phase k^2 time/2 + k sin(theta) (cos(phi) disx + sin(phi) disy).  The
 * length-gauge run refuses t-SURFF (sphTDSE.cpp:288).
 * qpc_tsurff_displacement returns {disx, disy, disz} (int A dt so far). */
int  qpc_upload_tsurff_xy(qpc_ctx* ctx, const double* pmat1, const double* pmat2);
int  qpc_tsurff_accumulate_xy(qpc_ctx* ctx, double ax0, double ax1, double ay0, double ay1);
int  qpc_tsurff_displacement(qpc_ctx* ctx, double* dis);
int  qpc_krylov_step(qpc_ctx* ctx, double az, double dt, int64_t kmax, double emin, double emax, int* crit,
                     int64_t* m_used, double* eps);
int  qpc_project(qpc_ctx* ctx, double* proj, int64_t stride);

/* Adaptive Krylov step.  No counterpart in the reference (propagate1 always
 * builds kmax vectors): the Lanczos space of qpc_krylov_step grows from
 * kstart vectors until the a-posteriori estimate |x_{m-1}|^2 of
 * la_1e.hpp:296-310 is <= tol, or kmax.  m_used / eps (may be NULL) get the
 * size and its estimate; ab (may be NULL, 2*kmax doubles) the Lanczos
 * tridiagonal a_i, b_{i+1} (i < m). */
int  qpc_krylov_step_adaptive(qpc_ctx* ctx, double az, double dt, int64_t kstart, int64_t kmax, double tol,
                              int64_t* m_used, double* eps, double* ab);

/* Krylov time propagation with a learned controller.  No counterpart in the
 * reference.  A(t) is field_component_sinen<2> (field.h:90-133):
 * A(t) = f0 sin(w0 t + pi c0) sin^2(pi (t - t_i)/tau), t_i = n0 T, tau = nc T,
 * T = 2 pi / w0, zero outside (t_i, t_i + tau). */
typedef struct { double f0, w0, nc, n0, c0; } qpc_sinen2;
typedef struct qpc_mlp qpc_mlp;  /* include/qpc_mlp.h */
typedef struct {
    int64_t steps;          /* accepted steps                                   */
    int64_t rejected;       /* steps redone at half the size (estimate > tol at kmax) */
    int64_t exponentials;   /* Krylov exponentials computed (accepted and rejected) */
    int64_t vectors;        /* Krylov vectors built in all of them              */
    double  worst_estimate; /* largest accepted estimate                        */
} qpc_krylov_stats;
/* Optional per-step record.  With ab (training data of the controller): for
 * step s < cap, features[s][7] (the network inputs below), h[s] the step size,
 * m[s] and ab[s][2*kmax] the Lanczos tridiagonal of the step's first
 * exponential, which then runs to kmax vectors (no early stop).  With ab ==
 * NULL: the accepted steps as taken, features[s], h[s] and the subspace size
 * m[s] of the first exponential, the propagation unchanged by the record.
 * n receives the count. */
typedef struct {
    int64_t cap, n;
    float* features;
    double* h;
    int64_t* m;
    double* ab;
} qpc_krylov_record;
/* Propagates the current state from t0 to t1 with 4th-order commutator-free
 * Magnus steps (Gauss points t + (1/2 -+ sqrt(3)/6) h; two exponentials per
 * step, exp(-i h/2 H(a)) with a = 2(a2 A1 + a1 A2) then 2(a1 A1 + a2 A2),
 * a1,2 = (3 -+ 2 sqrt(3))/12), each an adaptive Krylov step with tolerance
 * tol.  net == NULL: step h = dt_base, kstart = kmin.  Otherwise the network
 * (qpc_mlp.h) maps the features [A(t), 10 E(t), the estimates of the last 4
 * steps as log10(e)/10 + 1, log10(outer norm)/10] to (y0, y1):
 * kstart = clamp(round(8 + 32 y0), kmin, kmax), h = dt_base 2^clamp(y1,-1,1).
 * Either way a step whose estimate exceeds tol at kmax is undone and redone
 * at half the size, so the estimate decides; QPC_E_NUMERIC when the step
 * would fall below dt_base/256 (the estimate has a floor above tol: a state
 * with components the Lanczos space cannot resolve).  Needs the Krylov
 * operators and, with a network, the outer band. */
int  qpc_krylov_propagate(qpc_ctx* ctx, const qpc_sinen2* pulse, double t0, double t1, double dt_base, double tol,
                          int64_t kmin, int64_t kmax, qpc_mlp* net, qpc_krylov_stats* stats,
                          qpc_krylov_record* rec);

/* Batched shots.  No counterpart in the reference (it runs one shot per
 * process).  n_shots independent z-polarised runs of one radial grid (dims:
 * N, n_rank, dt; n_mdim 1, n_mmin 0; n_ldim = L, the largest channel count
 * of the batch) in one context: every launch covers the channels of all
 * shots.  Shot s owns channels [s*L, (s+1)*L) of the state (the reference's
 * layout per shot); a shot with fewer channels L_s is padded with decoupled
 * channels: zero eigenvalues, identity rows in its projectors, zero state.
 * qpc_upload_angular takes n_shots*L eigenvalues and n_shots*L*L projectors
 * (one block per shot); qpc_upload_hc the bands of l = 0..L-1.
 * time_batch (2..32) is the number of time points factorised ahead per table
 * set (memory: 2 sets x time_batch x n_shots*L tables).
 * qpc_run_batch is qpc_run with one A(t) row per shot: A of shot s at time
 * point it is atab[s*stride + it] (stride >= nt_field + 1); with-field steps
 * while it < nt_field (a shot whose pulse has ended carries A = 0: its step
 * then equals the field-free step to rounding), field-free afterwards.
 * qpc_observe_batch writes one value per shot.  qpc_step, the EP, Krylov and
 * projection entry points reject a batch context (QPC_E_ARG). */
int  qpc_create_batch(const qpc_dims* dims, int64_t n_shots, int64_t time_batch, qpc_ctx** out);
int  qpc_run_batch(qpc_ctx* ctx, const double* atab, int64_t stride, int64_t nt_field, int64_t it0, int64_t it1,
                   int use_mask);
int  qpc_observe_batch(qpc_ctx* ctx, int what, double* out);

/* Elliptically polarised (x-y) field: propagator_cn_hc_vgwr
 * (propagator/cn_1e.hpp:812-951), selected for gauge 0 with n_field_x and
 * n_field_y (binsrc/sphTDSE.cpp:574-586).  All M*L = ML channels of the
 * rectangular (m, l) grid are coupled by the angular operators.
 * qpc_upload_angular_ep takes the angular eigen-data of its initialize
 * (cn_1e.hpp:878-895) in the reference's layouts:
 *   isub          the sub-block selector of propagate<isub>
 *                 (dimension/1e.hpp:57-80): -1 (or 0, 1) all channels,
 *                 2 gathered even (propflag_eve), 3 gathered odd;
 *   eig_x1 .. y2  ML doubles each: prop_x1.zeig (Y11, mult 1), prop_x2.zeig
 *                 (Q11, mult -i), prop_y1.zeig (Y11, mult -i), prop_y2.zeig
 *                 (Q11, mult 1), in converter::catagorize order
 *                 (utilities/converter.hpp:27-107: even, zero, odd);
 *   tran_x1, tran_y2           ML*ML doubles, row-major;
 *   x1_2_x2, x2_2_y1, y1_2_y2  ML*ML complex, row-major (merge_two).
 * The radial operators, field-free bands and (for the absorber) the mask
 * come from qpc_upload_radial / qpc_upload_hc as for the z-polarised step.
 * qpc_step_ep is propagate<isub>(coef, ax_i, ax_f, ay_i, ay_f)
 * (cn_1e.hpp:898-945); qpc_run_ep the time loop of the EP task
 * (binsrc/sphTDSE.cpp:921-951): with-field steps with
 * (axtab[it], axtab[it+1], aytab[it], aytab[it+1]) while it < nt_field,
 * field-free steps afterwards, the absorber after each step when use_mask;
 * it returns when the work is queued (qpc_sync waits).  The eight coupling
 * sweeps of a step are factorised on a second stream one step ahead.
 * QPC_E_NUMERIC when A(t)*dt leaves the range of the D-part sweeps
 * (as qpc_step). */
int  qpc_upload_angular_ep(qpc_ctx* ctx, int isub, const double* eig_x1, const double* eig_x2,
                           const double* eig_y1, const double* eig_y2, const double* tran_x1,
                           const double* tran_y2, const double* x1_2_x2, const double* x2_2_y1,
                           const double* y1_2_y2);
int  qpc_step_ep(qpc_ctx* ctx, double ax_i, double ax_f, double ay_i, double ay_f);
int  qpc_run_ep(qpc_ctx* ctx, const double* axtab, const double* aytab, int64_t nt_field, int64_t it0,
                int64_t it1, int use_mask);

/* The same field in the length gauge: propagator_cn_hc_lgwr
 * (propagator/cn_1e.hpp:717-810), selected for gauge 1 with n_field_x and
 * n_field_y (binsrc/sphTDSE.cpp:585).  Two coupling sweeps S + i e r per
 * half step (prop_x1 on Y11, prop_y1 on Y11 with multiplier -i, both with
 * <Bi|r|Bj>) around the field-free solve.  qpc_upload_angular_lgwr takes the
 * data of its initialize (cn_1e.hpp:745-765), isub as above:
 *   eig_x1, eig_y1   ML doubles, prop_x1.zeig and prop_y1.zeig
 *                    (catagorize order);
 *   tran_x1          ML*ML doubles (copy_real of the conjugated x1 projector);
 *   x1_2_y1          ML*ML complex, U[y1] U[x1]^H (merge_two);
 *   tran_y1          ML*ML complex, U[y1] (the conjugated y1 projector).
 * The radial band <Bi|r|Bj> must have been uploaded (qpc_upload_radial's
 * rpow1).  qpc_step_lgwr is propagate<isub>(coef, ex_i, ex_f, ey_i, ey_f)
 * (cn_1e.hpp:767-803); qpc_run_lgwr the time loop of the EP task with E(t)
 * in place of A(t) (binsrc/sphTDSE.cpp:926-951), queued as qpc_run_ep's. */
int  qpc_upload_angular_lgwr(qpc_ctx* ctx, int isub, const double* eig_x1, const double* eig_y1,
                             const double* tran_x1, const double* x1_2_y1, const double* tran_y1);
/* The angular eigen-data of either elliptical step built on the device, in
 * place of qpc_setup_angular_ep / _lgwr + qpc_upload_angular_ep / _lgwr
 * (operator_y11/q11::solve_eigen, operators/1e.hpp:336-523, with
 * converter::catagorize and the transforms of cn_1e.hpp:878-895, 745-765):
 * the parity blocks of each angular operator are diagonalised by rocSOLVER
 * (dsyevd) and the merged products formed on the FP64 matrix cores.  As for
 * the host set-up, bases inside degenerate eigenspaces may differ from
 * zhbev's; the steps depend only on the spectral projectors.  Synchronous;
 * temporary device memory about 6 (ML)^2 complex. */
int  qpc_build_angular_ep(qpc_ctx* ctx, int isub);
int  qpc_build_angular_lgwr(qpc_ctx* ctx, int isub);
int  qpc_step_lgwr(qpc_ctx* ctx, double ex_i, double ex_f, double ey_i, double ey_f);
int  qpc_run_lgwr(qpc_ctx* ctx, const double* extab, const double* eytab, int64_t nt_field, int64_t it0,
                  int64_t it1, int use_mask);


/* Launch-shape autotuner.  No counterpart in the reference (its loops have no
 * launch shape).  The partitioned solves cut each channel's partitions into
 * segments, one work-group each; seg_add adds segments beyond the fewest
 * that hold them (more work-groups, each with the carry's halo partitions
 * again), and grids of at most stage_grid work-groups stage the transfer sets
 * in LDS with four waves per work-group.  Neither changes the result (the
 * segments are exact, the two paths share their arithmetic).
 * qpc_autotune runs the with-field steps [it0, it0+nsteps) of the field
 * table (as qpc_run, or qpc_run_batch with its stride for a batch context;
 * stride is ignored otherwise) once per candidate, seg_add 0..3 x stage_grid
 * 512 and 4096 (ms[k]: seg_add k % 4, stage_grid k / 4), from the current
 * state, which it restores, keeps the fastest in the context (it applies to
 * every later launch, the elliptical sweeps too) and returns it in best.
 * Needs the absorber band (the runs apply it); observables and t-SURFF are
 * off during the trials.  ms and best may be NULL. */
typedef struct {
    int32_t seg_add;
    int32_t serial;      /* 1: a run's steps wait for each table batch as soon as it is queued */
    int64_t stage_grid;
    int32_t share_r;     /* 1 (default): channels whose R-part eigenvalues have the same magnitude
                            share one table (conjugated for the opposite sign); 0: one table each */
    int32_t tables_r;    /* qpc_get_tune: R-part tables factorised per time point */
} qpc_tune;
int  qpc_set_tune(qpc_ctx* ctx, const qpc_tune* t);
int  qpc_get_tune(qpc_ctx* ctx, qpc_tune* t);
int  qpc_autotune(qpc_ctx* ctx, const double* atab, int64_t stride, int64_t nt_field, int64_t it0, int64_t nsteps,
                  qpc_tune* best, double* ms);

/* Kernel timing with HIP events on the stream the kernel runs on.
 * qpc_timing_begin arms event pairs around every launch of kernel class
 * `which` (QPC_K_*) queued afterwards, up to max_launches; qpc_timing_end
 * waits and returns the number of timed launches and their mean duration in
 * ms.  The solve classes time the right-hand side B z and the partitioned
 * triangular solve of one CN factor: one launch (k_bsolve forming B z while
 * it stages) on grids of at most 512 work-groups with distinct input and
 * output, else the banded product (k_rhs*) followed by the solve. */
#define QPC_K_CN_D     0   /* CN factor of a coupling step, D-part (prop_z2) */
#define QPC_K_CN_R     1   /* CN factor of a coupling step, R-part (prop_z1) */
#define QPC_K_FREE     2   /* field-free CN factor (prop_hc)         */
#define QPC_K_ROT      3   /* angular rotations (tran_*, z2_2_z1)     */
#define QPC_K_MASK     4   /* absorber (mask product and S solve)     */
#define QPC_K_FACTOR   5   /* batch of coupling-sweep factorisations (second stream) */
#define QPC_K_PROJ     6   /* projection onto field-free eigenstates   */
#define QPC_K_KRYLOV   7   /* one Krylov propagation step (all its launches) */
#define QPC_K_EP_ROT   8   /* one (m,l) rotation GEMM of the elliptically polarised step */
#define QPC_K_EP_SWEEP 9   /* one coupling sweep of the elliptically polarised step (in place: product, then solve) */
int  qpc_timing_begin(qpc_ctx* ctx, int which, int64_t max_launches);
int  qpc_timing_end(qpc_ctx* ctx, int64_t* launches, double* mean_ms);

/* Plan of the coupling sweeps' solves (host only, no device work).  No
 * counterpart in the reference, whose sweeps run sequentially over all rows
 * (propagate.hpp:3-98); exposed for tests and for callers that want to see
 * what the library will run.  Bands row-major N*(n_rank+1) (reference
 * layout).  qpc_coupling_reach: for A = S + i e X (asym 0, X symmetric) or
 * S + e X (asym 1, X antisymmetric), the rows a unit source takes to fall
 * below 1e-17 of its peak through L^-1 (*fwd) and U^-1 (*bwd) of the LU of A
 * without pivoting, measured on windows of up to 900 rows (a response that
 * has not decayed within the window reports the window's reach).
 * qpc_coupling_plan: the D-part (X = d/dr) carry depth *pk in partitions of
 * 24 rows (above 28: the sequential solve) and factorisation restart
 * *restart in rows for scalars |e| <= emax. */
int  qpc_coupling_reach(const double* S, const double* X, int64_t N, int64_t n_rank, double e, int asym,
                        int* fwd, int* bwd);
int  qpc_coupling_plan(const double* S, const double* D, int64_t N, int64_t n_rank, double emax, int* pk,
                       int* restart);

/* ------------------------------------------------------------------------
 * Host-side set-up (runs once, no device work).  Builds the operands above
 * from the config fields of the reference's config/<task>.cpp for a
 * hydrogenic atom: the host half of propagator_cn_hc_vgwz::initialize
 * (propagator/cn_1e.hpp:642-680), absorber_mask::initialize
 * (absorber.hpp:42-49) and coefficient::initialize (coefficient/1e.hpp:89-143).
 * A caller that already has the reference's own set-up uploads that instead.
 * ------------------------------------------------------------------------ */
typedef struct {
    qpc_dims dims;
    double para_rmin, para_rmax;   /* radial box (linear knots, basis_bsplines.h:450-459) */
    double para_zc;                /* hydrogenic charge (potential.hpp:9-17)             */
    double para_me;                /* mass                                                */
    double mask_rmin, mask_rmax, mask_fact;   /* cos^fact absorber (absorber.hpp:42-49)   */
} qpc_config;

typedef struct qpc_setup qpc_setup;

#define QPC_SETUP_S       0   /* N*ne            */
#define QPC_SETUP_CHOL    1   /* N*ne            */
#define QPC_SETUP_H       2   /* (lmax+1)*N*ne   */
#define QPC_SETUP_RINV    3   /* N*ne            */
#define QPC_SETUP_RDIF    4   /* N*ne            */
#define QPC_SETUP_RPOW1   5   /* N*ne            */
#define QPC_SETUP_MASK    6   /* N*ne            */
#define QPC_SETUP_UPPER   7   /* complex (lmax+1)*N*ne */
#define QPC_SETUP_LOWER   8   /* complex (lmax+1)*N*ne */
#define QPC_SETUP_YEIG    9   /* M*L             */
#define QPC_SETUP_YPRJ   10   /* M*L*L           */
#define QPC_SETUP_QEIG   11   /* M*L             */
#define QPC_SETUP_QPRJ   12   /* complex M*L*L   */
#define QPC_SETUP_MERGED 13   /* complex M*L*L   */
#define QPC_SETUP_Y10    14   /* M*L             */
#define QPC_SETUP_DVDR   15   /* N*ne  <Bi|Z/r^2|Bj> */
#define QPC_SETUP_Q10    16   /* M*L   off-diagonal of Q10 (operator_q10 data) */
#define QPC_SETUP_SOUTER 17   /* N*ne  <Bi|theta(r - mask_rmin)|Bj>, overlap of the absorbing region */
#define QPC_SETUP_KNOTS  18   /* N+n_rank+1: knots of the active intervals (the driver's /knot_sequence) */
/* the l-independent parts of integrate_hc (basisfunc.hpp:92-106), the
 * operands of qpc_build_hc: <Bi|V|Bj>, <Bi|1/r^2|Bj>, <Bi'|Bj'> */
#define QPC_SETUP_HV     19   /* N*ne            */
#define QPC_SETUP_HK     20   /* N*ne            */
#define QPC_SETUP_HT     21   /* N*ne            */

int     qpc_setup_build(const qpc_config* cfg, qpc_setup** out);
void    qpc_setup_free(qpc_setup* st);
int64_t qpc_setup_size(const qpc_setup* st, int what);          /* number of doubles */
int     qpc_setup_get(const qpc_setup* st, int what, double* dst);
int     qpc_setup_lu_info(const qpc_setup* st);                 /* tiny-pivot count */
/* Eigenvector `in` (ascending energy) of the (H_l, S) pencil of channel
 * (im,il), normalised v^T S v = 1 with its largest component positive, as a
 * full M*L*N complex coefficient array (coefficient::initialize, cf = 1). */
int     qpc_setup_eigenstate(const qpc_setup* st, int64_t im, int64_t il, int64_t in, double* energy,
                             double* coef);
/* Left-hand basis of eigenstate_sph::initialize (population_1e.hpp:57-118)
 * for the nsel lowest states of every l = 0..lmax: W[l][n][ir] = (S v_{n,l})_ir
 * ((lmax+1)*nsel*N doubles), E[l][n] the energies ((lmax+1)*nsel, may be
 * NULL).  v^T S v = 1, largest component positive (the reference keeps the
 * sign LAPACKE_dsbgv returns). */
int     qpc_setup_projection_basis(const qpc_setup* st, int64_t nsel, double* W, double* E);
/* Angular eigen-data of propagator_cn_hc_vgwr for an M x L grid from m0
 * (operator_y11/q11::solve_eigen, operators/1e.hpp:336-523, with
 * converter::catagorize, and the transforms of cn_1e.hpp:878-895), in the
 * layouts of qpc_upload_angular_ep.  The Hermitian angular matrices conserve
 * the parity of il; each parity block is diagonalised on its own as a real
 * symmetric matrix (the two with multiplier -i after the diagonal phase
 * change diag(i^-im)).  Within a degenerate eigenspace the basis may differ
 * from zhbev's; every step result depends only on the spectral projectors.
 * Output sizes: eig_* ML, tran_* ML*ML doubles, merged ML*ML complex. */
int     qpc_setup_angular_ep(int64_t M, int64_t L, int64_t m0, double* eig_x1, double* eig_x2, double* eig_y1,
                             double* eig_y2, double* tran_x1, double* tran_y2, double* x1_2_x2,
                             double* x2_2_y1, double* y1_2_y2);
/* Angular data of propagator_cn_hc_lgwr (cn_1e.hpp:745-765) in the layouts of
 * qpc_upload_angular_lgwr: eig_x1, eig_y1 ML doubles, tran_x1 ML*ML doubles,
 * x1_2_y1 and tran_y1 ML*ML complex.  The same decompositions as
 * qpc_setup_angular_ep. */
int     qpc_setup_angular_lgwr(int64_t M, int64_t L, int64_t m0, double* eig_x1, double* eig_y1,
                               double* tran_x1, double* x1_2_y1, double* tran_y1);
/* Photoelectron spectra by projection onto Coulomb waves, spectrum_c_pcs
 * (spectrum/momentum_1e.hpp:276-521) for a hydrogenic{zc} potential.
 * qpc_setup_pcs is its initialize with initialize_kr (:119-148, :455-486):
 * kr[nkr] linear on [krmin, krmax] (set_sequence_lin,
 * spectrum/inventory.h:15-26); ijl[ikr][l][j] = int B_j(r) F_l(eta, kr r) dr /
 * kr with eta = -zc/kr, F_l the regular Coulomb function (times
 * 1/(1 + exp(-5(r - rb))) when rb > 0, the plane-wave window), l = 0..lmax,
 * nkr*(lmax+1)*N doubles; cps[ikr][l] = exp(i(sigma_l(eta) - l pi/2))
 * (complex, nkr*(lmax+1)), sigma_l = arg Gamma(l+1+i eta).  The Coulomb
 * functions follow Steed's continued-fraction method (qpc_coulomb_f) on the
 * set-up's Gauss points, threaded over kr.  qpc_coulomb_f / qpc_coulomb_sigma
 * expose F_l(eta, rho) for l = 0..lmax and sigma_l (tests). */
int     qpc_setup_pcs(const qpc_setup* st, int64_t nkr, double krmin, double krmax, double zc, double rb,
                      double* kr, double* ijl, double* cps);
int     qpc_coulomb_f(int64_t lmax, double eta, double rho, double* F);
/* t-SURFF tables of spectrum_c_tsf::initialize (momentum_1e.hpp:779-880)
 * for qpc_upload_tsurff: axes kr[nkr] (flag 0: linear in k; 1: linear in
 * E = k^2/2), kt[nkt], kp[nkp]; the nb (<= n_rank+1) non-zero B-splines at
 * r0: indices ib, values b0, derivatives br; j0/jr[nkr][lmax+1] =
 * F_l(-zc/k, k r0)/k and F_l'; y0[nkt][M][L] = Y_lm(theta, 0);
 * e0[nkp][M] complex exp(i m phi); pmat0[M][L] = <l m|cos|l+1 m>. */
int     qpc_setup_tsurff(const qpc_setup* st, const qpc_tsurff_cfg* tc, double* kr, double* kt, double* kp,
                         int64_t* nb, int64_t* ib, double* b0, double* br, double* j0, double* jr, double* y0,
                         double* e0, double* pmat0);
int     qpc_coulomb_sigma(int64_t lmax, double eta, double* sig);
/* pmat1, pmat2 [M][L] of spectrum_c_tsf::initialize (momentum_1e.hpp:876-877):
 * <l m|sin(theta) e^(-i phi)|l+1 m+1> (integrate_sinexp1) and
 * <l m|sin(theta) e^(-i phi)|l-1 m+1> (integrate_sinexp2, basisfunc.hpp:241-259)
 * for qpc_upload_tsurff_xy. */
int     qpc_setup_tsurff_xy(const qpc_setup* st, double* pmat1, double* pmat2);
/* Uploads every operand of `st` into `ctx` (shapes must agree). */
int     qpc_upload_setup(qpc_ctx* ctx, const qpc_setup* st);

#ifdef __cplusplus
}
#endif
#endif
