// qpc_ctx.h -- internal: the propagator context shared by the C ABI sources
// (qpc_gpu.hip: set-up, z-polarised step, Krylov, projection; qpc_ep.hip:
// the elliptically polarised step) and their common host helpers.
#pragma once
#include "../../include/qpc_gpu.h"
#include "qpc_common.h"
#include "qpc_kernels.h"
#include "qpc_solve.h"


#include <string>
#include <vector>

using qpc::cplx;

// time points per batch of coupling-sweep factorisations
static constexpr int64_t kTimeBatch = 32;

struct qpc_ctx {
    qpc_dims d{};
    int64_t N = 0, p = 0, ne = 0, L = 0, M = 0, m0 = 0, lmax = 0, ML = 0;
    hipStream_t stream = nullptr;   // the step sequence
    hipStream_t fstream = nullptr;  // coupling-sweep factorisations, one batch ahead
    // the D-part factorisation of a per-step call, beside the R part's on fstream
    hipStream_t dstream = nullptr;
    hipEvent_t put_done = nullptr;  // the per-step scalars are on the device (fstream)
    int device = 0;
    std::string err;

    // radial operators (real bands N*ne, row-major as the reference stores them)
    double *S = nullptr, *chol = nullptr, *rinv = nullptr, *rdif = nullptr, *rpow1 = nullptr, *mask = nullptr;
    // diagonal-major copies read by the right-hand sides of the solves
    double *Sd = nullptr, *rinvd = nullptr, *rdifd = nullptr, *maskd = nullptr, *rpow1d = nullptr;
    // R-part tables shared by channels whose eigenvalues have the same
    // magnitude (the table of S + i e 1/r at -e is the complex conjugate of
    // the one at e, and none depends on m): rmap[ch] = table | kMapConj,
    // qeig_prim the n_rprim distinct eigenvalues factorised; n_rprim 0: every
    // channel its own table (also for batch contexts); pair_r switches it
    int* rmap = nullptr;
    double* qeig_prim = nullptr;
    int64_t n_rprim = 0;
    bool pair_r = true;
    std::vector<int64_t> rprim_ch;  // the channel each shared table is factorised for
    // the R-part solve's work-group order that keeps the channels of a shared
    // table on one XCD (qpc::solve_order), for rorder_nseg segments per channel
    std::vector<int> rtab_h, rorder_h;
    int* rorder = nullptr;
    int rorder_nseg = 0;
    // the same for the D-part and field-free solves (one table per channel,
    // per l): a channel's segments, and channels of one l, on one XCD
    struct WgOrder {
        std::vector<int> h;
        int* d = nullptr;
        int nseg = 0;
    } dorder, forder;
    // batch contexts (tables shared within a shot): the R-part scalars of
    // those channels per time point, d_eprim[t][n_rprim]
    double* d_eprim = nullptr;
    int64_t eprim_cap = 0;
    // {S, X} pairs of the coupling sweeps' right-hand sides, X = 1/r, d/dr, r
    cplx *sxR = nullptr, *sxD = nullptr, *sxL = nullptr;
    // t-SURFF (qpc_upload_tsurff): tables on the device, the accumulated
    // amplitudes work[kr][kt][kp], and the Volkov clock (time, int A dt)
    bool have_tsf = false, tsf_on = false, have_tsf_xy = false;
    int tsf_nkr = 0, tsf_nktp = 0, tsf_nb = 0;
    double tsf_time = 0.0, tsf_disz = 0.0, tsf_disx = 0.0, tsf_disy = 0.0;
    int* tsf_ib = nullptr;
    double *tsf_b0 = nullptr, *tsf_br = nullptr, *tsf_kr = nullptr, *tsf_ckt = nullptr, *tsf_j0 = nullptr,
           *tsf_jr = nullptr, *tsf_pm = nullptr;
    // elliptical polarisation: sin(theta) cos/sin(phi) per point, the
    // sin e^(-i phi) elements (qpc_upload_tsurff_xy)
    double *tsf_stx = nullptr, *tsf_sty = nullptr, *tsf_pm1 = nullptr, *tsf_pm2 = nullptr;
    cplx *tsf_Y = nullptr, *tsf_tmp = nullptr, *tsf_work = nullptr;
    // cn_component_hc bands (complex, (lmax+1)*N*ne) and the diagonal-major upper bands
    cplx *upper = nullptr, *lower = nullptr, *upperd = nullptr;
    // angular data
    double *yeig = nullptr, *qeig = nullptr, *yprj = nullptr, *y10 = nullptr;
    cplx *qprj = nullptr, *merged = nullptr;
    // prepared rotation operators P' (qpc_kernels.h): P_Y, P_Y^T; merged, merged^H; P_Q, P_Q^H
    double *rtY = nullptr, *rtYt = nullptr;
    cplx *rtM = nullptr, *rtMh = nullptr, *rtQ = nullptr, *rtQh = nullptr;
    // state ("coef.data") and cn_workspace_so::tmpc
    cplx *coef = nullptr, *tmpc = nullptr;
    // third state buffer of the out-of-place solves, and the right-hand sides B z
    cplx *spare = nullptr, *vbuf = nullptr;
    // solve tables (qpc_solve.h): field-free per l, absorber, and a batch of
    // coupling-sweep factorisations for kTB time points
    qpc::Geom geo{};
    cplx *tab_free = nullptr, *trf_free = nullptr;
    double *tab_mask = nullptr, *trf_mask = nullptr;
    int64_t tb_cap = 0;  // time points per batch
    // two table sets: batch k+1 is factorised (fstream) while batch k steps (stream)
    struct TableSet {
        double *tabD = nullptr, *trfD = nullptr;
        cplx *tabR = nullptr, *trfR = nullptr;
        hipEvent_t ready = nullptr;  // factorisation done (fstream)
        hipEvent_t ready_r = nullptr;  // its R-part (length gauge: all) tables and transfers done
        hipEvent_t freed = nullptr;  // last step that reads the set issued (stream)
        bool pending = false;        // freed has been recorded
    } ts[2];
    // per-time-point scalars of the current run: host copy (solve right-hand
    // sides) and device copy (factorisations), D-part and R-part
    std::vector<double> h_dD, h_dR;
    // the per-step entry points keep their last A(t_f) (or E(t_f)): its tables
    // are at position step_slot (0..2) of table set 0 (-1: none), its value
    // step_a; slot_read[k] marks the end of the last step that read position k
    int64_t step_slot = -1;
    double step_a = 0.0;
    int step_gauge = 0;
    hipEvent_t slot_read[3] = {nullptr, nullptr, nullptr};
    bool slot_rec[3] = {false, false, false};
    double yeig_max = 0.0;  // largest |eigenvalue| of the D-part rotation
    // host copies of S, d/dr and r (row-major bands) for the sweep plans
    std::vector<double> S_h, rdif_h, rpow1_h;
    // D-part sweeps: carry depth (partitions; > qpc::kPKMax runs the
    // sequential solve) and factorisation restart (rows), valid for scalars
    // |e| <= plan_emax (qpc::d_part_plan); the same for the length-gauge
    // sweeps S + i e r (pk_l, restart_l, plan_emax_l)
    int pk_d = qpc::kPKD, restart_d = qpc::kRestart;
    double plan_emax = -1.0;
    int pk_l = qpc::kPKS, restart_l = qpc::kRestart;
    double plan_emax_l = -1.0;
    // gauge of the run being queued: 0 velocity (vgwz), 1 length (lgwz)
    int gauge = 0;
    double* d_ifdt = nullptr;  // [2][run_cap]
    int64_t run_cap = 0;
    // reductions
    double* red = nullptr;
    // projection onto field-free eigenstates (eigenstate_sph)
    qpc::ProjGeom pg{};
    int64_t proj_nsel = 0;
    double *Wp = nullptr, *ppart = nullptr;
    int *pchan = nullptr, *pcnt = nullptr;
    cplx* pout = nullptr;
    // photoelectron spectra (spectrum_c_pcs::measure_lmd): the Coulomb-wave
    // integrals in GEMM operand order, the phase factors, the GEMM's split-K
    // partials and outputs
    qpc::ProjGeom sg{};
    int64_t pcs_nkr = 0;
    double *pcs_Wp = nullptr, *pcs_part = nullptr;
    int *pcs_chan = nullptr, *pcs_cnt = nullptr;
    cplx *pcs_cps = nullptr, *pcs_out = nullptr, *pcs_lmd = nullptr;
    // Krylov propagator (propagator_arnoldi): H_l diagonal-major, Q10
    // off-diagonal, Krylov vectors U[kry_cap][ML*N], dot partials, coefficients
    double *Hd = nullptr, *q10 = nullptr;
    double* souter = nullptr;  // overlap band of the outer region (QPC_OBS_OUTER)
    double* accel = nullptr;   // <Bi|dV/dr|Bj> of the dipole acceleration (QPC_OBS_ACCEL_Z)
    bool have_accel = false;
    // observables recorded before every step of the time loops (qpc_record):
    // hist[row][qpc::kObs] = norm, <z>, z acceleration, outer norm, <x>, <y>, x and
    // y acceleration; red4 the per-channel partials
    double* obs_tmp = nullptr;  // partials and sums of qpc_observe's observable_r kinds
    int rec_flags = 0;
    int64_t rec_cap = 0, rec_n = 0;
    double *hist = nullptr, *red4 = nullptr;
    // batched shots (qpc_create_batch): the M blocks are independent shots with
    // m = 0, each padded to L channels; the coupling-sweep scalars differ per
    // shot, so they are kept per channel: d_e[2][run_cap][ML] (D-part, R-part)
    bool batch = false;
    std::vector<double> yeig_h, qeig_h;  // host copies of the angular eigenvalues
    double *d_e = nullptr, *d_ones = nullptr;
    std::vector<double> h_e;             // host copy of d_e
    bool have_outer = false;
    cplx *kU = nullptr, *kpart = nullptr, *kcoef = nullptr;
    int64_t kry_cap = 0;
    bool have_krylov = false;
    // elliptically polarised step (propagator_cn_hc_vgwr, qpc_ep.hip)
    struct EpData {
        int isub = -1;
        int64_t s0 = 0, s1 = 0, p0 = 0, p1 = 0;  // rotation (msub) and sweep (mpro) channel ranges
        double* eig = nullptr;                    // [4][ML] zeig of x1, x2, y1, y2
        double emax[4] = {0, 0, 0, 0};            // largest |zeig| per operator
        double *tx1 = nullptr, *ty2 = nullptr;    // ML*ML real (tran_x1, tran_y2)
        cplx *x1x2 = nullptr, *x2y1 = nullptr, *y1y2 = nullptr;  // ML*ML complex (merge_two)
        // length gauge (propagator_cn_hc_lgwr): eig x1 and y1 in slots 0 and 2,
        // tran_x1 in tx1, x1_2_y1 in x1x2, the complex tran_y1 here; its four
        // sweeps (x1, y1) x (t_i, t_f) are R-type tables of S + i e r
        bool lg = false;
        cplx* ty1 = nullptr;
        // the natural channels in parity order (il even first, then odd, each
        // ascending): perm[k] = natural channel of position k, lmap[k] its l,
        // n_ev the even ones; ty2p / ty1p = tran_y2 / tran_y1 with their
        // columns in that order, so that the rotations around the field-free
        // solve skip the other parity's zero block
        int64_t n_ev = 0;
        int *perm = nullptr, *lmap = nullptr;
        // sweeps of S + i e X (the 1/r sweeps x2, y2; in the length gauge x1,
        // y1 with r) whose rows [p0, p1) hold eigenvalue pairs +-e share one
        // table per |e|: rmap[op][row - p0] = table | kMapConj, eprim[op] the
        // nprim[op] values factorised (rmap null: every row its own table)
        std::vector<double> eig_h[4];
        std::vector<int> rtab_h[4];  // table of each swept row (host)
        // the sweeps' work-group order (qpc::solve_order) per operator, for
        // ord_nseg segments per row
        std::vector<int> ord_h[4];
        int* ord[4] = {nullptr, nullptr, nullptr, nullptr};
        int ord_nseg[4] = {0, 0, 0, 0};
        int* rmap[4] = {nullptr, nullptr, nullptr, nullptr};
        double* eprim[4] = {nullptr, nullptr, nullptr, nullptr};
        int64_t nprim[4] = {0, 0, 0, 0};
        double* ty2p = nullptr;
        cplx* ty1p = nullptr;
        // tables of the eight sweeps of one step: D (x1, y1) x (t_i, t_f), R (x2, y2) x (t_i, t_f)
        struct Set {
            double *tabD = nullptr, *trfD = nullptr;
            cplx *tabR = nullptr, *trfR = nullptr;
            hipEvent_t ready = nullptr, freed = nullptr;
            bool pending = false;
        } set[2];
        std::vector<double> h_sc;    // per step: [x1 ti, x1 tf, y1 ti, y1 tf, x2 ti, x2 tf, y2 ti, y2 tf]
        double* d_sc = nullptr;      // device copy, [run_cap][8]
        int64_t run_cap = 0;
        bool have = false;
    } ep;
    bool have_radial = false, have_hc = false, have_ang = false, have_mask = false, have_state = false;
    // launch timing (qpc_timing_begin/end)
    int timed_kind = -1;
    // batch bookkeeping
    int64_t timed_max = 0;
    std::vector<hipEvent_t> ev;  // pairs
    int64_t timed_n = 0;

    void set_error(const std::string& s) { err = s; }
};

namespace {

template <class T>
int dev_alloc(qpc_ctx* c, T** p, size_t n) {
    QPC_HIP_CHECK(c, hipMalloc((void**)p, n * sizeof(T)));
    QPC_HIP_CHECK(c, hipMemsetAsync(*p, 0, n * sizeof(T), c->stream));
    return QPC_OK;
}
template <class T>
int upload(qpc_ctx* c, T* dst, const void* src, size_t n) {
    QPC_HIP_CHECK(c, hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
    return QPC_OK;
}

#define TRY(x)                 \
    do {                       \
        int _r = (x);          \
        if (_r != QPC_OK) return _r; \
    } while (0)

// device buffer freed when it goes out of scope (temporaries of the uploads)
struct DevTemp {
    void* p = nullptr;
    ~DevTemp() {
        if (p) (void)hipFree(p);
    }
};

inline int launch_check(qpc_ctx* c, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        c->set_error(std::string(what) + ": " + hipGetErrorString(e));
        return QPC_E_HIP;
    }
    return QPC_OK;
}

// event pairs around launches of the armed kernel class
struct TimedLaunch {
    qpc_ctx* c;
    hipStream_t s;
    bool on;
    TimedLaunch(qpc_ctx* c_, int kind) : TimedLaunch(c_, kind, c_->stream) {}
    TimedLaunch(qpc_ctx* c_, int kind, hipStream_t s_)
        : c(c_), s(s_), on(c_->timed_kind == kind && c_->timed_n < c_->timed_max) {
        if (on) (void)hipEventRecord(c->ev[2 * c->timed_n], s);
    }
    ~TimedLaunch() {
        if (on) {
            (void)hipEventRecord(c->ev[2 * c->timed_n + 1], s);
            ++c->timed_n;
        }
    }
};

}  // namespace

// pieces of the z-polarised path the elliptically polarised step reuses (qpc_gpu.hip)
namespace qpc_host {
int solve_free(qpc_ctx* c, const cplx* in, cplx* out);
// the same on channels in another order: channel k has l = lmap[k]
int solve_free_mapped(qpc_ctx* c, const cplx* in, cplx* out, const int* lmap);
int do_mask(qpc_ctx* c);
int check_ready(qpc_ctx* c, bool need_mask);
// observation of the current state before a step of a time loop (obsv::observe,
// binsrc/sphTDSE.cpp:794), when qpc_record is on; record_room checks that the
// history holds `steps` more rows
int record(qpc_ctx* c);
int record_room(qpc_ctx* c, int64_t steps);
// the D-part solve plan for sweep scalars |e| <= emax (measured on the uploaded
// S and d/dr; reused while emax stays within the planned range), and the
// length-gauge sweeps' (S and r)
int plan_d_part(qpc_ctx* c, double emax);
int plan_l_part(qpc_ctx* c, double emax);
// t-SURFF accumulate (xy) on the state after a step (spectrum_c_tsf,
// momentum_1e.hpp:716-776; all four zero: the field-free form)
int tsurff_step_xy(qpc_ctx* c, double ax0, double ax1, double ay0, double ay1);
// the device address of this GPU's copy of qpc_observe_noalloc's scratch
// array, resolved once per device (qpc_noalloc.hip); called at context creation
int resolve_scratch(qpc_ctx* c);
}  // namespace qpc_host
