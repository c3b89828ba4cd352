// qpc_gpu.hip -- MI355X Crank-Nicolson propagator for QPC-TDSE: context,
// C ABI (include/qpc_gpu.h) and the step sequence of propagator_cn_hc_vgwz
// (reference propagator/cn_1e.hpp:682-715) on device.
#include "qpc_ctx.h"
#include "qpc_krylov.h"
#include "qpc_setup.h"

#include <map>

#include <cmath>
#include <cstring>

namespace {

// ---- step pieces -----------------------------------------------------------
int rot_real(qpc_ctx* c, const double* Pt, const cplx* src, cplx* dst) {
    TimedLaunch t(c, QPC_K_ROT);
    qpc::launch_rot_real(c->stream, Pt, src, dst, c->M, c->L, c->N);
    return launch_check(c, "rot_real");
}
int rot_cplx(qpc_ctx* c, const cplx* Pt, const cplx* src, cplx* dst) {
    TimedLaunch t(c, QPC_K_ROT);
    qpc::launch_rot_cplx(c->stream, Pt, src, dst, c->M, c->L, c->N);
    return launch_check(c, "rot_cplx");
}

// per-batch table pointers of time point t
inline int64_t tabsz_r(const qpc_ctx* c) { return c->ML * c->geo.tab_elems; }   // real LU tables (D-part)
inline int64_t tabsz_c(const qpc_ctx* c) { return c->ML * c->geo.ctab_elems; }  // complex symmetric (R-part)
inline int64_t trfsz(const qpc_ctx* c) { return c->ML * c->geo.trf_elems; }

// cn_hl_propagate2 (D-part, propagate.hpp:388-441) on rotated channels:
// run time point t, whose tables are entry t - t0 of table set ts
// the work-group order of a partitioned solve whose channel ch reads table
// tab(ch) (qpc::solve_order), rebuilt when the segments per channel change
template <class TabOf>
int wg_order(qpc_ctx* c, qpc_ctx::WgOrder& o, int pk, TabOf tab_of, qpc::TabMap& tm) {
    const int nseg = qpc::solve_segments(c->geo, pk);
    if (nseg == 0) return QPC_OK;  // sequential solve
    if (nseg != o.nseg) {
        std::vector<int> tab((size_t)c->ML);
        for (int64_t ch = 0; ch < c->ML; ++ch) tab[ch] = tab_of(ch);
        o.h.assign((size_t)c->ML * nseg, 0);
        qpc::solve_order(tab.data(), (int)c->ML, nseg, o.h.data());
        if (o.d) QPC_HIP_CHECK(c, hipFree(o.d));
        o.d = nullptr;
        o.nseg = 0;
        TRY(dev_alloc(c, &o.d, o.h.size()));
        TRY(upload(c, o.d, o.h.data(), o.h.size()));
        o.nseg = nseg;
    }
    tm.order = o.d;
    tm.order_nseg = nseg;
    return QPC_OK;
}

int solve_D(qpc_ctx* c, int ts, int64_t t0, int64_t t, const cplx* in, cplx* out) {
    qpc::RhsArgs ra{};
    ra.sx = c->sxD;
    ra.eig = c->batch ? c->d_e + t * c->ML : c->yeig;
    ra.ifdt = c->batch ? 1.0 : c->h_dD[t];
    qpc::TabMap tm{0, (int)c->L, (int)c->m0};
    TRY(wg_order(c, c->dorder, c->pk_d, [](int64_t ch) { return (int)ch; }, tm));
    const int64_t k = t - t0;
    TimedLaunch tl(c, QPC_K_CN_D);
    qpc::launch_rsolve_r(c->stream, qpc::RHS_SX_ASY, ra, c->ts[ts].tabD + k * tabsz_r(c), c->ts[ts].trfD + k * trfsz(c),
                         tm, in, c->vbuf, out, (int)c->ML, c->geo, c->pk_d);
    return launch_check(c, "solve_D");
}
// cn_hl_propagate1 (R-part, propagate.hpp:333-386)
int solve_R(qpc_ctx* c, int ts, int64_t t0, int64_t t, const cplx* in, cplx* out) {
    qpc::RhsArgs ra{};
    ra.sx = c->sxR;
    ra.eig = c->batch ? c->d_e + (c->run_cap + t) * c->ML : c->qeig;
    ra.ifdt = c->batch ? 1.0 : c->h_dR[t];
    qpc::TabMap tm = c->n_rprim > 0 ? qpc::TabMap{4, (int)c->L, (int)c->m0, c->rmap}
                                    : qpc::TabMap{0, (int)c->L, (int)c->m0};
    if (c->n_rprim > 0) {
        // the channels of a shared table on one XCD, so that its rows are read once per L2
        const int nseg = qpc::solve_segments(c->geo, qpc::kPKS);
        if (nseg != c->rorder_nseg) {
            const size_t n = (size_t)c->ML * nseg;
            c->rorder_h.assign(n, 0);
            qpc::solve_order(c->rtab_h.data(), (int)c->ML, nseg, c->rorder_h.data());
            if (c->rorder) QPC_HIP_CHECK(c, hipFree(c->rorder));
            c->rorder = nullptr;
            TRY(dev_alloc(c, &c->rorder, n));
            TRY(upload(c, c->rorder, c->rorder_h.data(), n));
            c->rorder_nseg = nseg;
        }
        tm.order = c->rorder;
        tm.order_nseg = nseg;
    }
    const int64_t k = t - t0;
    TimedLaunch tl(c, QPC_K_CN_R);
    qpc::launch_rsolve_c(c->stream, qpc::RHS_SX_SYM, ra, c->ts[ts].tabR + k * tabsz_c(c), c->ts[ts].trfR + k * trfsz(c),
                         tm, in, c->vbuf, out, (int)c->ML, c->geo);
    return launch_check(c, "solve_R");
}
// length gauge: cn_hl_propagate1 (propagate.hpp:333-386) with X = r and the
// Y10 eigenvalues (prop_z1 of propagator_cn_hc_lgwz, cn_1e.hpp:587-590); its
// tables are the R-part tables of the set
int solve_L(qpc_ctx* c, int ts, int64_t t0, int64_t t, const cplx* in, cplx* out) {
    qpc::RhsArgs ra{};
    ra.sx = c->sxL;
    ra.eig = c->yeig;
    ra.ifdt = c->h_dR[t];
    const qpc::TabMap tm{0, (int)c->L, (int)c->m0};
    const int64_t k = t - t0;
    TimedLaunch tl(c, QPC_K_CN_R);
    qpc::launch_rsolve_c(c->stream, qpc::RHS_SX_SYM, ra, c->ts[ts].tabR + k * tabsz_c(c), c->ts[ts].trfR + k * trfsz(c),
                         tm, in, c->vbuf, out, (int)c->ML, c->geo, c->pk_l);
    return launch_check(c, "solve_L");
}
// cn_hc_propagate0 (propagate.hpp:311-331)
int solve_free(qpc_ctx* c, const cplx* in, cplx* out) {
    qpc::RhsArgs ra{};
    ra.cband = c->upperd;
    qpc::TabMap tm{c->batch ? 3 : 2, (int)c->L, (int)c->m0};
    const int64_t L = c->L, m0 = c->m0;
    const bool batch = c->batch;
    TRY(wg_order(c, c->forder, qpc::kPKS,
                 [=](int64_t ch) {
                     const int64_t m = m0 + ch / L;
                     return (int)(batch ? ch % L : ch % L + (m < 0 ? -m : m));
                 },
                 tm));
    TimedLaunch tl(c, QPC_K_FREE);
    qpc::launch_rsolve_c(c->stream, qpc::RHS_CBAND, ra, c->tab_free, c->trf_free, tm, in, c->vbuf, out, (int)c->ML,
                         c->geo);
    return launch_check(c, "solve_free");
}
// absorber_mask::operator() (absorber.hpp:52-64)
int solve_mask(qpc_ctx* c, const cplx* in, cplx* out) {
    qpc::RhsArgs ra{};
    ra.rband = c->maskd;
    const qpc::TabMap tm{1, (int)c->L, (int)c->m0};
    TimedLaunch tl(c, QPC_K_MASK);
    qpc::launch_rsolve_r(c->stream, qpc::RHS_RBAND, ra, c->tab_mask, c->trf_mask, tm, in, c->vbuf, out, (int)c->ML,
                         c->geo);
    return launch_check(c, "solve_mask");
}

// scalars of the run's time points from A(t) (cn_hl_propagate1/2,
// propagate.hpp:343,400; prop step dt/2, cn_1e.hpp:670-671; az = -A,
// cn_1e.hpp:689-690), host and device copies.  The streams are idle here.
int set_run_scalars(qpc_ctx* c, const double* a, int64_t n, int64_t stride = 0) {
    const double step = c->d.para_dt * 0.5;
    if (c->batch) {
        // per channel: e = eig * ifdt(shot) with the shot's own A(t) (row s of a)
        const int64_t ML = c->ML, L = c->L;
        if (n > c->run_cap) {
            void* old[] = {c->d_e, c->d_ones, c->d_ifdt};
            for (void* p : old)
                if (p) QPC_HIP_CHECK(c, hipFree(p));
            c->d_e = c->d_ones = c->d_ifdt = nullptr;
            c->run_cap = 0;
            QPC_HIP_CHECK(c, hipMalloc((void**)&c->d_e, sizeof(double) * 2 * n * ML));
            QPC_HIP_CHECK(c, hipMalloc((void**)&c->d_ones, sizeof(double) * n));
            QPC_HIP_CHECK(c, hipMalloc((void**)&c->d_ifdt, sizeof(double) * 2 * n));
            c->run_cap = n;
            std::vector<double> ones(n, 1.0);
            QPC_HIP_CHECK(c, hipMemcpy(c->d_ones, ones.data(), sizeof(double) * n, hipMemcpyHostToDevice));
        }
        c->h_e.assign((size_t)2 * c->run_cap * ML, 0.0);
        double emax = 0.0;
        for (int64_t t = 0; t < n; ++t)
            for (int64_t ch = 0; ch < ML; ++ch) {
                const double fzdt = (-a[(ch / L) * stride + t]) * step;
                const double dD = -fzdt / 2, dR = fzdt / 2;
                const double eD = c->yeig_h[ch] * dD;
                if (!std::isfinite(eD) || !std::isfinite(dR)) {
                    c->set_error("non-finite A(t)");
                    return QPC_E_NUMERIC;
                }
                emax = std::max(emax, std::fabs(eD));
                c->h_e[(size_t)t * ML + ch] = eD;
                c->h_e[(size_t)(c->run_cap + t) * ML + ch] = c->qeig_h[ch] * dR;
            }
        TRY(qpc_host::plan_d_part(c, emax));
        QPC_HIP_CHECK(c, hipMemcpyAsync(c->d_e, c->h_e.data(), sizeof(double) * c->h_e.size(), hipMemcpyHostToDevice,
                                        c->stream));
        if (c->n_rprim > 0) {
            // the shared R-part tables' scalars: their channels' own products
            const int64_t np = c->n_rprim;
            if (c->eprim_cap < c->run_cap * np) {
                QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
                if (c->d_eprim) QPC_HIP_CHECK(c, hipFree(c->d_eprim));
                c->d_eprim = nullptr;
                c->eprim_cap = 0;
                QPC_HIP_CHECK(c, hipMalloc((void**)&c->d_eprim, sizeof(double) * c->run_cap * np));
                c->eprim_cap = c->run_cap * np;
            }
            std::vector<double> hp((size_t)n * np);
            for (int64_t t = 0; t < n; ++t)
                for (int64_t j = 0; j < np; ++j)
                    hp[(size_t)t * np + j] = c->h_e[(size_t)(c->run_cap + t) * ML + c->rprim_ch[j]];
            QPC_HIP_CHECK(c, hipMemcpyAsync(c->d_eprim, hp.data(), sizeof(double) * hp.size(), hipMemcpyHostToDevice,
                                            c->stream));
            QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        }
        QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        return QPC_OK;
    }
    c->h_dD.resize(n);
    c->h_dR.resize(n);
    double emax = 0.0;
    if (c->gauge == 1) {
        // length gauge: prop_z1.propagate1(ez) with step dt/2 (cn_1e.hpp:587,
        // 608), ifdt = ez dt/4 (propagate.hpp:343); the R-part slot carries it
        for (int64_t t = 0; t < n; ++t) {
            c->h_dR[t] = a[t] * step / 2;
            c->h_dD[t] = 0.0;
            if (!std::isfinite(c->h_dR[t])) {
                c->set_error("non-finite E(t)");
                return QPC_E_NUMERIC;
            }
            emax = std::max(emax, std::fabs(c->h_dR[t]) * c->yeig_max);
        }
        TRY(qpc_host::plan_l_part(c, emax));
        if (n > c->run_cap) {
            if (c->d_ifdt) QPC_HIP_CHECK(c, hipFree(c->d_ifdt));
            c->d_ifdt = nullptr;
            c->run_cap = 0;
            QPC_HIP_CHECK(c, hipMalloc((void**)&c->d_ifdt, sizeof(double) * 2 * n));
            c->run_cap = n;
        }
        QPC_HIP_CHECK(c, hipMemcpyAsync(c->d_ifdt + c->run_cap, c->h_dR.data(), sizeof(double) * n,
                                        hipMemcpyHostToDevice, c->stream));
        QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        return QPC_OK;
    }
    for (int64_t t = 0; t < n; ++t) {
        const double fzdt = (-a[t]) * step;
        c->h_dD[t] = -fzdt / 2;
        c->h_dR[t] = fzdt / 2;
        if (!std::isfinite(fzdt)) {
            c->set_error("non-finite A(t)");
            return QPC_E_NUMERIC;
        }
        emax = std::max(emax, std::fabs(c->h_dD[t]) * c->yeig_max);
    }
    // carry depth and restart of the D-part solves for this run's scalars
    TRY(qpc_host::plan_d_part(c, emax));
    if (n > c->run_cap) {
        if (c->d_ifdt) QPC_HIP_CHECK(c, hipFree(c->d_ifdt));
        c->d_ifdt = nullptr;
        c->run_cap = 0;
        QPC_HIP_CHECK(c, hipMalloc((void**)&c->d_ifdt, sizeof(double) * 2 * n));
        c->run_cap = n;
    }
    QPC_HIP_CHECK(c, hipMemcpyAsync(c->d_ifdt, c->h_dD.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    QPC_HIP_CHECK(c, hipMemcpyAsync(c->d_ifdt + c->run_cap, c->h_dR.data(), sizeof(double) * n,
                                    hipMemcpyHostToDevice, c->stream));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return QPC_OK;
}

// factorisations of the coupling sweeps for run time points [t0, t0+n) into
// table set ts from position pos on, on fstream once the steps that read the
// set before are done
int factor_batch(qpc_ctx* c, int ts, int64_t t0, int64_t n, int64_t pos = 0, bool wait_set = true,
                 bool split = false) {
    if (pos + n > c->tb_cap) {
        c->set_error("factor batch larger than its capacity");
        return QPC_E_ARG;
    }
    const qpc_ctx::TableSet& S = c->ts[ts];
    qpc_ctx::TableSet T = S;  // the set's tables from position pos on
    T.tabD = S.tabD + pos * tabsz_r(c);
    T.trfD = S.trfD + pos * trfsz(c);
    T.tabR = S.tabR + pos * tabsz_c(c);
    T.trfR = S.trfR + pos * trfsz(c);
    if (wait_set && T.pending) QPC_HIP_CHECK(c, hipStreamWaitEvent(c->fstream, T.freed, 0));
    // the R-part tables and transfer sets first (ready_r), then the D-part's
    // (ready): a step reads the R part of its t_f tables before the D part
    {
        TimedLaunch tl(c, QPC_K_FACTOR, c->fstream);
        const int64_t ML = c->ML;
        if (c->batch) {
            if (c->n_rprim > 0)  // one table per shot and distinct |eigenvalue|
                qpc::launch_factor_sym(c->fstream, c->S, c->rinv, c->d_eprim + t0 * c->n_rprim, c->d_ones, (int)n,
                                       (int)c->n_rprim, c->geo, T.tabR, c->n_rprim, 0, ML);
            else
                qpc::launch_factor_sym(c->fstream, c->S, c->rinv, c->d_e + (c->run_cap + t0) * ML, c->d_ones,
                                       (int)n, (int)ML, c->geo, T.tabR, ML);
        } else if (c->gauge == 1) {
            // length gauge: S + i e r with the Y10 eigenvalues only
            const double* dL = c->d_ifdt + c->run_cap + t0;
            qpc::launch_factor_sym(c->fstream, c->S, c->rpow1, c->yeig, dL, (int)n, (int)ML, c->geo, T.tabR, 0,
                                   c->restart_l);
        } else {
            const double* dR = c->d_ifdt + c->run_cap + t0;
            if (c->n_rprim > 0)  // one table per distinct |eigenvalue|, in the first slots of each time point
                qpc::launch_factor_sym(c->fstream, c->S, c->rinv, c->qeig_prim, dR, (int)n, (int)c->n_rprim, c->geo,
                                       T.tabR, 0, 0, ML);
            else
                qpc::launch_factor_sym(c->fstream, c->S, c->rinv, c->qeig, dR, (int)n, (int)ML, c->geo, T.tabR);
        }
        if (c->gauge != 1 && c->n_rprim > 0)
            qpc::launch_transfer_c(c->fstream, T.tabR, T.trfR, (int)c->n_rprim, (int)n, ML, c->geo);
        else
            qpc::launch_transfer_c(c->fstream, T.tabR, T.trfR, (int)(n * ML), c->geo);
    }
    TRY(launch_check(c, "factor_batch"));
    QPC_HIP_CHECK(c, hipEventRecord(S.ready_r, c->fstream));
    // split (a per-step call's one time point): the D part on its own stream,
    // after everything queued on fstream before this call's R part
    hipStream_t ds = split && c->gauge != 1 ? c->dstream : c->fstream;
    if (ds != c->fstream) QPC_HIP_CHECK(c, hipStreamWaitEvent(ds, c->put_done, 0));
    if (c->gauge != 1) {
        TimedLaunch tl(c, QPC_K_FACTOR, ds);
        const int64_t ML = c->ML;
        if (c->batch)
            qpc::launch_factor_asy(ds, c->S, c->rdif, c->d_e + t0 * ML, c->d_ones, (int)n, (int)ML, c->geo, T.tabD,
                                   ML, c->restart_d);
        else
            qpc::launch_factor_asy(ds, c->S, c->rdif, c->yeig, c->d_ifdt + t0, (int)n, (int)ML, c->geo, T.tabD, 0,
                                   c->restart_d);
        qpc::launch_transfer_r(ds, T.tabD, T.trfD, (int)(n * ML), c->geo);
        TRY(launch_check(c, "factor_batch"));
    }
    QPC_HIP_CHECK(c, hipEventRecord(S.ready, ds));
    if (ds != c->fstream) {
        // later work on fstream (the next call's tables) comes after this D part
        QPC_HIP_CHECK(c, hipStreamWaitEvent(c->fstream, S.ready, 0));
    }
    return QPC_OK;
}

// the step stream waits for a table set, and later marks it free again
int use_set(qpc_ctx* c, int ts) {
    QPC_HIP_CHECK(c, hipStreamWaitEvent(c->stream, c->ts[ts].ready, 0));
    return QPC_OK;
}
int release_set(qpc_ctx* c, int ts) {
    QPC_HIP_CHECK(c, hipEventRecord(c->ts[ts].freed, c->stream));
    c->ts[ts].pending = true;
    return QPC_OK;
}

// two scalars written by the device (queued with the values, so the host
// can go on at once)
__global__ void k_put2(double* p, double v, double* q, double w) {
    if (threadIdx.x == 0) {
        *p = v;
        *q = w;
    }
}

// one with-field step (propagator_cn_hc_vgwz::propagate, cn_1e.hpp:682-709)
// with run time points ti (A(t_i)) and tf (A(t_f)) whose tables are in set ts
// from run time point t0 on; coef ends in c->coef.
//
// Inside a run the last rotation of a step (tran_z2.eliminate, P_Y^T) and
// the first of the next (tran_z2.construct, P_Y) enclose only the absorber,
// S^-1 M, which acts alike on every channel and so commutes with any mixing
// of the channels at fixed r: P_Y (S^-1 M) P_Y^T = S^-1 M.  A step that
// hands over to another with-field step (out_y) therefore applies the
// absorber in the D-part basis and leaves the state there; a step that
// starts from that basis (in_y) skips its first rotation.  Two of the six
// rotations per step drop out; the result differs from the sequence with
// both rotations by rounding only.
int do_step(qpc_ctx* c, int ts, int64_t t0, int64_t ti, int64_t tf, bool in_y = false, bool out_y = false,
            bool mask_in_y = false, hipEvent_t wait_r = nullptr, hipEvent_t wait_d = nullptr) {
    cplx* X = c->coef;
    cplx* Y = c->tmpc;
    cplx* Z = c->spare;
    if (in_y) {
        TRY(solve_D(c, ts, t0, ti, X, Z));           // prop_z2.propagate2(A_i), state already in P_Y's basis
    } else {
        TRY(rot_real(c, c->rtY, X, Y));              // tran_z2.construct
        TRY(solve_D(c, ts, t0, ti, Y, Z));           // prop_z2.propagate2(A_i)
    }
    TRY(rot_cplx(c, c->rtM, Z, Y));    // z2_2_z1.construct
    TRY(solve_R(c, ts, t0, ti, Y, Z));               // prop_z1.propagate1(A_i)
    TRY(rot_cplx(c, c->rtQh, Z, Y));      // tran_z1.eliminate
    TRY(solve_free(c, Y, Z));                // prop_hc.propagate0
    TRY(rot_cplx(c, c->rtQ, Z, Y));      // tran_z1.construct
    // (a per-step call: the t_f tables are factorised while the first half runs)
    if (wait_r) QPC_HIP_CHECK(c, hipStreamWaitEvent(c->stream, wait_r, 0));
    TRY(solve_R(c, ts, t0, tf, Y, Z));               // prop_z1.propagate1(A_f)
    TRY(rot_cplx(c, c->rtMh, Z, Y));    // z2_2_z1.eliminate
    if (wait_d) QPC_HIP_CHECK(c, hipStreamWaitEvent(c->stream, wait_d, 0));
    TRY(solve_D(c, ts, t0, tf, Y, Z));               // prop_z2.propagate2(A_f)
    if (out_y) {
        // stay in P_Y's basis for the next step, with the absorber applied there
        if (mask_in_y) {
            TRY(solve_mask(c, Z, X));
        } else {
            c->coef = Z;
            c->spare = X;
        }
        return QPC_OK;
    }
    TRY(rot_real(c, c->rtYt, Z, X));      // tran_z2.eliminate
    return QPC_OK;
}

// one length-gauge step (propagator_cn_hc_lgwz::propagate, cn_1e.hpp:598-613)
// with run time points ti (E(t_i)) and tf (E(t_f)); coef ends in c->coef
int do_step_lg(qpc_ctx* c, int ts, int64_t t0, int64_t ti, int64_t tf, hipEvent_t wait_r = nullptr) {
    cplx* X = c->coef;
    cplx* Y = c->tmpc;
    cplx* Z = c->spare;
    TRY(rot_real(c, c->rtY, X, Y));            // tran_z1.construct
    TRY(solve_L(c, ts, t0, ti, Y, Z));         // prop_z1.propagate1(ez_ti)
    TRY(rot_real(c, c->rtYt, Z, X));           // tran_z1.eliminate
    TRY(solve_free(c, X, Y));                  // prop_hc.propagate0
    TRY(rot_real(c, c->rtY, Y, Z));            // tran_z1.construct
    if (wait_r) QPC_HIP_CHECK(c, hipStreamWaitEvent(c->stream, wait_r, 0));
    TRY(solve_L(c, ts, t0, tf, Z, Y));         // prop_z1.propagate1(ez_tf)
    TRY(rot_real(c, c->rtYt, Y, X));           // tran_z1.eliminate
    return QPC_OK;
}

int do_free(qpc_ctx* c) {
    TRY(solve_free(c, c->coef, c->tmpc));
    std::swap(c->coef, c->tmpc);
    return QPC_OK;
}

int do_mask(qpc_ctx* c) {
    TRY(solve_mask(c, c->coef, c->tmpc));
    std::swap(c->coef, c->tmpc);
    return QPC_OK;
}

// spectrum_c_tsf::accumulate (momentum_1e.hpp:675-719; a field-free step is
// az0 = az1 = 0, :631-673) on the state after a step: the clock advances by
// dt and int A dt by the mean of the step's end values, then every momentum
// point gains the step's flux term
//
// xy: accumulate (xy) (:716-776) with A_x, A_y at the step's ends; with all
// four zero it is the field-free form of the elliptically polarised run.
int tsurff_launch(qpc_ctx* c, bool xy, double az1, double ax1, double ay1) {
    const double dt = c->d.para_dt;
    qpc::launch_tsf_temp(c->stream, c->coef, c->tsf_ib, c->tsf_b0, c->tsf_br, c->tsf_nb, c->tsf_tmp, c->ML, c->N);
    qpc::TsfArgs a{};
    a.xy = xy ? 1 : 0;
    a.ax1 = ax1;
    a.ay1 = ay1;
    a.stx = c->tsf_stx;
    a.sty = c->tsf_sty;
    a.pm1 = c->tsf_pm1;
    a.pm2 = c->tsf_pm2;
    a.disx = c->tsf_disx;
    a.disy = c->tsf_disy;
    a.kr = c->tsf_kr;
    a.ckt = c->tsf_ckt;
    a.j0 = c->tsf_j0;
    a.jr = c->tsf_jr;
    a.pm = c->tsf_pm;
    a.Y = c->tsf_Y;
    a.tmp = c->tsf_tmp;
    a.work = c->tsf_work;
    a.nkr = c->tsf_nkr;
    a.nktp = c->tsf_nktp;
    a.nl = (int)c->lmax + 1;
    a.M = (int)c->M;
    a.L = (int)c->L;
    a.m0 = (int)c->m0;
    a.az1 = az1;
    a.time = c->tsf_time;
    a.disz = c->tsf_disz;
    a.dt = dt;
    qpc::launch_tsf_acc(c->stream, a);
    return launch_check(c, "tsurff");
}
int tsurff_step(qpc_ctx* c, double az0, double az1) {
    const double dt = c->d.para_dt;
    c->tsf_time += dt;
    c->tsf_disz += 0.5 * (az0 + az1) * dt;
    return tsurff_launch(c, false, az1, 0.0, 0.0);
}
int tsurff_step_xy(qpc_ctx* c, double ax0, double ax1, double ay0, double ay1) {
    const double dt = c->d.para_dt;
    c->tsf_time += dt;
    c->tsf_disx += 0.5 * (ax0 + ax1) * dt;
    c->tsf_disy += 0.5 * (ay0 + ay1) * dt;
    return tsurff_launch(c, true, 0.0, ax1, ay1);
}

int ready(qpc_ctx* c, bool need_mask) {
    if (!c->have_radial || !c->have_hc || !c->have_ang || !c->have_state) {
        c->set_error("operands or state not uploaded");
        return QPC_E_STATE;
    }
    if (need_mask && !c->have_mask) {
        c->set_error("absorber band not uploaded");
        return QPC_E_STATE;
    }
    return QPC_OK;
}

// the two table sets of the z-polarised coupling sweeps, allocated at their
// first use (a context that only runs the elliptically polarised step or the
// Krylov propagator never holds them)
int alloc_zsets(qpc_ctx* c) {
    if (c->ts[0].tabD) return QPC_OK;
    const size_t ntab = (size_t)c->geo.tab_elems, nctab = (size_t)c->geo.ctab_elems;
    const size_t ntrf = (size_t)c->geo.trf_elems;
    const size_t nt = (size_t)c->tb_cap * c->ML;
    for (auto& T : c->ts) {
        int rc = QPC_OK;
        if ((rc = dev_alloc(c, &T.tabD, nt * ntab)) || (rc = dev_alloc(c, &T.trfD, nt * ntrf)) ||
            (rc = dev_alloc(c, &T.tabR, nt * nctab)) || (rc = dev_alloc(c, &T.trfR, nt * ntrf)))
            return rc;
    }
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return QPC_OK;
}

}  // namespace

namespace {
int run_impl(qpc_ctx* c, const double* atab, int64_t stride, int64_t nt_field, int64_t it0, int64_t it1,
             int use_mask);
int step_impl(qpc_ctx* c, int gauge, double a_i, double a_f);
}  // namespace

namespace qpc_host {
int solve_free(qpc_ctx* c, const cplx* in, cplx* out) { return ::solve_free(c, in, out); }
int solve_free_mapped(qpc_ctx* c, const cplx* in, cplx* out, const int* lmap) {
    qpc::RhsArgs ra{};
    ra.cband = c->upperd;
    const qpc::TabMap tm{4, (int)c->L, (int)c->m0, lmap};
    TimedLaunch tl(c, QPC_K_FREE);
    qpc::launch_rsolve_c(c->stream, qpc::RHS_CBAND, ra, c->tab_free, c->trf_free, tm, in, c->vbuf, out, (int)c->ML,
                         c->geo);
    return launch_check(c, "solve_free");
}
int do_mask(qpc_ctx* c) { return ::do_mask(c); }
int check_ready(qpc_ctx* c, bool need_mask) { return ::ready(c, need_mask); }
int tsurff_step_xy(qpc_ctx* c, double ax0, double ax1, double ay0, double ay1) {
    return ::tsurff_step_xy(c, ax0, ax1, ay0, ay1);
}
int record(qpc_ctx* c) {
    if (!c->rec_flags) return QPC_OK;
    const qpc::ObsBands b{c->S, c->rpow1, c->accel, c->souter, c->y10};
    qpc::launch_observe(c->stream, b, c->rec_flags, c->coef, c->red4, c->hist + qpc::kObs * c->rec_n, c->M, c->L,
                        c->m0, c->N, c->p);
    ++c->rec_n;
    return launch_check(c, "record");
}
int plan_l_part(qpc_ctx* c, double emax) {
    if (emax <= c->plan_emax_l) return QPC_OK;
    if ((int64_t)c->rpow1_h.size() != c->N * c->ne) {
        c->set_error("radial operators not uploaded");
        return QPC_E_STATE;
    }
    const double e = 1.25 * emax;
    const qpc::DPlan d = qpc::coupling_plan(c->S_h.data(), c->rpow1_h.data(), (int)c->N, (int)c->p, e, 0);
    c->pk_l = d.pk;
    c->restart_l = d.restart;
    c->plan_emax_l = e;
    return QPC_OK;
}
int plan_d_part(qpc_ctx* c, double emax) {
    if (emax <= c->plan_emax) return QPC_OK;
    if ((int64_t)c->S_h.size() != c->N * c->ne) {
        c->set_error("radial operators not uploaded");
        return QPC_E_STATE;
    }
    // plan for a little more than this run needs, so that runs of slightly
    // stronger fields reuse it
    const double e = 1.25 * emax;
    const qpc::DPlan d = qpc::d_part_plan(c->S_h.data(), c->rdif_h.data(), (int)c->N, (int)c->p, e);
    c->pk_d = d.pk;
    c->restart_d = d.restart;
    c->plan_emax = e;
    return QPC_OK;
}
int record_room(qpc_ctx* c, int64_t steps) {
    if (c->rec_flags && c->rec_n + steps > c->rec_cap) {
        c->set_error("observable history full (qpc_record capacity)");
        return QPC_E_ARG;
    }
    return QPC_OK;
}
}  // namespace qpc_host

extern "C" {

static int create_impl(const qpc_dims* d, int64_t n_shots, int64_t tb_cap, qpc_ctx** out);

int qpc_create(const qpc_dims* d, qpc_ctx** out) { return create_impl(d, 0, kTimeBatch, out); }

int qpc_create_batch(const qpc_dims* d, int64_t n_shots, int64_t time_batch, qpc_ctx** out) {
    if (!d || !out || n_shots < 1 || d->n_mdim != 1 || d->n_mmin != 0 || time_batch < 2 || time_batch > kTimeBatch)
        return QPC_E_ARG;
    return create_impl(d, n_shots, time_batch, out);
}

static int create_impl(const qpc_dims* d, int64_t n_shots, int64_t tb_cap, qpc_ctx** out) {
    if (!d || !out) return QPC_E_ARG;
    *out = nullptr;
    // the partitioned solver is built for the order-7 basis (n_rank = 7) and
    // at most kMaxGroups wave groups of partitions per channel
    if (!qpc::width_supported((int)d->n_rank) || d->n_ndim < 2 * (d->n_rank + 1) ||
        d->n_ndim > (int64_t)qpc::kMaxGroups * qpc::kLane * qpc::kPM || d->n_ldim < 1 || d->n_mdim < 1 ||
        !(d->para_dt > 0) || d->n_ldim * d->n_mdim > 65535 || d->n_ldim > qpc::kRotMaxL)
        return QPC_E_SHAPE;
    qpc_ctx* c = new qpc_ctx();
    c->d = *d;
    c->N = d->n_ndim;
    c->p = d->n_rank;
    c->ne = c->p + 1;
    c->L = d->n_ldim;
    c->M = d->n_mdim;
    c->m0 = d->n_mmin;
    if (n_shots > 0) {  // batched shots: M blocks of m = 0 channels, one per shot
        c->batch = true;
        c->M = n_shots;
        c->d.n_mdim = n_shots;
    }
    c->ML = c->M * c->L;
    if (c->ML > 65535) {
        delete c;
        return QPC_E_SHAPE;
    }
    int64_t m1 = c->m0 + c->M - 1;
    c->lmax = c->batch ? c->L - 1 : std::max(std::llabs(c->m0), std::llabs(m1)) + c->L - 1;  // dimension/1e.hpp:49-50
    int rc = QPC_OK;
    auto fail = [&](int r) {
        qpc_destroy(c);
        return r;
    };
    if (hipGetDevice(&c->device) != hipSuccess) return fail(QPC_E_HIP);
    if ((rc = qpc_host::resolve_scratch(c))) return fail(rc);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return fail(QPC_E_HIP);
    if (hipStreamCreateWithFlags(&c->fstream, hipStreamNonBlocking) != hipSuccess) return fail(QPC_E_HIP);
    if (hipStreamCreateWithFlags(&c->dstream, hipStreamNonBlocking) != hipSuccess) return fail(QPC_E_HIP);
    if (hipEventCreateWithFlags(&c->put_done, hipEventDisableTiming) != hipSuccess) return fail(QPC_E_HIP);
    for (auto& T : c->ts)
        if (hipEventCreateWithFlags(&T.ready, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&T.ready_r, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&T.freed, hipEventDisableTiming) != hipSuccess)
            return fail(QPC_E_HIP);
    const size_t band = (size_t)c->N * c->ne;
    const size_t nl = (size_t)c->lmax + 1;
    const size_t nstate = (size_t)c->ML * c->N;
    c->geo = qpc::make_geom((int)c->N, (int)c->p);
    c->tb_cap = tb_cap;
    const size_t ntab = (size_t)c->geo.tab_elems, nctab = (size_t)c->geo.ctab_elems;
    const size_t ntrf = (size_t)c->geo.trf_elems;
    if ((rc = dev_alloc(c, &c->spare, nstate)) || (rc = dev_alloc(c, &c->vbuf, nstate)) ||
        (rc = dev_alloc(c, &c->tab_free, nl * nctab)) ||
        (rc = dev_alloc(c, &c->trf_free, nl * ntrf)) || (rc = dev_alloc(c, &c->tab_mask, ntab)) ||
        (rc = dev_alloc(c, &c->trf_mask, ntrf)))
        return fail(rc);
    if ((rc = dev_alloc(c, &c->S, band)) || (rc = dev_alloc(c, &c->chol, band)) || (rc = dev_alloc(c, &c->rinv, band)) ||
        (rc = dev_alloc(c, &c->rdif, band)) || (rc = dev_alloc(c, &c->rpow1, band)) ||
        (rc = dev_alloc(c, &c->mask, band)) || (rc = dev_alloc(c, &c->upper, nl * band)) ||
        (rc = dev_alloc(c, &c->lower, nl * band)) || (rc = dev_alloc(c, &c->upperd, nl * band)) ||
        (rc = dev_alloc(c, &c->Sd, band)) || (rc = dev_alloc(c, &c->rinvd, band)) ||
        (rc = dev_alloc(c, &c->rdifd, band)) || (rc = dev_alloc(c, &c->maskd, band)) ||
        (rc = dev_alloc(c, &c->rpow1d, band)) || (rc = dev_alloc(c, &c->sxR, band)) ||
        (rc = dev_alloc(c, &c->sxD, band)) || (rc = dev_alloc(c, &c->sxL, band)) ||
        (rc = dev_alloc(c, &c->yeig, (size_t)c->ML)) ||
        (rc = dev_alloc(c, &c->qeig, (size_t)c->ML)) || (rc = dev_alloc(c, &c->y10, (size_t)c->ML)) ||
        (rc = dev_alloc(c, &c->yprj, (size_t)c->ML * c->L)) || (rc = dev_alloc(c, &c->qprj, (size_t)c->ML * c->L)) ||
        (rc = dev_alloc(c, &c->merged, (size_t)c->ML * c->L)) ||
        (rc = dev_alloc(c, &c->rtY, (size_t)(c->M * qpc::rot_prep_elems(c->L)))) ||
        (rc = dev_alloc(c, &c->rtYt, (size_t)(c->M * qpc::rot_prep_elems(c->L)))) ||
        (rc = dev_alloc(c, &c->rtM, (size_t)(c->M * qpc::rot_prep_elems(c->L)))) ||
        (rc = dev_alloc(c, &c->rtMh, (size_t)(c->M * qpc::rot_prep_elems(c->L)))) ||
        (rc = dev_alloc(c, &c->rtQ, (size_t)(c->M * qpc::rot_prep_elems(c->L)))) ||
        (rc = dev_alloc(c, &c->rtQh, (size_t)(c->M * qpc::rot_prep_elems(c->L)))) || (rc = dev_alloc(c, &c->coef, nstate)) ||
        (rc = dev_alloc(c, &c->tmpc, nstate)) || (rc = dev_alloc(c, &c->red, (size_t)2 * c->ML)))
        return fail(rc);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(QPC_E_HIP);
    *out = c;
    return QPC_OK;
}

int qpc_destroy(qpc_ctx* c) {
    if (!c) return QPC_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->fstream) (void)hipStreamSynchronize(c->fstream);
    for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->slot_read)
        if (e) (void)hipEventDestroy(e);
    if (c->rmap) (void)hipFree(c->rmap);
    if (c->rorder) (void)hipFree(c->rorder);
    if (c->dorder.d) (void)hipFree(c->dorder.d);
    if (c->forder.d) (void)hipFree(c->forder.d);
    if (c->qeig_prim) (void)hipFree(c->qeig_prim);
    if (c->d_eprim) (void)hipFree(c->d_eprim);
    for (auto& T : c->ts) {
        if (T.ready) (void)hipEventDestroy(T.ready);
        if (T.ready_r) (void)hipEventDestroy(T.ready_r);
        if (T.freed) (void)hipEventDestroy(T.freed);
        void* tp[] = {T.tabD, T.trfD, T.tabR, T.trfR};
        for (void* p : tp)
            if (p) (void)hipFree(p);
    }
    void* ptrs[] = {c->d_e, c->d_ones, c->S, c->chol, c->rinv, c->rdif, c->rpow1, c->mask, c->upper, c->lower, c->yeig,
                    c->qeig,
                    c->y10, c->yprj, c->qprj, c->merged, c->coef, c->tmpc, c->spare, c->vbuf, c->red,
                    c->tab_free, c->trf_free, c->tab_mask, c->trf_mask, c->d_ifdt, c->Sd, c->rinvd, c->rdifd, c->maskd, c->upperd,
                    c->rpow1d, c->sxR, c->sxD, c->sxL, c->tsf_ib, c->tsf_b0, c->tsf_br, c->tsf_kr, c->tsf_ckt,
                    c->tsf_j0, c->tsf_jr, c->tsf_pm, c->tsf_Y, c->tsf_tmp, c->tsf_work, c->tsf_stx, c->tsf_sty,
                    c->tsf_pm1, c->tsf_pm2, c->rtY, c->rtYt, c->rtM, c->rtMh, c->rtQ, c->rtQh,
                    c->Wp, c->ppart, c->pchan, c->pcnt, c->pout, c->Hd, c->q10, c->kU, c->kpart, c->kcoef,
                    c->souter, c->accel, c->hist, c->red4, c->obs_tmp, c->pcs_Wp, c->pcs_part, c->pcs_chan,
                    c->pcs_cnt, c->pcs_cps, c->pcs_out, c->pcs_lmd};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    {
        auto& E = c->ep;
        for (auto& T : E.set) {
            if (T.ready) (void)hipEventDestroy(T.ready);
            if (T.freed) (void)hipEventDestroy(T.freed);
            void* tp[] = {T.tabD, T.trfD, T.tabR, T.trfR};
            for (void* p : tp)
                if (p) (void)hipFree(p);
        }
        void* ep_ptrs[] = {E.eig,  E.tx1,  E.ty2,     E.x1x2,     E.x2y1,     E.y1y2,     E.ty1,      E.d_sc,
                           E.perm, E.lmap, E.ty2p,    E.ty1p,     E.rmap[0],  E.rmap[1],  E.rmap[2],  E.rmap[3],
                           E.eprim[0], E.eprim[1], E.eprim[2], E.eprim[3], E.ord[0], E.ord[1], E.ord[2],
                           E.ord[3]};
        for (void* p : ep_ptrs)
            if (p) (void)hipFree(p);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->fstream) (void)hipStreamDestroy(c->fstream);
    if (c->dstream) (void)hipStreamDestroy(c->dstream);
    if (c->put_done) (void)hipEventDestroy(c->put_done);
    delete c;
    return QPC_OK;
}

const char* qpc_last_error(const qpc_ctx* c) { return c ? c->err.c_str() : "null context"; }

int qpc_upload_radial(qpc_ctx* c, const double* S, const double* chol, const double* rinv, const double* rdif,
                      const double* rpow1, const double* mask) {
    if (!c || !S || !chol || !rinv || !rdif) return QPC_E_ARG;
    c->step_slot = -1;  // tables of the previous operators
    const size_t band = (size_t)c->N * c->ne;
    TRY(upload(c, c->S, S, band));
    TRY(upload(c, c->chol, chol, band));
    TRY(upload(c, c->rinv, rinv, band));
    TRY(upload(c, c->rdif, rdif, band));
    if (rpow1) TRY(upload(c, c->rpow1, rpow1, band));
    if (mask) {
        TRY(upload(c, c->mask, mask, band));
        c->have_mask = true;
    }
    const int n = (int)c->N, p = (int)c->p;
    qpc::launch_band_diag_r(c->stream, c->S, c->Sd, 1, n, p);
    qpc::launch_band_diag_r(c->stream, c->rinv, c->rinvd, 1, n, p);
    qpc::launch_band_diag_r(c->stream, c->rdif, c->rdifd, 1, n, p);
    qpc::launch_band_diag_r(c->stream, c->mask, c->maskd, 1, n, p);
    qpc::launch_band_diag_r(c->stream, c->rpow1, c->rpow1d, 1, n, p);
    qpc::launch_band_pair(c->stream, c->Sd, c->rinvd, c->sxR, n, p);
    qpc::launch_band_pair(c->stream, c->Sd, c->rdifd, c->sxD, n, p);
    qpc::launch_band_pair(c->stream, c->Sd, c->rpow1d, c->sxL, n, p);
    c->S_h.assign(S, S + band);
    c->rdif_h.assign(rdif, rdif + band);
    if (rpow1) c->rpow1_h.assign(rpow1, rpow1 + band);
    else c->rpow1_h.clear();
    c->plan_emax = c->plan_emax_l = -1.0;  // the plans belong to the previous operators
    c->pk_d = qpc::kPKD;
    c->restart_d = qpc::kRestart;
    c->pk_l = qpc::kPKS;
    c->restart_l = qpc::kRestart;
    // absorber solve tables from chol(S) and their partition transfers
    qpc::launch_tab_chol(c->stream, c->chol, c->tab_mask, c->geo);
    qpc::launch_transfer_r(c->stream, c->tab_mask, c->trf_mask, 1, c->geo);
    TRY(launch_check(c, "mask tables"));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    c->have_radial = true;
    return QPC_OK;
}

// field-free solve tables per l from the device's upper / lower bands
static int install_hc(qpc_ctx* c) {
    qpc::launch_band_diag_c(c->stream, c->upper, c->upperd, (int)(c->lmax + 1), (int)c->N, (int)c->p);
    // field-free solve tables per l from the LDL^T factors
    qpc::launch_tab_ldlt(c->stream, c->lower, c->tab_free, (int)(c->lmax + 1), c->geo);
    qpc::launch_transfer_c(c->stream, c->tab_free, c->trf_free, (int)(c->lmax + 1), c->geo);
    TRY(launch_check(c, "free tables"));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    c->have_hc = true;
    return QPC_OK;
}

int qpc_upload_hc(qpc_ctx* c, const double* upper, const double* lower) {
    if (!c || !upper || !lower) return QPC_E_ARG;
    c->step_slot = -1;  // tables of the previous operators
    const size_t n = (size_t)(c->lmax + 1) * c->N * c->ne;
    TRY(upload(c, c->upper, upper, n));
    TRY(upload(c, c->lower, lower, n));
    return install_hc(c);
}

int qpc_build_hc(qpc_ctx* c, const double* V, const double* K, const double* T, double mass) {
    if (!c || !V || !K || !T || !(mass > 0.0)) return QPC_E_ARG;
    if (!c->have_radial) {
        c->set_error("radial operators not uploaded");
        return QPC_E_STATE;
    }
    c->step_slot = -1;
    const int64_t band = c->N * c->ne;
    const int nl = (int)(c->lmax + 1);
    double* tmp = nullptr;
    QPC_HIP_CHECK(c, hipMalloc((void**)&tmp, sizeof(double) * 3 * band + 256));
    int rc = [&]() -> int {
        int* tiny = (int*)(tmp + 3 * band);
        TRY(upload(c, tmp, V, (size_t)band));
        TRY(upload(c, tmp + band, K, (size_t)band));
        TRY(upload(c, tmp + 2 * band, T, (size_t)band));
        const int zero = 0;
        TRY(upload(c, tiny, &zero, 1));
        qpc::launch_hc_assemble(c->stream, c->S, tmp, tmp + band, tmp + 2 * band, 0.5 / mass, 0.5 * c->d.para_dt,
                                c->upper, c->lower, nl, band);
        qpc::launch_ldlt_csym(c->stream, c->lower, nl, (int)c->N, (int)c->p, tiny);
        TRY(launch_check(c, "field-free bands"));
        int ntiny = 0;
        QPC_HIP_CHECK(c, hipMemcpyAsync(&ntiny, tiny, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        if (ntiny) {
            c->set_error("singular field-free CN band (lu_zsb pivot)");
            return QPC_E_NUMERIC;
        }
        return install_hc(c);
    }();
    (void)hipFree(tmp);
    return rc;
}

// the R-part table map (qpc_ctx::rmap) from the uploaded eigenvalues
static int plan_rpairs(qpc_ctx* c) {
    void* old[] = {c->rmap, c->qeig_prim, c->rorder};
    for (void* q : old)
        if (q) QPC_HIP_CHECK(c, hipFree(q));
    c->rmap = nullptr;
    c->qeig_prim = nullptr;
    c->rorder = nullptr;
    c->rorder_nseg = 0;
    c->n_rprim = 0;
    c->rprim_ch.clear();
    if (!c->pair_r || (int64_t)c->qeig_h.size() != c->ML) return QPC_OK;
    // |eigenvalue| (and, in a batch context, the shot: its own A(t)) -> table
    std::map<std::pair<int64_t, double>, int> idx;
    std::vector<double> prim;
    std::vector<int> map((size_t)c->ML);
    for (int64_t ch = 0; ch < c->ML; ++ch) {
        const double q = c->qeig_h[ch];
        const std::pair<int64_t, double> key{c->batch ? ch / c->L : 0, std::fabs(q)};
        const auto it = idx.find(key);
        if (it == idx.end()) {
            idx.emplace(key, (int)prim.size());
            map[ch] = (int)prim.size();
            prim.push_back(q);
            c->rprim_ch.push_back(ch);
        } else {
            map[ch] = it->second | (q != prim[it->second] ? qpc::kMapConj : 0);
        }
    }
    if ((int64_t)prim.size() == c->ML) return QPC_OK;  // nothing shared
    c->rtab_h.resize(map.size());
    for (size_t ch = 0; ch < map.size(); ++ch) c->rtab_h[ch] = map[ch] & qpc::kMapIdx;
    c->rorder_nseg = 0;
    TRY(dev_alloc(c, &c->rmap, map.size()));
    TRY(dev_alloc(c, &c->qeig_prim, prim.size()));
    TRY(upload(c, c->rmap, map.data(), map.size()));
    TRY(upload(c, c->qeig_prim, prim.data(), prim.size()));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    c->n_rprim = (int64_t)prim.size();
    return QPC_OK;
}

int qpc_upload_angular(qpc_ctx* c, const double* yeig, const double* yprj, const double* qeig, const double* qprj,
                       const double* merged, const double* y10) {
    if (!c || !yeig || !yprj || !qeig || !qprj || !merged) return QPC_E_ARG;
    c->step_slot = -1;  // tables of the previous operators
    const size_t ml = (size_t)c->ML, mll = (size_t)c->ML * c->L;
    TRY(upload(c, c->yeig, yeig, ml));
    c->yeig_h.assign(yeig, yeig + ml);
    c->qeig_h.assign(qeig, qeig + ml);
    c->yeig_max = 0.0;
    for (size_t k = 0; k < ml; ++k) c->yeig_max = std::max(c->yeig_max, std::fabs(yeig[k]));
    TRY(upload(c, c->qeig, qeig, ml));
    TRY(upload(c, c->yprj, yprj, mll));
    TRY(upload(c, c->qprj, qprj, mll));
    TRY(upload(c, c->merged, merged, mll));
    if (y10) TRY(upload(c, c->y10, y10, ml));
    qpc::launch_rot_prep_real(c->stream, c->yprj, c->rtY, c->M, c->L, 0);
    qpc::launch_rot_prep_real(c->stream, c->yprj, c->rtYt, c->M, c->L, 1);
    qpc::launch_rot_prep_cplx(c->stream, c->merged, c->rtM, c->M, c->L, 0);
    qpc::launch_rot_prep_cplx(c->stream, c->merged, c->rtMh, c->M, c->L, 1);
    qpc::launch_rot_prep_cplx(c->stream, c->qprj, c->rtQ, c->M, c->L, 0);
    qpc::launch_rot_prep_cplx(c->stream, c->qprj, c->rtQh, c->M, c->L, 1);
    TRY(launch_check(c, "rotation operators"));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    TRY(plan_rpairs(c));
    c->have_ang = true;
    return QPC_OK;
}

int qpc_set_state(qpc_ctx* c, const double* coef) {
    if (!c || !coef) return QPC_E_ARG;
    TRY(upload(c, c->coef, coef, (size_t)c->ML * c->N));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    c->have_state = true;
    return QPC_OK;
}

int qpc_get_state(qpc_ctx* c, double* coef) {
    if (!c || !coef) return QPC_E_ARG;
    QPC_HIP_CHECK(c, hipMemcpyAsync(coef, c->coef, sizeof(cplx) * c->ML * c->N, hipMemcpyDeviceToHost, c->stream));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return QPC_OK;
}

int qpc_step(qpc_ctx* c, double a_i, double a_f) {
    if (!c || c->batch) return QPC_E_ARG;
    return step_impl(c, 0, a_i, a_f);
}

int qpc_step_lg(qpc_ctx* c, double e_i, double e_f) {
    if (!c || c->batch) return QPC_E_ARG;
    if (c->rpow1_h.empty()) {
        c->set_error("<Bi|r|Bj> (rpow1) not uploaded");
        return QPC_E_STATE;
    }
    return step_impl(c, 1, e_i, e_f);
}

int qpc_step_free(qpc_ctx* c) {
    if (!c) return QPC_E_ARG;
    TRY(ready(c, false));
    return do_free(c);
}

int qpc_mask(qpc_ctx* c) {
    if (!c) return QPC_E_ARG;
    TRY(ready(c, true));
    return do_mask(c);
}

int qpc_run(qpc_ctx* c, const double* atab, int64_t nt_field, int64_t it0, int64_t it1, int use_mask) {
    if (!c || c->batch) return QPC_E_ARG;
    // earlier calls may still be reading the run scalars (the gauge selects them)
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->fstream));
    c->gauge = 0;
    return run_impl(c, atab, 0, nt_field, it0, it1, use_mask);
}

int qpc_run_lg(qpc_ctx* c, const double* etab, int64_t nt_field, int64_t it0, int64_t it1, int use_mask) {
    if (!c || c->batch) return QPC_E_ARG;
    if (c->tsf_on) {
        c->set_error("t-SURFF needs the velocity gauge");  // sphTDSE.cpp:288
        return QPC_E_STATE;
    }
    if (c->rpow1_h.empty()) {
        c->set_error("<Bi|r|Bj> (rpow1) not uploaded");
        return QPC_E_STATE;
    }
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->fstream));
    c->gauge = 1;
    return run_impl(c, etab, 0, nt_field, it0, it1, use_mask);
}

// spectrum_c_tsf::initialize's tables on the device (momentum_1e.hpp:779-880)
// and clearwork (:580-589); Y = y0 e0 (-i)^(l+1) channel-major, cos(theta)
// per (theta, phi) point
int qpc_upload_tsurff(qpc_ctx* c, const qpc_tsurff_cfg* tc, const double* kr, const double* kt, const double* kp,
                      int64_t nb, const int64_t* ib, const double* b0, const double* br, const double* j0,
                      const double* jr, const double* y0, const double* e0, const double* pmat0) {
    if (!c || !tc || !kr || !kt || !kp || !ib || !b0 || !br || !j0 || !jr || !y0 || !e0 || !pmat0) return QPC_E_ARG;
    if (c->batch || tc->nkr < 1 || tc->nkt < 1 || tc->nkp < 1 || nb < 1 || nb > c->p + 1) return QPC_E_ARG;
    if (c->ML > 4096 || tc->nkr > (1 << 30) || tc->nkt * tc->nkp > (1 << 30)) {
        c->set_error("t-SURFF is built for at most 4096 channels");
        return QPC_E_SHAPE;
    }
    for (int64_t i = 0; i < nb; ++i)
        if (ib[i] < 0 || ib[i] >= c->N) return QPC_E_ARG;
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    void* old[] = {c->tsf_ib, c->tsf_b0, c->tsf_br, c->tsf_kr,  c->tsf_ckt, c->tsf_j0,  c->tsf_jr, c->tsf_pm,
                   c->tsf_Y,  c->tsf_tmp, c->tsf_work, c->tsf_stx, c->tsf_sty, c->tsf_pm1, c->tsf_pm2};
    for (void* p : old)
        if (p) QPC_HIP_CHECK(c, hipFree(p));
    c->tsf_ib = nullptr;
    c->tsf_b0 = c->tsf_br = c->tsf_kr = c->tsf_ckt = c->tsf_j0 = c->tsf_jr = c->tsf_pm = nullptr;
    c->tsf_stx = c->tsf_sty = c->tsf_pm1 = c->tsf_pm2 = nullptr;
    c->tsf_Y = c->tsf_tmp = c->tsf_work = nullptr;
    c->have_tsf = c->tsf_on = c->have_tsf_xy = false;
    const int64_t nkr = tc->nkr, nkt = tc->nkt, nkp = tc->nkp, nktp = nkt * nkp, ML = c->ML, nl = c->lmax + 1;
    std::vector<int> ibh(ib, ib + nb);
    std::vector<double> ckt((size_t)nktp), stx((size_t)nktp), sty((size_t)nktp), pm((size_t)ML);
    std::vector<cplx> Y((size_t)ML * nktp);
    static const cplx pmi[4] = {{0.0, -1.0}, {-1.0, 0.0}, {0.0, 1.0}, {1.0, 0.0}};  // (-i)^(l+1)
    for (int64_t ikt = 0; ikt < nkt; ++ikt)
        for (int64_t ikp = 0; ikp < nkp; ++ikp) {
            ckt[ikt * nkp + ikp] = std::cos(kt[ikt]);
            stx[ikt * nkp + ikp] = std::sin(kt[ikt]) * std::cos(kp[ikp]);
            sty[ikt * nkp + ikp] = std::sin(kt[ikt]) * std::sin(kp[ikp]);
        }
    const cplx* e = (const cplx*)e0;
    for (int64_t im = 0; im < c->M; ++im)
        for (int64_t il = 0; il < c->L; ++il) {
            const int64_t ch = im * c->L + il, m = c->m0 + im, l = il + (m < 0 ? -m : m);
            pm[ch] = pmat0[ch];
            for (int64_t ikt = 0; ikt < nkt; ++ikt)
                for (int64_t ikp = 0; ikp < nkp; ++ikp)
                    Y[ch * nktp + ikt * nkp + ikp] = (y0[(ikt * c->M + im) * c->L + il] * e[ikp * c->M + im]) * pmi[l % 4];
        }
    std::vector<cplx> zero((size_t)(nkr * nktp), cplx{0.0, 0.0});
    TRY(dev_alloc(c, &c->tsf_ib, (size_t)nb));
    TRY(dev_alloc(c, &c->tsf_b0, (size_t)nb));
    TRY(dev_alloc(c, &c->tsf_br, (size_t)nb));
    TRY(dev_alloc(c, &c->tsf_kr, (size_t)nkr));
    TRY(dev_alloc(c, &c->tsf_ckt, (size_t)nktp));
    TRY(dev_alloc(c, &c->tsf_stx, (size_t)nktp));
    TRY(dev_alloc(c, &c->tsf_sty, (size_t)nktp));
    TRY(dev_alloc(c, &c->tsf_pm1, (size_t)ML));
    TRY(dev_alloc(c, &c->tsf_pm2, (size_t)ML));
    TRY(dev_alloc(c, &c->tsf_j0, (size_t)(nkr * nl)));
    TRY(dev_alloc(c, &c->tsf_jr, (size_t)(nkr * nl)));
    TRY(dev_alloc(c, &c->tsf_pm, (size_t)ML));
    TRY(dev_alloc(c, &c->tsf_Y, (size_t)(ML * nktp)));
    TRY(dev_alloc(c, &c->tsf_tmp, (size_t)(2 * ML)));
    TRY(dev_alloc(c, &c->tsf_work, (size_t)(nkr * nktp)));
    TRY(upload(c, c->tsf_ib, ibh.data(), (size_t)nb));
    TRY(upload(c, c->tsf_b0, b0, (size_t)nb));
    TRY(upload(c, c->tsf_br, br, (size_t)nb));
    TRY(upload(c, c->tsf_kr, kr, (size_t)nkr));
    TRY(upload(c, c->tsf_ckt, ckt.data(), (size_t)nktp));
    TRY(upload(c, c->tsf_stx, stx.data(), (size_t)nktp));
    TRY(upload(c, c->tsf_sty, sty.data(), (size_t)nktp));
    TRY(upload(c, c->tsf_j0, j0, (size_t)(nkr * nl)));
    TRY(upload(c, c->tsf_jr, jr, (size_t)(nkr * nl)));
    TRY(upload(c, c->tsf_pm, pm.data(), (size_t)ML));
    TRY(upload(c, c->tsf_Y, Y.data(), (size_t)(ML * nktp)));
    TRY(upload(c, c->tsf_work, zero.data(), zero.size()));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    c->tsf_nkr = (int)nkr;
    c->tsf_nktp = (int)nktp;
    c->tsf_nb = (int)nb;
    c->tsf_time = c->tsf_disz = c->tsf_disx = c->tsf_disy = 0.0;
    c->have_tsf = true;
    return QPC_OK;
}

// pmat1, pmat2 of spectrum_c_tsf::initialize (integrate_sinexp1/2,
// momentum_1e.hpp:876-877) for the elliptically polarised accumulate
int qpc_upload_tsurff_xy(qpc_ctx* c, const double* pmat1, const double* pmat2) {
    if (!c || !pmat1 || !pmat2) return QPC_E_ARG;
    if (!c->have_tsf) {
        c->set_error("t-SURFF tables not uploaded");
        return QPC_E_STATE;
    }
    TRY(upload(c, c->tsf_pm1, pmat1, (size_t)c->ML));
    TRY(upload(c, c->tsf_pm2, pmat2, (size_t)c->ML));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    c->have_tsf_xy = true;
    return QPC_OK;
}

int qpc_tsurff(qpc_ctx* c, int on) {
    if (!c) return QPC_E_ARG;
    if (on && !c->have_tsf) {
        c->set_error("t-SURFF tables not uploaded");
        return QPC_E_STATE;
    }
    c->tsf_on = on != 0;
    return QPC_OK;
}

int qpc_tsurff_accumulate(qpc_ctx* c, double az0, double az1) {
    if (!c) return QPC_E_ARG;
    if (!c->have_tsf || !c->have_state) {
        c->set_error("t-SURFF tables or state not uploaded");
        return QPC_E_STATE;
    }
    return tsurff_step(c, az0, az1);
}

int qpc_tsurff_accumulate_xy(qpc_ctx* c, double ax0, double ax1, double ay0, double ay1) {
    if (!c) return QPC_E_ARG;
    if (!c->have_tsf_xy || !c->have_state) {
        c->set_error("t-SURFF xy tables or state not uploaded");
        return QPC_E_STATE;
    }
    return qpc_host::tsurff_step_xy(c, ax0, ax1, ay0, ay1);
}

int qpc_tsurff_displacement(qpc_ctx* c, double* dis) {
    if (!c || !dis) return QPC_E_ARG;
    dis[0] = c->tsf_disx;
    dis[1] = c->tsf_disy;
    dis[2] = c->tsf_disz;
    return QPC_OK;
}

int qpc_tsurff_clear(qpc_ctx* c) {
    if (!c) return QPC_E_ARG;
    if (!c->have_tsf) return QPC_E_STATE;
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    std::vector<cplx> zero((size_t)c->tsf_nkr * c->tsf_nktp, cplx{0.0, 0.0});
    TRY(upload(c, c->tsf_work, zero.data(), zero.size()));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    c->tsf_time = c->tsf_disz = c->tsf_disx = c->tsf_disy = 0.0;
    return QPC_OK;
}

int qpc_tsurff_read(qpc_ctx* c, double* work, double* time, double* disz) {
    if (!c || !work) return QPC_E_ARG;
    if (!c->have_tsf) return QPC_E_STATE;
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    QPC_HIP_CHECK(c, hipMemcpy(work, c->tsf_work, sizeof(cplx) * (size_t)c->tsf_nkr * c->tsf_nktp,
                               hipMemcpyDeviceToHost));
    if (time) *time = c->tsf_time;
    if (disz) *disz = c->tsf_disz;
    return QPC_OK;
}

int qpc_run_batch(qpc_ctx* c, const double* atab, int64_t stride, int64_t nt_field, int64_t it0, int64_t it1,
                  int use_mask) {
    if (!c || !c->batch || stride < nt_field + 1) return QPC_E_ARG;
    c->gauge = 0;
    return run_impl(c, atab, stride, nt_field, it0, it1, use_mask);
}

// launch-shape autotuner (no counterpart in the reference): the solves'
// segments per channel and staged-grid limit change how the work is spread,
// not the result, so the fastest of them for this context's grid and field is
// kept.  Each candidate runs the same steps from the same state (saved and
// restored, with observables and t-SURFF off meanwhile), timed with events on
// the step stream after one untimed run.
static const int kTuneSegAdd[] = {0, 1, 2, 3};
static const int64_t kTuneStage[] = {qpc::kStageGrid, 4096};
constexpr int kTuneN = 8;

int qpc_set_tune(qpc_ctx* c, const qpc_tune* t) {
    if (!c || !t || t->seg_add < 0 || t->seg_add > 64 || t->stage_grid < 0 || t->serial < 0 || t->serial > 1 ||
        t->share_r < 0 || t->share_r > 1)
        return QPC_E_ARG;
    c->geo.seg_add = t->seg_add;
    c->geo.stage_grid = t->stage_grid;
    c->geo.serial = t->serial;
    if (c->pair_r != (t->share_r != 0)) {
        // the R-part tables of the steps to come follow the new map
        QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        QPC_HIP_CHECK(c, hipStreamSynchronize(c->fstream));
        c->pair_r = t->share_r != 0;
        c->step_slot = -1;
        TRY(plan_rpairs(c));
    }
    return QPC_OK;
}

int qpc_get_tune(qpc_ctx* c, qpc_tune* t) {
    if (!c || !t) return QPC_E_ARG;
    t->seg_add = c->geo.seg_add;
    t->serial = c->geo.serial;
    t->stage_grid = c->geo.stage_grid;
    t->share_r = c->pair_r ? 1 : 0;
    t->tables_r = c->n_rprim > 0 ? (int32_t)c->n_rprim : (int32_t)c->ML;
    return QPC_OK;
}

int qpc_autotune(qpc_ctx* c, const double* atab, int64_t stride, int64_t nt_field, int64_t it0, int64_t nsteps,
                 qpc_tune* best, double* ms) {
    if (!c || !atab || nsteps < 1 || it0 < 0 || it0 + nsteps > nt_field || (c->batch && stride < nt_field + 1))
        return QPC_E_ARG;
    TRY(ready(c, true));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->fstream));
    const size_t n = (size_t)c->ML * c->N;
    DevTemp saved;
    TRY(dev_alloc(c, (cplx**)&saved.p, n));
    QPC_HIP_CHECK(c, hipMemcpyAsync(saved.p, c->coef, n * sizeof(cplx), hipMemcpyDeviceToDevice, c->stream));
    const int rec = c->rec_flags;
    const bool tsf = c->tsf_on;
    const int gauge = c->gauge;
    c->rec_flags = 0;
    c->tsf_on = false;
    c->gauge = 0;
    const qpc::Geom g0 = c->geo;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double tms[kTuneN];
    int rc = QPC_OK;
    auto trial = [&](int k, bool timed) -> int {
        c->geo.seg_add = kTuneSegAdd[k % 4];
        c->geo.stage_grid = kTuneStage[k / 4];
        QPC_HIP_CHECK(c, hipMemcpyAsync(c->coef, saved.p, n * sizeof(cplx), hipMemcpyDeviceToDevice, c->stream));
        QPC_HIP_CHECK(c, hipEventRecord(e0, c->stream));
        TRY(run_impl(c, atab, stride, nt_field, it0, it0 + nsteps, 1));
        QPC_HIP_CHECK(c, hipEventRecord(e1, c->stream));
        QPC_HIP_CHECK(c, hipEventSynchronize(e1));
        float f = 0.f;
        QPC_HIP_CHECK(c, hipEventElapsedTime(&f, e0, e1));
        if (timed) tms[k] = f / (double)nsteps;
        return QPC_OK;
    };
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = QPC_E_HIP;
    if (rc == QPC_OK) rc = trial(0, false);
    for (int k = 0; k < kTuneN && rc == QPC_OK; ++k) rc = trial(k, true);
    int kb = 0;
    if (rc == QPC_OK)
        for (int k = 1; k < kTuneN; ++k)
            if (tms[k] < tms[kb]) kb = k;
    c->geo = g0;
    if (rc == QPC_OK) {
        c->geo.seg_add = kTuneSegAdd[kb % 4];
        c->geo.stage_grid = kTuneStage[kb / 4];
    }
    // the state as it was
    (void)hipMemcpyAsync(c->coef, saved.p, n * sizeof(cplx), hipMemcpyDeviceToDevice, c->stream);
    (void)hipStreamSynchronize(c->stream);
    c->rec_flags = rec;
    c->tsf_on = tsf;
    c->gauge = gauge;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc != QPC_OK) return rc;
    if (best) TRY(qpc_get_tune(c, best));
    if (ms)
        for (int k = 0; k < kTuneN; ++k) ms[k] = tms[k];
    return QPC_OK;
}

int qpc_observe_batch(qpc_ctx* c, int what, double* out) {
    if (!c || !out || !c->batch) return QPC_E_ARG;
    if (!c->have_radial || !c->have_state) return QPC_E_STATE;
    if (what == QPC_OBS_NORM)
        qpc::launch_norm(c->stream, c->S, c->coef, c->red, c->ML, c->N, c->p);
    else if (what == QPC_OBS_DIPOLE_Z)
        qpc::launch_dipole_z(c->stream, c->rpow1, c->y10, c->coef, c->red, c->M, c->L, c->N, c->p);
    else
        return QPC_E_ARG;
    TRY(launch_check(c, "observe"));
    std::vector<double> h(c->ML);
    QPC_HIP_CHECK(c, hipMemcpyAsync(h.data(), c->red, sizeof(double) * c->ML, hipMemcpyDeviceToHost, c->stream));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    for (int64_t s = 0; s < c->M; ++s) {
        double v = 0.0;
        for (int64_t il = 0; il < c->L; ++il) v += h[s * c->L + il];
        out[s] = v;
    }
    return QPC_OK;
}

}  // extern "C"

namespace {
// the driver loop of qpc_run / qpc_run_batch (atab row s at atab + s*stride for batches)
int run_impl(qpc_ctx* c, const double* atab, int64_t stride, int64_t nt_field, int64_t it0, int64_t it1,
             int use_mask) {
    if (!c || (nt_field > 0 && !atab) || it0 < 0 || it1 < it0) return QPC_E_ARG;
    TRY(ready(c, use_mask != 0));
    c->step_slot = -1;  // the run overwrites the table sets and the scalars
    // earlier calls may still be reading the run scalars and the table sets
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->fstream));
    TRY(qpc_host::record_room(c, it1 - it0));
    const int64_t itf = std::min(it1, nt_field);  // with-field steps [it0, itf)
    if (itf > it0) {
        // run time point k is A at driver time point it0 + k (k = 0..itf-it0)
        TRY(alloc_zsets(c));
        TRY(set_run_scalars(c, atab + it0, itf - it0 + 1, stride));
        const int64_t B = c->tb_cap - 1;  // steps per batch
        const int64_t nb = (itf - it0 + B - 1) / B;
        bool in_y = false;  // the state is in P_Y's basis (do_step)
        TRY(factor_batch(c, 0, 0, std::min(B, itf - it0) + 1));
        for (int64_t k = 0; k < nb; ++k) {
            const int ts = (int)(k & 1);
            const int64_t s0 = k * B, n = std::min(B, itf - it0 - s0);
            if (k + 1 < nb) {
                const int64_t s1 = (k + 1) * B;
                TRY(factor_batch(c, 1 - ts, s1, std::min(B, itf - it0 - s1) + 1));
                if (c->geo.serial) TRY(use_set(c, 1 - ts));  // the batch runs alone, before these steps
            }
            TRY(use_set(c, ts));
            for (int64_t j = 0; j < n; ++j) {
                TRY(qpc_host::record(c));
                if (c->gauge == 1) {
                    TRY(do_step_lg(c, ts, s0, s0 + j, s0 + j + 1));
                    if (use_mask) TRY(do_mask(c));
                    continue;
                }
                // hand over in P_Y's basis while the next step is a with-field
                // step of this run and nothing observes the state in between
                // (do_step)
                const bool next = (j + 1 < n || k + 1 < nb) && !c->rec_flags && !c->tsf_on;
                TRY(do_step(c, ts, s0, s0 + j, s0 + j + 1, in_y, next, use_mask != 0));
                if (c->tsf_on) {
                    const int64_t it = it0 + s0 + j;  // sphTDSE.cpp:798: after the step, before the mask
                    TRY(tsurff_step(c, atab[it], atab[it + 1]));
                }
                if (!next && use_mask) TRY(do_mask(c));
                in_y = next;
            }
            TRY(release_set(c, ts));
        }
    }
    for (int64_t it = std::max(it0, itf); it < it1; ++it) {
        TRY(qpc_host::record(c));
        TRY(do_free(c));
        if (c->tsf_on) TRY(tsurff_step(c, 0.0, 0.0));  // sphTDSE.cpp:803
        if (use_mask) TRY(do_mask(c));
    }
    return QPC_OK;
}
// the per-step entry points (qpc_step, qpc_step_lg): a driver loop hands
// each call A(t_i), A(t_f) with A(t_i) the previous call's A(t_f), so the
// tables of that time point are kept and a call factorises A(t_f) alone,
// into the one of three positions of table set 0 that the previous step did
// not read: that factorisation waits only for the step before the previous
// one (event slot_read) and so runs beside the previous step.  Nothing
// waits on the host: the new scalars are written and the tables factorised
// on the second stream, the step waits for them with an event.
int step_impl(qpc_ctx* c, int gauge, double a_i, double a_f) {
    TRY(ready(c, false));
    TRY(alloc_zsets(c));
    for (auto& e : c->slot_read)
        if (!e) QPC_HIP_CHECK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    const bool reuse = c->step_slot >= 0 && c->step_gauge == gauge && c->gauge == gauge && a_i == c->step_a &&
                       c->run_cap >= 3 && c->h_dD.size() >= 3;
    c->gauge = gauge;
    int ti = 0, tf = 1;
    if (reuse) {
        ti = (int)c->step_slot;
        tf = (ti + 1) % 3;
        const double step = c->d.para_dt * 0.5;
        double dD = 0.0, dR = 0.0;
        if (gauge == 1) {
            dR = a_f * step / 2;  // set_run_scalars' length-gauge scalar
            if (!std::isfinite(dR)) {
                c->set_error("non-finite E(t)");
                return QPC_E_NUMERIC;
            }
            TRY(qpc_host::plan_l_part(c, std::fabs(dR) * c->yeig_max));
        } else {
            const double fzdt = (-a_f) * step;
            dD = -fzdt / 2;
            dR = fzdt / 2;
            if (!std::isfinite(fzdt)) {
                c->set_error("non-finite A(t)");
                return QPC_E_NUMERIC;
            }
            TRY(qpc_host::plan_d_part(c, std::fabs(dD) * c->yeig_max));
        }
        c->h_dD[tf] = dD;
        c->h_dR[tf] = dR;
        if (c->slot_rec[tf]) QPC_HIP_CHECK(c, hipStreamWaitEvent(c->fstream, c->slot_read[tf], 0));
        hipLaunchKernelGGL(k_put2, dim3(1), dim3(64), 0, c->fstream, c->d_ifdt + tf, dD, c->d_ifdt + c->run_cap + tf,
                           dR);
        TRY(launch_check(c, "step scalars"));
        QPC_HIP_CHECK(c, hipEventRecord(c->put_done, c->fstream));
        TRY(factor_batch(c, 0, tf, 1, tf, false, true));
    } else {
        // a fresh pair: scalars of both time points (and room for a third;
        // the streams are idle after set_run_scalars), tables at positions 0
        // and 1 once every step that read the set is done
        const double a[3] = {a_i, a_f, a_f};
        TRY(set_run_scalars(c, a, 3));
        TRY(factor_batch(c, 0, 0, 2));
    }
    // a fresh pair waits for both time points' tables; with the t_i tables
    // kept from the previous call (already waited for), the step waits for
    // the t_f tables only where it first reads them
    hipEvent_t wr = nullptr, wd = nullptr;
    if (reuse) {
        wr = c->ts[0].ready_r;
        wd = c->ts[0].ready;
    } else {
        TRY(use_set(c, 0));
    }
    if (gauge == 1) TRY(do_step_lg(c, 0, 0, ti, tf, wr));
    else TRY(do_step(c, 0, 0, ti, tf, false, false, false, wr, wd));
    for (int k : {ti, tf}) {
        QPC_HIP_CHECK(c, hipEventRecord(c->slot_read[k], c->stream));
        c->slot_rec[k] = true;
    }
    TRY(release_set(c, 0));
    c->step_slot = tf;
    c->step_a = a_f;
    c->step_gauge = gauge;
    return QPC_OK;
}
}  // namespace

extern "C" {

int qpc_sync(qpc_ctx* c) {
    if (!c) return QPC_E_ARG;
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->fstream));
    return QPC_OK;
}

int qpc_observe(qpc_ctx* c, int what, double* out) {
    if (!c || !out) return QPC_E_ARG;
    if (!c->have_radial || !c->have_state) return QPC_E_STATE;
    if (what == QPC_OBS_NORM) {
        qpc::launch_norm(c->stream, c->S, c->coef, c->red, c->ML, c->N, c->p);
    } else if (what == QPC_OBS_OUTER) {
        if (!c->have_outer) {
            c->set_error("outer band not uploaded");
            return QPC_E_STATE;
        }
        qpc::launch_norm(c->stream, c->souter, c->coef, c->red, c->ML, c->N, c->p);
    } else if (what == QPC_OBS_DIPOLE_Z) {
        qpc::launch_dipole_z(c->stream, c->rpow1, c->y10, c->coef, c->red, c->M, c->L, c->N, c->p);
    } else if (what == QPC_OBS_ACCEL_Z) {
        if (!c->have_accel) {
            c->set_error("acceleration band not uploaded");
            return QPC_E_STATE;
        }
        qpc::launch_dipole_z(c->stream, c->accel, c->y10, c->coef, c->red, c->M, c->L, c->N, c->p);
    } else if (what >= QPC_OBS_DIPOLE_X && what <= QPC_OBS_ACCEL_Y) {
        const bool acc = what >= QPC_OBS_ACCEL_X;
        if (acc && !c->have_accel) {
            c->set_error("acceleration band not uploaded");
            return QPC_E_STATE;
        }
        // observable_r: per-channel partials and their fixed-order sum on the device
        if (!c->obs_tmp) TRY(dev_alloc(c, &c->obs_tmp, (size_t)qpc::kObs * (c->ML + 1)));
        const qpc::ObsBands b{c->S, c->rpow1, c->accel, c->souter, c->y10};
        double* out_d = c->obs_tmp + (size_t)qpc::kObs * c->ML;
        qpc::launch_observe(c->stream, b, acc ? QPC_REC_ACCEL_XY : QPC_REC_DIPOLE_XY, c->coef, c->obs_tmp, out_d, c->M,
                            c->L, c->m0, c->N, c->p);
        TRY(launch_check(c, "observe"));
        double h[qpc::kObs];
        QPC_HIP_CHECK(c, hipMemcpyAsync(h, out_d, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        *out = h[what];  // slots 4..7 are QPC_OBS_DIPOLE_X .. QPC_OBS_ACCEL_Y
        return QPC_OK;
    } else {
        return QPC_E_ARG;
    }
    TRY(launch_check(c, "observe"));
    std::vector<double> h(c->ML);
    QPC_HIP_CHECK(c, hipMemcpyAsync(h.data(), c->red, sizeof(double) * c->ML, hipMemcpyDeviceToHost, c->stream));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    double s = 0.0;
    for (int64_t i = 0; i < c->ML; ++i) s += h[i];  // fixed order: reproducible run to run
    *out = s;
    return QPC_OK;
}

int qpc_upload_outer(qpc_ctx* c, const double* band) {
    if (!c || !band) return QPC_E_ARG;
    const size_t n = (size_t)c->N * c->ne;
    if (!c->souter) TRY(dev_alloc(c, &c->souter, n));
    TRY(upload(c, c->souter, band, n));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    c->have_outer = true;
    return QPC_OK;
}

int qpc_upload_accel(qpc_ctx* c, const double* band) {
    if (!c || !band) return QPC_E_ARG;
    const size_t n = (size_t)c->N * c->ne;
    if (!c->accel) TRY(dev_alloc(c, &c->accel, n));
    TRY(upload(c, c->accel, band, n));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    c->have_accel = true;
    return QPC_OK;
}

int qpc_record(qpc_ctx* c, int flags, int64_t cap) {
    if (!c || flags < 0 || flags > QPC_REC_ALL || cap < 0 || (flags && cap < 1) || (flags && c->batch))
        return QPC_E_ARG;
    if (((flags & (QPC_REC_ACCEL | QPC_REC_ACCEL_XY)) && !c->have_accel) || ((flags & QPC_REC_OUTER) && !c->have_outer)) {
        c->set_error("band of a recorded observable not uploaded");
        return QPC_E_STATE;
    }
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    if (c->hist) QPC_HIP_CHECK(c, hipFree(c->hist));
    if (c->red4) QPC_HIP_CHECK(c, hipFree(c->red4));
    c->hist = c->red4 = nullptr;
    c->rec_flags = 0;
    c->rec_cap = c->rec_n = 0;
    if (!flags) return QPC_OK;
    TRY(dev_alloc(c, &c->hist, (size_t)qpc::kObs * cap));
    TRY(dev_alloc(c, &c->red4, (size_t)qpc::kObs * c->ML));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    c->rec_flags = flags;
    c->rec_cap = cap;
    return QPC_OK;
}

int qpc_record_read(qpc_ctx* c, double* out, int64_t* n) {
    if (!c || !n) return QPC_E_ARG;
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    *n = c->rec_n;
    if (c->rec_n && !out) return QPC_E_ARG;
    if (c->rec_n)
        QPC_HIP_CHECK(c, hipMemcpy(out, c->hist, sizeof(double) * qpc::kObs * c->rec_n, hipMemcpyDeviceToHost));
    return QPC_OK;
}

int qpc_upload_projection(qpc_ctx* c, int64_t nsel, const double* W) {
    if (!c || !W || nsel < 1 || nsel > c->N || c->batch) return QPC_E_ARG;
    const int nl = (int)(c->lmax + 1);
    // channels of every l: l = il + |m0 + im| (dimension/1e.hpp:109-112)
    std::vector<std::vector<int>> lst(nl);
    for (int64_t im = 0; im < c->M; ++im)
        for (int64_t il = 0; il < c->L; ++il) lst[il + std::llabs(c->m0 + im)].push_back((int)(im * c->L + il));
    int cmax = 1;
    for (auto& v : lst) cmax = std::max(cmax, (int)v.size());
    std::vector<int> chan((size_t)nl * cmax, 0), cnt(nl, 0);
    for (int l = 0; l < nl; ++l) {
        cnt[l] = (int)lst[l].size();
        for (size_t k = 0; k < lst[l].size(); ++k) chan[(size_t)l * cmax + k] = lst[l][k];
    }
    void* old[] = {c->Wp, c->ppart, c->pchan, c->pcnt, c->pout};
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    for (void* p : old)
        if (p) QPC_HIP_CHECK(c, hipFree(p));
    c->Wp = c->ppart = nullptr;
    c->pchan = c->pcnt = nullptr;
    c->pout = nullptr;
    c->proj_nsel = 0;
    c->pg = qpc::make_proj_geom(nl, (int)nsel, (int)c->N, cmax);
    double* Wtmp = nullptr;
    const size_t nW = (size_t)nl * nsel * c->N;
    int rc = QPC_OK;
    if ((rc = dev_alloc(c, &c->Wp, (size_t)qpc::proj_prep_elems(c->pg))) ||
        (rc = dev_alloc(c, &c->ppart, (size_t)qpc::proj_part_elems(c->pg))) ||
        (rc = dev_alloc(c, &c->pchan, chan.size())) || (rc = dev_alloc(c, &c->pcnt, cnt.size())) ||
        (rc = dev_alloc(c, &c->pout, (size_t)c->ML * nsel)) || (rc = dev_alloc(c, &Wtmp, nW)))
        return rc;
    DevTemp guard{Wtmp};
    TRY(upload(c, Wtmp, W, nW));
    TRY(upload(c, c->pchan, chan.data(), chan.size()));
    TRY(upload(c, c->pcnt, cnt.data(), cnt.size()));
    qpc::launch_proj_prep(c->stream, Wtmp, c->Wp, c->pg);
    TRY(launch_check(c, "projection basis"));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    c->proj_nsel = nsel;
    return QPC_OK;
}

int qpc_upload_pcs(qpc_ctx* c, int64_t nkr, const double* ijl, const double* cps) {
    if (!c || !ijl || !cps || nkr < 1 || c->batch) return QPC_E_ARG;
    const int nl = (int)(c->lmax + 1);
    std::vector<std::vector<int>> lst(nl);
    for (int64_t im = 0; im < c->M; ++im)
        for (int64_t il = 0; il < c->L; ++il) lst[il + std::llabs(c->m0 + im)].push_back((int)(im * c->L + il));
    int cmax = 1;
    for (auto& v : lst) cmax = std::max(cmax, (int)v.size());
    std::vector<int> chan((size_t)nl * cmax, 0), cnt(nl, 0);
    for (int l = 0; l < nl; ++l) {
        cnt[l] = (int)lst[l].size();
        for (size_t k = 0; k < lst[l].size(); ++k) chan[(size_t)l * cmax + k] = lst[l][k];
    }
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    void* old[] = {c->pcs_Wp, c->pcs_part, c->pcs_chan, c->pcs_cnt, c->pcs_cps, c->pcs_out, c->pcs_lmd};
    for (void* p : old)
        if (p) QPC_HIP_CHECK(c, hipFree(p));
    c->pcs_Wp = c->pcs_part = nullptr;
    c->pcs_chan = c->pcs_cnt = nullptr;
    c->pcs_cps = c->pcs_out = c->pcs_lmd = nullptr;
    c->pcs_nkr = 0;
    c->sg = qpc::make_proj_geom(nl, (int)nkr, (int)c->N, cmax);
    const size_t nW = (size_t)nkr * nl * c->N;
    double* Wtmp = nullptr;
    int rc = QPC_OK;
    if ((rc = dev_alloc(c, &c->pcs_Wp, (size_t)qpc::proj_prep_elems(c->sg))) ||
        (rc = dev_alloc(c, &c->pcs_part, (size_t)qpc::proj_part_elems(c->sg))) ||
        (rc = dev_alloc(c, &c->pcs_chan, chan.size())) || (rc = dev_alloc(c, &c->pcs_cnt, cnt.size())) ||
        (rc = dev_alloc(c, &c->pcs_cps, (size_t)nkr * nl)) || (rc = dev_alloc(c, &c->pcs_out, (size_t)c->ML * nkr)) ||
        (rc = dev_alloc(c, &c->pcs_lmd, (size_t)c->ML * nkr)) || (rc = dev_alloc(c, &Wtmp, nW)))
        return rc;
    DevTemp guard{Wtmp};
    TRY(upload(c, Wtmp, ijl, nW));
    TRY(upload(c, c->pcs_cps, cps, (size_t)nkr * nl));
    TRY(upload(c, c->pcs_chan, chan.data(), chan.size()));
    TRY(upload(c, c->pcs_cnt, cnt.data(), cnt.size()));
    // ijl[ikr][l][j] read with the l stride N and the ikr stride NL * N
    qpc::launch_proj_prep_strided(c->stream, Wtmp, c->pcs_Wp, c->sg, c->N, (int64_t)nl * c->N);
    TRY(launch_check(c, "pcs basis"));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    c->pcs_nkr = nkr;
    return QPC_OK;
}

int qpc_measure_lmd(qpc_ctx* c, double* lmd) {
    if (!c || !lmd) return QPC_E_ARG;
    if (!c->pcs_nkr || !c->have_state) {
        c->set_error("Coulomb-wave basis or state not uploaded");
        return QPC_E_STATE;
    }
    {
        TimedLaunch tl(c, QPC_K_PROJ);
        qpc::launch_project(c->stream, c->pcs_Wp, c->coef, c->pcs_chan, c->pcs_cnt, c->pcs_part, c->pcs_out, c->sg);
        qpc::launch_lmd_finish(c->stream, c->pcs_out, c->pcs_cps, c->pcs_lmd, c->pcs_nkr, c->M, c->L, c->m0,
                               c->lmax + 1);
    }
    TRY(launch_check(c, "measure_lmd"));
    QPC_HIP_CHECK(c, hipMemcpyAsync(lmd, c->pcs_lmd, sizeof(cplx) * c->ML * c->pcs_nkr, hipMemcpyDeviceToHost,
                                    c->stream));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return QPC_OK;
}

int qpc_project(qpc_ctx* c, double* proj, int64_t stride) {
    if (!c || !proj || stride < c->proj_nsel) return QPC_E_ARG;
    if (!c->proj_nsel || !c->have_state) {
        c->set_error("projection basis or state not uploaded");
        return QPC_E_STATE;
    }
    {
        TimedLaunch tl(c, QPC_K_PROJ);
        qpc::launch_project(c->stream, c->Wp, c->coef, c->pchan, c->pcnt, c->ppart, c->pout, c->pg);
    }
    TRY(launch_check(c, "project"));
    const int64_t ns = c->proj_nsel;
    std::vector<cplx> h((size_t)c->ML * ns);
    QPC_HIP_CHECK(c, hipMemcpyAsync(h.data(), c->pout, sizeof(cplx) * h.size(), hipMemcpyDeviceToHost, c->stream));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    for (int64_t ch = 0; ch < c->ML; ++ch) {
        double* o = proj + 2 * ch * stride;
        for (int64_t n = 0; n < ns; ++n) {
            o[2 * n] = h[(size_t)(ch * ns + n)].re;
            o[2 * n + 1] = h[(size_t)(ch * ns + n)].im;
        }
        for (int64_t n = ns; n < stride; ++n) o[2 * n] = o[2 * n + 1] = 0.0;
    }
    return QPC_OK;
}

int qpc_upload_krylov(qpc_ctx* c, const double* H, const double* y10, const double* q10) {
    if (!c || !H || !y10 || !q10 || c->batch) return QPC_E_ARG;
    const size_t nl = (size_t)c->lmax + 1, band = (size_t)c->N * c->ne, ml = (size_t)c->ML;
    int rc = QPC_OK;
    if (!c->Hd && ((rc = dev_alloc(c, &c->Hd, nl * band)) || (rc = dev_alloc(c, &c->q10, ml)))) return rc;
    double* tmp = nullptr;
    if ((rc = dev_alloc(c, &tmp, nl * band))) return rc;
    DevTemp guard{tmp};
    TRY(upload(c, tmp, H, nl * band));
    TRY(upload(c, c->y10, y10, ml));
    TRY(upload(c, c->q10, q10, ml));
    qpc::launch_band_diag_r(c->stream, tmp, c->Hd, (int)nl, (int)c->N, (int)c->p);
    TRY(launch_check(c, "krylov operators"));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    c->have_krylov = true;
    return QPC_OK;
}

namespace {
// sum_ml <v|S|v> with the per-channel partials of k_norm, summed in order
int snorm(qpc_ctx* c, const cplx* v, double* out) {
    qpc::launch_norm(c->stream, c->S, v, c->red, c->ML, c->N, c->p);
    TRY(launch_check(c, "snorm"));
    std::vector<double> h(c->ML);
    QPC_HIP_CHECK(c, hipMemcpyAsync(h.data(), c->red, sizeof(double) * c->ML, hipMemcpyDeviceToHost, c->stream));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    double s = 0.0;
    for (int64_t i = 0; i < c->ML; ++i) s += h[i];
    *out = s;
    return QPC_OK;
}

// propagate_sub (la_1e.hpp:157-182) for the leading m x m block of the
// Lanczos tridiagonal (a, b): xs = b0 Z e^{-i dt Lambda} Z^T e0
int lanczos_coefs(int64_t m, const std::vector<double>& a, const std::vector<double>& b, double dt,
                  std::vector<cplx>& xs) {
    std::vector<double> lam(m), z((size_t)m * m), off(m > 1 ? m - 1 : 1, 0.0);
    for (int64_t i = 0; i + 1 < m; ++i) off[i] = b[i + 1];
    if (qpc::tridiag_eigen((int)m, a.data(), off.data(), lam.data(), z.data())) return -1;
    std::vector<cplx> wk(m);
    xs.assign(m, cplx{0.0, 0.0});
    for (int64_t i = 0; i < m; ++i) {
        const double ph = -dt * lam[i];
        wk[i] = cplx{std::cos(ph), std::sin(ph)} * z[(size_t)i * m];
    }
    for (int64_t j = 0; j < m; ++j) {
        cplx t = {0.0, 0.0};
        for (int64_t i = 0; i < m; ++i) t += wk[i] * z[(size_t)i * m + j];
        xs[j] = t * b[0];
    }
    return 0;
}
inline double estimate(int64_t m, const std::vector<cplx>& xs) {
    return m > 1 ? xs[m - 1].re * xs[m - 1].re + xs[m - 1].im * xs[m - 1].im : 0.0;
}

// The Lanczos step of propagator_arnoldi::propagate1 (la_1e.hpp:312-335) on
// the current state.  tol < 0: the reference's loop (kmax vectors unless the
// S-norm breaks down).  tol >= 0: the subspace stops growing at the first
// m >= kstart whose estimate |x_{m-1}|^2 is <= tol.  ab (may be NULL)
// receives a_i, b_{i+1} for i < m.
int krylov_core(qpc_ctx* c, double az, double dt, int64_t kstart, int64_t kmax, double tol, int64_t* m_used,
                double* eps_out, double* ab) {
    if (kmax < 1 || kmax > 256) return QPC_E_ARG;
    if (!c->have_krylov || !c->have_radial || !c->have_state) {
        c->set_error("Krylov operators, radial operators or state not uploaded");
        return QPC_E_STATE;
    }
    const int64_t n = c->ML * c->N;
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->fstream));
    if (kmax + 2 > c->kry_cap) {
        QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        void* old[] = {c->kU, c->kpart, c->kcoef};
        for (void* p : old)
            if (p) QPC_HIP_CHECK(c, hipFree(p));
        c->kU = c->kpart = c->kcoef = nullptr;
        c->kry_cap = 0;
        int rc = QPC_OK;
        if ((rc = dev_alloc(c, &c->kU, (size_t)(kmax + 2) * n)) ||
            (rc = dev_alloc(c, &c->kpart, (size_t)(kmax + 2) * qpc::kDotBlocks)) ||
            (rc = dev_alloc(c, &c->kcoef, (size_t)(kmax + 2))))
            return rc;
        c->kry_cap = kmax + 2;
    }
    TimedLaunch tl(c, QPC_K_KRYLOV);
    auto U = [&](int64_t i) { return c->kU + i * n; };
    cplx* T = c->vbuf;  // H u_{k-1}
    std::vector<double> a(kmax + 2, 0.0), b(kmax + 2, 0.0);
    std::vector<cplx> x(kmax + 2), part((size_t)(kmax + 2) * qpc::kDotBlocks), xs;
    const qpc::TabMap tm{1, (int)c->L, (int)c->m0};
    // construct_sub (la_1e.hpp:134-154)
    QPC_HIP_CHECK(c, hipMemcpyAsync(U(0), c->coef, sizeof(cplx) * n, hipMemcpyDeviceToDevice, c->stream));
    double nrm = 0.0;
    TRY(snorm(c, U(0), &nrm));
    if (!(nrm > 0.0)) {
        c->set_error("zero state");
        return QPC_E_NUMERIC;
    }
    nrm = std::sqrt(nrm);
    qpc::launch_scale(c->stream, U(0), 1.0 / nrm, n);
    b[0] = nrm;
    // Lanczos loop (la_1e.hpp:318-328), orthogonalize_ip (:195-240)
    bool more = true;
    int64_t m = 0;
    while (more && m < kmax) {
        qpc::launch_hmul(c->stream, c->Hd, c->rinvd, c->rdifd, c->y10, c->q10, az, (int)c->M, (int)c->L, (int)c->m0,
                         (int)c->N, (int)c->p, U(m), T);
        const int64_t k = ++m;
        qpc::launch_dots(c->stream, c->kU, n, (int)k, T, n, c->kpart);
        TRY(launch_check(c, "krylov dots"));
        QPC_HIP_CHECK(c, hipMemcpyAsync(part.data(), c->kpart, sizeof(cplx) * k * qpc::kDotBlocks,
                                        hipMemcpyDeviceToHost, c->stream));
        // S \ (H u_{k-1}) into u_k, channel by channel (chol(S) tables)
        qpc::launch_bsolve_r(c->stream, qpc::RHS_RBAND, c->tab_mask, c->trf_mask, tm, T, U(k), (int)c->ML, c->geo);
        QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        std::vector<cplx> cf(k);
        for (int64_t i = 0; i < k; ++i) {
            cplx s = {0.0, 0.0};
            for (int q = 0; q < qpc::kDotBlocks; ++q) s += part[(size_t)i * qpc::kDotBlocks + q];
            x[i] = s;
            const bool use = i == k - 1 || (s.re * s.re + s.im * s.im) > 4e-27;
            cf[i] = use ? s : cplx{0.0, 0.0};
        }
        QPC_HIP_CHECK(c, hipMemcpyAsync(c->kcoef, cf.data(), sizeof(cplx) * k, hipMemcpyHostToDevice, c->stream));
        qpc::launch_axpy_multi(c->stream, c->kU, n, (int)k, c->kcoef, U(k), n);
        double nk = 0.0;
        TRY(snorm(c, U(k), &nk));  // synchronises: cf may be released after
        bool stop = false;
        if (nk < 1e-13) {
            nk = 0.0;
            stop = true;
        } else {
            nk = std::sqrt(nk);
            qpc::launch_scale(c->stream, U(k), 1.0 / nk, n);
        }
        a[k - 1] = x[k - 1].re;
        b[k] = nk;
        more = !stop;
        // subspace extension judged by the a-posteriori estimate
        if (more && tol >= 0.0 && k >= kstart) {
            if (lanczos_coefs(k, a, b, dt, xs)) {
                c->set_error("tridiagonal eigen-solve failed");
                return QPC_E_NUMERIC;
            }
            if (estimate(k, xs) <= tol) more = false;
        }
    }
    if (lanczos_coefs(m, a, b, dt, xs)) {
        c->set_error("tridiagonal eigen-solve failed");
        return QPC_E_NUMERIC;
    }
    // eliminate_sub (la_1e.hpp:185-193)
    QPC_HIP_CHECK(c, hipMemcpyAsync(c->kcoef, xs.data(), sizeof(cplx) * m, hipMemcpyHostToDevice, c->stream));
    qpc::launch_combine(c->stream, c->kU, n, (int)m, c->kcoef, c->coef, n);
    TRY(launch_check(c, "krylov combine"));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    if (m_used) *m_used = m;
    if (eps_out) *eps_out = estimate(m, xs);
    if (ab)
        for (int64_t i = 0; i < m; ++i) {
            ab[2 * i] = a[i];
            ab[2 * i + 1] = b[i + 1];
        }
    return QPC_OK;
}
}  // namespace

int qpc_krylov_step(qpc_ctx* c, double az, double dt, int64_t kmax, double emin, double emax, int* crit,
                    int64_t* m_used, double* eps) {
    if (!c || !crit) return QPC_E_ARG;
    int64_t m = 0;
    double e = 0.0;
    TRY(krylov_core(c, az, dt, kmax, kmax, -1.0, &m, &e, nullptr));
    // criterion (la_1e.hpp:296-310)
    *crit = e > emax ? 0 : (e > emin ? 1 : 2);
    if (m_used) *m_used = m;
    if (eps) *eps = e;
    return QPC_OK;
}

int qpc_krylov_step_adaptive(qpc_ctx* c, double az, double dt, int64_t kstart, int64_t kmax, double tol,
                             int64_t* m_used, double* eps, double* ab) {
    if (!c || kstart < 1 || kstart > kmax || !(tol >= 0.0)) return QPC_E_ARG;
    return krylov_core(c, az, dt, kstart, kmax, tol, m_used, eps, ab);
}

int qpc_upload_setup(qpc_ctx* c, const qpc_setup* st) {
    if (!c || !st) return QPC_E_ARG;
    const int64_t band = c->N * c->ne;
    if (qpc_setup_size(st, QPC_SETUP_S) != band || qpc_setup_size(st, QPC_SETUP_YEIG) != c->ML ||
        qpc_setup_size(st, QPC_SETUP_UPPER) != 2 * (c->lmax + 1) * band) {
        c->set_error("set-up shape does not match the context");
        return QPC_E_SHAPE;
    }
    auto get = [&](int what) {
        std::vector<double> v((size_t)qpc_setup_size(st, what));
        qpc_setup_get(st, what, v.data());
        return v;
    };
    std::vector<double> S = get(QPC_SETUP_S), ch = get(QPC_SETUP_CHOL), ri = get(QPC_SETUP_RINV),
                        rd = get(QPC_SETUP_RDIF), r1 = get(QPC_SETUP_RPOW1), mk = get(QPC_SETUP_MASK);
    TRY(qpc_upload_radial(c, S.data(), ch.data(), ri.data(), rd.data(), r1.data(), mk.data()));
    std::vector<double> up = get(QPC_SETUP_UPPER), lo = get(QPC_SETUP_LOWER);
    TRY(qpc_upload_hc(c, up.data(), lo.data()));
    std::vector<double> ye = get(QPC_SETUP_YEIG), yp = get(QPC_SETUP_YPRJ), qe = get(QPC_SETUP_QEIG),
                        qp = get(QPC_SETUP_QPRJ), mg = get(QPC_SETUP_MERGED), y10 = get(QPC_SETUP_Y10);
    TRY(qpc_upload_angular(c, ye.data(), yp.data(), qe.data(), qp.data(), mg.data(), y10.data()));
    std::vector<double> H = get(QPC_SETUP_H), q10 = get(QPC_SETUP_Q10), so = get(QPC_SETUP_SOUTER);
    TRY(qpc_upload_krylov(c, H.data(), y10.data(), q10.data()));
    TRY(qpc_upload_outer(c, so.data()));
    std::vector<double> dv = get(QPC_SETUP_DVDR);
    TRY(qpc_upload_accel(c, dv.data()));
    return QPC_OK;
}

int qpc_coupling_reach(const double* S, const double* X, int64_t N, int64_t p, double e, int asym, int* fwd,
                       int* bwd) {
    if (!S || !X || !fwd || !bwd || N < 2 * (p + 1) || p < 1 || p > 8 || !std::isfinite(e)) return QPC_E_ARG;
    const qpc::Reach r = qpc::solve_reach(S, X, (int)N, (int)p, e, asym ? 1 : 0);
    *fwd = r.fwd;
    *bwd = r.bwd;
    return QPC_OK;
}

int qpc_coupling_plan(const double* S, const double* D, int64_t N, int64_t p, double emax, int* pk, int* restart) {
    if (!S || !D || !pk || !restart || N < 2 * (p + 1) || p < 1 || p > 8 || !(emax >= 0.0)) return QPC_E_ARG;
    const qpc::DPlan d = qpc::d_part_plan(S, D, (int)N, (int)p, emax);
    *pk = d.pk;
    *restart = d.restart;
    return QPC_OK;
}

int qpc_timing_begin(qpc_ctx* c, int which, int64_t max_launches) {
    if (!c || which < 0 || which > QPC_K_EP_SWEEP || max_launches < 1) return QPC_E_ARG;
    while ((int64_t)c->ev.size() < 2 * max_launches) {
        hipEvent_t e;
        QPC_HIP_CHECK(c, hipEventCreate(&e));
        c->ev.push_back(e);
    }
    c->timed_kind = which;
    c->timed_max = max_launches;
    c->timed_n = 0;
    return QPC_OK;
}

int qpc_timing_end(qpc_ctx* c, int64_t* launches, double* mean_ms) {
    if (!c || !launches || !mean_ms) return QPC_E_ARG;
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->fstream));
    double tot = 0.0;
    for (int64_t i = 0; i < c->timed_n; ++i) {
        float ms = 0.f;
        QPC_HIP_CHECK(c, hipEventElapsedTime(&ms, c->ev[2 * i], c->ev[2 * i + 1]));
        tot += ms;
    }
    *launches = c->timed_n;
    *mean_ms = c->timed_n ? tot / (double)c->timed_n : 0.0;
    c->timed_kind = -1;
    c->timed_n = 0;
    return QPC_OK;
}

}  // extern "C"
