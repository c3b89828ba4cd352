// qpc_ep.hip -- the elliptically polarised (x-y) step of propagator_cn_hc_vgwr
// (reference propagator/cn_1e.hpp:812-951) on device: the (m,l)-space
// rotations are dense FP64 GEMMs on the matrix cores (rocBLAS), the eight
// coupling sweeps per step are the restarted-factorisation solves of
// qpc_solve.hip on the rotated channels, the field-free part and the
// absorber are those of the z-polarised path.
#include "qpc_ctx.h"

#include <cmath>

namespace {

// ---- row-major GEMMs through rocBLAS (column-major) --------------------------
// C(m x n) = op(A) B with all three row-major: the column-major view of a
// row-major matrix is its transpose, so C^T = B^T op(A)^T is one column-major
// call with the operands swapped.  op(A) = A (A is m x k) or A^H / A^T (A is
// k x m, stored row-major with leading dimension lda).
int gemm_z(qpc_ctx* c, bool herm, int64_t m, int64_t n, int64_t k, const cplx* A, int64_t lda, const cplx* B,
           int64_t ldb, cplx* C, int64_t ldc) {
    static_assert(sizeof(cplx) == sizeof(rocblas_double_complex), "complex layouts differ");
    const rocblas_double_complex one(1.0, 0.0), zero(0.0, 0.0);
    rocblas_status st = rocblas_zgemm(c->ep.blas, rocblas_operation_none,
                                      herm ? rocblas_operation_conjugate_transpose : rocblas_operation_none,
                                      (rocblas_int)n, (rocblas_int)m, (rocblas_int)k, &one,
                                      reinterpret_cast<const rocblas_double_complex*>(B), (rocblas_int)ldb,
                                      reinterpret_cast<const rocblas_double_complex*>(A), (rocblas_int)lda, &zero,
                                      reinterpret_cast<rocblas_double_complex*>(C), (rocblas_int)ldc);
    if (st != rocblas_status_success) {
        c->set_error(std::string("rocblas_zgemm: ") + rocblas_status_to_string(st));
        return QPC_E_HIP;
    }
    return QPC_OK;
}
// real operator on complex rows: each complex row is 2N doubles
int gemm_d(qpc_ctx* c, bool trans, int64_t m, int64_t k, const double* A, int64_t lda, const cplx* B, cplx* C) {
    const double one = 1.0, zero = 0.0;
    const int64_t n2 = 2 * c->N;
    rocblas_status st = rocblas_dgemm(c->ep.blas, rocblas_operation_none,
                                      trans ? rocblas_operation_transpose : rocblas_operation_none, (rocblas_int)n2,
                                      (rocblas_int)m, (rocblas_int)k, &one, reinterpret_cast<const double*>(B),
                                      (rocblas_int)n2, A, (rocblas_int)lda, &zero, reinterpret_cast<double*>(C),
                                      (rocblas_int)n2);
    if (st != rocblas_status_success) {
        c->set_error(std::string("rocblas_dgemm: ") + rocblas_status_to_string(st));
        return QPC_E_HIP;
    }
    return QPC_OK;
}

// cn_transform_hm::construct<-1,isub> / eliminate<isub,-1> of a real projector
// (cn_1e.hpp:222-273): rows [s0,s1) of dst from all rows of src, or all rows
// of dst from rows [s0,s1) of src
int rot_real_construct(qpc_ctx* c, const double* P, const cplx* src, cplx* dst) {
    const auto& E = c->ep;
    TimedLaunch t(c, QPC_K_EP_ROT);
    return gemm_d(c, false, E.s1 - E.s0, c->ML, P + E.s0 * c->ML, c->ML, src, dst + E.s0 * c->N);
}
int rot_real_eliminate(qpc_ctx* c, const double* P, const cplx* src, cplx* dst) {
    const auto& E = c->ep;
    TimedLaunch t(c, QPC_K_EP_ROT);
    return gemm_d(c, true, c->ML, E.s1 - E.s0, P + E.s0 * c->ML, c->ML, src + E.s0 * c->N, dst);
}
// construct<isub,isub> / eliminate<isub,isub> of a merged complex projector:
// the [s0,s1) x [s0,s1) block, P + s0 + s0*ML
int rot_cplx(qpc_ctx* c, bool herm, const cplx* P, const cplx* src, cplx* dst) {
    const auto& E = c->ep;
    const int64_t ns = E.s1 - E.s0, N = c->N;
    TimedLaunch t(c, QPC_K_EP_ROT);
    return gemm_z(c, herm, ns, N, ns, P + E.s0 + E.s0 * c->ML, c->ML, src + E.s0 * N, N, dst + E.s0 * N, N);
}

// operators of the sweeps: x1 and y1 are D-parts (Y11 with d/dr), x2 and y2
// R-parts (Q11 with 1/r); table index within a set: D tables [x1 ti, x1 tf,
// y1 ti, y1 tf], R tables [x2 ti, x2 tf, y2 ti, y2 tf]
enum EpOp { EP_X1 = 0, EP_X2 = 1, EP_Y1 = 2, EP_Y2 = 3 };
inline bool ep_is_d(int op) { return op == EP_X1 || op == EP_Y1; }
inline int ep_slot(int op, int tf) { return (op == EP_X1 || op == EP_X2 ? 0 : 2) + tf; }
// position of a sweep scalar within the 8 of a step
inline int ep_sc(int op, int tf) { return (ep_is_d(op) ? 0 : 4) + ep_slot(op, tf); }

// scalars of cn_hm_propagate1/2 (propagate.hpp:463, 525) for the run's steps:
// propagate<isub> halves A (cn_1e.hpp:905-908), f = -ax, +ax, -ay, -ay for
// x1, x2, y1, y2 (:911-942), frdt = f * dt/2 (prop step, :873-876);
// ifdt = -frdt/2 (d/dr sweeps) or frdt/2 (times i, 1/r sweeps).
int ep_set_scalars(qpc_ctx* c, const double* ax, const double* ay, int64_t nsteps) {
    auto& E = c->ep;
    const double step = c->d.para_dt * 0.5;
    E.h_sc.assign((size_t)nsteps * 8, 0.0);
    double emax = 0.0;
    for (int64_t k = 0; k < nsteps; ++k) {
        double* sc = &E.h_sc[(size_t)k * 8];
        for (int tf = 0; tf < 2; ++tf) {
            const double hx = ax[k + tf] / 2.0, hy = ay[k + tf] / 2.0;
            const double f[4] = {-hx, hx, -hy, -hy};
            for (int op = 0; op < 4; ++op) {
                const double frdt = f[op] * step;
                const double v = ep_is_d(op) ? -frdt / 2 : frdt / 2;
                sc[ep_sc(op, tf)] = v;
                if (!std::isfinite(v)) {
                    c->set_error("non-finite A(t)");
                    return QPC_E_NUMERIC;
                }
                if (ep_is_d(op)) emax = std::max(emax, std::fabs(v) * E.emax[op]);
            }
        }
    }
    // carry depth and restart of the D-part sweeps for this run's scalars
    TRY(qpc_host::plan_d_part(c, emax));
    if (nsteps > E.run_cap) {
        if (E.d_sc) QPC_HIP_CHECK(c, hipFree(E.d_sc));
        E.d_sc = nullptr;
        E.run_cap = 0;
        QPC_HIP_CHECK(c, hipMalloc((void**)&E.d_sc, sizeof(double) * 8 * nsteps));
        E.run_cap = nsteps;
    }
    QPC_HIP_CHECK(c, hipMemcpyAsync(E.d_sc, E.h_sc.data(), sizeof(double) * 8 * nsteps, hipMemcpyHostToDevice,
                                    c->stream));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return QPC_OK;
}

// factorisations of the eight sweeps of run step k into table set `set`, on
// the second stream once the step that last read the set is done
int ep_factor(qpc_ctx* c, int set, int64_t k) {
    auto& E = c->ep;
    auto& T = E.set[set];
    const int64_t np = E.p1 - E.p0, ML = c->ML;
    if (np <= 0) return QPC_OK;
    if (T.pending) QPC_HIP_CHECK(c, hipStreamWaitEvent(c->fstream, T.freed, 0));
    const double* sc = E.d_sc + k * 8;
    const int64_t te = c->geo.tab_elems, cte = c->geo.ctab_elems;
    {
        TimedLaunch tl(c, QPC_K_FACTOR, c->fstream);
        for (int op = 0; op < 4; ++op) {
            const double* eig = E.eig + op * ML + E.p0;
            const int64_t slot = ep_slot(op, 0);
            if (ep_is_d(op))
                qpc::launch_factor_asy(c->fstream, c->S, c->rdif, eig, sc + ep_sc(op, 0), 2, (int)np, c->geo,
                                       T.tabD + slot * np * te, 0, c->restart_d);
            else
                qpc::launch_factor_sym(c->fstream, c->S, c->rinv, eig, sc + ep_sc(op, 0), 2, (int)np, c->geo,
                                       T.tabR + slot * np * cte);
        }
        qpc::launch_transfer_r(c->fstream, T.tabD, T.trfD, (int)(4 * np), c->geo);
        qpc::launch_transfer_c(c->fstream, T.tabR, T.trfR, (int)(4 * np), c->geo);
    }
    TRY(launch_check(c, "ep factor"));
    QPC_HIP_CHECK(c, hipEventRecord(T.ready, c->fstream));
    return QPC_OK;
}

// cn_hm_propagate1/2<isub> (propagate.hpp:443-568) in place on the rotated
// channels [p0,p1) of buf: v = (S - e X) z into vbuf, then the solve back
int ep_sweep(qpc_ctx* c, int set, int64_t k, int op, int tf, cplx* buf) {
    auto& E = c->ep;
    const int64_t np = E.p1 - E.p0;
    if (np <= 0) return QPC_OK;
    const auto& T = E.set[set];
    const int64_t te = c->geo.tab_elems, cte = c->geo.ctab_elems, tr = c->geo.trf_elems;
    const int64_t slot = ep_slot(op, tf);
    qpc::RhsArgs ra{};
    ra.sx = ep_is_d(op) ? c->sxD : c->sxR;
    ra.eig = E.eig + op * c->ML + E.p0;
    ra.ifdt = E.h_sc[(size_t)k * 8 + ep_sc(op, tf)];
    const qpc::TabMap tm{0, (int)c->L, (int)c->m0};
    cplx* z = buf + E.p0 * c->N;
    const int kind = ep_is_d(op) ? qpc::RHS_SX_ASY : qpc::RHS_SX_SYM;
    // in place (z in, z out): the right-hand side goes through vbuf
    TimedLaunch tl(c, QPC_K_EP_SWEEP);
    if (ep_is_d(op))
        qpc::launch_rsolve_r(c->stream, kind, ra, T.tabD + slot * np * te, T.trfD + slot * np * tr, tm, z, c->vbuf, z,
                             (int)np, c->geo, c->pk_d);
    else
        qpc::launch_rsolve_c(c->stream, kind, ra, T.tabR + slot * np * cte, T.trfR + slot * np * tr, tm, z, c->vbuf, z,
                             (int)np, c->geo);
    return launch_check(c, "ep sweep");
}

// propagator_cn_hc_vgwr::propagate<isub> (cn_1e.hpp:898-945) for run step k
// whose tables are in `set`; the state ends in c->coef
int ep_do_step(qpc_ctx* c, int set, int64_t k) {
    auto& E = c->ep;
    cplx* X = c->coef;
    cplx* Y = c->tmpc;
    cplx* Z = c->spare;
    TRY(rot_real_construct(c, E.tx1, X, Y));   // tran_x1.construct<-1,isub>
    TRY(ep_sweep(c, set, k, EP_X1, 0, Y));     // prop_x1.propagate2(-ax_ti)
    TRY(rot_cplx(c, false, E.x1x2, Y, Z));     // x1_2_x2.construct
    TRY(ep_sweep(c, set, k, EP_X2, 0, Z));     // prop_x2.propagate1(ax_ti)
    TRY(rot_cplx(c, false, E.x2y1, Z, Y));     // x2_2_y1.construct
    TRY(ep_sweep(c, set, k, EP_Y1, 0, Y));     // prop_y1.propagate2(-ay_ti)
    TRY(rot_cplx(c, false, E.y1y2, Y, Z));     // y1_2_y2.construct
    TRY(ep_sweep(c, set, k, EP_Y2, 0, Z));     // prop_y2.propagate1(-ay_ti)
    TRY(rot_real_eliminate(c, E.ty2, Z, X));   // tran_y2.eliminate<isub,-1>
    TRY(qpc_host::solve_free(c, X, Y));        // prop_hc.propagate0
    TRY(rot_real_construct(c, E.ty2, Y, Z));   // tran_y2.construct<-1,isub>
    TRY(ep_sweep(c, set, k, EP_Y2, 1, Z));     // prop_y2.propagate1(-ay_tf)
    TRY(rot_cplx(c, true, E.y1y2, Z, Y));      // y1_2_y2.eliminate
    TRY(ep_sweep(c, set, k, EP_Y1, 1, Y));     // prop_y1.propagate2(-ay_tf)
    TRY(rot_cplx(c, true, E.x2y1, Y, Z));      // x2_2_y1.eliminate
    TRY(ep_sweep(c, set, k, EP_X2, 1, Z));     // prop_x2.propagate1(ax_tf)
    TRY(rot_cplx(c, true, E.x1x2, Z, Y));      // x1_2_x2.eliminate
    TRY(ep_sweep(c, set, k, EP_X1, 1, Y));     // prop_x1.propagate2(-ax_tf)
    TRY(rot_real_eliminate(c, E.tx1, Y, X));   // tran_x1.eliminate<isub,-1>
    return QPC_OK;
}

int ep_ready(qpc_ctx* c, bool need_mask) {
    TRY(qpc_host::check_ready(c, need_mask));
    if (!c->ep.have) {
        c->set_error("elliptically polarised angular data not uploaded");
        return QPC_E_STATE;
    }
    return QPC_OK;
}

}  // namespace

extern "C" {

int qpc_upload_angular_ep(qpc_ctx* c, int isub, const double* eig_x1, const double* eig_x2, const double* eig_y1,
                          const double* eig_y2, const double* tran_x1, const double* tran_y2, const double* x1_2_x2,
                          const double* x2_2_y1, const double* y1_2_y2) {
    if (!c || !eig_x1 || !eig_x2 || !eig_y1 || !eig_y2 || !tran_x1 || !tran_y2 || !x1_2_x2 || !x2_2_y1 || !y1_2_y2)
        return QPC_E_ARG;
    if (isub < -1 || isub > 3 || c->batch) return QPC_E_ARG;
    auto& E = c->ep;
    const int64_t ML = c->ML, M = c->M, L = c->L;
    // sub-block ranges of dimension_1e (dimension/1e.hpp:52-80)
    const int64_t n_zero = M % 2 == 0 ? 0 : L;
    const int64_t n_even = (M / 2 * 2) * ((L + 1) / 2);
    E.s0 = 0, E.s1 = ML, E.p0 = 0, E.p1 = ML;
    if (isub == 2) {
        E.s1 = n_even + n_zero;
        E.p1 = n_even;
    } else if (isub == 3) {
        E.s0 = n_even;
        E.p0 = n_even + n_zero;
    }
    E.isub = isub;
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->fstream));
    const size_t mm = (size_t)ML * ML;
    const int64_t np = E.p1 - E.p0;
    const size_t te = (size_t)c->geo.tab_elems, cte = (size_t)c->geo.ctab_elems, tr = (size_t)c->geo.trf_elems;
    int rc = QPC_OK;
    if (!E.eig) {
        if ((rc = dev_alloc(c, &E.eig, 4 * (size_t)ML)) || (rc = dev_alloc(c, &E.tx1, mm)) ||
            (rc = dev_alloc(c, &E.ty2, mm)) || (rc = dev_alloc(c, &E.x1x2, mm)) || (rc = dev_alloc(c, &E.x2y1, mm)) ||
            (rc = dev_alloc(c, &E.y1y2, mm)))
            return rc;
        for (auto& T : E.set) {
            QPC_HIP_CHECK(c, hipEventCreateWithFlags(&T.ready, hipEventDisableTiming));
            QPC_HIP_CHECK(c, hipEventCreateWithFlags(&T.freed, hipEventDisableTiming));
        }
        if (rocblas_create_handle(&E.blas) != rocblas_status_success ||
            rocblas_set_stream(E.blas, c->stream) != rocblas_status_success) {
            c->set_error("rocBLAS handle");
            return QPC_E_HIP;
        }
    }
    // sweep tables for the rotated channels of this sub-block (re-sized when isub changes)
    for (auto& T : E.set) {
        void* tp[] = {T.tabD, T.trfD, T.tabR, T.trfR};
        for (void* p : tp)
            if (p) QPC_HIP_CHECK(c, hipFree(p));
        T.tabD = T.trfD = nullptr;
        T.tabR = T.trfR = nullptr;
        T.pending = false;
        if (np > 0 && ((rc = dev_alloc(c, &T.tabD, 4 * np * te)) || (rc = dev_alloc(c, &T.trfD, 4 * np * tr)) ||
                       (rc = dev_alloc(c, &T.tabR, 4 * np * cte)) || (rc = dev_alloc(c, &T.trfR, 4 * np * tr))))
            return rc;
    }
    const double* e[4] = {eig_x1, eig_x2, eig_y1, eig_y2};
    for (int op = 0; op < 4; ++op) {
        TRY(upload(c, E.eig + op * ML, e[op], (size_t)ML));
        E.emax[op] = 0.0;
        for (int64_t i = E.p0; i < E.p1; ++i) E.emax[op] = std::max(E.emax[op], std::fabs(e[op][i]));
    }
    TRY(upload(c, E.tx1, tran_x1, mm));
    TRY(upload(c, E.ty2, tran_y2, mm));
    TRY(upload(c, E.x1x2, x1_2_x2, mm));
    TRY(upload(c, E.x2y1, x2_2_y1, mm));
    TRY(upload(c, E.y1y2, y1_2_y2, mm));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    E.have = true;
    return QPC_OK;
}

int qpc_step_ep(qpc_ctx* c, double ax_i, double ax_f, double ay_i, double ay_f) {
    if (!c) return QPC_E_ARG;
    TRY(ep_ready(c, false));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->fstream));
    const double ax[2] = {ax_i, ax_f}, ay[2] = {ay_i, ay_f};
    TRY(ep_set_scalars(c, ax, ay, 1));
    TRY(ep_factor(c, 0, 0));
    QPC_HIP_CHECK(c, hipStreamWaitEvent(c->stream, c->ep.set[0].ready, 0));
    TRY(ep_do_step(c, 0, 0));
    QPC_HIP_CHECK(c, hipEventRecord(c->ep.set[0].freed, c->stream));
    c->ep.set[0].pending = true;
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return QPC_OK;
}

int qpc_run_ep(qpc_ctx* c, const double* axtab, const double* aytab, int64_t nt_field, int64_t it0, int64_t it1,
               int use_mask) {
    if (!c || (nt_field > 0 && (!axtab || !aytab)) || it0 < 0 || it1 < it0) return QPC_E_ARG;
    TRY(ep_ready(c, use_mask != 0));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->fstream));
    TRY(qpc_host::record_room(c, it1 - it0));
    const int64_t itf = std::min(it1, nt_field);  // with-field steps [it0, itf)
    auto& E = c->ep;
    if (itf > it0) {
        const int64_t n = itf - it0;
        TRY(ep_set_scalars(c, axtab + it0, aytab + it0, n));
        TRY(ep_factor(c, 0, 0));
        for (int64_t k = 0; k < n; ++k) {
            const int set = (int)(k & 1);
            if (k + 1 < n) TRY(ep_factor(c, 1 - set, k + 1));  // one step ahead on the second stream
            TRY(qpc_host::record(c));
            QPC_HIP_CHECK(c, hipStreamWaitEvent(c->stream, E.set[set].ready, 0));
            TRY(ep_do_step(c, set, k));
            QPC_HIP_CHECK(c, hipEventRecord(E.set[set].freed, c->stream));
            E.set[set].pending = true;
            if (use_mask) TRY(qpc_host::do_mask(c));
        }
    }
    for (int64_t it = std::max(it0, itf); it < it1; ++it) {
        TRY(qpc_host::record(c));
        TRY(qpc_host::solve_free(c, c->coef, c->tmpc));
        std::swap(c->coef, c->tmpc);
        if (use_mask) TRY(qpc_host::do_mask(c));
    }
    return QPC_OK;
}

}  // extern "C"
