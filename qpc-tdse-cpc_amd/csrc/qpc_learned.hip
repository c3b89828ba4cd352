// qpc_learned.hip -- Krylov time propagation under 4th-order commutator-free
// Magnus steps, with the step size and starting subspace chosen by the
// learned controller (include/qpc_mlp.h) and judged by the Lanczos
// a-posteriori estimate (reference propagator/la_1e.hpp:296-310).  Host
// orchestration over qpc_krylov_step_adaptive; the state stays on the device.
#include "qpc_ctx.h"
#include "../../include/qpc_mlp.h"

#include <algorithm>
#include <cmath>

namespace {

constexpr double kPiL = 3.14159265358979323846;

// field_component_sinen<2> (field.h:99-123) and its time derivative
double pulse_a(const qpc_sinen2& p, double t, double* dadt) {
    const double pd = 2.0 * kPiL / p.w0, tau = p.nc * pd, ti = p.n0 * pd;
    if (!(t > ti && t < ti + tau)) {
        if (dadt) *dadt = 0.0;
        return 0.0;
    }
    const double ph = p.w0 * t + kPiL * p.c0, u = kPiL * (t - ti) / tau;
    const double s = std::sin(u), co = std::cos(u);
    if (dadt) *dadt = p.f0 * (p.w0 * std::cos(ph) * s * s + std::sin(ph) * 2.0 * s * co * kPiL / tau);
    return p.f0 * std::sin(ph) * s * s;
}

double log_feature(double e) { return std::max(-40.0, std::log10(std::max(e, 1e-300))) / 10.0 + 1.0; }

}  // namespace

extern "C" {

int qpc_krylov_propagate(qpc_ctx* c, const qpc_sinen2* pulse, double t0, double t1, double dt_base, double tol,
                         int64_t kmin, int64_t kmax, qpc_mlp* net, qpc_krylov_stats* stats, qpc_krylov_record* rec) {
    if (!c || !pulse || !(t1 >= t0) || !(dt_base > 0.0) || !(tol > 0.0) || kmin < 1 || kmax < kmin || kmax > 256)
        return QPC_E_ARG;
    if (net && !c->have_outer) {
        c->set_error("the controller's outer-norm feature needs qpc_upload_outer");
        return QPC_E_STATE;
    }
    qpc_krylov_stats st{};
    if (rec) rec->n = 0;
    const int64_t n = c->ML * c->N;
    const double r3 = std::sqrt(3.0);
    const double g1 = 0.5 - r3 / 6.0, g2 = 0.5 + r3 / 6.0;
    const double a1 = (3.0 - 2.0 * r3) / 12.0, a2 = (3.0 + 2.0 * r3) / 12.0;
    double hist[4] = {tol, tol, tol, tol};  // estimates of the last 4 steps, newest first
    std::vector<double> ab((size_t)2 * kmax);
    double t = t0;
    while (t1 - t > 1e-12 * std::max(1.0, std::fabs(t1))) {
        int64_t kstart = kmin;
        double h = dt_base;
        float feat[QPC_MLP_IN];
        if (net || rec) {
            double dadt = 0.0;
            const double a = pulse_a(*pulse, t, &dadt);
            double outer = 0.0;
            if (c->have_outer) TRY(qpc_observe(c, QPC_OBS_OUTER, &outer));
            feat[0] = (float)a;
            feat[1] = (float)(-10.0 * dadt);
            for (int k = 0; k < 4; ++k) feat[2 + k] = (float)log_feature(hist[k]);
            feat[6] = (float)(std::log10(std::max(outer, 1e-30)) / 10.0);
        }
        if (net) {
            float y[QPC_MLP_OUT];
            if (qpc_mlp_forward(net, feat, 1, y) != QPC_OK) {
                c->set_error(std::string("controller: ") + qpc_mlp_last_error(net));
                return QPC_E_HIP;
            }
            const double k = std::round(8.0 + 32.0 * (double)y[0]);
            kstart = (int64_t)std::min<double>(std::max<double>(k, (double)kmin), (double)kmax);
            h = dt_base * std::exp2(std::min(1.0, std::max(-1.0, (double)y[1])));
        }
        const bool recording = rec && rec->n < rec->cap;
        int64_t m_first = 0;  // subspace size of the step's first exponential
        // the state before the step (undo on rejection)
        QPC_HIP_CHECK(c, hipMemcpyAsync(c->spare, c->coef, sizeof(cplx) * n, hipMemcpyDeviceToDevice, c->stream));
        for (int attempt = 0;; ++attempt) {
            h = std::min(h, t1 - t);
            double d1 = 0.0, d2 = 0.0;
            const double A1 = pulse_a(*pulse, t + g1 * h, &d1), A2 = pulse_a(*pulse, t + g2 * h, &d2);
            const double aeff[2] = {2.0 * (a2 * A1 + a1 * A2), 2.0 * (a1 * A1 + a2 * A2)};
            bool ok = true;
            double worst = 0.0;
            for (int e = 0; e < 2 && ok; ++e) {
                int64_t m = 0;
                double eps = 0.0;
                // training data (rec->ab given): the first exponential of the
                // step's first attempt runs to kmax vectors
                const bool full = recording && rec->ab && e == 0 && attempt == 0;
                int rc = full ? qpc_krylov_step_adaptive(c, aeff[e], 0.5 * h, kmax, kmax, 0.0, &m, &eps, ab.data())
                              : qpc_krylov_step_adaptive(c, aeff[e], 0.5 * h, kstart, kmax, tol, &m, &eps, nullptr);
                if (rc) return rc;
                ++st.exponentials;
                st.vectors += m;
                if (e == 0) m_first = m;
                if (full) {
                    const int64_t s = rec->n;
                    std::copy(feat, feat + QPC_MLP_IN, rec->features + s * QPC_MLP_IN);
                    rec->h[s] = h;
                    rec->m[s] = m;
                    std::copy(ab.begin(), ab.end(), rec->ab + s * 2 * kmax);
                    ++rec->n;
                }
                if (eps > tol) ok = false;
                worst = std::max(worst, eps);
            }
            if (ok) {
                if (recording && !rec->ab) {
                    // the accepted step as taken (the exponentials as without a record)
                    const int64_t s = rec->n;
                    std::copy(feat, feat + QPC_MLP_IN, rec->features + s * QPC_MLP_IN);
                    rec->h[s] = h;
                    rec->m[s] = m_first;
                    ++rec->n;
                }
                t += h;
                ++st.steps;
                st.worst_estimate = std::max(st.worst_estimate, worst);
                for (int k = 3; k > 0; --k) hist[k] = hist[k - 1];
                hist[0] = worst;
                break;
            }
            // estimate above the tolerance at kmax: undo, halve the step.  A
            // step below dt_base/256 means the estimate has a floor above tol
            // (unresolved stiff components of the state): an error, not a crawl.
            ++st.rejected;
            QPC_HIP_CHECK(c, hipMemcpyAsync(c->coef, c->spare, sizeof(cplx) * n, hipMemcpyDeviceToDevice, c->stream));
            h *= 0.5;
            if (h < dt_base / 256.0 && h < t1 - t) {
                QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
                if (stats) *stats = st;
                c->set_error("Krylov step: the estimate stays above the tolerance down to dt_base/256");
                return QPC_E_NUMERIC;
            }
        }
    }
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    if (stats) *stats = st;
    return QPC_OK;
}

}  // extern "C"
