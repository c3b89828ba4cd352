// qpc_ep.hip -- the elliptically polarised (x-y) steps of propagator_cn_hc_vgwr
// (reference propagator/cn_1e.hpp:812-951) and, in the length gauge,
// propagator_cn_hc_lgwr (:717-810) on device: the (m,l)-space rotations are
// dense FP64 GEMMs on the matrix cores (k_gemm), the eight (four) coupling
// sweeps per step are the restarted-factorisation solves of qpc_solve.hip on
// the rotated channels, the field-free part and the absorber are those of the
// z-polarised path.
#include "qpc_ctx.h"
#include "qpc_setup.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cmath>
#include <map>

namespace {

// ---- (m,l)-space rotations on the FP64 matrix cores ---------------------------
// C (M x Nc doubles, row-major, ldc) = A' B', K terms, one 64 x 64 tile of C
// per work-group of four waves (each 2 x 2 tiles of v_mfma_f64_16x16x4f64,
// operand layout as k_rot's), K staged through LDS 8 at a time,
// double-buffered: the next slice's global loads are in flight while the
// current one is multiplied, and each k-step's LDS operands are read while
// the previous k-step's products run.  Small tiles and slices keep many
// work-groups per CU (the staging, not the matrix cores, sets the pace: the
// same loop without global loads runs 1.4x faster), and slices that lie
// wholly inside the operands take a load path with per-thread base pointers
// and contiguous elements.  A' (M x K):
//   AM 0: A[(ai + i) * lda + ak + k]            (real, row-major)
//   AM 1: A[(ak + k) * lda + ai + i]            (real, transposed)
//   AM 2: k < K0: Re P[(ai + i) * lda + ak + k], k >= K0: Im P[.. + k - K0]   (complex P)
//   AM 3: k < K0: Re P[(ak + k) * lda + ai + i], k >= K0: -Im P[(ak + k - K0) * lda + ai + i]   (P^H)
// B' (K x Nc doubles): row k of B for k < K0; for k >= K0 the complex row
// k - K0 times i ({re, im} -> {-im, re}), so that with AM 2 / 3 and K = 2 K0
// the product is the complex C = P B (or P^H B) with B, C complex rows
// (re P B + im P (i B)).  Each sum runs over k in order.
typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int kGM = 64, kGN = 64, kGK = 8;  // tile rows, tile columns, K slice

template <int AM>
__global__ __launch_bounds__(256) void k_gemm(const double* __restrict__ A, int64_t lda, int64_t ai, int64_t ak,
                                               const double* __restrict__ B, int64_t ldb, double* __restrict__ C,
                                               int64_t ldc, int M, int Nc, int K, int K0) {
    constexpr int PA = kGM + 4, PB = kGN + 4;
    constexpr int EA = kGM * kGK / 256, EB = kGK * kGN / 256;  // staged elements per thread
    static_assert(EB % 2 == 0 && kGK % EA == 0 && kGM % EA == 0 && kGK * (kGN / EB) == 256, "staging map");
    __shared__ double sA[2][kGK][PA];  // [k][i]
    __shared__ double sB[2][kGK][PB];  // [k][j]
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
    const int bm = blockIdx.y * kGM, bn = blockIdx.x * kGN;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    constexpr bool TRANS = AM == 1 || AM == 3;
    constexpr bool CPLX = AM >= 2;
    // thread -> staged elements: A row-wise (row r, EA consecutive k) or
    // column-wise (k, EA consecutive i); B (k, EB consecutive j)
    constexpr int KPR = kGK / EA;  // threads per A row (row-wise)
    const int ar = TRANS ? (tid % (kGM / EA)) * EA : tid / KPR;
    const int akk = TRANS ? tid / (kGM / EA) : (tid % KPR) * EA;
    constexpr int TPR = kGN / EB;
    const int bk = tid / TPR, bj = (tid % TPR) * EB;
    double ra[EA], rb[EB];
    // load() only issues the global loads (addresses clamped into the
    // operands, the element's validity and sign kept aside); store() applies
    // them when it writes LDS after the slice's products, so no load is
    // waited for before the multiplications that hide it
    unsigned va = 0, na = 0, vb = 0;  // per A element: in range, negated (-Im P^H); per B pair: in range
    bool rotb = false;
    const bool tile_in = bm + kGM <= M && bn + kGN <= Nc;
    const double* pa = CPLX ? (TRANS ? A + 2 * ((ak + akk) * lda + ai + bm + ar) : A + 2 * ((ai + bm + ar) * lda + ak + akk))
                            : (TRANS ? A + ((ak + akk) * lda + ai + bm + ar) : A + ((ai + bm + ar) * lda + ak + akk));
    const double* pb = B + bn + bj;
    auto load = [&](int k0) {
        if (tile_in && k0 + kGK <= K && (!CPLX || k0 + kGK <= K0 || k0 >= K0)) {
            // the whole slice inside the operands (and inside one half of a complex K)
            const bool imh = CPLX && k0 >= K0;
            const int kq = imh ? k0 - K0 : k0;
            va = (1u << EA) - 1;
            na = (CPLX && TRANS && imh) ? va : 0;
#pragma unroll
            for (int e = 0; e < EA; ++e) {
                if constexpr (!CPLX) ra[e] = TRANS ? pa[(int64_t)kq * lda + e] : pa[kq + e];
                else ra[e] = TRANS ? pa[2 * ((int64_t)kq * lda + e) + (imh ? 1 : 0)] : pa[2 * (kq + e) + (imh ? 1 : 0)];
            }
            rotb = imh;
            const double* br = pb + (int64_t)(kq + bk) * ldb;
            vb = (1u << (EB / 2)) - 1;
#pragma unroll
            for (int e = 0; e < EB; ++e) rb[e] = br[e];
            return;
        }
        va = na = 0;
#pragma unroll
        for (int e = 0; e < EA; ++e) {
            const int i = bm + ar + (TRANS ? e : 0), k = k0 + akk + (TRANS ? 0 : e);
            const bool ok = i < M && k < K;
            const int ic = ok ? i : 0, kc = ok ? k : 0;
            const double* q;
            if constexpr (!CPLX) {
                q = TRANS ? A + ((ak + kc) * lda + ai + ic) : A + ((ai + ic) * lda + ak + kc);
            } else {
                const bool im = kc >= K0;
                const int kk = im ? kc - K0 : kc;
                q = (TRANS ? A + 2 * ((ak + kk) * lda + ai + ic) : A + 2 * ((ai + ic) * lda + ak + kk)) + (im ? 1 : 0);
                if (im && TRANS) na |= 1u << e;
            }
            if (ok) va |= 1u << e;
            ra[e] = *q;
        }
        const int k = k0 + bk;
        rotb = k >= K0;
        const int row = k < K ? (rotb ? k - K0 : k) : 0;
        const double* br = B + (int64_t)row * ldb;
        vb = 0;
#pragma unroll
        for (int e = 0; e < EB; e += 2) {
            const int j = bn + bj + e;
            const bool v = k < K && j < Nc;  // Nc even: pairs stay whole
            if (v) vb |= 1u << (e / 2);
            const double* q = br + (v ? j : 0);
            rb[e] = q[0];
            rb[e + 1] = q[1];
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int e = 0; e < EA; ++e) {
            double v = (va >> e) & 1 ? ra[e] : 0.0;
            if ((na >> e) & 1) v = -v;
            if (TRANS) sA[buf][akk][ar + e] = v;
            else sA[buf][akk + e][ar] = v;
        }
#pragma unroll
        for (int e = 0; e < EB; e += 2) {
            const bool v = (vb >> (e / 2)) & 1;
            const double re = v ? rb[e] : 0.0, im = v ? rb[e + 1] : 0.0;
            sB[buf][bk][bj + e] = rotb ? -im : re;
            sB[buf][bk][bj + e + 1] = rotb ? re : im;
        }
    };
    v4d acc[2][2];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < 2; ++u) acc[t][u] = v4d{0.0, 0.0, 0.0, 0.0};
    load(0);
    store(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = 0; k0 < K; k0 += kGK) {
        const bool more = k0 + kGK < K;
        if (more) load(k0 + kGK);
        double a0[2], b0[2], a1[2], b1[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) a0[t] = sA[buf][(l >> 4)][wr + t * 16 + (l & 15)];
#pragma unroll
        for (int u = 0; u < 2; ++u) b0[u] = sB[buf][(l >> 4)][wc + u * 16 + (l & 15)];
#pragma unroll
        for (int ks = 0; ks < kGK / 4; ks += 2) {
#pragma unroll
            for (int t = 0; t < 2; ++t) a1[t] = sA[buf][(ks + 1) * 4 + (l >> 4)][wr + t * 16 + (l & 15)];
#pragma unroll
            for (int u = 0; u < 2; ++u) b1[u] = sB[buf][(ks + 1) * 4 + (l >> 4)][wc + u * 16 + (l & 15)];
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int u = 0; u < 2; ++u) acc[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[t], b0[u], acc[t][u], 0, 0, 0);
            if (ks + 2 < kGK / 4) {
#pragma unroll
                for (int t = 0; t < 2; ++t) a0[t] = sA[buf][(ks + 2) * 4 + (l >> 4)][wr + t * 16 + (l & 15)];
#pragma unroll
                for (int u = 0; u < 2; ++u) b0[u] = sB[buf][(ks + 2) * 4 + (l >> 4)][wc + u * 16 + (l & 15)];
            }
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int u = 0; u < 2; ++u) acc[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[t], b1[u], acc[t][u], 0, 0, 0);
        }
        if (more) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int j = bn + wc + u * 16 + (l & 15);
            if (j >= Nc) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = bm + wr + t * 16 + 4 * r + (l >> 4);
                if (i < M) C[(int64_t)i * ldc + j] = acc[t][u][r];
            }
        }
}

template <int AM>
int gemm(qpc_ctx* c, const double* A, int64_t lda, int64_t ai, int64_t ak, const double* B, int64_t ldb, double* C,
         int64_t ldc, int64_t M, int64_t Nc, int64_t K, int64_t K0) {
    if (M <= 0 || Nc <= 0) return QPC_OK;
    if (Nc % 2 || M > (1 << 30) || Nc > (1 << 30) || K > (1 << 30)) {
        c->set_error("rotation GEMM shape");
        return QPC_E_SHAPE;
    }
    hipLaunchKernelGGL((k_gemm<AM>), dim3((unsigned)((Nc + kGN - 1) / kGN), (unsigned)((M + kGM - 1) / kGM)), dim3(256),
                       0, c->stream, A, lda, ai, ak, B, ldb, C, ldc, (int)M, (int)Nc, (int)K, (int)K0);
    return launch_check(c, "rotation GEMM");
}

// cn_transform_hm::construct<-1,isub> / eliminate<isub,-1> of a real projector
// (cn_1e.hpp:222-273): rows [s0,s1) of dst from all rows of src, or all rows
// of dst from rows [s0,s1) of src
int rot_real_construct(qpc_ctx* c, const double* P, const cplx* src, cplx* dst) {
    const auto& E = c->ep;
    const int64_t n2 = 2 * c->N;
    TimedLaunch t(c, QPC_K_EP_ROT);
    // dst rows [s0, s1) = P[s0:s1, :] src
    return gemm<0>(c, P, c->ML, E.s0, 0, (const double*)src, n2, (double*)(dst + E.s0 * c->N), n2, E.s1 - E.s0, n2,
                   c->ML, c->ML);
}
int rot_real_eliminate(qpc_ctx* c, const double* P, const cplx* src, cplx* dst) {
    const auto& E = c->ep;
    const int64_t n2 = 2 * c->N, ns = E.s1 - E.s0;
    TimedLaunch t(c, QPC_K_EP_ROT);
    // dst = P[s0:s1, :]^T src[s0:s1]
    return gemm<1>(c, P, c->ML, 0, E.s0, (const double*)(src + E.s0 * c->N), n2, (double*)dst, n2, c->ML, n2, ns, ns);
}
// ---- the rotations around the field-free solve, on the parity-ordered state --
// Rows of tran_y2 / tran_y1 in the even category (catagorize's first n_even)
// touch only the even-parity channels and those of the odd category only the
// odd ones; the zero category (odd M) mixes both.  With the natural state in
// parity order (E.perm) and the projector's columns in the same order
// (ty2p / ty1p), each row group multiplies only its own column block: the
// terms left out are exact zeros, so the sums are those of the full product.
struct RowGroups {
    int64_t ev0, ev1, z1;  // even category [ev0, ev1), zero [ev1, z1), odd [z1, ML)
};
inline RowGroups row_groups(const qpc_ctx* c) {
    const int64_t M = c->M, L = c->L;
    const int64_t n_zero = M % 2 == 0 ? 0 : L, n_even = (M / 2 * 2) * ((L + 1) / 2);
    return {0, n_even, n_even + n_zero};
}
// dst rows [s0, s1) = P[s0:s1, :] src (src, P's columns in parity order);
// CPLX: P complex (construct of tran_y1)
template <bool CPLX>
int rot_construct_p(qpc_ctx* c, const void* P, const cplx* src, cplx* dst) {
    const auto& E = c->ep;
    const RowGroups g = row_groups(c);
    const int64_t ML = c->ML, N = c->N, n2 = 2 * N;
    // (row range, column block [k0, k1)) of the three groups
    const int64_t r[3][4] = {{g.ev0, g.ev1, 0, E.n_ev}, {g.ev1, g.z1, 0, ML}, {g.z1, ML, E.n_ev, ML}};
    TimedLaunch t(c, QPC_K_EP_ROT);
    for (const auto& q : r) {
        const int64_t a = std::max(q[0], E.s0), b = std::min(q[1], E.s1), k0 = q[2], kn = q[3] - q[2];
        if (b <= a) continue;
        if (CPLX)
            TRY(gemm<2>(c, (const double*)P, ML, a, k0, (const double*)(src + k0 * N), n2, (double*)(dst + a * N), n2,
                        b - a, n2, 2 * kn, kn));
        else
            TRY(gemm<0>(c, (const double*)P, ML, a, k0, (const double*)(src + k0 * N), n2, (double*)(dst + a * N), n2,
                        b - a, n2, kn, kn));
    }
    return QPC_OK;
}
// dst (parity order, all ML rows) = P[s0:s1, :]^H src[s0:s1] (^T when real):
// the even block from the even and zero rows, the odd block from the zero and
// odd rows (contiguous row ranges of catagorize's order)
template <bool CPLX>
int rot_eliminate_p(qpc_ctx* c, const void* P, const cplx* src, cplx* dst) {
    const auto& E = c->ep;
    const RowGroups g = row_groups(c);
    const int64_t ML = c->ML, N = c->N, n2 = 2 * N;
    const int64_t blk[2][4] = {{0, E.n_ev, g.ev0, g.z1}, {E.n_ev, ML, g.ev1, ML}};  // out rows, source rows
    TimedLaunch t(c, QPC_K_EP_ROT);
    for (const auto& q : blk) {
        const int64_t o0 = q[0], on = q[1] - q[0];
        const int64_t a = std::max(q[2], E.s0), b = std::min(q[3], E.s1), kn = std::max<int64_t>(0, b - a);
        if (on <= 0) continue;
        // an empty source range leaves the block zero (the GEMM with K = 0 writes zeros)
        const int64_t k0 = kn > 0 ? a : 0;
        if (CPLX)
            TRY(gemm<3>(c, (const double*)P, ML, o0, k0, (const double*)(src + k0 * N), n2, (double*)(dst + o0 * N),
                        n2, on, n2, 2 * kn, kn));
        else
            TRY(gemm<1>(c, (const double*)P, ML, o0, k0, (const double*)(src + k0 * N), n2, (double*)(dst + o0 * N),
                        n2, on, n2, kn, kn));
    }
    return QPC_OK;
}

// construct<isub,isub> / eliminate<isub,isub> of a merged complex projector:
// the [s0,s1) x [s0,s1) block, P + s0 + s0*ML
int rot_cplx(qpc_ctx* c, bool herm, const cplx* P, const cplx* src, cplx* dst) {
    const auto& E = c->ep;
    const int64_t ns = E.s1 - E.s0, N = c->N;
    TimedLaunch t(c, QPC_K_EP_ROT);
    const double* B = (const double*)(src + E.s0 * N);
    double* C = (double*)(dst + E.s0 * N);
    if (herm) return gemm<3>(c, (const double*)P, c->ML, E.s0, E.s0, B, 2 * N, C, 2 * N, ns, 2 * N, 2 * ns, ns);
    return gemm<2>(c, (const double*)P, c->ML, E.s0, E.s0, B, 2 * N, C, 2 * N, ns, 2 * N, 2 * ns, ns);
}

// operators of the sweeps: x1 and y1 are D-parts (Y11 with d/dr), x2 and y2
// R-parts (Q11 with 1/r); table index within a set: D tables [x1 ti, x1 tf,
// y1 ti, y1 tf], R tables [x2 ti, x2 tf, y2 ti, y2 tf].  In the length gauge
// x1 and y1 are sweeps of S + i e r whose tables take the R slots [x1 ti,
// x1 tf, y1 ti, y1 tf]; ep_is_d then only names their scalar positions.
enum EpOp { EP_X1 = 0, EP_X2 = 1, EP_Y1 = 2, EP_Y2 = 3 };
inline bool ep_is_d(int op) { return op == EP_X1 || op == EP_Y1; }
inline int ep_slot(int op, int tf) { return (op == EP_X1 || op == EP_X2 ? 0 : 2) + tf; }
// position of a sweep scalar within the 8 of a step
inline int ep_sc(int op, int tf) { return (ep_is_d(op) ? 0 : 4) + ep_slot(op, tf); }

// scalars of cn_hm_propagate1/2 (propagate.hpp:463, 525) for the run's steps:
// propagate<isub> halves A (cn_1e.hpp:905-908), f = -ax, +ax, -ay, -ay for
// x1, x2, y1, y2 (:911-942), frdt = f * dt/2 (prop step, :873-876);
// ifdt = -frdt/2 (d/dr sweeps) or frdt/2 (times i, 1/r sweeps).
// Length gauge (cn_1e.hpp:775-778, 780-800): f = ex/2, ey/2 for x1, y1, both
// sweeps cn_hm_propagate1 with X = r, ifdt = frdt/2 (times i).
int ep_set_scalars(qpc_ctx* c, const double* ax, const double* ay, int64_t nsteps) {
    auto& E = c->ep;
    const double step = c->d.para_dt * 0.5;
    E.h_sc.assign((size_t)nsteps * 8, 0.0);
    double emax = 0.0;
    for (int64_t k = 0; k < nsteps; ++k) {
        double* sc = &E.h_sc[(size_t)k * 8];
        for (int tf = 0; tf < 2; ++tf) {
            const double hx = ax[k + tf] / 2.0, hy = ay[k + tf] / 2.0;
            const double f[4] = {E.lg ? hx : -hx, hx, E.lg ? hy : -hy, -hy};
            for (int op = 0; op < 4; ++op) {
                if (E.lg && !ep_is_d(op)) continue;
                const double frdt = f[op] * step;
                const double v = ep_is_d(op) && !E.lg ? -frdt / 2 : frdt / 2;
                sc[ep_sc(op, tf)] = v;
                if (!std::isfinite(v)) {
                    c->set_error(E.lg ? "non-finite E(t)" : "non-finite A(t)");
                    return QPC_E_NUMERIC;
                }
                if (ep_is_d(op)) emax = std::max(emax, std::fabs(v) * E.emax[op]);
            }
        }
    }
    // carry depth and restart of the D-part (length gauge: S + i e r) sweeps
    // for this run's scalars
    if (E.lg) TRY(qpc_host::plan_l_part(c, emax));
    else TRY(qpc_host::plan_d_part(c, emax));
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
    // a complex symmetric sweep's tables (ts: tables of time point t at
    // position t * np of its block; shared tables: the nprim primaries)
    const int64_t trs = c->geo.trf_elems;
    auto sym_tables = [&](int op, const double* X, int restart) {
        const int64_t slot = ep_slot(op, 0);
        cplx* tab = T.tabR + slot * np * cte;
        cplx* trf = T.trfR + slot * np * trs;
        if (E.rmap[op]) {
            qpc::launch_factor_sym(c->fstream, c->S, X, E.eprim[op], sc + ep_sc(op, 0), 2, (int)E.nprim[op], c->geo,
                                   tab, 0, restart, np);
            qpc::launch_transfer_c(c->fstream, tab, trf, (int)E.nprim[op], 2, np, c->geo);
        } else {
            qpc::launch_factor_sym(c->fstream, c->S, X, E.eig + op * ML + E.p0, sc + ep_sc(op, 0), 2, (int)np, c->geo,
                                   tab, 0, restart);
            qpc::launch_transfer_c(c->fstream, tab, trf, (int)(2 * np), c->geo);
        }
    };
    if (E.lg) {
        TimedLaunch tl(c, QPC_K_FACTOR, c->fstream);
        for (int op : {EP_X1, EP_Y1}) sym_tables(op, c->rpow1, c->restart_l);
    } else {
        TimedLaunch tl(c, QPC_K_FACTOR, c->fstream);
        for (int op = 0; op < 4; ++op) {
            const double* eig = E.eig + op * ML + E.p0;
            const int64_t slot = ep_slot(op, 0);
            if (ep_is_d(op))
                qpc::launch_factor_asy(c->fstream, c->S, c->rdif, eig, sc + ep_sc(op, 0), 2, (int)np, c->geo,
                                       T.tabD + slot * np * te, 0, c->restart_d);
            else
                sym_tables(op, c->rinv, 0);
        }
        qpc::launch_transfer_r(c->fstream, T.tabD, T.trfD, (int)(4 * np), c->geo);
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
    ra.sx = E.lg ? c->sxL : ep_is_d(op) ? c->sxD : c->sxR;
    ra.eig = E.eig + op * c->ML + E.p0;
    ra.ifdt = E.h_sc[(size_t)k * 8 + ep_sc(op, tf)];
    qpc::TabMap tm = E.rmap[op] ? qpc::TabMap{4, (int)c->L, (int)c->m0, E.rmap[op]}
                                : qpc::TabMap{0, (int)c->L, (int)c->m0};
    // the work-groups of a row's segments and of a shared table's rows on one XCD
    const int pk = E.lg ? c->pk_l : ep_is_d(op) ? c->pk_d : qpc::kPKS;
    const int nseg = qpc::solve_segments(c->geo, pk);
    if (nseg > 0 && nseg != E.ord_nseg[op] && (int64_t)E.rtab_h[op].size() == np) {
        E.ord_h[op].assign((size_t)np * nseg, 0);
        qpc::solve_order(E.rtab_h[op].data(), (int)np, nseg, E.ord_h[op].data());
        if (E.ord[op]) QPC_HIP_CHECK(c, hipFree(E.ord[op]));
        E.ord[op] = nullptr;
        TRY(dev_alloc(c, &E.ord[op], E.ord_h[op].size()));
        TRY(upload(c, E.ord[op], E.ord_h[op].data(), E.ord_h[op].size()));
        E.ord_nseg[op] = nseg;
    }
    if (nseg > 0 && E.ord_nseg[op] == nseg) {
        tm.order = E.ord[op];
        tm.order_nseg = nseg;
    }
    cplx* z = buf + E.p0 * c->N;
    const int kind = ep_is_d(op) && !E.lg ? qpc::RHS_SX_ASY : qpc::RHS_SX_SYM;
    // in place (z in, z out): the right-hand side goes through vbuf
    TimedLaunch tl(c, QPC_K_EP_SWEEP);
    if (E.lg)
        qpc::launch_rsolve_c(c->stream, kind, ra, T.tabR + slot * np * cte, T.trfR + slot * np * tr, tm, z, c->vbuf, z,
                             (int)np, c->geo, c->pk_l);
    else if (ep_is_d(op))
        qpc::launch_rsolve_r(c->stream, kind, ra, T.tabD + slot * np * te, T.trfD + slot * np * tr, tm, z, c->vbuf, z,
                             (int)np, c->geo, c->pk_d);
    else
        qpc::launch_rsolve_c(c->stream, kind, ra, T.tabR + slot * np * cte, T.trfR + slot * np * tr, tm, z, c->vbuf, z,
                             (int)np, c->geo);
    return launch_check(c, "ep sweep");
}

// propagator_cn_hc_vgwr::propagate<isub> (cn_1e.hpp:898-945) for run step k
// whose tables are in `set`; the state ends in c->coef.  Inside a run the
// rotation out of x1's eigenbasis that ends a step and the one back that
// begins the next cancel around the absorber, which acts alike on every
// channel (P S^-1 M P^T = S^-1 M on the rows [s0, s1) that the construct
// keeps): out_rot leaves the state in x1's basis (in c->coef), in_rot takes
// it from there.
int ep_do_step(qpc_ctx* c, int set, int64_t k, bool in_rot, bool out_rot) {
    auto& E = c->ep;
    if (in_rot) std::swap(c->coef, c->tmpc);  // the rotated state is Y
    cplx* X = c->coef;
    cplx* Y = c->tmpc;
    cplx* Z = c->spare;
    if (!in_rot) TRY(rot_real_construct(c, E.tx1, X, Y));   // tran_x1.construct<-1,isub>
    TRY(ep_sweep(c, set, k, EP_X1, 0, Y));     // prop_x1.propagate2(-ax_ti)
    TRY(rot_cplx(c, false, E.x1x2, Y, Z));     // x1_2_x2.construct
    TRY(ep_sweep(c, set, k, EP_X2, 0, Z));     // prop_x2.propagate1(ax_ti)
    TRY(rot_cplx(c, false, E.x2y1, Z, Y));     // x2_2_y1.construct
    TRY(ep_sweep(c, set, k, EP_Y1, 0, Y));     // prop_y1.propagate2(-ay_ti)
    TRY(rot_cplx(c, false, E.y1y2, Y, Z));     // y1_2_y2.construct
    TRY(ep_sweep(c, set, k, EP_Y2, 0, Z));     // prop_y2.propagate1(-ay_ti)
    TRY(rot_eliminate_p<false>(c, E.ty2p, Z, X));          // tran_y2.eliminate<isub,-1> (parity order)
    TRY(qpc_host::solve_free_mapped(c, X, Y, E.lmap));     // prop_hc.propagate0
    TRY(rot_construct_p<false>(c, E.ty2p, Y, Z));          // tran_y2.construct<-1,isub>
    TRY(ep_sweep(c, set, k, EP_Y2, 1, Z));     // prop_y2.propagate1(-ay_tf)
    TRY(rot_cplx(c, true, E.y1y2, Z, Y));      // y1_2_y2.eliminate
    TRY(ep_sweep(c, set, k, EP_Y1, 1, Y));     // prop_y1.propagate2(-ay_tf)
    TRY(rot_cplx(c, true, E.x2y1, Y, Z));      // x2_2_y1.eliminate
    TRY(ep_sweep(c, set, k, EP_X2, 1, Z));     // prop_x2.propagate1(ax_tf)
    TRY(rot_cplx(c, true, E.x1x2, Z, Y));      // x1_2_x2.eliminate
    TRY(ep_sweep(c, set, k, EP_X1, 1, Y));     // prop_x1.propagate2(-ax_tf)
    if (out_rot) {
        std::swap(c->coef, c->tmpc);
        return QPC_OK;
    }
    TRY(rot_real_eliminate(c, E.tx1, Y, X));   // tran_x1.eliminate<isub,-1>
    return QPC_OK;
}

// propagator_cn_hc_lgwr::propagate<isub> (cn_1e.hpp:767-803) for run step k;
// in_rot / out_rot as ep_do_step's
int lg_do_step(qpc_ctx* c, int set, int64_t k, bool in_rot, bool out_rot) {
    auto& E = c->ep;
    if (in_rot) std::swap(c->coef, c->tmpc);
    cplx* X = c->coef;
    cplx* Y = c->tmpc;
    cplx* Z = c->spare;
    if (!in_rot) TRY(rot_real_construct(c, E.tx1, X, Y));    // tran_x1.construct<-1,isub>
    TRY(ep_sweep(c, set, k, EP_X1, 0, Y));      // prop_x1.propagate1(ex_ti)
    TRY(rot_cplx(c, false, E.x1x2, Y, Z));      // x1_2_y1.construct
    TRY(ep_sweep(c, set, k, EP_Y1, 0, Z));      // prop_y1.propagate1(ey_ti)
    TRY(rot_eliminate_p<true>(c, E.ty1p, Z, X));           // tran_y1.eliminate<isub,-1> (parity order)
    TRY(qpc_host::solve_free_mapped(c, X, Y, E.lmap));     // prop_hc.propagate0
    TRY(rot_construct_p<true>(c, E.ty1p, Y, Z));           // tran_y1.construct<-1,isub>
    TRY(ep_sweep(c, set, k, EP_Y1, 1, Z));      // prop_y1.propagate1(ey_tf)
    TRY(rot_cplx(c, true, E.x1x2, Z, Y));       // x1_2_y1.eliminate
    TRY(ep_sweep(c, set, k, EP_X1, 1, Y));      // prop_x1.propagate1(ex_tf)
    if (out_rot) {
        std::swap(c->coef, c->tmpc);
        return QPC_OK;
    }
    TRY(rot_real_eliminate(c, E.tx1, Y, X));    // tran_x1.eliminate<isub,-1>
    return QPC_OK;
}

int ep_ready(qpc_ctx* c, bool need_mask, bool lg) {
    TRY(qpc_host::check_ready(c, need_mask));
    if (!c->ep.have || c->ep.lg != lg) {
        c->set_error(lg ? "length-gauge elliptically polarised angular data not uploaded"
                        : "elliptically polarised angular data not uploaded");
        return QPC_E_STATE;
    }
    if (lg && c->rpow1_h.empty()) {
        c->set_error("<Bi|r|Bj> (rpow1) not uploaded");
        return QPC_E_STATE;
    }
    return QPC_OK;
}

// the sweeps that are complex symmetric (S + i e X): the 1/r sweeps x2, y2,
// or in the length gauge x1, y1 with r
inline bool ep_sym(const qpc_ctx* c, int op) { return c->ep.lg ? ep_is_d(op) : !ep_is_d(op); }

// shared tables of a complex symmetric sweep: rows [p0, p1) whose
// eigenvalues have one magnitude (exact pairs, see qpc::pair_spectrum) use
// one table, the negative one conjugated (S + i (-e) X = conj(S + i e X))
int ep_plan_share(qpc_ctx* c) {
    auto& E = c->ep;
    const int64_t np = E.p1 - E.p0;
    for (int op = 0; op < 4; ++op) {
        if (E.rmap[op]) QPC_HIP_CHECK(c, hipFree(E.rmap[op]));
        if (E.eprim[op]) QPC_HIP_CHECK(c, hipFree(E.eprim[op]));
        E.rmap[op] = nullptr;
        E.eprim[op] = nullptr;
        E.nprim[op] = np;
        E.ord_nseg[op] = 0;  // the sweep orders follow the new rows and tables
        E.rtab_h[op].resize((size_t)std::max<int64_t>(np, 0));
        for (int64_t r = 0; r < np; ++r) E.rtab_h[op][r] = (int)r;
        if (!ep_sym(c, op) || np <= 0 || (int64_t)E.eig_h[op].size() != c->ML) continue;
        std::map<double, int> idx;
        std::vector<double> prim;
        std::vector<int> map((size_t)np);
        for (int64_t r = 0; r < np; ++r) {
            const double e = E.eig_h[op][E.p0 + r];
            const auto it = idx.find(std::fabs(e));
            if (it == idx.end()) {
                idx.emplace(std::fabs(e), (int)prim.size());
                map[r] = (int)prim.size();
                prim.push_back(e);
            } else {
                map[r] = it->second | (e != prim[it->second] ? qpc::kMapConj : 0);
            }
        }
        if ((int64_t)prim.size() == np) continue;  // no pairs
        for (int64_t r = 0; r < np; ++r) E.rtab_h[op][r] = map[r] & qpc::kMapIdx;
        TRY(dev_alloc(c, &E.rmap[op], map.size()));
        TRY(dev_alloc(c, &E.eprim[op], prim.size()));
        TRY(upload(c, E.rmap[op], map.data(), map.size()));
        TRY(upload(c, E.eprim[op], prim.data(), prim.size()));
        QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        E.nprim[op] = (int64_t)prim.size();
    }
    return QPC_OK;
}

// out[r][k] = in[r][perm[k]] (the columns of a projector in parity order)
template <class T>
__global__ __launch_bounds__(256) void k_perm_cols(const T* __restrict__ in, const int* __restrict__ perm,
                                                   T* __restrict__ out, int ML) {
    const int r = blockIdx.y, k = blockIdx.x * 256 + threadIdx.x;
    if (k < ML) out[(int64_t)r * ML + k] = in[(int64_t)r * ML + perm[k]];
}

// the parity order of the natural channels and the column-permuted copy of
// tran_y2 (tran_y1 in the length gauge), once the projectors are in place
int ep_permute(qpc_ctx* c) {
    auto& E = c->ep;
    const int64_t ML = c->ML, L = c->L;
    std::vector<int> perm, lmap;
    perm.reserve(ML);
    for (int pty = 0; pty < 2; ++pty)
        for (int64_t ch = 0; ch < ML; ++ch)
            if ((ch % L) % 2 == pty) perm.push_back((int)ch);
    E.n_ev = 0;
    for (int64_t ch = 0; ch < ML; ++ch) E.n_ev += (ch % L) % 2 == 0;
    for (int ch : perm) {
        const int64_t m = c->m0 + ch / L;
        lmap.push_back((int)(ch % L + (m < 0 ? -m : m)));
    }
    TRY(upload(c, E.perm, perm.data(), (size_t)ML));
    TRY(upload(c, E.lmap, lmap.data(), (size_t)ML));
    const dim3 grid((unsigned)((ML + 255) / 256), (unsigned)ML);
    if (E.lg)
        hipLaunchKernelGGL(k_perm_cols<cplx>, grid, dim3(256), 0, c->stream, E.ty1, E.perm, E.ty1p, (int)ML);
    else
        hipLaunchKernelGGL(k_perm_cols<double>, grid, dim3(256), 0, c->stream, E.ty2, E.perm, E.ty2p, (int)ML);
    TRY(launch_check(c, "projector columns in parity order"));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return QPC_OK;
}

// sub-block ranges, buffers and sweep tables of either gauge
int ep_prepare(qpc_ctx* c, int isub, bool lg) {
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
    E.have = false;
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->fstream));
    const size_t mm = (size_t)ML * ML;
    const int64_t np = E.p1 - E.p0;
    const size_t te = (size_t)c->geo.tab_elems, cte = (size_t)c->geo.ctab_elems, tr = (size_t)c->geo.trf_elems;
    int rc = QPC_OK;
    if (!E.eig) {
        if ((rc = dev_alloc(c, &E.eig, 4 * (size_t)ML)) || (rc = dev_alloc(c, &E.tx1, mm)) ||
            (rc = dev_alloc(c, &E.x1x2, mm)))
            return rc;
        for (auto& T : E.set) {
            QPC_HIP_CHECK(c, hipEventCreateWithFlags(&T.ready, hipEventDisableTiming));
            QPC_HIP_CHECK(c, hipEventCreateWithFlags(&T.freed, hipEventDisableTiming));
        }
    }
    if (!E.perm && ((rc = dev_alloc(c, &E.perm, (size_t)ML)) || (rc = dev_alloc(c, &E.lmap, (size_t)ML)))) return rc;
    if (!lg && !E.ty2 &&
        ((rc = dev_alloc(c, &E.ty2, mm)) || (rc = dev_alloc(c, &E.x2y1, mm)) || (rc = dev_alloc(c, &E.y1y2, mm)) ||
         (rc = dev_alloc(c, &E.ty2p, mm))))
        return rc;
    if (lg && !E.ty1 && ((rc = dev_alloc(c, &E.ty1, mm)) || (rc = dev_alloc(c, &E.ty1p, mm)))) return rc;
    // sweep tables for the rotated channels of this sub-block (re-sized when
    // isub or the gauge changes; the length gauge has no D-part tables)
    for (auto& T : E.set) {
        void* tp[] = {T.tabD, T.trfD, T.tabR, T.trfR};
        for (void* p : tp)
            if (p) QPC_HIP_CHECK(c, hipFree(p));
        T.tabD = T.trfD = nullptr;
        T.tabR = T.trfR = nullptr;
        T.pending = false;
        if (np > 0 && !lg && ((rc = dev_alloc(c, &T.tabD, 4 * np * te)) || (rc = dev_alloc(c, &T.trfD, 4 * np * tr))))
            return rc;
        if (np > 0 && ((rc = dev_alloc(c, &T.tabR, 4 * np * cte)) || (rc = dev_alloc(c, &T.trfR, 4 * np * tr))))
            return rc;
    }
    E.lg = lg;
    return QPC_OK;
}

// ---- angular eigen-data on the device ------------------------------------------
// Projector rows from the per-parity eigenvectors (eigenvector k of parity pb
// at Z[pb] + k * n_pb, component q for the q-th channel of that parity):
// prj[r][j] = ph(j) Z[pb(r)][k(r)][loc(j)] where par(j) = pb(r), else 0, with
// ph(j) = i^(im_j) for the operators with multiplier -i.  Each output is
// optional: re = Re prj, cj = conj(prj), tr = prj^T (row-major, ML x ML).
__global__ __launch_bounds__(256) void k_ang_proj(const double* __restrict__ Z0, const double* __restrict__ Z1,
                                                  int n0, int n1, const int* __restrict__ rpb,
                                                  const int* __restrict__ rk, const int* __restrict__ par,
                                                  const int* __restrict__ loc, int L, int ML, int phase,
                                                  double* __restrict__ re, cplx* __restrict__ cj,
                                                  cplx* __restrict__ tr) {
    const int r = blockIdx.y;
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= ML) return;
    const int pb = rpb[r], k = rk[r];
    double v = 0.0;
    if (par[j] == pb) v = pb == 0 ? Z0[(int64_t)k * n0 + loc[j]] : Z1[(int64_t)k * n1 + loc[j]];
    cplx x = {v, 0.0};
    if (phase) {
        const int q = (j / L) & 3;  // i^q
        x = q == 0 ? cplx{v, 0.0} : q == 1 ? cplx{0.0, v} : q == 2 ? cplx{-v, 0.0} : cplx{0.0, -v};
    }
    const int64_t o = (int64_t)r * ML + j;
    if (re) re[o] = x.re;
    if (cj) cj[o] = cplx{x.re, -x.im};
    if (tr) tr[(int64_t)j * ML + r] = x;
}

// operator_y11/q11::solve_eigen + converter::catagorize + the transforms of
// propagator_cn_hc_vgwr / _lgwr::initialize (cn_1e.hpp:878-895, 745-765) on
// the device: each operator's two parity blocks are dense real symmetric
// problems (the host builds them, qpc::angular_blocks), solved by rocSOLVER's
// divide-and-conquer syevd batched over the operators; the eigenvalues come
// back to order the rows (qpc::angular_rows), k_ang_proj forms the projectors
// and the merged products P_b P_a^H are FP64 matrix-core GEMMs (k_gemm) of
// conj(prj_b) and prj_a^T written straight into the context.
int ep_build_angular(qpc_ctx* c, int isub, bool lg) {
    const int M = (int)c->M, L = (int)c->L, ML = (int)c->ML;
    if (c->ML > 65535) {
        c->set_error("angular set-up for at most 65535 channels");
        return QPC_E_SHAPE;
    }
    TRY(ep_prepare(c, isub, lg));
    auto& E = c->ep;
    qpc::AngBlocks ab;
    qpc::angular_blocks(M, L, (int)c->m0, ab);
    const std::vector<int> ops = lg ? std::vector<int>{0, 2} : std::vector<int>{0, 1, 2, 3};
    const int nops = (int)ops.size();
    const int64_t n[2] = {(int64_t)ab.idx[0].size(), (int64_t)ab.idx[1].size()};
    const size_t mm = (size_t)ML * ML;
    // device buffers of this set-up, freed on every path out
    std::vector<void*> bufs;
    auto dalloc = [&](void** p, size_t bytes) -> int {
        QPC_HIP_CHECK(c, hipMalloc(p, bytes));
        bufs.push_back(*p);
        return QPC_OK;
    };
    rocblas_handle h = nullptr;
    int rc = [&]() -> int {
        double *dZ[2] = {nullptr, nullptr}, *dW[2] = {nullptr, nullptr}, *dE[2] = {nullptr, nullptr};
        int* dinfo = nullptr;
        for (int pb = 0; pb < 2; ++pb) {
            if (!n[pb]) continue;
            TRY(dalloc((void**)&dZ[pb], sizeof(double) * nops * n[pb] * n[pb]));
            TRY(dalloc((void**)&dW[pb], sizeof(double) * nops * n[pb]));
            TRY(dalloc((void**)&dE[pb], sizeof(double) * nops * n[pb]));
            for (int s = 0; s < nops; ++s)
                QPC_HIP_CHECK(c, hipMemcpyAsync(dZ[pb] + s * n[pb] * n[pb], ab.B[ops[s]][pb].data(),
                                                sizeof(double) * n[pb] * n[pb], hipMemcpyHostToDevice, c->stream));
        }
        TRY(dalloc((void**)&dinfo, sizeof(int) * 2 * nops));
        if (rocblas_create_handle(&h) != rocblas_status_success || rocblas_set_stream(h, c->stream) != rocblas_status_success) {
            c->set_error("rocBLAS handle");
            return QPC_E_SETUP;
        }
        for (int pb = 0; pb < 2; ++pb) {
            if (!n[pb]) continue;
            const rocblas_int nn = (rocblas_int)n[pb];
            if (rocsolver_dsyevd_strided_batched(h, rocblas_evect_original, rocblas_fill_upper, nn, dZ[pb], nn,
                                                 (rocblas_stride)nn * nn, dW[pb], nn, dE[pb], nn, dinfo + pb * nops,
                                                 nops) != rocblas_status_success) {
                c->set_error("rocsolver_dsyevd");
                return QPC_E_SETUP;
            }
        }
        // eigenvalues back to the host to order the rows
        std::vector<double> w[2];
        std::vector<int> info(2 * nops, 0);
        for (int pb = 0; pb < 2; ++pb) {
            w[pb].resize((size_t)nops * n[pb]);
            if (n[pb])
                QPC_HIP_CHECK(c, hipMemcpyAsync(w[pb].data(), dW[pb], sizeof(double) * nops * n[pb],
                                                hipMemcpyDeviceToHost, c->stream));
        }
        QPC_HIP_CHECK(c, hipMemcpyAsync(info.data(), dinfo, sizeof(int) * 2 * nops, hipMemcpyDeviceToHost, c->stream));
        QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        for (int i = 0; i < 2 * nops; ++i)
            if (info[i] != 0 && n[i / nops] > 0) {
                c->set_error("rocsolver_dsyevd did not converge");
                return QPC_E_SETUP;
            }
        int *dpar = nullptr, *dloc = nullptr, *drows = nullptr;
        TRY(dalloc((void**)&dpar, sizeof(int) * ML));
        TRY(dalloc((void**)&dloc, sizeof(int) * ML));
        TRY(dalloc((void**)&drows, sizeof(int) * 2 * ML * nops));
        TRY(upload(c, dpar, ab.par.data(), (size_t)ML));
        TRY(upload(c, dloc, ab.loc.data(), (size_t)ML));
        std::vector<int> hrows((size_t)2 * ML * nops);
        std::vector<double> eig(ML);
        for (int s = 0; s < nops; ++s) {
            const int op = ops[s];
            std::vector<double> ws[2];
            for (int pb = 0; pb < 2; ++pb) ws[pb].assign(w[pb].begin() + s * n[pb], w[pb].begin() + (s + 1) * n[pb]);
            std::vector<qpc::AngRow> rows;
            if (qpc::angular_rows(M, L, ws, rows)) {
                c->set_error("angular eigenvalue categories");
                return QPC_E_SETUP;
            }
            E.emax[op] = 0.0;
            for (int r = 0; r < ML; ++r) {
                eig[r] = ws[rows[r].pb][rows[r].k];
                hrows[(size_t)(2 * s) * ML + r] = rows[r].pb;
                hrows[(size_t)(2 * s + 1) * ML + r] = rows[r].k;
            }
            qpc::pair_spectrum(eig.data(), M, L);
            for (int64_t r = E.p0; r < E.p1; ++r) E.emax[op] = std::max(E.emax[op], std::fabs(eig[r]));
            E.eig_h[op] = eig;
            TRY(upload(c, E.eig + (size_t)op * ML, eig.data(), (size_t)ML));
            QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));  // eig is reused
        }
        TRY(upload(c, drows, hrows.data(), hrows.size()));
        // projectors: conj(prj) of the merges' left operands, prj^T of their right ones
        cplx *cj[4] = {nullptr, nullptr, nullptr, nullptr}, *tr[4] = {nullptr, nullptr, nullptr, nullptr};
        const int mg[3][2] = {{1, 0}, {2, 1}, {3, 2}};  // (b, a): x1_2_x2, x2_2_y1, y1_2_y2
        for (int s = 0; s < nops; ++s) {
            const int op = ops[s];
            double* re = op == 0 ? E.tx1 : (!lg && op == 3) ? E.ty2 : nullptr;
            const bool need_cj = lg ? op == 2 : op >= 1, need_tr = lg ? op == 0 : op <= 2;
            if (lg && op == 2) cj[op] = E.ty1;  // tran_y1 = conj(prj_y1), also the merge's left operand
            else if (need_cj) TRY(dalloc((void**)&cj[op], sizeof(cplx) * mm));
            if (need_tr) TRY(dalloc((void**)&tr[op], sizeof(cplx) * mm));
            hipLaunchKernelGGL(k_ang_proj, dim3((unsigned)((ML + 255) / 256), (unsigned)ML), dim3(256), 0, c->stream,
                               dZ[0] + s * n[0] * n[0], dZ[1] ? dZ[1] + s * n[1] * n[1] : dZ[0], (int)n[0], (int)n[1],
                               drows + (size_t)(2 * s) * ML,
                               drows + (size_t)(2 * s + 1) * ML, dpar, dloc, L, ML, qpc::kAngImagMult[op] ? 1 : 0,
                               re, cj[op], tr[op]);
            TRY(launch_check(c, "angular projectors"));
        }
        // merged products P_b P_a^H = conj(prj_b) prj_a^T (merge_two, cn_1e.hpp:161-181)
        if (lg) {
            TRY(gemm<2>(c, (const double*)cj[2], ML, 0, 0, (const double*)tr[0], 2 * ML, (double*)E.x1x2, 2 * ML, ML,
                        2 * ML, 2 * ML, ML));
        } else {
            cplx* out[3] = {E.x1x2, E.x2y1, E.y1y2};
            for (int t = 0; t < 3; ++t)
                TRY(gemm<2>(c, (const double*)cj[mg[t][0]], ML, 0, 0, (const double*)tr[mg[t][1]], 2 * ML,
                            (double*)out[t], 2 * ML, ML, 2 * ML, 2 * ML, ML));
        }
        QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        TRY(ep_plan_share(c));
        return ep_permute(c);
    }();
    if (h) rocblas_destroy_handle(h);
    for (void* p : bufs) (void)hipFree(p);
    if (rc == QPC_OK) E.have = true;
    return rc;
}

// the time loop of the EP task (binsrc/sphTDSE.cpp:921-951) for either gauge
int ep_run(qpc_ctx* c, bool lg, const double* atx, const double* aty, int64_t nt_field, int64_t it0, int64_t it1,
           int use_mask) {
    if (!c || (nt_field > 0 && (!atx || !aty)) || it0 < 0 || it1 < it0) return QPC_E_ARG;
    TRY(ep_ready(c, use_mask != 0, lg));
    if (c->tsf_on && (lg || !c->have_tsf_xy)) {
        // sphTDSE.cpp:288 (velocity gauge only); the xy accumulate needs pmat1/pmat2
        c->set_error(lg ? "t-SURFF needs the velocity gauge" : "t-SURFF xy tables not uploaded");
        return QPC_E_STATE;
    }
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->fstream));
    TRY(qpc_host::record_room(c, it1 - it0));
    const int64_t itf = std::min(it1, nt_field);  // with-field steps [it0, itf)
    auto& E = c->ep;
    if (itf > it0) {
        const int64_t n = itf - it0;
        TRY(ep_set_scalars(c, atx + it0, aty + it0, n));
        TRY(ep_factor(c, 0, 0));
        // between with-field steps the state stays in x1's eigenbasis unless
        // an observer reads it after every step
        const bool keep = !c->rec_flags && !c->tsf_on;
        bool rot = false;
        for (int64_t k = 0; k < n; ++k) {
            const int set = (int)(k & 1);
            if (k + 1 < n) TRY(ep_factor(c, 1 - set, k + 1));  // one step ahead on the second stream
            TRY(qpc_host::record(c));
            QPC_HIP_CHECK(c, hipStreamWaitEvent(c->stream, E.set[set].ready, 0));
            const bool out = keep && k + 1 < n;
            TRY(lg ? lg_do_step(c, set, k, rot, out) : ep_do_step(c, set, k, rot, out));
            rot = out;
            QPC_HIP_CHECK(c, hipEventRecord(E.set[set].freed, c->stream));
            E.set[set].pending = true;
            // spec::accumulate(fx, gx, fy, gy, dt, coef) (sphTDSE.cpp:940)
            if (c->tsf_on) TRY(qpc_host::tsurff_step_xy(c, atx[it0 + k], atx[it0 + k + 1], aty[it0 + k], aty[it0 + k + 1]));
            if (use_mask) TRY(qpc_host::do_mask(c));
        }
    }
    for (int64_t it = std::max(it0, itf); it < it1; ++it) {
        TRY(qpc_host::record(c));
        TRY(qpc_host::solve_free(c, c->coef, c->tmpc));
        std::swap(c->coef, c->tmpc);
        if (c->tsf_on) TRY(qpc_host::tsurff_step_xy(c, 0.0, 0.0, 0.0, 0.0));  // sphTDSE.cpp:944
        if (use_mask) TRY(qpc_host::do_mask(c));
    }
    return QPC_OK;
}

int ep_step(qpc_ctx* c, bool lg, double ax_i, double ax_f, double ay_i, double ay_f) {
    if (!c) return QPC_E_ARG;
    TRY(ep_ready(c, false, lg));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->fstream));
    const double ax[2] = {ax_i, ax_f}, ay[2] = {ay_i, ay_f};
    TRY(ep_set_scalars(c, ax, ay, 1));
    TRY(ep_factor(c, 0, 0));
    QPC_HIP_CHECK(c, hipStreamWaitEvent(c->stream, c->ep.set[0].ready, 0));
    TRY(lg ? lg_do_step(c, 0, 0, false, false) : ep_do_step(c, 0, 0, false, false));
    QPC_HIP_CHECK(c, hipEventRecord(c->ep.set[0].freed, c->stream));
    c->ep.set[0].pending = true;
    QPC_HIP_CHECK(c, hipStreamSynchronize(c->stream));
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
    TRY(ep_prepare(c, isub, false));
    auto& E = c->ep;
    const int64_t ML = c->ML;
    const size_t mm = (size_t)ML * ML;
    const double* e[4] = {eig_x1, eig_x2, eig_y1, eig_y2};
    for (int op = 0; op < 4; ++op) {
        TRY(upload(c, E.eig + op * ML, e[op], (size_t)ML));
        E.eig_h[op].assign(e[op], e[op] + ML);
        E.emax[op] = 0.0;
        for (int64_t i = E.p0; i < E.p1; ++i) E.emax[op] = std::max(E.emax[op], std::fabs(e[op][i]));
    }
    TRY(ep_plan_share(c));
    TRY(upload(c, E.tx1, tran_x1, mm));
    TRY(upload(c, E.ty2, tran_y2, mm));
    TRY(upload(c, E.x1x2, x1_2_x2, mm));
    TRY(upload(c, E.x2y1, x2_2_y1, mm));
    TRY(upload(c, E.y1y2, y1_2_y2, mm));
    TRY(ep_permute(c));
    E.have = true;
    return QPC_OK;
}

int qpc_upload_angular_lgwr(qpc_ctx* c, int isub, const double* eig_x1, const double* eig_y1, const double* tran_x1,
                            const double* x1_2_y1, const double* tran_y1) {
    if (!c || !eig_x1 || !eig_y1 || !tran_x1 || !x1_2_y1 || !tran_y1) return QPC_E_ARG;
    if (isub < -1 || isub > 3 || c->batch) return QPC_E_ARG;
    TRY(ep_prepare(c, isub, true));
    auto& E = c->ep;
    const int64_t ML = c->ML;
    const size_t mm = (size_t)ML * ML;
    const double* e[4] = {eig_x1, nullptr, eig_y1, nullptr};
    for (int op = 0; op < 4; ++op) {
        E.emax[op] = 0.0;
        E.eig_h[op].clear();
        if (!e[op]) continue;
        TRY(upload(c, E.eig + op * ML, e[op], (size_t)ML));
        E.eig_h[op].assign(e[op], e[op] + ML);
        for (int64_t i = E.p0; i < E.p1; ++i) E.emax[op] = std::max(E.emax[op], std::fabs(e[op][i]));
    }
    TRY(ep_plan_share(c));
    TRY(upload(c, E.tx1, tran_x1, mm));
    TRY(upload(c, E.x1x2, x1_2_y1, mm));
    TRY(upload(c, E.ty1, tran_y1, mm));
    TRY(ep_permute(c));
    E.have = true;
    return QPC_OK;
}

int qpc_build_angular_ep(qpc_ctx* c, int isub) {
    if (!c || isub < -1 || isub > 3 || c->batch) return QPC_E_ARG;
    return ep_build_angular(c, isub, false);
}
int qpc_build_angular_lgwr(qpc_ctx* c, int isub) {
    if (!c || isub < -1 || isub > 3 || c->batch) return QPC_E_ARG;
    return ep_build_angular(c, isub, true);
}

int qpc_step_ep(qpc_ctx* c, double ax_i, double ax_f, double ay_i, double ay_f) {
    return ep_step(c, false, ax_i, ax_f, ay_i, ay_f);
}
int qpc_step_lgwr(qpc_ctx* c, double ex_i, double ex_f, double ey_i, double ey_f) {
    return ep_step(c, true, ex_i, ex_f, ey_i, ey_f);
}

int qpc_run_ep(qpc_ctx* c, const double* axtab, const double* aytab, int64_t nt_field, int64_t it0, int64_t it1,
               int use_mask) {
    return ep_run(c, false, axtab, aytab, nt_field, it0, it1, use_mask);
}
int qpc_run_lgwr(qpc_ctx* c, const double* extab, const double* eytab, int64_t nt_field, int64_t it0, int64_t it1,
                 int use_mask) {
    return ep_run(c, true, extab, eytab, nt_field, it0, it1, use_mask);
}

}  // extern "C"
