// qpc_kernels.hip -- HIP kernels (gfx950) of the Crank-Nicolson propagator.
//
// Reference operation order is kept inside each recurrence so that results
// agree with the reference's FP64 CPU run to rounding.  Citations are
// file:line in zzh535897/QPC-TDSE-CPC.
#include "qpc_kernels.h"

namespace qpc {
namespace {

// ---------------------------------------------------------------------------
// angular rotations: cn_transform_hl::construct / eliminate (cn_1e.hpp:101-147)
//   dst[m][k][x] = sum_i P'[k][i] src[m][i][x]   (complex state, x < N)
// on the FP64 matrix cores.  One wave owns one 16x16 output tile (16 rows k,
// 16 columns x) and accumulates it with v_mfma_f64_16x16x4f64 over K = L in
// steps of 4:  A = P' (16 x 4), B = src (4 x 16), D (16 x 16), with lane l
// holding A[l%16][l/16], B[l/16][l%16] and D[4*r+l/16][l%16] in register r
// (layout measured on gfx950).
// P' is prepared once per upload in exactly that operand order
// (Pp[m][row tile][k step][lane]), so a wave's A fetch is one contiguous run.
// A complex P' takes four real products per step (re*re - im*im, re*im + im*re).
typedef double v4d __attribute__((ext_vector_type(4)));

__device__ inline double rzero(double) { return 0.0; }
__device__ inline cplx rzero(cplx) { return {0.0, 0.0}; }

__device__ inline v4d mfma4(double a, double b, v4d c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }

// A work-group owns one 16-column tile of one m and up to NW row tiles (one
// per wave): the tile's src block (K = 4 KS rows x 16 columns, zero outside
// L x N) is fetched into LDS once and every wave reads its B operands from
// there.  Row tiles beyond NW go to further work-groups (blockIdx.y), so that
// L = 81 (6 row tiles, NW = 3) and N = 3000 make 2 x 188 work-groups; the
// grid's x extent is a multiple of 8 so that the work-groups of a column tile
// run on the same XCD and share its L2.  KR is the largest L of the build.
template <class TP, int KR, int NW>
__global__ __launch_bounds__(64 * NW) void k_rot(const TP* __restrict__ Pp, const cplx* __restrict__ src,
                                                 cplx* __restrict__ dst, int L, int N) {
    constexpr int NTH = 64 * NW, EPT = KR * 16 / NTH;
    static_assert(KR * 16 % NTH == 0, "whole staged elements per thread");
    __shared__ cplx sB[KR * 16];
    const int m = blockIdx.z;
    const int RT = (L + 15) / 16, KS = (L + 3) / 4;
    const int c0 = blockIdx.x * 16;
    if (c0 >= N) return;  // padding column tile: the whole work-group leaves before any barrier
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    const int rt = blockIdx.y * NW + w;
    const int kr = l / 16, col = c0 + (l % 16);
    // src rows [0, KR) x columns [c0, c0 + 16): unconditional clamped loads,
    // zero outside L x N (the A operand is zero for k >= L, and 0 * 0 stays 0)
    const cplx* Bm = src + (int64_t)m * L * N;
    {
        cplx bv[EPT];
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            const int e = threadIdx.x + j * NTH;
            bv[j] = Bm[(int64_t)min(e >> 4, L - 1) * N + min(c0 + (e & 15), N - 1)];
        }
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            const int e = threadIdx.x + j * NTH;
            sB[e] = (e >> 4) < L && c0 + (e & 15) < N ? bv[j] : cplx{0.0, 0.0};
        }
    }
    __syncthreads();
    if (rt >= RT) return;
    const TP* A = Pp + ((int64_t)(m * RT + rt) * KS) * 64 + l;
    constexpr int KSMAX = KR / 4;
    TP av[KSMAX];
#pragma unroll
    for (int ks = 0; ks < KSMAX; ++ks) av[ks] = A[(int64_t)min(ks, KS - 1) * 64];
    v4d dre = {0.0, 0.0, 0.0, 0.0}, dim = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < KSMAX; ++ks) {
        if (ks < KS) {
            const TP a = av[ks];
            const cplx b = sB[(ks * 4 + kr) * 16 + (l % 16)];
            if constexpr (sizeof(TP) == sizeof(double)) {
                dre = mfma4(a, b.re, dre);
                dim = mfma4(a, b.im, dim);
            } else {
                dre = mfma4(a.re, b.re, dre);
                dre = mfma4(-a.im, b.im, dre);
                dim = mfma4(a.re, b.im, dim);
                dim = mfma4(a.im, b.re, dim);
            }
        }
    }
    if (col >= N) return;
    cplx* D = dst + (int64_t)m * L * N + col;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = rt * 16 + 4 * r + kr;
        if (row < L) D[(int64_t)row * N] = cplx{dre[r], dim[r]};
    }
}

// Pp[m][rt][ks][l] = P'[rt*16 + l%16][ks*4 + l/16], P' = P or P^H, zero outside L x L
template <class T>
__global__ void k_rot_prep(const T* __restrict__ P, T* __restrict__ Pp, int L, int64_t total, int trans) {
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= total) return;
    const int RT = (L + 15) / 16, KS = (L + 3) / 4;
    const int l = (int)(id % 64);
    const int ks = (int)((id / 64) % KS);
    const int rt = (int)((id / (64 * (int64_t)KS)) % RT);
    const int64_t m = id / (64 * (int64_t)KS * RT);
    const int k = rt * 16 + l % 16, i = ks * 4 + l / 16;
    const T* Pm = P + m * L * L;
    T v = rzero(T{});
    if (k < L && i < L) {
        if (trans) {
            v = Pm[(int64_t)i * L + k];
            if constexpr (sizeof(T) == sizeof(cplx)) v = conjc(v);
        } else {
            v = Pm[(int64_t)k * L + i];
        }
    }
    Pp[id] = v;
}

// ---------------------------------------------------------------------------
// projection onto field-free eigenstates: eigenstate_sph::projection
// (population_1e.hpp:29-55).  For every l, the columns are the (channel,
// re/im) pairs of the channels with that l, and
//   P_l[n][col] = sum_r W_l[n][r] * coef[ch(col)][r].(re|im)
// is a GEMM on the FP64 matrix cores: one wave per 16x16 tile and K chunk
// (split K, partial tiles summed in a fixed order by k_proj_reduce).
// W_l is prepared in operand order Wp[l][rt][ks][lane] = W_l[rt*16+lane%16][ks*4+lane/16].
__global__ void k_proj_prep(const double* __restrict__ W, double* __restrict__ Wp, int nsel, int N, int RT,
                            int KS, int64_t total, int64_t sl, int64_t sn) {
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= total) return;
    const int li = (int)(id % 64);
    const int ks = (int)((id / 64) % KS);
    const int rt = (int)((id / (64 * (int64_t)KS)) % RT);
    const int64_t l = id / (64 * (int64_t)KS * RT);
    const int n = rt * 16 + li % 16, r = ks * 4 + li / 16;
    Wp[id] = (n < nsel && r < N) ? W[l * sl + (int64_t)n * sn + r] : 0.0;
}

// spectrum_c::measure_lmd (spectrum/momentum_1e.hpp:384-401) after the GEMM:
// the Coulomb phase factor, sqrt(2/pi) and the (ikr, im, il) layout
__global__ void k_lmd_finish(const cplx* __restrict__ proj, const cplx* __restrict__ cps, cplx* __restrict__ lmd,
                             int nsel, int M, int L, int m0, int NL) {
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)nsel * M * L;
    if (id >= total) return;
    const int il = (int)(id % L), im = (int)((id / L) % M);
    const int n = (int)(id / ((int64_t)M * L));
    const int m = im + m0, l = il + (m < 0 ? -m : m);
    const cplx v = proj[((int64_t)im * L + il) * nsel + n] * cps[(int64_t)n * NL + l];
    lmd[id] = v * 0.797884560802865;  // sqrt(2/pi)
}

__global__ __launch_bounds__(64) void k_project(const double* __restrict__ Wp, const cplx* __restrict__ coef,
                                                const int* __restrict__ chan, const int* __restrict__ cnt,
                                                double* __restrict__ part, ProjGeom pg) {
    const int rt = blockIdx.x % pg.RT, ct = blockIdx.x / pg.RT;
    const int l = blockIdx.y, kc = blockIdx.z;
    const int ncol = 2 * cnt[l];
    if (ct * 16 >= ncol) return;  // uniform over the wave
    const int li = threadIdx.x;
    const int j = ct * 16 + li % 16, kr = li / 16;
    const bool valid = j < ncol;
    const int ch = valid ? chan[l * pg.CMAX + j / 2] : 0;
    const int pt = j & 1;
    const double* cf = reinterpret_cast<const double*>(coef + (int64_t)ch * pg.N) + pt;
    const double* A = Wp + ((int64_t)(l * pg.RT + rt) * pg.KS) * 64 + li;
    const int ks0 = kc * pg.KCH, ks1 = min(pg.KS, ks0 + pg.KCH);
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    constexpr int PB = 16;  // k steps whose operands are fetched together
    for (int kb = ks0; kb < ks1; kb += PB) {
        double av[PB], bv[PB];
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            const int ks = kb + q, r = ks * 4 + kr;
            av[q] = ks < ks1 ? A[(int64_t)ks * 64] : 0.0;
            bv[q] = (ks < ks1 && valid && r < pg.N) ? cf[2 * (int64_t)r] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < PB; ++q) acc = mfma4(av[q], bv[q], acc);
    }
    double* o = part + ((((int64_t)kc * pg.NL + l) * pg.RT + rt) * pg.CT + ct) * 256 + li;
#pragma unroll
    for (int r = 0; r < 4; ++r) o[r * 64] = acc[r];
}

// sum of the K chunks and scatter to out[ch * nsel + n] (complex)
__global__ void k_proj_reduce(const double* __restrict__ part, const int* __restrict__ chan,
                              const int* __restrict__ cnt, cplx* __restrict__ out, ProjGeom pg) {
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // (l, rt, ct, r, li)
    const int64_t total = (int64_t)pg.NL * pg.RT * pg.CT * 256;
    if (id >= total) return;
    const int li = (int)(id % 64), r = (int)((id / 64) % 4);
    const int64_t tile = id / 256;
    const int ct = (int)(tile % pg.CT), rt = (int)((tile / pg.CT) % pg.RT);
    const int l = (int)(tile / ((int64_t)pg.CT * pg.RT));
    const int n = rt * 16 + 4 * r + li / 16, j = ct * 16 + li % 16;
    if (n >= pg.nsel || j >= 2 * cnt[l]) return;
    double s = 0.0;
    for (int kc = 0; kc < pg.KSPLIT; ++kc)
        s += part[((((int64_t)kc * pg.NL + l) * pg.RT + rt) * pg.CT + ct) * 256 + r * 64 + li];
    const int ch = chan[l * pg.CMAX + j / 2];
    double* o = reinterpret_cast<double*>(out + (int64_t)ch * pg.nsel + n);
    o[j & 1] = s;
}

// ---------------------------------------------------------------------------
// bilinear forms x^H A y with a real symmetric band (symb_prj_vecd,
// blas2.hpp:1110-1127); one block per channel, fixed-shape tree reduction.
constexpr int kRedThreads = 256;
__device__ inline cplx band_row(const double* A, const cplx* y, int i, int N, int ne) {
    cplx t = {0.0, 0.0};
    for (int j = 1; j < ne; ++j)
        if (j <= i) t += A[(int64_t)(i - j) * ne + j] * y[i - j];
    for (int k = 0; k < ne; ++k)
        if (i + k < N) t += A[(int64_t)i * ne + k] * y[i + k];
    return t;
}
__device__ inline double block_sum(double v) {
    __shared__ double sh[kRedThreads];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = kRedThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    return sh[0];
}

__global__ __launch_bounds__(kRedThreads) void k_norm(const double* __restrict__ S, const cplx* __restrict__ coef,
                                                      double* __restrict__ red, int N, int ne) {
    const int iml = blockIdx.x;
    const cplx* x = coef + (int64_t)iml * N;
    double acc = 0.0;
    for (int i = threadIdx.x; i < N; i += blockDim.x) acc += mulc(band_row(S, x, i, N, ne), x[i]).re;
    double s = block_sum(acc);
    if (threadIdx.x == 0) red[iml] = s;
}

// observable_z::observe (observable/1e.hpp:26-49), real part per channel
__global__ __launch_bounds__(kRedThreads) void k_dipole_z(const double* __restrict__ R, const double* __restrict__ y10,
                                                          const cplx* __restrict__ coef, double* __restrict__ red,
                                                          int L, int N, int ne) {
    const int iml = blockIdx.x;
    const int im = iml / L, il = iml % L;
    const cplx* lhs = coef + (int64_t)iml * N;
    const cplx* base = coef + (int64_t)im * L * N;
    double acc = 0.0;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        if (il > 0) acc += mulc(band_row(R, base + (int64_t)(il - 1) * N, i, N, ne), lhs[i]).re * y10[im * L + il - 1];
        if (il + 1 < L) acc += mulc(band_row(R, base + (int64_t)(il + 1) * N, i, N, ne), lhs[i]).re * y10[im * L + il];
    }
    double s = block_sum(acc);
    if (threadIdx.x == 0) red[iml] = s;
}

// closed forms of integrator_angu::integrate_sinexp1..4 (utilities/
// basisfunc.hpp:243-279): <l',m+-1| sin(theta) e^(-+i phi) |l,m> couplings
// of operator_y11<-1> (1 outer, 2 inner) and operator_y11<+1> (3 outer, 4 inner)
__device__ inline double sinexp(int which, int m, int l) {
    auto q = [](double a, double b) { return b != 0.0 ? a / b : 0.0; };
    if (which == 1) return -sqrt(q((double)(l + m + 2) * (l + m + 1), (double)(2 * l + 1) * (2 * l + 3)));
    if (which == 2) return +sqrt(q((double)(l - m) * (l - m - 1), (double)(2 * l + 1) * (2 * l - 1)));
    if (which == 3) return -sqrt(q((double)(l + m) * (l + m - 1), (double)(2 * l + 1) * (2 * l - 1)));
    return +sqrt(q((double)(l - m + 2) * (l - m + 1), (double)(2 * l + 1) * (2 * l + 3)));
}
__device__ inline cplx cmulc(cplx a, cplx b) { return mulc(a, b); }  // a * conj(b)

// the observables of launch_observe for one channel per block: partial q of
// channel iml at red[q * ML + iml].  observable_r (observable/1e.hpp:76-126):
// channel (m, l) couples to (m+1, l+-1) in observe_rm (sinexp1 outer,
// sinexp2 inner) and to (m-1, l+-1) in observe_rp (sinexp4 for l+1, sinexp3
// for l-1), the partner channel existing when in_check finds it in the grid
// (dimension/1e.hpp:97-100); the channel's share of <x> is Re(rm + rp)/2 and
// of <y> Im(rp - rm)/2.
__global__ __launch_bounds__(kRedThreads) void k_observe(ObsBands b, int flags, const cplx* __restrict__ coef,
                                                         double* __restrict__ red, int M, int L, int m0, int N,
                                                         int ne, int ML) {
    const int iml = blockIdx.x;
    const int im = iml / L, il = iml % L;
    const int m = im + m0, l = il + (m < 0 ? -m : m);
    const cplx* x = coef + (int64_t)iml * N;
    const cplx* base = coef + (int64_t)im * L * N;
    const cplx* lo = il > 0 ? base + (int64_t)(il - 1) * N : nullptr;
    const cplx* hi = il + 1 < L ? base + (int64_t)(il + 1) * N : nullptr;
    const double ylo = il > 0 ? b.y10[im * L + il - 1] : 0.0, yhi = il + 1 < L ? b.y10[im * L + il] : 0.0;
    // observable_r partners: k = 0, 1 (rm: m+1, l+1 / l-1), k = 2, 3 (rp: m-1, l+1 / l-1)
    const cplx* nb[4] = {nullptr, nullptr, nullptr, nullptr};
    double nc[4] = {0.0, 0.0, 0.0, 0.0};
    if (flags & 48) {
        const int dm[4] = {+1, +1, -1, -1}, dl[4] = {+1, -1, +1, -1}, fw[4] = {1, 2, 4, 3};
        for (int k = 0; k < 4; ++k) {
            const int im2 = im + dm[k];
            if (im2 < 0 || im2 >= M) continue;
            const int m2 = im2 + m0;
            const int il2 = l + dl[k] - (m2 < 0 ? -m2 : m2);
            if (il2 < 0 || il2 >= L) continue;
            nb[k] = coef + ((int64_t)im2 * L + il2) * N;
            nc[k] = sinexp(fw[k], m, l);
        }
    }
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    cplx rr[2][2] = {{{0.0, 0.0}, {0.0, 0.0}}, {{0.0, 0.0}, {0.0, 0.0}}};  // [R or A][rm or rp]
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        const cplx xi = x[i];
        if (flags & 1) acc[0] += mulc(band_row(b.S, x, i, N, ne), xi).re;
        if (flags & 2) {
            if (lo) acc[1] += mulc(band_row(b.R, lo, i, N, ne), xi).re * ylo;
            if (hi) acc[1] += mulc(band_row(b.R, hi, i, N, ne), xi).re * yhi;
        }
        if (flags & 4) {
            if (lo) acc[2] += mulc(band_row(b.A, lo, i, N, ne), xi).re * ylo;
            if (hi) acc[2] += mulc(band_row(b.A, hi, i, N, ne), xi).re * yhi;
        }
        if (flags & 8) acc[3] += mulc(band_row(b.W, x, i, N, ne), xi).re;
        for (int f = 0; f < 2; ++f) {
            if (!(flags & (16 << f))) continue;
            const double* band = f ? b.A : b.R;
            for (int k = 0; k < 4; ++k)
                if (nb[k]) rr[f][k / 2] += cmulc(band_row(band, nb[k], i, N, ne), xi) * nc[k];
        }
    }
    for (int q = 0; q < 4; ++q)
        if (flags & (1 << q)) {
            const double s = block_sum(acc[q]);
            if (threadIdx.x == 0) red[(int64_t)q * ML + iml] = s;
            __syncthreads();  // block_sum's buffer is reused by the next sum
        }
    for (int f = 0; f < 2; ++f)
        if (flags & (16 << f)) {
            double v[4];
            v[0] = block_sum(rr[f][0].re);
            __syncthreads();
            v[1] = block_sum(rr[f][0].im);
            __syncthreads();
            v[2] = block_sum(rr[f][1].re);
            __syncthreads();
            v[3] = block_sum(rr[f][1].im);
            __syncthreads();
            if (threadIdx.x == 0) {
                red[(int64_t)(4 + 2 * f) * ML + iml] = 0.5 * (v[0] + v[2]);  // Re(rm + rp) / 2
                red[(int64_t)(5 + 2 * f) * ML + iml] = 0.5 * (v[3] - v[1]);  // Im(rp - rm) / 2
            }
        }
}

// out[q] = sum over channels of red[q * ML + ch] (strided partials, then a
// fixed tree); 0 where the observable is not requested
__device__ inline bool obs_slot_on(int flags, int q) { return q < 4 ? (flags >> q) & 1 : (flags >> (4 + (q - 4) / 2)) & 1; }
__global__ __launch_bounds__(kRedThreads) void k_observe_sum(const double* __restrict__ red, int flags, int ML,
                                                             double* __restrict__ out) {
    for (int q = 0; q < kObs; ++q) {
        double a = 0.0;
        if (obs_slot_on(flags, q))
            for (int i = threadIdx.x; i < ML; i += blockDim.x) a += red[(int64_t)q * ML + i];
        const double s = block_sum(a);
        if (threadIdx.x == 0) out[q] = s;
        __syncthreads();
    }
}

}  // namespace

int64_t rot_prep_elems(int64_t L) { return ((L + 15) / 16) * ((L + 3) / 4) * 64; }

template <class TP, int KR, int NW>
static void rot_launch_k(hipStream_t s, const TP* Pp, const cplx* src, cplx* dst, int64_t M, int64_t L, int64_t N) {
    const int64_t RT = (L + 15) / 16;
    dim3 grid((unsigned)(((N + 15) / 16 + 7) / 8 * 8), (unsigned)((RT + NW - 1) / NW), (unsigned)M);
    hipLaunchKernelGGL((k_rot<TP, KR, NW>), grid, dim3(64 * NW), 0, s, Pp, src, dst, (int)L, (int)N);
}
template <class TP>
static void rot_launch(hipStream_t s, const TP* Pp, const cplx* src, cplx* dst, int64_t M, int64_t L, int64_t N) {
    if (L <= 96) rot_launch_k<TP, 96, 3>(s, Pp, src, dst, M, L, N);
    else rot_launch_k<TP, kRotMaxL, 4>(s, Pp, src, dst, M, L, N);
}

void launch_rot_real(hipStream_t s, const double* Pp, const cplx* src, cplx* dst, int64_t M, int64_t L, int64_t N) {
    rot_launch<double>(s, Pp, src, dst, M, L, N);
}

void launch_rot_cplx(hipStream_t s, const cplx* Pp, const cplx* src, cplx* dst, int64_t M, int64_t L, int64_t N) {
    rot_launch<cplx>(s, Pp, src, dst, M, L, N);
}

void launch_rot_prep_real(hipStream_t s, const double* P, double* Pp, int64_t M, int64_t L, int trans) {
    const int64_t total = M * rot_prep_elems(L);
    hipLaunchKernelGGL(k_rot_prep<double>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, P, Pp, (int)L,
                       total, trans);
}
void launch_rot_prep_cplx(hipStream_t s, const cplx* P, cplx* Pp, int64_t M, int64_t L, int trans) {
    const int64_t total = M * rot_prep_elems(L);
    hipLaunchKernelGGL(k_rot_prep<cplx>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, P, Pp, (int)L,
                       total, trans);
}

ProjGeom make_proj_geom(int nl, int nsel, int N, int cmax) {
    ProjGeom g;
    g.NL = nl;
    g.nsel = nsel;
    g.N = N;
    g.CMAX = cmax;
    g.RT = (nsel + 15) / 16;
    g.KS = (N + 3) / 4;
    g.CT = (2 * cmax + 15) / 16;
    g.KCH = 64;
    g.KSPLIT = (g.KS + g.KCH - 1) / g.KCH;
    return g;
}
int64_t proj_prep_elems(const ProjGeom& g) { return (int64_t)g.NL * g.RT * g.KS * 64; }
int64_t proj_part_elems(const ProjGeom& g) { return (int64_t)g.KSPLIT * g.NL * g.RT * g.CT * 256; }

void launch_proj_prep(hipStream_t s, const double* W, double* Wp, const ProjGeom& g) {
    launch_proj_prep_strided(s, W, Wp, g, (int64_t)g.nsel * g.N, g.N);
}
void launch_proj_prep_strided(hipStream_t s, const double* W, double* Wp, const ProjGeom& g, int64_t sl, int64_t sn) {
    const int64_t total = proj_prep_elems(g);
    hipLaunchKernelGGL(k_proj_prep, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, W, Wp, g.nsel, g.N, g.RT,
                       g.KS, total, sl, sn);
}
void launch_lmd_finish(hipStream_t s, const cplx* proj, const cplx* cps, cplx* lmd, int64_t nsel, int64_t M,
                       int64_t L, int64_t m0, int64_t NL) {
    const int64_t total = nsel * M * L;
    hipLaunchKernelGGL(k_lmd_finish, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, proj, cps, lmd, (int)nsel,
                       (int)M, (int)L, (int)m0, (int)NL);
}
void launch_project(hipStream_t s, const double* Wp, const cplx* coef, const int* chan, const int* cnt, double* part,
                    cplx* out, const ProjGeom& g) {
    hipLaunchKernelGGL(k_project, dim3(g.RT * g.CT, g.NL, g.KSPLIT), dim3(64), 0, s, Wp, coef, chan, cnt, part, g);
    const int64_t total = (int64_t)g.NL * g.RT * g.CT * 256;
    hipLaunchKernelGGL(k_proj_reduce, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, part, chan, cnt, out, g);
}

void launch_norm(hipStream_t s, const double* S, const cplx* coef, double* red, int64_t nch, int64_t N, int64_t p) {
    hipLaunchKernelGGL(k_norm, dim3((unsigned)nch), dim3(kRedThreads), 0, s, S, coef, red, (int)N, (int)(p + 1));
}

void launch_dipole_z(hipStream_t s, const double* rband, const double* y10, const cplx* coef, double* red, int64_t M,
                     int64_t L, int64_t N, int64_t p) {
    hipLaunchKernelGGL(k_dipole_z, dim3((unsigned)(M * L)), dim3(kRedThreads), 0, s, rband, y10, coef, red, (int)L,
                       (int)N, (int)(p + 1));
}

void launch_observe(hipStream_t s, const ObsBands& b, int flags, const cplx* coef, double* red, double* out,
                    int64_t M, int64_t L, int64_t m0, int64_t N, int64_t p) {
    const int ML = (int)(M * L);
    hipLaunchKernelGGL(k_observe, dim3((unsigned)ML), dim3(kRedThreads), 0, s, b, flags, coef, red, (int)M, (int)L,
                       (int)m0, (int)N, (int)(p + 1), ML);
    hipLaunchKernelGGL(k_observe_sum, dim3(1), dim3(kRedThreads), 0, s, red, flags, ML, out);
}


// ---------------------------------------------------------------------------
// field-free bands on the device (cn_component_hc::initialize, cn_1e.hpp:642-680)
namespace {
__global__ __launch_bounds__(256) void k_hc_assemble(const double* __restrict__ S, const double* __restrict__ V,
                                                     const double* __restrict__ K, const double* __restrict__ T,
                                                     double c2m, double hdt, cplx* __restrict__ upper,
                                                     cplx* __restrict__ lower, int64_t band) {
    const int l = blockIdx.y;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= band) return;
    const double cent = (double)l * (l + 1) * c2m;
    const double h = hc_entry(V[i], cent, K[i], c2m, T[i]);
    upper[(int64_t)l * band + i] = cplx{S[i], -hdt * h};
    lower[(int64_t)l * band + i] = cplx{S[i], hdt * h};
}

// LDL^T without pivoting of one complex symmetric band per thread (row j:
// D_j at [0], L(j+c, j) at [c]); the P rows above the current one are kept
// in registers
template <int P>
__global__ __launch_bounds__(64) void k_ldlt_csym(cplx* __restrict__ a, int nl, int N, int* __restrict__ tiny) {
    constexpr int NE = P + 1;
    const int l = blockIdx.x * 64 + threadIdx.x;
    if (l >= nl) return;
    cplx* al = a + (int64_t)l * N * NE;
    cplx w[P][NE];  // rows j-P .. j-1 (w[P-1] = row j-1)
#pragma unroll
    for (int q = 0; q < P; ++q)
#pragma unroll
        for (int i = 0; i < NE; ++i) w[q][i] = cplx{0.0, 0.0};
    int small = 0;
    for (int j = 0; j < N; ++j) {
        cplx aj[NE];
#pragma unroll
        for (int i = 0; i < NE; ++i) aj[i] = i < N - j ? al[(int64_t)j * NE + i] : cplx{0.0, 0.0};
#pragma unroll
        for (int k = 1; k <= P; ++k) {
            if (k > j) break;
            const cplx* ak = w[P - k];  // row j-k
            const cplx f = ldlt_mul(ak[k], ak[0]);
#pragma unroll
            for (int i = 0; i + k < NE; ++i) ldlt_sub(aj[i], f, ak[k + i]);
        }
        if (ldlt_norm2(aj[0]) < 1e-31) ++small;
        const cplx r = ldlt_recip(aj[0]);
#pragma unroll
        for (int i = 1; i < NE; ++i) aj[i] = ldlt_mul(aj[i], r);
#pragma unroll
        for (int i = 0; i < NE; ++i) al[(int64_t)j * NE + i] = aj[i];
#pragma unroll
        for (int q = 0; q + 1 < P; ++q)
#pragma unroll
            for (int i = 0; i < NE; ++i) w[q][i] = w[q + 1][i];
#pragma unroll
        for (int i = 0; i < NE; ++i) w[P - 1][i] = aj[i];
    }
    if (small) atomicAdd(tiny, small);
}
}  // namespace

void launch_hc_assemble(hipStream_t s, const double* S, const double* V, const double* K, const double* T, double c2m,
                        double hdt, cplx* upper, cplx* lower, int nl, int64_t band) {
    hipLaunchKernelGGL(k_hc_assemble, dim3((unsigned)((band + 255) / 256), (unsigned)nl), dim3(256), 0, s, S, V, K, T,
                       c2m, hdt, upper, lower, band);
}
void launch_ldlt_csym(hipStream_t s, cplx* a, int nl, int N, int p, int* tiny) {
    if (p == 8) hipLaunchKernelGGL(k_ldlt_csym<8>, dim3((unsigned)((nl + 63) / 64)), dim3(64), 0, s, a, nl, N, tiny);
    else hipLaunchKernelGGL(k_ldlt_csym<7>, dim3((unsigned)((nl + 63) / 64)), dim3(64), 0, s, a, nl, N, tiny);
}

// ---------------------------------------------------------------------------
// t-SURFF (spectrum_c_tsf, momentum_1e.hpp:609-776)
namespace {
__global__ __launch_bounds__(256) void k_tsf_temp(const cplx* __restrict__ coef, const int* __restrict__ ib,
                                                  const double* __restrict__ b0, const double* __restrict__ br, int nb,
                                                  cplx* __restrict__ tmp, int64_t ML, int64_t N) {
    const int64_t ch = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (ch >= ML) return;
    const cplx* c = coef + ch * N;
    cplx s0 = {0.0, 0.0}, s1 = {0.0, 0.0};
    for (int i = 0; i < nb; ++i) s0 += c[ib[i]] * b0[i];
    for (int i = 0; i < nb; ++i) s1 += c[ib[i]] * br[i];
    tmp[ch] = s0;
    tmp[ML + ch] = s1;
}

// one work-group per kr: its threads form T(kr, ch) of every channel in LDS,
// then each thread sums over the channels for its (theta, phi) points
constexpr int kTsfMaxCh = 4096;
__global__ __launch_bounds__(256) void k_tsf_acc(TsfArgs a) {
    __shared__ cplx sT[kTsfMaxCh];
    const int ikr = blockIdx.x;
    const int ML = a.M * a.L;
    const cplx* temp = a.tmp;
    const cplx* temq = a.tmp + ML;
    for (int ch = threadIdx.x; ch < ML; ch += 256) {
        const int im = ch / a.L, il = ch % a.L;
        const int m = a.m0 + im, l = il + (m < 0 ? -m : m);
        const double j0 = a.j0[(int64_t)ikr * a.nl + l], jr = a.jr[(int64_t)ikr * a.nl + l];
        const cplx bt = temp[ch], rt = temq[ch];
        const cplx bp = il > 0 ? temp[ch - 1] : cplx{0.0, 0.0};
        const double pp = il > 0 ? a.pm[ch - 1] : 0.0;
        const cplx bn = il + 1 < a.L ? temp[ch + 1] : cplx{0.0, 0.0};
        const double pn = il + 1 < a.L ? a.pm[ch] : 0.0;
        const cplx jdb = -0.5 * (j0 * rt - jr * bt);
        if (a.xy) {
            // neighbours (m + dm, l + dl) inside the grid (dimension/1e.hpp:97-100)
            const cplx iap = cplx{-a.ay1 * 0.5, a.ax1 * 0.5}, ias = cplx{a.ay1 * 0.5, a.ax1 * 0.5};
            cplx t = jdb;
            for (int dm = -1; dm <= 1; dm += 2) {
                const int jm = im + dm;
                if (jm < 0 || jm >= a.M) continue;
                const int mj = a.m0 + jm, amj = mj < 0 ? -mj : mj;
                for (int dl = -1; dl <= 1; dl += 2) {
                    const int jl = l + dl - amj;
                    if (jl < 0 || jl >= a.L) continue;
                    const int n = jm * a.L + jl;
                    const double e = dm < 0 ? (dl < 0 ? a.pm1[n] : a.pm2[n]) : (dl < 0 ? a.pm2[ch] : a.pm1[ch]);
                    t -= (dm < 0 ? ias : iap) * j0 * temp[n] * e;
                }
            }
            sT[ch] = t;
            continue;
        }
        const cplx iaz = {0.0, a.az1};
        const cplx jbp = j0 * bp, jbn = j0 * bn;
        sT[ch] = jdb - iaz * jbp * pp - iaz * jbn * pn;
    }
    __syncthreads();
    const double k = a.kr[ikr];
    for (int t = threadIdx.x; t < a.nktp; t += 256) {
        cplx sum = {0.0, 0.0};
        for (int ch = 0; ch < ML; ++ch) sum += a.Y[(int64_t)ch * a.nktp + t] * sT[ch];
        const double ph = 0.5 * k * k * a.time + k * a.ckt[t] * a.disz + k * a.stx[t] * a.disx + k * a.sty[t] * a.disy;
        double sn, cs;
        sincos(ph, &sn, &cs);
        cplx* w = a.work + (int64_t)ikr * a.nktp + t;
        *w += sum * cplx{cs, sn} * a.dt * 0.797884560802865;
    }
}
}  // namespace

void launch_tsf_temp(hipStream_t s, const cplx* coef, const int* ib, const double* b0, const double* br, int nb,
                     cplx* tmp, int64_t ML, int64_t N) {
    hipLaunchKernelGGL(k_tsf_temp, dim3((unsigned)((ML + 255) / 256)), dim3(256), 0, s, coef, ib, b0, br, nb, tmp, ML,
                       N);
}
void launch_tsf_acc(hipStream_t s, const TsfArgs& a) {
    hipLaunchKernelGGL(k_tsf_acc, dim3((unsigned)a.nkr), dim3(256), 0, s, a);
}
}  // namespace qpc
