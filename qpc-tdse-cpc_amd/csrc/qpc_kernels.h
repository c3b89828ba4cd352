// qpc_kernels.h -- launchers of the HIP kernels of the propagator.
#pragma once
#include "qpc_common.h"

namespace qpc {

// rotations (FP64 matrix cores): largest l-dimension (two kernel builds: L
// up to 96, and up to 176 for the reference's l_max = 160 grids)
constexpr int kRotMaxL = 176;
// elements per m of a prepared operator
int64_t rot_prep_elems(int64_t L);
// prepared operator of P' = P (trans = 0) or P^H (trans = 1; P^T for real P)
void launch_rot_prep_real(hipStream_t s, const double* P, double* Pp, int64_t M, int64_t L, int trans);
void launch_rot_prep_cplx(hipStream_t s, const cplx* P, cplx* Pp, int64_t M, int64_t L, int trans);
// dst[m][k][x] = sum_i P'[m][k][i] src[m][i][x], x < N, with Pp from launch_rot_prep_*
void launch_rot_real(hipStream_t s, const double* Pp, const cplx* src, cplx* dst, int64_t M, int64_t L, int64_t N);
void launch_rot_cplx(hipStream_t s, const cplx* Pp, const cplx* src, cplx* dst, int64_t M, int64_t L, int64_t N);
// projection onto field-free eigenstates (eigenstate_sph::projection)
struct ProjGeom {
    int NL;      // l tables (lmax + 1)
    int nsel;    // states per l
    int N;       // radial functions
    int CMAX;    // largest number of channels sharing one l
    int RT, KS, CT;  // 16-row tiles of states, 4-row k steps, 16-column tiles
    int KCH, KSPLIT; // k steps per chunk, chunks
};
ProjGeom make_proj_geom(int nl, int nsel, int N, int cmax);
int64_t proj_prep_elems(const ProjGeom& g);
int64_t proj_part_elems(const ProjGeom& g);
// W[l][n][r] (left-hand basis S v) -> operand order; the strided form reads
// W[l * sl + n * sn + r] (the Coulomb-wave integrals ijl[n][l][r] of the
// photoelectron spectra: sl = N, sn = NL * N)
void launch_proj_prep(hipStream_t s, const double* W, double* Wp, const ProjGeom& g);
void launch_proj_prep_strided(hipStream_t s, const double* W, double* Wp, const ProjGeom& g, int64_t sl, int64_t sn);
// measure_lmd's finish: lmd[(n * M + im) * L + il] = sqrt(2/pi) cps[n * NL + l] proj[(im * L + il) * nsel + n],
// l = il + |m0 + im|
void launch_lmd_finish(hipStream_t s, const cplx* proj, const cplx* cps, cplx* lmd, int64_t nsel, int64_t M,
                       int64_t L, int64_t m0, int64_t NL);
// out[ch * nsel + n] = sum_r W_l(ch)[n][r] coef[ch][r]; chan[l * CMAX + c] lists
// the cnt[l] channels of l; part holds proj_part_elems doubles
void launch_project(hipStream_t s, const double* Wp, const cplx* coef, const int* chan, const int* cnt, double* part,
                    cplx* out, const ProjGeom& g);
// per-channel partial sums of <c|S|c> into red[nch]
void launch_norm(hipStream_t s, const double* S, const cplx* coef, double* red, int64_t nch, int64_t N, int64_t p);
// per-channel partial sums of Re observable_z into red[M*L]
void launch_dipole_z(hipStream_t s, const double* rband, const double* y10, const cplx* coef, double* red, int64_t M,
                     int64_t L, int64_t N, int64_t p);
// Observables recorded in the time loop, one value each for the whole state
// (kObs slots): bit 0 <c|S|c>, bit 1 Re observable_z with the band R (f =
// r), bit 2 the same with the band A (f = dV/dr), bit 3 <c|W|c>
// (outer-region overlap W), bit 4 observable_r's <x> and <y> with R (slots 4,
// 5), bit 5 the same with A (slots 6, 7).  launch_observe writes
// out[0..kObs) (device, 0 where the bit is clear) with red as scratch
// (kObs * M * L doubles); the per-channel partials are summed in a fixed
// order, so the values are reproducible run to run.
constexpr int kObs = 8;
// t-SURFF (spectrum_c_tsf::prepare / accumulate, momentum_1e.hpp:609-719).
// tmp[ch] = sum_i coef[ch][ib[i]] b0[i], tmp[M*L + ch] = the same with br.
void launch_tsf_temp(hipStream_t s, const cplx* coef, const int* ib, const double* b0, const double* br, int nb,
                     cplx* tmp, int64_t ML, int64_t N);
// Per (kr, theta/phi point): sum over channels of Y[ch][ktp] T(kr, ch) with
// T = -(j0 temq - jr temp)/2 - i az1 j0 (temp(l-1) pm(l-1) + temp(l+1) pm(l)),
// or with xy set (accumulate (xy), :716-776) the couplings to the (m-+1,
// l-+1) neighbours: - ias j0 temp(n) pm1/pm2(n) for m-1, - iap j0 temp(n)
// pm2/pm1(ch) for m+1 (l-1 / l+1), iap = i (ax1 + i ay1)/2, ias = i (ax1 -
// i ay1)/2; then work[kr][ktp] += sum exp(i (kr^2 time/2 + kr (ckt disz +
// stx disx + sty disy))) dt sqrt(2/pi).  Y = y0 e0 (-i)^(l+1) (channel-major,
// nktp = nkt * nkp points), j0/jr [nkr][nl], pm[ch] = <l|cos|l+1> of the
// channel, pm1/pm2[ch] = <l m|sin e^(-i ph)|l+-1 m+1>, ckt = cos(theta),
// stx/sty = sin(theta) cos/sin(phi) per point.
struct TsfArgs {
    const double *kr, *ckt, *stx, *sty, *j0, *jr, *pm, *pm1, *pm2;
    const cplx *Y, *tmp;
    cplx* work;
    int nkr, nktp, nl, M, L, m0, xy;
    double az1, ax1, ay1, time, disx, disy, disz, dt;
};
void launch_tsf_acc(hipStream_t s, const TsfArgs& a);
// integrate_hc's assembly for every l (basisfunc.hpp:92-106): h = V + l(l+1) c2m K
// + c2m T, upper = {S, -hdt h}, lower = {S, hdt h}, nl bands of `band` entries
void launch_hc_assemble(hipStream_t s, const double* S, const double* V, const double* K, const double* T, double c2m,
                        double hdt, cplx* upper, cplx* lower, int nl, int64_t band);
// lu_zsb (algorithm/ed.h:208-289) of the nl complex symmetric bands in place,
// one l per thread; tiny counts the pivots with |D|^2 < 1e-31
void launch_ldlt_csym(hipStream_t s, cplx* a, int nl, int N, int p, int* tiny);
struct ObsBands {
    const double *S, *R, *A, *W, *y10;
};
void launch_observe(hipStream_t s, const ObsBands& b, int flags, const cplx* coef, double* red, double* out,
                    int64_t M, int64_t L, int64_t m0, int64_t N, int64_t p);

}  // namespace qpc
