"""Photoelectron spectra by projection onto Coulomb waves (spectrum_c_pcs,
spectrum/momentum_1e.hpp:276-521): the library's host set-up
(qpc_setup_pcs: Coulomb functions by Steed's method, the Coulomb phase, the
B-spline integrals ijl) and the oracle's measure_lmd, pinned by references
that share none of their code: mpmath's Coulomb functions, scipy's log-gamma,
scipy's spherical Bessel functions with scipy B-splines for the Z = 0 case,
and closed-form spectra of a Gaussian wavepacket (sine and Hankel transforms)."""
import numpy as np
import pytest
import scipy.special as sp

import pyoracle as po
import qpc_gpu

mp = pytest.importorskip("mpmath")


@pytest.mark.parametrize("eta", [-0.4, -2.0, -10.0, -60.0, 0.0, 1.5])
def test_coulomb_f_matches_mpmath(eta):
    mp.mp.dps = 30
    for rho in (1e-6, 1e-3, 0.03, 0.08, 1.0, 9.5, 120.0, 900.0):
        F = qpc_gpu.coulomb_f(40, eta, rho)
        for L in (0, 1, 7, 20, 40):
            ref = float(mp.coulombf(L, eta, rho))
            if abs(ref) < 1e-250:
                assert abs(F[L]) < 1e-240
                continue
            assert abs(F[L] - ref) <= 1e-10 * abs(ref), (eta, rho, L, F[L], ref)


def test_coulomb_sigma_matches_scipy_loggamma():
    for eta in (-100.0, -5.0, -0.4, 0.0, 2.0):
        s = qpc_gpu.coulomb_sigma(60, eta)
        for l in (0, 1, 10, 60):
            ref = sp.loggamma(l + 1 + 1j * eta).imag
            d = np.angle(np.exp(1j * (s[l] - ref)))
            assert abs(d) < 1e-12, (eta, l)


def _gaussian_state(o, l, im=0, il=None):
    """c of u(r) = r^(l+1) exp(-r^2/2) in channel (im, il): the least-squares
    B-spline expansion S c = <B|u>, with <B|u> by scipy quadrature."""
    from test_oracle_pins import _scipy_bands  # noqa: F401  (same knots and quadrature)
    from scipy.interpolate import BSpline
    N, p = o.N, o.p
    kn = o.get(po.QOR_KNOTS)
    xg, wg = np.polynomial.legendre.leggauss(2 * p + 2)
    pts, wts = [], []
    for a, b in zip(kn[:-1], kn[1:]):
        if b > a:
            pts.append(0.5 * (a + b) + 0.5 * (b - a) * xg)
            wts.append(0.5 * (b - a) * wg)
    x, w = np.concatenate(pts), np.concatenate(wts)
    B = np.nan_to_num(np.array([BSpline(kn, np.eye(N + 2)[i], p, extrapolate=False)(x) for i in range(1, N + 1)]))
    rhs = B @ (w * x ** (l + 1) * np.exp(-x * x / 2))
    S = po.band_to_dense(o.get(po.QOR_S), N, p)
    c = np.zeros(o.M * o.L * N, np.complex128)
    il = l - abs(o.m0 + im) if il is None else il
    c[(im * o.L + il) * N:(im * o.L + il + 1) * N] = np.linalg.solve(S, rhs)
    return c


def test_planewave_lmd_of_gaussian_packets():
    """Z = 0 (eta = 0, F_l(0, kr) = kr j_l(kr)): for u = r exp(-r^2/2) (l = 0)
    lmd(k) = sqrt(2/pi) int u sin(kr)/k dr = exp(-k^2/2); for
    u = r^2 exp(-r^2/2) (l = 1) lmd(k) = sqrt(2/pi) e^(-i pi/2) int u j_1(kr) r dr
    = -i k exp(-k^2/2) (Hankel transforms of Gaussians)."""
    N, L, R = 160, 3, 16.0
    st = qpc_gpu.Setup(N, 7, L, 1, 0, 0.05, 0.0, R, 0.0, 1.0, R - 2, R, 0.125)
    o = po.Oracle(N=N, p=7, L=L, rmax=R, zc=0.0)
    kr, ijl, cps = st.pcs(40, 0.05, 3.0, zc=0.0)
    for l, want in ((0, lambda k: np.exp(-k * k / 2)), (1, lambda k: -1j * k * np.exp(-k * k / 2))):
        lmd = o.measure_lmd(_gaussian_state(o, l), ijl, cps)
        got = lmd[:, 0, l]
        assert np.abs(got - want(kr)).max() < 2e-7, l
        others = np.delete(lmd[:, 0, :], l, axis=1)
        assert np.abs(others).max() == 0.0


def test_ijl_zero_charge_matches_scipy_spherical_bessel():
    """ijl at Z = 0 against int B_j(r) k r j_l(k r) dr / k from scipy B-splines and
    scipy spherical Bessel functions on the same Gauss points."""
    from scipy.interpolate import BSpline
    N, L, R = 60, 4, 12.0
    st = qpc_gpu.Setup(N, 7, L, 1, 0, 0.05, 0.0, R, 0.0, 1.0, R - 2, R, 0.125)
    o = po.Oracle(N=N, p=7, L=L, rmax=R, zc=0.0)
    kr, ijl, cps = st.pcs(5, 0.3, 2.1, zc=0.0)
    kn = o.get(po.QOR_KNOTS)
    xg, wg = np.polynomial.legendre.leggauss(16)
    pts, wts = [], []
    for a, b in zip(kn[:-1], kn[1:]):
        if b > a:
            pts.append(0.5 * (a + b) + 0.5 * (b - a) * xg)
            wts.append(0.5 * (b - a) * wg)
    x, w = np.concatenate(pts), np.concatenate(wts)
    B = np.nan_to_num(np.array([BSpline(kn, np.eye(N + 2)[i], 7, extrapolate=False)(x) for i in range(1, N + 1)]))
    for ik, k in enumerate(kr):
        for l in range(L):
            ref = B @ (w * k * x * sp.spherical_jn(l, k * x)) / k
            assert np.abs(ijl[ik, l] - ref).max() < 1e-13 * max(1.0, np.abs(ref).max())
        assert np.allclose(cps[ik], np.exp(-0.5j * np.pi * np.arange(L)), atol=1e-15)


def test_hydrogen_bound_state_has_no_continuum_amplitude():
    """Z = 1: the hydrogen 1s state projects to nearly zero on the Coulomb
    waves of every k (orthogonality of the bound and continuum states of the
    same Hamiltonian; a unit-norm continuum packet has amplitudes of order
    one), which pins the charge and sign conventions of eta = -Z/k."""
    N, L, R = 400, 2, 80.0
    st = qpc_gpu.Setup(N, 7, L, 1, 0, 0.05, 0.0, R, 1.0, 1.0, R - 5, R, 0.125)
    o = po.Oracle(N=N, p=7, L=L, rmax=R)
    kr, ijl, cps = st.pcs(30, 0.4, 1.6, zc=1.0)
    _, s1 = o.eigenstate(0, 0, 0)
    lmd = o.measure_lmd(s1, ijl, cps)
    # the bound state's continuum amplitudes are small against those of a
    # unit-norm wavepacket (order 1) at every k
    assert np.abs(lmd).max() < 1e-10
