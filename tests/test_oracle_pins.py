"""Pins of the CPU oracle against references that do not share its code:
numpy/scipy (Gauss-Legendre, B-spline evaluation, dense linear algebra),
closed forms (hydrogen levels, Golub-Welsch, Cayley phase, sin^2 pulse) and
invariants (unitarity, time reversal).  The reference repository ships no
tests or golden data, so these are the pins (see tests/golden/README.md)."""
import numpy as np
import pytest
import scipy.linalg as sl
from scipy.interpolate import BSpline

import pyoracle as po

N_SMALL, P, L_SMALL, R_SMALL = 30, 7, 4, 12.0


@pytest.fixture(scope="module")
def small():
    return po.Oracle(N=N_SMALL, p=P, L=L_SMALL, rmax=R_SMALL, dt=0.05, mask_rmin=9.0, mask_rmax=12.0,
                     mask_fact=0.125)


def test_gauss_legendre_matches_numpy():
    for n in (2, 3, 7, 16, 20, 33):
        x, w = po.gauss_legendre(n)
        xr, wr = np.polynomial.legendre.leggauss(n)
        assert np.abs(x - xr).max() < 2e-15
        assert np.abs(w - wr).max() < 2e-15


def _scipy_bands(o, N, p):
    """Bands <Bi^(mi)|f|Bj^(mj)> from scipy B-splines on the oracle's knots and
    numpy Gauss-Legendre (2p+2 points per knot interval)."""
    kn = o.get(po.QOR_KNOTS)  # knots -1..N+p+1
    nb = N + 2                 # including the two dropped end splines
    basis = [BSpline(kn, np.eye(nb)[i], p, extrapolate=False) for i in range(1, N + 1)]
    dbasis = [b.derivative() for b in basis]
    xg, wg = np.polynomial.legendre.leggauss(2 * p + 2)
    pts, wts = [], []
    for a, b in zip(kn[:-1], kn[1:]):
        if b > a:
            pts.append(0.5 * (a + b) + 0.5 * (b - a) * xg)
            wts.append(0.5 * (b - a) * wg)
    x = np.concatenate(pts)
    w = np.concatenate(wts)
    B = np.nan_to_num(np.array([b(x) for b in basis]))
    dB = np.nan_to_num(np.array([b(x) for b in dbasis]))

    def band(f, mi, mj):
        A = (dB if mi else B) * (w * f(x))
        M = A @ (dB if mj else B).T
        out = np.zeros((N, p + 1))
        for i in range(N):
            for j in range(p + 1):
                if i + j < N:
                    out[i, j] = M[i, i + j]
        return out.ravel()
    return band


def test_radial_bands_match_scipy(small):
    o = small
    band = _scipy_bands(o, N_SMALL, P)
    one = lambda r: np.ones_like(r)
    checks = {
        po.QOR_S: band(one, 0, 0),
        po.QOR_RINV: band(lambda r: 1.0 / (r + 1e-50), 0, 0),
        po.QOR_RDIF: band(one, 0, 1),
        po.QOR_RPOW1: band(lambda r: r, 0, 0),
    }
    for what, ref in checks.items():
        got = o.get(what)
        assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max()), what
    # H_l = <V> + l(l+1)/2 <1/r^2> + 1/2 <B'|B'>, hydrogen V = -1/r
    H = o.get(po.QOR_H).reshape(-1, N_SMALL * (P + 1))
    for l in (0, 1, 3):
        ref = band(lambda r: -1.0 / (r + 1e-50), 0, 0) + 0.5 * l * (l + 1) * band(lambda r: 1.0 / (r * r + 1e-50), 0, 0) \
            + 0.5 * band(one, 1, 1)
        assert np.abs(H[l] - ref).max() <= 1e-12 * np.abs(ref).max()
    # d/dr is antisymmetric (dropped end splines: no boundary term), so the
    # stored upper band determines the whole operator: <Bi|Bj'> = -<Bj|Bi'>
    assert np.abs(o.get(po.QOR_RDIF).reshape(N_SMALL, P + 1)[:, 0]).max() < 1e-14
    assert o.get(po.QOR_KNOTS).size == N_SMALL + P + 3


def test_hydrogen_levels_closed_form():
    o = po.Oracle(N=1000, p=7, L=3, rmax=200.0, dt=0.05)
    for l in range(3):
        for k in range(3):
            n = l + 1 + k
            e, c = o.eigenstate(0, l, k)
            assert abs(e - (-0.5 / n ** 2)) < 1e-10, (l, k, e)
            assert abs(o.norm(c) - 1.0) < 1e-13


def test_lu_zsb_reconstructs_band():
    n, p = 40, 7
    rng = np.random.default_rng(3)
    A = np.zeros((n, n), complex)
    for i in range(n):
        for j in range(i, min(n, i + p + 1)):
            A[i, j] = A[j, i] = rng.standard_normal() + 1j * rng.standard_normal()
        A[i, i] += 12.0 + 3j
    band = np.zeros((n, p + 1), complex)
    for i in range(n):
        for j in range(p + 1):
            if i + j < n:
                band[i, j] = A[i, i + j]
    f, info = po.lu_zsb(band.ravel(), n, p)
    assert info == 0
    f = f.reshape(n, p + 1)
    Lm = np.eye(n, dtype=complex)
    D = np.diag(f[:, 0])
    for i in range(n):
        for j in range(1, p + 1):
            if i + j < n:
                Lm[i + j, i] = f[i, j]
    assert np.abs(Lm @ D @ Lm.T - A).max() < 1e-12 * np.abs(A).max()  # transpose, not conjugate


def test_lu_zsb_counts_tiny_pivots():
    n, p = 20, 7
    band = np.zeros((n, p + 1), complex)
    band[:, 0] = 1.0
    band[5, 0] = 0.0  # singular pivot
    _, info = po.lu_zsb(band.ravel(), n, p)
    assert info >= 1


def test_cholesky_of_overlap(small):
    o = small
    S = po.band_to_dense(o.get(po.QOR_S), N_SMALL, P)
    ch = o.get(po.QOR_CHOL).reshape(N_SMALL, P + 1)
    Lc = np.zeros((N_SMALL, N_SMALL))
    for i in range(N_SMALL):
        for j in range(P + 1):
            if i + j < N_SMALL:
                Lc[i + j, i] = ch[i, j]
    assert np.abs(Lc @ Lc.T - S).max() < 1e-14


@pytest.mark.parametrize("symflag", [0, 1])
def test_cn_kernel_matches_dense_solve(small, symflag):
    o = small
    n = N_SMALL
    S = po.band_to_dense(o.get(po.QOR_S), n, P)
    if symflag == 0:
        R = po.band_to_dense(o.get(po.QOR_RINV), n, P, 1)
        ifet = 0.3j
    else:
        R = po.band_to_dense(o.get(po.QOR_RDIF), n, P, -1)
        ifet = 0.013
    rng = np.random.default_rng(5)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    got = po.cn_kernel(n, P, symflag, o.get(po.QOR_S), o.get(po.QOR_RINV if symflag == 0 else po.QOR_RDIF), ifet, z)
    if symflag == 0:
        ref = np.linalg.solve(S + ifet * R, (S - ifet * R) @ z)
    else:
        # the antisymmetric sweep solves (S + a D) x = (S - a D) z with D = <Bi|Bj'> (propagate.hpp:47-49)
        ref = np.linalg.solve(S + ifet * R, (S - ifet * R) @ z)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


def test_fieldfree_ldl_solve_matches_dense(small):
    o = small
    n = N_SMALL
    up = o.get(po.QOR_UPPER).reshape(-1, n * (P + 1))
    lo = o.get(po.QOR_LOWER).reshape(-1, n * (P + 1))
    S = po.band_to_dense(o.get(po.QOR_S), n, P)
    H = po.band_to_dense(o.get(po.QOR_H).reshape(-1, n * (P + 1))[2], n, P)
    rng = np.random.default_rng(2)
    c = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    w = po.symb_mul(n, P, up[2], c)
    assert np.abs(w - (S - 0.5j * 0.05 * H) @ c).max() < 1e-13
    x = po.ldl_solve(n, P, lo[2], w)
    ref = np.linalg.solve(S + 0.5j * 0.05 * H, w)
    assert np.abs(x - ref).max() < 1e-12 * np.abs(ref).max()


def test_tridiagonal_eigen_and_golub_welsch(small):
    # Y10 at m = 0: Jacobi matrix of Legendre polynomials -> eigenvalues are
    # the Gauss-Legendre nodes of order L (Golub-Welsch)
    for L in (5, 31, 81):
        o = po.Oracle(N=40, p=7, L=L, rmax=10.0)
        x, _ = np.polynomial.legendre.leggauss(L)
        assert np.abs(np.sort(o.get(po.QOR_YEIG)) - x).max() < 5e-15 * L
    rng = np.random.default_rng(9)
    d, e = rng.standard_normal(12), rng.standard_normal(11)
    ev, z = po.tridiag_eig(d, e)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    assert np.abs(ev - np.linalg.eigvalsh(T)).max() < 1e-13
    assert np.abs(z @ T @ z.T - np.diag(ev)).max() < 1e-13


def test_angular_projectors_reproduce_operators():
    L = 9
    o = po.Oracle(N=40, p=7, L=L, rmax=10.0)
    y10 = o.get(po.QOR_Y10)
    q10 = o.get(po.QOR_Q10)
    l = np.arange(L)
    assert np.allclose(y10, np.sqrt(((l + 1) ** 2) / ((2 * l + 1) * (2 * l + 3))), rtol=0, atol=1e-15)
    assert np.allclose(q10, y10 * (l + 1), rtol=0, atol=1e-14)
    Y = np.diag(y10[:-1], 1) + np.diag(y10[:-1], -1)
    Py = o.get(po.QOR_YPRJ).reshape(L, L)
    assert np.abs(Py.T @ np.diag(o.get(po.QOR_YEIG)) @ Py - Y).max() < 1e-14
    # Hermitian i*Q10 (lower entries -i q_l, operators/1e.hpp:436-441)
    HQ = np.diag(-1j * q10[:-1], -1) + np.diag(1j * q10[:-1], 1)
    Pq = o.get(po.QOR_QPRJ).reshape(L, L)
    assert np.abs(Pq.conj().T @ np.diag(o.get(po.QOR_QEIG)) @ Pq - HQ).max() < 1e-13
    assert np.abs(Pq @ Pq.conj().T - np.eye(L)).max() < 1e-14
    M = o.get(po.QOR_MERGED).reshape(L, L)
    assert np.abs(M - Pq @ Py.conj().T).max() < 1e-14


def test_fieldfree_cayley_phase():
    o = po.Oracle(N=400, p=7, L=2, rmax=60.0, dt=0.05)
    for (il, k) in ((0, 0), (1, 1)):
        e, c0 = o.eigenstate(0, il, k)
        c = c0.copy()
        for _ in range(20):
            o.step_free(c)
        ph = ((1 - 0.5j * 0.05 * e) / (1 + 0.5j * 0.05 * e)) ** 20
        assert np.abs(c - ph * c0).max() < 1e-12


def test_time_reversal_and_unitarity():
    fwd = po.Oracle(N=200, p=7, L=5, rmax=40.0, dt=0.05)
    bwd = po.Oracle(N=200, p=7, L=5, rmax=40.0, dt=-0.05)
    rng = np.random.default_rng(4)
    c0 = rng.standard_normal(5 * 200) + 1j * rng.standard_normal(5 * 200)
    c0 /= np.sqrt(fwd.norm(c0))
    c = c0.copy()
    fwd.step(c, 0.7, 0.9)
    assert abs(fwd.norm(c) - 1.0) < 1e-13  # S-unitarity of every factor
    bwd.step(c, 0.9, 0.7)  # the Strang sequence is symmetric: reversed fields undo it
    assert np.abs(c - c0).max() < 1e-11


def _dense_step(o, N, L, dt, a_i, a_f):
    """Dense numpy brute force of one vgwz step from closed-form angular
    matrices and the radial bands (no eigen-decomposition, no banded solver)."""
    n = N
    S = po.band_to_dense(o.get(po.QOR_S), n, P)
    R = po.band_to_dense(o.get(po.QOR_RINV), n, P, 1)
    D = po.band_to_dense(o.get(po.QOR_RDIF), n, P, -1)
    Hl = o.get(po.QOR_H).reshape(-1, n * (P + 1))
    l = np.arange(L)
    y = np.sqrt(((l + 1) ** 2) / ((2 * l + 1) * (2 * l + 3)))
    Y = np.diag(y[:-1], 1) + np.diag(y[:-1], -1)
    HQ = np.diag(-1j * (y * (l + 1))[:-1], -1) + np.diag(1j * (y * (l + 1))[:-1], 1)
    I_L = np.eye(L)
    S2 = np.kron(I_L, S)

    def cayley(Kp):
        return np.linalg.solve(S2 + Kp, S2 - Kp)

    # f = -A; D-part a = -f dt/4 Y@D (real), R-part i f dt/4 HQ@R (propagate.hpp:343,400)
    UD = lambda A: cayley((A * dt / 4) * np.kron(Y, D))
    UR = lambda A: cayley(1j * (-A) * dt / 4 * np.kron(HQ, R))
    U0 = np.zeros((L * n, L * n), complex)
    for li in range(L):
        H = po.band_to_dense(Hl[li], n, P)
        U0[li * n:(li + 1) * n, li * n:(li + 1) * n] = np.linalg.solve(S + 0.5j * dt * H, S - 0.5j * dt * H)
    return UD(a_f) @ UR(a_f) @ U0 @ UR(a_i) @ UD(a_i)


def test_vgwz_step_matches_dense_strang_product():
    N, L, dt = 24, 4, 0.05
    o = po.Oracle(N=N, p=P, L=L, rmax=10.0, dt=dt)
    U = _dense_step(o, N, L, dt, 0.8, 1.1)
    rng = np.random.default_rng(8)
    c0 = rng.standard_normal(L * N) + 1j * rng.standard_normal(L * N)
    c = c0.copy()
    o.step(c, 0.8, 1.1)
    ref = U @ c0
    assert np.abs(c - ref).max() < 1e-12 * np.abs(ref).max()


def test_mask_identity_and_dense(small):
    o = po.Oracle(N=N_SMALL, p=P, L=2, rmax=R_SMALL, dt=0.05, mask_rmin=9.0, mask_rmax=12.0, mask_fact=0.125)
    rng = np.random.default_rng(1)
    c0 = rng.standard_normal(2 * N_SMALL) + 1j * rng.standard_normal(2 * N_SMALL)
    # S^-1 M c against dense numpy
    S = po.band_to_dense(o.get(po.QOR_S), N_SMALL, P)
    Mk = po.band_to_dense(o.get(po.QOR_MASK), N_SMALL, P)
    c = c0.copy()
    o.mask(c)
    ref = np.concatenate([np.linalg.solve(S, Mk @ c0[:N_SMALL]), np.linalg.solve(S, Mk @ c0[N_SMALL:])])
    assert np.abs(c - ref).max() < 1e-12 * np.abs(ref).max()
    # mask function 1 everywhere -> M = S -> identity
    o.set(po.QOR_MASK, o.get(po.QOR_S))
    c = c0.copy()
    o.mask(c)
    assert np.abs(c - c0).max() < 1e-12


def test_laser_pulse_and_time_grid():
    # field_component_sinen<2>: A(t) = F0 sin(w t + pi c0) sin^2(pi t / tau) on (0, tau)
    e0, w0, nc = 1.32, 0.057, 8.0
    tau = nc * 2 * np.pi / w0
    for t in (0.0, 13.3, tau / 2, tau - 1e-3, tau + 1.0):
        ref = e0 * np.sin(w0 * t + np.pi * 0.25) * (np.sin(np.pi * t / tau) ** 2 if 0 < t < tau else 0.0)
        assert abs(po.sinen2(e0, w0, nc, 0.0, 0.25, t) - ref) < 1e-13
    assert po.para_nt(0.057, 8.0, 0.0, 0.05) == 17638
    assert po.para_nt(0.057, 4.0, 0.0, 0.05) == 8819


def test_observables_norm_and_parity():
    o = po.Oracle(N=300, p=7, L=3, rmax=60.0, dt=0.05)
    e1, c1 = o.eigenstate(0, 0, 0)
    e2, c2 = o.eigenstate(0, 1, 0)
    assert abs(o.norm(c1) - 1) < 1e-13
    assert abs(o.dipole_z(c1)) < 1e-15  # single-l state has no dipole
    c = (c1 + c2) / np.sqrt(2)
    # <1s|z|2p0> = 2^7.5/3^5 for hydrogen (closed form)
    ref = 2 ** 7.5 / 3 ** 5
    assert abs(o.dipole_z(c) - ref) < 1e-8


def test_projection_of_eigenstates_is_kronecker():
    """eigenstate_sph projection (population_1e.hpp:29-55) of the basis's own
    eigenstates: <S v_n,l | v_n0,l0> = delta (v^T S v = 1), with the hydrogen
    levels -1/(2 n^2) on the diagonal of the basis energies; and a dense
    numpy contraction of a random state."""
    o = po.Oracle(N=300, p=7, L=4, M=2, m0=0, rmax=60.0)
    nsel = 5
    W, E = o.projection_basis(nsel)
    for l in range(2):  # principal quantum numbers up to 3 fit the 60 a.u. box
        np.testing.assert_allclose(E[l, :2], [-0.5 / (n + l + 1) ** 2 for n in range(2)], rtol=1e-7)
    for (im, il, n0) in [(0, 0, 0), (0, 2, 3), (1, 1, 1)]:
        _, c = o.eigenstate(im, il, n0)
        p = o.projection(c, W).reshape(o.M, o.L, nsel)
        want = np.zeros_like(p)
        want[im, il, n0] = 1.0
        np.testing.assert_allclose(p, want, atol=1e-10)
    rng = np.random.default_rng(4)
    c = rng.standard_normal(o.M * o.L * o.N) + 1j * rng.standard_normal(o.M * o.L * o.N)
    p = o.projection(c, W, stride=o.N).reshape(o.M, o.L, o.N)
    for im in range(o.M):
        for il in range(o.L):
            l = il + abs(im + o.m0)
            dense = W[l] @ c.reshape(o.M, o.L, o.N)[im, il]
            np.testing.assert_allclose(p[im, il, :nsel], dense, rtol=1e-12, atol=1e-12)
            assert np.all(p[im, il, nsel:] == 0)


def _dense_band(band, N, ne, sign=1.0):
    """Dense matrix of a band stored upper (A(i,i+k) = band[i*ne+k]); the lower
    triangle is sign * the transposed upper (sign -1: antisymmetric)."""
    A = np.zeros((N, N))
    for i in range(N):
        for k in range(ne):
            if i + k < N:
                A[i, i + k] = band[i * ne + k]
                if k:
                    A[i + k, i] = sign * band[i * ne + k]
    return A


def test_krylov_step_matches_dense_exponential():
    """propagator_arnoldi::propagate1 (la_1e.hpp:312-335) with the vgwz
    multiply (:480-497): the Lanczos step equals exp(-i dt S^-1 H(az)) built
    densely from a generalized Hermitian eigen-decomposition (scipy eigh),
    and stops early (m = 1) on a field-free eigenstate, giving the exact
    phase exp(-i E dt)."""
    import scipy.linalg as sl
    N, L = 60, 3
    o = po.Oracle(N=N, p=7, L=L, rmax=16.0)
    ne = 8
    Hb, S = o.get(po.QOR_H).reshape(-1, N * ne), _dense_band(o.get(po.QOR_S), N, ne)
    F = _dense_band(o.get(po.QOR_RINV), N, ne)
    G = _dense_band(o.get(po.QOR_RDIF), N, ne, sign=-1.0)
    y, q = o.get(po.QOR_Y10), o.get(po.QOR_Q10)
    az, dt = 0.7, 0.05
    Hd = np.zeros((L * N, L * N), complex)
    Sd = np.zeros((L * N, L * N))
    for il in range(L):
        sl_ = slice(il * N, (il + 1) * N)
        Hd[sl_, sl_] = _dense_band(Hb[il], N, ne)
        Sd[sl_, sl_] = S
        if il > 0:
            Hd[sl_, (il - 1) * N:il * N] = 1j * az * q[il - 1] * F - 1j * az * y[il - 1] * G
        if il < L - 1:
            Hd[sl_, (il + 1) * N:(il + 2) * N] = -1j * az * y[il] * G - 1j * az * q[il] * F
    assert np.abs(Hd - Hd.conj().T).max() < 1e-12  # H(az) is Hermitian
    lam, V = sl.eigh(Hd, Sd)
    rng = np.random.default_rng(2)
    c = rng.standard_normal(L * N) + 1j * rng.standard_normal(L * N)
    c /= np.sqrt(np.real(c.conj() @ Sd @ c))
    exact = V @ (np.exp(-1j * dt * lam) * (V.conj().T @ (Sd @ c)))
    g = c.copy()
    crit, m, eps, _ = o.krylov_step(g, az, dt, kmax=60, emin=1e-26, emax=1e-20)
    assert np.linalg.norm(g - exact) / np.linalg.norm(exact) < 1e-11, m
    assert eps < 1e-20 and crit == 2
    # too small a subspace is visibly wrong, and the estimate says so
    g2 = c.copy()
    crit2, m2, eps2, _ = o.krylov_step(g2, az, dt, kmax=3, emin=1e-14, emax=1e-10)
    assert m2 == 3 and crit2 == 0 and eps2 > 1e-10
    assert np.linalg.norm(g2 - exact) / np.linalg.norm(exact) > 1e-8
    # field-free eigenstate: one Krylov vector, exact phase
    e, v = o.eigenstate(0, 1, 0)
    g3 = v.copy()
    _, m3, _, _ = o.krylov_step(g3, 0.0, dt, kmax=20)
    assert m3 == 1
    np.testing.assert_allclose(g3, np.exp(-1j * e * dt) * v, atol=1e-12)


def test_outer_norm_closed_form_and_quadrature(small):
    """Windowed norm sum_ml <c|S_w|c>, S_w = <Bi|theta(r - r0)|Bj> (the
    overlap of observable_molec with rmax beyond the box, observable/1e.hpp:
    188-192): the band against scipy B-splines with the step function, the
    hydrogen 1s closed form int_r0^inf 4 r^2 exp(-2r) dr =
    exp(-2 r0)(2 r0^2 + 2 r0 + 1) for r0 on a knot, the full norm for r0 at
    the origin and zero beyond the box."""
    o = small
    band = _scipy_bands(o, N_SMALL, P)
    kn = o.get(po.QOR_KNOTS)
    rng = np.random.default_rng(11)
    c = rng.standard_normal(L_SMALL * N_SMALL) + 1j * rng.standard_normal(L_SMALL * N_SMALL)
    for r0 in (kn[P + 5], kn[P + 17], 0.5 * (kn[P + 9] + kn[P + 10])):
        Sw = po.band_to_dense(band(lambda r, r0=r0: (r >= r0).astype(float), 0, 0), N_SMALL, P)
        ref = sum(np.real(v.conj() @ Sw @ v) for v in c.reshape(L_SMALL, N_SMALL))
        assert abs(o.outer_norm(c, r0) - ref) <= 1e-13 * abs(ref), r0
    assert abs(o.outer_norm(c, -1.0) - o.norm(c)) <= 1e-13 * o.norm(c)
    assert o.outer_norm(c, R_SMALL + 1.0) == 0.0
    h = po.Oracle(N=300, p=7, L=3, rmax=60.0, dt=0.05)
    _, s1 = h.eigenstate(0, 0, 0)
    kn = h.get(po.QOR_KNOTS)
    for k in (P + 2, P + 5, P + 11):  # r0 = 0.20, 0.81, 2.03
        r0 = kn[k]
        assert abs(h.outer_norm(s1, r0) - np.exp(-2 * r0) * (2 * r0 ** 2 + 2 * r0 + 1)) < 1e-11, r0


def test_acceleration_band_closed_form_and_ehrenfest(small):
    """Dipole acceleration of observable_z with f(r) = dV/dr = Z/r^2
    (observable/1e.hpp:26-49 with the config's func_obsv_z): the band
    against scipy B-splines, <1s|cos(theta)/r^2|2p0> = 4/(27 sqrt 2) for
    hydrogen (closed form of the radial integral (1/sqrt 6) int r e^{-3r/2}
    dr times <Y00|cos|Y10> = 1/sqrt 3), and Ehrenfest's relation
    d^2<z>/dt^2 = -<dV/dz>: for the (1s + 2p0)/sqrt 2 beat,
    dE^2 <z> = <cos/r^2> with dE = 3/8."""
    o = small
    band = _scipy_bands(o, N_SMALL, P)
    ref = band(lambda r: 1.0 / (r * r + 1e-50), 0, 0)
    got = o.get(po.QOR_DVDR)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    h = po.Oracle(N=300, p=7, L=3, rmax=60.0, dt=0.05)
    e1, s1 = h.eigenstate(0, 0, 0)
    e2, s2 = h.eigenstate(0, 1, 0)
    c = (s1 + s2) / np.sqrt(2)
    acc = h.dipole_z(c, 1)
    assert abs(acc - 4.0 / (27.0 * np.sqrt(2.0))) < 1e-10
    assert abs((e2 - e1) ** 2 * h.dipole_z(c, 0) - acc) < 1e-10
    assert abs(h.dipole_z(s1, 1)) < 1e-15  # one l: no coupling


def _dense_step_lg(o, N, L, dt, e_i, e_f):
    """Dense numpy brute force of one lgwz step (cn_1e.hpp:598-613): the
    Cayley factors of E z = E r cos(theta) (Y10 closed form, <Bi|r|Bj> band)
    around the field-free factor, no eigen-decomposition, no banded solver."""
    n = N
    S = po.band_to_dense(o.get(po.QOR_S), n, P)
    Rr = po.band_to_dense(o.get(po.QOR_RPOW1), n, P, 1)
    Hl = o.get(po.QOR_H).reshape(-1, n * (P + 1))
    l = np.arange(L)
    y = np.sqrt(((l + 1) ** 2) / ((2 * l + 1) * (2 * l + 3)))
    Y = np.diag(y[:-1], 1) + np.diag(y[:-1], -1)
    S2 = np.kron(np.eye(L), S)

    def cayley(Kp):
        return np.linalg.solve(S2 + Kp, S2 - Kp)

    UL = lambda E: cayley(1j * E * dt / 4 * np.kron(Y, Rr))
    U0 = np.zeros((L * n, L * n), complex)
    for li in range(L):
        H = po.band_to_dense(Hl[li], n, P)
        U0[li * n:(li + 1) * n, li * n:(li + 1) * n] = np.linalg.solve(S + 0.5j * dt * H, S - 0.5j * dt * H)
    return UL(e_f) @ U0 @ UL(e_i)


def test_lgwz_step_matches_dense_strang_product():
    """propagator_cn_hc_lgwz against the dense product of Cayley factors of
    the length-gauge Hamiltonian H0 + E z; a planted sign error in the
    coupling (E -> -E) fails it."""
    N, L, dt = 24, 5, 0.05
    o = po.Oracle(N=N, p=P, L=L, rmax=10.0, dt=dt)
    U = _dense_step_lg(o, N, L, dt, 0.3, 0.45)
    rng = np.random.default_rng(12)
    c0 = rng.standard_normal(L * N) + 1j * rng.standard_normal(L * N)
    c = c0.copy()
    o.step_lg(c, 0.3, 0.45)
    ref = U @ c0
    assert np.abs(c - ref).max() < 1e-12 * np.abs(ref).max()
    wrong = _dense_step_lg(o, N, L, dt, -0.3, -0.45) @ c0
    assert np.abs(wrong - ref).max() > 1e-4 * np.abs(ref).max()


def test_lgwz_unitarity_time_reversal_and_weak_field_gauge_agreement():
    """S-unitarity of the lgwz step, a -dt step with reversed fields undoing
    it, and gauge invariance: over a weak 2-cycle pulse the length-gauge run
    (E = -dA/dt at the time points) and the velocity-gauge run (A) leave the
    same 1s population to the CN time-discretisation error."""
    fwd = po.Oracle(N=200, p=7, L=5, rmax=40.0, dt=0.05)
    bwd = po.Oracle(N=200, p=7, L=5, rmax=40.0, dt=-0.05)
    rng = np.random.default_rng(6)
    c0 = rng.standard_normal(5 * 200) + 1j * rng.standard_normal(5 * 200)
    c0 /= np.sqrt(fwd.norm(c0))
    c = c0.copy()
    fwd.step_lg(c, 0.05, 0.07)
    assert abs(fwd.norm(c) - 1.0) < 1e-13
    bwd.step_lg(c, 0.07, 0.05)
    assert np.abs(c - c0).max() < 1e-11
    o = po.Oracle(N=300, p=7, L=6, rmax=60.0, dt=0.02, mask_rmin=55.0, mask_rmax=60.0)
    _, g = o.eigenstate(0, 0, 0)
    w0, nc, e0 = 0.5, 2.0, 0.02
    tau = nc * 2 * np.pi / w0
    nt = int(tau / 0.02) + 1
    t = 0.02 * np.arange(nt + 1)
    A = np.array([po.sinen2(e0 / w0, w0, nc, 0.0, 0.0, x) for x in t])
    E = -np.gradient(A, t, edge_order=2)
    E[0] = E[-1] = 0.0
    cv, cl = g.copy(), g.copy()
    o.run(cv, A, nt, 0, nt, use_mask=False)
    o.run_lg(cl, E, nt, 0, nt, use_mask=False)
    pv, pl = abs(np.vdot(po.band_to_dense(o.get(po.QOR_S), 300, 7) @ g[:300], cv[:300])) ** 2, \
        abs(np.vdot(po.band_to_dense(o.get(po.QOR_S), 300, 7) @ g[:300], cl[:300])) ** 2
    assert 1 - pv > 1e-5  # the pulse does something
    assert abs(pv - pl) < 2e-3 * (1 - pv)


def test_dipole_xy_closed_forms():
    """observable_r (observable/1e.hpp:53-129): with the Condon-Shortley
    phase, p_x = (Y_1,-1 - Y_1,1)/sqrt 2 and p_y = i (Y_1,-1 + Y_1,1)/sqrt 2,
    so for hydrogen <x> of (1s + 2p_x)/sqrt 2 and <y> of (1s + 2p_y)/sqrt 2
    are <1s|z|2p0> = 2^7.5/3^5, with zero in the other two directions; the
    acceleration form (f = dV/dr) gives 4/(27 sqrt 2) alike."""
    o = po.Oracle(N=300, p=7, L=3, M=3, m0=-1, rmax=60.0, dt=0.05)
    _, s1 = o.eigenstate(1, 0, 0)   # m = 0, l = 0
    _, pm = o.eigenstate(0, 0, 0)   # m = -1, l = 1
    _, pp = o.eigenstate(2, 0, 0)   # m = +1, l = 1
    _, pz = o.eigenstate(1, 1, 0)   # m = 0, l = 1
    ref, acc = 2 ** 7.5 / 3 ** 5, 4.0 / (27.0 * np.sqrt(2.0))
    for st, want in (((pm - pp) / np.sqrt(2), (1, 0, 0)), (1j * (pm + pp) / np.sqrt(2), (0, 1, 0)), (pz, (0, 0, 1))):
        c = (s1 + st) / np.sqrt(2)
        x, y = o.dipole_xy(c, 0)
        assert np.allclose([x, y, o.dipole_z(c, 0)], np.array(want) * ref, atol=1e-9)
        ax, ay = o.dipole_xy(c, 1)
        assert np.allclose([ax, ay, o.dipole_z(c, 1)], np.array(want) * acc, atol=1e-9)
