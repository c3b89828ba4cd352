"""Pins of the elliptically polarised (m,l) propagator propagator_cn_hc_vgwr
(cn_1e.hpp:812-951) in the CPU oracle, against references that do not share
its code: the angular closed forms against sphere quadrature of scipy's
spherical harmonics and against the L^2 commutator identity, the eigen-data
against its defining relations, and the split step against a dense numpy
product of Cayley factors built without any eigen-decomposition."""
import numpy as np
import pytest
import scipy.special as ss

import pyoracle as po

P = 7


def _sphere_element(l1, m1, l2, m2, f, n=48):
    """<Y_{l1,m1}| f(theta, phi) |Y_{l2,m2}> by Gauss-Legendre in cos(theta) and
    the trapezoid rule in phi (exact for these band-limited integrands)."""
    x, w = np.polynomial.legendre.leggauss(n)
    th = np.arccos(x)
    ph = 2 * np.pi * np.arange(2 * n) / (2 * n)
    T, F = np.meshgrid(th, ph, indexing="ij")
    y1 = ss.sph_harm_y(l1, m1, T, F)
    y2 = ss.sph_harm_y(l2, m2, T, F)
    g = np.conj(y1) * f(T, F) * y2
    return np.sum(w[:, None] * g) * (2 * np.pi / (2 * n))


def test_y11_closed_forms_match_sphere_quadrature():
    """integrate_sinexp1/2 (basisfunc.hpp:243-262) = <l,m| sin(th) e^{-i ph} |l+-1, m+1>
    with the Condon-Shortley phase of the reference's Y_lm (basisfunc.hpp:200-207)."""
    f = lambda t, p: np.sin(t) * np.exp(-1j * p)
    for m in range(-4, 4):
        for l in range(abs(m), 7):
            se1, se2, _, _ = po.ep_closed_forms(m, l)
            assert abs(_sphere_element(l, m, l + 1, m + 1, f) - se1) < 1e-13
            if l - 1 >= abs(m + 1):
                assert abs(_sphere_element(l, m, l - 1, m + 1, f) - se2) < 1e-13


def test_q11_closed_forms_follow_commutator_identity():
    """-Q1-1 = -cos(th) L_- - sin(th) e^{-i ph} (1 - L_z) = grad f . grad - f for
    f = sin(th) e^{-i ph} (basisfunc.hpp:283-302).  With [L^2, f] = 2 f - 2 grad f . grad
    (f is an l = 1 harmonic) its elements are f_12 (l2(l2+1) - l1(l1+1)) / 2:
    outer (l2 = l+1) f_12 (l+1), inner (l2 = l-1) -f_12 l."""
    for m in range(-5, 5):
        for l in range(abs(m), 9):
            se1, se2, cl1, cl2 = po.ep_closed_forms(m, l)
            assert abs(cl1 - se1 * (l + 1)) < 1e-14
            assert abs(cl2 + se2 * l) < 1e-14


@pytest.mark.parametrize("M,L,m0", [(3, 4, -1), (4, 3, 0), (5, 5, -2), (2, 6, 1)])
def test_ep_eigen_data(M, L, m0):
    """zhbev + catagorize restated: P^H diag(eig) P reproduces each angular
    matrix, P is unitary, each eigenvector lives on one parity of il, the
    categories have the sizes of dimension_1e (dimension/1e.hpp:52-55), and
    the transforms are the products named in cn_1e.hpp:890-895."""
    ep = po.ep_angular(M, L, m0)
    ML = M * L
    par = np.array([il % 2 for im in range(M) for il in range(L)])
    n_zero = L if M % 2 else 0
    n_even = (M // 2 * 2) * ((L + 1) // 2)
    for k in ("x1", "x2", "y1", "y2"):
        A, e, Pk = ep["mat"][k], ep["eig_" + k], ep["P"][k]
        assert np.abs(A - A.conj().T).max() == 0.0
        assert np.abs(Pk.conj().T @ np.diag(e) @ Pk - A).max() < 1e-13
        assert np.abs(Pk @ Pk.conj().T - np.eye(ML)).max() < 1e-13
        for i in range(ML):
            sup = np.abs(Pk[i]) > 0
            assert len(set(par[sup])) == 1
            if i < n_even:
                assert set(par[sup]) == {0}
            elif i < n_even + n_zero:
                assert abs(e[i]) < 1e-8
            else:
                assert set(par[sup]) == {1}
    assert np.abs(ep["tran_x1"] - ep["P"]["x1"].real).max() == 0.0
    assert np.abs(ep["P"]["x1"].imag).max() == 0.0 and np.abs(ep["P"]["y2"].imag).max() == 0.0
    assert np.abs(ep["x1_2_x2"] - ep["P"]["x2"] @ ep["P"]["x1"].conj().T).max() < 1e-15


def _dense_vgwr(o, ep, N, ML, dt, ax_i, ax_f, ay_i, ay_f):
    """Dense brute force of one vgwr step: the Strang sequence of Cayley
    factors (S + K)^-1 (S - K) with K = ifdt (Z (x) X) for the angular
    matrices Z themselves (no eigen-decomposition, no banded solver).  The
    scalars follow propagate<isub> (cn_1e.hpp:898-945): A/2, f = -ax, +ax, -ay,
    -ay for x1, x2, y1, y2, step dt/2; ifdt = -f step / 2 for the d/dr sweeps
    and i f step / 2 for the 1/r sweeps (propagate.hpp:463, 525)."""
    S = po.band_to_dense(o.get(po.QOR_S), N, P)
    R = po.band_to_dense(o.get(po.QOR_RINV), N, P, 1)
    D = po.band_to_dense(o.get(po.QOR_RDIF), N, P, -1)
    Hl = o.get(po.QOR_H).reshape(-1, N * (P + 1))
    S2 = np.kron(np.eye(ML), S)
    step = dt / 2
    Z = ep["mat"]

    def cayley(K):
        return np.linalg.solve(S2 + K, S2 - K)

    UD = lambda Zk, f: cayley((-(f * step) / 2) * np.kron(Zk, D))
    UR = lambda Zk, f: cayley((1j * (f * step) / 2) * np.kron(Zk, R))
    U0 = np.zeros((ML * N, ML * N), complex)
    L, M, m0 = o.L, o.M, o.m0
    for im in range(M):
        for il in range(L):
            ch = im * L + il
            H = po.band_to_dense(Hl[il + abs(m0 + im)], N, P)
            U0[ch * N:(ch + 1) * N, ch * N:(ch + 1) * N] = np.linalg.solve(S + 0.5j * dt * H, S - 0.5j * dt * H)
    ax_i, ax_f, ay_i, ay_f = ax_i / 2, ax_f / 2, ay_i / 2, ay_f / 2
    first = UR(Z["y2"], -ay_i) @ UD(Z["y1"], -ay_i) @ UR(Z["x2"], ax_i) @ UD(Z["x1"], -ax_i)
    second = UD(Z["x1"], -ax_f) @ UR(Z["x2"], ax_f) @ UD(Z["y1"], -ay_f) @ UR(Z["y2"], -ay_f)
    return second @ U0 @ first


@pytest.mark.parametrize("M,L,m0", [(3, 3, -1), (2, 4, 0)])
def test_vgwr_step_matches_dense_strang_product(M, L, m0):
    N, dt = 24, 0.05
    o = po.Oracle(N=N, p=P, L=L, M=M, m0=m0, rmax=10.0, dt=dt)
    ep = po.ep_angular(M, L, m0)
    ML = M * L
    U = _dense_vgwr(o, ep, N, ML, dt, 0.8, 1.1, -0.6, 0.4)
    rng = np.random.default_rng(M * 10 + L)
    c0 = rng.standard_normal(ML * N) + 1j * rng.standard_normal(ML * N)
    ref = U @ c0
    # all channels (isub = -1)
    o.set_ep(ep, isub=-1)
    c = c0.copy()
    o.step_ep(c, 0.8, 1.1, -0.6, 0.4)
    assert np.abs(c - ref).max() < 1e-12 * np.abs(ref).max()
    # gathered even / odd sub-blocks on states of one parity of il (propflag_eve / _odd)
    par = np.repeat(np.array([il % 2 for im in range(M) for il in range(L)]), N)
    for isub, pty in ((2, 0), (3, 1)):
        cp = np.where(par == pty, c0, 0.0)
        o.set_ep(ep, isub=isub)
        c = cp.copy()
        o.step_ep(c, 0.8, 1.1, -0.6, 0.4)
        refp = U @ cp
        assert np.abs(refp[par != pty]).max() < 1e-13 * np.abs(refp).max()
        assert np.abs(c - refp).max() < 1e-12 * np.abs(refp).max(), isub


def test_vgwr_step_is_unitary_and_reversible():
    """S-unitarity of every factor, and the symmetric Strang sequence undone by
    the step with -dt and the fields swapped."""
    M, L, m0, N = 3, 4, -1, 60
    fwd = po.Oracle(N=N, p=P, L=L, M=M, m0=m0, rmax=20.0, dt=0.05)
    bwd = po.Oracle(N=N, p=P, L=L, M=M, m0=m0, rmax=20.0, dt=-0.05)
    ep = po.ep_angular(M, L, m0)
    fwd.set_ep(ep)
    bwd.set_ep(ep)
    rng = np.random.default_rng(9)
    c0 = rng.standard_normal(M * L * N) + 1j * rng.standard_normal(M * L * N)
    c0 /= np.sqrt(fwd.norm(c0))
    c = c0.copy()
    fwd.step_ep(c, 0.7, 0.9, -0.3, 0.2)
    assert abs(fwd.norm(c) - 1.0) < 1e-13
    bwd.step_ep(c, 0.9, 0.7, 0.2, -0.3)
    assert np.abs(c - c0).max() < 1e-11


def _dense_lgwr(o, lg, N, ML, dt, ex_i, ex_f, ey_i, ey_f):
    """Dense brute force of one lgwr step (cn_1e.hpp:767-803): Cayley factors
    (S + K)^-1 (S - K) with K = i (f step / 2) (Z (x) r) for the angular
    matrices Z = Y11 (x1) and -i Y11 (y1) themselves, f = e / 2, step dt/2
    (propagate.hpp:463), around the field-free solve."""
    S = po.band_to_dense(o.get(po.QOR_S), N, P)
    Rr = po.band_to_dense(o.get(po.QOR_RPOW1), N, P)
    Hl = o.get(po.QOR_H).reshape(-1, N * (P + 1))
    S2 = np.kron(np.eye(ML), S)
    step = dt / 2
    Z = lg["mat"]
    UR = lambda Zk, f: np.linalg.solve(S2 + (1j * (f * step) / 2) * np.kron(Zk, Rr),
                                       S2 - (1j * (f * step) / 2) * np.kron(Zk, Rr))
    U0 = np.zeros((ML * N, ML * N), complex)
    L, M, m0 = o.L, o.M, o.m0
    for im in range(M):
        for il in range(L):
            ch = im * L + il
            H = po.band_to_dense(Hl[il + abs(m0 + im)], N, P)
            U0[ch * N:(ch + 1) * N, ch * N:(ch + 1) * N] = np.linalg.solve(S + 0.5j * dt * H, S - 0.5j * dt * H)
    first = UR(Z["y1"], ey_i / 2) @ UR(Z["x1"], ex_i / 2)
    second = UR(Z["x1"], ex_f / 2) @ UR(Z["y1"], ey_f / 2)
    return second @ U0 @ first


def test_lgwr_angular_data():
    """tran_x1 is the real x1 projector, tran_y1 the conjugated y1 rows, and
    x1_2_y1 their merge product; both eigen-decompositions reproduce Y11."""
    M, L, m0 = 4, 5, -2
    lg = po.lgwr_angular(M, L, m0)
    ep = po.ep_angular(M, L, m0)
    assert np.abs(lg["eig_x1"] - ep["eig_x1"]).max() == 0.0
    assert np.abs(lg["eig_y1"] - ep["eig_y1"]).max() == 0.0
    assert np.abs(lg["tran_y1"] - ep["P"]["y1"]).max() == 0.0
    assert np.abs(lg["tran_x1"] - ep["tran_x1"]).max() == 0.0
    U = lg["tran_y1"]
    assert np.abs(U.conj().T @ np.diag(lg["eig_y1"]) @ U - lg["mat"]["y1"]).max() < 1e-13
    assert np.abs(lg["x1_2_y1"] @ lg["tran_x1"] - U).max() < 1e-13


@pytest.mark.parametrize("M,L,m0", [(3, 3, -1), (2, 4, 0)])
def test_lgwr_step_matches_dense_strang_product(M, L, m0):
    N, dt = 24, 0.05
    o = po.Oracle(N=N, p=P, L=L, M=M, m0=m0, rmax=10.0, dt=dt)
    lg = po.lgwr_angular(M, L, m0)
    ML = M * L
    U = _dense_lgwr(o, lg, N, ML, dt, 0.8, 1.1, -0.6, 0.4)
    rng = np.random.default_rng(M * 10 + L + 1)
    c0 = rng.standard_normal(ML * N) + 1j * rng.standard_normal(ML * N)
    ref = U @ c0
    o.set_lgwr(lg, isub=-1)
    c = c0.copy()
    o.step_lgwr(c, 0.8, 1.1, -0.6, 0.4)
    assert np.abs(c - ref).max() < 1e-12 * np.abs(ref).max()
    par = np.repeat(np.array([il % 2 for im in range(M) for il in range(L)]), N)
    for isub, pty in ((2, 0), (3, 1)):
        cp = np.where(par == pty, c0, 0.0)
        o.set_lgwr(lg, isub=isub)
        c = cp.copy()
        o.step_lgwr(c, 0.8, 1.1, -0.6, 0.4)
        refp = U @ cp
        assert np.abs(refp[par != pty]).max() < 1e-13 * np.abs(refp).max()
        assert np.abs(c - refp).max() < 1e-12 * np.abs(refp).max(), isub


def test_lgwr_step_is_unitary_and_reversible():
    M, L, m0, N = 3, 4, -1, 60
    fwd = po.Oracle(N=N, p=P, L=L, M=M, m0=m0, rmax=20.0, dt=0.05)
    bwd = po.Oracle(N=N, p=P, L=L, M=M, m0=m0, rmax=20.0, dt=-0.05)
    lg = po.lgwr_angular(M, L, m0)
    fwd.set_lgwr(lg)
    bwd.set_lgwr(lg)
    rng = np.random.default_rng(19)
    c0 = rng.standard_normal(M * L * N) + 1j * rng.standard_normal(M * L * N)
    c0 /= np.sqrt(fwd.norm(c0))
    c = c0.copy()
    fwd.step_lgwr(c, 0.7, 0.9, -0.3, 0.2)
    assert abs(fwd.norm(c) - 1.0) < 1e-13
    bwd.step_lgwr(c, 0.9, 0.7, 0.2, -0.3)
    assert np.abs(c - c0).max() < 1e-11
