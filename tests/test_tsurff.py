"""t-SURFF (spectrum_c_tsf, momentum_1e.hpp:531-880): the host tables of
qpc_setup_tsurff pinned against scipy (B-splines at r0, Riccati-Bessel and
mpmath Coulomb functions, spherical harmonics, <l m|cos|l+1 m> by quadrature),
and the oracle's accumulator (qor_tsurff_accumulate) pinned by physics:
 - an outgoing free wave packet: the t-SURFF spectrum after the packet has
   crossed r0 equals |<k|psi(0)>|^2, the plane-wave projection of the initial
   state (2/pi |Y00|^2 |int sin(kr)/k u(r) dr|^2);
 - a constant vector potential in the velocity gauge leaves the canonical
   momentum distribution unchanged: the spectrum with A = 0.05 along z equals
   the field-free one at every polar angle (this exercises the i A terms with
   the <l|cos|l+1> couplings and the Volkov phase k cos(theta) int A dt);
 - the z- and field-free branches agree at A = 0 (the accumulator is linear
   in the state)."""
import numpy as np
import pytest
from scipy.interpolate import BSpline
from scipy.special import spherical_jn, sph_harm_y

import pyoracle as po


def _setup(N, p, L, rmax, M=1, m0=0, dt=0.05):
    import qpc_gpu
    return qpc_gpu.Setup(N, p, L, M, m0, dt, 0.0, rmax, 1.0, 1.0, 0.9 * rmax, rmax, 0.125)


def _knots(N, p, rmin, rmax):
    """The radial knot sequence (basis_bsplines.h:450-459): basis function i
    is the B-spline on knots[i : i+p+2]."""
    nk = N + p + 3
    h = (rmax - rmin) / (N + 2 - p)
    t = np.array([rmin if k < p else (rmin + (k - p) * h if k < N + 2 else rmax) for k in range(nk)])
    return t[1:]


def test_tsurff_tables_against_scipy():
    N, p, L, rmax, r0 = 300, 7, 6, 100.0, 47.3
    st = _setup(N, p, L, rmax)
    t = st.tsurff(r0, 0.05, 2.0, 12, 0.1, 3.0, 5, 0.0, 6.0, 3)
    kn = _knots(N, p, 0.0, rmax)
    # every basis function that is non-zero at r0, with its derivative
    want = {}
    for i in range(N):
        b = BSpline.basis_element(kn[i:i + p + 2], extrapolate=False)
        v = b(r0)
        if np.isfinite(v) and v > 0:
            want[i] = (float(v), float(b.derivative()(r0)))
    assert sorted(want) == list(t["ib"])
    for i, v, d in zip(t["ib"], t["b0"], t["br"]):
        assert abs(v - want[i][0]) < 1e-14 and abs(d - want[i][1]) < 1e-13
    # zc = 0: j0 = F_l(k r0)/k = r0 j_l(k r0), jr = d/dx (x j_l(x)) at x = k r0
    for ik, k in enumerate(t["kr"]):
        x = k * r0
        for l in range(L):
            assert abs(t["j0"][ik, l] - r0 * spherical_jn(l, x)) < 1e-12 * r0
            assert abs(t["jr"][ik, l] - (spherical_jn(l, x) + x * spherical_jn(l, x, derivative=True))) < 1e-12
    for it, th in enumerate(t["kt"]):
        for l in range(L):
            assert abs(t["y0"][it, 0, l] - sph_harm_y(l, 0, th, 0.0).real) < 1e-14
    assert np.allclose(t["e0"], 1.0)
    # <l 0|cos|l+1 0> by quadrature on the sphere
    x, w = np.polynomial.legendre.leggauss(40)
    th = np.arccos(x)
    for l in range(L):
        q = 2 * np.pi * np.sum(w * sph_harm_y(l, 0, th, 0.0).real * x * sph_harm_y(l + 1, 0, th, 0.0).real)
        assert abs(t["pmat0"][0, l] - q) < 1e-13


def test_tsurff_tables_m_and_coulomb():
    mp = pytest.importorskip("mpmath")
    N, p, L, rmax, r0, zc = 200, 7, 4, 80.0, 40.0, 1.0
    st = _setup(N, p, L, rmax, M=3, m0=-1)
    t = st.tsurff(r0, 0.2, 1.5, 4, 0.3, 2.5, 3, 0.5, 5.0, 2, zc=zc, flag=1)
    # flag 1: linear in energy
    e = 0.5 * t["kr"] ** 2
    assert np.allclose(np.diff(e), (0.5 * 1.5 ** 2 - 0.5 * 0.2 ** 2) / 3, rtol=0, atol=1e-14)
    lmax = t["j0"].shape[1] - 1
    for ik, k in enumerate(t["kr"]):
        eta, x = -zc / k, k * r0
        for l in range(lmax + 1):
            f = float(mp.coulombf(l, eta, x))
            fd = float(mp.diff(lambda z: mp.coulombf(l, eta, z), x))
            assert abs(t["j0"][ik, l] - f / k) < 1e-11 * max(1.0, abs(f / k))
            assert abs(t["jr"][ik, l] - fd) < 1e-11
    for it, th in enumerate(t["kt"]):
        for im, m in enumerate((-1, 0, 1)):
            for il in range(L):
                l = il + abs(m)
                assert abs(t["y0"][it, im, il] - sph_harm_y(l, m, th, 0.0).real) < 1e-14
    for ip, ph in enumerate(t["kp"]):
        assert np.allclose(t["e0"][ip], np.exp(1j * np.array([-1, 0, 1]) * ph), atol=1e-15)
    x, w = np.polynomial.legendre.leggauss(40)
    th = np.arccos(x)
    for im, m in enumerate((-1, 0, 1)):
        for il in range(L):
            l = il + abs(m)
            q = 2 * np.pi * np.sum(w * sph_harm_y(l, m, th, 0.0).real * x * sph_harm_y(l + 1, m, th, 0.0).real)
            assert abs(t["pmat0"][im, il] - q) < 1e-13


def _packet(o, N, p, rmax, rc, sigma, k0):
    """S-projection of u(r) = exp(-(r-rc)^2/(2 sigma^2) + i k0 r), S-normalised, into l = 0."""
    kn = _knots(N, p, 0.0, rmax)
    xg, wg = np.polynomial.legendre.leggauss(2 * p + 2)
    b = np.zeros(N, np.complex128)
    u = lambda r: np.exp(-(r - rc) ** 2 / (2 * sigma ** 2) + 1j * k0 * r)
    for i in range(N):
        s = BSpline.basis_element(kn[i:i + p + 2], extrapolate=False)
        for a0, a1 in zip(kn[i:i + p + 1], kn[i + 1:i + p + 2]):
            if a1 > a0:
                r = 0.5 * (a1 + a0) + 0.5 * (a1 - a0) * xg
                b[i] += 0.5 * (a1 - a0) * np.sum(wg * s(r) * u(r))
    S = po.band_to_dense(o.get(po.QOR_S), N, p)
    c = np.linalg.solve(S, b)
    c /= np.sqrt(np.real(np.conj(c) @ S @ c))
    return c, u, np.real(np.conj(c) @ S @ c)


def _projection_l0(u, k, rc, sigma):
    """<k|psi> for psi = u(r)/r Y00 normalised: sqrt(2/pi) Y00 int sin(kr)/k u(r) dr."""
    r = np.linspace(max(0.0, rc - 12 * sigma), rc + 12 * sigma, 20001)
    f = u(r)[None, :] * np.sin(np.outer(k, r)) / k[:, None]
    return np.sqrt(2 / np.pi) / np.sqrt(4 * np.pi) * np.trapezoid(f, r, axis=1)


@pytest.mark.parametrize("A", [0.0, 0.05])
def test_tsurff_free_packet_and_constant_field(A):
    """|t-SURFF amplitude|^2 after the packet has crossed r0 against the
    plane-wave projection of the initial state; with a constant A along z the
    canonical momentum distribution is unchanged at every polar angle."""
    N, p, rmax, r0, dt = 400, 7, 200.0, 100.0, 0.05
    # the packet drifts by A T along z: channels up to l ~ k A T (+ margin)
    L = 1 if A == 0.0 else 18
    rc, sigma, k0 = 60.0, 5.0, 1.5
    st = _setup(N, p, L, rmax, dt=dt)
    # a free particle (no Coulomb potential): the plane waves are the continuum states
    o = po.Oracle(N=N, p=p, L=L, rmax=rmax, zc=0.0, dt=dt, mask_rmin=180.0, mask_rmax=200.0, mask_fact=0.125)
    c0, u, _ = _packet(o, N, p, rmax, rc, sigma, k0)
    c = np.zeros(L * N, np.complex128)
    c[:N] = c0[:N]
    assert abs(o.norm(c) - 1.0) < 1e-12
    t = st.tsurff(r0, 1.2, 1.8, 7, 0.0, np.pi, 4 if A else 1)
    state, work = np.zeros(2), np.zeros((7, len(t["kt"]), 1), np.complex128)
    nsteps = int(110.0 / dt)
    for _ in range(nsteps):
        if A:
            o.step(c, A, A)
        else:
            o.step_free(c)
        o.tsurff_accumulate(c, A, A, dt, t, state, work)
        o.mask(c)
    # c0 is the S-normalised projection of u: compare with u / |u|
    r = np.linspace(1e-6, rmax - 1e-6, 40001)
    scale = np.sqrt(np.trapezoid(np.abs(u(r)) ** 2, r))
    want = np.abs(_projection_l0(u, t["kr"], rc, sigma) / scale) ** 2
    got = np.abs(work[:, :, 0]) ** 2
    print(f"A={A}: time {state[0]:.2f}, disz {state[1]:.3f}")
    for ikt in range(got.shape[1]):
        rel = np.abs(got[:, ikt] - want) / want.max()
        print(f"  theta={t['kt'][ikt]:.3f}: max rel dev {rel.max():.2e}")
        assert rel.max() < 2e-3
    assert abs(state[0] - nsteps * dt) < 1e-9 and abs(state[1] - A * nsteps * dt) < 1e-9


def test_tsurff_xy_tables_match_sphere_quadrature_closed_forms():
    """pmat1/pmat2 (qpc_setup_tsurff_xy, integrate_sinexp1/2) equal the
    <l m|sin e^(-i ph)|l+-1 m+1> closed forms pinned by sphere quadrature in
    test_ep_oracle."""
    M, L, m0 = 5, 4, -2
    st = _setup(120, 7, L, 60.0, M=M, m0=m0)
    t = st.tsurff(30.0, 0.5, 1.0, 2)
    for im in range(M):
        for il in range(L):
            m = m0 + im
            se1, se2, _, _ = po.ep_closed_forms(m, il + abs(m))
            assert t["pmat1"][im, il] == se1 and t["pmat2"][im, il] == se2


def _sphere(n=24):
    x, w = np.polynomial.legendre.leggauss(n)
    ph = 2 * np.pi * np.arange(2 * n) / (2 * n)
    T, F = np.meshgrid(np.arccos(x), ph, indexing="ij")
    return T, F, w[:, None] * np.full((1, 2 * n), 2 * np.pi / (2 * n))


def test_tsurff_xy_couplings_against_sphere_quadrature():
    """accumulate (xy) minus its field-free form, with the step's mean field
    zero (no Volkov displacement) and A(t_f) = (ax, ay, 0): the difference is
    sqrt(2/pi) dt e^(i k^2 dt/2) sum_lm Y_lm(k) (-i)^(l+1) (-i) j_l(k) Q_lm
    with Q_lm = <Y_lm| A.r |psi(r0)> computed by quadrature on the sphere of
    psi(r0, .) = sum temp_lm Y_lm (no closed-form matrix element involved)."""
    N, p, L, M, m0, rmax, r0, dt = 160, 7, 5, 5, -2, 80.0, 40.0, 0.05
    st = _setup(N, p, L, rmax, M=M, m0=m0, dt=dt)
    t = st.tsurff(r0, 0.3, 1.5, 4, 0.2, 2.9, 3, 0.3, 5.9, 4)
    o = po.Oracle(N=N, p=p, L=L, M=M, m0=m0, rmax=rmax, dt=dt)
    rng = np.random.default_rng(61)
    c = rng.standard_normal(M * L * N) + 1j * rng.standard_normal(M * L * N)
    ax, ay = 0.7, -0.4
    s1, w1 = np.zeros(3), np.zeros((4, 3, 4), np.complex128)
    s0, w0 = np.zeros(3), np.zeros((4, 3, 4), np.complex128)
    o.tsurff_accumulate_xy(c, -ax, ax, -ay, ay, dt, t, t["pmat1"], t["pmat2"], s1, w1)
    o.tsurff_accumulate_xy(c, 0.0, 0.0, 0.0, 0.0, dt, t, t["pmat1"], t["pmat2"], s0, w0)
    assert np.array_equal(s1, [dt, 0.0, 0.0]) and np.array_equal(s0, s1)
    temp = c.reshape(M * L, N)[:, t["ib"]] @ t["b0"]
    T, F, W = _sphere()
    lm = [(il + abs(m0 + im), m0 + im) for im in range(M) for il in range(L)]
    Y = np.array([sph_harm_y(l, m, T, F) for l, m in lm])
    psi = np.tensordot(temp, Y, axes=1)
    ar = np.sin(T) * (ax * np.cos(F) + ay * np.sin(F))
    Q = np.array([np.sum(W * np.conj(y) * ar * psi) for y in Y])
    want = np.zeros_like(w1)
    for ik, k in enumerate(t["kr"]):
        for it, th in enumerate(t["kt"]):
            for ip, ph in enumerate(t["kp"]):
                s = sum(sph_harm_y(l, m, th, ph) * (-1j) ** (l + 1) * (-1j) * t["j0"][ik, l] * Q[ch]
                        for ch, (l, m) in enumerate(lm))
                want[ik, it, ip] = np.sqrt(2 / np.pi) * dt * np.exp(0.5j * k * k * dt) * s
    got = w1 - w0
    assert np.abs(want).max() > 1e-3
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


def test_tsurff_xy_volkov_phase_is_the_rotated_z_phase():
    """For a state in the (l, m) = (0, 0) channel the flux term is isotropic, so
    the amplitude at k along x (y) after a step with int A_x dt = d (int A_y dt
    = d) and A(t_f) = 0 equals the z-polarised accumulate's (pinned by the
    free-packet physics above) at k along z with int A_z dt = d; along z the
    xy form keeps the field-free phase."""
    N, p, L, M, m0, rmax, r0, dt, d = 160, 7, 3, 3, -1, 80.0, 40.0, 0.05, 0.8
    st = _setup(N, p, L, rmax, M=M, m0=m0, dt=dt)
    t = st.tsurff(r0, 0.3, 1.5, 3, 0.0, np.pi / 2, 2, 0.0, np.pi / 2, 2)
    o = po.Oracle(N=N, p=p, L=L, M=M, m0=m0, rmax=rmax, dt=dt)
    rng = np.random.default_rng(67)
    c = np.zeros(M * L * N, np.complex128)
    ch = (-m0) * L  # m = 0, l = 0
    c[ch * N:(ch + 1) * N] = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    shape = (3, 2, 2)
    sz, wz = np.zeros(2), np.zeros(shape, np.complex128)
    o.tsurff_accumulate(c, 2 * d / dt, 0.0, dt, t, sz, wz)
    assert abs(sz[1] - d) < 1e-15
    sx, wx = np.zeros(3), np.zeros(shape, np.complex128)
    o.tsurff_accumulate_xy(c, 2 * d / dt, 0.0, 0.0, 0.0, dt, t, t["pmat1"], t["pmat2"], sx, wx)
    sy, wy = np.zeros(3), np.zeros(shape, np.complex128)
    o.tsurff_accumulate_xy(c, 0.0, 0.0, 2 * d / dt, 0.0, dt, t, t["pmat1"], t["pmat2"], sy, wy)
    s0, w0 = np.zeros(3), np.zeros(shape, np.complex128)
    o.tsurff_accumulate_xy(c, 0.0, 0.0, 0.0, 0.0, dt, t, t["pmat1"], t["pmat2"], s0, w0)
    assert abs(sx[1] - d) < 1e-15 and abs(sy[2] - d) < 1e-15
    ref = wz[:, 0, 0]  # k along z
    sc = np.abs(ref).max()
    assert sc > 1e-3
    assert np.abs(wx[:, 1, 0] - ref).max() < 1e-13 * sc  # k along x
    assert np.abs(wy[:, 1, 1] - ref).max() < 1e-13 * sc  # k along y
    assert np.abs(wx[:, 0, :] - w0[:, 0, :]).max() < 1e-13 * sc  # k along z: no x displacement phase
    assert np.abs(wx[:, 1, 0] - w0[:, 1, 0]).max() > 1e-3 * sc
