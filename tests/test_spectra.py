"""The spectrum helpers against closed forms: plot_hhs.m's transform of a
windowed harmonic signal peaks at its frequency with the amplitude the
window fixes; the dipole and acceleration forms agree for d'' = -w^2 d; the
PES of a Gaussian momentum packet integrates to its norm; measure_prj's
angular sum against scipy's spherical harmonics and Parseval's identity."""
import numpy as np

import spectra


def test_hhg_peak_and_forms_agree():
    dt, n = 0.05, 20000
    w0 = 143 * 2 * np.pi / (dt * n)  # on the FFT grid (no scalloping loss)
    t = dt * np.arange(n)
    tmin, tmax = 100.0, 900.0
    d = np.cos(w0 * t)
    a = -w0 ** 2 * np.cos(w0 * t)
    w, sd = spectra.hhg_spectrum(t, d, dt, tmin, tmax, kind="dipole")
    w2, sa = spectra.hhg_spectrum(t, a, dt, tmin, tmax, kind="acceleration")
    i = int(np.argmax(sd))
    assert abs(w[i] - w0) < 2 * (w[1] - w[0])
    # |d(w0)| = (1/2) int window dt / (2 pi) = (T/4) / (2 pi), T = tmax - tmin
    amp = (tmax - tmin) / 4 / (2 * np.pi)
    ref = 4 * np.pi ** 2 * w0 ** 2 / spectra.C_AU ** 2 * amp ** 2
    assert abs(sd.max() - ref) < 0.02 * ref
    # at the signal's frequency the two forms coincide (a = -w0^2 d)
    assert abs(sd[i] - sa[i]) <= 1e-9 * sd[i]


def test_pes_of_gaussian_amplitude_integrates_to_norm():
    k = np.linspace(1e-3, 8.0, 4001)
    lmd = np.exp(-k * k / 2)[:, None, None]  # l = 0 amplitude of u = r exp(-r^2/2)
    dpdk = spectra.pes_momentum(k, lmd)
    norm = np.sqrt(np.pi) / 4  # int r^2 exp(-r^2) dr
    assert abs(np.trapezoid(dpdk, k) - norm) < 1e-6
    e, dpde = spectra.pes_energy(k, lmd)
    assert abs(np.trapezoid(dpde, e) - norm) < 1e-5


def test_pmd_from_lmd_against_scipy_harmonics_and_parseval():
    """measure_prj's angular sum: on the set-up's momentum-grid tables, the
    amplitudes equal sum_(m,l) lmd Y_lm(theta, phi) with scipy's spherical
    harmonics (l = il + |m|), and the angle integral of |pmd|^2 equals
    sum |lmd|^2 (the harmonics are orthonormal on the sphere)."""
    import qpc_gpu
    from scipy.special import sph_harm_y
    M, m0, L = 3, -1, 4
    st = qpc_gpu.Setup(120, 7, L, M, m0, 0.05, 0.0, 60.0, 1.0, 1.0, 54.0, 60.0, 0.125)
    nkt, nkp = 801, 64
    t = st.tsurff(30.0, 0.5, 1.0, 2, 0.0, np.pi, nkt, 0.0, 2 * np.pi * (1 - 1 / nkp), nkp)
    rng = np.random.default_rng(5)
    lmd = rng.standard_normal((2, M, L)) + 1j * rng.standard_normal((2, M, L))
    pmd = spectra.pmd_from_lmd(lmd, t["y0"], t["e0"])
    assert pmd.shape == (2, nkt, nkp)
    for it in (0, 137, 400, 800):
        for ip in (0, 9, 41):
            th, ph = t["kt"][it], t["kp"][ip]
            want = [sum(lmd[k, im, il] * sph_harm_y(il + abs(m0 + im), m0 + im, th, ph)
                        for im in range(M) for il in range(L)) for k in range(2)]
            assert np.allclose(pmd[:, it, ip], want, rtol=0, atol=1e-13)
    # Parseval: trapezoid in theta (sin weight), exact rectangle rule in phi
    w_t = np.sin(t["kt"]) * np.gradient(t["kt"])  # the end points carry sin = 0
    integ = np.einsum("ktp,t->k", np.abs(pmd) ** 2, w_t) * (2 * np.pi / nkp)
    assert np.allclose(integ, np.sum(np.abs(lmd) ** 2, axis=(1, 2)), rtol=1e-4)
