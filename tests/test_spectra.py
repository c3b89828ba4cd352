"""The spectrum helpers against closed forms: plot_hhs.m's transform of a
windowed harmonic signal peaks at its frequency with the amplitude the
window fixes; the dipole and acceleration forms agree for d'' = -w^2 d; the
PES of a Gaussian momentum packet integrates to its norm."""
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
