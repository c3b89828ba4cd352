"""Spectra from the propagator's outputs (host post-processing of device
results).

hhg_spectrum restates the reference's HHG analysis (analysis/plot_hhs.m:1-60)
on a recorded dipole or dipole-acceleration history (qpc_record); the PES
helpers turn the /lmd1 partial momentum amplitudes of spectrum_c::
measure_lmd (qpc_measure_lmd, spectrum/momentum_1e.hpp:384-401) into
momentum and energy distributions and, with the angular tables, into the
angle-resolved amplitudes /pmd1 of measure_prj.
"""
import numpy as np

C_AU = 137.036  # speed of light in atomic units, as the analysis script uses


def hhg_spectrum(t, dip, dt, tmin, tmax, kind="acceleration", charge=-1.0):
    """analysis/plot_hhs.m: the windowed Fourier transform of the dipole
    signal of the electron (charge -1 times <z>, or times the acceleration
    <dV/dr cos theta>), window (tmin < t < tmax) sin^2(pi (t - tmin) /
    (tmax - tmin)); returns (omega, S(omega)) on the first ceil(n/2) FFT
    frequencies with
        S = 4 pi^2 omega^(+2) / c^2 |d(omega)|^2   (dipole moment)
        S = 4 pi^2 omega^(-2) / c^2 |a(omega)|^2   (dipole acceleration)
    and the transform d(omega) = fft(d * window) dt / (2 pi)."""
    t = np.asarray(t, np.float64)
    d = charge * np.asarray(dip, np.float64)
    window = ((t > tmin) & (t < tmax)) * np.sin(np.pi * (t - tmin) / (tmax - tmin)) ** 2
    n = d.size
    nh = int(np.ceil(n / 2))
    dw = 2 * np.pi / dt / n
    w = np.arange(1, nh + 1) * dw - dw
    df = (np.fft.fft(d * window) * dt / (2 * np.pi))[:nh]
    with np.errstate(divide="ignore"):
        if kind == "dipole":
            s = 4 * np.pi ** 2 * w ** 2 / C_AU ** 2 * np.abs(df) ** 2
        elif kind == "acceleration":
            s = np.where(w > 0, 4 * np.pi ** 2 / np.where(w > 0, w, 1.0) ** 2 / C_AU ** 2 * np.abs(df) ** 2, 0.0)
        else:
            raise ValueError("kind is 'dipole' or 'acceleration'")
    return w, s


def pes_momentum(kr, lmd):
    """dP/dk = k^2 sum over (m, l) of |lmd(k, m, l)|^2 (lmd normalised as a
    momentum-space amplitude, d^3k = k^2 dk dOmega)."""
    kr = np.asarray(kr, np.float64)
    p = np.sum(np.abs(np.asarray(lmd).reshape(kr.size, -1)) ** 2, axis=1)
    return kr ** 2 * p


def pmd_from_lmd(lmd, y0, e0):
    """spectrum_c_pcs::measure_prj (spectrum/momentum_1e.hpp:337-356) from
    the /lmd1 amplitudes of measure_lmd (which already carry sqrt(2/pi), the
    Coulomb phase and the radial integrals): pmd[kr][kt][kp] = sum over (m, l)
    of Y_lm(kt, 0) e^(i m kp) lmd[kr][m][l], the driver's /pmd1.  y0[kt][m][l]
    and e0[kp][m] are the spherical-harmonic tables of the momentum grid
    (spherical-harmonic normalisation, as initialize_th / initialize_ph fill
    them, momentum_1e.hpp:12-46; the t-SURFF set-up returns the same tables)."""
    lmd = np.asarray(lmd, np.complex128)
    return np.einsum("tml,pm,kml->ktp", np.asarray(y0, np.float64), np.asarray(e0, np.complex128), lmd)


def pes_energy(kr, lmd):
    """(E, dP/dE) with E = k^2/2 and dP/dE = dP/dk / k."""
    kr = np.asarray(kr, np.float64)
    return 0.5 * kr ** 2, pes_momentum(kr, lmd) / kr


def harmonic_table(w, s, w0, orders):
    """S at the harmonic orders (the nearest FFT frequency to q w0)."""
    out = {}
    for q in orders:
        i = int(np.argmin(np.abs(w - q * w0)))
        out[int(q)] = float(s[i])
    return out
