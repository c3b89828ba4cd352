"""The learned Krylov controller: training data, labels and training.

The controller (include/qpc_mlp.h, a 7 -> 64 -> 64 -> 2 GELU perceptron run
by the library's HIP kernels) chooses, per Magnus step of
qpc_krylov_propagate, a starting subspace size and a step multiplier.  Its
labels come from the plain method itself: along seeded random pulses the
plain adaptive Lanczos step is recorded with a full subspace, and from each
recorded Lanczos tridiagonal the a-posteriori estimate |x_{m-1}|^2
(reference propagator/la_1e.hpp:296-310) is evaluated for every smaller
subspace and every candidate step.  The estimate sees only the error of the
exponential, not the time-discretisation error of the Magnus step, so each
candidate step size is also checked by step doubling from the recorded
state against a local error budget proportional to the step.  The label is
the admissible candidate with the least Krylov work per unit of time.  At
run time the estimate still judges every exponential (the subspace is
extended, or the step halved, when it exceeds the tolerance).

Pulses: sin^2 envelope on A(t), intensity log-uniform in [5e13, 5e14] W/cm^2
(E0 = sqrt(I / 3.51e16)), CEP uniform in [0, 2 pi), 2 to 8 cycles,
omega = 0.057, drawn from numpy's PCG64 with a fixed seed.  Training steps
start from the field-free initial state at seeded times inside each pulse.
Errors are measured in the wavefunction's L2 norm (the overlap metric S):
in the plain coefficient norm the ill-conditioned B-spline overlap lets
physically negligible high-energy components dominate.
"""
import numpy as np
from scipy.linalg import eigh_tridiagonal

MU_GRID = 2.0 ** np.array([-1.0, -0.5, 0.0, 0.5, 1.0])
WORK_OFFSET = 4.0  # per-exponential overhead in units of one Krylov vector
W0 = 0.057


def random_pulses(n, seed):
    """(f0, w0, nc, n0, c0) of field_component_sinen<2> for n seeded pulses."""
    rng = np.random.default_rng(seed)
    inten = np.exp(rng.uniform(np.log(5e13), np.log(5e14), n))
    e0 = np.sqrt(inten / 3.51e16)
    cep = rng.uniform(0.0, 2 * np.pi, n)
    nc = rng.uniform(2.0, 8.0, n)
    return [(float(e0[k] / W0), W0, float(nc[k]), 0.0, float(cep[k] / np.pi)) for k in range(n)]


def sinen2(p, t):
    """A(t) of field_component_sinen<2> (reference field.h:99-123)."""
    f0, w0, nc, n0, c0 = p
    pd = 2 * np.pi / w0
    tau, ti = nc * pd, n0 * pd
    if not (ti < t < ti + tau):
        return 0.0
    return f0 * np.sin(w0 * t + np.pi * c0) * np.sin(np.pi * (t - ti) / tau) ** 2


def min_subspace(ab, hh, tol, b0=1.0):
    """Smallest m' whose estimate |x_{m'-1}|^2 <= tol for the exponential
    exp(-i hh T) of the leading m' x m' Lanczos block, x = b0 Z e^{-i hh L} Z^T e0
    (la_1e.hpp:157-182); None when even the whole recorded space misses."""
    a, b = ab[:, 0], ab[:, 1]
    m = len(a)
    for mp in range(2, m + 1):
        lam, Z = eigh_tridiagonal(a[:mp], b[:mp - 1])
        x = b0 * (Z @ (np.exp(-1j * hh * lam) * Z[0, :]))
        if abs(x[-1]) ** 2 <= tol:
            return mp
    return None


def snorm(S, x):
    """sqrt(sum_ch <x_ch|S|x_ch>) for a state of channels of N coefficients and
    the reference band layout S[i*ne + j] = S(i, i+j): the L2 norm of the
    wavefunction."""
    ne = 8
    N = S.size // ne
    B = S.reshape(N, ne)
    X = x.reshape(-1, N)
    Y = B[:, 0] * X
    for j in range(1, ne):
        Y[:, :N - j] += B[:N - j, j] * X[:, j:]
        Y[:, j:] += B[:N - j, j] * X[:, :N - j]
    return float(np.sqrt(max(np.real(np.vdot(X, Y)), 0.0)))


LOCAL_TOL = 1e-26  # exponentials of the step-doubling probe: far below the Magnus error measured


def local_errors(prop, S, state, pulse, t, h, kmax):
    """Local error of one CF4 step of size mu*h from `state` at time t, for
    every mu on MU_GRID, by step doubling: |U(mu h) c - U(mu h/2)^2 c|_S * 16/15
    (4th order), relative to |c|_S.  The exponentials run to LOCAL_TOL so
    that their own error does not mask the time-discretisation error."""
    out = np.zeros(len(MU_GRID))
    nrm = snorm(S, state)
    for i, mu in enumerate(MU_GRID):
        hs = mu * h
        prop.set_state(state)
        prop.krylov_propagate(pulse, t, t + hs, hs, LOCAL_TOL, kmin=4, kmax=kmax)
        full = prop.get_state()
        prop.set_state(state)
        prop.krylov_propagate(pulse, t, t + hs, 0.5 * hs, LOCAL_TOL, kmin=4, kmax=kmax)
        out[i] = snorm(S, full - prop.get_state()) * 16.0 / 15.0 / nrm
    return out


def label(ab, m, h, tol, kmax, loc_err, err_rate):
    """Target (y0, y1) = ((k - 8)/32, log2 mu) of one recorded step: among the
    multipliers mu on MU_GRID whose local Magnus error is within the budget
    err_rate * mu * h and whose exponential meets `tol` inside the recorded
    subspace, the one with the least Krylov work per unit time
    (m*(mu) + WORK_OFFSET) / mu; k = m*(mu).  The smallest multiplier when
    none qualifies."""
    ab = ab[:2 * m].reshape(-1, 2)
    best = None
    for i, mu in enumerate(MU_GRID):
        if loc_err[i] > err_rate * mu * h:
            continue
        k = min_subspace(ab, 0.5 * mu * h, tol)
        if k is None:
            continue
        cost = (k + WORK_OFFSET) / mu
        if best is None or cost < best[0]:
            best = (cost, k, mu)
    if best is None:
        k = min_subspace(ab, 0.5 * MU_GRID[0] * h, tol) or kmax
        best = (0.0, k, MU_GRID[0])
    return (min(max(best[1], 8), kmax) - 8) / 32.0, float(np.log2(best[2]))


def generate(prop, S, state0, pulses, windows, steps_per_window, dt_base, tol, kmax, err_rate, seed, log=None):
    """Training samples: for each pulse, `windows` seeded start times inside
    it; from `state0` (the field-free initial state) `steps_per_window` plain
    Krylov steps are recorded, and each recorded step is labelled from its
    Lanczos tridiagonal and the step-doubling errors at its own start state.
    Returns (features, targets)."""
    rng = np.random.default_rng(seed)
    feats, targs = [], []
    for ip, p in enumerate(pulses):
        tau = p[2] * 2 * np.pi / p[1]
        for t in np.sort(rng.uniform(0.05 * tau, 0.95 * tau, windows)):
            prop.set_state(state0)
            _, rec = prop.krylov_propagate(p, t, t + steps_per_window * dt_base, dt_base, tol, kmin=8, kmax=kmax,
                                           record=steps_per_window)
            # start state of each recorded step: replay the plain steps before it
            prop.set_state(state0)
            ts = t
            for s in range(rec["h"].size):
                cs = prop.get_state()
                le = local_errors(prop, S, cs, p, ts, rec["h"][s], 2 * kmax)
                feats.append(rec["features"][s])
                targs.append(label(rec["ab"][s], int(rec["m"][s]), rec["h"][s], tol, kmax, le, err_rate))
                prop.set_state(cs)
                prop.krylov_propagate(p, ts, ts + rec["h"][s], rec["h"][s], tol, kmin=8, kmax=kmax)
                ts += rec["h"][s]
            if log:
                log(f"pulse {ip + 1}/{len(pulses)}, window at t={t:.1f}: {len(feats)} samples")
    return np.array(feats, np.float32), np.array(targs, np.float32)
