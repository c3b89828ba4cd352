"""The coupling-sweep LU tables are built by a factorisation restarted
kRestart = 64 rows above each 24-row partition (qpc_solve.h).  This checks,
on the real operators (S with 1/r and with d/dr from the oracle's set-up),
that the restarted rows agree with the sequential reference-order
elimination to rounding over the scalar range the library accepts
(|e| <= 0.2 for the D-part; any e for the R-part), and that the agreement
does fail far outside that range (so the check has teeth)."""
import numpy as np
import pytest

import pyoracle as po

NE, ND = 8, 7
RESTART = 64


def _eliminate(S, X, e, sym, N, start, stop):
    """Reference-order banded LU (propagate.hpp:33-75) of S + e*X over rows
    [start, stop) of the matrix restarted at `start`; returns the U rows and
    pivots."""
    u = np.zeros((N, ND), complex)
    piv = np.zeros(N, complex)
    for i in range(start, stop):
        b = np.zeros(2 * ND + 1, complex)
        for j in range(1, NE):
            k = (i - j) * NE + j
            if i - j >= start:
                b[ND - j] = S[k] + 1j * e * X[k] if sym == 0 else S[k] - e * X[k]
        for j in range(NE):
            k = i * NE + j
            b[ND + j] = S[k] + 1j * e * X[k] if sym == 0 else S[k] + e * X[k]
        for j in range(max(0, ND - (i - start)), ND):
            b[j + 1:j + 1 + ND] -= b[j] * u[i + j - 7]
        u[i] = b[ND + 1:] / b[ND]
        piv[i] = b[ND]
    return u, piv


@pytest.fixture(scope="module")
def ops():
    o = po.Oracle(N=640, p=7, L=2, rmax=128.0, dt=0.05, mask_rmin=115.0, mask_rmax=128.0)
    return o.get(po.QOR_S), o.get(po.QOR_RINV), o.get(po.QOR_RDIF)


def _restart_error(S, X, e, sym, r0=400, W=RESTART):
    N = S.size // NE
    uf, pf = _eliminate(S, X, e, sym, N, 0, r0 + 24)
    ur, pr = _eliminate(S, X, e, sym, N, r0 - W, r0 + 24)
    sl = slice(r0, r0 + 24)
    du = np.abs(ur[sl] - uf[sl]).max() / np.abs(uf[sl]).max()
    dp = np.abs(pr[sl] - pf[sl]).max() / np.abs(pf[sl]).max()
    return max(du, dp)


@pytest.mark.parametrize("e", [-30.0, -1.2, 0.0, 0.3, 5.0])
def test_restart_r_part(ops, e):
    S, R, _ = ops
    assert _restart_error(S, R, e, 0) <= 1e-14


@pytest.mark.parametrize("e", [-0.2, -0.0165, 0.05, 0.2])
def test_restart_d_part_in_range(ops, e):
    S, _, D = ops
    assert _restart_error(S, D, e, 1) <= 1e-14


def test_restart_has_teeth(ops):
    S, R, D = ops
    # too short a restart, or a D-part scalar far outside the accepted range, is visibly wrong
    assert _restart_error(S, R, 0.3, 0, W=16) > 1e-7
    assert _restart_error(S, D, 1.0, 1) > 1e-9
