"""The partitioned solve truncates each carry after PK partitions of 24 rows
(qpc_solve.h): the incoming state of a partition is assembled from the end
states of the PK partitions before it, and what lies further back is
dropped.  What is dropped is the response of the triangular factors'
inverses across PK*24 rows.  This measures that response directly on the
real operators (dense LU without pivoting of a window of the banded
matrices from the oracle's set-up, the factors' inverses by triangular
solves): PK = 6 (144 rows) for the R-part S + i e R (|e| up to 30), the
field-free S + i dt/2 H_l and the absorber's chol(S); PK = 8 (192 rows) for
the D-part S + e D over the scalar range the library accepts
(|e| <= kMaxDScalar = 0.035).  The bar, 1e-17 relative, keeps the truncation
below FP64 rounding; just beyond the accepted D-part range the 192-row tail
exceeds it, which is why the range is limited."""
import numpy as np
import pytest
import scipy.linalg as sl

import pyoracle as po

NE = 8
DEPTH_S = 6 * 24  # rows behind the horizon of the R-part, field-free and absorber carries
DEPTH_D = 8 * 24  # rows behind the horizon of the D-part carries
D_MAX = 0.035     # kMaxDScalar
WIN = 420


@pytest.fixture(scope="module")
def ops():
    o = po.Oracle(N=640, p=7, L=31, rmax=128.0, dt=0.05, mask_rmin=115.0, mask_rmax=128.0)
    return o


def _dense(band, w0, n, lower=None):
    """Rows/columns [w0, w0+n) of the banded matrix with upper band
    B(i, i+j) = band[i*NE + j]; the lower triangle is `lower` (a band of the
    same layout, read transposed) or the symmetric image."""
    A = np.zeros((n, n), complex)
    lo = band if lower is None else lower
    for i in range(n):
        for j in range(NE):
            if i + j < n:
                A[i, i + j] = band[(w0 + i) * NE + j]
                if j:
                    A[i + j, i] = lo[(w0 + i) * NE + j]
    return A


def _lu_nopivot(A):
    A = A.copy()
    n = A.shape[0]
    for k in range(n - 1):
        hi = min(n, k + NE)
        A[k + 1:hi, k] /= A[k, k]
        A[k + 1:hi, k + 1:hi] -= np.outer(A[k + 1:hi, k], A[k, k + 1:hi])
    return np.tril(A, -1) + np.eye(n), np.triu(A)


def _tail(A, depth):
    """Largest response of L^-1 and U^-1 (unit diagonal) `depth` rows away
    from a unit source in the window interior, relative to the response's
    peak."""
    L, U = _lu_nopivot(A)
    Un = U / np.diag(U)[:, None]
    r = 60
    worst = 0.0
    for M, lower, src in ((L, True, r), (Un, False, A.shape[0] - 1 - r)):
        e = np.zeros(A.shape[0], complex)
        e[src] = 1.0
        y = sl.solve_triangular(M, e, lower=lower, unit_diagonal=True)
        far = y[src + depth:] if lower else y[:src - depth + 1]
        worst = max(worst, np.abs(far).max() / np.abs(y).max())
    return worst


@pytest.mark.parametrize("w0", [0, 200])
@pytest.mark.parametrize("e", [-30.0, -1.0, 0.05, 0.6, 30.0])
def test_carry_tail_r_part(ops, w0, e):
    S, R = ops.get(po.QOR_S), ops.get(po.QOR_RINV)
    assert _tail(_dense(S + 1j * e * R, w0, WIN), DEPTH_S) <= 1e-17


@pytest.mark.parametrize("w0", [0, 200])
@pytest.mark.parametrize("e", [-D_MAX, -0.017, 0.01, D_MAX])
def test_carry_tail_d_part(ops, w0, e):
    S, D = ops.get(po.QOR_S), ops.get(po.QOR_RDIF)
    # S + e D with D antisymmetric: upper B(i,i+j) = S + e D(i,i+j), lower B(i+j,i) = S - e D(i,i+j)
    assert _tail(_dense(S + e * D, w0, WIN, lower=S - e * D), DEPTH_D) <= 1e-17


@pytest.mark.parametrize("w0", [0, 200])
@pytest.mark.parametrize("l", [0, 10, 30])
def test_carry_tail_field_free(ops, w0, l):
    upper = ops.get(po.QOR_UPPER).reshape(-1, 640 * NE)[l]
    assert _tail(_dense(upper, w0, WIN), DEPTH_S) <= 1e-17


@pytest.mark.parametrize("w0", [0, 200])
def test_carry_tail_absorber(ops, w0):
    S = ops.get(po.QOR_S)
    assert _tail(_dense(S.astype(complex), w0, WIN), DEPTH_S) <= 1e-17


def test_tail_measure_has_teeth(ops):
    """A 48-row horizon leaves a response far above rounding, and so does
    the D-part's 192-row horizon just beyond the accepted scalar range."""
    S, R, D = ops.get(po.QOR_S), ops.get(po.QOR_RINV), ops.get(po.QOR_RDIF)
    assert _tail(_dense(S + 0.6j * R, 200, WIN), 48) > 1e-8
    assert _tail(_dense(S + 0.06 * D, 200, WIN, lower=S - 0.06 * D), DEPTH_D) > 1e-15
