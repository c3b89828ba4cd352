"""CPU-side checks of the C-ABI library: it loads, exports every entry point
declared in include/qpc_gpu.h, and its host set-up matches the oracle's."""
import ctypes
import os
import re

import numpy as np
import pytest

import pyoracle as po
import qpc_gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(headers):
    txt = "".join(open(os.path.join(ROOT, "include", h)).read() for h in headers)
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(qpc_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import glob
    # every public header is served by one of the two libraries
    assert sorted(os.path.basename(f) for f in glob.glob(os.path.join(ROOT, "include", "*.h"))) == \
        ["qpc_gpu.h", "qpc_io.h", "qpc_mlp.h"]
    lib = ctypes.CDLL(qpc_gpu.LIB_PATH)
    names = declared_symbols(["qpc_gpu.h", "qpc_mlp.h"])
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_io_library_exports_every_declared_symbol():
    import qpc_io
    if not os.path.exists(qpc_io.IO_PATH):
        pytest.skip("libqpc_io.so not built (no HDF5)")
    lib = ctypes.CDLL(qpc_io.IO_PATH)
    names = declared_symbols(["qpc_io.h"])
    assert len(names) == 6
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_fails_loudly_without_library(monkeypatch):
    monkeypatch.setattr(qpc_gpu, "LIB_PATH", "/nonexistent/libqpc_gpu.so")
    monkeypatch.setattr(qpc_gpu, "_LIB", None)
    with pytest.raises(qpc_gpu.QpcError):
        qpc_gpu.lib()


@pytest.mark.parametrize("N,L,rmax", [(400, 7, 80.0), (1000, 31, 200.0)])
def test_host_setup_matches_oracle(N, L, rmax):
    o = po.Oracle(N=N, p=7, L=L, rmax=rmax, dt=0.05, mask_rmin=rmax - 20, mask_rmax=rmax, mask_fact=0.125)
    st = qpc_gpu.Setup(N, 7, L, 1, 0, 0.05, 0.0, rmax, 1.0, 1.0, rmax - 20, rmax, 0.125)
    pairs = [(po.QOR_S, qpc_gpu.SETUP_S), (po.QOR_CHOL, qpc_gpu.SETUP_CHOL), (po.QOR_H, qpc_gpu.SETUP_H),
             (po.QOR_RINV, qpc_gpu.SETUP_RINV), (po.QOR_RDIF, qpc_gpu.SETUP_RDIF),
             (po.QOR_RPOW1, qpc_gpu.SETUP_RPOW1), (po.QOR_MASK, qpc_gpu.SETUP_MASK),
             (po.QOR_UPPER, qpc_gpu.SETUP_UPPER), (po.QOR_LOWER, qpc_gpu.SETUP_LOWER),
             (po.QOR_YEIG, qpc_gpu.SETUP_YEIG), (po.QOR_QEIG, qpc_gpu.SETUP_QEIG), (po.QOR_Y10, qpc_gpu.SETUP_Y10),
             (po.QOR_DVDR, qpc_gpu.SETUP_DVDR)]
    for a, b in pairs:
        x, y = o.get(a), st.get(b)
        assert x.shape == y.shape
        assert np.abs(x - y).max() <= 1e-13 * max(1.0, np.abs(x).max()), (a, b)
    # eigenvector phases are free; the operators P^H diag(e) P are not
    for wp, we, op, oe in [(qpc_gpu.SETUP_YPRJ, qpc_gpu.SETUP_YEIG, po.QOR_YPRJ, po.QOR_YEIG),
                           (qpc_gpu.SETUP_QPRJ, qpc_gpu.SETUP_QEIG, po.QOR_QPRJ, po.QOR_QEIG)]:
        A = st.get(wp).reshape(L, L)
        B = o.get(op).reshape(L, L)
        assert np.abs(A.conj().T @ np.diag(st.get(we)) @ A - B.conj().T @ np.diag(o.get(oe)) @ B).max() < 1e-12
    e1, c1 = st.eigenstate(0, 0, 0)
    e2, c2 = o.eigenstate(0, 0, 0)
    assert abs(e1 - e2) < 1e-13
    assert np.abs(c1 - c2).max() < 1e-12
    assert st.lu_info == o.lu_info == 0


def test_host_projection_basis_matches_oracle():
    """qpc_setup_projection_basis vs the oracle's eigenstate_sph basis: the
    same energies and the same left-hand vectors S v (sign convention shared)."""
    import pyoracle as po
    import qpc_gpu
    N, L, M, m0 = 400, 5, 2, -1
    st = qpc_gpu.Setup(N, 7, L, M, m0, 0.05, 0.0, 80.0, 1.0, 1.0, 72.0, 80.0, 0.125)
    W, E = st.projection_basis(6)
    o = po.Oracle(N=N, p=7, L=L, M=M, m0=m0, rmax=80.0, mask_rmin=72.0, mask_rmax=80.0)
    W2, E2 = o.projection_basis(6)
    np.testing.assert_allclose(E, E2, rtol=0, atol=1e-13)
    np.testing.assert_allclose(W, W2, rtol=0, atol=1e-12 * np.abs(W2).max())


def test_host_setup_q10_matches_oracle():
    """Q10 off-diagonals (operator_q10) exported by the host set-up for the Krylov H(t)."""
    import pyoracle as po
    import qpc_gpu
    N, L, M, m0 = 200, 6, 3, -1
    st = qpc_gpu.Setup(N, 7, L, M, m0, 0.05, 0.0, 40.0, 1.0, 1.0, 36.0, 40.0, 0.125)
    o = po.Oracle(N=N, p=7, L=L, M=M, m0=m0, rmax=40.0, mask_rmin=36.0, mask_rmax=40.0)
    np.testing.assert_allclose(st.get(qpc_gpu.SETUP_Q10), o.get(po.QOR_Q10), rtol=0, atol=1e-14)
    np.testing.assert_allclose(st.get(qpc_gpu.SETUP_Y10), o.get(po.QOR_Y10), rtol=0, atol=1e-14)


@pytest.mark.parametrize("M,L,m0", [(3, 4, -1), (4, 3, 0), (5, 5, -2), (9, 7, -4)])
def test_host_angular_ep_reproduces_angular_operators(M, L, m0):
    """qpc_setup_angular_ep (the library's host set-up of propagator_cn_hc_vgwr)
    against the angular matrices of operator_y11/q11 (closed forms pinned by
    sphere quadrature in test_ep_oracle): the projectors recovered from its
    outputs (P_x1 = tran_x1, P_x2 = x1_2_x2 P_x1, P_y1 = x2_2_y1 P_x2,
    P_y2 = y1_2_y2 P_y1 = tran_y2) diagonalise them with its eigenvalues,
    which equal the restatement's; every projector is unitary."""
    g = qpc_gpu.angular_ep(M, L, m0)
    r = po.ep_angular(M, L, m0)
    ML = M * L
    P = {"x1": g["tran_x1"].astype(complex)}
    P["x2"] = g["x1_2_x2"] @ P["x1"]
    P["y1"] = g["x2_2_y1"] @ P["x2"]
    P["y2"] = g["y1_2_y2"] @ P["y1"]
    assert np.abs(P["y2"] - g["tran_y2"]).max() < 1e-14
    for k in ("x1", "x2", "y1", "y2"):
        assert np.abs(g["eig_" + k] - r["eig_" + k]).max() < 1e-13
        assert np.abs(P[k] @ P[k].conj().T - np.eye(ML)).max() < 1e-13
        assert np.abs(P[k].conj().T @ np.diag(g["eig_" + k]) @ P[k] - r["mat"][k]).max() < 1e-13


@pytest.mark.parametrize("M,L,m0", [(3, 4, -1), (4, 3, 0), (6, 5, -2)])
def test_host_angular_lgwr_reproduces_angular_operators(M, L, m0):
    """qpc_setup_angular_lgwr (the host set-up of propagator_cn_hc_lgwr): its
    eigenvalues equal the restatement's, tran_y1 diagonalises -i Y11 and is
    unitary, and x1_2_y1 carries the x1 projector (tran_x1) to it."""
    g = qpc_gpu.angular_lgwr(M, L, m0)
    r = po.lgwr_angular(M, L, m0)
    ML = M * L
    Px, Py = g["tran_x1"].astype(complex), g["tran_y1"]
    for k, Pk in (("x1", Px), ("y1", Py)):
        assert np.abs(g["eig_" + k] - r["eig_" + k]).max() < 1e-13
        assert np.abs(Pk @ Pk.conj().T - np.eye(ML)).max() < 1e-13
        assert np.abs(Pk.conj().T @ np.diag(g["eig_" + k]) @ Pk - r["mat"][k]).max() < 1e-13
    assert np.abs(g["x1_2_y1"] @ Px - Py).max() < 1e-14


def test_host_setup_hc_parts_match_scipy_and_rebuild_h():
    """QPC_SETUP_HV/HK/HT (the l-independent parts of integrate_hc,
    basisfunc.hpp:92-106, the operands of qpc_build_hc) against scipy
    B-splines on the same knots, and H_l = V + l(l+1)/2 K + T/2 for every l
    equal to the set-up's own QPC_SETUP_H."""
    from test_oracle_pins import _scipy_bands
    N, L, rmax = 120, 5, 24.0
    o = po.Oracle(N=N, p=7, L=L, rmax=rmax)
    st = qpc_gpu.Setup(N, 7, L, 1, 0, 0.05, 0.0, rmax, 1.0, 1.0, 20.0, rmax, 0.125)
    band = _scipy_bands(o, N, 7)
    V, K, T = (st.get(w) for w in (qpc_gpu.SETUP_HV, qpc_gpu.SETUP_HK, qpc_gpu.SETUP_HT))
    for got, want in ((V, band(lambda r: -1.0 / r, 0, 0)), (K, band(lambda r: 1.0 / r ** 2, 0, 0)),
                      (T, band(lambda r: np.ones_like(r), 1, 1))):
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    H = st.get(qpc_gpu.SETUP_H).reshape(L, -1)
    for l in range(L):
        assert np.abs(H[l] - (V + l * (l + 1) * 0.5 * K + 0.5 * T)).max() <= 1e-15 * np.abs(H[l]).max()


def test_host_setup_outer_band_matches_scipy():
    """QPC_SETUP_SOUTER = <Bi|theta(r - mask_rmin)|Bj> against scipy B-splines
    on the same knots and Gauss points (the window overlap of
    observable_molec, observable/1e.hpp:188-192)."""
    from test_oracle_pins import _scipy_bands
    N, rmax, r0 = 120, 24.0, 20.0
    o = po.Oracle(N=N, p=7, L=2, rmax=rmax, mask_rmin=r0, mask_rmax=rmax)
    st = qpc_gpu.Setup(N, 7, 2, 1, 0, 0.05, 0.0, rmax, 1.0, 1.0, r0, rmax, 0.125)
    ref = _scipy_bands(o, N, 7)(lambda r: (r >= r0).astype(float), 0, 0)
    got = st.get(qpc_gpu.SETUP_SOUTER)
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


def test_angular_eigenvectors_agree_to_their_conditioning():
    """The one set-up operand that differs from the oracle's by more than a
    few ulps is the eigenvector matrix of Y10 (yprj, and the merged
    projector built from it): two backward-stable tridiagonal solvers give
    eigenvectors that agree only to about eps |Y| / gap, and the eigenvalues
    of Y10 (cos theta in the l basis) cluster like Gauss-Legendre nodes,
    gaps ~ 1/L^2.  Both bases are orthonormal and diagonalise Y10 to
    rounding; their difference per eigenvector stays within the Davis-Kahan
    bound.  This difference carries the 2e-10 whole-run gap between the
    library's own set-up and the oracle's operators on the N=3000, L=81 run."""
    import pyoracle as po
    import qpc_gpu
    L = 81
    st = qpc_gpu.Setup(300, 7, L, 1, 0, 0.05, 0.0, 600.0, 1.0, 1.0, 580.0, 600.0, 0.125)
    o = po.Oracle(N=300, p=7, L=L, rmax=600.0, mask_rmin=580.0, mask_rmax=600.0)
    off = st.get(qpc_gpu.SETUP_Y10)[:L - 1]
    Y = np.diag(off, 1) + np.diag(off, -1)
    e = st.get(qpc_gpu.SETUP_YEIG)
    A = st.get(qpc_gpu.SETUP_YPRJ).reshape(L, L)
    B = o.get(po.QOR_YPRJ).reshape(L, L)
    eps = np.finfo(float).eps
    worst = 0.0
    for V in (A, B):
        assert np.abs(V @ V.T - np.eye(L)).max() < 1e-13
        assert np.abs(V @ Y @ V.T - np.diag(e)).max() < 1e-13
    for k in range(L):
        gap = np.min(np.abs(np.delete(e, k) - e[k]))
        d = min(np.linalg.norm(A[k] - B[k]), np.linalg.norm(A[k] + B[k]))
        bound = 50 * eps * np.linalg.norm(Y, 2) / gap
        worst = max(worst, d / bound)
        assert d <= bound, (k, d, bound)
    assert np.abs(np.abs(A) - np.abs(B)).max() > 1e-14  # the difference is real, not rounding of a copy
