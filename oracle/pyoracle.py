"""ctypes binding of the CPU oracle (oracle/qpc_oracle.h).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product library.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libqpc_oracle.so")

QOR_S, QOR_CHOL, QOR_H, QOR_RINV, QOR_RDIF, QOR_RPOW1, QOR_MASK, QOR_UPPER, QOR_LOWER, \
    QOR_YEIG, QOR_YPRJ, QOR_QEIG, QOR_QPRJ, QOR_MERGED, QOR_KNOTS, QOR_GSX, QOR_GSW, \
    QOR_Y10, QOR_Q10, QOR_DVDR = range(20)
_COMPLEX = {QOR_UPPER, QOR_LOWER, QOR_QPRJ, QOR_MERGED}


class Dims(C.Structure):
    _fields_ = [("n_ndim", C.c_int64), ("n_rank", C.c_int64), ("n_ldim", C.c_int64),
                ("n_mdim", C.c_int64), ("n_mmin", C.c_int64),
                ("para_rmin", C.c_double), ("para_rmax", C.c_double),
                ("para_zc", C.c_double), ("para_me", C.c_double), ("para_dt", C.c_double),
                ("mask_rmin", C.c_double), ("mask_rmax", C.c_double), ("mask_fact", C.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.qor_create.argtypes = [C.POINTER(Dims), C.POINTER(C.c_void_p)]
        L.qor_destroy.argtypes = [C.c_void_p]
        L.qor_size.argtypes = [C.c_void_p, C.c_int]
        L.qor_size.restype = C.c_int64
        L.qor_get.argtypes = [C.c_void_p, C.c_int, dp]
        L.qor_set.argtypes = [C.c_void_p, C.c_int, dp]
        L.qor_lu_info.argtypes = [C.c_void_p]
        L.qor_set_threads.argtypes = [C.c_int]
        L.qor_eigenstate.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_int64, dp, dp]
        L.qor_projection_basis.argtypes = [C.c_void_p, C.c_int64, dp, dp]
        L.qor_projection.argtypes = [C.c_void_p, dp, C.c_int64, dp, dp, C.c_int64]
        L.qor_krylov_step.argtypes = [C.c_void_p, dp, C.c_double, C.c_double, C.c_int64, C.c_double, C.c_double,
                                      C.POINTER(C.c_int64), dp, dp]
        L.qor_step_vgwz.argtypes = [C.c_void_p, dp, C.c_double, C.c_double]
        L.qor_step_free.argtypes = [C.c_void_p, dp]
        L.qor_step_lgwz.argtypes = [C.c_void_p, dp, C.c_double, C.c_double]
        L.qor_run_lg.argtypes = [C.c_void_p, dp, dp, C.c_int64, C.c_int64, C.c_int64, C.c_int]
        L.qor_set_ep.argtypes = [C.c_void_p, C.c_int] + [dp] * 9
        L.qor_mlp_forward.argtypes = [dp, dp, C.c_int64, dp]
        L.qor_mlp_loss_grad.argtypes = [dp, dp, dp, C.c_int64, dp, dp]
        L.qor_step_vgwr.argtypes = [C.c_void_p, dp, C.c_double, C.c_double, C.c_double, C.c_double]
        L.qor_run_ep.argtypes = [C.c_void_p, dp, dp, dp, C.c_int64, C.c_int64, C.c_int64, C.c_int]
        L.qor_tsurff_accumulate_xy.argtypes = [C.c_void_p, dp] + [C.c_double] * 5 + [C.c_int64] * 3 + [dp] * 3 + \
            [C.c_int64, C.POINTER(C.c_int64)] + [dp] * 10
        L.qor_set_lgwr.argtypes = [C.c_void_p, C.c_int] + [dp] * 5
        L.qor_step_lgwr.argtypes = [C.c_void_p, dp, C.c_double, C.c_double, C.c_double, C.c_double]
        L.qor_run_lgwr.argtypes = [C.c_void_p, dp, dp, dp, C.c_int64, C.c_int64, C.c_int64, C.c_int]
        L.qor_mask.argtypes = [C.c_void_p, dp]
        L.qor_run.argtypes = [C.c_void_p, dp, dp, C.c_int64, C.c_int64, C.c_int64, C.c_int, dp]
        L.qor_run_record.argtypes = [C.c_void_p, dp, dp, C.c_int64, C.c_int64, C.c_int64, C.c_int, C.c_double, dp]
        L.qor_norm.argtypes = [C.c_void_p, dp]
        L.qor_norm.restype = C.c_double
        L.qor_outer_norm.argtypes = [C.c_void_p, dp, C.c_double]
        L.qor_outer_norm.restype = C.c_double
        L.qor_dipole_xy.argtypes = [C.c_void_p, dp, C.c_int, dp, dp]
        L.qor_measure_lmd.argtypes = [C.c_void_p, dp, C.c_int64, dp, dp, dp]
        L.qor_tsurff_accumulate.argtypes = [C.c_void_p, dp, C.c_double, C.c_double, C.c_double, C.c_int64,
                                            C.c_int64, C.c_int64, dp, dp, C.c_int64, C.POINTER(C.c_int64), dp,
                                            dp, dp, dp, dp, dp, dp, dp, dp]
        L.qor_dipole_z.argtypes = [C.c_void_p, dp, C.c_int]
        L.qor_dipole_z.restype = C.c_double
        L.qor_sinen2.argtypes = [C.c_double] * 6
        L.qor_sinen2.restype = C.c_double
        L.qor_para_nt.argtypes = [C.c_double] * 4
        L.qor_para_nt.restype = C.c_int64
        L.qor_lu_zsb.argtypes = [C.c_int64, C.c_int64, dp]
        L.qor_cn_kernel.argtypes = [C.c_int64, C.c_int64, C.c_int, dp, dp, C.c_double, C.c_double, dp]
        L.qor_symb_mul.argtypes = [C.c_int64, C.c_int64, dp, dp, dp]
        L.qor_ldl_solve.argtypes = [C.c_int64, C.c_int64, dp, dp]
        L.qor_chol.argtypes = [C.c_int64, C.c_int64, dp]
        L.qor_tridiag_eig.argtypes = [C.c_int64, dp, dp, dp, dp]
        L.qor_gauss_legendre.argtypes = [C.c_int64, dp, dp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a):
    a = np.ascontiguousarray(a)
    if np.iscomplexobj(a):
        return np.ascontiguousarray(a.astype(np.complex128)).view(np.float64)
    return a.astype(np.float64)


class Oracle:
    """Owns a qor_ctx: setup arrays + the reference-order propagator."""

    def __init__(self, N, p, L, M=1, m0=0, rmin=0.0, rmax=200.0, zc=1.0, me=1.0, dt=0.05,
                 mask_rmin=None, mask_rmax=None, mask_fact=0.125):
        self.L_ = lib()
        if mask_rmin is None:
            mask_rmin = rmax - 20.0
        if mask_rmax is None:
            mask_rmax = rmax
        self.dims = Dims(N, p, L, M, m0, rmin, rmax, zc, me, dt, mask_rmin, mask_rmax, mask_fact)
        h = C.c_void_p()
        rc = self.L_.qor_create(C.byref(self.dims), C.byref(h))
        if rc != 0:
            raise RuntimeError(f"qor_create failed: {rc}")
        self.h = h
        self.N, self.p, self.L, self.M, self.m0 = N, p, L, M, m0
        self.ne = p + 1
        self.lmax = max(abs(m0), abs(m0 + M - 1)) + L - 1

    def __del__(self):
        if getattr(self, "h", None):
            self.L_.qor_destroy(self.h)
            self.h = None

    def get(self, what):
        n = self.L_.qor_size(self.h, what)
        out = np.empty(n, np.float64)
        if self.L_.qor_get(self.h, what, _p(out)) != 0:
            raise RuntimeError("qor_get failed")
        if what in _COMPLEX:
            return out.view(np.complex128)
        return out

    def set(self, what, arr):
        a = _f64(arr)
        n = self.L_.qor_size(self.h, what)
        assert a.size == n, (a.size, n)
        if self.L_.qor_set(self.h, what, _p(a)) != 0:
            raise RuntimeError("qor_set failed")

    def eigenstate(self, im, il, n):
        c = np.zeros(self.M * self.L * self.N, np.complex128)
        e = np.zeros(1)
        rc = self.L_.qor_eigenstate(self.h, im, il, n, _p(e), _p(c.view(np.float64)))
        if rc != 0:
            raise RuntimeError(f"qor_eigenstate failed: {rc}")
        return float(e[0]), c

    def projection_basis(self, nsel):
        """(W, E): W[l][n][ir] = (S v_{n,l})_ir, E[l][n], for the nsel lowest states of every l."""
        nl = self.lmax + 1
        W = np.zeros((nl, nsel, self.N))
        E = np.zeros((nl, nsel))
        if self.L_.qor_projection_basis(self.h, nsel, _p(W), _p(E)) != 0:
            raise RuntimeError("qor_projection_basis failed")
        return W, E

    def projection(self, c, W, stride=None):
        nsel = W.shape[1]
        stride = nsel if stride is None else stride
        out = np.zeros(self.M * self.L * stride, np.complex128)
        cc = np.ascontiguousarray(c, np.complex128)
        Wc = np.ascontiguousarray(W, np.float64)
        if self.L_.qor_projection(self.h, _p(cc.view(np.float64)), nsel, _p(Wc), _p(out.view(np.float64)),
                                    stride) != 0:
            raise RuntimeError("qor_projection failed")
        return out

    def krylov_step(self, c, az, dt, kmax=40, emin=1e-14, emax=1e-10):
        """In place: c <- exp(-i dt S^-1 H(az)) c (Lanczos); returns (criterion, m, eps, ab)."""
        m = C.c_int64(0)
        eps = np.zeros(1)
        ab = np.zeros(2 * kmax)
        rc = self.L_.qor_krylov_step(self.h, _p(c.view(np.float64)), az, dt, kmax, emin, emax, C.byref(m), _p(eps),
                                     _p(ab))
        if rc < 0:
            raise RuntimeError(f"qor_krylov_step failed: {rc}")
        return rc, int(m.value), float(eps[0]), ab[:2 * m.value].reshape(-1, 2)

    def step(self, c, a_i, a_f):
        assert c.dtype == np.complex128 and c.flags.c_contiguous
        self.L_.qor_step_vgwz(self.h, _p(c.view(np.float64)), a_i, a_f)

    def set_ep(self, ep, isub=-1):
        """Angular eigen-data of propagator_cn_hc_vgwr (a dict from ep_angular)."""
        self._ep = [np.ascontiguousarray(ep[k], np.float64) for k in ("eig_x1", "eig_x2", "eig_y1", "eig_y2",
                                                                      "tran_x1", "tran_y2")]
        self._ep += [np.ascontiguousarray(ep[k], np.complex128).view(np.float64)
                     for k in ("x1_2_x2", "x2_2_y1", "y1_2_y2")]
        if self.L_.qor_set_ep(self.h, isub, *[_p(a) for a in self._ep]) != 0:
            raise RuntimeError("qor_set_ep failed")

    def step_ep(self, c, ax_i, ax_f, ay_i, ay_f):
        assert c.dtype == np.complex128 and c.flags.c_contiguous
        if self.L_.qor_step_vgwr(self.h, _p(c.view(np.float64)), ax_i, ax_f, ay_i, ay_f) != 0:
            raise RuntimeError("qor_step_vgwr failed")

    def run_ep(self, c, axtab, aytab, nt_field, it0, it1, use_mask=True):
        axtab = np.ascontiguousarray(axtab, np.float64)
        aytab = np.ascontiguousarray(aytab, np.float64)
        if self.L_.qor_run_ep(self.h, _p(c.view(np.float64)), _p(axtab), _p(aytab), nt_field, it0, it1,
                              1 if use_mask else 0) != 0:
            raise RuntimeError("qor_run_ep failed")

    def set_lgwr(self, lg, isub=-1):
        """Angular eigen-data of propagator_cn_hc_lgwr (a dict from lgwr_angular)."""
        self._lr = [np.ascontiguousarray(lg[k], np.float64) for k in ("eig_x1", "eig_y1", "tran_x1")]
        self._lr += [np.ascontiguousarray(lg[k], np.complex128).view(np.float64) for k in ("x1_2_y1", "tran_y1")]
        if self.L_.qor_set_lgwr(self.h, isub, *[_p(a) for a in self._lr]) != 0:
            raise RuntimeError("qor_set_lgwr failed")

    def step_lgwr(self, c, ex_i, ex_f, ey_i, ey_f):
        assert c.dtype == np.complex128 and c.flags.c_contiguous
        if self.L_.qor_step_lgwr(self.h, _p(c.view(np.float64)), ex_i, ex_f, ey_i, ey_f) != 0:
            raise RuntimeError("qor_step_lgwr failed")

    def run_lgwr(self, c, extab, eytab, nt_field, it0, it1, use_mask=True):
        extab = np.ascontiguousarray(extab, np.float64)
        eytab = np.ascontiguousarray(eytab, np.float64)
        if self.L_.qor_run_lgwr(self.h, _p(c.view(np.float64)), _p(extab), _p(eytab), nt_field, it0, it1,
                                1 if use_mask else 0) != 0:
            raise RuntimeError("qor_run_lgwr failed")

    def step_lg(self, c, e_i, e_f):
        """propagator_cn_hc_lgwz::propagate (length gauge, z polarisation)."""
        self.L_.qor_step_lgwz(self.h, _p(c.view(np.float64)), e_i, e_f)

    def run_lg(self, c, etab, nt_field, it0, it1, use_mask=True):
        etab = np.ascontiguousarray(etab, np.float64)
        self.L_.qor_run_lg(self.h, _p(c.view(np.float64)), _p(etab), nt_field, it0, it1, 1 if use_mask else 0)

    def step_free(self, c):
        self.L_.qor_step_free(self.h, _p(c.view(np.float64)))

    def mask(self, c):
        self.L_.qor_mask(self.h, _p(c.view(np.float64)))

    def run(self, c, atab, nt_field, it0, it1, use_mask=True, observe=False):
        atab = np.ascontiguousarray(atab, np.float64)
        obs = np.zeros(2 * (it1 - it0)) if observe else None
        self.L_.qor_run(self.h, _p(c.view(np.float64)), _p(atab), nt_field, it0, it1,
                        1 if use_mask else 0, _p(obs) if observe else None)
        return obs.reshape(-1, 2) if observe else None

    def run_record(self, c, atab, nt_field, it0, it1, r0, use_mask=True):
        """qor_run_record: the run, with {norm, <z>, z acceleration, outer norm
        over r >= r0, <x>, <y>, x and y acceleration} observed before every
        step (rows of the returned array)."""
        atab = np.ascontiguousarray(atab, np.float64)
        obs = np.zeros(8 * (it1 - it0))
        self.L_.qor_run_record(self.h, _p(c.view(np.float64)), _p(atab), nt_field, it0, it1,
                               1 if use_mask else 0, r0, _p(obs))
        return obs.reshape(-1, 8)

    def norm(self, c):
        return self.L_.qor_norm(self.h, _p(c.view(np.float64)))

    def outer_norm(self, c, r0):
        return self.L_.qor_outer_norm(self.h, _p(c.view(np.float64)), r0)

    def measure_lmd(self, c, ijl, cps):
        """spectrum_c::measure_lmd: lmd[ikr][im][il] (complex) from ijl[ikr][l][j], cps[ikr][l]."""
        ijl = np.ascontiguousarray(ijl, np.float64)
        cps = np.ascontiguousarray(cps, np.complex128)
        nkr = cps.shape[0]
        out = np.zeros(nkr * self.M * self.L, np.complex128)
        self.L_.qor_measure_lmd(self.h, _p(np.ascontiguousarray(c).view(np.float64)), nkr, _p(ijl),
                                _p(cps.view(np.float64)), _p(out.view(np.float64)))
        return out.reshape(nkr, self.M, self.L)

    def tsurff_accumulate(self, c, az0, az1, dt, t, state, work):
        """spectrum_c_tsf::accumulate on state c (complex) with the tables t (a dict as
        qpc_gpu.Setup.tsurff returns); state = [time, disz] and work (complex
        [nkr][nkt][nkp]) are updated in place."""
        assert state.dtype == np.float64 and work.dtype == np.complex128 and work.flags.c_contiguous
        nkr, nkt, nkp = len(t["kr"]), len(t["kt"]), t["e0"].shape[0]
        ib = np.ascontiguousarray(t["ib"], np.int64)
        rc = self.L_.qor_tsurff_accumulate(
            self.h, _p(np.ascontiguousarray(c).view(np.float64)), az0, az1, dt, nkr, nkt, nkp, _p(t["kr"]),
            _p(t["kt"]), len(ib), ib.ctypes.data_as(C.POINTER(C.c_int64)), _p(t["b0"]), _p(t["br"]),
            _p(np.ascontiguousarray(t["j0"])), _p(np.ascontiguousarray(t["jr"])), _p(np.ascontiguousarray(t["y0"])),
            _p(np.ascontiguousarray(t["e0"]).view(np.float64)), _p(np.ascontiguousarray(t["pmat0"])), _p(state),
            _p(work.view(np.float64)))
        assert rc == 0

    def tsurff_accumulate_xy(self, c, ax0, ax1, ay0, ay1, dt, t, pmat1, pmat2, state, work):
        """spectrum_c_tsf::accumulate (xy) on state c with the tables t and the
        sin(th) e^(-i ph) elements pmat1/pmat2 [M][L]; state = [time, disx, disy]
        and work are updated in place."""
        assert state.dtype == np.float64 and work.dtype == np.complex128 and work.flags.c_contiguous
        nkr, nkt, nkp = len(t["kr"]), len(t["kt"]), t["e0"].shape[0]
        ib = np.ascontiguousarray(t["ib"], np.int64)
        rc = self.L_.qor_tsurff_accumulate_xy(
            self.h, _p(np.ascontiguousarray(c).view(np.float64)), ax0, ax1, ay0, ay1, dt, nkr, nkt, nkp,
            _p(t["kr"]), _p(t["kt"]), _p(np.ascontiguousarray(t["kp"], np.float64)), len(ib),
            ib.ctypes.data_as(C.POINTER(C.c_int64)), _p(t["b0"]), _p(t["br"]),
            _p(np.ascontiguousarray(t["j0"])), _p(np.ascontiguousarray(t["jr"])), _p(np.ascontiguousarray(t["y0"])),
            _p(np.ascontiguousarray(t["e0"]).view(np.float64)), _p(np.ascontiguousarray(pmat1, np.float64)),
            _p(np.ascontiguousarray(pmat2, np.float64)), _p(state), _p(work.view(np.float64)))
        assert rc == 0

    def dipole_xy(self, c, which=0):
        x, y = np.zeros(1), np.zeros(1)
        self.L_.qor_dipole_xy(self.h, _p(c.view(np.float64)), which, _p(x), _p(y))
        return float(x[0]), float(y[0])

    def dipole_z(self, c, which=0):
        return self.L_.qor_dipole_z(self.h, _p(c.view(np.float64)), which)

    @property
    def lu_info(self):
        return self.L_.qor_lu_info(self.h)


def set_threads(n=0):
    """OpenMP threads of the oracle's later calls (n <= 0: leave them); returns the count."""
    return lib().qor_set_threads(int(n))


def sinen2(e0, w0, nc, n0, c0, t):
    return lib().qor_sinen2(e0, w0, nc, n0, c0, t)


def para_nt(w0, nc, n0, dt):
    return lib().qor_para_nt(w0, nc, n0, dt)


def lu_zsb(band, n, p):
    b = np.ascontiguousarray(band, np.complex128).copy()
    info = lib().qor_lu_zsb(n, p, _p(b.view(np.float64)))
    return b, info


def cn_kernel(n, p, symflag, smat, rmat, ifet, z):
    z = np.ascontiguousarray(z, np.complex128).copy()
    s = np.ascontiguousarray(smat, np.float64)
    r = np.ascontiguousarray(rmat, np.float64)
    lib().qor_cn_kernel(n, p, symflag, _p(s), _p(r), float(np.real(ifet)), float(np.imag(ifet)),
                        _p(z.view(np.float64)))
    return z


def symb_mul(n, p, cband, x):
    cb = np.ascontiguousarray(cband, np.complex128)
    x = np.ascontiguousarray(x, np.complex128)
    y = np.zeros(n, np.complex128)
    lib().qor_symb_mul(n, p, _p(cb.view(np.float64)), _p(x.view(np.float64)), _p(y.view(np.float64)))
    return y


def ldl_solve(n, p, factor, x):
    f = np.ascontiguousarray(factor, np.complex128)
    x = np.ascontiguousarray(x, np.complex128).copy()
    lib().qor_ldl_solve(n, p, _p(f.view(np.float64)), _p(x.view(np.float64)))
    return x


def chol(n, p, band):
    b = np.ascontiguousarray(band, np.float64).copy()
    info = lib().qor_chol(n, p, _p(b))
    return b, info


def tridiag_eig(diag, off):
    n = len(diag)
    d = np.ascontiguousarray(diag, np.float64)
    o = np.zeros(n)
    o[:len(off)] = off
    e = np.zeros(n)
    z = np.zeros(n * n)
    rc = lib().qor_tridiag_eig(n, _p(d), _p(o), _p(e), _p(z))
    if rc != 0:
        raise RuntimeError("tridiag_eig failed")
    return e, z.reshape(n, n)


def gauss_legendre(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().qor_gauss_legendre(n, _p(x), _p(w))
    return x, w


def band_to_dense(band, n, p, sym=1):
    """Reference band layout band[i*ne+j] = A(i,i+j) -> dense; sym=+1 symmetric, -1 antisymmetric."""
    ne = p + 1
    b = np.asarray(band).reshape(n, ne)
    A = np.zeros((n, n), dtype=b.dtype)
    for i in range(n):
        for j in range(ne):
            if i + j < n:
                A[i, i + j] = b[i, j]
                if j > 0:
                    A[i + j, i] = sym * b[i, j]
    return A


# ---------------------------------------------------------------------------
# Angular eigen-data of propagator_cn_hc_vgwr (cn_1e.hpp:850-896)

def ep_closed_forms(m, l):
    """Closed forms of integrator_angu (utilities/basisfunc.hpp:243-302) for the
    (m, l) -> (m+1, l+1) "outer" and (m+1, l-1) "inner" couplings:
    (sinexp1, sinexp2, coslm1, coslm2)."""
    def q(a, b):
        return a / b if b != 0 else 0.0
    se1 = -np.sqrt(q((l + m + 2) * (l + m + 1), (2 * l + 1) * (2 * l + 3)))
    se2 = +np.sqrt(q((l - m) * (l - m - 1), (2 * l + 1) * (2 * l - 1)))
    cl1 = -np.sqrt(q((l + m + 1) * (l + m + 2), (2 * l + 1) * (2 * l + 3))) * (l + 1)
    cl2 = -np.sqrt(q((l - m) * (l - m - 1), (2 * l + 1) * (2 * l - 1))) * l
    return se1, se2, cl1, cl2


def ep_matrices(M, L, m0):
    """The four Hermitian (M*L)^2 angular matrices that operator_y11/q11::
    solve_eigen (operators/1e.hpp:363-405, 478-522) hand to zhbev, with the
    lower panel A[(m+1, l+-1), (m, l)] = value * mult and its conjugate above:
    x1 = Y11 (mult 1), x2 = Q11 (mult -i), y1 = Y11 (mult -i), y2 = Q11
    (mult 1); Q11<-1> stores minus the closed form (:500-513)."""
    ML = M * L
    out = {k: np.zeros((ML, ML), np.complex128) for k in ("x1", "x2", "y1", "y2")}
    mult = {"x1": 1.0, "x2": -1j, "y1": -1j, "y2": 1.0}
    for im in range(M):
        m = m0 + im
        for il in range(L):
            l = il + abs(m)
            i = im * L + il
            if im + 1 >= M:
                continue
            se1, se2, cl1, cl2 = ep_closed_forms(m, l)
            m2 = m + 1
            for lsh, vy, vq in ((+1, se1, cl1), (-1, se2, cl2)):
                jl = l + lsh - abs(m2)  # dimension_1e::in_check (dimension/1e.hpp:93-97)
                if jl < 0 or jl >= L:
                    continue
                j = (im + 1) * L + jl
                for k in out:
                    v = (vy if k in ("x1", "y1") else -vq) * mult[k]
                    out[k][j, i] = v
                    out[k][i, j] = np.conj(v)
    return out


def ep_eigen(A, M, L, real):
    """zhbev + converter::catagorize (utilities/converter.hpp:27-107): eigenvalues
    and row-major eigenvectors (row k = eigenvector k, as zhbev's column-major
    z read row-major) ordered even / zero / odd.  A conserves the parity of
    il, so each parity block is diagonalised on its own (ascending order
    within a category, as zhbev's ascending output filtered by catagorize)."""
    ML = M * L
    par = np.array([il % 2 for im in range(M) for il in range(L)])
    vals, vecs, kind = [], [], []
    for pty in (0, 1):
        idx = np.nonzero(par == pty)[0]
        sub = A[np.ix_(idx, idx)]
        w, v = np.linalg.eigh(sub.real if real else sub)
        for k in range(len(w)):
            full = np.zeros(ML, np.float64 if real else np.complex128)
            full[idx] = v[:, k]
            vals.append(w[k])
            vecs.append(full)
            kind.append("zero" if (M % 2 == 1 and abs(w[k]) < 1e-8) else ("even" if pty == 0 else "odd"))
    n_zero = L if M % 2 == 1 else 0
    n_even = (M // 2 * 2) * ((L + 1) // 2)
    n_odds = (M // 2 * 2) * (L // 2)
    order = []
    for cat in ("even", "zero", "odd"):
        sel = [k for k in range(ML) if kind[k] == cat]
        sel.sort(key=lambda k: vals[k])
        order.append(sel)
    if (len(order[0]), len(order[1]), len(order[2])) != (n_even, n_zero, n_odds):
        raise RuntimeError("catagorize: unexpected category sizes")
    perm = order[0] + order[1] + order[2]
    return np.array([vals[k] for k in perm]), np.array([vecs[k] for k in perm])


def ep_angular(M, L, m0):
    """Operands of propagator_cn_hc_vgwr::initialize (cn_1e.hpp:878-895):
    eigenvalues, tran_x1 / tran_y2 (real parts of the unconjugated
    projectors), and the merge_two products P_b P_a^H of the conjugated
    projectors P = conj(rows of eigenvectors)."""
    A = ep_matrices(M, L, m0)
    out = {"mat": A}
    prj = {}
    for k in ("x1", "x2", "y1", "y2"):
        e, z = ep_eigen(A[k], M, L, real=k in ("x1", "y2"))
        out["eig_" + k] = e
        prj[k] = z.astype(np.complex128)
    out["tran_x1"] = prj["x1"].real.copy()
    out["tran_y2"] = prj["y2"].real.copy()
    P = {k: np.conj(v) for k, v in prj.items()}
    out["x1_2_x2"] = P["x2"] @ P["x1"].conj().T
    out["x2_2_y1"] = P["y1"] @ P["x2"].conj().T
    out["y1_2_y2"] = P["y2"] @ P["y1"].conj().T
    out["P"] = P
    return out


def lgwr_angular(M, L, m0):
    """Operands of propagator_cn_hc_lgwr::initialize (cn_1e.hpp:745-765): the
    eigen-data of Y11 (x1, mult 1) and Y11 with mult -i (y1), tran_x1 (real
    part of the unconjugated x1 projector), tran_y1 = U[y1] and x1_2_y1 =
    U[y1] U[x1]^H with U = conj(rows of eigenvectors)."""
    A = ep_matrices(M, L, m0)
    out = {"mat": A}
    e1, z1 = ep_eigen(A["x1"], M, L, real=True)
    e2, z2 = ep_eigen(A["y1"], M, L, real=False)
    out["eig_x1"], out["eig_y1"] = e1, e2
    out["tran_x1"] = z1.real.copy()
    Ux, Uy = np.conj(z1.astype(np.complex128)), np.conj(z2.astype(np.complex128))
    out["tran_y1"] = Uy
    out["x1_2_y1"] = Uy @ Ux.conj().T
    return out


# ---------------------------------------------------------------------------
# The learned Krylov controller's network in FP64 (the checker of the device
# kernels of include/qpc_mlp.h)
MLP_NPARAM = 4802


def mlp_forward(params, X):
    P = np.ascontiguousarray(params, np.float64)
    X = np.ascontiguousarray(X, np.float64)
    Y = np.zeros((X.shape[0], 2))
    if lib().qor_mlp_forward(_p(P), _p(X), X.shape[0], _p(Y)) != 0:
        raise RuntimeError("qor_mlp_forward failed")
    return Y


def mlp_loss_grad(params, X, T):
    P = np.ascontiguousarray(params, np.float64)
    X = np.ascontiguousarray(X, np.float64)
    T = np.ascontiguousarray(T, np.float64)
    loss = np.zeros(1)
    G = np.zeros(MLP_NPARAM)
    if lib().qor_mlp_loss_grad(_p(P), _p(X), _p(T), X.shape[0], _p(loss), _p(G)) != 0:
        raise RuntimeError("qor_mlp_loss_grad failed")
    return float(loss[0]), G
