"""ctypes binding of libqpc_gpu.so (include/qpc_gpu.h).

Mirrors the member API of the reference propagator object the driver uses
(propagator_cn_hc_vgwz, reference propagator/cn_1e.hpp:621-715, and
absorber_mask, absorber.hpp:52-64).  The library is the only compute path:
if it is missing, importing this module raises, there is no fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libqpc_gpu.so")

QPC_OBS_NORM = 0
QPC_OBS_DIPOLE_Z = 1
QPC_OBS_OUTER = 2
QPC_OBS_ACCEL_Z = 3
QPC_OBS_DIPOLE_X, QPC_OBS_DIPOLE_Y, QPC_OBS_ACCEL_X, QPC_OBS_ACCEL_Y = 4, 5, 6, 7
REC_NORM, REC_DIPOLE, REC_ACCEL, REC_OUTER, REC_DIPOLE_XY, REC_ACCEL_XY, REC_ALL = 1, 2, 4, 8, 16, 32, 63
REC_COLUMNS = 8  # norm, <z>, z acceleration, outer norm, <x>, <y>, x and y acceleration
K_CN_D, K_CN_R, K_FREE, K_ROT, K_MASK, K_FACTOR, K_PROJ, K_KRYLOV, K_EP_ROT, K_EP_SWEEP = range(10)
_EP_KEYS = ("eig_x1", "eig_x2", "eig_y1", "eig_y2", "tran_x1", "tran_y2", "x1_2_x2", "x2_2_y1", "y1_2_y2")

(SETUP_S, SETUP_CHOL, SETUP_H, SETUP_RINV, SETUP_RDIF, SETUP_RPOW1, SETUP_MASK, SETUP_UPPER, SETUP_LOWER,
 SETUP_YEIG, SETUP_YPRJ, SETUP_QEIG, SETUP_QPRJ, SETUP_MERGED, SETUP_Y10, SETUP_DVDR, SETUP_Q10,
 SETUP_SOUTER, SETUP_KNOTS, SETUP_HV, SETUP_HK, SETUP_HT) = range(22)
_SETUP_COMPLEX = {SETUP_UPPER, SETUP_LOWER, SETUP_QPRJ, SETUP_MERGED}


class Dims(C.Structure):
    _fields_ = [("n_ndim", C.c_int64), ("n_rank", C.c_int64), ("n_ldim", C.c_int64),
                ("n_mdim", C.c_int64), ("n_mmin", C.c_int64), ("para_dt", C.c_double)]


class TsurffCfg(C.Structure):
    _fields_ = [("r0", C.c_double), ("zc", C.c_double), ("krmin", C.c_double), ("krmax", C.c_double),
                ("ktmin", C.c_double), ("ktmax", C.c_double), ("kpmin", C.c_double), ("kpmax", C.c_double),
                ("nkr", C.c_int64), ("nkt", C.c_int64), ("nkp", C.c_int64), ("flag", C.c_int)]


class Config(C.Structure):
    _fields_ = [("dims", Dims), ("para_rmin", C.c_double), ("para_rmax", C.c_double),
                ("para_zc", C.c_double), ("para_me", C.c_double), ("mask_rmin", C.c_double),
                ("mask_rmax", C.c_double), ("mask_fact", C.c_double)]


class QpcError(RuntimeError):
    pass


class Sinen2(C.Structure):
    """A(t) of field_component_sinen<2> (reference field.h:90-133)."""
    _fields_ = [("f0", C.c_double), ("w0", C.c_double), ("nc", C.c_double), ("n0", C.c_double),
                ("c0", C.c_double)]


class Tune(C.Structure):
    _fields_ = [("seg_add", C.c_int32), ("serial", C.c_int32), ("stage_grid", C.c_int64), ("share_r", C.c_int32),
                ("tables_r", C.c_int32)]


TUNE_CANDIDATES = [(k % 4, (512, 4096)[k // 4]) for k in range(8)]  # (seg_add, stage_grid) of qpc_autotune's ms[k]


class KrylovStats(C.Structure):
    _fields_ = [("steps", C.c_int64), ("rejected", C.c_int64), ("exponentials", C.c_int64),
                ("vectors", C.c_int64), ("worst_estimate", C.c_double)]


class KrylovRecord(C.Structure):
    _fields_ = [("cap", C.c_int64), ("n", C.c_int64), ("features", C.POINTER(C.c_float)),
                ("h", C.POINTER(C.c_double)), ("m", C.POINTER(C.c_int64)), ("ab", C.POINTER(C.c_double))]


def _load():
    if not os.path.exists(LIB_PATH):
        raise QpcError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    L = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    fp = C.POINTER(C.c_float)
    vp = C.c_void_p
    sig = {
        "qpc_create": ([C.POINTER(Dims), C.POINTER(vp)], C.c_int),
        "qpc_destroy": ([vp], C.c_int),
        "qpc_last_error": ([vp], C.c_char_p),
        "qpc_upload_radial": ([vp, dp, dp, dp, dp, dp, dp], C.c_int),
        "qpc_upload_hc": ([vp, dp, dp], C.c_int),
        "qpc_upload_angular": ([vp, dp, dp, dp, dp, dp, dp], C.c_int),
        "qpc_set_state": ([vp, dp], C.c_int),
        "qpc_get_state": ([vp, dp], C.c_int),
        "qpc_step": ([vp, C.c_double, C.c_double], C.c_int),
        "qpc_step_free": ([vp], C.c_int),
        "qpc_step_lg": ([vp, C.c_double, C.c_double], C.c_int),
        "qpc_run_lg": ([vp, dp, C.c_int64, C.c_int64, C.c_int64, C.c_int], C.c_int),
        "qpc_mask": ([vp], C.c_int),
        "qpc_run": ([vp, dp, C.c_int64, C.c_int64, C.c_int64, C.c_int], C.c_int),
        "qpc_sync": ([vp], C.c_int),
        "qpc_observe": ([vp, C.c_int, dp], C.c_int),
        "qpc_setup_build": ([C.POINTER(Config), C.POINTER(vp)], C.c_int),
        "qpc_setup_free": ([vp], None),
        "qpc_setup_size": ([vp, C.c_int], C.c_int64),
        "qpc_setup_get": ([vp, C.c_int, dp], C.c_int),
        "qpc_setup_lu_info": ([vp], C.c_int),
        "qpc_setup_eigenstate": ([vp, C.c_int64, C.c_int64, C.c_int64, dp, dp], C.c_int),
        "qpc_upload_setup": ([vp, vp], C.c_int),
        "qpc_timing_begin": ([vp, C.c_int, C.c_int64], C.c_int),
        "qpc_set_tune": ([vp, C.POINTER(Tune)], C.c_int),
        "qpc_get_tune": ([vp, C.POINTER(Tune)], C.c_int),
        "qpc_autotune": ([vp, dp, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.POINTER(Tune), dp], C.c_int),
        "qpc_upload_projection": ([vp, C.c_int64, dp], C.c_int),
        "qpc_project": ([vp, dp, C.c_int64], C.c_int),
        "qpc_upload_pcs": ([vp, C.c_int64, dp, dp], C.c_int),
        "qpc_measure_lmd": ([vp, dp], C.c_int),
        "qpc_setup_projection_basis": ([vp, C.c_int64, dp, dp], C.c_int),
        "qpc_setup_pcs": ([vp, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double, dp, dp, dp], C.c_int),
        "qpc_coulomb_f": ([C.c_int64, C.c_double, C.c_double, dp], C.c_int),
        "qpc_setup_tsurff": ([vp, C.POINTER(TsurffCfg), dp, dp, dp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), dp, dp,
                              dp, dp, dp, dp, dp], C.c_int),
        "qpc_coulomb_sigma": ([C.c_int64, C.c_double, dp], C.c_int),
        "qpc_upload_krylov": ([vp, dp, dp, dp], C.c_int),
        "qpc_krylov_step": ([vp, C.c_double, C.c_double, C.c_int64, C.c_double, C.c_double, C.POINTER(C.c_int),
                             C.POINTER(C.c_int64), dp], C.c_int),
        "qpc_timing_end": ([vp, C.POINTER(C.c_int64), dp], C.c_int),
        "qpc_upload_angular_ep": ([vp, C.c_int] + [dp] * 9, C.c_int),
        "qpc_step_ep": ([vp, C.c_double, C.c_double, C.c_double, C.c_double], C.c_int),
        "qpc_run_ep": ([vp, dp, dp, C.c_int64, C.c_int64, C.c_int64, C.c_int], C.c_int),
        "qpc_setup_angular_ep": ([C.c_int64, C.c_int64, C.c_int64] + [dp] * 9, C.c_int),
        "qpc_setup_tsurff_xy": ([vp, dp, dp], C.c_int),
        "qpc_upload_tsurff_xy": ([vp, dp, dp], C.c_int),
        "qpc_tsurff_accumulate_xy": ([vp, C.c_double, C.c_double, C.c_double, C.c_double], C.c_int),
        "qpc_tsurff_displacement": ([vp, dp], C.c_int),
        "qpc_build_hc": ([vp, dp, dp, dp, C.c_double], C.c_int),
        "qpc_build_angular_ep": ([vp, C.c_int], C.c_int),
        "qpc_build_angular_lgwr": ([vp, C.c_int], C.c_int),
        "qpc_upload_angular_lgwr": ([vp, C.c_int] + [dp] * 5, C.c_int),
        "qpc_step_lgwr": ([vp, C.c_double, C.c_double, C.c_double, C.c_double], C.c_int),
        "qpc_run_lgwr": ([vp, dp, dp, C.c_int64, C.c_int64, C.c_int64, C.c_int], C.c_int),
        "qpc_setup_angular_lgwr": ([C.c_int64, C.c_int64, C.c_int64] + [dp] * 5, C.c_int),
        "qpc_upload_outer": ([vp, dp], C.c_int),
        "qpc_upload_accel": ([vp, dp], C.c_int),
        "qpc_record": ([vp, C.c_int, C.c_int64], C.c_int),
        "qpc_record_read": ([vp, dp, C.POINTER(C.c_int64)], C.c_int),
        "qpc_observe_noalloc": ([vp, C.c_int, C.c_void_p], C.c_int),
        "qpc_upload_tsurff": ([vp, C.POINTER(TsurffCfg), dp, dp, dp, C.c_int64, C.POINTER(C.c_int64), dp, dp, dp, dp,
                               dp, dp, dp], C.c_int),
        "qpc_tsurff": ([vp, C.c_int], C.c_int),
        "qpc_tsurff_accumulate": ([vp, C.c_double, C.c_double], C.c_int),
        "qpc_tsurff_clear": ([vp], C.c_int),
        "qpc_tsurff_read": ([vp, dp, dp, dp], C.c_int),
        "qpc_coupling_reach": ([dp, dp, C.c_int64, C.c_int64, C.c_double, C.c_int, C.POINTER(C.c_int),
                                C.POINTER(C.c_int)], C.c_int),
        "qpc_coupling_plan": ([dp, dp, C.c_int64, C.c_int64, C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_int)],
                              C.c_int),
        "qpc_create_batch": ([C.POINTER(Dims), C.c_int64, C.c_int64, C.POINTER(vp)], C.c_int),
        "qpc_run_batch": ([vp, dp, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int], C.c_int),
        "qpc_observe_batch": ([vp, C.c_int, dp], C.c_int),
        "qpc_krylov_step_adaptive": ([vp, C.c_double, C.c_double, C.c_int64, C.c_int64, C.c_double,
                                      C.POINTER(C.c_int64), dp, dp], C.c_int),
        "qpc_krylov_propagate": ([vp, C.POINTER(Sinen2), C.c_double, C.c_double, C.c_double, C.c_double, C.c_int64,
                                  C.c_int64, vp, C.POINTER(KrylovStats), C.POINTER(KrylovRecord)], C.c_int),
        "qpc_mlp_create": ([C.c_uint64, C.POINTER(vp)], C.c_int),
        "qpc_mlp_destroy": ([vp], C.c_int),
        "qpc_mlp_last_error": ([vp], C.c_char_p),
        "qpc_mlp_get_params": ([vp, fp], C.c_int),
        "qpc_mlp_set_params": ([vp, fp], C.c_int),
        "qpc_mlp_forward": ([vp, fp, C.c_int64, fp], C.c_int),
        "qpc_mlp_loss_grad": ([vp, fp, fp, C.c_int64, fp, fp], C.c_int),
        "qpc_mlp_train": ([vp, fp, fp, C.c_int64, C.c_int64, C.c_int64, C.c_float, C.c_uint64, fp], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    return L


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _load()
    return _LIB


def _dp(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _as_f64(a):
    a = np.ascontiguousarray(a)
    if np.iscomplexobj(a):
        return np.ascontiguousarray(a, np.complex128).view(np.float64)
    return np.ascontiguousarray(a, np.float64)


class Setup:
    """Host-side operator set-up (qpc_setup_build): basis, bands, factors, angular eigen-data."""

    def __init__(self, n_ndim, n_rank, n_ldim, n_mdim=1, n_mmin=0, para_dt=0.05, para_rmin=0.0,
                 para_rmax=200.0, para_zc=1.0, para_me=1.0, mask_rmin=None, mask_rmax=None, mask_fact=0.125):
        self._lib = lib()
        if mask_rmin is None:
            mask_rmin = para_rmax - 20.0
        if mask_rmax is None:
            mask_rmax = para_rmax
        self.cfg = Config(Dims(n_ndim, n_rank, n_ldim, n_mdim, n_mmin, para_dt), para_rmin, para_rmax,
                          para_zc, para_me, mask_rmin, mask_rmax, mask_fact)
        h = C.c_void_p()
        rc = self._lib.qpc_setup_build(C.byref(self.cfg), C.byref(h))
        if rc != 0:
            raise QpcError(f"qpc_setup_build failed ({rc})")
        self.h = h
        self.N, self.p, self.L, self.M = n_ndim, n_rank, n_ldim, n_mdim

    def close(self):
        if getattr(self, "h", None):
            self._lib.qpc_setup_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def get(self, what):
        n = self._lib.qpc_setup_size(self.h, what)
        out = np.empty(n, np.float64)
        if self._lib.qpc_setup_get(self.h, what, _dp(out)) != 0:
            raise QpcError("qpc_setup_get failed")
        return out.view(np.complex128) if what in _SETUP_COMPLEX else out

    @property
    def lu_info(self):
        return self._lib.qpc_setup_lu_info(self.h)

    def projection_basis(self, nsel):
        """(W, E) of eigenstate_sph::initialize for the nsel lowest states of every l."""
        nl = max(abs(self.cfg.dims.n_mmin), abs(self.cfg.dims.n_mmin + self.M - 1)) + self.L
        W = np.zeros((nl, nsel, self.N))
        E = np.zeros((nl, nsel))
        rc = self._lib.qpc_setup_projection_basis(self.h, nsel, _dp(W), _dp(E))
        if rc != 0:
            raise QpcError(f"qpc_setup_projection_basis failed ({rc})")
        return W, E

    def pcs(self, nkr, krmin, krmax, zc=None, rb=0.0):
        """spectrum_c_pcs::initialize (qpc_setup_pcs): (kr[nkr], ijl[nkr][lmax+1][N], cps[nkr][lmax+1])."""
        zc = self.cfg.para_zc if zc is None else zc
        nl = max(abs(self.cfg.dims.n_mmin), abs(self.cfg.dims.n_mmin + self.M - 1)) + self.L
        kr = np.zeros(nkr)
        ijl = np.zeros((nkr, nl, self.N))
        cps = np.zeros((nkr, nl), np.complex128)
        rc = self._lib.qpc_setup_pcs(self.h, nkr, krmin, krmax, zc, rb, _dp(kr), _dp(ijl), _dp(cps.view(np.float64)))
        if rc != 0:
            raise QpcError(f"qpc_setup_pcs failed ({rc})")
        return kr, ijl, cps

    def tsurff(self, r0, krmin, krmax, nkr, ktmin=0.0, ktmax=np.pi, nkt=1, kpmin=0.0, kpmax=2 * np.pi, nkp=1,
               zc=0.0, flag=0):
        """spectrum_c_tsf::initialize (qpc_setup_tsurff): a dict of the t-SURFF tables
        (cfg, kr, kt, kp, ib, b0, br, j0, jr, y0, e0, pmat0)."""
        tc = TsurffCfg(r0, zc, krmin, krmax, ktmin, ktmax, kpmin, kpmax, nkr, nkt, nkp, flag)
        nl = max(abs(self.cfg.dims.n_mmin), abs(self.cfg.dims.n_mmin + self.M - 1)) + self.L
        p1 = self.cfg.dims.n_rank + 1
        t = dict(cfg=tc, kr=np.zeros(nkr), kt=np.zeros(nkt), kp=np.zeros(nkp), ib=np.zeros(p1, np.int64),
                 b0=np.zeros(p1), br=np.zeros(p1), j0=np.zeros((nkr, nl)), jr=np.zeros((nkr, nl)),
                 y0=np.zeros((nkt, self.M, self.L)), e0=np.zeros((nkp, self.M), np.complex128),
                 pmat0=np.zeros((self.M, self.L)))
        nb = C.c_int64(0)
        rc = self._lib.qpc_setup_tsurff(self.h, C.byref(tc), _dp(t["kr"]), _dp(t["kt"]), _dp(t["kp"]), C.byref(nb),
                                        t["ib"].ctypes.data_as(C.POINTER(C.c_int64)), _dp(t["b0"]), _dp(t["br"]),
                                        _dp(t["j0"]), _dp(t["jr"]), _dp(t["y0"]), _dp(t["e0"].view(np.float64)),
                                        _dp(t["pmat0"]))
        if rc != 0:
            raise QpcError(f"qpc_setup_tsurff failed ({rc})")
        n = nb.value
        t["ib"], t["b0"], t["br"] = t["ib"][:n].copy(), t["b0"][:n].copy(), t["br"][:n].copy()
        t["pmat1"], t["pmat2"] = np.zeros((self.M, self.L)), np.zeros((self.M, self.L))
        rc = self._lib.qpc_setup_tsurff_xy(self.h, _dp(t["pmat1"]), _dp(t["pmat2"]))
        if rc != 0:
            raise QpcError(f"qpc_setup_tsurff_xy failed ({rc})")
        return t

    def eigenstate(self, im, il, n):
        c = np.zeros(self.M * self.L * self.N, np.complex128)
        e = np.zeros(1)
        rc = self._lib.qpc_setup_eigenstate(self.h, im, il, n, _dp(e), _dp(c.view(np.float64)))
        if rc != 0:
            raise QpcError(f"qpc_setup_eigenstate failed ({rc})")
        return float(e[0]), c


def angular_ep(M, L, m0):
    """Angular eigen-data of propagator_cn_hc_vgwr (qpc_setup_angular_ep) as a
    dict: eig_x1..eig_y2 (M*L), tran_x1/tran_y2 (M*L)^2 real, x1_2_x2,
    x2_2_y1, y1_2_y2 (M*L)^2 complex, row-major."""
    ML = M * L
    out = {k: np.zeros(ML) for k in _EP_KEYS[:4]}
    out["tran_x1"] = np.zeros((ML, ML))
    out["tran_y2"] = np.zeros((ML, ML))
    for k in _EP_KEYS[6:]:
        out[k] = np.zeros((ML, ML), np.complex128)
    ptrs = [_dp(out[k].view(np.float64) if np.iscomplexobj(out[k]) else out[k]) for k in _EP_KEYS]
    rc = lib().qpc_setup_angular_ep(M, L, m0, *ptrs)
    if rc != 0:
        raise QpcError(f"qpc_setup_angular_ep failed ({rc})")
    return out


_LG_KEYS = ("eig_x1", "eig_y1", "tran_x1", "x1_2_y1", "tran_y1")


def angular_lgwr(M, L, m0):
    """Angular data of propagator_cn_hc_lgwr (qpc_setup_angular_lgwr) as a
    dict: eig_x1, eig_y1 (M*L), tran_x1 (M*L)^2 real, x1_2_y1 and tran_y1
    (M*L)^2 complex, row-major."""
    ML = M * L
    out = {"eig_x1": np.zeros(ML), "eig_y1": np.zeros(ML), "tran_x1": np.zeros((ML, ML)),
           "x1_2_y1": np.zeros((ML, ML), np.complex128), "tran_y1": np.zeros((ML, ML), np.complex128)}
    ptrs = [_dp(out[k].view(np.float64) if np.iscomplexobj(out[k]) else out[k]) for k in _LG_KEYS]
    rc = lib().qpc_setup_angular_lgwr(M, L, m0, *ptrs)
    if rc != 0:
        raise QpcError(f"qpc_setup_angular_lgwr failed ({rc})")
    return out


def coupling_reach(S, X, N, p, e, asym):
    """(fwd, bwd) reach in rows of the recurrences of S + i e X (asym 0) or S + e X (asym 1)."""
    f, b = C.c_int(0), C.c_int(0)
    S, X = _as_f64(S), _as_f64(X)
    if lib().qpc_coupling_reach(_dp(S), _dp(X), N, p, e, asym, C.byref(f), C.byref(b)) != 0:
        raise QpcError("qpc_coupling_reach failed")
    return f.value, b.value


def coupling_plan(S, D, N, p, emax):
    """(carry depth in partitions, factorisation restart in rows) of the D-part sweeps."""
    pk, rs = C.c_int(0), C.c_int(0)
    S, D = _as_f64(S), _as_f64(D)
    if lib().qpc_coupling_plan(_dp(S), _dp(D), N, p, emax, C.byref(pk), C.byref(rs)) != 0:
        raise QpcError("qpc_coupling_plan failed")
    return pk.value, rs.value


def coulomb_f(lmax, eta, rho):
    """Regular Coulomb functions F_l(eta, rho), l = 0..lmax (qpc_coulomb_f)."""
    F = np.zeros(lmax + 1)
    if lib().qpc_coulomb_f(lmax, eta, rho, _dp(F)) != 0:
        raise QpcError("qpc_coulomb_f failed")
    return F


def coulomb_sigma(lmax, eta):
    s = np.zeros(lmax + 1)
    if lib().qpc_coulomb_sigma(lmax, eta, _dp(s)) != 0:
        raise QpcError("qpc_coulomb_sigma failed")
    return s


MLP_NPARAM = 4802


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Mlp:
    """The learned Krylov controller's network (include/qpc_mlp.h): FP32
    parameters on the device, forward / backward / Adam in HIP kernels."""

    def __init__(self, seed=0):
        self._lib = lib()
        h = C.c_void_p()
        rc = self._lib.qpc_mlp_create(seed, C.byref(h))
        if rc != 0:
            raise QpcError(f"qpc_mlp_create failed ({rc})")
        self.h = h

    def _chk(self, rc, what):
        if rc != 0:
            raise QpcError(f"{what}: {self._lib.qpc_mlp_last_error(self.h).decode()} ({rc})")

    def close(self):
        if getattr(self, "h", None):
            self._lib.qpc_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def params(self):
        p = np.zeros(MLP_NPARAM, np.float32)
        self._chk(self._lib.qpc_mlp_get_params(self.h, _fp(p)), "get_params")
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == MLP_NPARAM
        self._chk(self._lib.qpc_mlp_set_params(self.h, _fp(p)), "set_params")

    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((X.shape[0], 2), np.float32)
        self._chk(self._lib.qpc_mlp_forward(self.h, _fp(X), X.shape[0], _fp(Y)), "forward")
        return Y

    def loss_grad(self, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        loss = np.zeros(1, np.float32)
        G = np.zeros(MLP_NPARAM, np.float32)
        self._chk(self._lib.qpc_mlp_loss_grad(self.h, _fp(X), _fp(T), X.shape[0], _fp(loss), _fp(G)), "loss_grad")
        return float(loss[0]), G

    def train(self, X, T, epochs, batch=256, lr=1e-3, seed=0):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        loss = np.zeros(1, np.float32)
        self._chk(self._lib.qpc_mlp_train(self.h, _fp(X), _fp(T), X.shape[0], epochs, batch, lr, seed, _fp(loss)),
                  "train")
        return float(loss[0])


class Propagator:
    """Device-resident vgwz propagator + absorber (one context, one stream)."""

    def __init__(self, n_ndim, n_rank, n_ldim, n_mdim=1, n_mmin=0, para_dt=0.05):
        self._lib = lib()
        self.dims = Dims(n_ndim, n_rank, n_ldim, n_mdim, n_mmin, para_dt)
        h = C.c_void_p()
        rc = self._lib.qpc_create(C.byref(self.dims), C.byref(h))
        if rc != 0:
            raise QpcError(f"qpc_create failed ({rc})")
        self.h = h
        self.N, self.p, self.L, self.M = n_ndim, n_rank, n_ldim, n_mdim

    def _chk(self, rc, what):
        if rc != 0:
            msg = self._lib.qpc_last_error(self.h)
            raise QpcError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None):
            self._lib.qpc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # --- operand upload (the reference's set-up output) ---
    def upload_radial(self, S, chol, rinv, rdif, rpow1=None, mask=None):
        keep = [_as_f64(x) if x is not None else None for x in (S, chol, rinv, rdif, rpow1, mask)]
        self._chk(self._lib.qpc_upload_radial(self.h, *[_dp(x) for x in keep]), "upload_radial")

    def upload_hc(self, upper, lower):
        u, l = _as_f64(upper), _as_f64(lower)
        self._chk(self._lib.qpc_upload_hc(self.h, _dp(u), _dp(l)), "upload_hc")

    def upload_angular(self, yeig, yprj, qeig, qprj, merged, y10=None):
        keep = [_as_f64(x) if x is not None else None for x in (yeig, yprj, qeig, qprj, merged, y10)]
        self._chk(self._lib.qpc_upload_angular(self.h, *[_dp(x) for x in keep]), "upload_angular")

    def upload_setup(self, setup):
        self._chk(self._lib.qpc_upload_setup(self.h, setup.h), "upload_setup")

    # --- state ---
    def set_state(self, coef):
        c = _as_f64(np.asarray(coef, np.complex128))
        self._chk(self._lib.qpc_set_state(self.h, _dp(c)), "set_state")

    def get_state(self):
        out = np.zeros(self.M * self.L * self.N, np.complex128)
        self._chk(self._lib.qpc_get_state(self.h, _dp(out.view(np.float64))), "get_state")
        return out

    # --- propagation (propagator_cn_hc_vgwz / absorber_mask API) ---
    def propagate(self, a_i=None, a_f=None):
        if a_i is None:
            self._chk(self._lib.qpc_step_free(self.h), "step_free")
        else:
            self._chk(self._lib.qpc_step(self.h, a_i, a_f), "step")

    def mask(self):
        self._chk(self._lib.qpc_mask(self.h), "mask")

    # --- length gauge (propagator_cn_hc_lgwz) ---
    def propagate_lg(self, e_i, e_f):
        self._chk(self._lib.qpc_step_lg(self.h, e_i, e_f), "step_lg")

    def run_lg(self, etab, nt_field, it0, it1, use_mask=True, sync=True):
        e = np.ascontiguousarray(etab, np.float64)
        self._chk(self._lib.qpc_run_lg(self.h, _dp(e), nt_field, it0, it1, 1 if use_mask else 0), "run_lg")
        if sync:
            self.sync()

    def run(self, atab, nt_field, it0, it1, use_mask=True, sync=True):
        a = np.ascontiguousarray(atab, np.float64)
        self._chk(self._lib.qpc_run(self.h, _dp(a), nt_field, it0, it1, 1 if use_mask else 0), "run")
        if sync:
            self.sync()

    def sync(self):
        self._chk(self._lib.qpc_sync(self.h), "sync")

    def observe(self, what):
        out = np.zeros(1)
        self._chk(self._lib.qpc_observe(self.h, what, _dp(out)), "observe")
        return float(out[0])

    # --- projection onto field-free eigenstates (eigenstate_sph) ---
    def upload_projection(self, W):
        W = np.ascontiguousarray(W, np.float64)
        self._proj_nsel = W.shape[-2]
        self._chk(self._lib.qpc_upload_projection(self.h, self._proj_nsel, _dp(W)), "upload_projection")

    def project(self, stride=None):
        stride = self._proj_nsel if stride is None else stride
        out = np.zeros(self.M * self.L * stride, np.complex128)
        self._chk(self._lib.qpc_project(self.h, _dp(out.view(np.float64)), stride), "project")
        return out

    # --- photoelectron spectra (spectrum_c_pcs::measure_lmd) ---
    def upload_pcs(self, ijl, cps):
        ijl = np.ascontiguousarray(ijl, np.float64)
        cps = np.ascontiguousarray(cps, np.complex128)
        self._pcs_nkr = cps.shape[0]
        self._chk(self._lib.qpc_upload_pcs(self.h, self._pcs_nkr, _dp(ijl), _dp(cps.view(np.float64))), "upload_pcs")

    def measure_lmd(self):
        """lmd[ikr][im][il] (complex): the /lmd1 partial momentum amplitudes."""
        out = np.zeros(self._pcs_nkr * self.M * self.L, np.complex128)
        self._chk(self._lib.qpc_measure_lmd(self.h, _dp(out.view(np.float64))), "measure_lmd")
        return out.reshape(self._pcs_nkr, self.M, self.L)

    # --- Krylov propagator (propagator_arnoldi with the vgwz H(t)) ---
    def upload_krylov(self, H, y10, q10):
        keep = [_as_f64(x) for x in (H, y10, q10)]
        self._chk(self._lib.qpc_upload_krylov(self.h, *[_dp(x) for x in keep]), "upload_krylov")

    def krylov_step(self, az, dt, kmax=40, emin=1e-14, emax=1e-10):
        """state <- exp(-i dt S^-1 H(az)) state; returns (criterion, m, estimate)."""
        crit = C.c_int(0)
        m = C.c_int64(0)
        eps = np.zeros(1)
        self._chk(self._lib.qpc_krylov_step(self.h, az, dt, kmax, emin, emax, C.byref(crit), C.byref(m), _dp(eps)),
                  "krylov_step")
        return int(crit.value), int(m.value), float(eps[0])

    # --- elliptically polarised step (propagator_cn_hc_vgwr) ---
    def upload_angular_ep(self, ep, isub=-1):
        keep = [_as_f64(ep[k]) for k in _EP_KEYS]
        self._chk(self._lib.qpc_upload_angular_ep(self.h, isub, *[_dp(x) for x in keep]), "upload_angular_ep")

    def propagate_ep(self, ax_i, ax_f, ay_i, ay_f):
        self._chk(self._lib.qpc_step_ep(self.h, ax_i, ax_f, ay_i, ay_f), "step_ep")

    def run_ep(self, axtab, aytab, nt_field, it0, it1, use_mask=True, sync=True):
        ax = np.ascontiguousarray(axtab, np.float64)
        ay = np.ascontiguousarray(aytab, np.float64)
        self._chk(self._lib.qpc_run_ep(self.h, _dp(ax), _dp(ay), nt_field, it0, it1, 1 if use_mask else 0), "run_ep")
        if sync:
            self.sync()

    def build_hc(self, V, K, T, mass=1.0):
        """The field-free bands built on the device from integrate_hc's
        l-independent parts (qpc_build_hc), in place of upload_hc."""
        v, k, t = _as_f64(V), _as_f64(K), _as_f64(T)
        self._chk(self._lib.qpc_build_hc(self.h, _dp(v), _dp(k), _dp(t), mass), "build_hc")

    def build_angular_ep(self, isub=-1):
        """The vgwr angular eigen-data built on the device (qpc_build_angular_ep)."""
        self._chk(self._lib.qpc_build_angular_ep(self.h, isub), "build_angular_ep")

    def build_angular_lgwr(self, isub=-1):
        """The lgwr angular data built on the device (qpc_build_angular_lgwr)."""
        self._chk(self._lib.qpc_build_angular_lgwr(self.h, isub), "build_angular_lgwr")

    # --- the same field in the length gauge (propagator_cn_hc_lgwr) ---
    def upload_angular_lgwr(self, lg, isub=-1):
        keep = [_as_f64(lg[k]) for k in _LG_KEYS]
        self._chk(self._lib.qpc_upload_angular_lgwr(self.h, isub, *[_dp(x) for x in keep]), "upload_angular_lgwr")

    def propagate_lgwr(self, ex_i, ex_f, ey_i, ey_f):
        self._chk(self._lib.qpc_step_lgwr(self.h, ex_i, ex_f, ey_i, ey_f), "step_lgwr")

    def run_lgwr(self, extab, eytab, nt_field, it0, it1, use_mask=True, sync=True):
        ex = np.ascontiguousarray(extab, np.float64)
        ey = np.ascontiguousarray(eytab, np.float64)
        self._chk(self._lib.qpc_run_lgwr(self.h, _dp(ex), _dp(ey), nt_field, it0, it1, 1 if use_mask else 0),
                  "run_lgwr")
        if sync:
            self.sync()

    def upload_outer(self, band):
        b = _as_f64(band)
        self._chk(self._lib.qpc_upload_outer(self.h, _dp(b)), "upload_outer")

    def upload_accel(self, band):
        b = _as_f64(band)
        self._chk(self._lib.qpc_upload_accel(self.h, _dp(b)), "upload_accel")

    def record(self, flags=REC_ALL, cap=0):
        """Observe the state before every step of run / run_ep into a device
        history of cap rows (flags 0: off)."""
        self._rec_cap = cap
        self._chk(self._lib.qpc_record(self.h, flags, cap), "record")

    # --- t-SURFF (spectrum_c_tsf) ---
    def upload_tsurff(self, t):
        """Tables of Setup.tsurff; clears the amplitudes and the Volkov clock."""
        self._tsf_shape = (len(t["kr"]), len(t["kt"]), t["e0"].shape[0])
        ib = np.ascontiguousarray(t["ib"], np.int64)
        self._chk(self._lib.qpc_upload_tsurff(
            self.h, C.byref(t["cfg"]), _dp(t["kr"]), _dp(t["kt"]), _dp(t["kp"]), len(ib),
            ib.ctypes.data_as(C.POINTER(C.c_int64)), _dp(t["b0"]), _dp(t["br"]), _dp(np.ascontiguousarray(t["j0"])),
            _dp(np.ascontiguousarray(t["jr"])), _dp(np.ascontiguousarray(t["y0"])),
            _dp(np.ascontiguousarray(t["e0"]).view(np.float64)), _dp(np.ascontiguousarray(t["pmat0"]))),
            "upload_tsurff")
        if "pmat1" in t:  # the elliptically polarised accumulate's elements
            p1, p2 = _as_f64(t["pmat1"]), _as_f64(t["pmat2"])
            self._chk(self._lib.qpc_upload_tsurff_xy(self.h, _dp(p1), _dp(p2)), "upload_tsurff_xy")

    def tsurff(self, on=True):
        """Accumulate t-SURFF inside run (after every step, before the absorber)."""
        self._chk(self._lib.qpc_tsurff(self.h, 1 if on else 0), "tsurff")

    def tsurff_accumulate(self, az0=0.0, az1=0.0):
        self._chk(self._lib.qpc_tsurff_accumulate(self.h, az0, az1), "tsurff_accumulate")

    def tsurff_accumulate_xy(self, ax0=0.0, ax1=0.0, ay0=0.0, ay1=0.0):
        self._chk(self._lib.qpc_tsurff_accumulate_xy(self.h, ax0, ax1, ay0, ay1), "tsurff_accumulate_xy")

    def tsurff_displacement(self):
        """(disx, disy, disz): int A dt accumulated so far."""
        d = np.zeros(3)
        self._chk(self._lib.qpc_tsurff_displacement(self.h, _dp(d)), "tsurff_displacement")
        return tuple(float(v) for v in d)

    def tsurff_clear(self):
        self._chk(self._lib.qpc_tsurff_clear(self.h), "tsurff_clear")

    def tsurff_read(self):
        """(work complex [nkr][nkt][nkp], time, int A dt)."""
        w = np.zeros(self._tsf_shape, np.complex128)
        t, d = np.zeros(1), np.zeros(1)
        self._chk(self._lib.qpc_tsurff_read(self.h, _dp(w.view(np.float64)), _dp(t), _dp(d)), "tsurff_read")
        return w, float(t[0]), float(d[0])

    def observe_noalloc(self, flags, out_ptr):
        """qpc_observe_noalloc into the device address out_ptr (REC_COLUMNS doubles), queued."""
        self._chk(self._lib.qpc_observe_noalloc(self.h, flags, C.c_void_p(out_ptr)), "observe_noalloc")

    def record_read(self):
        """Rows recorded so far, REC_COLUMNS per step: norm, <z>, z acceleration,
        outer norm, <x>, <y>, x and y acceleration."""
        out = np.zeros(REC_COLUMNS * max(1, getattr(self, "_rec_cap", 0)))
        n = C.c_int64(0)
        self._chk(self._lib.qpc_record_read(self.h, _dp(out), C.byref(n)), "record_read")
        return out[:REC_COLUMNS * n.value].reshape(-1, REC_COLUMNS)

    def krylov_step_adaptive(self, az, dt, kstart, kmax, tol):
        """Lanczos step grown from kstart until the estimate <= tol; returns (m, estimate, ab[m][2])."""
        m = C.c_int64(0)
        eps = np.zeros(1)
        ab = np.zeros(2 * kmax)
        self._chk(self._lib.qpc_krylov_step_adaptive(self.h, az, dt, kstart, kmax, tol, C.byref(m), _dp(eps),
                                                      _dp(ab)), "krylov_step_adaptive")
        return int(m.value), float(eps[0]), ab[:2 * m.value].reshape(-1, 2)

    def krylov_propagate(self, pulse, t0, t1, dt_base, tol, kmin=8, kmax=64, net=None, record=0, record_ab=True):
        """CF4-Magnus Krylov propagation from t0 to t1 (qpc_krylov_propagate);
        pulse = (f0, w0, nc, n0, c0).  Returns (stats dict, record dict or None).
        record_ab=False records the accepted steps (features, h, m) without
        changing the propagation."""
        p = Sinen2(*pulse)
        st = KrylovStats()
        rec = None
        if record:
            buf = dict(features=np.zeros((record, 7), np.float32), h=np.zeros(record),
                       m=np.zeros(record, np.int64))
            if record_ab:
                buf["ab"] = np.zeros((record, 2 * kmax))
            rec = KrylovRecord(record, 0, _fp(buf["features"]), _dp(buf["h"]),
                               buf["m"].ctypes.data_as(C.POINTER(C.c_int64)),
                               _dp(buf["ab"]) if record_ab else None)
        self._chk(self._lib.qpc_krylov_propagate(self.h, C.byref(p), t0, t1, dt_base, tol, kmin, kmax,
                                                 net.h if net is not None else None, C.byref(st),
                                                 C.byref(rec) if rec is not None else None), "krylov_propagate")
        stats = {k: getattr(st, k) for k, _ in KrylovStats._fields_}
        if rec is None:
            return stats, None
        n = rec.n
        return stats, {k: v[:n] for k, v in buf.items()}

    # --- launch-shape autotuner ---
    def set_tune(self, seg_add=0, stage_grid=512, serial=0, share_r=1):
        self._chk(self._lib.qpc_set_tune(self.h, C.byref(Tune(seg_add, serial, stage_grid, share_r, 0))), "set_tune")

    def get_tune(self):
        t = Tune()
        self._chk(self._lib.qpc_get_tune(self.h, C.byref(t)), "get_tune")
        return {"seg_add": t.seg_add, "stage_grid": t.stage_grid, "serial": t.serial, "share_r": t.share_r,
                "tables_r": t.tables_r}

    def autotune(self, atab, nt_field, it0, nsteps, stride=0):
        """qpc_autotune: times the candidates over the with-field steps
        [it0, it0+nsteps), keeps the fastest; returns (choice, {candidate: ms
        per step})."""
        a = np.ascontiguousarray(atab, np.float64)
        t = Tune()
        ms = np.zeros(8)
        self._chk(self._lib.qpc_autotune(self.h, _dp(a), stride, nt_field, it0, nsteps, C.byref(t), _dp(ms)),
                  "autotune")
        return ({"seg_add": t.seg_add, "stage_grid": t.stage_grid, "serial": t.serial, "share_r": t.share_r,
                 "tables_r": t.tables_r},
                {f"seg_add={sa},stage_grid={sg}": float(ms[k]) for k, (sa, sg) in enumerate(TUNE_CANDIDATES)})

    def timing_begin(self, which, max_launches):
        self._chk(self._lib.qpc_timing_begin(self.h, which, max_launches), "timing_begin")

    def timing_end(self):
        n = C.c_int64(0)
        ms = np.zeros(1)
        self._chk(self._lib.qpc_timing_end(self.h, C.byref(n), _dp(ms)), "timing_end")
        return int(n.value), float(ms[0])


class BatchPropagator(Propagator):
    """Batched shots (qpc_create_batch): n_shots z-polarised runs of one radial
    grid in one context, each padded to n_ldim channels."""

    def __init__(self, n_ndim, n_rank, n_ldim, n_shots, para_dt=0.05, time_batch=8):
        self._lib = lib()
        self.N, self.p, self.L, self.M, self.m0 = n_ndim, n_rank, n_ldim, n_shots, 0
        self.dims = Dims(n_ndim, n_rank, n_ldim, 1, 0, para_dt)
        h = C.c_void_p()
        rc = self._lib.qpc_create_batch(C.byref(self.dims), n_shots, time_batch, C.byref(h))
        if rc != 0:
            raise QpcError(f"qpc_create_batch failed ({rc})")
        self.h = h

    def upload_angular_shots(self, shots):
        """shots: per shot a dict yeig, yprj (Ls x Ls), qeig, qprj, merged (complex
        Ls x Ls), y10 (Ls) for its own Ls <= L; padded here with decoupled channels."""
        L, S = self.L, self.M
        ye, qe, y10 = np.zeros(S * L), np.zeros(S * L), np.zeros(S * L)
        yp = np.zeros((S, L, L))
        qp = np.zeros((S, L, L), np.complex128)
        mg = np.zeros((S, L, L), np.complex128)
        for s, d in enumerate(shots):
            ls = len(d["yeig"])
            assert ls <= L
            ye[s * L:s * L + ls] = d["yeig"]
            qe[s * L:s * L + ls] = d["qeig"]
            y10[s * L:s * L + ls - 1] = np.asarray(d["y10"])[:ls - 1]
            for arr, src in ((yp, d["yprj"]), (qp, d["qprj"]), (mg, d["merged"])):
                arr[s, :ls, :ls] = np.asarray(src).reshape(ls, ls)
                arr[s, ls:, ls:] = np.eye(L - ls)
        self.upload_angular(ye, yp, qe, qp, mg, y10)

    def run_batch(self, atab, nt_field, it0, it1, use_mask=True, sync=True):
        """atab: n_shots x stride array of A at the time points (zeros after a shot's pulse)."""
        a = np.ascontiguousarray(atab, np.float64)
        self._chk(self._lib.qpc_run_batch(self.h, _dp(a), a.shape[1], nt_field, it0, it1, 1 if use_mask else 0),
                  "run_batch")
        if sync:
            self.sync()

    def autotune_batch(self, atab, nt_field, it0, nsteps):
        """qpc_autotune over the batch's steps (atab as run_batch)."""
        a = np.ascontiguousarray(atab, np.float64)
        return self.autotune(a, nt_field, it0, nsteps, stride=a.shape[1])

    def observe_batch(self, what):
        out = np.zeros(self.M)
        self._chk(self._lib.qpc_observe_batch(self.h, what, _dp(out)), "observe_batch")
        return out


def bucket_shots(ldims, n_buckets):
    """Split shots of channel counts `ldims` into at most n_buckets groups that
    are contiguous in the order of increasing channel count and minimise the
    padded channel count, sum over groups of (shots in the group) x (largest
    count in it).  Exact (dynamic programming over the sorted counts).
    Returns the groups as lists of shot indices, each in increasing count."""
    order = sorted(range(len(ldims)), key=lambda k: (ldims[k], k))
    ls = [ldims[k] for k in order]
    n = len(ls)
    K = max(1, min(int(n_buckets), n))
    inf = float("inf")
    cost = [[inf] * (n + 1) for _ in range(K + 1)]
    cut = [[0] * (n + 1) for _ in range(K + 1)]
    cost[0][0] = 0
    for k in range(1, K + 1):
        for j in range(1, n + 1):
            for i in range(k - 1, j):
                v = cost[k - 1][i] + (j - i) * ls[j - 1]
                if v < cost[k][j]:
                    cost[k][j], cut[k][j] = v, i
    best = min(range(1, K + 1), key=lambda k: cost[k][n])
    groups, j = [], n
    for k in range(best, 0, -1):
        i = cut[k][j]
        groups.append(order[i:j])
        j = i
    return groups[::-1]


class ShotBuckets:
    """Shots of uneven channel counts in a few batch contexts
    (qpc_create_batch), each padded only to its own largest channel count
    (bucket_shots) instead of every shot to the largest one.  Each context has
    its own streams and run_batch only queues work, so the buckets run
    concurrently on the GPU.  The per-shot interface follows BatchPropagator:
    shot k keeps index k in every argument and result."""

    def __init__(self, n_ndim, n_rank, shot_ldim, n_buckets=4, para_dt=0.05, time_batch=8):
        self.N, self.ldim = n_ndim, [int(x) for x in shot_ldim]
        self.groups = bucket_shots(self.ldim, n_buckets)
        self.Lb = [max(self.ldim[k] for k in g) for g in self.groups]
        self.props = []
        try:
            for g, Lb in zip(self.groups, self.Lb):
                self.props.append(BatchPropagator(n_ndim, n_rank, Lb, len(g), para_dt, time_batch))
        except Exception:
            self.close()
            raise

    @property
    def channels_padded(self):
        return sum(len(g) * Lb for g, Lb in zip(self.groups, self.Lb))

    def channels_of(self, q):
        """Padded channels of bucket q."""
        return len(self.groups[q]) * self.Lb[q]

    def upload_radial(self, *args):
        for p in self.props:
            p.upload_radial(*args)

    def upload_hc(self, upper, lower):
        """upper/lower: the field-free bands of l = 0 .. max(shot_ldim) - 1; a
        bucket takes those of its own l range."""
        for p, Lb in zip(self.props, self.Lb):
            n = Lb * self.N * (p.p + 1)
            p.upload_hc(np.asarray(upper, np.complex128).ravel()[:n], np.asarray(lower, np.complex128).ravel()[:n])

    def upload_angular_shots(self, shots):
        for p, g in zip(self.props, self.groups):
            p.upload_angular_shots([shots[k] for k in g])

    def set_states(self, states):
        """states[k]: shot k's coefficients, ldim[k] x N."""
        for p, g, Lb in zip(self.props, self.groups, self.Lb):
            blk = np.zeros(len(g) * Lb * self.N, np.complex128)
            for q, k in enumerate(g):
                blk[q * Lb * self.N:q * Lb * self.N + self.ldim[k] * self.N] = states[k]
            p.set_state(blk)

    def get_states(self):
        out = [None] * len(self.ldim)
        for p, g, Lb in zip(self.props, self.groups, self.Lb):
            blk = p.get_state()
            for q, k in enumerate(g):
                out[k] = blk[q * Lb * self.N:q * Lb * self.N + self.ldim[k] * self.N].copy()
        return out

    def run_batch(self, atab, nt_field, it0, it1, use_mask=True, sync=True):
        """atab: one row per shot (zeros after a shot's pulse)."""
        a = np.asarray(atab, np.float64)
        for p, g in zip(self.props, self.groups):
            p.run_batch(np.ascontiguousarray(a[g]), nt_field, it0, it1, use_mask=use_mask, sync=False)
        if sync:
            self.sync()

    def sync(self):
        for p in self.props:
            p.sync()

    def autotune(self, atab, nt_field, it0, nsteps):
        """Each bucket context tuned on its own shots' steps; returns the
        per-bucket (choice, candidate times)."""
        a = np.asarray(atab, np.float64)
        return [p.autotune_batch(np.ascontiguousarray(a[g]), nt_field, it0, nsteps)
                for p, g in zip(self.props, self.groups)]

    def observe_batch(self, what):
        out = np.zeros(len(self.ldim))
        for p, g in zip(self.props, self.groups):
            out[g] = p.observe_batch(what)
        return out

    def close(self):
        for p in self.props:
            p.close()
        self.props = []
