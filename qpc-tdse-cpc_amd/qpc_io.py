"""The QPC-TDSE driver's output layout (binsrc/sphTDSE.cpp:168-200, 206-209,
243-277, 364-403, 450-472, 545-561, 837-864) and its ground-state reader
(load_initial, :119-160), over lib/libqpc_io.so (include/qpc_io.h: 1-D HDF5
datasets of native double, size_t, long and int, as support_hdf.h writes them).

save_realtime writes the datasets realtime_lp saves for the z-polarised
velocity- or length-gauge run; a dataset is written only when its data is
given (the driver's calculate_* switches).  load_initial reads a file back
into a run's channel layout with the reference's checks and return codes.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
IO_PATH = os.path.join(HERE, "lib", "libqpc_io.so")

F64, U64, I64, I32 = 0, 1, 2, 3
_NP = {F64: np.float64, U64: np.uint64, I64: np.int64, I32: np.int32}


class H5Error(RuntimeError):
    pass


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(IO_PATH):
            raise H5Error(f"{IO_PATH} is not built (needs the HDF5 C library; make -C qpc-tdse-cpc_amd)")
        L = C.CDLL(IO_PATH)
        vp = C.c_void_p
        L.qpc_h5_open.argtypes = [C.c_char_p, C.c_char, C.POINTER(vp)]
        L.qpc_h5_close.argtypes = [vp]
        L.qpc_h5_save.argtypes = [vp, C.c_char_p, C.c_int, vp, C.c_int64]
        L.qpc_h5_load.argtypes = [vp, C.c_char_p, C.c_int, vp, C.c_int64]
        L.qpc_h5_info.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_int)]
        L.qpc_h5_list.argtypes = [vp, C.c_char_p, C.c_int64]
        _LIB = L
    return _LIB


class H5File:
    """hdf5_file (support_hdf.h): mode 'w', 'r' or 'u'."""

    def __init__(self, path, mode="r"):
        self._lib = lib()
        h = C.c_void_p()
        rc = self._lib.qpc_h5_open(path.encode(), mode.encode(), C.byref(h))
        if rc:
            raise H5Error(f"cannot open {path} ({mode}): {rc}")
        self.h = h

    def close(self):
        if self.h:
            self._lib.qpc_h5_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def save(self, name, data, typ):
        a = np.ascontiguousarray(np.atleast_1d(data), _NP[typ])
        if np.iscomplexobj(data):
            raise H5Error("complex data is saved as its interleaved doubles")
        rc = self._lib.qpc_h5_save(self.h, name.encode(), typ, a.ctypes.data, a.size)
        if rc:
            raise H5Error(f"cannot save {name}: {rc}")

    def info(self, name):
        n, t = C.c_int64(0), C.c_int(0)
        rc = self._lib.qpc_h5_info(self.h, name.encode(), C.byref(n), C.byref(t))
        return None if rc else (int(n.value), int(t.value))

    def load(self, name, typ, n=None):
        inf = self.info(name)
        if inf is None:
            raise H5Error(f"no dataset {name}")
        n = inf[0] if n is None else n
        a = np.zeros(n, _NP[typ])
        rc = self._lib.qpc_h5_load(self.h, name.encode(), typ, a.ctypes.data, n)
        if rc:
            raise H5Error(f"cannot load {name}: {rc}")
        return a

    def names(self):
        buf = C.create_string_buffer(1 << 16)
        self._lib.qpc_h5_list(self.h, buf, len(buf))
        return [s for s in buf.value.decode().split("\n") if s]


def _cplx(a):
    return np.ascontiguousarray(a, np.complex128).view(np.float64).ravel()


def save_realtime(path, *, N, n_rank, L, M=1, m0=0, rmin=0.0, rmax, para_dt, para_nt, knots, fieldx=(), fieldy=(),
                  fieldz=(), coef=None, lmd1=None, pmd1=None, pmd2=None, axes=None, tsurff=None, proj=None,
                  t_list=None, x_list=None, y_list=None, z_list=None, fx_list=None, fy_list=None, fz_list=None,
                  gauge=0):
    """realtime_lp's save block (sphTDSE.cpp:837-864) with the datasets of
    util/knot/disp/spec/proj/obsv::save_* (names and element types as the
    driver writes them).  fieldz etc.: the (e0, w0, nc, n0, cep) arrays of
    each laser component; axes = (pr, pt, pp) of the spectra; tsurff =
    (zc, r0); proj = (val complex, nnz); coef, lmd1, pmd1, pmd2 complex."""
    with H5File(path, "w") as f:
        f.save("/para_dt", para_dt, F64)
        f.save("/para_nt", para_nt, U64)
        # util::save_basic (:168-199)
        f.save("/nr", N, U64)
        f.save("/nl", L, U64)
        f.save("/nm", M, U64)
        f.save("/m0", m0, I64)
        f.save("/mr", n_rank, U64)
        f.save("/rmin", rmin, F64)
        f.save("/rmax", rmax, F64)
        for tag, comps in (("x", fieldx), ("y", fieldy), ("z", fieldz)):
            for i, p in enumerate(comps):
                f.save(f"/field{tag}_{i}", np.asarray(p, np.float64), F64)
        # spec::save_basic (:383-402)
        if axes is not None:
            for name, a in zip(("/pr", "/pt", "/pp"), axes):
                f.save(name, np.asarray(a, np.float64), F64)
        if tsurff is not None:
            f.save("/tsurff_zc", tsurff[0], F64)
            f.save("/tsurff_r0", tsurff[1], F64)
        # knot::save_knots (:206-209)
        f.save("/knot_sequence", np.asarray(knots, np.float64), F64)
        # disp::save_disp (:243-262)
        if coef is not None:
            f.save("/coef1", _cplx(coef), F64)
        # spec::save_spec (:364-380)
        if lmd1 is not None:
            f.save("/lmd1", _cplx(lmd1), F64)
        if pmd1 is not None:
            f.save("/pmd1", _cplx(pmd1), F64)
        if pmd2 is not None:
            f.save("/pmd2", _cplx(pmd2), F64)
        # proj::save_proj (:450-472)
        if proj is not None:
            f.save("/proj_val", _cplx(proj[0]), F64)
            f.save("/proj_nnz", np.asarray(proj[1], np.uint64), U64)
        # obsv::save_time / save_obsv / save_field (:545-561)
        if t_list is not None:
            f.save("/t_list", t_list, F64)
        for name, a in (("/x_list", x_list), ("/y_list", y_list), ("/z_list", z_list)):
            if a is not None and len(a):
                f.save(name, a, F64)
        f.save("gauge", gauge, I32)
        for name, a in (("/fx_list", fx_list), ("/fy_list", fy_list), ("/fz_list", fz_list)):
            if a is not None and len(a):
                f.save(name, a, F64)


def load_initial(path, N, n_rank, L, M=1, m0=0, rmin=0.0, rmax=None):
    """init::load_initial (sphTDSE.cpp:119-160): (code, coef) with coef the
    run's M*L*N complex coefficients.  Codes as the reference: -1 radial
    grid differs (N, rmin, rmax), -2 n_rank differs, -3 the file's m range is
    not inside the run's, -4 the file has more l channels.  The file's
    channel (m, l index) goes to the run's channel of the same m and l index;
    the reference's offset (:150, size_t(n_mmin - m_init)) has the opposite
    sign and only places the channels right when both m ranges start at the
    same m."""
    with H5File(path, "r") as f:
        nr = int(f.load("/nr", U64, 1)[0])
        mr = int(f.load("/mr", U64, 1)[0])
        nl = int(f.load("/nl", U64, 1)[0])
        nm = int(f.load("/nm", U64, 1)[0])
        fm0 = int(f.load("/m0", I64, 1)[0])
        frmin = float(f.load("/rmin", F64, 1)[0])
        frmax = float(f.load("/rmax", F64, 1)[0])
        if nr != N or frmin != rmin or (rmax is not None and frmax != rmax):
            return -1, None
        if mr != n_rank:
            return -2, None
        if fm0 < m0 or fm0 + nm > m0 + M:
            return -3, None
        if nl > L:
            return -4, None
        tmp = f.load("/coef1", F64, 2 * nr * nl * nm).view(np.complex128).reshape(nm, nl, nr)
    coef = np.zeros((M, L, N), np.complex128)
    coef[fm0 - m0:fm0 - m0 + nm, :nl, :] = tmp
    return 0, coef.ravel()
