"""qpc_observe_noalloc (the observables without any device allocation,
partial sums in the library's static scratch array) against qpc_observe
(per-channel partials in context memory, summed on the host) on two calls in
a row on one GPU, the second on a smaller context, through the C ABI."""
import ctypes

import numpy as np
import pytest

from _cfg import make_oracle, mixed_state, shape, upload_from_oracle

import pyoracle as po

pytestmark = pytest.mark.gpu


def _ctx(cfg, seed):
    import qpc_gpu
    o = make_oracle(cfg)
    prop = qpc_gpu.Propagator(cfg["N"], cfg["p"], cfg["L"], cfg.get("M", 1), cfg.get("m0", 0), cfg["dt"])
    upload_from_oracle(prop, o)
    st = qpc_gpu.Setup(cfg["N"], cfg["p"], cfg["L"], cfg.get("M", 1), cfg.get("m0", 0), cfg["dt"], 0.0, cfg["rmax"],
                       1.0, 1.0, cfg["mask_rmin"], cfg["mask_rmax"], cfg["mask_fact"])
    prop.upload_outer(st.get(qpc_gpu.SETUP_SOUTER))
    prop.upload_accel(o.get(po.QOR_DVDR))
    prop.set_state(mixed_state(o, seed=seed, amp=0.4))
    return prop


def _hip_runtime():
    """The HIP runtime libqpc_gpu.so is linked against, found among the
    objects the process has mapped (a second runtime in the process, such as
    the one torch bundles, cannot open the device after the first has)."""
    with open("/proc/self/maps") as f:
        for line in f:
            path = line.split()[-1]
            if "libamdhip64.so" in path:
                hip = ctypes.CDLL(path)
                hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
                hip.hipFree.argtypes = [ctypes.c_void_p]
                hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
                return hip
    raise RuntimeError("libqpc_gpu.so did not map a HIP runtime")


class _DevBuf:
    """n doubles of device memory, filled with -1 so unwritten words show."""

    def __init__(self, hip, n):
        self.hip, self.n = hip, n
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), 8 * n) == 0
        self.put(np.full(n, -1.0))

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data, 8 * self.n, 1) == 0  # hipMemcpyHostToDevice

    def get(self):
        a = np.empty(self.n)
        assert self.hip.hipMemcpy(a.ctypes.data, self.ptr, 8 * self.n, 2) == 0  # hipMemcpyDeviceToHost
        return a

    def free(self):
        assert self.hip.hipFree(self.ptr) == 0


def test_observe_noalloc_matches_observe_on_two_calls():
    import qpc_gpu
    qpc_gpu.lib()
    out = _DevBuf(_hip_runtime(), qpc_gpu.REC_COLUMNS)
    kinds = (qpc_gpu.QPC_OBS_NORM, qpc_gpu.QPC_OBS_DIPOLE_Z, qpc_gpu.QPC_OBS_ACCEL_Z, qpc_gpu.QPC_OBS_OUTER,
             qpc_gpu.QPC_OBS_DIPOLE_X, qpc_gpu.QPC_OBS_DIPOLE_Y, qpc_gpu.QPC_OBS_ACCEL_X, qpc_gpu.QPC_OBS_ACCEL_Y)
    # the second call runs on a smaller context (fewer channels and rows)
    for cfg, seed in ((shape(600, 24, 120.0, M=3, m0=-1), 3), (shape(240, 5, 48.0), 4)):
        prop = _ctx(cfg, seed)
        ref = [prop.observe(k) for k in kinds]
        print(f"case: observe_noalloc N={cfg['N']} L={cfg['L']} M={cfg.get('M', 1)}", flush=True)
        out.put(np.full(out.n, -1.0))
        prop.observe_noalloc(qpc_gpu.REC_ALL, out.ptr.value)
        prop.sync()
        got = out.get()
        print(f"  got {got}, qpc_observe {np.array(ref)}", flush=True)
        for g, r in zip(got, ref):
            assert abs(g - r) <= 1e-13 * max(1.0, abs(r))
        # a subset of the flags leaves the others at zero
        prop.observe_noalloc(qpc_gpu.REC_NORM | qpc_gpu.REC_OUTER, out.ptr.value)
        prop.sync()
        got = out.get()
        assert np.all(got[[1, 2, 4, 5, 6, 7]] == 0.0)
        assert abs(got[0] - ref[0]) <= 1e-13 and abs(got[3] - ref[3]) <= 1e-13 * max(1.0, ref[3])
        prop.close()
    out.free()
