"""Set-up on the device.  The field-free bands (qpc_build_hc: integrate_hc's
assembly and lu_zsb for every l) in place of the host's upper / lower, and
the elliptical steps' angular eigen-data (qpc_build_angular_ep / _lgwr:
rocSOLVER dsyevd on the parity blocks, the merged products on the FP64
matrix cores) in place of the oracle's or the host set-up's: runs and steps
agree to 1e-12 (the angular steps depend only on the spectral projectors, so
bases that differ inside degenerate eigenspaces, such as the zero
eigenvalues of odd M, give the same step)."""
import time

import numpy as np
import pytest

import pyoracle as po
from _cfg import make_oracle, mixed_state, rel_l2, shape, upload_from_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("gauge", ["vgwr", "lgwr"])
@pytest.mark.parametrize("isub", [-1, 2, 3])
def test_device_angular_step_matches_oracle(gauge, isub):
    import qpc_gpu
    cfg = shape(240, 6, 48.0, M=5, m0=-2)
    print(f"case: device angular set-up {gauge} N=240 L=6 M=5 isub={isub}", flush=True)
    o = make_oracle(cfg)
    prop = qpc_gpu.Propagator(cfg["N"], cfg["p"], cfg["L"], cfg["M"], cfg["m0"], cfg["dt"])
    upload_from_oracle(prop, o)
    c0 = mixed_state(o, seed=71 + isub, amp=0.3)
    par = np.repeat(np.array([il % 2 for im in range(cfg["M"]) for il in range(cfg["L"])]), cfg["N"])
    if isub >= 2:
        c0 = np.where(par == isub - 2, c0, 0.0)
        c0 /= np.sqrt(o.norm(c0))
    c = c0.copy()
    if gauge == "vgwr":
        o.set_ep(po.ep_angular(cfg["M"], cfg["L"], cfg["m0"]), isub=isub)
        o.step_ep(c, 0.9, 1.1, -0.5, 0.3)
        prop.build_angular_ep(isub)
        prop.set_state(c0)
        prop.propagate_ep(0.9, 1.1, -0.5, 0.3)
    else:
        o.set_lgwr(po.lgwr_angular(cfg["M"], cfg["L"], cfg["m0"]), isub=isub)
        o.step_lgwr(c, 0.05, 0.06, -0.03, 0.02)
        prop.build_angular_lgwr(isub)
        prop.set_state(c0)
        prop.propagate_lgwr(0.05, 0.06, -0.03, 0.02)
    err = rel_l2(prop.get_state(), c)
    print(f"  rel L2 {err:.3e}", flush=True)
    assert err <= 1e-12


@pytest.mark.timeout(600)
def test_device_angular_full_size_matches_host_setup():
    """N=1500, M=81, L=41 (3321 channels), even sub-block: one vgwr step with
    the device-built eigen-data against one with the host set-up's, and the
    set-up times of both."""
    import qpc_gpu
    N, L, M, m0 = 1500, 41, 81, -40
    print("case: device angular set-up N=1500 M=81 L=41 isub=2", flush=True)
    st = qpc_gpu.Setup(N, 7, L, M, m0, 0.05, 0.0, 300.0, 1.0, 1.0, 280.0, 300.0, 0.125)
    prop = qpc_gpu.Propagator(N, 7, L, M, m0, 0.05)
    prop.upload_setup(st)
    _, c = st.eigenstate(40, 0, 0)
    rng = np.random.default_rng(73)
    par = np.repeat(np.array([il % 2 for im in range(M) for il in range(L)]), N)
    c = c + 1e-3 * np.where(par == 0, rng.standard_normal(c.size) + 1j * rng.standard_normal(c.size), 0.0)
    a0 = 0.0534 / 0.057 / np.sqrt(1.25)
    f = (a0 * 0.9, a0 * 0.91, 0.5 * a0 * 0.1, 0.5 * a0 * 0.12)
    t0 = time.perf_counter()
    ep = qpc_gpu.angular_ep(M, L, m0)
    t_host = time.perf_counter() - t0
    prop.upload_angular_ep(ep, isub=2)
    prop.set_state(c)
    prop.propagate_ep(*f)
    ref = prop.get_state()
    prop.build_angular_ep(2)  # first call: rocBLAS / rocSOLVER initialisation
    t0 = time.perf_counter()
    prop.build_angular_ep(2)
    t_dev = time.perf_counter() - t0
    prop.set_state(c)
    prop.propagate_ep(*f)
    err = rel_l2(prop.get_state(), ref)
    print(f"  host set-up {t_host:.2f} s, device set-up {t_dev:.3f} s; step rel L2 {err:.3e}", flush=True)
    assert err <= 1e-11


@pytest.mark.parametrize("p,L,M,m0", [(7, 31, 1, 0), (8, 6, 3, -1)])
def test_device_hc_bands_match_host_bands(p, L, M, m0):
    """qpc_build_hc from the library set-up's integrate_hc parts gives the
    set-up's own upper / lower bands bit for bit (the same uncontracted
    arithmetic on host and device): a with-field run across the pulse end
    with the absorber is identical to the one with the host's bands (whose
    operands are pinned against the oracle's in test_library.py)."""
    import qpc_gpu
    from _cfg import atable
    cfg = shape(600, L, 120.0, M=M, m0=m0, nc=1.0, p=p)
    print(f"case: device field-free bands p={p} N=600 L={L} M={M}", flush=True)
    st = qpc_gpu.Setup(cfg["N"], p, L, M, m0, cfg["dt"], 0.0, cfg["rmax"], 1.0, 1.0, cfg["mask_rmin"],
                       cfg["mask_rmax"], cfg["mask_fact"])
    o = make_oracle(cfg)
    nt, a = atable(cfg)
    c0 = mixed_state(o, seed=83, amp=0.3)  # the oracle's 1s state: any fixed state serves
    it0, it1 = nt - 40, nt + 10
    parts = [st.get(w) for w in (qpc_gpu.SETUP_HV, qpc_gpu.SETUP_HK, qpc_gpu.SETUP_HT)]
    finals = []
    for dev in (False, True):
        prop = qpc_gpu.Propagator(cfg["N"], p, L, M, m0, cfg["dt"])
        prop.upload_setup(st)
        if dev:
            prop.build_hc(*parts, 1.0)
        prop.set_state(c0)
        prop.run(a, nt, it0, it1, use_mask=True)
        finals.append(prop.get_state())
        prop.close()
    print(f"  device vs host bands: identical {np.array_equal(finals[0], finals[1])}, "
          f"rel L2 {rel_l2(finals[1], finals[0]):.3e}", flush=True)
    assert np.array_equal(finals[0], finals[1])

