"""GPU parity of the observables: the outer-region norm and the dipole
acceleration of qpc_observe, and the in-loop recording of qpc_run (norm,
<z>, acceleration, outer norm before every step, as the driver's
obsv::observe, binsrc/sphTDSE.cpp:793-795), against the CPU oracle on the
same operands and states.  The oracle's outer norm and acceleration band are
pinned in test_oracle_pins.py (scipy quadrature, hydrogen closed forms,
Ehrenfest's relation)."""
import numpy as np
import pytest

from _cfg import H_N1000_L31, atable, make_oracle, mixed_state, rel_l2, shape, upload_from_oracle

import pyoracle as po

pytestmark = pytest.mark.gpu


def _prop(o, cfg):
    import qpc_gpu
    prop = qpc_gpu.Propagator(cfg["N"], cfg["p"], cfg["L"], cfg.get("M", 1), cfg.get("m0", 0), cfg["dt"])
    upload_from_oracle(prop, o)
    st = qpc_gpu.Setup(cfg["N"], cfg["p"], cfg["L"], cfg.get("M", 1), cfg.get("m0", 0), cfg["dt"], 0.0,
                       cfg["rmax"], 1.0, 1.0, cfg["mask_rmin"], cfg["mask_rmax"], cfg["mask_fact"])
    prop.upload_outer(st.get(qpc_gpu.SETUP_SOUTER))  # <Bi|theta(r - mask_rmin)|Bj>
    prop.upload_accel(o.get(po.QOR_DVDR))
    return prop


@pytest.mark.parametrize("N,L,rmax,M,m0", [(1000, 31, 200.0, 1, 0), (240, 5, 48.0, 3, -1), (300, 6, 60.0, 5, -2)])
def test_observe_outer_and_acceleration(N, L, rmax, M, m0):
    import qpc_gpu
    cfg = shape(N, L, rmax, M=M, m0=m0)
    print(f"case: observables N={N} L={L} M={M}", flush=True)
    o = make_oracle(cfg)
    prop = _prop(o, cfg)
    # a state with weight beyond the window: 1s plus a broad admixture
    c = mixed_state(o, seed=21, amp=0.5)
    prop.set_state(c)
    out_g = prop.observe(qpc_gpu.QPC_OBS_OUTER)
    out_o = o.outer_norm(c, cfg["mask_rmin"])
    acc_g = prop.observe(qpc_gpu.QPC_OBS_ACCEL_Z)
    acc_o = o.dipole_z(c, 1)
    print(f"  outer {out_g:.6e} vs {out_o:.6e}; accel {acc_g:.6e} vs {acc_o:.6e}", flush=True)
    assert out_o > 1e-4
    assert abs(out_g - out_o) <= 1e-13 * max(1.0, out_o)
    assert abs(acc_g - acc_o) <= 1e-12 * max(1.0, abs(acc_o))
    # observable_r: <x>, <y> with f = r and f = dV/dr
    for which, kx, ky in ((0, qpc_gpu.QPC_OBS_DIPOLE_X, qpc_gpu.QPC_OBS_DIPOLE_Y),
                          (1, qpc_gpu.QPC_OBS_ACCEL_X, qpc_gpu.QPC_OBS_ACCEL_Y)):
        xo, yo = o.dipole_xy(c, which)
        xg, yg = prop.observe(kx), prop.observe(ky)
        print(f"  observable_r f={'dV/dr' if which else 'r'}: x {xg:.6e} vs {xo:.6e}, y {yg:.6e} vs {yo:.6e}",
              flush=True)
        assert abs(xg - xo) <= 1e-12 * max(1.0, abs(xo)) and abs(yg - yo) <= 1e-12 * max(1.0, abs(yo))
    prop.close()


def test_acceleration_closed_form_on_device():
    """<1s|cos(theta)/r^2|2p0> = 4/(27 sqrt 2) through the device observable."""
    import qpc_gpu
    cfg = shape(300, 3, 60.0)
    o = make_oracle(cfg)
    prop = _prop(o, cfg)
    _, s1 = o.eigenstate(0, 0, 0)
    _, s2 = o.eigenstate(0, 1, 0)
    prop.set_state((s1 + s2) / np.sqrt(2))
    acc = prop.observe(qpc_gpu.QPC_OBS_ACCEL_Z)
    print(f"case: accel closed form: {acc:.15f}", flush=True)
    assert abs(acc - 4.0 / (27.0 * np.sqrt(2.0))) < 1e-10
    prop.close()


def test_recorded_history_against_oracle():
    """400 steps of the N=1000, L=31 run (pulse on, absorber every step) from
    a state with outer weight: the device history recorded inside qpc_run
    against the oracle's observations before every step."""
    import qpc_gpu
    cfg = H_N1000_L31
    o = make_oracle(cfg)
    prop = _prop(o, cfg)
    nt, a = atable(cfg)
    c = mixed_state(o, seed=9, amp=0.3)
    prop.set_state(c)
    steps = (2000, 2400)
    n = steps[1] - steps[0]
    print(f"case: N=1000 L=31 recorded run steps {steps}", flush=True)
    prop.record(qpc_gpu.REC_ALL, n + 5)
    prop.run(a, nt, steps[0], steps[0] + 150, use_mask=True)   # two calls fill one history
    prop.run(a, nt, steps[0] + 150, steps[1], use_mask=True)
    h = prop.record_read()
    ref = o.run_record(c, a, nt, steps[0], steps[1], cfg["mask_rmin"], use_mask=True)
    assert h.shape == (n, 8)
    err = rel_l2(prop.get_state(), c)
    d = np.abs(h - ref).max(axis=0)
    scale = np.abs(ref).max(axis=0)
    assert np.abs(h[:, 4:]).max() <= 1e-12  # one m: observable_r vanishes
    print(f"  final rel L2 {err:.3e}; max diffs norm {d[0]:.2e} z {d[1]:.2e} (max {scale[1]:.3f}) "
          f"accel {d[2]:.2e} (max {scale[2]:.3f}) outer {d[3]:.2e} (max {scale[3]:.3e})", flush=True)
    assert err <= 1e-10
    assert d[0] <= 1e-12
    assert d[1] <= 1e-10 * scale[1]
    assert d[2] <= 1e-10 * scale[2]
    assert d[3] <= 1e-12
    # a run that would overflow the history is refused before it starts
    with pytest.raises(qpc_gpu.QpcError):
        prop.run(a, nt, steps[1], steps[1] + 10, use_mask=True)
    prop.record(0, 0)
    prop.run(a, nt, steps[1], steps[1] + 10, use_mask=True)
    assert prop.record_read().shape == (0, 8)
    prop.close()
