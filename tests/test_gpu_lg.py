"""GPU parity of the length-gauge propagator propagator_cn_hc_lgwz
(cn_1e.hpp:549-619, gauge 1 of the LP task) through the C ABI against the
CPU oracle (pinned by the dense Strang product and gauge invariance in
test_oracle_pins.py): single steps, a driver-loop run across factorisation
batches with the absorber, and the end of the pulse."""
import numpy as np
import pytest

from _cfg import H_N1000_L31, atable, make_oracle, mixed_state, rel_l2, shape, upload_from_oracle

pytestmark = pytest.mark.gpu


def _prop(o, cfg):
    import qpc_gpu
    prop = qpc_gpu.Propagator(cfg["N"], cfg["p"], cfg["L"], cfg.get("M", 1), cfg.get("m0", 0), cfg["dt"])
    upload_from_oracle(prop, o)
    return prop


def _etable(cfg):
    """E = -dA/dt at the time points of the sin^2 pulse (length-gauge laser values)."""
    nt, a = atable(cfg)
    t = cfg["dt"] * np.arange(nt + 1)
    e = -np.gradient(a, t, edge_order=2)
    e[0] = e[-1] = 0.0
    return nt, e


@pytest.mark.parametrize("N,L,rmax,M,m0,p", [(400, 10, 80.0, 1, 0, 7), (240, 5, 48.0, 3, -1, 7),
                                             (400, 10, 80.0, 1, 0, 8), (4000, 6, 800.0, 1, 0, 7)])
def test_lg_single_step(N, L, rmax, M, m0, p):
    cfg = shape(N, L, rmax, M=M, m0=m0, p=p)
    print(f"case: length-gauge step N={N} L={L} M={M} p={p}", flush=True)
    o = make_oracle(cfg)
    prop = _prop(o, cfg)
    c0 = mixed_state(o, seed=N + p)
    c = c0.copy()
    o.step_lg(c, 0.05, 0.07)
    prop.set_state(c0)
    prop.propagate_lg(0.05, 0.07)
    err = rel_l2(prop.get_state(), c)
    print(f"  rel L2 = {err:.3e}", flush=True)
    assert err <= 1e-12
    prop.close()


def test_lg_driver_run_n1000_l31():
    """300 steps of the N=1000, L=31 grid in length gauge around the pulse
    peak (E = -dA/dt of the 4-cycle pulse), absorber every step, recorded
    observables on the way."""
    import qpc_gpu
    cfg = H_N1000_L31
    o = make_oracle(cfg)
    prop = _prop(o, cfg)
    nt, e = _etable(cfg)
    c = mixed_state(o, seed=23)
    prop.set_state(c)
    steps = (4000, 4300)
    print(f"case: length gauge N=1000 L=31 driver loop steps {steps}", flush=True)
    prop.record(qpc_gpu.REC_NORM | qpc_gpu.REC_DIPOLE, steps[1] - steps[0])
    o.run_lg(c, e, nt, steps[0], steps[1], use_mask=True)
    prop.run_lg(e, nt, steps[0], steps[1], use_mask=True)
    err = rel_l2(prop.get_state(), c)
    h = prop.record_read()
    print(f"  rel L2 = {err:.3e}; recorded {h.shape[0]} rows, last norm {h[-1, 0]:.12f}", flush=True)
    assert err <= 1e-10
    assert h.shape == (steps[1] - steps[0], qpc_gpu.REC_COLUMNS)
    prop.close()


def test_lg_run_across_pulse_end():
    cfg = shape(400, 10, 80.0, nc=1.0)
    o = make_oracle(cfg)
    prop = _prop(o, cfg)
    nt, e = _etable(cfg)
    steps = (nt - 40, nt + 15)
    print(f"case: length gauge N=400 L=10 steps {steps}, pulse ends at {nt}", flush=True)
    c = mixed_state(o, seed=29)
    prop.set_state(c)
    o.run_lg(c, e, nt, steps[0], steps[1], use_mask=True)
    prop.run_lg(e, nt, steps[0], steps[1], use_mask=True)
    err = rel_l2(prop.get_state(), c)
    print(f"  rel L2 = {err:.3e}", flush=True)
    assert err <= 1e-12
    prop.close()
