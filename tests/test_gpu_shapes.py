"""GPU parity at the shapes the reference's own configurations use and at
fields beyond any fixed carry depth, through the C ABI against the CPU oracle
on the same operands and states:

* n_rank = 8, the B-spline degree of every shipped config
  (config/example01_a.cpp:28, "do not change unless you know what may
  happen"): single steps and a driver-loop run;
* L = 161 channels (l_max = 160, config/example01_c.cpp:24), the wide
  rotation build;
* a fine grid (knot spacing 0.05) under a strong field, where the D-part
  sweeps' carries reach 16-20 partitions (planned from the operators), and a
  field so strong that no partitioned carry reaches far enough and the
  sweeps run sequentially."""
import numpy as np
import pytest

from _cfg import atable, make_oracle, mixed_state, rel_l2, shape, upload_from_oracle

import pyoracle as po

pytestmark = pytest.mark.gpu


def _prop(o, cfg):
    import qpc_gpu
    prop = qpc_gpu.Propagator(cfg["N"], cfg["p"], cfg["L"], cfg.get("M", 1), cfg.get("m0", 0), cfg["dt"])
    upload_from_oracle(prop, o)
    return prop


def _pieces(o, prop, label, a, seed, bar=1e-12):
    c0 = mixed_state(o, seed=seed)
    for name, fn_o, fn_g in [
        ("vgwz", lambda c: o.step(c, *a), lambda: prop.propagate(*a)),
        ("free", lambda c: o.step_free(c), lambda: prop.propagate()),
        ("mask", lambda c: o.mask(c), lambda: prop.mask()),
    ]:
        c = c0.copy()
        fn_o(c)
        prop.set_state(c0)
        fn_g()
        err = rel_l2(prop.get_state(), c)
        print(f"  {label} {name}: rel L2 = {err:.3e}", flush=True)
        assert err <= bar, (label, name, err)


@pytest.mark.parametrize("N,L,rmax,M,m0", [(400, 10, 80.0, 1, 0), (96, 4, 20.0, 1, 0), (240, 5, 48.0, 3, -1),
                                           (4000, 6, 800.0, 1, 0)])
def test_nrank8_single_step_pieces(N, L, rmax, M, m0):
    cfg = shape(N, L, rmax, M=M, m0=m0, p=8)
    print(f"case: n_rank=8 single step pieces N={N} L={L} M={M}", flush=True)
    o = make_oracle(cfg)
    prop = _prop(o, cfg)
    _pieces(o, prop, f"p=8 N={N} L={L} M={M}", (0.9, 1.1), seed=N)
    prop.close()


def test_nrank8_driver_run():
    """300 steps of a 4-cycle pulse on an N=1000, L=31, n_rank=8 grid, absorber
    every step, across three factorisation batches."""
    cfg = shape(1000, 31, 200.0, p=8)
    o = make_oracle(cfg)
    prop = _prop(o, cfg)
    nt, a = atable(cfg)
    c = mixed_state(o, seed=17)
    prop.set_state(c)
    steps = (1500, 1800)
    print(f"case: n_rank=8 N=1000 L=31 driver loop steps {steps}", flush=True)
    o.run(c, a, nt, steps[0], steps[1], use_mask=True)
    prop.run(a, nt, steps[0], steps[1], use_mask=True)
    err = rel_l2(prop.get_state(), c)
    print(f"  rel L2 = {err:.3e}", flush=True)
    assert err <= 1e-10
    prop.close()


def test_l161_single_step_pieces():
    """L = 161: the rotation build for up to 176 channels per m."""
    cfg = shape(240, 161, 48.0)
    print("case: L=161 single step pieces N=240", flush=True)
    o = make_oracle(cfg)
    prop = _prop(o, cfg)
    _pieces(o, prop, "N=240 L=161", (0.8, 0.9), seed=5)
    prop.close()


def _fine(L=8, p=7):
    # knot spacing 0.05: rmax = 0.05 (N + 2 - p)
    N = 1000
    return shape(N, L, 0.05 * (N + 2 - p), p=p)


@pytest.mark.parametrize("amp", [0.3, 1.0, 2.0])
def test_fine_grid_strong_field_deep_carry(amp):
    """Knot spacing 0.05 and |A| up to 2 a.u. (D-part scalar e/h up to ~0.5):
    carries of up to ~16 partitions, planned from the operators, against the
    oracle's sequential sweeps; single steps and a 20-step run of a field
    ramp."""
    import qpc_gpu
    cfg = _fine()
    o = make_oracle(cfg)
    pk, rs = qpc_gpu.coupling_plan(o.get(po.QOR_S), o.get(po.QOR_RDIF), cfg["N"], 7,
                                   1.25 * amp * 0.05 / 4 * np.abs(o.get(po.QOR_YEIG)).max())
    print(f"case: fine grid h=0.05 |A|={amp}: D-part plan pk={pk} restart={rs}", flush=True)
    assert pk <= 28
    prop = _prop(o, cfg)
    _pieces(o, prop, f"h=0.05 A={amp}", (amp * 0.9, amp), seed=3)
    atab = amp * np.sin(np.linspace(0.0, 1.5, 22))
    c = mixed_state(o, seed=4)
    prop.set_state(c)
    o.run(c, atab, 21, 0, 20, use_mask=True)
    prop.run(atab, 21, 0, 20, use_mask=True)
    err = rel_l2(prop.get_state(), c)
    print(f"  20-step ramp: rel L2 = {err:.3e}", flush=True)
    assert err <= 1e-12


def test_fine_grid_field_beyond_partitioned_reach():
    """|A| = 12 a.u. on the 0.05 grid (e/h ~ 3): the D-part sweeps run the
    sequential solve; single steps and a short run against the oracle."""
    import qpc_gpu
    cfg = _fine()
    o = make_oracle(cfg)
    amp = 12.0
    pk, rs = qpc_gpu.coupling_plan(o.get(po.QOR_S), o.get(po.QOR_RDIF), cfg["N"], 7,
                                   1.25 * amp * 0.05 / 4 * np.abs(o.get(po.QOR_YEIG)).max())
    print(f"case: fine grid h=0.05 |A|={amp}: D-part plan pk={pk} restart={rs}", flush=True)
    assert pk > 28
    prop = _prop(o, cfg)
    _pieces(o, prop, f"h=0.05 A={amp}", (amp, -amp * 0.8), seed=6)
    atab = amp * np.cos(np.linspace(0.0, 2.0, 8))
    c = mixed_state(o, seed=8)
    prop.set_state(c)
    o.run(c, atab, 7, 0, 7, use_mask=True)
    prop.run(atab, 7, 0, 7, use_mask=True)
    err = rel_l2(prop.get_state(), c)
    print(f"  7-step run: rel L2 = {err:.3e}", flush=True)
    assert err <= 1e-12
    prop.close()


def test_nrank8_fine_grid_strong_field():
    """n_rank = 8 with the deep planned carries (knot spacing 0.05, |A| = 1.5)."""
    import qpc_gpu
    cfg = _fine(L=6, p=8)
    o = make_oracle(cfg)
    pk, _ = qpc_gpu.coupling_plan(o.get(po.QOR_S), o.get(po.QOR_RDIF), cfg["N"], 8,
                                  1.25 * 1.5 * 0.05 / 4 * np.abs(o.get(po.QOR_YEIG)).max())
    print(f"case: n_rank=8 fine grid |A|=1.5: pk={pk}", flush=True)
    prop = _prop(o, cfg)
    _pieces(o, prop, "p=8 h=0.05 A=1.5", (1.2, 1.5), seed=9)
    prop.close()
