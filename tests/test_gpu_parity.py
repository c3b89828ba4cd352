"""GPU parity: the HIP propagator against the CPU oracle on the same operands
and the same initial vector.  Bar: relative L2 of the coefficients <= 1e-10
for the CN integrator, tighter per single step."""
import sys

import numpy as np
import pytest

import pyoracle as po
from _cfg import CFG1, CFG2, atable, make_oracle, mixed_state, rel_l2, upload_from_oracle

pytestmark = pytest.mark.gpu


def _gpu_prop(o, cfg):
    import qpc_gpu
    prop = qpc_gpu.Propagator(cfg["N"], cfg["p"], cfg["L"], 1, 0, cfg["dt"])
    upload_from_oracle(prop, o)
    return prop


@pytest.fixture(scope="module")
def cfg1():
    o = make_oracle(CFG1)
    return o, _gpu_prop(o, CFG1)


def test_single_step_pieces(cfg1):
    o, prop = cfg1
    print("case: config-1 shape single with-field step / free step / mask", flush=True)
    c0 = mixed_state(o, seed=3)
    for name, fn_o, fn_g in [
        ("vgwz", lambda c: o.step(c, 0.7, 0.8), lambda: prop.propagate(0.7, 0.8)),
        ("free", lambda c: o.step_free(c), lambda: prop.propagate()),
        ("mask", lambda c: o.mask(c), lambda: prop.mask()),
    ]:
        c = c0.copy()
        fn_o(c)
        prop.set_state(c0)
        fn_g()
        g = prop.get_state()
        err = rel_l2(g, c)
        print(f"  {name}: rel L2 = {err:.3e}", flush=True)
        assert err <= 1e-12, (name, err)


def test_observables(cfg1):
    o, prop = cfg1
    c = mixed_state(o, seed=5, amp=0.3)
    prop.set_state(c)
    import qpc_gpu
    n_g = prop.observe(qpc_gpu.QPC_OBS_NORM)
    z_g = prop.observe(qpc_gpu.QPC_OBS_DIPOLE_Z)
    assert abs(n_g - o.norm(c)) <= 1e-13
    assert abs(z_g - o.dipole_z(c, 0)) <= 1e-12 * max(1.0, abs(o.dipole_z(c, 0)))


def test_driver_loop_config1_prefix(cfg1):
    """First 400 steps of config 1 (pulse on), mask every step."""
    o, prop = cfg1
    nt, a = atable(CFG1)
    assert nt == 8819
    c = mixed_state(o, seed=7)
    prop.set_state(c)
    steps = (0, 400)
    print(f"case: config-1 driver loop steps {steps}", flush=True)
    o.run(c, a, nt, steps[0], steps[1], use_mask=True)
    prop.run(a, nt, steps[0], steps[1], use_mask=True)
    g = prop.get_state()
    err = rel_l2(g, c)
    print(f"  rel L2 = {err:.3e}", flush=True)
    assert err <= 1e-10


def test_config2_shape_steps():
    """Config-2 operators (N=3000, L=81): 3 with-field steps around the pulse peak + mask."""
    o = make_oracle(CFG2)
    prop = _gpu_prop(o, CFG2)
    nt, a = atable(CFG2)
    assert nt == 17638
    c = mixed_state(o, seed=11)
    prop.set_state(c)
    it0 = nt // 2
    print(f"case: config-2 driver loop steps [{it0},{it0 + 3})", flush=True)
    o.run(c, a, nt, it0, it0 + 3, use_mask=True)
    prop.run(a, nt, it0, it0 + 3, use_mask=True)
    err = rel_l2(prop.get_state(), c)
    print(f"  rel L2 = {err:.3e}", flush=True)
    assert err <= 1e-12
