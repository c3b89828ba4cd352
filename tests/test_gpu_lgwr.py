"""GPU parity of the elliptically polarised (m,l) step in the length gauge,
propagator_cn_hc_lgwr (cn_1e.hpp:717-810): the HIP path (FP64 matrix-core
rotations, device S + i e r sweeps) through the C ABI against the CPU oracle
on the same operands, angular data and initial vector.  Bar: relative L2 <=
1e-12 for single steps and short runs; at a large (m,l) grid the S-norm is
conserved by the unitary step to 1e-12."""
import numpy as np
import pytest

import pyoracle as po
from _cfg import make_oracle, mixed_state, rel_l2, shape, upload_from_oracle

pytestmark = pytest.mark.gpu


def _setup(cfg, isub):
    import qpc_gpu
    o = make_oracle(cfg)
    lg = po.lgwr_angular(cfg["M"], cfg["L"], cfg["m0"])
    o.set_lgwr(lg, isub=isub)
    prop = qpc_gpu.Propagator(cfg["N"], cfg["p"], cfg["L"], cfg["M"], cfg["m0"], cfg["dt"])
    upload_from_oracle(prop, o)
    prop.upload_angular_lgwr(lg, isub=isub)
    return o, prop


def _parity_state(o, cfg, pty, seed):
    c = mixed_state(o, seed=seed, amp=0.3)
    if pty is not None:
        par = np.repeat(np.array([il % 2 for im in range(cfg["M"]) for il in range(cfg["L"])]), cfg["N"])
        c = np.where(par == pty, c, 0.0)
        c /= np.sqrt(o.norm(c))
    return np.ascontiguousarray(c)


def _e_tables(cfg, ell=0.5):
    """E_x, E_y on the time grid: sin^2 envelope, ellipticity ell, CEP 0 and
    pi/2 (field.h:90-133 shapes, used here as the electric field)."""
    nt = po.para_nt(cfg["w0"], cfg["nc"], 0.0, cfg["dt"])
    e0 = cfg["e0"] / np.sqrt(1 + ell * ell)
    ex = np.array([po.sinen2(e0, cfg["w0"], cfg["nc"], 0.0, 0.0, k * cfg["dt"]) for k in range(nt + 1)])
    ey = np.array([po.sinen2(ell * e0, cfg["w0"], cfg["nc"], 0.0, 0.5, k * cfg["dt"]) for k in range(nt + 1)])
    return nt, ex, ey


@pytest.mark.parametrize("isub,pty", [(-1, None), (2, 0), (3, 1)])
def test_lgwr_single_step_matches_oracle(isub, pty):
    cfg = shape(240, 5, 48.0, M=3, m0=-1)
    print(f"case: lgwr single step N=240 L=5 M=3 m0=-1 isub={isub}", flush=True)
    o, prop = _setup(cfg, isub)
    c0 = _parity_state(o, cfg, pty, seed=37 + isub)
    for fields in [(0.05, 0.06, -0.03, 0.02), (-0.02, -0.01, 0.04, 0.035)]:
        c = c0.copy()
        o.step_lgwr(c, *fields)
        prop.set_state(c0)
        prop.propagate_lgwr(*fields)
        err = rel_l2(prop.get_state(), c)
        print(f"  fields {fields}: rel L2 = {err:.3e}", flush=True)
        assert err <= 1e-12


def test_lgwr_run_across_pulse_end():
    """The EP time loop (binsrc/sphTDSE.cpp:921-951) with E(t) over the end of
    a 1-cycle elliptical pulse, absorber every step, even sub-block."""
    cfg = shape(300, 6, 60.0, M=5, m0=-2, nc=1.0)
    o, prop = _setup(cfg, 2)
    nt, ex, ey = _e_tables(cfg)
    steps = (nt - 24, nt + 8)
    print(f"case: lgwr driver loop N=300 L=6 M=5 steps {steps}, pulse ends at {nt}", flush=True)
    c = _parity_state(o, cfg, 0, seed=43)
    prop.set_state(c)
    o.run_lgwr(c, ex, ey, nt, steps[0], steps[1], use_mask=True)
    prop.run_lgwr(ex, ey, nt, steps[0], steps[1], use_mask=True)
    err = rel_l2(prop.get_state(), c)
    print(f"  rel L2 = {err:.3e}", flush=True)
    assert err <= 1e-12


def test_lgwr_run_all_channels_m9_l9():
    """40 with-field steps at the pulse peak, all channels, M=9, L=9."""
    cfg = shape(400, 9, 80.0, M=9, m0=-4, nc=2.0)
    o, prop = _setup(cfg, -1)
    nt, ex, ey = _e_tables(cfg)
    it0 = nt // 2
    print(f"case: lgwr driver loop N=400 L=9 M=9 steps [{it0},{it0 + 40})", flush=True)
    c = mixed_state(o, seed=47, amp=0.1)
    prop.set_state(c)
    o.run_lgwr(c, ex, ey, nt, it0, it0 + 40, use_mask=True)
    prop.run_lgwr(ex, ey, nt, it0, it0 + 40, use_mask=True)
    err = rel_l2(prop.get_state(), c)
    print(f"  rel L2 = {err:.3e}", flush=True)
    assert err <= 1e-12


def test_lgwr_library_setup_matches_oracle_step():
    """The library's host set-up (qpc_setup_angular_lgwr) in place of the
    oracle's angular data: one step on the same state agrees to 1e-12 (the
    step depends only on the spectral projectors)."""
    import qpc_gpu
    cfg = shape(240, 6, 48.0, M=5, m0=-2)
    o, prop = _setup(cfg, 2)
    prop.upload_angular_lgwr(qpc_gpu.angular_lgwr(cfg["M"], cfg["L"], cfg["m0"]), isub=2)
    print("case: lgwr N=240 L=6 M=5 library angular set-up", flush=True)
    c0 = _parity_state(o, cfg, 0, seed=53)
    c = c0.copy()
    o.step_lgwr(c, 0.05, 0.055, 0.02, 0.025)
    prop.set_state(c0)
    prop.propagate_lgwr(0.05, 0.055, 0.02, 0.025)
    err = rel_l2(prop.get_state(), c)
    print(f"  rel L2 = {err:.3e}", flush=True)
    assert err <= 1e-12


@pytest.mark.timeout(600)
def test_lgwr_large_grid_is_unitary():
    """N=1500 (R=300), M=41 (m0=-20), L=41, even sub-block: three with-field
    steps conserve <c|S|c> (no absorber) and keep the odd sub-block empty."""
    import qpc_gpu
    N, L, M, m0 = 1500, 41, 41, -20
    print("case: lgwr N=1500 M=41 L=41 isub=2: host set-up", flush=True)
    st = qpc_gpu.Setup(N, 7, L, M, m0, 0.05, 0.0, 300.0, 1.0, 1.0, 280.0, 300.0, 0.125)
    lg = qpc_gpu.angular_lgwr(M, L, m0)
    prop = qpc_gpu.Propagator(N, 7, L, M, m0, 0.05)
    prop.upload_setup(st)
    prop.upload_angular_lgwr(lg, isub=2)
    _, c = st.eigenstate(40, 0, 0)
    rng = np.random.default_rng(59)
    par = np.repeat(np.array([il % 2 for im in range(M) for il in range(L)]), N)
    c = c + 1e-3 * np.where(par == 0, rng.standard_normal(c.size) + 1j * rng.standard_normal(c.size), 0.0)
    prop.set_state(c)
    n0 = prop.observe(qpc_gpu.QPC_OBS_NORM)
    print("  steps", flush=True)
    e0 = 0.0534 / np.sqrt(1.25)
    for k in range(3):
        prop.propagate_lgwr(e0 * (0.9 + 0.01 * k), e0 * (0.91 + 0.01 * k), 0.5 * e0 * 0.1, 0.5 * e0 * 0.12)
    n1 = prop.observe(qpc_gpu.QPC_OBS_NORM)
    g = prop.get_state()
    print(f"  norm {n0:.15f} -> {n1:.15f}, odd part {np.abs(g[par == 1]).max():.2e}", flush=True)
    assert abs(n1 - n0) <= 1e-12 * n0
    assert np.abs(g[par == 1]).max() < 1e-13


@pytest.mark.parametrize("isub,pty", [(-1, None), (2, 0)])
def test_lgwr_single_step_n_rank_8(isub, pty):
    """The reference's own basis order (n_rank = 8 in every config/*.cpp:28)."""
    cfg = shape(240, 5, 48.0, M=3, m0=-1, p=8)
    print(f"case: lgwr single step n_rank=8 N=240 L=5 M=3 isub={isub}", flush=True)
    o, prop = _setup(cfg, isub)
    c0 = _parity_state(o, cfg, pty, seed=91 + isub)
    c = c0.copy()
    o.step_lgwr(c, 0.05, 0.06, -0.03, 0.02)
    prop.set_state(c0)
    prop.propagate_lgwr(0.05, 0.06, -0.03, 0.02)
    err = rel_l2(prop.get_state(), c)
    print(f"  rel L2 = {err:.3e}", flush=True)
    assert err <= 1e-12
