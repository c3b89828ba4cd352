"""GPU parity of the elliptically polarised (m,l) step propagator_cn_hc_vgwr
(cn_1e.hpp:812-951): the HIP path (FP64 matrix-core GEMM rotations, device
coupling sweeps) through the C ABI against the CPU oracle on the same
operands, angular eigen-data and initial vector.  Bar: relative L2 <= 1e-12
for single steps and short runs; at the full elliptical-polarisation size
(N=1500, M=81, L=41, even sub-block) the S-norm is conserved by the unitary
step to 1e-12."""
import numpy as np
import pytest

import pyoracle as po
from _cfg import make_oracle, mixed_state, rel_l2, shape, upload_from_oracle

pytestmark = pytest.mark.gpu


def _setup(cfg, isub):
    import qpc_gpu
    o = make_oracle(cfg)
    ep = po.ep_angular(cfg["M"], cfg["L"], cfg["m0"])
    o.set_ep(ep, isub=isub)
    prop = qpc_gpu.Propagator(cfg["N"], cfg["p"], cfg["L"], cfg["M"], cfg["m0"], cfg["dt"])
    upload_from_oracle(prop, o)
    prop.upload_angular_ep(ep, isub=isub)
    return o, prop


def _parity_state(o, cfg, pty, seed):
    """mixed_state restricted to one parity of il (the sub-block a gathered
    propagate<2|3> keeps), S-normalised."""
    c = mixed_state(o, seed=seed, amp=0.3)
    if pty is not None:
        par = np.repeat(np.array([il % 2 for im in range(cfg["M"]) for il in range(cfg["L"])]), cfg["N"])
        c = np.where(par == pty, c, 0.0)
        c /= np.sqrt(o.norm(c))
    return np.ascontiguousarray(c)


def _ep_tables(cfg, ell=0.5):
    """A_x, A_y on the time grid: sin^2 pulse, ellipticity ell, x/y components
    E0/sqrt(1+ell^2) and ell*E0/sqrt(1+ell^2), CEP 0 and pi/2 (field.h:90-133)."""
    nt = po.para_nt(cfg["w0"], cfg["nc"], 0.0, cfg["dt"])
    a0 = cfg["e0"] / np.sqrt(1 + ell * ell)
    ax = np.array([po.sinen2(a0, cfg["w0"], cfg["nc"], 0.0, 0.0, k * cfg["dt"]) for k in range(nt + 1)])
    ay = np.array([po.sinen2(ell * a0, cfg["w0"], cfg["nc"], 0.0, 0.5, k * cfg["dt"]) for k in range(nt + 1)])
    return nt, ax, ay


@pytest.mark.parametrize("isub,pty", [(-1, None), (2, 0), (3, 1)])
def test_ep_single_step_matches_oracle(isub, pty):
    cfg = shape(240, 5, 48.0, M=3, m0=-1)
    print(f"case: EP single step N=240 L=5 M=3 m0=-1 isub={isub}", flush=True)
    o, prop = _setup(cfg, isub)
    c0 = _parity_state(o, cfg, pty, seed=17 + isub)
    for fields in [(0.9, 1.1, -0.5, 0.3), (-0.4, -0.2, 0.8, 0.7)]:
        c = c0.copy()
        o.step_ep(c, *fields)
        prop.set_state(c0)
        prop.propagate_ep(*fields)
        err = rel_l2(prop.get_state(), c)
        print(f"  fields {fields}: rel L2 = {err:.3e}", flush=True)
        assert err <= 1e-12


def test_ep_run_across_pulse_end():
    """The EP time loop (binsrc/sphTDSE.cpp:921-951) over the end of a 1-cycle
    elliptical pulse, absorber every step, even sub-block (propflag_eve)."""
    cfg = shape(300, 6, 60.0, M=5, m0=-2, nc=1.0)
    o, prop = _setup(cfg, 2)
    nt, ax, ay = _ep_tables(cfg)
    steps = (nt - 24, nt + 8)
    print(f"case: EP driver loop N=300 L=6 M=5 steps {steps}, pulse ends at {nt}", flush=True)
    c = _parity_state(o, cfg, 0, seed=23)
    prop.set_state(c)
    o.run_ep(c, ax, ay, nt, steps[0], steps[1], use_mask=True)
    prop.run_ep(ax, ay, nt, steps[0], steps[1], use_mask=True)
    err = rel_l2(prop.get_state(), c)
    print(f"  rel L2 = {err:.3e}", flush=True)
    assert err <= 1e-12


def test_ep_run_all_channels_m9_l9():
    """40 with-field steps at the pulse peak, all channels (isub = -1), M=9, L=9."""
    cfg = shape(400, 9, 80.0, M=9, m0=-4, nc=2.0)
    o, prop = _setup(cfg, -1)
    nt, ax, ay = _ep_tables(cfg)
    it0 = nt // 2
    print(f"case: EP driver loop N=400 L=9 M=9 steps [{it0},{it0 + 40})", flush=True)
    c = mixed_state(o, seed=29, amp=0.1)
    prop.set_state(c)
    o.run_ep(c, ax, ay, nt, it0, it0 + 40, use_mask=True)
    prop.run_ep(ax, ay, nt, it0, it0 + 40, use_mask=True)
    err = rel_l2(prop.get_state(), c)
    print(f"  rel L2 = {err:.3e}", flush=True)
    assert err <= 1e-12


@pytest.mark.timeout(600)
def test_ep_full_size_is_unitary():
    """N=1500 (R=300), M=81 (m0=-40), L=41, even sub-block: 3321 channels,
    1721-row rotations; three with-field steps at the pulse peak conserve
    <c|S|c> (no absorber) and keep the odd sub-block empty."""
    import qpc_gpu
    N, L, M, m0 = 1500, 41, 81, -40
    print("case: EP full size N=1500 M=81 L=41 isub=2: host set-up", flush=True)
    st = qpc_gpu.Setup(N, 7, L, M, m0, 0.05, 0.0, 300.0, 1.0, 1.0, 280.0, 300.0, 0.125)
    ep = qpc_gpu.angular_ep(M, L, m0)
    prop = qpc_gpu.Propagator(N, 7, L, M, m0, 0.05)
    prop.upload_setup(st)
    prop.upload_angular_ep(ep, isub=2)
    _, c = st.eigenstate(40, 0, 0)  # 1s: m = 0, l = 0
    rng = np.random.default_rng(31)
    par = np.repeat(np.array([il % 2 for im in range(M) for il in range(L)]), N)
    c = c + 1e-3 * np.where(par == 0, rng.standard_normal(c.size) + 1j * rng.standard_normal(c.size), 0.0)
    prop.set_state(c)
    n0 = prop.observe(qpc_gpu.QPC_OBS_NORM)
    print("  steps", flush=True)
    a0 = 0.0534 / 0.057 / np.sqrt(1.25)
    for k in range(3):
        prop.propagate_ep(a0 * (0.9 + 0.01 * k), a0 * (0.91 + 0.01 * k), 0.5 * a0 * 0.1, 0.5 * a0 * 0.12)
    n1 = prop.observe(qpc_gpu.QPC_OBS_NORM)
    g = prop.get_state()
    print(f"  norm {n0:.15f} -> {n1:.15f}, odd part {np.abs(g[par == 1]).max():.2e}", flush=True)
    assert abs(n1 - n0) <= 1e-12 * n0
    assert np.abs(g[par == 1]).max() < 1e-13


@pytest.mark.parametrize("isub,pty", [(-1, None), (2, 0)])
def test_ep_single_step_n_rank_8(isub, pty):
    """The reference's own basis order (n_rank = 8 in every config/*.cpp:28)."""
    cfg = shape(240, 5, 48.0, M=3, m0=-1, p=8)
    print(f"case: ep single step n_rank=8 N=240 L=5 M=3 isub={isub}", flush=True)
    o, prop = _setup(cfg, isub)
    c0 = _parity_state(o, cfg, pty, seed=91 + isub)
    c = c0.copy()
    o.step_ep(c, 0.9, 1.1, -0.5, 0.3)
    prop.set_state(c0)
    prop.propagate_ep(0.9, 1.1, -0.5, 0.3)
    err = rel_l2(prop.get_state(), c)
    print(f"  rel L2 = {err:.3e}", flush=True)
    assert err <= 1e-12


def test_ep_field_beyond_partitioned_reach():
    """|A| = 12 a.u. on a knot spacing of 0.05 (N=1000, D-part scalar e/h ~ 3):
    the x1 and y1 D-type sweeps plan a carry deeper than the partitioned solve
    takes and run the sequential solve (the work-group order is not used
    there); single steps against the oracle's sequential sweeps."""
    import qpc_gpu
    N, amp = 1000, 12.0
    cfg = shape(N, 4, 0.05 * (N + 2 - 7), M=3, m0=-1)
    o = make_oracle(cfg)
    ep = po.ep_angular(cfg["M"], cfg["L"], cfg["m0"])
    lam = max(np.abs(ep["eig_x1"]).max(), np.abs(ep["eig_y1"]).max())
    pk, rs = qpc_gpu.coupling_plan(o.get(po.QOR_S), o.get(po.QOR_RDIF), N, 7, 1.25 * amp * cfg["dt"] / 4 * lam)
    print(f"case: EP h=0.05 |A|={amp} N={N} L=4 M=3: D-part plan pk={pk} restart={rs}", flush=True)
    assert pk > 28
    o.set_ep(ep, isub=-1)
    prop = qpc_gpu.Propagator(cfg["N"], cfg["p"], cfg["L"], cfg["M"], cfg["m0"], cfg["dt"])
    upload_from_oracle(prop, o)
    prop.upload_angular_ep(ep, isub=-1)
    c0 = mixed_state(o, seed=41, amp=0.3)
    for fields in [(amp, 0.9 * amp, -0.8 * amp, -0.7 * amp), (-0.6 * amp, -amp, 0.3 * amp, 0.5 * amp)]:
        c = c0.copy()
        o.step_ep(c, *fields)
        prop.set_state(c0)
        prop.propagate_ep(*fields)
        err = rel_l2(prop.get_state(), c)
        print(f"  fields {fields}: rel L2 = {err:.3e}", flush=True)
        assert err <= 1e-12
    prop.close()
