"""qpc_run keeps the state in the D-part eigenbasis between consecutive
with-field steps (the closing rotation of one step and the opening rotation
of the next enclose only the absorber, which acts alike on every channel);
a run that records observables before every step rotates back each step.
The two schedules are the same propagator and must agree to rounding, on
masked and unmasked runs that cross table batches (31 steps) and the pulse
end; the oracle comparison of the elided schedule is in test_gpu_parity.py."""
import numpy as np
import pytest

from _cfg import atable, make_oracle, mixed_state, rel_l2, shape, upload_from_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("use_mask", [True, False])
@pytest.mark.parametrize("N,L,rmax,M,m0", [(240, 5, 48.0, 3, -1), (1000, 31, 200.0, 1, 0)])
def test_elided_rotations_match_per_step_rotations(N, L, rmax, M, m0, use_mask):
    import qpc_gpu
    cfg = shape(N, L, rmax, M=M, m0=m0, nc=1.0)
    print(f"case: elision N={N} L={L} M={M} mask={use_mask}", flush=True)
    o = make_oracle(cfg)
    nt, a = atable(cfg)
    c0 = mixed_state(o, seed=5, amp=0.3)
    prop = qpc_gpu.Propagator(cfg["N"], cfg["p"], cfg["L"], M, m0, cfg["dt"])
    upload_from_oracle(prop, o)
    for it0, it1 in ((0, 70), (nt - 45, nt + 5)):
        finals = []
        for rec in (0, qpc_gpu.REC_NORM):
            prop.record(rec, it1 - it0 if rec else 0)
            prop.set_state(c0)
            prop.run(a, nt, it0, it1, use_mask=use_mask)
            finals.append(prop.get_state())
        err = rel_l2(finals[0], finals[1])
        print(f"  steps {it0}..{it1}: rel L2 elided vs per-step {err:.3e}", flush=True)
        assert err <= 1e-12
        assert np.all(np.isfinite(finals[0]))
    prop.close()


@pytest.mark.parametrize("gauge", ["vgwr", "lgwr"])
@pytest.mark.parametrize("isub", [-1, 2])
def test_ep_elided_x1_rotations_match_per_step(gauge, isub):
    """The elliptical steps inside qpc_run_ep / qpc_run_lgwr keep the state in
    x1's eigenbasis between with-field steps (tran_x1's eliminate and the next
    construct enclose only the absorber); with observables recorded every
    step they rotate back each step.  Both schedules agree to rounding, and
    the elided one with the oracle, across the pulse end with the absorber."""
    import pyoracle as po
    import qpc_gpu
    cfg = shape(300, 6, 60.0, M=5, m0=-2, nc=1.0)
    print(f"case: EP elision {gauge} N=300 L=6 M=5 isub={isub}", flush=True)
    o = make_oracle(cfg)
    prop = qpc_gpu.Propagator(cfg["N"], cfg["p"], cfg["L"], cfg["M"], cfg["m0"], cfg["dt"])
    upload_from_oracle(prop, o)
    nt = po.para_nt(cfg["w0"], cfg["nc"], 0.0, cfg["dt"])
    e0 = cfg["e0"] / np.sqrt(1.25) * (1.0 if gauge == "vgwr" else 0.057)
    fx = np.array([po.sinen2(e0, cfg["w0"], cfg["nc"], 0.0, 0.0, k * cfg["dt"]) for k in range(nt + 1)])
    fy = np.array([po.sinen2(0.5 * e0, cfg["w0"], cfg["nc"], 0.0, 0.5, k * cfg["dt"]) for k in range(nt + 1)])
    if gauge == "vgwr":
        ang = po.ep_angular(cfg["M"], cfg["L"], cfg["m0"])
        o.set_ep(ang, isub=isub)
        prop.upload_angular_ep(ang, isub=isub)
        run, orun = prop.run_ep, o.run_ep
    else:
        ang = po.lgwr_angular(cfg["M"], cfg["L"], cfg["m0"])
        o.set_lgwr(ang, isub=isub)
        prop.upload_angular_lgwr(ang, isub=isub)
        run, orun = prop.run_lgwr, o.run_lgwr
    c0 = mixed_state(o, seed=17, amp=0.3)
    if isub == 2:
        par = np.repeat(np.array([il % 2 for im in range(cfg["M"]) for il in range(cfg["L"])]), cfg["N"])
        c0 = np.where(par == 0, c0, 0.0)
        c0 /= np.sqrt(o.norm(c0))
    it0, it1 = nt - 30, nt + 6
    finals = []
    for rec in (0, qpc_gpu.REC_NORM):
        prop.record(rec, it1 - it0 if rec else 0)
        prop.set_state(c0)
        run(fx, fy, nt, it0, it1, use_mask=True)
        finals.append(prop.get_state())
    c = c0.copy()
    orun(c, fx, fy, nt, it0, it1, use_mask=True)
    err = rel_l2(finals[0], finals[1])
    eo = rel_l2(finals[0], c)
    print(f"  rel L2 elided vs per-step {err:.3e}, elided vs oracle {eo:.3e}", flush=True)
    assert err <= 1e-12 and eo <= 1e-12
    prop.close()
