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
