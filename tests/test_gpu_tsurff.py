"""GPU parity of t-SURFF (spectrum_c_tsf::accumulate, momentum_1e.hpp:609-719)
against the oracle's restatement (qor_tsurff_accumulate, pinned by the
free-packet and constant-field physics checks of test_tsurff.py): single
accumulates on seeded states (with and without a field term, several m), and
whole runs across the pulse end with t-SURFF inside qpc_run (after every
step, before the absorber, as binsrc/sphTDSE.cpp:798-804)."""
import numpy as np
import pytest

from _cfg import atable, make_oracle, mixed_state, rel_l2, shape, upload_from_oracle

pytestmark = pytest.mark.gpu


def _ctx(cfg):
    import qpc_gpu
    o = make_oracle(cfg)
    prop = qpc_gpu.Propagator(cfg["N"], cfg["p"], cfg["L"], cfg.get("M", 1), cfg.get("m0", 0), cfg["dt"])
    upload_from_oracle(prop, o)
    st = qpc_gpu.Setup(cfg["N"], cfg["p"], cfg["L"], cfg.get("M", 1), cfg.get("m0", 0), cfg["dt"], 0.0, cfg["rmax"],
                       1.0, 1.0, cfg["mask_rmin"], cfg["mask_rmax"], cfg["mask_fact"])
    return o, prop, st


@pytest.mark.parametrize("N,L,rmax,M,m0", [(400, 8, 80.0, 1, 0), (240, 5, 48.0, 3, -1)])
def test_tsurff_accumulate_matches_oracle(N, L, rmax, M, m0):
    cfg = shape(N, L, rmax, M=M, m0=m0)
    print(f"case: tsurff accumulate N={N} L={L} M={M}", flush=True)
    o, prop, st = _ctx(cfg)
    t = st.tsurff(0.6 * rmax, 0.1, 2.0, 33, 0.0, np.pi, 7, 0.0, 5.0, 2 if M > 1 else 1)
    prop.upload_tsurff(t)
    state, work = np.zeros(2), np.zeros((33, 7, 2 if M > 1 else 1), np.complex128)
    for seed, (a0, a1) in ((3, (0.0, 0.0)), (4, (0.7, -0.4)), (5, (1.3, 1.1))):
        c = mixed_state(o, seed=seed, amp=0.8)
        prop.set_state(c)
        prop.tsurff_accumulate(a0, a1)
        o.tsurff_accumulate(c, a0, a1, cfg["dt"], t, state, work)
    got, tim, disz = prop.tsurff_read()
    err = rel_l2(got.ravel(), work.ravel())
    print(f"  rel L2 {err:.3e}, time {tim} disz {disz}", flush=True)
    assert err <= 1e-12
    assert tim == state[0] and disz == state[1]


def test_tsurff_run_matches_oracle_across_pulse_end():
    cfg = shape(400, 8, 80.0, nc=1.0)
    print("case: tsurff run N=400 L=8 across the pulse end", flush=True)
    o, prop, st = _ctx(cfg)
    nt, a = atable(cfg)
    t = st.tsurff(50.0, 0.2, 2.5, 40, 0.0, np.pi, 9)
    prop.upload_tsurff(t)
    prop.tsurff(True)
    c = mixed_state(o, seed=11, amp=0.3)
    it0, it1 = nt - 150, nt + 60
    prop.set_state(c)
    prop.run(a, nt, it0, it1, use_mask=True)
    got, tim, disz = prop.tsurff_read()
    state, work = np.zeros(2), np.zeros((40, 9, 1), np.complex128)
    for it in range(it0, it1):
        if it < nt:
            o.step(c, a[it], a[it + 1])
            o.tsurff_accumulate(c, a[it], a[it + 1], cfg["dt"], t, state, work)
        else:
            o.step_free(c)
            o.tsurff_accumulate(c, 0.0, 0.0, cfg["dt"], t, state, work)
        o.mask(c)
    err = rel_l2(got.ravel(), work.ravel())
    serr = rel_l2(prop.get_state(), c)
    print(f"  work rel L2 {err:.3e}, state rel L2 {serr:.3e}", flush=True)
    assert serr <= 1e-11 and err <= 1e-10
    assert abs(tim - state[0]) < 1e-12 and abs(disz - state[1]) < 1e-12
    # the length gauge refuses t-SURFF (sphTDSE.cpp:288)
    import qpc_gpu
    with pytest.raises(qpc_gpu.QpcError):
        prop.run_lg(a, nt, it0, it0 + 1)
    prop.tsurff(False)
    prop.close()
