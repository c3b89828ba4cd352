"""The launch-shape autotuner (qpc_set_tune / qpc_autotune, no counterpart in
the reference): every launch shape it chooses from gives the same bits (the
segments are exact and the staged and two-kernel solve paths share their
arithmetic), the tuner restores the state and keeps a candidate, and the
tuned run still matches the oracle."""
import numpy as np
import pytest

from _cfg import H_N1000_L31, atable, make_oracle, mixed_state, rel_l2, upload_from_oracle

pytestmark = pytest.mark.gpu


def _prop(o, cfg):
    import qpc_gpu
    prop = qpc_gpu.Propagator(cfg["N"], cfg["p"], cfg["L"], 1, 0, cfg["dt"])
    upload_from_oracle(prop, o)
    return prop


def test_launch_shapes_give_identical_bits_and_tuner_restores_state():
    import qpc_gpu
    cfg = H_N1000_L31
    print("case: launch shapes N=1000 L=31, 20 with-field steps each", flush=True)
    o = make_oracle(cfg)
    prop = _prop(o, cfg)
    nt, a = atable(cfg)
    c0 = mixed_state(o, seed=17, amp=0.05)
    it0 = nt // 2
    ref = None
    # stage_grid 0: every solve reads its transfer sets from L2 with the
    # right-hand side as a separate launch
    for sa in (0, 1, 2, 3):
        for sg in (512, 0):
            prop.set_tune(sa, sg)
            t = prop.get_tune()
            assert (t["seg_add"], t["stage_grid"]) == (sa, sg)
            prop.set_state(c0)
            prop.run(a, nt, it0, it0 + 20, use_mask=True)
            got = prop.get_state()
            if ref is None:
                ref = got
            print(f"  seg_add {sa}, stage_grid {sg}: max |diff| {np.abs(got - ref).max():.1e}", flush=True)
            assert np.array_equal(got, ref)
    c = c0.copy()
    o.run(c, a, nt, it0, it0 + 20, use_mask=True)
    assert rel_l2(ref, c) <= 1e-12
    # the tuner: state untouched, a candidate kept, every candidate timed
    prop.set_tune(0, 512)
    prop.set_state(c0)
    choice, ms = prop.autotune(a, nt, it0, 16)
    print(f"  autotune: {choice}, ms per step {ms}", flush=True)
    assert np.array_equal(prop.get_state(), c0)
    assert (choice["seg_add"], choice["stage_grid"]) in qpc_gpu.TUNE_CANDIDATES
    assert prop.get_tune() == choice
    assert all(v > 0 for v in ms.values()) and len(ms) == 8
    prop.run(a, nt, it0, it0 + 20, use_mask=True)
    assert np.array_equal(prop.get_state(), ref)
    prop.close()


def test_batch_autotune_keeps_results():
    """A batch context tuned on its own shots (qpc_autotune with the field
    table's stride) runs the same bits afterwards."""
    import qpc_gpu
    import pyoracle as po
    N, R, DT = 300, 60.0, 0.05
    shots = [dict(L=4, a0=0.0534 / 0.057, nc=1.0, c0=0.0), dict(L=6, a0=0.08 / 0.057, nc=1.5, c0=0.5)]
    print("case: batch autotune, 2 shots N=300", flush=True)
    orc = [po.Oracle(N=N, p=7, L=s["L"], rmax=R, dt=DT, mask_rmin=54.0, mask_rmax=60.0) for s in shots]
    big = orc[1]
    prop = qpc_gpu.BatchPropagator(N, 7, 6, 2, DT, time_batch=8)
    prop.upload_radial(big.get(po.QOR_S), big.get(po.QOR_CHOL), big.get(po.QOR_RINV), big.get(po.QOR_RDIF),
                       big.get(po.QOR_RPOW1), big.get(po.QOR_MASK))
    prop.upload_hc(big.get(po.QOR_UPPER), big.get(po.QOR_LOWER))
    prop.upload_angular_shots([dict(yeig=o.get(po.QOR_YEIG), yprj=o.get(po.QOR_YPRJ), qeig=o.get(po.QOR_QEIG),
                                    qprj=o.get(po.QOR_QPRJ), merged=o.get(po.QOR_MERGED), y10=o.get(po.QOR_Y10))
                               for o in orc])
    nts = [po.para_nt(0.057, s["nc"], 0.0, DT) for s in shots]
    atab = np.zeros((2, max(nts) + 1))
    for k, s in enumerate(shots):
        atab[k, :nts[k] + 1] = [po.sinen2(s["a0"], 0.057, s["nc"], 0.0, s["c0"], j * DT) for j in range(nts[k] + 1)]
    state = np.zeros(2 * 6 * N, np.complex128)
    cs = [mixed_state(o, seed=80 + k, amp=0.05) for k, o in enumerate(orc)]
    for k, s in enumerate(shots):
        state[k * 6 * N:k * 6 * N + s["L"] * N] = cs[k]
    it0 = min(nts) // 2
    prop.set_state(state)
    prop.run_batch(atab, max(nts), it0, it0 + 12)
    ref = prop.get_state()
    prop.set_state(state)
    choice, ms = prop.autotune_batch(atab, max(nts), it0, 12)
    print(f"  batch autotune: {choice}", flush=True)
    assert np.array_equal(prop.get_state(), state)
    prop.run_batch(atab, max(nts), it0, it0 + 12)
    assert np.array_equal(prop.get_state(), ref)
    prop.close()
