"""GPU parity of batched shots (qpc_create_batch / qpc_run_batch): three
independent z-polarised runs of uneven size (L = 4, 6, 3) and different
pulses in one context, each against its own oracle run (a separate
reference process per shot), across the end of the shortest pulse with the
absorber.  Bar: relative L2 <= 1e-12 per shot; the padding channels stay
exactly zero."""
import numpy as np
import pytest

import pyoracle as po
from _cfg import mixed_state, rel_l2

pytestmark = pytest.mark.gpu

N, R, DT = 300, 60.0, 0.05
SHOTS = [dict(L=4, a0=0.0534 / 0.057, nc=1.0, c0=0.0), dict(L=6, a0=0.08 / 0.057, nc=1.5, c0=0.5),
         dict(L=3, a0=0.04 / 0.057, nc=1.2, c0=1.3)]


def test_batch_matches_per_shot_oracle():
    import qpc_gpu
    Lb = max(s["L"] for s in SHOTS)
    S = len(SHOTS)
    print(f"case: batch of {S} shots N={N} L={[s['L'] for s in SHOTS]}", flush=True)
    orc = [po.Oracle(N=N, p=7, L=s["L"], rmax=R, dt=DT, mask_rmin=54.0, mask_rmax=60.0) for s in SHOTS]
    big = orc[[s["L"] for s in SHOTS].index(Lb)]
    prop = qpc_gpu.BatchPropagator(N, 7, Lb, S, DT, time_batch=8)
    prop.upload_radial(big.get(po.QOR_S), big.get(po.QOR_CHOL), big.get(po.QOR_RINV), big.get(po.QOR_RDIF),
                       big.get(po.QOR_RPOW1), big.get(po.QOR_MASK))
    prop.upload_hc(big.get(po.QOR_UPPER), big.get(po.QOR_LOWER))
    prop.upload_angular_shots([dict(yeig=o.get(po.QOR_YEIG), yprj=o.get(po.QOR_YPRJ), qeig=o.get(po.QOR_QEIG),
                                    qprj=o.get(po.QOR_QPRJ), merged=o.get(po.QOR_MERGED), y10=o.get(po.QOR_Y10))
                               for o in orc])
    nts = [po.para_nt(0.057, s["nc"], 0.0, DT) for s in SHOTS]
    ntf = max(nts)
    atab = np.zeros((S, ntf + 1))
    for k, s in enumerate(SHOTS):
        atab[k, :nts[k] + 1] = [po.sinen2(s["a0"], 0.057, s["nc"], 0.0, s["c0"], j * DT) for j in range(nts[k] + 1)]
    cs = [mixed_state(o, seed=50 + k, amp=0.05) for k, o in enumerate(orc)]
    state = np.zeros(S * Lb * N, np.complex128)
    for k, s in enumerate(SHOTS):
        state[k * Lb * N:k * Lb * N + s["L"] * N] = cs[k]
    prop.set_state(state)
    it0, it1 = min(nts) - 20, min(nts) + 30
    print(f"  steps [{it0},{it1}), pulse ends {nts}", flush=True)
    for k, o in enumerate(orc):
        o.run(cs[k], atab[k], nts[k], it0, it1, use_mask=True)
    prop.run_batch(atab, ntf, it0, it1, use_mask=True)
    g = prop.get_state()
    norms = prop.observe_batch(qpc_gpu.QPC_OBS_NORM)
    for k, s in enumerate(SHOTS):
        blk = g[k * Lb * N:(k + 1) * Lb * N]
        err = rel_l2(blk[:s["L"] * N], cs[k])
        print(f"  shot {k} (L={s['L']}): rel L2 = {err:.3e}, norm {norms[k]:.12f} / {orc[k].norm(cs[k]):.12f}",
              flush=True)
        assert err <= 1e-12
        assert np.abs(blk[s["L"] * N:]).max(initial=0.0) == 0.0
        assert abs(norms[k] - orc[k].norm(cs[k])) <= 1e-12
