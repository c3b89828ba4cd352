"""Shared R-part tables: the coupling sweep (S + i e 1/r) at -e is the complex
conjugate of the one at e, and neither depends on m, so channels whose
eigenvalues have the same magnitude share one factorised table (a channel
of the opposite sign is solved as conj(A^-1 B conj(z))).  Complex
arithmetic commutes with conjugation exactly, so sharing changes no bit:
runs with shared tables against one table per channel, in the solve that
forms its right-hand side in the launch and in the two-launch path, on one
m block (L = 8: 4 tables) and on three (m = -1, 0, 1: the m = +-1 blocks
share as well), with the eigenvalues made exact pairs as the library's
set-up makes them, and against the oracle."""
import numpy as np
import pytest

import pyoracle as po
from _cfg import atable, make_oracle, mixed_state, rel_l2, shape

pytestmark = pytest.mark.gpu


def _paired(q, L):
    """The ascending eigenvalues of each m block as exact +- pairs."""
    q = q.copy().reshape(-1, L)
    for b in q:
        o = np.argsort(b)
        for k in range(L // 2):
            v = 0.5 * (b[o[L - 1 - k]] - b[o[k]])
            b[o[k]], b[o[L - 1 - k]] = -v, v
        if L % 2:
            b[o[L // 2]] = 0.0
    return q.ravel()


@pytest.mark.parametrize("M,m0", [(1, 0), (3, -1)])
def test_shared_tables_give_identical_bits(M, m0):
    import qpc_gpu
    L = 8
    cfg = shape(400, L, 80.0, M=M, m0=m0)
    print(f"case: shared R-part tables N=400 L={L} M={M} m0={m0}", flush=True)
    o = make_oracle(cfg)
    prop = qpc_gpu.Propagator(cfg["N"], cfg["p"], L, M, m0, cfg["dt"])
    prop.upload_radial(o.get(po.QOR_S), o.get(po.QOR_CHOL), o.get(po.QOR_RINV), o.get(po.QOR_RDIF),
                       o.get(po.QOR_RPOW1), o.get(po.QOR_MASK))
    prop.upload_hc(o.get(po.QOR_UPPER), o.get(po.QOR_LOWER))
    q = _paired(o.get(po.QOR_QEIG), L)
    assert np.abs(q - o.get(po.QOR_QEIG)).max() < 1e-12
    prop.upload_angular(o.get(po.QOR_YEIG), o.get(po.QOR_YPRJ), q, o.get(po.QOR_QPRJ), o.get(po.QOR_MERGED),
                        o.get(po.QOR_Y10))
    t = prop.get_tune()
    distinct = len(set(np.abs(q)))
    print(f"  {t['tables_r']} R-part tables for {M * L} channels", flush=True)
    assert t["share_r"] == 1 and t["tables_r"] == distinct < M * L
    nt, a = atable(cfg)
    c0 = mixed_state(o, seed=23, amp=0.1)
    it0 = nt // 2 - 20
    out = {}
    for share in (1, 0):
        for sg in (512, 0):
            prop.set_tune(0, sg, 0, share)
            prop.set_state(c0)
            prop.run(a, nt, it0, it0 + 40, use_mask=True)
            for k in range(3):  # and the per-step entry point
                prop.propagate(a[it0 + 40 + k], a[it0 + 41 + k])
                prop.mask()
            out[(share, sg)] = prop.get_state()
    assert prop.get_tune()["tables_r"] == M * L  # sharing off
    ref = out[(0, 512)]
    for key, got in out.items():
        print(f"  share {key[0]}, stage_grid {key[1]}: max |diff| {np.abs(got - ref).max():.1e}", flush=True)
        assert np.array_equal(got, ref), key
    # against the oracle's run with its own (unpaired, to ~1e-14) eigenvalues
    c = c0.copy()
    o.run(c, a, nt, it0, it0 + 40, use_mask=True)
    for k in range(3):
        o.step(c, a[it0 + 40 + k], a[it0 + 41 + k])
        o.mask(c)
    err = rel_l2(ref, c)
    print(f"  vs oracle: rel L2 = {err:.3e}", flush=True)
    assert err <= 1e-12
    prop.close()


def test_shared_tables_in_a_batch_give_identical_bits():
    """Batch contexts share tables within a shot (its own A(t)): two shots of
    L = 5 and 6 in one context with exact-pair eigenvalues, shared against
    per-channel tables, bit for bit."""
    import qpc_gpu
    N, R, DT = 300, 60.0, 0.05
    shots = [dict(L=5, a0=0.0534 / 0.057, nc=1.0, c0=0.0), dict(L=6, a0=0.08 / 0.057, nc=1.5, c0=0.5)]
    print("case: shared R-part tables in a batch of 2 shots N=300", flush=True)
    orc = [po.Oracle(N=N, p=7, L=s["L"], rmax=R, dt=DT, mask_rmin=54.0, mask_rmax=60.0) for s in shots]
    big = orc[1]
    prop = qpc_gpu.BatchPropagator(N, 7, 6, 2, DT, time_batch=8)
    prop.upload_radial(big.get(po.QOR_S), big.get(po.QOR_CHOL), big.get(po.QOR_RINV), big.get(po.QOR_RDIF),
                       big.get(po.QOR_RPOW1), big.get(po.QOR_MASK))
    prop.upload_hc(big.get(po.QOR_UPPER), big.get(po.QOR_LOWER))
    prop.upload_angular_shots([dict(yeig=o.get(po.QOR_YEIG), yprj=o.get(po.QOR_YPRJ),
                                    qeig=_paired(o.get(po.QOR_QEIG), s["L"]), qprj=o.get(po.QOR_QPRJ),
                                    merged=o.get(po.QOR_MERGED), y10=o.get(po.QOR_Y10))
                               for o, s in zip(orc, shots)])
    t = prop.get_tune()
    print(f"  {t['tables_r']} R-part tables for 12 channels", flush=True)
    assert t["tables_r"] < 12
    nts = [po.para_nt(0.057, s["nc"], 0.0, DT) for s in shots]
    atab = np.zeros((2, max(nts) + 1))
    for k, s in enumerate(shots):
        atab[k, :nts[k] + 1] = [po.sinen2(s["a0"], 0.057, s["nc"], 0.0, s["c0"], j * DT) for j in range(nts[k] + 1)]
    state = np.zeros(2 * 6 * N, np.complex128)
    for k, (o, s) in enumerate(zip(orc, shots)):
        state[k * 6 * N:k * 6 * N + s["L"] * N] = mixed_state(o, seed=90 + k, amp=0.05)
    it0 = min(nts) // 2 - 10
    out = []
    for share in (1, 0):
        prop.set_tune(0, 512, 0, share)
        prop.set_state(state)
        prop.run_batch(atab, max(nts), it0, it0 + 30)
        out.append(prop.get_state())
    print(f"  max |diff| {np.abs(out[0] - out[1]).max():.1e}", flush=True)
    assert np.array_equal(out[0], out[1])
    prop.close()
