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
