"""The benchmark workload end to end: all 17638 steps of the N=3000, L=81
run (8-cycle pulse, absorber every step) from the 1s state on the device,
with the oracle's operators uploaded, against the oracle's own run stored in
tests/golden/full_run_h-N3000-L81.npz (written by
tests/golden/make_full_run_h_n3000.py): the final coefficients and the norm
and <z> recorded before every step (qpc_record, the driver's obsv::observe
order)."""
import os

import numpy as np
import pytest

from _cfg import H_N3000_L81, atable, make_oracle, rel_l2, upload_from_oracle

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "full_run_h-N3000-L81.npz")


@pytest.mark.skipif(not os.path.exists(GOLDEN), reason="fixture not generated")
def test_full_run_n3000_l81_against_golden():
    import qpc_gpu
    cfg = H_N3000_L81
    g = np.load(GOLDEN)
    o = make_oracle(cfg)
    prop = qpc_gpu.Propagator(cfg["N"], cfg["p"], cfg["L"], 1, 0, cfg["dt"])
    upload_from_oracle(prop, o)
    nt, a = atable(cfg)
    print(f"case: N=3000 L=81 whole run, {nt} steps from the 1s state", flush=True)
    prop.record(qpc_gpu.REC_NORM | qpc_gpu.REC_DIPOLE, nt)
    prop.set_state(g["c0"])
    prop.run(a, nt, 0, nt, use_mask=True)
    hist = prop.record_read()
    err = rel_l2(prop.get_state(), g["final"])
    dn = np.abs(hist[:, 0] - g["norm"]).max()
    dz = np.abs(hist[:, 1] - g["dipole_z"]).max()
    print(f"  final rel L2 {err:.3e}, norm history max diff {dn:.3e}, <z> history max diff {dz:.3e} "
          f"(max |<z>| {np.abs(g['dipole_z']).max():.3f}), final norm {g['norm'][-1]:.6f}", flush=True)
    assert hist.shape[0] == nt
    assert err <= 1e-10
    assert dn <= 1e-11
    assert dz <= 1e-10 * np.abs(g["dipole_z"]).max()
    prop.close()
