"""GPU parity of the photoelectron amplitudes /lmd1 (spectrum_c::measure_lmd,
spectrum/momentum_1e.hpp:384-401): qpc_measure_lmd (one FP64 matrix-core
GEMM per l over the Coulomb-wave integrals) against the oracle's
restatement on the same Coulomb-wave data (qpc_setup_pcs, pinned against
mpmath and closed forms in test_pcs.py) and the same state."""
import numpy as np
import pytest

from _cfg import make_oracle, mixed_state, rel_l2, shape, upload_from_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,L,rmax,M,m0,nkr", [(300, 5, 60.0, 3, -1, 40), (1000, 31, 200.0, 1, 0, 100)])
def test_measure_lmd_matches_oracle(N, L, rmax, M, m0, nkr):
    import qpc_gpu
    cfg = shape(N, L, rmax, M=M, m0=m0)
    print(f"case: measure_lmd N={N} L={L} M={M} nkr={nkr}", flush=True)
    o = make_oracle(cfg)
    st = qpc_gpu.Setup(N, 7, L, M, m0, cfg["dt"], 0.0, rmax, 1.0, 1.0, cfg["mask_rmin"], rmax, 0.125)
    kr, ijl, cps = st.pcs(nkr, 0.02, 2.0)
    prop = qpc_gpu.Propagator(N, 7, L, M, m0, cfg["dt"])
    upload_from_oracle(prop, o)
    prop.upload_pcs(ijl, cps)
    c = mixed_state(o, seed=N + nkr, amp=0.5)
    prop.set_state(c)
    got = prop.measure_lmd()
    ref = o.measure_lmd(c, ijl, cps)
    err = rel_l2(got.ravel(), ref.ravel())
    print(f"  rel L2 = {err:.3e} (|lmd| max {np.abs(ref).max():.3e})", flush=True)
    assert err <= 1e-12
    prop.close()
