"""Host logic of the learned Krylov controller's training pipeline
(qpc-tdse-cpc_amd/krylov_controller.py) against the oracle: the smallest
subspace meeting a tolerance read off one recorded Lanczos tridiagonal equals
the one the oracle's propagate1 reaches by building that many vectors; the
S-norm equals the oracle's norm; the seeded pulses cover the stated ranges."""
import numpy as np

import krylov_controller as kc
import pyoracle as po
from _cfg import mixed_state, shape, make_oracle


def test_min_subspace_matches_oracle_estimates():
    cfg = shape(200, 5, 40.0)
    o = make_oracle(cfg)
    c = mixed_state(o, seed=3, amp=0.1)
    c /= np.sqrt(o.norm(c))
    for az, hh, tol in [(0.7, 0.01, 1e-20), (-1.2, 0.02, 1e-16), (0.3, 0.005, 1e-24)]:
        g = c.copy()
        _, m, _, ab = o.krylov_step(g, az, hh, kmax=40)
        k = kc.min_subspace(ab, hh, tol, b0=np.sqrt(o.norm(c)))
        assert k is not None and 2 <= k < 40
        for kk, below in ((k, True), (k - 1, False)):
            g = c.copy()
            _, _, eps, _ = o.krylov_step(g, az, hh, kmax=kk)
            assert (eps <= tol) == below, (az, hh, tol, kk, eps)


def test_snorm_is_the_oracle_norm():
    o = po.Oracle(N=80, p=7, L=4, rmax=16.0)
    rng = np.random.default_rng(2)
    x = rng.standard_normal(320) + 1j * rng.standard_normal(320)
    assert abs(kc.snorm(o.get(po.QOR_S), x) ** 2 - o.norm(x)) < 1e-12 * o.norm(x)


def test_random_pulses_cover_the_ranges():
    p = kc.random_pulses(256, seed=5)
    assert p == kc.random_pulses(256, seed=5)
    e0 = np.array([q[0] * q[1] for q in p])
    nc = np.array([q[2] for q in p])
    cep = np.array([q[4] * np.pi for q in p])
    assert np.sqrt(5e13 / 3.51e16) <= e0.min() and e0.max() <= np.sqrt(5e14 / 3.51e16)
    assert 2.0 <= nc.min() and nc.max() <= 8.0 and 0.0 <= cep.min() and cep.max() < 2 * np.pi
    assert e0.max() / e0.min() > 2.5 and nc.max() - nc.min() > 5.0
