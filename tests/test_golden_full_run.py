"""The whole-run fixture (tests/golden/full_run_h-N1000-L31.npz, written by
tests/golden/make_full_run_h_n1000.py) against the oracle and the physics
that fixes parts of it: the initial vector is the oracle's 1s state, the
first steps replay exactly, the 1s state carries no dipole, the norm
starts at 1 and only the absorber lowers it."""
import os

import numpy as np

from _cfg import H_N1000_L31, atable, make_oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "full_run_h-N1000-L31.npz")


def test_fixture_against_oracle_and_physics():
    g = np.load(GOLDEN)
    o = make_oracle(H_N1000_L31)
    nt, a = atable(H_N1000_L31)
    assert g["norm"].size == nt == 8819
    e, c0 = o.eigenstate(0, 0, 0)
    assert abs(e + 0.5) < 1e-9
    assert np.array_equal(c0, g["c0"])
    # norm 1 at the start, never rising (unitary steps, absorber), <z> = 0 for the 1s state
    assert abs(g["norm"][0] - 1.0) < 1e-12
    assert np.all(np.diff(g["norm"]) <= 1e-13)
    assert abs(g["dipole_z"][0]) < 1e-12
    # the first 200 steps replay bit for bit
    c = np.ascontiguousarray(g["c0"].copy())
    obs = o.run(c, a, nt, 0, 200, use_mask=True, observe=True)
    assert np.array_equal(obs[:, 0], g["norm"][:200])
    assert np.array_equal(obs[:, 1], g["dipole_z"][:200])
    # the final state is the one whose norm the history approaches
    assert abs(o.norm(g["final"]) - g["norm"][-1]) < 1e-4
