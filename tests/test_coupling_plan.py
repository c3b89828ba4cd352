"""The D-part coupling sweeps (S + e d/dr, propagate.hpp:3-98 with
symflag 1) run as partitioned solves whose carry depth and factorisation
restart the library plans from the operators (qpc_coupling_plan): the reach
of the factors' inverses grows with |e|/h (h the knot spacing), so a fixed
depth is exact only on a range of fields.  These checks hold the plan to the
same bars as the fixed depths of test_carry_depth.py and
test_restart_factor.py, on grids of three spacings and scalars from weak to
far beyond any laser field: the dropped carry tail <= 1e-17 at the planned
depth (measured densely in numpy, independent of the library's own
measurement), the restarted factorisation within 1e-14 of the sequential
elimination at the planned restart, and the sequential solve chosen when no
partitioned carry reaches far enough."""
import numpy as np
import pytest

import pyoracle as po
import qpc_gpu
from test_carry_depth import _dense, _tail
from test_restart_factor import _restart_error

PKMAX = 28  # deepest partitioned carry (qpc_solve.h kPKMax)


def _ops(h, N):
    o = po.Oracle(N=N, p=7, L=2, rmax=h * (N + 2 - 7), dt=0.05)
    return o.get(po.QOR_S), o.get(po.QOR_RDIF)


@pytest.mark.parametrize("h,N", [(0.2, 640), (0.1, 640), (0.05, 1000)])
@pytest.mark.parametrize("emax", [0.005, 0.017, 0.035, 0.07])
def test_plan_covers_the_tail_and_the_restart(h, N, emax):
    S, D = _ops(h, N)
    pk, restart = qpc_gpu.coupling_plan(S, D, N, 7, emax)
    assert restart >= 64 and restart % 16 == 0
    if pk > PKMAX:
        # sequential solve: exact; the restart still has to meet the bar
        depth = None
    else:
        depth = pk * 24
    W = min(N, 900)
    for e in (emax, -emax, emax / 2):
        if depth is not None:
            A = _dense(S + e * D, (N - W) // 2, W, lower=S - e * D)
            assert _tail(A, depth) <= 1e-17, (h, emax, e, pk)
        assert _restart_error(S, D, e, 1, r0=min(400, N - 48), W=min(restart, min(400, N - 48))) <= 1e-14


def test_plan_matches_the_fixed_depth_where_that_was_exact():
    """Config 2's grid (h = 0.2) at its strongest D-part scalar keeps the
    former fixed depth of 8 partitions and the 64-row... restart or more."""
    S, D = _ops(0.2, 640)
    pk, restart = qpc_gpu.coupling_plan(S, D, 640, 7, 1.25 * 0.0171)
    assert pk == 8 and restart >= 64


def test_plan_grows_with_field_over_spacing_and_goes_sequential():
    S, D = _ops(0.05, 1000)
    pks = [qpc_gpu.coupling_plan(S, D, 1000, 7, e)[0] for e in (0.005, 0.02, 0.035, 0.2)]
    assert pks == sorted(pks) and pks[0] < pks[-1]
    assert pks[-1] > PKMAX  # e/h = 4: no partitioned carry reaches far enough


def test_reach_matches_dense_numpy():
    S, D = _ops(0.1, 640)
    for e in (0.01, 0.05):
        f, b = qpc_gpu.coupling_reach(S, D, 640, 7, e, 1)
        A = _dense(S + e * D, 0, 640, lower=S - e * D)
        # the tail just beyond the reported reach is below 1e-17, and well above it just before
        assert _tail(A, b + 1) <= 1e-17 and _tail(A, max(b - 24, 1)) > 1e-17
