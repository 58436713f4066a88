"""grid::Statistics on the device against the oracle's restatement of
grid.cpp:902-982 over the same run: two steps of a two-way coupled heated
case with variable density, and of a one-way case (whose concentration bins
must be refreshed inside the step).  Sums are compared to 1e-12 relative
(their order differs), extrema to 1e-14."""
import numpy as np
import pytest

import tpx
from helpers import oracle_from_case, small_case, solver_from_case

pytestmark = pytest.mark.gpu

SUMS = ("Tg_mean", "Tp_mean", "HT_mean", "TKE2", "TKE_U", "TKE_V", "TKE_W", "C_mean", "Rho_mean", "P0")


@pytest.mark.parametrize("heat,twc", [(True, True), (False, False)])
def test_statistics_match_the_oracle(gpu_available, heat, twc):
    case = small_case(n=16, npart=3000, heat=heat, twc=twc, variable_rho=heat, seed=8)
    o, s = oracle_from_case(case), solver_from_case(case)
    s.statistics_enable(True)
    o.step(2)
    s.step(2)
    so, ss = o.statistics(case.params), s.statistics()
    assert set(ss) == set(tpx.STATS) == set(so)
    for k in tpx.STATS:
        a, b = ss[k], so[k]
        tol = 1e-12 if k in SUMS else 1e-14
        scale = max(abs(b), 1e-300)
        print(f"{k}: device {a:.16e} oracle {b:.16e}")
        assert abs(a - b) <= tol * scale or (a == b), k


def test_statistics_need_the_capture(gpu_available):
    case = small_case(n=16, npart=500, seed=9)
    s = solver_from_case(case)
    s.step(1)
    with pytest.raises(tpx.TpxError, match="statistics are off"):
        s.statistics()
