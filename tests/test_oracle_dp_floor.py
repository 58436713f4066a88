"""Why the benchmark cases stop the Poisson iteration at max|dp| <= 1e-10 and
not at the 1e-13 the benchmark configurations name.

The stopping rule (poisson.cpp:391-411) compares the absolute max|dp| of an
iteration with epsilon.  Once the iteration has converged, dp is the zero mode
of the residual, which ccp_fft passes through undivided (poisson.cpp:442):
the round-off of the residual's mean.  On the 128^3 configuration with
heated particles that floor is about 2e-13, so epsilon = 1e-13 is never met
and every solve would run all Iteration (50) passes; 1e-10 stops at the
floor's first plateau.  The oracle is run one iteration at a time
(Iteration = 1, P kept between calls) to read the max|dp| sequence."""
import numpy as np
import pytest

from helpers import oracle_from_case
from oracle.oracle import threads
from tpx import ic


@pytest.mark.slow
def test_max_dp_floor_at_128_with_heat_exceeds_1e13():
    case = ic.make_case(128, 1_000_000, nu=0.005, seed=1, St=1.0, phi=0.5, heat=True, eps_range=(0.5, 1.0))
    p = case.params
    p.Iteration1 = p.Iteration2 = 1
    threads(0)
    o = oracle_from_case(case)
    o.params_update()
    o.set_rk4_count(0)
    for m in ("update_rho", "update_p0", "update_particle", "update_ru_wop", "compute_div_u_new",
              "compute_rhs_pois"):
        getattr(o, m)()
    P = o.get("P")
    dps = []
    for _ in range(8):
        o.solve_poisson()
        Pn = o.get("P")
        dps.append(float(np.max(np.abs(Pn - P))))
        P = Pn
    o.close()
    print("max|dp| per iteration:", " ".join(f"{d:.3e}" for d in dps))
    floor = min(dps[3:])
    assert dps[0] > 1.0                    # the first iteration solves for the whole pressure
    assert 1e-13 < floor < 1e-11, dps      # converged: a plateau above 1e-13 ...
    assert max(dps[3:]) < 1e-10            # ... and below the benchmark's epsilon
