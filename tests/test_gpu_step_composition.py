"""The whole-step closed forms and the threshold switch (params::update,
params.cpp:75-86) on the device, through the C ABI.

The closed-form cases are the ones that pin the oracle's composition
(tests/test_oracle_step_composition.py): shear-mode decay, the heat pair
(particle temperature and P0 with the lagged heat source), the drag pair
(two-way coupling with make_mean_U0 once per step) and the body force on a
z-varying density; the device must meet each closed form itself, not only
match the oracle.  The threshold tests cross params::update's switch with
every switched coefficient different (A, g, I0, Iteration)."""
import numpy as np
import pytest

from helpers import oracle_from_case, particle_errors, rel_err, small_case, solver_from_case
from test_oracle_step_composition import PIN_CHECKS, start
from tpx import ic

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("pin", sorted(PIN_CHECKS))
def test_closed_form_on_device(gpu_available, pin):
    err = PIN_CHECKS[pin]("gpu")
    print(pin, f"{err:.2e}")
    assert err < 1e-12, (pin, err)


def test_threshold_crossing_matches_the_oracle(gpu_available):
    case = small_case(n=16, npart=2500, heat=True, variable_rho=True, gravity=True, eps_range=(0.5, 1.0),
                      seed=31)
    p = case.params
    p.A2 = -0.5 * p.A1
    p.gx2, p.gy2, p.gz2 = -0.2, 0.4, 0.1
    p.I02 = 2.5 * p.I01
    p.Iteration1, p.Iteration2 = 50, 4
    p.threshold = 2.5 * p.dt
    o, s = oracle_from_case(case), solver_from_case(case)
    its_o, its_s = [], []
    for _ in range(5):
        o.step(1)
        s.step(1)
        its_o.append(int(o.scalars()["num_iteration"]))
        its_s.append(int(s.scalars()["num_iteration"]))
    assert its_s == its_o, (its_s, its_o)
    assert max(its_o[3:]) <= 4 and max(its_o[:3]) > 4, its_o
    so, ss = o.scalars(), s.scalars()
    for k in ("A", "I0", "Iteration", "gx", "gy", "gz"):
        assert ss[k] == so[k], (k, ss[k], so[k])
    assert so["A"] == p.A2 and so["I0"] == p.I02 and so["gy"] == p.gy2
    e = {f"RU{c}": rel_err(s.get("RU")[c], o.get("RU")[c]) for c in range(3)}
    e["Rho"] = rel_err(s.get("Rho"), o.get("Rho"))
    e["P"] = rel_err(s.get("P"), o.get("P"))
    e["P0"] = abs(ss["P0"] - so["P0"]) / abs(so["P0"])
    e.update({f"p_{k}": v for k, v in particle_errors(s.particles(), o.particles(),
                                                       (p.Lx, p.Ly, p.Lz)).items()})
    print({k: f"{v:.2e}" for k, v in e.items()}, "iterations", its_o)
    assert max(e.values()) <= 1e-9, e


def test_threshold_switches_the_iteration_cap_on_device(gpu_available):
    case = ic.make_case(16, 0, nu=0.02, seed=4, St=1.0, phi=0.3, heat=False, twc=False)
    p = case.params
    rng = np.random.default_rng(4)
    case.Rho = 1.0 + 0.1 * rng.random((16, 16, 16))
    p.epsilon = -1.0
    p.Iteration1, p.Iteration2 = 3, 5
    p.threshold = 1.5 * p.dt
    s = start(p, case.RU, case.Rho, lib_path="gpu")
    seen = []
    for _ in range(4):
        s.step(1)
        seen.append(int(s.scalars()["num_iteration"]))
    s.close()
    assert seen == [3, 3, 5, 5], seen
