"""The passive vector on the device (grid.cpp:616-715), through the C ABI:

* the oracle's closed forms (tests/test_oracle_passive_vector.py) met by the
  device itself: a diffusing mode, the type-0 source from rest, transport by
  a uniform gas velocity, the type-1 source on that flow, the projection;
* method by method against the oracle at 16^3 with heated two-way coupled
  particles and a variable density, both source forms (<= 1e-11 after each
  method), then five fused steps (<= 1e-9), with the gas bit for bit what it
  is with the vector off (the vector is passive);
* its statistics TKEV_V1..3 and VV against the oracle's restatement.
"""
import numpy as np
import pytest

from helpers import oracle_from_case, rel_err, small_case, solver_from_case
from test_oracle_passive_vector import CHECKS

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", sorted(CHECKS))
def test_closed_form_on_device(gpu_available, name):
    fn, tol = CHECKS[name]
    err = fn("gpu")
    print(name, f"{err:.2e}")
    assert err < 10 * tol, (name, err)


def vector_case(s2t, seed):
    case = small_case(n=16, npart=2000, heat=True, variable_rho=True, eps_range=(0.5, 1.0), seed=seed)
    p = case.params
    p.solve_for_vector, p.eta, p.Rho_forV, p.s2t = 1, 0.04, 1.3, s2t
    rng = np.random.default_rng(seed + 1)
    n = p.Nx_tot
    x = (np.arange(n) + 0.5) * (p.Lx / n)
    case.RV = 0.2 * rng.standard_normal((3, n, n, n)) + np.sin(x)[None, None, None, :]
    case.Q = np.zeros((n, n, n))
    return case


@pytest.mark.parametrize("s2t", [0.0, 1.0])
def test_methods_and_steps_match_the_oracle(gpu_available, s2t):
    case = vector_case(s2t, seed=41 + int(s2t))
    o, s = oracle_from_case(case), solver_from_case(case)
    errs = {}
    o.params_update(); s.params_update()
    for rk in range(4):
        o.set_rk4_count(rk)
        o.update_rho(); s.update_rho(rk)
        o.update_p0(); s.update_p0(rk)
        o.update_particle(); s.update_particle(rk)
        o.update_rv_woq(); s.update_rv_woq(rk)
        errs[f"{rk}.RV_np1"] = max(rel_err(a, b) for a, b in zip(s.get("RV_np1"), o.get("RV_np1")))
        errs[f"{rk}.RV_new"] = max(rel_err(a, b) for a, b in zip(s.get("RV_new"), o.get("RV_new")))
        o.compute_rhs_pois_q(); s.compute_rhs_pois_q(rk)
        errs[f"{rk}.RHS_Pois_Q"] = rel_err(s.get("RHS_Pois_Q"), o.get("RHS_Pois_Q"))
        o.solve_poisson_q(); it = s.solve_poisson_q(rk)
        assert it == o.scalars()["num_iteration"], (rk, it)
        errs[f"{rk}.Q"] = rel_err(s.get("Q"), o.get("Q"))
        o.update_rv_wq(); s.update_rv_wq(rk)
        errs[f"{rk}.RV_np1_wq"] = max(rel_err(a, b) for a, b in zip(s.get("RV_np1"), o.get("RV_np1")))
        for m in ("update_ru_wop", "compute_div_u_new", "compute_rhs_pois", "solve_poisson", "update_ru_wp"):
            getattr(o, m)(); getattr(s, m)(rk)
        o.time_advance_rk4(); s.time_advance_rk4()
    a = [case.params.U0 * case.params.Rho0, case.params.V0 * case.params.Rho0, case.params.W0 * case.params.Rho0]
    o.make_mean_u0(a); s.make_mean_u0(a)
    o.time_advance(); s.time_advance()
    print("methods:", max(errs.values()), max(errs, key=errs.get))
    assert max(errs.values()) <= 1e-11, errs
    # fused steps
    o.step(4); s.step(4)
    e = {"RV": rel_err(s.get("RV"), o.get("RV")), "Q": rel_err(s.get("Q"), o.get("Q")),
         "RU": rel_err(s.get("RU"), o.get("RU")), "P": rel_err(s.get("P"), o.get("P"))}
    print("5 steps:", {k: f"{v:.2e}" for k, v in e.items()})
    assert max(e.values()) <= 1e-9, e
    # the gas is what it is without the vector
    off = vector_case(s2t, seed=41 + int(s2t))
    off.params.solve_for_vector = 0
    g = solver_from_case(off)
    g.step(5)
    for f in ("RU", "Rho", "P"):
        assert np.array_equal(g.get(f), s.get(f)), f


def test_vector_statistics(gpu_available):
    case = vector_case(0.0, seed=47)
    o, s = oracle_from_case(case), solver_from_case(case)
    s.statistics_enable(True)
    o.step(2); s.step(2)
    so, ss = o.statistics(case.params), s.statistics()
    for k in ("TKEV_V1", "TKEV_V2", "TKEV_V3", "VV"):
        assert so[k] != 0.0
        assert abs(ss[k] - so[k]) <= 1e-11 * abs(so[k]), (k, ss[k], so[k])
