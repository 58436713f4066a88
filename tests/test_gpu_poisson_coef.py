"""poisson::Solve with the caller's face coefficient on the device
(tpx_poisson_solve, poisson.h:79) against the oracle's restatement
(or_poisson_solve), on a cubic and a non-cubic box; and the grid's own state
is left untouched."""
import numpy as np
import pytest

from helpers import rel_err, solver_from_case, small_case
from oracle.oracle import Oracle
from test_oracle_poisson_coef import manufactured

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(32, 32, 32), (16, 8, 64), (8, 512, 16)])
def test_coefficient_solve_matches_the_oracle(gpu_available, shape):
    from tpx import Solver
    p, Pstar, C, RHS = manufactured(shape, seed=sum(shape), spread=0.5)
    p.epsilon = 1e-10
    p.Iteration1 = p.Iteration2 = 50
    rng = np.random.default_rng(7)
    P0 = 0.1 * rng.standard_normal(shape)
    o = Oracle(p.struct_fields())
    Po, ito = o.poisson_solve(C, P0, RHS)
    o.close()
    s = Solver(p, np_capacity=1)
    Ps, its = s.poisson_solve(C, P0, RHS)
    s.close()
    e = rel_err(Ps, Po)
    print(shape, "iterations", its, ito, f"rel err {e:.2e}")
    assert its == ito
    assert e <= 1e-10


def test_coefficient_solve_leaves_the_grid_alone(gpu_available):
    case = small_case(n=16, npart=800, heat=True, variable_rho=True, seed=19)
    s = solver_from_case(case)
    s.step(1)
    before = {f: s.get(f) for f in ("RU", "Rho", "P")}
    it0 = s.scalars()["total_iterations"]
    rng = np.random.default_rng(1)
    C = 1.0 + 0.3 * rng.random((3, 16, 16, 16))
    RHS = rng.standard_normal((16, 16, 16))
    RHS -= RHS.mean()
    s.poisson_solve(C, np.zeros((16, 16, 16)), RHS)
    for f, v in before.items():
        assert np.array_equal(s.get(f), v), f
    assert s.scalars()["total_iterations"] == it0
    a = solver_from_case(case)
    a.step(2)
    s.step(1)
    for f in ("RU", "Rho", "P"):
        assert np.array_equal(s.get(f), a.get(f)), f
