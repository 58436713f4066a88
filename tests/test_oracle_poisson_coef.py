"""Pins of poisson::Solve with the caller's face coefficient
(poisson.cpp:115-117, 391-462; or_poisson_solve), CPU only.

* A manufactured problem: P* and a positive, varying face coefficient C give
  RHS = div_F2C(C grad_C2F P*) (formed here in numpy); the iteration converges
  to P* up to the constant the periodic problem leaves free.
* A uniform coefficient c0: cc = -1/c0 makes one iteration the exact inverse
  of the 7-point operator, so P* is reached in one iteration.
* With C = 1/I_C2F(Rho_new) and the grid's RHS_Pois the solve equals the
  grid's own Solve_Poisson (grid.cpp:851-855) bit for bit.
"""
import math

import numpy as np

from oracle.oracle import Oracle
from tpx.params import Params


def params(shape, **kw):
    nz, ny, nx = shape
    p = Params(Nx_tot=nx, Ny_tot=ny, Nz_tot=nz, Lx=2 * math.pi, Ly=2 * math.pi * ny / nx, Lz=2 * math.pi * nz / nx,
               bs=2, dt=1e-2, Iteration1=200, Iteration2=200, epsilon=1e-14, TWC=0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def grad_c2f(P, d):
    return np.stack([(P - np.roll(P, 1, axis=2 - c)) / d[c] for c in range(3)])


def div_f2c(F, d):
    return sum((np.roll(F[c], -1, axis=2 - c) - F[c]) / d[c] for c in range(3))


def manufactured(shape, seed, spread):
    rng = np.random.default_rng(seed)
    p = params(shape)
    d = (p.Lx / p.Nx_tot, p.Ly / p.Ny_tot, p.Lz / p.Nz_tot)
    Pstar = rng.standard_normal(shape)
    C = 1.0 + spread * rng.random((3,) + shape)
    RHS = div_f2c(C * grad_c2f(Pstar, d), d)
    return p, Pstar, C, RHS


def test_manufactured_variable_coefficient_converges():
    shape = (16, 8, 32)
    p, Pstar, C, RHS = manufactured(shape, seed=3, spread=0.5)
    o = Oracle(p.struct_fields())
    P, its = o.poisson_solve(C, np.zeros(shape), RHS)
    o.close()
    err = np.abs((P - P.mean()) - (Pstar - Pstar.mean())).max() / np.abs(Pstar).max()
    assert its > 3 and err < 1e-9, (its, err)


def test_uniform_coefficient_is_exact_in_one_iteration():
    shape = (8, 16, 16)
    p, Pstar, _, _ = manufactured(shape, seed=4, spread=0.0)
    p.Iteration1 = p.Iteration2 = 1
    d = (p.Lx / p.Nx_tot, p.Ly / p.Ny_tot, p.Lz / p.Nz_tot)
    c0 = 0.37
    C = np.full((3,) + shape, c0)
    RHS = div_f2c(C * grad_c2f(Pstar, d), d)
    o = Oracle(p.struct_fields())
    P, its = o.poisson_solve(C, np.zeros(shape), RHS)
    o.close()
    assert its == 1
    err = np.abs((P - P.mean()) - (Pstar - Pstar.mean())).max() / np.abs(Pstar).max()
    assert err < 1e-12, err


def test_equals_the_grids_own_solve():
    from helpers import oracle_from_case, small_case
    case = small_case(n=16, npart=500, heat=True, variable_rho=True, seed=8)
    o = oracle_from_case(case)
    o.set_rk4_count(0)
    for m in ("update_rho", "update_p0", "update_particle", "update_ru_wop", "compute_div_u_new",
              "compute_rhs_pois"):
        getattr(o, m)()
    P0 = o.get("P")
    rhs = o.get("RHS_Pois")
    rf = o.get("Rho_face")
    C = 1.0 / rf
    P, its = o.poisson_solve(C, P0, rhs)
    o.solve_poisson()
    assert its == o.scalars()["num_iteration"]
    assert np.array_equal(P, o.get("P"))
