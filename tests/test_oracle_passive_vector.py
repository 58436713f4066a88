"""Pins of the oracle's passive vector (grid.cpp:616-715: Update_RV_WOQ,
Compute_RHS_Pois_Q, Solve_Poisson_Q, Update_RV_WQ, V_Source), CPU only,
through whole steps of or_step with solve_for_vector on:

* a diffusing single mode RV_x = e sin(k y) with the gas at rest and the
  type-1 source (which vanishes with RU = 0): every step multiplies RV by
  R(-dt eta lam_y / Rho_forV), R the RK4 stability polynomial;
* the type-0 source (vector_source.h: cos(X) on the y component, minus its
  mean) acting on RV_y(x) from rest: the affine RK4 map of
  y' = -(eta lam_x / Rho_forV) y + cos(X);
* a mode RV_y = e sin(k x) carried by a uniform gas velocity u0 along x: the
  transport term div(I_C2F(RV_int.y) I_y(U)) is the central difference, so
  each step multiplies the complex amplitude by R(dt(-i u0 sin(k dx)/dx -
  eta lam_x / Rho_forV)) -- transport of RV by U, not of U by RV;
* the type-1 source on that uniform flow: -Iy_C2F(RU_int.x / Rho_face.x) is
  -u0 on the y component, so RV_y falls by u0 dt per step exactly;
* a start with a gradient part: after a step the discrete divergence of RV
  is zero (the Q projection with the RK4 post/pre pairing) and Q holds the
  pressure that removed it.
The in-place Iy_C2F of the type-1 source (its recurrence along y) is pinned
on constant fields only.
"""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle
from tpx.params import Params

L = 2 * math.pi
N = 16


def params(**kw):
    p = Params(Nx_tot=N, Ny_tot=N, Nz_tot=N, Lx=L, Ly=L, Lz=L, bs=2, dt=0.05, Mu0=0.02, A1=0.0, A2=0.0,
               TWC=0, Nu=0.0, I01=0.0, I02=0.0, k=0.0, Iteration1=50, Iteration2=50, epsilon=1e-12,
               solve_for_vector=1, eta=0.07, Rho_forV=1.3, s2t=1.0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def run(p, RV, RU=None, nsteps=3, impl=None):
    """Whole steps on the oracle, or on the device library when impl is
    "gpu" (the device tests reuse these closed forms)."""
    if impl == "gpu":
        from tpx import Solver
        o = Solver(p, np_capacity=1)
    else:
        o = Oracle(p.struct_fields())
    o.set("RU", np.zeros((3, N, N, N)) if RU is None else RU)
    o.set("Rho", np.full((N, N, N), p.Rho0))
    o.set("P", np.zeros((N, N, N)))
    o.set("RV", RV)
    o.set("Q", np.zeros((N, N, N)))
    o.set_restart_numbers(p.P0, 0.0, 0)
    o.initialize()
    o.step(nsteps)
    out = o.get("RV"), o.get("Q")
    o.close()
    return out


def R(z):
    return 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24


d = L / N
centres = (np.arange(N) + 0.5) * d      # x of a y-face, y of an x-face


def check_diffusing_mode(impl=None):
    p = params()
    k, eps = 2.0, 0.4
    RV = np.zeros((3, N, N, N))
    RV[0] = eps * np.sin(k * centres)[None, :, None]
    got, Q = run(p, RV, impl=impl)
    lam = (2 / d ** 2) * (1 - math.cos(k * d))
    amp = R(-p.dt * p.eta * lam / p.Rho_forV) ** 3
    return max(np.max(np.abs(got[0] - amp * RV[0])) / eps, np.max(np.abs(got[1:])), np.max(np.abs(Q)) * 1e-2)


def check_type0_source_from_rest(impl=None):
    p = params(s2t=0.0)
    got, _ = run(p, np.zeros((3, N, N, N)), nsteps=4, impl=impl)
    X = centres
    src = np.cos(X) - np.cos(X).mean()
    lam = (2 / d ** 2) * (1 - math.cos(d))       # cos(X) is the k = 1 mode along x
    a = -p.eta * lam / p.Rho_forV
    # per x value: y' = a y + src, from y = 0
    h = p.dt
    P, Q = R(h * a), 1 + h * a / 2 + (h * a) ** 2 / 6 + (h * a) ** 3 / 24
    y = np.zeros(N)
    for _ in range(4):
        y = P * y + h * Q * src
    return max(np.max(np.abs(got[1] - y[None, None, :])) / np.max(np.abs(y)), np.max(np.abs(got[0])),
               np.max(np.abs(got[2])))


def uniform_flow(p, u0):
    RU = np.zeros((3, N, N, N))
    RU[0] = p.Rho0 * u0
    return RU


def check_transport_by_the_gas(impl=None):
    u0 = 0.8
    p = params(U0=u0)
    k, eps = 3.0, 0.25
    RV = np.zeros((3, N, N, N))
    RV[1] = eps * np.sin(k * centres)[None, None, :]
    got, _ = run(p, RV, RU=uniform_flow(p, u0), nsteps=3, impl=impl)
    lam = (2 / d ** 2) * (1 - math.cos(k * d))
    z = p.dt * (-1j * u0 * math.sin(k * d) / d - p.eta * lam / p.Rho_forV)
    # the type-1 source adds -u0 per unit time on the y component
    want = np.imag(eps * R(z) ** 3 * np.exp(1j * k * centres)) - 3 * p.dt * u0
    return max(np.max(np.abs(got[1] - want[None, None, :])), np.max(np.abs(got[0])), np.max(np.abs(got[2])))


def check_type1_source_on_a_uniform_flow(impl=None):
    u0 = -0.6
    p = params(U0=u0)
    got, _ = run(p, np.zeros((3, N, N, N)), RU=uniform_flow(p, u0), nsteps=5, impl=impl)
    return max(np.max(np.abs(got[1] - (-5 * p.dt * u0))), np.max(np.abs(got[0])), np.max(np.abs(got[2])))


def check_projection(impl=None):
    """max|div RV| after one step relative to before, and whether Q moved"""
    p = params(s2t=0.0)
    rng = np.random.default_rng(3)
    RV = 0.3 * rng.standard_normal((3, N, N, N))
    got, Q = run(p, RV, nsteps=1, impl=impl)

    def div(F):
        return sum((np.roll(F[c], -1, axis=2 - c) - F[c]) / d for c in range(3))
    assert np.max(np.abs(div(RV))) > 1.0 and np.max(np.abs(Q)) > 1e-3
    return np.max(np.abs(div(got))) / np.max(np.abs(div(RV)))


CHECKS = {"diffusing_mode": (check_diffusing_mode, 1e-13), "type0_source": (check_type0_source_from_rest, 1e-13),
          "transport": (check_transport_by_the_gas, 1e-13), "type1_source": (check_type1_source_on_a_uniform_flow, 1e-14),
          "projection": (check_projection, 1e-11)}


@pytest.mark.parametrize("name", sorted(CHECKS))
def test_closed_form(name):
    fn, tol = CHECKS[name]
    err = fn()
    assert err < tol, (name, err)
