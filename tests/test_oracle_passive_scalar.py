"""Pins of the oracle's passive scalar (grid.cpp:686-745, Update_Passive_Scalar
and C_Source, with the stage copies of Initialize, TimeAdvance_RK4 and
TimeAdvance) against what the mathematics fixes:

* diffusion alone: a single Fourier mode is multiplied every step by the RK4
  polynomial R(z) = 1 + z + z^2/2 + z^3/6 + z^4/24 of z = dt D_M lambda, lambda
  the 7-point Laplacian's eigenvalue;
* transport by a uniform velocity Uc: the face interpolation and flux
  divergence give the central-difference eigenvalue z = -i dt Uc sin(k dx)/dx,
  and the second source form (-Ix_F2C(U.x) = -Uc) adds -Uc t exactly;
* the first source form: cos(X) minus its mean, at the cell centres, added at
  a constant rate;
* the transport, diffusion and zero-mean source keep sum(C) to round-off in a
  random flow of variable density.
A wrong stage (for instance the face value not taken from the previous
substep's new value) or a wrong sign breaks the first two."""
import math

import numpy as np

from oracle.oracle import Oracle
from tpx.params import Params

N = 16
L = 2 * math.pi


def params(**kw):
    p = Params(Nx_tot=N, Ny_tot=N, Nz_tot=N, Lx=L, Ly=L, Lz=L, bs=2, dt=2e-2, Mu0=0.0, A1=0.0, A2=0.0,
               TWC=0, Nu=0.0, I01=0.0, I02=0.0, Iteration1=50, Iteration2=50, epsilon=1e-12,
               solve_for_scalar=1, D_M=0.0, s1t=1.0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def centres():
    d = L / N
    k, j, i = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    return (i + 0.5) * d, (j + 0.5) * d, (k + 0.5) * d


def run(p, C0, RU, Rho, nsteps):
    o = Oracle(p.struct_fields(), np_capacity=1)
    o.set("RU", RU)
    o.set("Rho", Rho)
    o.set("P", np.zeros_like(Rho))
    o.set("Passive_Scalar", C0)
    o.set_restart_numbers(p.P0)
    o.initialize()
    o.step(nsteps)
    C = o.get("Passive_Scalar").copy()
    o.close()
    return C


def rk4(z):
    return 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24


def test_diffusion_of_a_mode_is_rk4_of_its_eigenvalue():
    p = params(D_M=0.15)
    X, Y, Z = centres()
    kx, ky, kz = 2, 3, 1
    C0 = np.sin(kx * X) * np.cos(ky * Y) * np.cos(kz * Z)
    d = L / N
    lam = sum((2 * math.cos(k * d) - 2) / (d * d) for k in (kx, ky, kz))
    n = 5
    C = run(p, C0, np.zeros((3, N, N, N)), np.ones((N, N, N)), n)
    expect = rk4(p.dt * p.D_M * lam) ** n * C0
    assert np.max(np.abs(C - expect)) <= 1e-13 * np.max(np.abs(C0))


def test_uniform_transport_is_rk4_of_the_central_eigenvalue():
    Uc = 0.7
    p = params(U0=Uc, Rho0=1.0)  # make_mean_U0 keeps the uniform flow
    X, _, _ = centres()
    k = 3
    d = L / N
    C0 = np.cos(k * X)
    RU = np.zeros((3, N, N, N))
    RU[0] = Uc
    n = 6
    C = run(p, C0, RU, np.ones((N, N, N)), n)
    z = -1j * p.dt * Uc * math.sin(k * d) / d
    expect = np.real(rk4(z) ** n * np.exp(1j * k * X)) - Uc * n * p.dt  # S1 form 1: -Ix_F2C(U.x) = -Uc
    assert np.max(np.abs(C - expect)) <= 1e-13


def test_cosine_source_grows_linearly():
    p = params(s1t=0.0)
    X, _, _ = centres()
    n = 4
    C = run(p, np.zeros((N, N, N)), np.zeros((3, N, N, N)), np.ones((N, N, N)), n)
    S = np.cos(X)
    expect = n * p.dt * (S - S.mean())
    assert np.max(np.abs(C - expect)) <= 1e-14


def test_sum_is_conserved_in_a_random_flow():
    rng = np.random.default_rng(5)
    p = params(s1t=0.0, D_M=0.05, Mu0=0.02)
    RU = 0.5 * rng.standard_normal((3, N, N, N))
    Rho = 1.0 + 0.2 * rng.random((N, N, N))
    C0 = rng.random((N, N, N))
    C = run(p, C0, RU, Rho, 3)
    assert abs(C.sum() - C0.sum()) <= 1e-12 * np.abs(C0).sum()
    assert np.max(np.abs(C - C0)) > 1e-3  # it did move
