"""Pins of the oracle's substep methods (rows G1, G2, G4, G12, P1-P9, B1, B2)
against what the mathematics fixes, not against a retyped formula:

* G1  mass: the density update conserves sum(Rho) to round-off.
* G2  P0: with a held heat source the RK4 update of P0 is exact (linear ODE).
* G4  momentum/energy: the convective term telescopes (sum = 0) and does no
      work on a discretely divergence-free field at uniform density; the
      viscous, forcing, body-force and drag terms are separated by linearity
      and checked against single-mode closed forms.
* G5  divergence: (k/Cp) lap(1/Rho_new) - R/(Cp P0_new Vcell) Q on a mode.
* G6  Poisson RHS: single-mode face divergence, the divergence subtracted,
      the 1/(dt pre) scaling and the mean removed.
* G10 projection: single-mode gradient with the RK4 post/pre coefficients.
* G12 the shifted mean equals rho0 U0 exactly.
* G14 Initialize's heat source from a known gas/particle temperature gap.
* P1  velocity gather: 1-D linear along its own axis, piecewise constant
      across (grid.cpp / particle.cpp:340-360), checked on separable fields.
* P6  the velocity deposit is the exact transpose of the P1 gather.
* P5  concentration is a partition of unity (sum C = Np).
* P8/P9 the heat deposit is the exact transpose of the temperature gather,
      and the gather reproduces constants.
* P2-P4 frozen uniform gas: the four substeps must equal classical RK4 of the
      linear system x' = u, u' = (u_g - u)/tau + g, whose one-step map is the
      matrix polynomial sum_k (hM)^k / k!; the temperature rises linearly.
* B2  per-particle tau / eps equal to the scalars reproduce the scalar run
      bit for bit.
* B1  the spectral initial condition is discretely divergence-free with u' = 1.
"""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle
from tpx import ic
from tpx.params import PI, Params

N = 16
L = 2 * math.pi


def grid_coords(n=N, Lx=L):
    d = Lx / n
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    return (i + 0.5) * d, (j + 0.5) * d, (k + 0.5) * d, i * d, j * d, k * d


def new_oracle(p: Params, npart=1):
    return Oracle(p.struct_fields(), np_capacity=max(npart, 1))


def base_params(**kw):
    p = Params(Nx_tot=N, Ny_tot=N, Nz_tot=N, Lx=L, Ly=L, Lz=L, bs=2, dt=1e-2, Mu0=0.0, A1=0.0, A2=0.0,
               TWC=0, Nu=0.0, I01=0.0, I02=0.0, Iteration1=50, Iteration2=50, epsilon=1e-12)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def load_grid(o, RU, Rho, P=None):
    o.set("RU", RU)
    o.set("Rho", Rho)
    o.set("P", np.zeros_like(Rho) if P is None else P)


def rhs_of_wop(p, RU, Rho, particles=None, with_particles=False):
    """RHS of Update_RU_WOP at the first substep, as (RU_new - RU)/(dt pre)."""
    o = new_oracle(p, 0 if particles is None else len(particles["x"]))
    load_grid(o, RU, Rho)
    if particles is not None:
        pp = particles
        o.set_particles(pp["x"], pp["y"], pp["z"], pp["u"], pp["v"], pp["w"], pp["T"], pp["flag"])
    o.set_restart_numbers(p.P0)
    o.initialize()
    o.params_update()
    o.set_rk4_count(0)
    o.update_rho()
    o.update_p0()
    o.update_particle()
    S = o.get("RHS_RU").copy()
    o.update_ru_wop()
    rhs = (o.get("RU_new") - RU) / (p.dt * 0.5)
    o.close()
    return rhs, S


def div_free_field(seed=1):
    return ic.spectral_velocity((N, N, N), (L, L, L), seed=seed, k0=2.0, uprime=1.0)


# ---------------------------------------------------------------- G1 / G2 / G12
def test_g1_density_update_conserves_mass():
    p = base_params()
    o = new_oracle(p, 0)
    rng = np.random.default_rng(0)
    RU = rng.standard_normal((3, N, N, N))
    Rho = 1 + 0.1 * rng.random((N, N, N))
    load_grid(o, RU, Rho)
    o.set_restart_numbers(p.P0)
    o.initialize()
    o.set_rk4_count(1)
    o.update_rho()
    m0 = math.fsum(Rho.ravel())
    for name in ("Rho_new", "Rho_np1"):
        assert abs(math.fsum(o.get(name).ravel()) - m0) < 1e-12 * m0, name
    # and the change is -dt*pre*div(RU) with the closed-form divergence
    assert not np.allclose(o.get("Rho_new"), Rho)


@pytest.mark.parametrize("cooling", [0, 1])
def test_g2_p0_rk4_is_exact_for_a_held_source(cooling):
    p = base_params(cooling=cooling, dt=0.37)
    o = new_oracle(p, 0)
    load_grid(o, np.zeros((3, N, N, N)), np.ones((N, N, N)))
    o.set_restart_numbers(p.P0)
    o.initialize()
    q = 0.8
    vcell = (L / N) ** 3
    for rk in range(4):
        o.set_rk4_count(rk)
        o.set("RHS_Part_Temp", np.full((N, N, N), q))
        o.update_p0()
        o.time_advance_rk4()
    o.make_mean_u0([0.0, 0.0, 0.0])
    o.time_advance()
    rate = 0.0 if cooling else -(p.R / p.Cv) * q / vcell
    # the mean of the held source is a 4096-term sum: round-off of order 1e-13
    assert abs(o.scalars()["P0"] - (p.P0 + p.dt * rate)) < 1e-12 * (abs(p.P0) + abs(p.dt * rate))


def test_g12_make_mean_u0_sets_the_mean_exactly():
    p = base_params()
    o = new_oracle(p, 0)
    rng = np.random.default_rng(2)
    RU = rng.standard_normal((3, N, N, N)) + 0.7
    load_grid(o, RU, np.ones((N, N, N)))
    o.set_restart_numbers(p.P0)
    o.initialize()
    target = [0.25, -1.5, 0.0]
    o.make_mean_u0(target)
    got = o.get("RU_np1")
    for c in range(3):
        assert abs(got[c].mean() - target[c]) < 1e-14
        # the shift is uniform: differences between cells are untouched
        assert np.allclose(got[c] - got[c].ravel()[0], RU[c] - RU[c].ravel()[0], atol=1e-14)


# ---------------------------------------------------------------- G4
def test_g4_convection_conserves_momentum_and_energy():
    """Inviscid, unforced, uniform density: RHS = -convection.  Its sum
    vanishes (flux form) and it does no work on a discretely divergence-free
    velocity (skew symmetry of the staggered scheme)."""
    p = base_params()
    U = div_free_field()
    assert np.max(np.abs(ic.discrete_divergence(U, (L, L, L)))) < 1e-12
    rhs, _ = rhs_of_wop(p, U, np.ones((N, N, N)))
    scale = float(np.max(np.abs(rhs)))
    assert scale > 1e-2
    for c in range(3):
        assert abs(rhs[c].sum()) < 1e-12 * scale * N ** 3
    work = float(np.sum(U * rhs))
    assert abs(work) < 1e-12 * scale * N ** 3


def test_g4_viscous_term_single_mode():
    """The viscous operator mu lap U + mu/3 grad div U is linear: the
    antisymmetric part (RHS(eps U) - RHS(-eps U)) / (2 eps) removes the
    quadratic convection exactly.  For U_x = sin(k x_face) it must equal
    -(4/3) mu lambda_x U_x, and the other components stay zero."""
    mu = 0.3
    p = base_params(Mu0=mu)
    _, _, _, xf, _, _ = grid_coords()
    kx = 3.0
    U = np.zeros((3, N, N, N))
    U[0] = np.sin(kx * xf)
    eps = 1e-3
    rp, _ = rhs_of_wop(p, eps * U, np.ones((N, N, N)))
    rm, _ = rhs_of_wop(p, -eps * U, np.ones((N, N, N)))
    lin = (rp - rm) / (2 * eps)
    dx = L / N
    lam = (2 / dx ** 2) * (1 - math.cos(kx * dx))
    assert np.max(np.abs(lin[0] + (4.0 / 3.0) * mu * lam * U[0])) < 1e-9
    assert np.max(np.abs(lin[1:])) < 1e-9
    # shear mode U_x = sin(k y_centre): div-free, only mu lap acts
    _, yc, _, _, _, _ = grid_coords()
    U2 = np.zeros((3, N, N, N))
    U2[0] = np.sin(2.0 * yc)
    rp, _ = rhs_of_wop(p, eps * U2, np.ones((N, N, N)))
    rm, _ = rhs_of_wop(p, -eps * U2, np.ones((N, N, N)))
    lin = (rp - rm) / (2 * eps)
    lam_y = (2 / dx ** 2) * (1 - math.cos(2.0 * dx))
    assert np.max(np.abs(lin[0] + mu * lam_y * U2[0])) < 1e-9


def test_g4_forcing_and_body_force_are_additive():
    rng = np.random.default_rng(4)
    U = div_free_field(seed=3)
    Rho = 1 + 0.2 * rng.random((N, N, N))
    RU = U.copy()
    r0, _ = rhs_of_wop(base_params(Mu0=0.05), RU, Rho)
    A = 0.37
    rA, _ = rhs_of_wop(base_params(Mu0=0.05, A1=A, A2=A), RU, Rho)
    assert np.max(np.abs((rA - r0) - A * RU)) < 1e-11
    g = (0.0, -9.81, 2.0)
    rg, _ = rhs_of_wop(base_params(Mu0=0.05, gx1=g[0], gy1=g[1], gz1=g[2]), RU, Rho)
    # body force (I_c^{C2F} rho - rho0) g_c on the c faces
    rho_face = [0.5 * (Rho + np.roll(Rho, 1, axis=2)), 0.5 * (Rho + np.roll(Rho, 1, axis=1)),
                0.5 * (Rho + np.roll(Rho, 1, axis=0))]
    for c in range(3):
        assert np.max(np.abs((rg[c] - r0[c]) - (rho_face[c] - 1.0) * g[c])) < 1e-11, c


def test_g4_drag_is_minus_mp_over_vcell_times_the_deposit():
    p = base_params(Mu0=0.05, TWC=1, Dp=0.05, Rhop=50.0)
    npart = 300
    rng = np.random.default_rng(5)
    parts = dict(x=rng.random(npart) * L, y=rng.random(npart) * L, z=rng.random(npart) * L,
                 u=rng.standard_normal(npart), v=rng.standard_normal(npart), w=rng.standard_normal(npart),
                 T=np.ones(npart), flag=np.arange(npart, dtype=np.int32))
    U = div_free_field(seed=6)
    r1, S = rhs_of_wop(p, U, np.ones((N, N, N)), parts)
    p0 = base_params(Mu0=0.05, TWC=0, Dp=0.05, Rhop=50.0)
    r0, S0 = rhs_of_wop(p0, U, np.ones((N, N, N)), parts)
    assert np.max(np.abs(S0)) == 0.0 and np.max(np.abs(S)) > 0
    vcell = (L / N) ** 3
    assert np.max(np.abs((r1 - r0) + (p.mp / vcell) * S)) < 1e-11 * max(1.0, np.max(np.abs(r1)))


# ---------------------------------------------------------------- particles
def particle_oracle(p, parts, U=None, Rho=None, tau=None, eps=None):
    n = len(parts["x"])
    o = new_oracle(p, n)
    load_grid(o, np.zeros((3, N, N, N)) if U is None else U, np.ones((N, N, N)) if Rho is None else Rho)
    o.set_particles(parts["x"], parts["y"], parts["z"], parts["u"], parts["v"], parts["w"], parts["T"],
                    parts["flag"], tau=tau, eps=eps)
    o.set_restart_numbers(p.P0)
    o.initialize()
    return o


def random_particles(n, seed, margin=0.0):
    rng = np.random.default_rng(seed)
    lo, hi = margin, L - margin
    return dict(x=lo + rng.random(n) * (hi - lo), y=lo + rng.random(n) * (hi - lo), z=lo + rng.random(n) * (hi - lo),
                u=rng.standard_normal(n), v=rng.standard_normal(n), w=rng.standard_normal(n),
                T=1 + rng.random(n), flag=np.arange(n, dtype=np.int32))


def test_p1_gather_is_linear_along_its_axis_and_constant_across():
    p = base_params()
    parts = random_particles(500, 7)
    o = particle_oracle(p, parts)
    dx = L / N
    i = np.arange(N)
    f = np.cos(0.3 * i) + 0.1 * i    # face profiles along each axis
    g = np.sin(0.7 * i)
    h = 0.05 * i ** 2
    U = np.zeros((3, N, N, N))
    U[0] = f[None, None, :] + g[None, :, None] + h[:, None, None]   # x faces
    U[1] = g[None, :, None] * 2 + h[:, None, None] + f[None, None, :] * 0.5
    U[2] = h[:, None, None] + 3 * f[None, None, :] - g[None, :, None]
    o.set("U", U)
    o.p_gather_velocity()
    x, y, z = parts["x"], parts["y"], parts["z"]
    ii, jj, kk = (x / dx).astype(int), (y / dx).astype(int), (z / dx).astype(int)
    xl, yl, zl = x / dx - ii, y / dx - jj, z / dx - kk

    def lin(a, idx, fr):
        return fr * a[(idx + 1) % N] + (1 - fr) * a[idx]
    ug = lin(f, ii, xl) + g[jj] + h[kk]
    vg = lin(2 * g, jj, yl) + h[kk] + 0.5 * f[ii]
    wg = lin(h, kk, zl) + 3 * f[ii] - g[jj]
    assert np.max(np.abs(o.aux("ug") - ug)) < 1e-13
    assert np.max(np.abs(o.aux("vg") - vg)) < 1e-13
    assert np.max(np.abs(o.aux("wg") - wg)) < 1e-13


def test_p6_velocity_deposit_is_the_transpose_of_the_gather():
    p = base_params(TWC=1)
    parts = random_particles(700, 8)
    o = particle_oracle(p, parts)
    rng = np.random.default_rng(9)
    U = rng.standard_normal((3, N, N, N))
    r = rng.standard_normal((3, 700))
    o.set("U", U)
    o.p_gather_velocity()
    lhs = sum(float(np.dot(o.aux(n), r[c])) for c, n in enumerate(("ug", "vg", "wg")))
    for c, n in enumerate(("RHS_u", "RHS_v", "RHS_w")):
        o.set_aux(n, r[c])
    o.p_deposit_velocity()
    S = o.get("RHS_RU")
    rhs = float(np.sum(U * S))
    assert abs(lhs - rhs) < 1e-11 * (abs(lhs) + 1)


def test_p5_concentration_partition_of_unity():
    p = base_params()
    parts = random_particles(900, 10)
    parts["x"][:3] = [0.0, L, L * (1 - 1e-15)]
    o = particle_oracle(p, parts)
    o.p_deposit_concentration()
    C = o.get("C")
    assert abs(C.sum() - 900) < 1e-10
    assert C.min() >= 0


def test_p8_p9_heat_deposit_is_the_transpose_of_the_temperature_gather():
    p = base_params()
    parts = random_particles(800, 11)
    o = particle_oracle(p, parts)
    rng = np.random.default_rng(12)
    T = rng.standard_normal((N, N, N))
    o.set("T", T)
    o.p_gather_temp("int")
    gathered = o.aux("Tg")
    r = rng.standard_normal(800)
    o.set_aux("RHS_T", r)
    o.p_deposit_temp("int")
    Q = o.get("RHS_Part_Temp")
    lhs, rhs = float(np.dot(gathered, r)), float(np.sum(T * Q))
    assert abs(lhs - rhs) < 1e-11 * (abs(lhs) + 1)
    # constants are reproduced and the deposit of ones is a partition of unity
    o.set("T", np.full((N, N, N), 2.5))
    o.p_gather_temp("int")
    assert np.max(np.abs(o.aux("Tg") - 2.5)) < 1e-13
    o.set_aux("RHS_T", np.ones(800))
    o.p_deposit_temp("int")
    assert abs(o.get("RHS_Part_Temp").sum() - 800) < 1e-10


def rk4_linear_map(M, b, h):
    """One classical RK4 step of y' = M y + b: y -> P y + Q h b with
    P = sum_{k<=4} (hM)^k/k!, Q = sum_{k<=3} (hM)^k/(k+1)!."""
    A = h * M
    I = np.eye(M.shape[0])
    P, Q, term = I.copy(), I.copy(), I.copy()
    for k in range(1, 5):
        term = term @ A
        P = P + term / math.factorial(k)
        if k <= 3:
            Q = Q + term / math.factorial(k + 1)
    return P, Q


def run_frozen_gas(p, parts, ug, nsteps, tau=None, eps=None):
    U = np.zeros((3, N, N, N))
    for c in range(3):
        U[c] = ug[c]
    o = particle_oracle(p, parts, U=U, tau=tau, eps=eps)
    for _ in range(nsteps):
        o.params_update()
        for rk in range(4):
            o.set_rk4_count(rk)
            o.update_p0()
            o.set("RU_new", U)
            o.set("Rho_new", np.ones((N, N, N)))
            o.update_particle()
            o.time_advance_rk4()
        o.set("RU_np1", U)
        o.set("Rho_np1", np.ones((N, N, N)))
        o.time_advance()
    out = o.particles("base")
    o.close()
    return out


@pytest.mark.parametrize("gravity", [0, 1])
def test_p2_p4_frozen_uniform_gas_is_classical_rk4(gravity):
    g = (0.0, -2.0, 0.5)
    p = base_params(Mu0=0.02, Dp=0.01, Rhop=360.0, dt=0.02, ParticleGravity=gravity,
                    gx1=g[0], gy1=g[1], gz1=g[2], I01=3.0, I02=3.0, epsilonp=0.7, Cvp=2.0)
    tau = p.Tp
    assert 0.05 < p.dt / tau < 0.5
    parts = random_particles(200, 13, margin=1.0)
    ug = (0.3, -0.2, 0.1)
    nsteps = 6
    got = run_frozen_gas(p, parts, ug, nsteps)
    order = np.argsort(got["flag"])
    got = {k: v[order] for k, v in got.items()}
    M = np.array([[0.0, 1.0], [0.0, -1.0 / tau]])
    P, Q = rk4_linear_map(M, None, p.dt)
    for c, (xq, uq) in enumerate((("x", "u"), ("y", "v"), ("z", "w"))):
        b = np.array([0.0, ug[c] / tau + (g[c] if gravity else 0.0)])
        y = np.stack([parts[xq], parts[uq]])
        for _ in range(nsteps):
            y = P @ y + (Q @ b)[:, None] * p.dt
        dxw = got[xq] - y[0]
        dxw -= L * np.round(dxw / L)
        assert np.max(np.abs(dxw)) < 1e-12, xq
        assert np.max(np.abs(got[uq] - y[1])) < 1e-12, uq
    # heat off (Nu = 0): dT/dt = alpha / (mp Cvp), alpha = eps pi Dp^2 I0 / 4
    alpha = 0.25 * p.epsilonp * PI * p.Dp * p.Dp * p.I01
    T_expect = parts["T"] + nsteps * p.dt * alpha / (p.mp * p.Cvp)
    assert np.max(np.abs(got["T"] - T_expect)) < 1e-12 * np.max(T_expect)


def test_b2_per_particle_tau_and_eps_equal_to_scalars_reproduce_scalar_run():
    p = base_params(Mu0=0.02, Dp=0.01, Rhop=1800.0, dt=0.02, I01=3.0, I02=3.0, epsilonp=0.7, Cvp=2.0)
    parts = random_particles(150, 14, margin=1.0)
    ug = (0.3, -0.2, 0.1)
    a = run_frozen_gas(p, parts, ug, 3)
    b = run_frozen_gas(p, parts, ug, 3, tau=np.full(150, p.Tp), eps=np.full(150, p.epsilonp))
    for q in "xyzuvwT":
        assert np.array_equal(a[q], b[q]), q


def test_b2_per_particle_tau_follows_each_particles_own_closed_form():
    p = base_params(Mu0=0.02, Dp=0.01, Rhop=1800.0, dt=0.02)
    n = 120
    parts = random_particles(n, 15, margin=1.5)
    rng = np.random.default_rng(16)
    tau = 0.05 + 0.3 * rng.random(n)
    ug = (0.1, 0.2, -0.3)
    got = run_frozen_gas(p, parts, ug, 4, tau=tau)
    order = np.argsort(got["flag"])
    for q_i, (xq, uq) in enumerate((("x", "u"), ("y", "v"), ("z", "w"))):
        for m in range(0, n, 17):
            M = np.array([[0.0, 1.0], [0.0, -1.0 / tau[m]]])
            P, Q = rk4_linear_map(M, None, p.dt)
            b = np.array([0.0, ug[q_i] / tau[m]])
            y = np.array([parts[xq][m], parts[uq][m]])
            for _ in range(4):
                y = P @ y + Q @ b * p.dt
            assert abs(got[uq][order][m] - y[1]) < 1e-12


def test_p7_wrap_keeps_ids_and_shifts_by_one_period():
    p = base_params(dt=0.05)
    parts = random_particles(400, 17)
    rng = np.random.default_rng(18)
    parts["u"] = rng.choice([-1.0, 1.0], 400) * 30.0   # many particles cross a face
    ug = (0.0, 0.0, 0.0)
    got = run_frozen_gas(p, parts, ug, 1)
    assert np.array_equal(np.sort(got["flag"]), np.arange(400))
    for q in "xyz":
        assert np.all(got[q] >= 0) and np.all(got[q] <= L)


# ---------------------------------------------------------------- B1
def test_b1_spectral_initial_condition():
    u = ic.spectral_velocity((N, N, N), (L, L, L), seed=4, k0=4.0, uprime=1.0)
    div = ic.discrete_divergence(u, (L, L, L))
    assert np.max(np.abs(div)) < 1e-12 * np.max(np.abs(u)) * N
    uprime = math.sqrt(float(np.mean(u[0] ** 2 + u[1] ** 2 + u[2] ** 2)) / 3)
    assert abs(uprime - 1.0) < 1e-12
    for c in range(3):
        assert abs(u[c].mean()) < 1e-13
    # the oracle's own divergence operator agrees
    o = new_oracle(base_params(), 0)
    assert np.max(np.abs(o.apply("div_f2c", u[0], u[1], u[2]))) < 1e-11


def test_b1_case_on_a_box_of_unequal_sides():
    """make_case with shape = (nz, ny, nx): one grid spacing L / nx, side
    lengths to match, a discretely solenoidal u' = 1 field, particles inside
    the box, the parameter file's sizes and lengths; shape (n, n, n) is the
    cube bit for bit."""
    a = ic.make_case(16, 300, nu=0.02, seed=2, St=2.0, phi=0.5, heat=True, twc=True, nsteps=4,
                     eps_range=(0.5, 1.0))
    b = ic.make_case(16, 300, nu=0.02, seed=2, St=2.0, phi=0.5, heat=True, twc=True, nsteps=4,
                     eps_range=(0.5, 1.0), shape=(16, 16, 16))
    assert np.array_equal(a.RU, b.RU) and a.params.dt == b.params.dt and a.params.Dp == b.params.Dp
    for q in "xyzuvw":
        assert np.array_equal(a.particles[q], b.particles[q])
    c = ic.make_case(0, 500, nu=0.02, seed=3, St=1.0, phi=0.3, shape=(8, 16, 32))
    p = c.params
    assert (p.Nz_tot, p.Ny_tot, p.Nx_tot) == (8, 16, 32)
    assert (p.Lx, p.Ly, p.Lz) == (L, L / 2, L / 4)
    assert c.RU.shape == (3, 8, 16, 32) and c.Rho.shape == (8, 16, 32)
    div = ic.discrete_divergence(c.RU, (p.Lx, p.Ly, p.Lz))
    assert np.max(np.abs(div)) < 1e-12 * np.max(np.abs(c.RU)) * 32
    uprime = math.sqrt(float(np.mean(c.RU[0] ** 2 + c.RU[1] ** 2 + c.RU[2] ** 2)) / 3)
    assert abs(uprime - 1.0) < 1e-12
    for q, side in (("x", p.Lx), ("y", p.Ly), ("z", p.Lz)):
        assert np.all(c.particles[q] >= 0) and np.all(c.particles[q] < side)
    # mass loading over the box's own volume
    mp = p.Rhop * math.pi * p.Dp ** 3 / 6
    assert abs(mp * 500 / (p.Lx * p.Ly * p.Lz) - 0.3) < 1e-12


# ---------------------------------------------------------------- G5 / G6 / G10 / G14
def prepared(p, npart=0, parts=None, Rho=None):
    o = new_oracle(p, npart)
    load_grid(o, np.zeros((3, N, N, N)), np.ones((N, N, N)) if Rho is None else Rho)
    if parts is not None:
        o.set_particles(parts["x"], parts["y"], parts["z"], parts["u"], parts["v"], parts["w"], parts["T"],
                        parts["flag"])
    o.set_restart_numbers(p.P0)
    o.initialize()
    return o


@pytest.mark.parametrize("cooling", [0, 1])
def test_g5_divergence_closed_form(cooling):
    """divergence = (k/Cp) lap(1/Rho_new) - R/(Cp P0_new Vcell) Q, with Q made
    zero-mean when cooling: checked on 1/Rho_new = 1 + e cos(k x)."""
    p = base_params(k=0.3, cooling=cooling)
    o = prepared(p)
    o.set_rk4_count(0)
    o.update_p0()
    P0n = o.scalars()["P0_new"]
    assert P0n == p.P0
    xc, yc, zc, _, _, _ = grid_coords()
    kv = (2.0, 1.0, 3.0)
    ph = kv[0] * xc + kv[1] * yc + kv[2] * zc
    e = 0.2
    o.set("Rho_new", 1.0 / (1.0 + e * np.cos(ph)))
    Q = 0.4 + 0.1 * np.sin(ph)
    o.set("RHS_Part_Temp", Q)
    o.compute_div_u_new()
    dx = L / N
    lam = sum((2 / dx ** 2) * (1 - math.cos(k * dx)) for k in kv)
    vcell = dx ** 3
    Qe = Q - Q.mean() if cooling else Q
    expect = (p.k / p.Cp) * (-lam * e * np.cos(ph)) - p.R / (p.Cp * P0n * vcell) * Qe
    assert np.max(np.abs(o.get("divergence") - expect)) < 1e-11 * np.max(np.abs(expect))


def test_g6_poisson_rhs_closed_form():
    """RHS_Pois = (div_F2C(RU_new / Rho_face) - divergence) / (dt pre) made
    zero-mean, for a single face mode at uniform density."""
    p = base_params(dt=0.01)
    o = prepared(p)
    o.set_rk4_count(1)     # pre = 1/2
    _, _, _, xf, _, _ = grid_coords()
    xc, yc, zc, _, _, _ = grid_coords()
    kx = 3.0
    RU = np.zeros((3, N, N, N))
    RU[0] = np.sin(kx * xf)
    o.set("RU_new", RU)
    o.set("Rho_new", np.ones((N, N, N)))
    d = 0.3 * np.cos(yc) + 0.05
    o.set("divergence", d)
    o.compute_rhs_pois()
    dx = L / N
    div = (2 / dx) * math.sin(kx * dx / 2) * np.cos(kx * xc)
    raw = (div - d) / (p.dt * 0.5)
    assert np.max(np.abs(o.get("RHS_Pois") - (raw - raw.mean()))) < 1e-10


@pytest.mark.parametrize("rk", [1, 3])
def test_g10_projection_closed_form(rk):
    """RU_np1 -= dt post grad P, RU_new -= dt pre grad P (RU_new = RU_np1 at
    the last substep), with the single-mode gradient in closed form."""
    p = base_params(dt=0.02)
    o = prepared(p)
    o.set_rk4_count(rk)
    xc, yc, zc, xf, yf, zf = grid_coords()
    kv = (1.0, 2.0, 2.0)
    o.set("P", np.cos(kv[0] * xc + kv[1] * yc + kv[2] * zc))
    rng = np.random.default_rng(20)
    np1, new = rng.standard_normal((3, N, N, N)), rng.standard_normal((3, N, N, N))
    o.set("RU_np1", np1)
    o.set("RU_new", new)
    o.update_ru_wp()
    dx = L / N
    pos = [(xf, yc, zc), (xc, yf, zc), (xc, yc, zf)]
    G = np.stack([-(2 / dx) * math.sin(kv[a] * dx / 2) *
                  np.sin(kv[0] * pos[a][0] + kv[1] * pos[a][1] + kv[2] * pos[a][2]) for a in range(3)])
    post = [1 / 6, 1 / 3, 1 / 3, 1 / 6][rk]
    pre = [0.5, 0.5, 1.0][rk] if rk < 3 else None
    e_np1 = np1 - p.dt * post * G
    assert np.max(np.abs(o.get("RU_np1") - e_np1)) < 1e-12
    e_new = e_np1 if rk == 3 else new - p.dt * pre * G
    assert np.max(np.abs(o.get("RU_new") - e_new)) < 1e-12


def test_g14_initial_heat_source():
    """Initialize's tail: T = P0/(R Rho) gathered at x_int, RHS_T =
    (Tg - T_p) Nu pi Dp k, deposited: particles 1 degree colder than a uniform
    gas give sum(RHS_Part_Temp) = Np Nu pi Dp k; particles at the gas
    temperature give zero."""
    p = base_params(Nu=2.0, k=0.05, Dp=0.01)
    parts = random_particles(600, 21)
    T_gas = p.P0 / (p.R * p.Rho0)
    parts["T"] = np.full(600, T_gas - 1.0)
    o = prepared(p, 600, parts)
    expect = 600 * p.Nu * PI * p.Dp * p.k
    assert abs(o.get("RHS_Part_Temp").sum() - expect) < 1e-12 * expect
    parts["T"] = np.full(600, T_gas)
    o = prepared(p, 600, parts)
    assert np.max(np.abs(o.get("RHS_Part_Temp"))) < 1e-15


def test_diagnostics_spectrum_parseval_and_shells():
    """E(k) sums to <u.u>/2, and a single Fourier mode lands in its shell."""
    rng = np.random.default_rng(30)
    u = rng.standard_normal((3, N, N, N))
    e = ic.energy_spectrum(u, (L, L, L))
    assert abs(e.sum() - 0.5 * np.mean(np.sum(u * u, axis=0))) < 1e-12 * e.sum()
    xc, yc, zc, _, _, _ = grid_coords()
    m = np.zeros((3, N, N, N))
    m[0] = np.cos(3 * xc + 4 * yc)        # |k| = 5
    e = ic.energy_spectrum(m, (L, L, L))
    assert np.argmax(e) == 5 and abs(e[5] - 0.25) < 1e-12 and e.sum() - e[5] < 1e-20
    parts = {"u": np.full(10, 1.0), "v": np.full(10, 2.0), "w": np.full(10, 2.0)}
    assert abs(ic.particle_rms(parts) - math.sqrt(3.0)) < 1e-15


def test_oracle_is_bitwise_independent_of_its_thread_count():
    """The oracle's deposits run in parallel over z-planes of the target with
    each cell's contributions added in particle order, so a whole heated,
    two-way coupled step gives the same bits on 1 and on several threads."""
    from helpers import oracle_from_case, small_case
    from oracle.oracle import threads
    case = small_case(n=16, npart=6000, heat=True, variable_rho=True, eps_range=(0.5, 1.0), seed=12)
    out = []
    n0 = threads(0)
    try:
        for nt in (1, 4):
            threads(nt)
            o = oracle_from_case(case)
            o.step(2)
            out.append({k: o.get(k) for k in ("RU", "Rho", "P", "RHS_Part_Temp", "C", "RHS_RU")})
            out[-1].update({"p_" + k: v for k, v in o.particles().items()})
            o.close()
    finally:
        threads(n0)
    for k in out[0]:
        assert np.array_equal(out[0][k], out[1][k]), k
