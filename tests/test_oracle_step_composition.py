"""Pins of the oracle's whole-step composition (or_step), CPU only.

The operator pins elsewhere fix each substep method on its own; these fix how
simulation.cpp:57-93 composes them: which stage value each term reads, where
the heat source lags, where Rho_face is rebuilt, where make_mean_U0 runs and
when params::update switches the *1 coefficients for the *2 ones.  Each case
is chosen so that the whole RK4 step collapses to a linear system whose
one-step map is known in closed form; the classical RK4 map of y' = M y + b
(y -> P y + h Q b, P = sum_{k<=4} (hM)^k/k!, Q = sum_{k<=3} (hM)^k/(k+1)!) is
the expected value, and the oracle must reproduce it to round-off.

* shear mode  RU_x = e sin(k y), uniform density, no particles: convection,
  divergence and pressure vanish exactly, so every step multiplies RU by
  R(dt (A - mu lam_y)), R the RK4 stability polynomial (grid.cpp:776-814,
  RU_new built from RU, the forcing on RU_int).
* heat pair   one particle at rest at every cell centre, gas at rest: the
  particle temperature and the thermodynamic pressure follow the 2x2 affine
  system T_p' = iS (K (P0/(R rho0) - T_p) + alpha),
  P0' = -(R/Cv) K (P0/(R rho0) - T_p) / V_cell, which holds only if
  Update_P0 reads the heat source left by the previous substep
  (grid.cpp:816-830 before Update_Particle, :865-888).
* drag pair   uniform gas and one particle per cell moving together, two-way
  coupled: (RU, u_p) follow their 2x2 linear system within the step and
  make_mean_U0 puts RU back to rho0 U0 once, after the RK4 loop
  (simulation.cpp:80, tensor1.cpp:406-411).  It shifts RU_np1 only: RU_int,
  which TimeAdvance_RK4 set to RU_new = RU_np1 at the last substep and which
  TimeAdvance does not touch (grid.cpp:595-606, 890-900), keeps the unshifted
  value, so the next step's first stage derivative is taken there while its
  updates start from the shifted RU.
* body force  density varying along z, gravity along x, no conduction: RU_x(z)
  follows y' = mu D2 (y / rho) + (rho - rho0) g, so the division by
  Rho_face = I_C2F(Rho) and the body force on Rho_face - rho0
  (grid.cpp:805-809, rebuilt by Compute_RHS_Pois :841-849) are fixed.
* threshold   A, g, I0 and Iteration take their *2 values from the first step
  whose start time exceeds `threshold` (params.cpp:75-86).
* projection  a divergence-free start stays discretely divergence-free at
  the end of a whole step (the RK4 post/pre coefficients of Update_RU_WP pair
  with those of Update_RU_WOP).

The last test builds copies of the oracle with a planted composition mistake
and checks that these pins catch each one.
"""
import math
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from oracle.oracle import Oracle
from tpx import ic
from tpx.params import PI, Params

ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
L = 2 * math.pi


def rk4_affine_map(M, h):
    """P, Q of one classical RK4 step of y' = M y + b: y1 = P y0 + h Q b."""
    A = h * np.atleast_2d(M)
    n = A.shape[0]
    P, Q, term = np.eye(n), np.eye(n), np.eye(n)
    for k in range(1, 5):
        term = term @ A
        P = P + term / math.factorial(k)
        if k <= 3:
            Q = Q + term / math.factorial(k + 1)
    return P, Q


def rk4_affine_step(M, h, y_base, b, y_first=None):
    """One classical RK4 step of y' = M y + b from y_base whose first stage
    derivative is taken at y_first instead of y_base: a change d of the first
    stage's argument moves the result by h (I/6 + hM/6 + (hM)^2/12 +
    (hM)^3/24) M d (k1 enters with weight h/6 directly and through the later
    stages' arguments y + h k/2, y + h k/2, y + h k)."""
    P, Q = rk4_affine_map(M, h)
    y = P @ y_base + h * (Q @ b)
    if y_first is not None:
        A = h * np.atleast_2d(M)
        n = A.shape[0]
        W = np.eye(n) / 6 + A / 6 + A @ A / 12 + A @ A @ A / 24
        y = y + h * (W @ (np.atleast_2d(M) @ (y_first - y_base)))
    return y


def params(n, **kw):
    p = Params(Nx_tot=n, Ny_tot=n, Nz_tot=n, Lx=L, Ly=L, Lz=L, bs=2, dt=1e-2, Mu0=0.02, A1=0.0, A2=0.0,
               TWC=0, Nu=0.0, I01=0.0, I02=0.0, k=0.0, Iteration1=50, Iteration2=50, epsilon=1e-12)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def start(p, RU, Rho, parts=None, lib_path=None, P=None):
    """A run of the case on the oracle at `lib_path` (None: the in-tree one),
    or on the device library through its C ABI when lib_path is "gpu" (the
    device tests reuse these checks unchanged)."""
    n = 0 if parts is None else len(parts["x"])
    if lib_path == "gpu":
        from tpx import Solver
        o = Solver(p, np_capacity=max(n, 1))
    else:
        o = Oracle(p.struct_fields(), np_capacity=max(n, 1), lib_path=lib_path)
    o.set("RU", RU)
    o.set("Rho", Rho)
    o.set("P", np.zeros_like(Rho) if P is None else P)
    if parts is not None:
        o.set_particles(parts["x"], parts["y"], parts["z"], parts["u"], parts["v"], parts["w"], parts["T"],
                        parts["flag"])
    o.set_restart_numbers(p.P0, 0.0, 0)
    o.initialize()
    return o


def cell_centre_particles(n, T=1.0, u=(0.0, 0.0, 0.0), frac=(0.5, 0.5, 0.5)):
    """One particle per cell, at the same fractional position in every cell."""
    d = L / n
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    m = n ** 3
    return {"x": ((i + frac[0]) * d).ravel(), "y": ((j + frac[1]) * d).ravel(), "z": ((k + frac[2]) * d).ravel(),
            "u": np.full(m, u[0]), "v": np.full(m, u[1]), "w": np.full(m, u[2]), "T": np.full(m, T),
            "flag": np.arange(m, dtype=np.int32)}


def by_flag(parts):
    order = np.argsort(parts["flag"], kind="stable")
    return {k: v[order] for k, v in parts.items()}


# ------------------------------------------------------------------ checks
# Each check runs one case on an oracle library (the in-tree one, or a mutant)
# and returns the worst relative deviation from the closed form.

def check_shear_mode(lib_path=None, A=0.0, nsteps=3):
    n = 16
    ky = 2.0
    p = params(n, A1=A, A2=A, dt=0.05, Mu0=0.05)
    dy = L / n
    j = np.arange(n)
    ycell = (j + 0.5) * dy                   # RU.x sits at y = (j + 1/2) dy
    eps = 0.3
    RU = np.zeros((3, n, n, n))
    RU[0] = eps * np.sin(ky * ycell)[None, :, None]
    o = start(p, RU, np.ones((n, n, n)), lib_path=lib_path)
    o.step(nsteps)
    lam = (2.0 / dy ** 2) * (1 - math.cos(ky * dy))
    z = p.dt * (A - p.Mu0 * lam)
    amp = (1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24) ** nsteps
    got = o.get("RU")
    err = float(np.max(np.abs(got[0] - amp * RU[0]))) / eps
    err = max(err, float(np.max(np.abs(got[1:]))) / eps)
    err = max(err, float(np.max(np.abs(o.get("Rho") - 1.0))))
    o.close()
    return err


def heat_pair_params(n, I01=10.0, I02=10.0, threshold=1e9):
    return params(n, TWC=1, Nu=2.0, k=0.1, Dp=0.1, Rhop=10.0, Cvp=2.5, Cv=2.5, Cp=3.5, epsilonp=0.8,
                  I01=I01, I02=I02, dt=0.05, threshold=threshold, epsilon=1e-9)


def heat_pair_expected(p, Tp0, nsteps):
    """RK4 of the particle temperature / P0 pair, one particle per cell."""
    n = p.Nx_tot
    vcell = (p.Lx / n) * (p.Ly / n) * (p.Lz / n)
    K = p.Nu * PI * p.Dp * p.k
    iS = 1.0 / (p.mp * p.Cvp)
    R = p.R
    M = np.array([[-iS * K, iS * K / (R * p.Rho0)],
                  [R * K / (p.Cv * vcell), -K / (p.Cv * p.Rho0 * vcell)]])
    P, Q = rk4_affine_map(M, p.dt)
    y = np.array([Tp0, p.P0])
    t = 0.0
    for _ in range(nsteps):
        I0 = p.I02 if t > p.threshold else p.I01
        alpha = 0.25 * p.epsilonp * PI * p.Dp * p.Dp * I0
        y = P @ y + p.dt * (Q @ np.array([iS * alpha, 0.0]))
        t += p.dt
    return y


def check_heat_pair(lib_path=None, nsteps=3, I01=10.0, I02=10.0, threshold=1e9):
    n = 8
    p = heat_pair_params(n, I01, I02, threshold)
    Tp0 = 1.3
    parts = cell_centre_particles(n, T=Tp0)
    o = start(p, np.zeros((3, n, n, n)), np.full((n, n, n), p.Rho0), parts, lib_path=lib_path)
    o.step(nsteps)
    Tp, P0 = heat_pair_expected(p, Tp0, nsteps)
    got = o.particles("base")
    err = float(np.max(np.abs(got["T"] - Tp))) / Tp
    err = max(err, abs(o.scalars()["P0"] - P0) / P0)
    # the gas stays at rest (a uniform source is removed with the mean)
    err = max(err, float(np.max(np.abs(o.get("RU")))))
    o.close()
    return err


def check_drag_pair(lib_path=None, nsteps=3, U0=0.2, gravity=True):
    n = 8
    g = (0.7, -0.4, 0.3) if gravity else (0.0, 0.0, 0.0)
    p = params(n, TWC=1, Dp=0.1, Rhop=50.0, Mu0=0.05, dt=0.05, U0=U0, V0=-U0, W0=0.5 * U0,
               ParticleGravity=int(gravity), gx1=g[0], gy1=g[1], gz1=g[2], gx2=g[0], gy2=g[1], gz2=g[2])
    up0 = np.array([-0.5, 0.4, 0.25])
    parts = cell_centre_particles(n, u=tuple(up0), frac=(0.3, 0.6, 0.45))
    U0v = np.array([p.U0, p.V0, p.W0])
    RU = np.broadcast_to((p.Rho0 * U0v)[:, None, None, None], (3, n, n, n)).copy()
    o = start(p, RU, np.full((n, n, n), p.Rho0), parts, lib_path=lib_path)
    o.step(nsteps)
    tau = p.Tp
    vcell = (L / n) ** 3
    c = p.mp / vcell
    M = np.array([[-c / (tau * p.Rho0), c / tau],        # RU' = -(mp/Vcell) (RU/rho0 - u_p)/tau
                  [1.0 / (tau * p.Rho0), -1.0 / tau]])   # u_p' = (RU/rho0 - u_p)/tau + g
    got = by_flag(o.particles("base"))
    x0 = by_flag(parts)
    err = float(np.max(np.abs(o.get("RU") - RU)))
    M3 = np.zeros((3, 3))            # (RU, u_p, x): x' = u_p
    M3[:2, :2] = M
    M3[2, 1] = 1.0
    for a, (q, xq) in enumerate((("u", "x"), ("v", "y"), ("w", "z"))):
        y = np.array([p.Rho0 * U0v[a], up0[a], 0.0])
        first = None
        b = np.array([0.0, g[a], 0.0])
        for _ in range(nsteps):
            y = rk4_affine_step(M3, p.dt, y, b, first)
            first = y.copy()                 # RU_int keeps the unshifted RU_np1
            y[0] = p.Rho0 * U0v[a]           # make_mean_U0 after the RK4 loop
        err = max(err, float(np.max(np.abs(got[q] - y[1]))) / max(abs(y[1]), 1e-3))
        d = got[xq] - (x0[xq] + y[2])
        d -= L * np.round(d / L)
        err = max(err, float(np.max(np.abs(d))) / L)
    o.close()
    return err


def check_body_force(lib_path=None, nsteps=3, threshold=1e9):
    n = 16
    gx1, gx2 = 0.8, -0.5
    p = params(n, Mu0=0.05, dt=0.05, gx1=gx1, gx2=gx2, threshold=threshold, k=0.0)
    dz = L / n
    zc = (np.arange(n) + 0.5) * dz
    rho_z = 1.0 + 0.3 * np.cos(zc) + 0.1 * np.sin(3 * zc)
    Rho = np.broadcast_to(rho_z[:, None, None], (n, n, n)).copy()
    RU = np.zeros((3, n, n, n))
    ru0 = 0.2 * np.sin(2 * zc)
    RU[0] = ru0[:, None, None]
    o = start(p, RU, Rho, lib_path=lib_path)
    o.step(nsteps)
    D2 = (np.diag(np.full(n, -2.0)) + np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1))
    D2[0, -1] = D2[-1, 0] = 1.0
    D2 /= dz * dz
    M = p.Mu0 * D2 @ np.diag(1.0 / rho_z)
    P, Q = rk4_affine_map(M, p.dt)
    y = ru0.copy()
    t = 0.0
    for _ in range(nsteps):
        g = gx2 if t > threshold else gx1
        y = P @ y + p.dt * (Q @ ((rho_z - p.Rho0) * g))
        y -= y.mean()                      # make_mean_U0 (U0 = 0)
        t += p.dt
    got = o.get("RU")
    err = float(np.max(np.abs(got[0] - y[:, None, None]))) / float(np.max(np.abs(y)))
    err = max(err, float(np.max(np.abs(got[1:]))))
    err = max(err, float(np.max(np.abs(o.get("Rho") - Rho))))
    o.close()
    return err


# ------------------------------------------------------------------ tests
@pytest.mark.parametrize("A", [0.0, 0.8])
def test_shear_mode_is_rk4_of_its_decay(A):
    assert check_shear_mode(A=A) < 1e-13


def test_heat_pair_follows_rk4_of_the_lagged_source():
    assert check_heat_pair() < 1e-13


@pytest.mark.parametrize("gravity", [False, True])
def test_drag_pair_and_make_mean_u0_once_per_step(gravity):
    assert check_drag_pair(gravity=gravity) < 1e-12


def test_body_force_and_rho_face():
    assert check_body_force() < 1e-12


def test_threshold_switches_a_g_i0_from_the_first_late_step():
    # threshold between the 2nd and 3rd step start: t = 0, dt, 2 dt, 3 dt
    dt = 0.05
    thr = 1.5 * dt
    assert check_heat_pair(nsteps=4, I01=10.0, I02=25.0, threshold=thr) < 1e-13
    assert check_body_force(nsteps=4, threshold=thr) < 1e-12
    # the forcing coefficient on the shear mode: A1 for two steps, then A2
    n, ky, eps = 16, 2.0, 0.3
    p = params(n, A1=0.8, A2=-1.5, dt=dt, Mu0=0.05, threshold=thr)
    dy = L / n
    RU = np.zeros((3, n, n, n))
    RU[0] = eps * np.sin(ky * (np.arange(n) + 0.5) * dy)[None, :, None]
    o = start(p, RU, np.ones((n, n, n)))
    o.step(4)
    lam = (2.0 / dy ** 2) * (1 - math.cos(ky * dy))
    R = lambda z: 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24  # noqa: E731
    amp = R(dt * (0.8 - p.Mu0 * lam)) ** 2 * R(dt * (-1.5 - p.Mu0 * lam)) ** 2
    assert np.max(np.abs(o.get("RU")[0] - amp * RU[0])) < 1e-13 * eps
    o.close()


def test_threshold_switches_the_iteration_cap():
    """With epsilon < 0 the stopping rule never fires, so every solve runs
    exactly Iteration iterations: Iteration1 before the switch, Iteration2
    from the first step that starts after the threshold."""
    case = ic.make_case(16, 0, nu=0.02, seed=4, St=1.0, phi=0.3, heat=False, twc=False)
    p = case.params
    rng = np.random.default_rng(4)
    case.Rho = 1.0 + 0.1 * rng.random((16, 16, 16))
    p.epsilon = -1.0
    p.Iteration1, p.Iteration2 = 3, 5
    p.threshold = 1.5 * p.dt
    o = start(p, case.RU, case.Rho)
    seen = []
    for _ in range(4):
        o.step(1)
        seen.append(int(o.scalars()["num_iteration"]))
    o.close()
    assert seen == [3, 3, 5, 5], seen


def test_whole_step_keeps_a_divergence_free_field_divergence_free():
    n = 16
    RU = ic.spectral_velocity((n, n, n), (L, L, L), seed=8, k0=2.0, uprime=1.0)
    p = params(n, Mu0=0.02, dt=0.02, A1=0.3, A2=0.3)
    o = start(p, RU, np.ones((n, n, n)))
    div0 = np.abs(o.apply("div_f2c", *RU)).max()
    o.step(2)
    RUe = o.get("RU")
    div = np.abs(o.apply("div_f2c", *RUe)).max()
    scale = np.abs(RUe).max() / (L / n)
    o.close()
    assert div0 < 1e-12 * scale
    assert div < 1e-11 * scale, (div, scale)


# ------------------------------------------------------------------ mutations
MUTANTS = {
    # RU_new built from RU_int instead of the step's base RU (grid.cpp:812)
    "ru_new_from_ru_int": (
        "ew_lc2(s, T1(s, OR_RU_NEW)[c], 1, T1(s, OR_RU)[c], dt * s->pre[s->rk], RHS[c]);",
        "ew_lc2(s, T1(s, OR_RU_NEW)[c], 1, T1(s, OR_RU_INT)[c], dt * s->pre[s->rk], RHS[c]);"),
    # Update_P0 after Update_Particle: the heat source is no longer the lagged one
    "p0_after_particle": (
        "      or_update_p0(s);\n      or_update_particle(s);\n",
        "      or_update_particle(s);\n      or_update_p0(s);\n"),
    # make_mean_U0 inside the RK4 loop instead of once per step
    "mean_u0_every_substep": (
        "      or_time_advance_rk4(s);\n    }\n    or_make_mean_u0(s, a);\n",
        "      or_time_advance_rk4(s);\n      or_make_mean_u0(s, a);\n    }\n"),
    # Rho_face left as Update_RU_WOP's (Rho_face - rho0) instead of rebuilt
    "rho_face_not_rebuilt": (
        "  i_c2f(s, Rf, T0(s, OR_RHO_NEW));\n  for (int c = 0; c < 3; ++c) ew_div(s, U[c], T1(s, OR_RU_NEW)[c], Rf[c]);",
        "  for (int c = 0; c < 3; ++c) ew_div(s, U[c], T1(s, OR_RU_NEW)[c], Rf[c]);"),
    # params::update never switching to the *2 values
    "no_threshold_switch": ("if (s->T_cur > s->p.threshold) {", "if (0) {"),
}


def build_mutant(tmp_path, name):
    old, new = MUTANTS[name]
    src = open(os.path.join(ORACLE_DIR, "or_core.c")).read()
    assert src.count(old) == 1, f"mutation {name}: anchor not found once in or_core.c"
    d = tmp_path / name
    d.mkdir()
    (d / "or_core.c").write_text(src.replace(old, new))
    for f in ("or_fft.c", "oracle.h"):
        shutil.copy(os.path.join(ORACLE_DIR, f), d / f)
    out = d / "liboracle_mutant.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-fopenmp", "-ffp-contract=off", "-std=gnu11", "-shared", "-o", str(out),
                    str(d / "or_core.c"), str(d / "or_fft.c"), "-lm"], check=True)
    return str(out)


PIN_CHECKS = {
    "body_force": lambda lib: check_body_force(lib),
    "shear": lambda lib: check_shear_mode(lib, A=0.8),
    "drag": lambda lib: check_drag_pair(lib),
    "threshold_body": lambda lib: check_body_force(lib, nsteps=4, threshold=1.5 * 0.05),
    "heat": lambda lib: check_heat_pair(lib),
    "threshold_heat": lambda lib: check_heat_pair(lib, nsteps=4, I01=10.0, I02=25.0, threshold=1.5 * 0.05),
}


def run_pins(lib_path):
    """Run every pin on `lib_path`, printing one line per pin as it ends."""
    for name, check in PIN_CHECKS.items():
        print("PIN", name, repr(float(check(lib_path))), flush=True)


@pytest.mark.parametrize("mutant", sorted(MUTANTS))
def test_each_planted_composition_mistake_fails_a_pin(tmp_path, mutant):
    """The pins run on the mutant in a child process: a mutant whose state
    goes non-finite may index outside the oracle's arrays and end the child,
    so the pins that ended before are read from its output line by line."""
    path = build_mutant(tmp_path, mutant)
    code = ("import sys; sys.path[:0] = %r; import conftest; "
            "from test_oracle_step_composition import run_pins; run_pins(%r)") % (sys.path[:4], path)
    r = subprocess.run([sys.executable, "-c", code], cwd=os.path.dirname(os.path.abspath(__file__)),
                       capture_output=True, text=True, timeout=300)
    errs = {ln.split()[1]: float(ln.split()[2]) for ln in r.stdout.splitlines() if ln.startswith("PIN ")}
    caught = {k: v for k, v in errs.items() if not v < 1e-10}
    assert caught, f"mutation {mutant} passed every composition pin that ran: {errs} (exit {r.returncode})"
