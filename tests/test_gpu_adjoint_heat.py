"""The adjoint of the step with the heat exchange on (Nu pi Dp k != 0), two-
and one-way coupled: the particle temperature joins the reversed state, the
exchange's gather of T = P0/(R Rho) and its deposit RHS_Part_Temp are
reversed, and so are the divergence's deposit term and Update_P0 (the
thermodynamic pressure P0 evolves; one case with cooling, where it is held
and the deposit term is made mean-free).  dJ/deps_p (per-particle
absorption efficiency), dJ/dT_p(0), dJ/dP0(0), and the gradients the runs
without heat exchange have (tau, A, RU, Rho, u_p), of J = Var(C), are checked
against central finite differences of the full solver along seeded
directions; kept and recomputed gas tapes agree and leave the run where it
ended."""
import copy

import numpy as np
import pytest

from helpers import by_flag, solver_from_case
from tpx import Solver, ic

pytestmark = pytest.mark.gpu

FD_REL = 1e-6  # |FD - adjoint| / max(|FD|, |adjoint|, median |adjoint|) per direction
NSTEPS, EVERY = 6, 2


def _case(twc=True):
    case = ic.make_case(16, 1000, nu=0.02, seed=11, St=2.0, phi=0.5, heat=True, twc=twc, nsteps=NSTEPS,
                        eps_range=(0.5, 1.0))
    tau0 = case.params.Rhop * case.params.Dp ** 2 / (18 * case.params.Mu0)
    case.tau = np.full(len(case.particles["x"]), tau0)
    return case, tau0


def _solver(case, P0=None):
    s = solver_from_case(case)
    if P0 is not None:  # the same case from another initial thermodynamic pressure
        s.close()
        s = Solver(case.params, np_capacity=len(case.particles["x"]))
        s.set("RU", case.RU)
        s.set("Rho", case.Rho)
        s.set("P", case.P)
        pp = case.particles
        s.set_particles(pp["x"], pp["y"], pp["z"], pp["u"], pp["v"], pp["w"], pp["T"], pp["flag"],
                        tau=case.tau, eps=case.eps)
        s.set_restart_numbers(P0, 0.0, 0)
        s.initialize()
    return s


def _J(case, P0=None):
    s = _solver(case, P0)
    J = s.adjoint_record(NSTEPS, EVERY)
    s.adjoint_release()
    s.close()
    return J


def _gradients(case, mode):
    s = solver_from_case(case)
    s.adjoint_set_gas_tape(mode)
    J = s.adjoint_record(NSTEPS, EVERY)
    g = s.adjoint_gradient_full()
    flag = s.particles()["flag"]
    s.adjoint_release()
    s.close()
    # particle gradients back in upload order through the ids
    for q in ("tau", "T", "eps"):
        v = np.empty_like(g[q])
        v[flag] = g[q]
        g[q] = v
    st = np.empty_like(g["state"])
    st[:, flag] = g["state"]
    g["state"] = st
    return J, g


@pytest.fixture(scope="module")
def kept(gpu_available):
    case, tau0 = _case()
    J, g = _gradients(case, "keep")
    return case, tau0, J, g


def _worst(ads, fds, label):
    scale = float(np.median(np.abs(ads)))
    assert scale > 0.0 and np.isfinite(scale), (label, ads)
    worst = 0.0
    for k, (ad, fd) in enumerate(zip(ads, fds)):
        rel = abs(fd - ad) / max(abs(fd), abs(ad), scale)
        worst = max(worst, rel)
        print(f"{label} direction {k}: FD {fd:.10e} adjoint {ad:.10e} rel {rel:.2e}")
    return worst


def _particle_fd(case, g, key, field, h, ndirs, seed, label):
    rng = np.random.default_rng(seed)
    dirs = [rng.standard_normal(len(case.tau)) for _ in range(ndirs)]
    fds = []
    for d in dirs:
        cp, cm = copy.deepcopy(case), copy.deepcopy(case)
        if field == "eps":
            cp.eps = case.eps + h * d
            cm.eps = case.eps - h * d
        else:
            cp.particles[field] = case.particles[field] + h * d
            cm.particles[field] = case.particles[field] - h * d
        fds.append((_J(cp) - _J(cm)) / (2 * h))
    return _worst([float(g[key] @ d) for d in dirs], fds, label)


def test_absorption_efficiency_gradient(kept):
    case, _, J, g = kept
    assert np.all(np.isfinite(g["eps"])) and np.any(g["eps"] != 0.0)
    assert _particle_fd(case, g, "eps", "eps", 1e-4, 8, 43, "eps") <= FD_REL


def test_initial_particle_temperature_gradient(kept):
    case, _, J, g = kept
    assert _particle_fd(case, g, "T", "T", 1e-4, 4, 47, "T_p") <= FD_REL


def test_initial_thermodynamic_pressure_gradient(kept):
    case, _, J, g = kept
    P0 = case.params.P0
    h = 1e-4 * P0
    fd = (_J(case, P0 + h) - _J(case, P0 - h)) / (2 * h)
    assert _worst([g["P0"]], [fd], "P0") <= FD_REL


def test_tau_and_forcing_gradients_with_heat(kept):
    case, tau0, J, g = kept
    rng = np.random.default_rng(53)
    h = 1e-6 * tau0
    dirs = [rng.standard_normal(len(case.tau)) for _ in range(8)]
    fds = []
    for d in dirs:
        cp, cm = copy.deepcopy(case), copy.deepcopy(case)
        cp.tau = case.tau + h * d
        cm.tau = case.tau - h * d
        fds.append((_J(cp) - _J(cm)) / (2 * h))
    worst = _worst([float(g["tau"] @ d) for d in dirs], fds, "tau")
    hA = 1e-4 * case.params.A1
    cp, cm = copy.deepcopy(case), copy.deepcopy(case)
    cp.params.A1 = cp.params.A2 = case.params.A1 + hA
    cm.params.A1 = cm.params.A2 = case.params.A1 - hA
    worst = max(worst, _worst([g["A"]], [(_J(cp) - _J(cm)) / (2 * hA)], "A"))
    assert worst <= FD_REL


@pytest.mark.parametrize("field", ["RU", "Rho"])
def test_initial_gas_gradient_with_heat(kept, field):
    """Rho reaches J also through the first exchange's gather of T."""
    case, _, J, g = kept
    rng = np.random.default_rng(59)
    h = 1e-6
    dirs = [rng.standard_normal(getattr(case, field).shape) for _ in range(3)]
    fds = []
    for d in dirs:
        cp, cm = copy.deepcopy(case), copy.deepcopy(case)
        setattr(cp, field, getattr(case, field) + h * d)
        setattr(cm, field, getattr(case, field) - h * d)
        fds.append((_J(cp) - _J(cm)) / (2 * h))
    assert _worst([float(np.sum(g[field] * d)) for d in dirs], fds, field) <= FD_REL


def test_particle_velocity_gradient_with_heat(kept):
    case, _, J, g = kept
    rng = np.random.default_rng(61)
    h = 1e-6
    dirs = [rng.standard_normal(len(case.tau)) for _ in range(3)]
    fds = []
    for d in dirs:
        cp, cm = copy.deepcopy(case), copy.deepcopy(case)
        cp.particles["v"] = case.particles["v"] + h * d
        cm.particles["v"] = case.particles["v"] - h * d
        fds.append((_J(cp) - _J(cm)) / (2 * h))
    assert _worst([float(g["state"][4] @ d) for d in dirs], fds, "v_p") <= FD_REL


@pytest.mark.parametrize("mode", ["recompute", "segments"])
def test_recomputed_tape_with_heat(kept, mode):
    """The gas recomputed step by step, or replayed segment by segment from
    its checkpoints (with the recorded deposits and P0), gives the kept
    tape's gradients to the rounding of the atomic deposits."""
    case, _, J, g = kept
    Jr, gr = _gradients(case, mode)
    assert Jr == J
    for q in ("tau", "T", "eps"):
        np.testing.assert_allclose(gr[q], g[q], rtol=0, atol=1e-10 * float(np.max(np.abs(g[q]))), err_msg=q)
    scale = max(float(np.max(np.abs(g[q]))) for q in ("RU", "Rho", "P"))
    for q in ("RU", "Rho", "P"):
        np.testing.assert_allclose(gr[q], g[q], rtol=0, atol=1e-10 * scale, err_msg=q)
    for q in ("A", "P0"):
        assert abs(gr[q] - g[q]) <= 1e-10 * abs(g[q]), q


@pytest.mark.parametrize("mode", ["keep", "recompute", "segments"])
def test_sweep_is_bitwise_reproducible(kept, mode):
    """The transposes of the gather and of the exchange's corner gather are
    pulled per cell over deterministic bins (no floating-point atomics), so a
    second recording and sweep, and the sweeps of the recomputed and
    segment-replayed gas tapes, give the kept tape's gradients bit for bit."""
    case, _, J, g = kept
    Jr, gr = _gradients(case, mode)
    assert Jr == J
    for q in ("tau", "T", "eps", "state", "RU", "Rho", "P"):
        assert np.array_equal(gr[q], g[q]), q
    for q in ("A", "P0"):
        assert gr[q] == g[q], q


@pytest.mark.parametrize("mode", ["keep", "recompute", "segments"])
def test_sweep_leaves_the_run_where_it_ended(gpu_available, mode):
    """After the sweep one more step equals a plain run of NSTEPS + 1 steps,
    bit for bit, with the thermodynamic pressure and the deposits restored."""
    case, _ = _case()
    a = solver_from_case(case)
    a.adjoint_set_gas_tape(mode)
    a.adjoint_record(NSTEPS, EVERY)
    a.adjoint_gradient_full()
    a.adjoint_release()
    a.step(1)
    b = solver_from_case(case)
    b.step(NSTEPS + 1)
    for f in ("RU", "Rho", "P"):
        assert np.array_equal(a.get(f), b.get(f)), f
    assert a.scalars()["P0"] == b.scalars()["P0"]
    pa, pb = by_flag(a.particles()), by_flag(b.particles())
    for q in ("x", "y", "z", "u", "v", "w", "T"):
        assert np.array_equal(pa[q], pb[q]), q


def test_one_way_coupling_with_heat(gpu_available):
    """Without the drag deposit the particles still act on the gas through
    the heat exchange, so dJ/deps is nonzero and checked the same way."""
    case, _ = _case(twc=False)
    J, g = _gradients(case, "auto")
    assert np.any(g["eps"] != 0.0)
    assert _particle_fd(case, g, "eps", "eps", 1e-4, 4, 67, "eps (one-way)") <= FD_REL


def test_cooling_branch(gpu_available):
    """cooling = 1: the divergence's deposit term is made mean-free and
    Update_P0 holds P0 (grid.cpp:816-839), so the reversal takes the mean
    removal's transpose and no P0 update; dJ/deps, dJ/dP0 and dJ/dRho
    against central differences."""
    case, _ = _case()
    case.params.cooling = 1
    J, g = _gradients(case, "auto")
    worst = _particle_fd(case, g, "eps", "eps", 1e-4, 4, 71, "eps (cooling)")
    P0 = case.params.P0
    h = 1e-4 * P0
    worst = max(worst, _worst([g["P0"]], [(_J(case, P0 + h) - _J(case, P0 - h)) / (2 * h)], "P0 (cooling)"))
    rng = np.random.default_rng(73)
    dirs = [rng.standard_normal(case.Rho.shape) for _ in range(2)]
    fds = []
    for d in dirs:
        cp, cm = copy.deepcopy(case), copy.deepcopy(case)
        cp.Rho = case.Rho + 1e-6 * d
        cm.Rho = case.Rho - 1e-6 * d
        fds.append((_J(cp) - _J(cm)) / 2e-6)
    worst = max(worst, _worst([float(np.sum(g["Rho"] * d)) for d in dirs], fds, "Rho (cooling)"))
    assert worst <= FD_REL
