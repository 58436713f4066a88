"""The adjoint of the two-way coupled step (drag deposit on): the particles'
stages and the gas substeps are reversed interleaved, the gas's adjoint of
its drag term returning to each particle through the gather.  dJ/dtau_p,
dJ/dA, dJ/dRU(t=0) and dJ/du_p(0) of J = Var(C) are checked against central
finite differences of the full two-way coupled solver along seeded
directions (16 for tau), with the gas tape kept and recomputed."""
import copy

import numpy as np
import pytest

from helpers import solver_from_case
from tpx import ic

pytestmark = pytest.mark.gpu

FD_REL = 1e-6  # |FD - adjoint| / max(|FD|, |adjoint|, median |adjoint|) per direction
NSTEPS, EVERY = 6, 2


def _case():
    case = ic.make_case(16, 1000, nu=0.02, seed=7, St=2.0, phi=0.5, heat=False, twc=True, nsteps=NSTEPS)
    tau0 = case.params.Rhop * case.params.Dp ** 2 / (18 * case.params.Mu0)
    case.tau = np.full(len(case.particles["x"]), tau0)
    return case, tau0


def _J(case, mode="auto"):
    s = solver_from_case(case)
    s.adjoint_set_gas_tape(mode)
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
    tau = np.empty_like(g["tau"])
    tau[flag] = g["tau"]
    st = np.empty_like(g["state"])
    st[:, flag] = g["state"]
    g["tau"], g["state"] = tau, st
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


def test_tau_gradient_two_way(kept):
    case, tau0, J, g = kept
    rng = np.random.default_rng(31)
    h = 1e-6 * tau0
    dirs = [rng.standard_normal(len(case.tau)) for _ in range(16)]
    fds = []
    for d in dirs:
        cp, cm = copy.deepcopy(case), copy.deepcopy(case)
        cp.tau = case.tau + h * d
        cm.tau = case.tau - h * d
        fds.append((_J(cp) - _J(cm)) / (2 * h))
    assert _worst([float(g["tau"] @ d) for d in dirs], fds, "tau") <= FD_REL


def test_forcing_gradient_two_way(kept):
    case, _, J, g = kept
    h = 1e-4 * case.params.A1
    cp, cm = copy.deepcopy(case), copy.deepcopy(case)
    cp.params.A1 = cp.params.A2 = case.params.A1 + h
    cm.params.A1 = cm.params.A2 = case.params.A1 - h
    fd = (_J(cp) - _J(cm)) / (2 * h)
    assert _worst([g["A"]], [fd], "A") <= FD_REL


def test_initial_velocity_field_gradient_two_way(kept):
    case, _, J, g = kept
    rng = np.random.default_rng(37)
    h = 1e-6
    dirs = [rng.standard_normal(case.RU.shape) for _ in range(4)]
    fds = []
    for d in dirs:
        cp, cm = copy.deepcopy(case), copy.deepcopy(case)
        cp.RU = case.RU + h * d
        cm.RU = case.RU - h * d
        fds.append((_J(cp) - _J(cm)) / (2 * h))
    assert _worst([float(np.sum(g["RU"] * d)) for d in dirs], fds, "RU") <= FD_REL


def test_particle_velocity_gradient_two_way(kept):
    case, _, J, g = kept
    rng = np.random.default_rng(41)
    h = 1e-6
    dirs = [rng.standard_normal(len(case.tau)) for _ in range(4)]
    fds = []
    for d in dirs:
        cp, cm = copy.deepcopy(case), copy.deepcopy(case)
        cp.particles["u"] = case.particles["u"] + h * d
        cm.particles["u"] = case.particles["u"] - h * d
        fds.append((_J(cp) - _J(cm)) / (2 * h))
    assert _worst([float(g["state"][3] @ d) for d in dirs], fds, "u_p") <= FD_REL


@pytest.mark.parametrize("mode", ["recompute", "segments"])
def test_recomputed_tape_two_way(kept, mode):
    """The gas recomputed step by step, or replayed segment by segment from
    its checkpoints (with the recorded drag deposits), gives the kept tape's
    gradients to the rounding of the atomic deposits."""
    case, _, J, g = kept
    Jr, gr = _gradients(case, mode)
    assert Jr == J
    np.testing.assert_allclose(gr["tau"], g["tau"], rtol=0, atol=1e-10 * float(np.max(np.abs(g["tau"]))))
    # one scale for the gas fields: at uniform density dJ/dP0 is rounding alone
    scale = max(float(np.max(np.abs(g[q]))) for q in ("RU", "Rho", "P"))
    for q in ("RU", "Rho", "P"):
        np.testing.assert_allclose(gr[q], g[q], rtol=0, atol=1e-10 * scale, err_msg=q)
    assert abs(gr["A"] - g["A"]) <= 1e-10 * abs(g["A"])


def test_particle_only_call_matches_full_sweep(kept):
    """Under two-way coupling the particle-only call runs the gas sweep too."""
    case, _, J, g = kept
    s = solver_from_case(case)
    s.adjoint_set_gas_tape("keep")
    s.adjoint_record(NSTEPS, EVERY)
    gt = s.adjoint_gradient()
    flag = s.particles()["flag"]
    s.adjoint_release()
    s.close()
    tau = np.empty_like(gt)
    tau[flag] = gt
    np.testing.assert_allclose(tau, g["tau"], rtol=0, atol=1e-10 * float(np.max(np.abs(g["tau"]))))
