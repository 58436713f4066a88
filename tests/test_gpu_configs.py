"""Every benchmark configuration at its full size on the device, through the C
ABI, against the oracle where the oracle finishes in about a minute and
through size-independent properties beyond that.

* heat256 (256^3, 1.6e7 heated two-way coupled particles, St = 1, mass
  loading 1.0): one whole step against the oracle (per field
  max|a-b|/max|b| <= 1e-10, particles by flag), then four more device steps
  that must conserve mass, keep every particle id and position, and meet the
  low-Mach constraint.
* plain128 (128^3, 1e6 particles, St ~ logU[0.1, 10], mass loading 0.5):
  twenty steps against the oracle (<= 1e-8).
* ensemble (64 boxes of 64^3, 1e4..2e5 particles each): two batched steps,
  every box bit for bit equal to its own solo run, two boxes against the
  oracle (<= 1e-10).
* adjoint (128^3, 2e5 particles, St = 2, two-way coupling at mass loading
  0.5): dJ/dtau, dJ/dA and dJ/dRU(t=0) of J = Var(C) against central finite
  differences of the whole solver over 6 steps with a particle checkpoint
  every 2 (so the sweep recomputes three checkpoint segments).  Over the
  configuration's 50 steps no finite difference converges: the reference's
  velocity gather and drag deposit are linear along their own axis and
  piecewise constant across it (particle.cpp:340-360, 494-520), so J jumps
  whenever a perturbed stage position crosses a cell face; with 2e5 particles
  such crossings dominate every central difference beyond about ten steps
  (the difference changes sign with the step size), while the adjoint
  differentiates the map between the jumps.
"""
import copy
import os

import numpy as np
import pytest

import tpx
from helpers import oracle_from_case, particle_errors, rel_err, solver_from_case
from oracle.oracle import threads
from test_gpu_parity import step_end_constraint_residual
from tpx import ic

pytestmark = pytest.mark.gpu

HOST_THREADS = min(16, os.cpu_count() or 1)


def state_errors(o, s, p):
    e = {f"RU{'xyz'[c]}": rel_err(a, b) for c, (a, b) in enumerate(zip(s.get("RU"), o.get("RU")))}
    e["Rho"] = rel_err(s.get("Rho"), o.get("Rho"))
    e["P"] = rel_err(s.get("P"), o.get("P"))
    e["P0"] = abs(s.scalars()["P0"] - o.scalars()["P0"]) / abs(o.scalars()["P0"])
    e.update({f"p_{k}": v for k, v in particle_errors(s.particles(), o.particles(), (p.Lx, p.Ly, p.Lz)).items()})
    return e


@pytest.mark.timeout(900)
def test_heat256_one_step_against_the_oracle_then_invariants(gpu_available):
    case = ic.config_case("heat256", seed=0)
    p = case.params
    npart = len(case.particles["x"])
    print(f"heat256: {p.Nx_tot}^3, {npart} particles, dt {p.dt:.4g}", flush=True)
    s = solver_from_case(case)
    m0 = s.field_sum("Rho")
    s.step(1)
    threads(HOST_THREADS)
    o = oracle_from_case(case)
    o.step(1)
    e = state_errors(o, s, p)
    its = (int(s.scalars()["num_iteration"]), int(o.scalars()["num_iteration"]))
    o.close()
    del o
    print("heat256 one step:", {k: f"{v:.2e}" for k, v in e.items()}, "iterations (device, oracle)", its, flush=True)
    assert its[0] == its[1]
    assert max(e.values()) <= 1e-10, e
    s.step(4)
    m1 = s.field_sum("Rho")
    assert abs(m1 - m0) <= 1e-12 * abs(m0), (m0, m1)
    parts = s.particles()
    assert np.array_equal(np.sort(parts["flag"]), np.arange(npart))
    for q, Lq in (("x", p.Lx), ("y", p.Ly), ("z", p.Lz)):
        assert np.all((parts[q] >= 0) & (parts[q] <= Lq)), q
    assert np.all(np.isfinite(parts["T"])) and np.all(np.isfinite(parts["u"]))
    r = step_end_constraint_residual(s, p)
    print(f"heat256 after 5 steps: mass drift {abs(m1 - m0) / m0:.1e}, low-Mach residual {r:.1e}", flush=True)
    assert r <= 1e-6, r


@pytest.mark.timeout(900)
def test_plain128_twenty_steps_against_the_oracle(gpu_available):
    case = ic.config_case("plain128", seed=0)
    p = case.params
    print(f"plain128: {p.Nx_tot}^3, {len(case.particles['x'])} particles, St in "
          f"[{case.tau.min() / case.info['tau_eta']:.3g}, {case.tau.max() / case.info['tau_eta']:.3g}]", flush=True)
    s = solver_from_case(case)
    threads(HOST_THREADS)
    o = oracle_from_case(case)
    for n in range(20):
        s.step(1)
        o.step(1)
        if n == 0:
            e1 = state_errors(o, s, p)
            print("plain128 step 1:", {k: f"{v:.2e}" for k, v in e1.items()}, flush=True)
            assert max(e1.values()) <= 1e-10, e1
    e = state_errors(o, s, p)
    print("plain128 step 20:", {k: f"{v:.2e}" for k, v in e.items()}, flush=True)
    assert s.scalars()["total_iterations"] > 0
    assert max(e.values()) <= 1e-8, e


@pytest.mark.timeout(900)
def test_ensemble_64_boxes_batched_equal_solo_and_oracle(gpu_available):
    cases = ic.ensemble_cases(64, 64, seed=0)
    counts = [len(c.particles["x"]) for c in cases]
    print(f"ensemble: 64 boxes of 64^3, particles {min(counts)}..{max(counts)}, total {sum(counts)}", flush=True)
    bat = [solver_from_case(c) for c in cases]
    batch = tpx.Batch(bat)
    batch.step(2)
    batch.close()
    for b, c in enumerate(cases):
        s = solver_from_case(c)
        s.step(2)
        for f in ("RU", "Rho", "P"):
            assert np.array_equal(bat[b].get(f), s.get(f)), (b, f)
        pa, ps = bat[b].particles(), s.particles()
        for q in ("x", "y", "z", "u", "v", "w", "T", "flag"):
            assert np.array_equal(pa[q], ps[q]), (b, q)
        assert bat[b].scalars()["total_iterations"] == s.scalars()["total_iterations"], b
        s.close()
    print("ensemble: every box equals its solo run bit for bit", flush=True)
    threads(HOST_THREADS)
    for b in (int(np.argmin(counts)), int(np.argmax(counts))):
        o = oracle_from_case(cases[b])
        o.step(2)
        e = state_errors(o, bat[b], cases[b].params)
        o.close()
        print(f"ensemble box {b} ({counts[b]} particles) vs oracle after 2 steps:",
              {k: f"{v:.2e}" for k, v in e.items()}, flush=True)
        assert max(e.values()) <= 1e-10, (b, e)
    for s in bat:
        s.close()


# ---------------------------------------------------------------- adjoint (A)
A_STEPS, A_EVERY = 6, 2
FD_REL = 1e-6


def _adjoint_case():
    case = ic.make_case(128, 200_000, nu=0.01, seed=0, St=2.0, phi=0.5, heat=False, twc=True, nsteps=A_STEPS)
    p = case.params
    tau0 = p.Rhop * p.Dp ** 2 / (18 * p.Mu0)
    case.tau = np.full(len(case.particles["x"]), tau0)
    return case, tau0


def _J(case):
    s = solver_from_case(case)
    J = s.adjoint_record(A_STEPS, A_EVERY)
    s.adjoint_release()
    s.close()
    return J


def _worst(ads, fds, label):
    scale = float(np.median(np.abs(ads)))
    assert scale > 0.0 and np.isfinite(scale), (label, ads)
    worst = 0.0
    for k, (ad, fd) in enumerate(zip(ads, fds)):
        rel = abs(fd - ad) / max(abs(fd), abs(ad), scale)
        worst = max(worst, rel)
        print(f"A {label} direction {k}: FD {fd:.10e} adjoint {ad:.10e} rel {rel:.2e}", flush=True)
    return worst


@pytest.mark.timeout(900)
def test_adjoint_config_against_finite_differences(gpu_available):
    case, tau0 = _adjoint_case()
    s = solver_from_case(case)
    J = s.adjoint_record(A_STEPS, A_EVERY)
    g = s.adjoint_gradient_full()
    flag = s.particles()["flag"]
    s.adjoint_release()
    s.close()
    gt = np.empty_like(g["tau"])
    gt[flag] = g["tau"]
    print(f"A: J = {J:.12e}, |dJ/dtau| = {np.linalg.norm(gt):.3e}, dJ/dA = {g['A']:.6e}", flush=True)
    rng = np.random.default_rng(2024)
    h = 1e-4 * tau0
    dirs = [rng.standard_normal(len(case.tau)) for _ in range(8)]
    fds = []
    for d in dirs:
        cp, cm = copy.deepcopy(case), copy.deepcopy(case)
        cp.tau, cm.tau = case.tau + h * d, case.tau - h * d
        fds.append((_J(cp) - _J(cm)) / (2 * h))
    w_tau = _worst([float(gt @ d) for d in dirs], fds, "tau")
    hA = 1e-3 * case.params.A1  # dJ/dA is small: a smaller step is rounding-limited
    cp, cm = copy.deepcopy(case), copy.deepcopy(case)
    cp.params.A1 = cp.params.A2 = case.params.A1 + hA
    cm.params.A1 = cm.params.A2 = case.params.A1 - hA
    w_A = _worst([g["A"]], [(_J(cp) - _J(cm)) / (2 * hA)], "A")
    dirs = [rng.standard_normal(case.RU.shape) for _ in range(4)]
    fds = []
    for d in dirs:
        cp, cm = copy.deepcopy(case), copy.deepcopy(case)
        cp.RU, cm.RU = case.RU + 1e-5 * d, case.RU - 1e-5 * d
        fds.append((_J(cp) - _J(cm)) / 2e-5)
    w_RU = _worst([float(np.sum(g["RU"] * d)) for d in dirs], fds, "RU")
    assert max(w_tau, w_A, w_RU) <= FD_REL, (w_tau, w_A, w_RU)
    # the configuration's own scheme, the gas checkpointed and each segment
    # replayed in the sweep, gives the same gradients bit for bit
    s = solver_from_case(case)
    s.adjoint_set_gas_tape("segments")
    assert s.adjoint_record(A_STEPS, A_EVERY) == J
    gs = s.adjoint_gradient_full()
    s.adjoint_release()
    s.close()
    assert gs["A"] == g["A"]
    for q in ("tau", "state", "RU", "Rho", "P"):
        assert np.array_equal(gs[q], g[q]), q
