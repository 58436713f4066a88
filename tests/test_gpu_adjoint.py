"""The hand-written adjoint of the particle path: dJ/dtau_p of J = Var(C)
(C the cloud-in-cell concentration of the final positions) for one-way
coupled runs, checked against central finite differences of the full
forward solver along seeded directions, through checkpoints and recomputed
segments."""
import numpy as np
import pytest

import tpx
from helpers import solver_from_case
from tpx import ic

pytestmark = pytest.mark.gpu

FD_REL = 1e-6  # |FD - adjoint| / max(|FD|, |adjoint|, median |adjoint|) per direction


def _case(tau=None, n=16, npart=400, seed=2):
    case = ic.make_case(n, npart, nu=0.02, seed=seed, St=2.0, phi=0.5, heat=False, twc=False, nsteps=6)
    tau0 = case.params.Rhop * case.params.Dp ** 2 / (18 * case.params.Mu0)
    case.tau = np.full(npart, tau0) if tau is None else np.asarray(tau, dtype=float)
    return case, tau0


def _by_flag(s, g):
    """The gradient comes in the solver's particle order (cell order after
    the initial sort); put it back in upload order through the ids."""
    out = np.empty_like(g)
    out[s.particles()["flag"]] = g
    return out


def _J(tau, nsteps, every):
    case, _ = _case(tau)
    s = solver_from_case(case)
    J = s.adjoint_record(nsteps, every)
    s.adjoint_release()
    return J


def test_recording_leaves_the_run_unchanged(gpu_available):
    case, _ = _case()
    a, b = solver_from_case(case), solver_from_case(case)
    a.adjoint_record(5, 2)
    b.step(5)
    for f in ("RU", "Rho", "P"):
        assert np.array_equal(a.get(f), b.get(f)), f
    pa, pb = a.particles(), b.particles()
    for q in ("x", "y", "z", "u", "v", "w", "flag"):
        assert np.array_equal(pa[q], pb[q]), q


@pytest.mark.parametrize("nsteps,every", [(6, 2), (5, 10)])
def test_adjoint_matches_central_differences(gpu_available, nsteps, every):
    case, tau0 = _case()
    s = solver_from_case(case)
    J0 = s.adjoint_record(nsteps, every)
    g = _by_flag(s, s.adjoint_gradient())
    s.adjoint_release()
    assert np.all(np.isfinite(g)) and np.any(g != 0.0)
    rng = np.random.default_rng(11)
    h = 1e-6 * tau0
    dirs = [rng.standard_normal(len(g)) for _ in range(8)]
    ads = [float(g @ d) for d in dirs]
    # the difference quotient carries an absolute floor (~1e-11 here), so a
    # direction whose derivative happens to be small is measured against the
    # median derivative instead of its own
    scale = float(np.median(np.abs(ads)))
    worst = 0.0
    for k, (d, ad) in enumerate(zip(dirs, ads)):
        tau = case.tau
        fd = (_J(tau + h * d, nsteps, every) - _J(tau - h * d, nsteps, every)) / (2 * h)
        rel = abs(fd - ad) / max(abs(fd), abs(ad), scale)
        worst = max(worst, rel)
        print(f"direction {k}: FD {fd:.10e} adjoint {ad:.10e} rel {rel:.2e}")
    print("J", J0, "worst", worst)
    assert worst <= FD_REL


def test_initial_velocity_gradient_matches_central_differences(gpu_available):
    """dJ/du_p(0): the adjoint of the recorded run's initial particle
    velocities against central differences of the full solver."""
    case, _ = _case()
    s = solver_from_case(case)
    s.adjoint_record(4, 2)
    _, gs = s.adjoint_gradient(state=True)
    gu = _by_flag(s, gs[3])
    rng = np.random.default_rng(5)
    h = 1e-6
    dirs = [rng.standard_normal(len(gu)) for _ in range(4)]
    ads = [float(gu @ d) for d in dirs]
    scale = float(np.median(np.abs(ads)))
    for k, (d, ad) in enumerate(zip(dirs, ads)):
        Js = []
        for sign in (1, -1):
            c, _ = _case()
            c.particles["u"] = c.particles["u"] + sign * h * d
            t = solver_from_case(c)
            Js.append(t.adjoint_record(4, 2))
            t.adjoint_release()
        fd = (Js[0] - Js[1]) / (2 * h)
        rel = abs(fd - ad) / max(abs(fd), abs(ad), scale)
        print(f"u direction {k}: FD {fd:.10e} adjoint {ad:.10e} rel {rel:.2e}")
        assert rel <= FD_REL


def test_scalar_tau_gives_the_same_gradient(gpu_available):
    """With one response time for all particles (no per-particle array) the
    per-particle gradient is the same as with an equal per-particle array."""
    case, _ = _case()
    s1 = solver_from_case(case)
    s1.adjoint_record(4, 2)
    g1 = s1.adjoint_gradient()
    case.tau = None
    s2 = solver_from_case(case)
    s2.adjoint_record(4, 2)
    g2 = s2.adjoint_gradient()
    assert np.allclose(g1, g2, rtol=1e-12, atol=0.0)


@pytest.mark.parametrize("nsteps,every", [(5, 2), (4, 4)])
def test_segment_tape_gives_the_kept_gradient(gpu_available, nsteps, every):
    """With the gas checkpointed every `every` steps and each segment
    replayed in the sweep (a short last segment included), the particle-only
    call and the full call give the kept tape's gradients, and the run ends
    where it ended."""
    case, _ = _case()
    out = {}
    for mode in ("keep", "segments"):
        s = solver_from_case(case)
        s.adjoint_set_gas_tape(mode)
        J = s.adjoint_record(nsteps, every)
        g_tau = _by_flag(s, s.adjoint_gradient())
        full = s.adjoint_gradient_full()
        s.adjoint_release()
        s.step(1)
        out[mode] = (J, g_tau, full, s.get("RU"), s.get("P"))
        s.close()
    (Ja, ta, fa, rua, pa), (Jb, tb, fb, rub, pb) = out["keep"], out["segments"]
    assert Ja == Jb
    assert np.array_equal(rua, rub) and np.array_equal(pa, pb)
    # the particle-only call of the kept tape reverses whole steps, the
    # segment mode's the stages between the gas substeps: equal to rounding
    np.testing.assert_allclose(tb, ta, rtol=0, atol=1e-12 * float(np.max(np.abs(ta))))
    np.testing.assert_allclose(fb["tau"], fa["tau"], rtol=0, atol=1e-12 * float(np.max(np.abs(ta))))
    assert abs(fa["A"] - fb["A"]) <= 1e-10 * abs(fa["A"])
    scale = max(float(np.max(np.abs(fa[q]))) for q in ("RU", "Rho", "P"))
    for q in ("RU", "Rho", "P"):
        np.testing.assert_allclose(fb[q], fa[q], rtol=0, atol=1e-10 * scale, err_msg=q)


def test_torch_shim_backward(gpu_available):
    torch = pytest.importorskip("torch")
    from tpx.adjoint import var_c
    case, tau0 = _case()

    def make(tau_np):
        c, _ = _case(tau_np)
        return solver_from_case(c)

    tau = torch.tensor(case.tau, dtype=torch.float64, requires_grad=True)
    J = var_c(tau, make, 4, 2)
    J.backward()
    s = solver_from_case(case)
    s.adjoint_record(4, 2)
    g = _by_flag(s, s.adjoint_gradient())
    assert abs(float(J.detach()) - _J(case.tau, 4, 2)) == 0.0
    assert np.array_equal(tau.grad.numpy(), g)


def test_heat_gradient_needs_a_heat_exchange_run(gpu_available):
    case = ic.make_case(16, 200, nu=0.02, seed=3, St=2.0, heat=False, twc=False)
    s = solver_from_case(case)
    s.adjoint_record(2, 1)
    with pytest.raises(tpx.TpxError, match="no heat exchange"):
        s.adjoint_gradient_heat()
    s.adjoint_release()
