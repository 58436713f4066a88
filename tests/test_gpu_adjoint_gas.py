"""The hand-written adjoint of the gas step: dJ/dA (forcing coefficient) and
dJ/d of the initial RU, Rho and P for J = Var(C) of one-way coupled runs,
checked against central finite differences of the full forward solver along
seeded directions.  The variable-density cases run every Poisson solve to a
fixed iteration count (epsilon = 0), so that the finite differences and the
adjoint differentiate the same map: six iterations, and two, where the
solve's warm start P still matters (after six it has decayed below what a
difference quotient resolves); the uniform-density case keeps the adaptive
stopping rule (its solves stop after two iterations).  The difference steps
for A (1e-4 A) and P (5e-2) are larger than for RU and Rho: J is less
sensitive to them, and a smaller step drowns the quotient in J's rounding."""
import copy

import numpy as np
import pytest

from helpers import by_flag, solver_from_case
from tpx import ic

pytestmark = pytest.mark.gpu

FD_REL = 1e-6  # |FD - adjoint| / max(|FD|, |adjoint|, median |adjoint|) per direction
NSTEPS, EVERY = 4, 2


def _variable_density_case(iterations=6, shape=None):
    """Non-uniform initial density (so the Poisson coefficient, the
    conduction term k/cp del2(1/Rho) and the body force are all active),
    gravity on the gas, fixed Poisson iteration counts; a 16^3 box, or
    shape = (nz, ny, nx) cells of one spacing."""
    case = ic.make_case(16, 400, nu=0.02, seed=2, St=2.0, phi=0.5, heat=False, twc=False, nsteps=6, shape=shape)
    nz, ny, nx = case.Rho.shape
    d = 2 * np.pi / nx
    z, y, x = ((np.arange(m) + 0.5) * d for m in (nz, ny, nx))
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    # one period along each side
    Y, Z = Y * (nx / ny), Z * (nx / nz)
    case.Rho = 1.0 + 0.1 * np.sin(X + 0.3) * np.cos(2 * Y) + 0.05 * np.sin(Z - 0.7)
    case.RU = case.RU * 1.0
    rng = np.random.default_rng(4)
    case.P = 0.05 * rng.standard_normal(case.Rho.shape)
    p = case.params
    p.gx1 = p.gx2 = 0.4
    p.gz1 = p.gz2 = -0.3
    p.epsilon = 0.0
    p.Iteration1 = p.Iteration2 = iterations
    return case


def _smooth(rng, shape, kmax=2):
    """A random field of the Fourier modes |k| <= kmax only."""
    f = np.fft.fftn(rng.standard_normal(shape))
    k = [np.fft.fftfreq(n, 1.0 / n) for n in shape]
    K = np.sqrt(sum(np.meshgrid(*[q * q for q in k], indexing="ij")))
    f[K > kmax] = 0.0
    r = np.real(np.fft.ifftn(f))
    return r / np.std(r)


def _J(case):
    s = solver_from_case(case)
    J = s.adjoint_record(NSTEPS, EVERY)
    s.adjoint_release()
    s.close()
    return J


def _perturbed(case, which, h, d):
    c = copy.deepcopy(case)
    if which == "A":
        c.params.A1 = c.params.A1 + h * d
        c.params.A2 = c.params.A2 + h * d
    else:
        setattr(c, which, getattr(c, which) + h * d)
    return c


def _check(case, g, which, h, ndirs, seed, smooth=False):
    rng = np.random.default_rng(seed)
    if which == "A":
        dirs = [1.0]
        ads = [g["A"]]
    else:
        draw = (lambda: _smooth(rng, g[which].shape)) if smooth else (lambda: rng.standard_normal(g[which].shape))
        dirs = [draw() for _ in range(ndirs)]
        ads = [float(np.sum(g[which] * d)) for d in dirs]
    scale = float(np.median(np.abs(ads)))
    assert scale > 0.0 and np.isfinite(scale), (which, ads)
    worst = 0.0
    for k, (d, ad) in enumerate(zip(dirs, ads)):
        fd = (_J(_perturbed(case, which, h, d)) - _J(_perturbed(case, which, -h, d))) / (2 * h)
        rel = abs(fd - ad) / max(abs(fd), abs(ad), scale)
        worst = max(worst, rel)
        print(f"{which} direction {k}: FD {fd:.10e} adjoint {ad:.10e} rel {rel:.2e}")
    return worst


def _record(case):
    s = solver_from_case(case)
    J = s.adjoint_record(NSTEPS, EVERY)
    g = s.adjoint_gradient_full()
    s.adjoint_release()
    s.close()
    return case, J, g


@pytest.fixture(scope="module")
def six_iterations(gpu_available):
    return _record(_variable_density_case(6))


@pytest.fixture(scope="module")
def two_iterations(gpu_available):
    return _record(_variable_density_case(2))


@pytest.mark.parametrize("which", ["A", "RU", "Rho"])
def test_gas_gradient_matches_central_differences(six_iterations, which):
    case, J, g = six_iterations
    for q in ("RU", "Rho", "P"):
        assert np.all(np.isfinite(g[q])), q
    h = 1e-4 * case.params.A1 if which == "A" else 1e-6
    worst = _check(case, g, which, h, 3, seed=17)
    print("J", J, "worst", worst)
    assert worst <= FD_REL


@pytest.mark.parametrize("which", ["A", "RU", "Rho", "P"])
def test_gas_gradient_two_iteration_solves(two_iterations, which):
    case, J, g = two_iterations
    # dJ/dP0 is small (the solve damps its warm start), so its step is large;
    # P enters the solve affinely, so the quotient stays accurate
    h = {"A": 1e-4 * case.params.A1, "RU": 1e-6, "Rho": 1e-6, "P": 5e-2}[which]
    worst = _check(case, g, which, h, 3, seed=29, smooth=which == "P")
    print("J", J, "worst", worst)
    assert worst <= FD_REL


@pytest.mark.parametrize("shape", [(16, 16, 32), (8, 16, 32)])
def test_gas_gradient_on_a_box_of_unequal_sides(gpu_available, shape):
    """The momentum right-hand side's adjoint marches tiles of 32 x 8 cells
    along z (all of x and more rows when nx < 32); on boxes of unequal sides
    (the three-pass transform at 16 x 16 x 32, the five-pass one where an
    axis is 8 long) dJ/dA, dJ/dRU and dJ/dRho match central differences."""
    case, J, g = _record(_variable_density_case(6, shape))
    for which, h in (("A", 1e-4 * case.params.A1), ("RU", 1e-6), ("Rho", 1e-6)):
        worst = _check(case, g, which, h, 2, seed=31)
        print(shape, which, "J", J, "worst", worst)
        assert worst <= FD_REL


def test_gas_gradient_uniform_density_adaptive_solve(gpu_available):
    case = ic.make_case(16, 400, nu=0.02, seed=5, St=1.0, phi=0.5, heat=False, twc=False, nsteps=6)
    s = solver_from_case(case)
    s.adjoint_record(NSTEPS, EVERY)
    g = s.adjoint_gradient_full()
    s.adjoint_release()
    worst = max(_check(case, g, "A", 1e-4 * case.params.A1, 1, seed=3),
                _check(case, g, "RU", 1e-6, 3, seed=3))
    assert worst <= FD_REL


def test_full_gradient_keeps_the_particle_gradient_and_the_run(gpu_available):
    """The gas sweep (particle stages interleaved with the gas substeps) gives
    dJ/dtau as the particle-only sweep does, to rounding (the stage form adds
    the four stages' terms one at a time), and leaves the solver where the
    recorded run ended: one more step equals a plain run of NSTEPS + 1 steps,
    bit for bit."""
    case = _variable_density_case()
    a = solver_from_case(case)
    a.adjoint_record(NSTEPS, EVERY)
    g_full = a.adjoint_gradient_full()
    g_tau = a.adjoint_gradient()
    np.testing.assert_allclose(g_full["tau"], g_tau, rtol=0, atol=1e-12 * float(np.max(np.abs(g_tau))))
    a.adjoint_release()
    a.step(1)
    b = solver_from_case(case)
    b.step(NSTEPS + 1)
    for f in ("RU", "RU_int", "Rho", "P"):
        assert np.array_equal(a.get(f), b.get(f)), f
    pa, pb = by_flag(a.particles()), by_flag(b.particles())
    for q in ("x", "y", "z", "u", "v", "w"):
        assert np.array_equal(pa[q], pb[q]), q


@pytest.mark.parametrize("iterations", [6, 2])
@pytest.mark.parametrize("other", ["recompute", "segments"])
def test_kept_and_recomputed_tapes_agree(gpu_available, iterations, other):
    """The sweep from a tape that keeps every substep, the sweep that
    recomputes each step from its start and the sweep that replays each
    checkpoint segment's gas from its checkpoint give the same gradients, and
    all leave the run as it ended."""
    case = _variable_density_case(iterations)
    out = {}
    for mode in ("keep", other):
        s = solver_from_case(case)
        s.adjoint_set_gas_tape(mode)
        J = s.adjoint_record(NSTEPS, EVERY)
        out[mode] = (J, s.adjoint_gradient_full())
        s.adjoint_release()
        s.step(1)
        out[mode] += (s.get("RU"), s.get("P"))
        s.close()
    (Ja, ga, rua, pa), (Jb, gb, rub, pb) = out["keep"], out[other]
    assert Ja == Jb
    assert np.array_equal(rua, rub) and np.array_equal(pa, pb)
    assert np.array_equal(ga["tau"], gb["tau"])
    assert abs(ga["A"] - gb["A"]) <= 1e-10 * abs(ga["A"])
    # one scale for the three fields: after six iterations dJ/dP0 is itself
    # rounding (~1e-18), so its own maximum is no measure of agreement
    scale = max(float(np.max(np.abs(ga[q]))) for q in ("RU", "Rho", "P"))
    assert scale > 0.0
    for q in ("RU", "Rho", "P"):
        np.testing.assert_allclose(ga[q], gb[q], rtol=0, atol=1e-10 * scale, err_msg=q)


@pytest.mark.parametrize("other", ["recompute", "segments"])
def test_tapes_agree_bitwise_on_the_five_pass_form(gpu_available, other):
    """On a box with an axis 8 long (the five-pass transform, whose last pass
    the replayed iterations run out of place: P + dp into the next slot of the
    tape) the recomputed and segment-replayed sweeps give the kept tape's
    gradients bit for bit and leave the run where it ended."""
    case = _variable_density_case(6, (8, 16, 32))
    out = {}
    for mode in ("keep", other):
        s = solver_from_case(case)
        s.adjoint_set_gas_tape(mode)
        J = s.adjoint_record(NSTEPS, EVERY)
        g = s.adjoint_gradient_full()
        s.adjoint_release()
        s.step(1)
        out[mode] = (J, g, s.get("RU"), s.get("P"))
        s.close()
    (Ja, ga, rua, pa), (Jb, gb, rub, pb) = out["keep"], out[other]
    assert Ja == Jb
    assert np.array_equal(rua, rub) and np.array_equal(pa, pb)
    assert ga["A"] == gb["A"]
    for q in ("tau", "state", "RU", "Rho", "P"):
        assert np.array_equal(ga[q], gb[q]), q


def test_torch_shim_flow_backward(gpu_available):
    torch = pytest.importorskip("torch")
    from tpx.adjoint import var_c_flow
    case = _variable_density_case()

    def make(ru_np):
        c = copy.deepcopy(case)
        c.RU = ru_np
        return solver_from_case(c)

    ru = torch.tensor(case.RU, dtype=torch.float64, requires_grad=True)
    J = var_c_flow(ru, make, NSTEPS, EVERY)
    J.backward()
    s = solver_from_case(case)
    J_ref = s.adjoint_record(NSTEPS, EVERY)
    g = s.adjoint_gradient_full()["RU"]
    assert float(J.detach()) == J_ref
    # the particles' gather adjoint is deposited with atomics, so two sweeps
    # agree to rounding rather than bit for bit
    np.testing.assert_allclose(ru.grad.numpy(), g, rtol=1e-9, atol=1e-12 * float(np.max(np.abs(g))))
