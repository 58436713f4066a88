"""Transpose (dot-product) checks of the Poisson solve's adjoint kernels, one
iteration at a time, as the reverse sweep runs them (tpx_adjoint_check).

The sweep reverses each forward iteration P_1 = P + M(cc (div(C grad P) - RHS))
(poisson.cpp:391-462, C = 1/I_C2F(Rho_new), cc = -1/I_F2C(C)) by applying the
spectral step M to lam_P as its own transpose and then the transposes of the
stencil, of cc and of C.  Checked here:
  * M is symmetric: <M x, y> = <x, M y> for every transform form (the
    three-pass form, with the packed real x-rows at x length 128, and the
    five-pass form where an axis is 8 long);
  * the iteration's map of P (linear at a zero right-hand side) against its
    reversal: <A x, y> = <x, A^T y>;
  * the reversal's adjoint of Rho against central differences of <A x, y>
    in Rho along seeded directions (the map is smooth in Rho).
"""
import numpy as np
import pytest

from helpers import solver_from_case
from tpx import ic
from tpx.params import Params

pytestmark = pytest.mark.gpu


def _case(shape, seed):
    """Gas-only case on an nz x ny x nx box with Rho in [0.9, 1.1]."""
    nz, ny, nx = shape
    p = Params(Nx_tot=nx, Ny_tot=ny, Nz_tot=nz, Lx=2 * np.pi, Ly=2 * np.pi * ny / nx, Lz=2 * np.pi * nz / nx,
               bs=2, dt=2e-3, Mu0=0.02, A1=0.1, A2=0.1, TWC=0, Nu=0.0, np0=0.0, Iteration1=50, Iteration2=50,
               epsilon=1e-10)
    rng = np.random.default_rng(seed)
    Rho = 1.0 + 0.2 * (rng.random(shape) - 0.5)
    empty = {q: np.zeros(0) for q in "xyzuvwT"}
    empty["flag"] = np.zeros(0, dtype=np.int32)
    return ic.Case(p, np.zeros((3,) + shape), Rho, np.zeros(shape), empty)


SHAPES = [(16, 16, 16), (16, 32, 128), (32, 16, 64), (16, 8, 16)]


@pytest.mark.parametrize("shape", SHAPES)
def test_spectral_step_is_symmetric(gpu_available, shape):
    s = solver_from_case(_case(shape, 3))
    rng = np.random.default_rng(5)
    x, y = rng.standard_normal(shape), rng.standard_normal(shape)
    mx, my = s.adjoint_check(0, x, y)
    a, b = float(np.vdot(mx, y)), float(np.vdot(x, my))
    scale = float(np.linalg.norm(mx) * np.linalg.norm(y))
    print(shape, a, b, abs(a - b) / scale)
    assert abs(a) > 1e-6 * scale  # not a trivially small product
    assert abs(a - b) <= 1e-13 * scale
    s.close()


@pytest.mark.parametrize("shape", SHAPES)
def test_iteration_reversal_is_the_transpose_in_P(gpu_available, shape):
    s = solver_from_case(_case(shape, 7))
    rng = np.random.default_rng(9)
    x, y = rng.standard_normal(shape), rng.standard_normal(shape)
    ax, aty, _ = s.adjoint_check(1, x, y)
    # linear in x at a zero right-hand side
    ax2, _, _ = s.adjoint_check(1, 2.0 * x, y)
    np.testing.assert_allclose(ax2, 2.0 * ax, rtol=0, atol=1e-13 * float(np.max(np.abs(ax))))
    a, b = float(np.vdot(ax, y)), float(np.vdot(x, aty))
    scale = float(np.linalg.norm(ax) * np.linalg.norm(y))
    print(shape, a, b, abs(a - b) / scale)
    assert abs(a - b) <= 1e-12 * scale
    s.close()


@pytest.mark.parametrize("shape", [(16, 16, 16), (16, 32, 128)])
def test_iteration_reversal_gives_the_density_derivative(gpu_available, shape):
    case = _case(shape, 11)
    s = solver_from_case(case)
    rng = np.random.default_rng(13)
    x, y = rng.standard_normal(shape), rng.standard_normal(shape)
    _, _, atr = s.adjoint_check(1, x, y)
    h = 1e-6
    dirs = [rng.standard_normal(shape) for _ in range(6)]
    ads = [float(np.vdot(atr, d)) for d in dirs]
    # the difference quotient's rounding floor (~|<Ax, y>| 1e-16 / h) is
    # absolute: a direction with a small derivative is measured against the
    # median one
    scale = float(np.median(np.abs(ads)))
    worst = 0.0
    for k, (d, ad) in enumerate(zip(dirs, ads)):
        f = []
        for sign in (1, -1):
            s.set("Rho", case.Rho + sign * h * d)
            ax, _, _ = s.adjoint_check(1, x, y)
            f.append(float(np.vdot(ax, y)))
        fd = (f[0] - f[1]) / (2 * h)
        rel = abs(fd - ad) / max(abs(fd), abs(ad), scale)
        worst = max(worst, rel)
        print(f"{shape} direction {k}: FD {fd:.12e} adjoint {ad:.12e} rel {rel:.2e}")
    s.set("Rho", case.Rho)
    assert worst <= 1e-7
    s.close()
