"""The adjoint with every term it reverses switched on together, on a box of
unequal sides: two-way coupling with per-particle response times, the heat
exchange with per-particle absorption efficiency (P0 evolving), gravity on
the gas and on the particles, a variable initial density and a mean flow, on
a 16 x 32 x 16 box (nz, ny, nx).  The gradients of the segment-replay sweep
(the gas checkpointed every 2 steps) — dJ/dtau_p, dJ/deps_p, dJ/dA, dJ/dRU,
dJ/dRho — against central finite differences of the full forward solver
along seeded directions."""
import copy

import numpy as np
import pytest

from helpers import solver_from_case
from tpx import ic

pytestmark = pytest.mark.gpu

FD_REL = 1e-6  # |FD - adjoint| / max(|FD|, |adjoint|, median |adjoint|) per direction
NSTEPS, EVERY = 4, 2
SHAPE = (16, 32, 16)


def _case(seed=29, npart=1500):
    nz, ny, nx = SHAPE
    rng = np.random.default_rng(seed)
    st = ic.loguniform(np.random.default_rng(seed + 5), 0.5, 4.0, npart)
    case = ic.make_case(nx, npart, nu=0.02, seed=seed, St=st, phi=0.5, heat=True, twc=True,
                        eps_range=(0.5, 1.0), nsteps=NSTEPS, shape=SHAPE)
    p = case.params
    d = p.Lx / nx
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    X, Y, Z = (i + 0.5) * d, (j + 0.5) * d, (k + 0.5) * d
    case.Rho = 1.0 + 0.1 * np.sin(X + 0.3) * np.cos(4 * np.pi * Y / p.Ly) + 0.05 * np.sin(2 * np.pi * Z / p.Lz)
    p.gx1 = p.gx2 = 0.3
    p.gz1 = p.gz2 = -0.4
    p.ParticleGravity = 1
    p.U0 = 0.2
    return case


def _J(case):
    s = solver_from_case(case)
    J = s.adjoint_record(NSTEPS, EVERY)
    s.adjoint_release()
    s.close()
    return J


@pytest.fixture(scope="module")
def swept(gpu_available):
    case = _case()
    s = solver_from_case(case)
    s.adjoint_set_gas_tape("segments")
    J = s.adjoint_record(NSTEPS, EVERY)
    g = s.adjoint_gradient_full()
    flag = s.particles()["flag"]
    s.adjoint_release()
    s.close()
    for q in ("tau", "eps"):  # particle gradients back in upload order
        v = np.empty_like(g[q])
        v[flag] = g[q]
        g[q] = v
    return case, J, g


def _perturbed(case, which, h, d):
    c = copy.deepcopy(case)
    if which == "A":
        c.params.A1 = c.params.A1 + h * d
        c.params.A2 = c.params.A2 + h * d
    else:
        setattr(c, which, getattr(c, which) + h * d)
    return c


@pytest.mark.parametrize("which", ["tau", "eps", "A", "RU", "Rho"])
def test_all_terms_gradient_matches_central_differences(swept, which):
    case, J, g = swept
    rng = np.random.default_rng(37)
    if which == "A":
        dirs, ads = [1.0], [g["A"]]
    else:
        shape = np.shape(getattr(case, which))
        dirs = [rng.standard_normal(shape) for _ in range(2)]
        ads = [float(np.sum(g[which] * d)) for d in dirs]
    h = {"tau": 1e-6 * float(np.mean(case.tau)), "eps": 1e-6, "A": 1e-4 * case.params.A1,
         "RU": 1e-6, "Rho": 1e-6}[which]
    scale = float(np.median(np.abs(ads)))
    assert scale > 0.0 and np.isfinite(scale), (which, ads)
    worst = 0.0
    for k, (d, ad) in enumerate(zip(dirs, ads)):
        fd = (_J(_perturbed(case, which, h, d)) - _J(_perturbed(case, which, -h, d))) / (2 * h)
        rel = abs(fd - ad) / max(abs(fd), abs(ad), scale)
        worst = max(worst, rel)
        print(f"{which} direction {k}: FD {fd:.10e} adjoint {ad:.10e} rel {rel:.2e}")
    print("J", J, "worst", worst)
    assert worst <= FD_REL
