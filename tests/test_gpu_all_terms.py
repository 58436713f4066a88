"""Every term of the step switched on together, on a box of unequal sides,
against the oracle: heat exchange with per-particle absorption efficiency,
two-way coupling with per-particle response times, gravity on the gas and on
the particles, a variable initial density, a mean flow, the passive scalar
(its type-1 source) and the passive vector (its type-0 source), on a
16 x 32 x 16 box (nz, ny, nx) through the three-pass transform.  Three fused
steps, every field and particle to 1e-9 (the threshold the single-term
variants use)."""
import numpy as np
import pytest

from helpers import oracle_from_case, particle_errors, rel_err, solver_from_case
from tpx import ic

pytestmark = pytest.mark.gpu

SHAPE = (16, 32, 16)


def _case(seed=23, npart=3000):
    nz, ny, nx = SHAPE
    rng = np.random.default_rng(seed)
    st = ic.loguniform(np.random.default_rng(seed + 5), 0.1, 10.0, npart)
    case = ic.make_case(nx, npart, nu=0.02, seed=seed, St=st, phi=0.3, heat=True, twc=True,
                        eps_range=(0.5, 1.0), nsteps=20, shape=SHAPE)
    p = case.params
    d = p.Lx / nx
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    X, Y, Z = (i + 0.5) * d, (j + 0.5) * d, (k + 0.5) * d
    # one period along each side
    Yp, Zp = 2 * np.pi * Y / p.Ly, 2 * np.pi * Z / p.Lz
    case.Rho = 1.0 + 0.2 * np.sin(X) * np.cos(2 * Yp) + 0.05 * np.sin(Zp) + 0.02 * rng.standard_normal(SHAPE)
    p.gx1 = p.gx2 = 0.3
    p.gz1 = p.gz2 = -0.5
    p.ParticleGravity = 1
    p.U0 = 0.3
    p.solve_for_scalar, p.D_M, p.s1t = 1, 0.05, 1.0
    case.PS = np.sin(X) * np.sin(Yp) + 0.1 * rng.standard_normal(SHAPE)
    p.solve_for_vector, p.eta, p.Rho_forV, p.s2t = 1, 0.04, 1.3, 0.0
    case.RV = 0.2 * rng.standard_normal((3,) + SHAPE) + np.sin(X)[None]
    case.Q = np.zeros(SHAPE)
    return case


def test_all_terms_on_a_box_of_unequal_sides(gpu_available):
    case = _case()
    p = case.params
    o, s = oracle_from_case(case), solver_from_case(case)
    o.step(3)
    s.step(3)
    ru_o, ru_s = o.get("RU"), s.get("RU")
    e = {f"RU{'xyz'[c]}": rel_err(ru_s[c], ru_o[c]) for c in range(3)}
    for f in ("Rho", "P", "Passive_Scalar", "Q"):
        e[f] = rel_err(s.get(f), o.get(f))
    e["RV"] = max(rel_err(a, b) for a, b in zip(s.get("RV"), o.get("RV")))
    e["P0"] = abs(s.scalars()["P0"] - o.scalars()["P0"]) / abs(o.scalars()["P0"])
    e.update({f"p_{k}": v for k, v in particle_errors(s.particles(), o.particles(), (p.Lx, p.Ly, p.Lz)).items()})
    assert s.scalars()["num_iteration"] == o.scalars()["num_iteration"]
    print({k: f"{v:.2e}" for k, v in e.items()})
    assert max(e.values()) <= 1e-9, e
