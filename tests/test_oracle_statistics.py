"""The oracle's restatement of grid::Statistics (grid.cpp:902-982), pinned by
closed forms: a uniform flow in a uniform density gives U exactly, so the
TKE_* means, the CFL numbers and u_max follow from the inputs; Tg, the
density extrema and the diffusive CFL follow from rho0 and P0; the particle
maxima and means from the uploaded arrays; and the concentration's
partition of unity fixes C_mean = Np / ncell after a step."""
import numpy as np
import pytest

from helpers import oracle_from_case
from tpx import ic


def _uniform_case(n=8, npart=200, U=(0.7, -0.3, 0.2)):
    case = ic.make_case(n, npart, nu=0.02, seed=4, St=1.0, phi=0.2, heat=False, twc=False, nsteps=4)
    p = case.params
    case.RU = np.stack([np.full((n, n, n), p.Rho0 * u) for u in U])
    case.particles["u"] = np.linspace(-1.0, 2.0, npart)
    case.particles["T"] = np.linspace(0.5, 1.5, npart)
    return case


def test_statistics_of_a_uniform_flow():
    U = (0.7, -0.3, 0.2)
    case = _uniform_case(U=U)
    p = case.params
    o = oracle_from_case(case)
    st = o.statistics(p)
    dx = p.Lx / p.Nx_tot
    assert st["Rho_mean"] == pytest.approx(p.Rho0, rel=1e-15)
    assert st["Rho_max"] == p.Rho0 and st["Rho_min"] == p.Rho0
    assert st["Tg_mean"] == pytest.approx(p.P0 / (p.R * p.Rho0), rel=1e-14)
    for c, name in enumerate(("TKE_U", "TKE_V", "TKE_W")):
        assert st[name] == pytest.approx(p.Rho0 * U[c] ** 2, rel=1e-14)
    assert st["u_max"] == pytest.approx(max(U), rel=1e-15)
    assert st["Gas_CFL_max"] == pytest.approx(max(U) * p.dt / dx, rel=1e-14)
    assert st["Gas_diff_CFL_max"] == pytest.approx(p.dt / (dx * dx * p.Rho0 / p.Mu0) * 6, rel=1e-14)
    assert st["Vp_max"] == 2.0 and st["Tp_max"] == 1.5
    assert st["Tp_mean"] == pytest.approx(1.0, rel=1e-14)
    assert st["balance_index"] == pytest.approx(len(case.particles["x"]) / p.N0, rel=1e-15)
    for k in ("Passive_Scalar_mean", "TKEV_V1", "TKEV_V2", "TKEV_V3", "VV", "Particle_CFL_max"):
        assert st[k] == 0.0


def test_concentration_mean_is_the_particle_count_per_cell():
    case = _uniform_case()
    o = oracle_from_case(case)
    o.step(1)
    st = o.statistics(case.params)
    n = case.params.Nx_tot ** 3
    assert st["C_mean"] == pytest.approx(len(case.particles["x"]) / n, rel=1e-13)
    assert st["C_min"] >= 0.0 and st["C_max"] >= st["C_mean"]
