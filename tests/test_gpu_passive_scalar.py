"""The passive scalar (grid::Update_Passive_Scalar, grid.cpp:686-745) on the
device against the oracle: one and twenty steps of heated, two-way coupled
runs with the scalar on (both source forms, molecular diffusion, a
non-cubic box), the method-by-method path against the fused step, the gas
untouched by it, its statistic, and the batch's refusal."""
import numpy as np
import pytest

import tpx
from helpers import oracle_from_case, rel_err, solver_from_case
from tpx import ic

pytestmark = pytest.mark.gpu

TOL_1, TOL_20 = 1e-10, 1e-8


def _case(s1t, shape=(16, 16, 16), seed=3, npart=2000):
    nz, ny, nx = shape
    case = ic.make_case(nx, npart, nu=0.02, seed=seed, St=1.0, phi=0.3, heat=True, twc=True, nsteps=20)
    p = case.params
    if shape != (nx, nx, nx):  # trim the cubic case to the box
        p.Ny_tot, p.Nz_tot = ny, nz
        p.Ly, p.Lz = p.Lx * ny / nx, p.Lx * nz / nx
        case.RU, case.Rho, case.P = case.RU[:, :nz, :ny, :], case.Rho[:nz, :ny, :], case.P[:nz, :ny, :]
        pp = case.particles
        keep = (pp["y"] < p.Ly) & (pp["z"] < p.Lz)
        case.particles = {k: v[keep] for k, v in pp.items()}
        case.particles["flag"] = np.arange(int(keep.sum()), dtype=np.int32)
        if case.tau is not None:
            case.tau = case.tau[keep]
        if case.eps is not None:
            case.eps = case.eps[keep]
    p.solve_for_scalar, p.D_M, p.s1t = 1, 0.05, float(s1t)
    dx = p.Lx / nx
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    X, Y = (i + 0.5) * dx, (j + 0.5) * dx
    rng = np.random.default_rng(seed)
    case.PS = np.sin(X) * np.sin(2 * np.pi * Y / p.Ly) + 0.1 * rng.standard_normal(shape)
    return case


@pytest.mark.parametrize("s1t", [0, 1])
def test_passive_scalar_matches_the_oracle(gpu_available, s1t):
    case = _case(s1t)
    o, s = oracle_from_case(case), solver_from_case(case)
    o.step(1)
    s.step(1)
    e1 = rel_err(s.get("Passive_Scalar"), o.get("Passive_Scalar"))
    o.step(19)
    s.step(19)
    e20 = rel_err(s.get("Passive_Scalar"), o.get("Passive_Scalar"))
    eru = rel_err(s.get("RU"), o.get("RU"))
    print(f"S1_type {s1t}: scalar 1 step {e1:.2e}, 20 steps {e20:.2e}; RU 20 steps {eru:.2e}")
    assert e1 <= TOL_1 and e20 <= TOL_20 and eru <= TOL_20
    s.close()
    o.close()


def test_passive_scalar_non_cubic_box(gpu_available):
    case = _case(0, shape=(8, 16, 32), seed=9, npart=1500)
    o, s = oracle_from_case(case), solver_from_case(case)
    o.step(2)
    s.step(2)
    e = rel_err(s.get("Passive_Scalar"), o.get("Passive_Scalar"))
    print("non-cubic scalar", e)
    assert e <= TOL_1


def test_method_path_equals_fused_step_and_gas_is_untouched(gpu_available):
    case = _case(1)
    a, b = solver_from_case(case), solver_from_case(case)
    a.step(2)
    b.method_step()
    b.method_step()
    assert np.array_equal(a.get("Passive_Scalar"), b.get("Passive_Scalar"))
    off = _case(1)
    off.params.solve_for_scalar = 0
    c = solver_from_case(off)
    c.step(2)
    for f in ("RU", "Rho", "P"):
        assert np.array_equal(a.get(f), c.get(f)), f


def test_passive_scalar_statistic(gpu_available):
    case = _case(0)
    o, s = oracle_from_case(case), solver_from_case(case)
    s.statistics_enable(True)
    o.step(3)
    s.step(3)
    got = s.statistics()["Passive_Scalar_mean"]
    want = o.statistics(case.params)["Passive_Scalar_mean"]
    assert got > 0 and abs(got - want) <= 1e-12 * want


def test_batch_refuses_the_passive_scalar(gpu_available):
    boxes = [solver_from_case(_case(0, seed=q)) for q in range(2)]
    with pytest.raises(tpx.TpxError, match="passive scalar"):
        tpx.Batch(boxes)


@pytest.mark.parametrize("mode", ["keep", "recompute"])
def test_adjoint_sweep_leaves_the_scalar_where_the_run_ended(gpu_available, mode):
    """The recording advances the scalar with the step; the gas sweep (which
    recomputes the gas only) leaves it alone: one more step equals a plain run
    of three steps, bit for bit."""
    case = _case(0)
    a = solver_from_case(case)
    a.adjoint_set_gas_tape(mode)
    a.adjoint_record(2, 1)
    a.adjoint_gradient_full()
    a.adjoint_release()
    a.step(1)
    b = solver_from_case(case)
    b.step(3)
    assert np.array_equal(a.get("Passive_Scalar"), b.get("Passive_Scalar"))
    assert np.array_equal(a.get("RU"), b.get("RU"))
