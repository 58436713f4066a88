"""The deposit kernels' tile shapes and their online tuner: any valid shape
gives the same results bit for bit (each cell sums its own particles in bin
order), so tuning inside real steps leaves the run unchanged."""
import numpy as np
import pytest

import tpx
from helpers import by_flag, small_case, solver_from_case

pytestmark = pytest.mark.gpu


def _state(s):
    out = {f: s.get(f) for f in ("RU", "Rho", "P", "RHS_Part_Temp")}
    out.update({f"p_{k}": v for k, v in by_flag(s.particles()).items()})
    return out


def _same(a, b):
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_tuning_leaves_the_run_unchanged(gpu_available):
    case = small_case(n=16, npart=3000, heat=True, twc=True, seed=12)
    a = solver_from_case(case)
    chosen = a.tune_tiles(max_steps=12)
    assert not chosen["tuning"]
    assert chosen["heat"] != (0, 0) and chosen["velocity"] != (0, 0)
    ran = a.scalars()["num_timestep"]
    a.step(2)
    b = solver_from_case(case)
    b.step(int(ran) + 2)
    _same(_state(a), _state(b))


@pytest.mark.parametrize("heat,vel", [((1, 2), (1, 4)), ((2, 16), (4, 2)), ((4, 8), (2, 16))])
def test_every_shape_gives_the_same_step(gpu_available, heat, vel):
    case = small_case(n=16, npart=2000, heat=True, twc=True, seed=13)
    a, b = solver_from_case(case), solver_from_case(case)
    a.set_tiles(heat=heat, velocity=vel)
    a.step(2)
    b.step(2)
    _same(_state(a), _state(b))


def test_invalid_shapes_are_refused(gpu_available):
    case = small_case(n=16, npart=200, seed=14)
    s = solver_from_case(case)
    with pytest.raises(tpx.TpxError, match="tile"):
        s.set_tiles(heat=(3, 4))
    with pytest.raises(tpx.TpxError, match="tile"):
        s.set_tiles(velocity=(1, 5))
