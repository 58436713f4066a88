"""The ensemble entry points (independent boxes stepped concurrently, or by
batched kernels): every box must equal its own plain run bit for bit."""
import numpy as np
import pytest

import tpx
from helpers import solver_from_case
from tpx import ic

pytestmark = pytest.mark.gpu


def test_ensemble_boxes_equal_their_own_runs(gpu_available):
    cases = ic.ensemble_cases(n_boxes=5, n=16, seed=3, np_range=(200, 3000))
    assert len({len(c.particles["x"]) for c in cases}) > 1
    ens = [solver_from_case(c) for c in cases]
    solo = [solver_from_case(c) for c in cases]
    tpx.ensemble_step(ens, 3, concurrency=3)
    for s in solo:
        s.step(3)
    for b, (a, s) in enumerate(zip(ens, solo)):
        for f in ("RU", "Rho", "P"):
            assert np.array_equal(a.get(f), s.get(f)), (b, f)
        pa, ps = a.particles(), s.particles()
        for q in ("x", "u", "T", "flag"):
            assert np.array_equal(pa[q], ps[q]), (b, q)
        assert a.scalars()["num_timestep"] == 3


def test_ensemble_rejects_repeated_contexts(gpu_available):
    cases = ic.ensemble_cases(n_boxes=2, n=8, seed=4, np_range=(100, 200))
    s = [solver_from_case(c) for c in cases]
    with pytest.raises(tpx.TpxError):
        tpx.ensemble_step([s[0], s[1], s[0]], 1)


def test_tuner_returns_a_candidate(gpu_available):
    cases = ic.ensemble_cases(n_boxes=4, n=16, seed=5, np_range=(200, 1000))
    s = [solver_from_case(c) for c in cases]
    c = tpx.tune_concurrency(s, candidates=(1, 2, 4), steps=1)
    assert c in (1, 2, 4)


@pytest.mark.parametrize("graphs", ["1", "0"])
@pytest.mark.parametrize("heat", [False, True])
def test_batched_boxes_equal_their_own_runs(gpu_available, heat, graphs, monkeypatch):
    """One launch per kernel for all boxes (box in the grid's third dimension,
    arguments from a table): boxes with different particle counts, viscosities,
    Stokes numbers and time steps each equal their own solo run bit for bit,
    Poisson iteration counts included.  With graphs "1" (the default) the
    segments between the solves are replayed from captured graphs (each
    captured once, replayed at every later step, with the host's own work and
    waits between the replays); with "0" they are launched kernel by kernel."""
    monkeypatch.setenv("TPX_BATCH_GRAPH", graphs)
    cases = ic.ensemble_cases(n_boxes=5, n=16, seed=7, np_range=(200, 3000))
    if heat:
        cases = [ic.make_case(16, npart, nu=0.01 + 0.004 * b, seed=20 + b, St=0.5 + b, heat=True, nsteps=4)
                 for b, npart in enumerate((300, 1200, 700, 2500))]
    bat = [solver_from_case(c) for c in cases]
    solo = [solver_from_case(c) for c in cases]
    batch = tpx.Batch(bat)
    batch.step(2)
    batch.step(2)
    batch.close()
    for s in solo:
        s.step(4)
    for b, (a, s) in enumerate(zip(bat, solo)):
        for f in ("RU", "Rho", "P"):
            assert np.array_equal(a.get(f), s.get(f)), (b, f)
        pa, ps = a.particles(), s.particles()
        for q in ("x", "y", "z", "u", "v", "w", "T", "flag"):
            assert np.array_equal(pa[q], ps[q]), (b, q)
        sa, ss = a.scalars(), s.scalars()
        assert sa["num_timestep"] == ss["num_timestep"] == 4
        assert sa["total_iterations"] == ss["total_iterations"], b


def test_batch_rejects_mismatched_boxes(gpu_available):
    a = ic.make_case(16, 200, nu=0.01, seed=1, heat=False)
    b = ic.make_case(16, 200, nu=0.01, seed=2, heat=True)
    with pytest.raises(tpx.TpxError, match="heat"):
        tpx.Batch([solver_from_case(a), solver_from_case(b)])
