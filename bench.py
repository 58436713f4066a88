"""Benchmark of the RK4 time step on one MI355X.

Workload (the configuration the headline metric is quoted on): 256^3 FP64
staggered grid, 1.6e7 two-way-coupled particles with Stokes number 1 and mass
loading 1.0, low-Mach heating (Nusselt exchange, absorption efficiency drawn
uniformly in [0.5, 1]), gamma 1.4, Pr 0.7, seeded spectral initial condition
(k0 = 4, u' = 1), dt from CFL 0.3.  A step is one full RK4 step: 4 substeps,
each with the particle path, the momentum/density update and the FFT Poisson
projection, followed by make_mean_U0 and TimeAdvance.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]

Prints one JSON line.  `value` is steps/s with all state resident in HBM.
`roofline` reports the dominant kernel (the one with the most device time
per step among the particle advance, the momentum right-hand side and the
three Poisson-iteration passes) with its algorithmic bytes per launch over
its event-timed launch duration; `cpu_baseline` times the oracle (the CPU restatement of the
reference step) on a bounded sample of the same workload.
"""
import argparse
import json

import numpy as np
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "triplyperiodic-modified_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "RK4 steps/s at 256^3+16M particles"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def grid_rhs_bytes_per_launch(ncell, rk, has_s=True):
    """Algorithmic bytes of one density + momentum RHS launch (k_rhs_ru).
    Per cell it reads RU_int(3), Rho_int, RU(3), RU_np1(3), Rho, Rho_np1, S(3)
    and writes RU_new(3), RU_np1(3), Rho_new, Rho_np1 (23 FP64 values); at
    the first substep RU_np1/Rho_np1 are the base state (read once), at the
    last one RU_new/Rho_new are RU_np1/Rho_np1 and RU/Rho are not read."""
    s = 3 if has_s else 0
    if rk == 0:
        n = 3 + 1 + 3 + 1 + s + 8
    elif rk == 3:
        n = 3 + 1 + 3 + 1 + s + 4
    else:
        n = 3 + 1 + 3 + 3 + 1 + 1 + s + 8
    return 8.0 * ncell * n


def advance_bytes_per_launch(ncell, npart, rk, per_eps, per_tau, twc, counted):
    """Algorithmic bytes of one particle advance launch (k_advance).  Per
    particle it reads the int stage (7 FP64), the base stage (7; x, y, z only
    at the last substep), the np1 stage (7; it is the base stage at the first
    substep), RHS_T and the per-particle eps / tau, and writes the np1 and new
    stages (14), the 6 drag shares with their sorted position (4 B read), the
    cell key and the counting slot (4 + 4 B, plus the 8 B atomic update of the
    cell count).  The gas it gathers is RU_int (3 fields) and Rho_int, read
    once per cell."""
    rd = 7 + (3 if rk == 3 else 7) + (0 if rk == 0 else 7) + 1 + per_eps + per_tau
    wr = 14
    b = 8.0 * (rd + wr)
    if twc:
        b += 6 * 8 + 4
    b += 4 + (16 if counted else 0)
    return npart * b + 8.0 * ncell * 4


def step_bytes(ncell, npart, n_it):
    """Algorithmic bytes of one RK4 step: 4 substeps of 46 grid field passes,
    26 per Poisson iteration plus 4 per extra iteration, and 65 FP64 per
    particle (state reads and writes, face/corner gathers and deposits)."""
    per_sub = 8.0 * ncell * (46 + 26 * n_it + 4 * (n_it - 1)) + 8.0 * npart * 65
    return 4 * per_sub


def step_kernel_bytes(ncell, npart, its_per_step, per_eps, per_tau, twc, heat, counted):
    """Algorithmic bytes of one RK4 step as the library's kernels move them:
    per substep the particle advance, the fused density/momentum update, the
    divergence and Poisson right-hand side (6 FP64 a cell), the projection
    (13 FP64 a cell; 7 at the last substep, where RU_new is RU_np1), the
    deposits (heat: 44 B a particle + 36 B a cell, then 32 B a cell for the
    y shifts; drag: 48 B a particle + 28 B a cell) and the binning (16 B a
    particle for the placement, 12 B a cell + 12 B a particle for the
    per-cell sort); per Poisson iteration its three passes (P, Rho_new, RHS
    and the half spectrum; the spectrum twice; the spectrum, P read and
    written); per step the re-sort of the particle arrays (140 B a
    particle)."""
    b = 0.0
    for rk in range(4):
        b += advance_bytes_per_launch(ncell, npart, rk, per_eps, per_tau, twc, counted) if npart else 0.0
        b += grid_rhs_bytes_per_launch(ncell, rk, twc and npart > 0)
        b += 8.0 * ncell * (6 + (7 if rk == 3 else 13))
        if heat and npart:
            b += 44.0 * npart + 36.0 * ncell + 32.0 * ncell
        if twc and npart:
            b += 48.0 * npart + 28.0 * ncell
        if counted:
            b += 16.0 * npart + 12.0 * ncell + 12.0 * npart
    spec = (ncell ** (1.0 / 3.0) // 2 + 1) * ncell ** (2.0 / 3.0) * 16.0
    b += its_per_step * ((3 * 8.0 * ncell + spec) + 2 * spec + (spec + 2 * 8.0 * ncell))
    if counted:
        b += 140.0 * npart
    return b


def load_traffic(workload):
    """Per-kernel HBM bytes per dispatch measured with rocprofv3 FETCH_SIZE /
    WRITE_SIZE (scaled by a plain copy of known size) for this workload."""
    path = os.path.join(ROOT, "profiles", f"rk4_step_{workload}_traffic.json")
    if not os.path.exists(path):
        return {}
    with open(path) as fh:
        kernels = json.load(fh).get("kernels", {})
    # by kernel name without template arguments: the dispatch-weighted mean
    # over its instances (e.g. the momentum kernel's rk 0-2 and rk 3 forms)
    out = {}
    for name, e in kernels.items():
        base = name.split("<")[0]
        a = out.setdefault(base, {"dispatches": 0, "traffic_bytes": 0.0})
        a["traffic_bytes"] += e["traffic_bytes"] * e["dispatches"]
        a["dispatches"] += e["dispatches"]
    for a in out.values():
        a["traffic_bytes"] /= max(a["dispatches"], 1)
    return out


def cpu_baseline(case, threads, full=False):
    """The oracle (the CPU restatement of the reference step) on this host's
    cores, on the same workload from the same initial state, initialisation
    excluded.  All threads (OpenMP over z-slabs, standing in for the
    reference's MPI ranks): one warm-up step, then one timed step (three with
    full=True); its Poisson iterations per step are reported beside the
    device's.  One thread (the reference's per-rank code path): the next
    substep (a whole step with full=True), timed alone."""
    from oracle.oracle import Oracle, threads as set_threads
    n_all = set_threads(threads)
    p = case.params
    o = Oracle(p.struct_fields(), np_capacity=len(case.particles["x"]))
    o.set("RU", case.RU)
    o.set("Rho", case.Rho)
    o.set("P", case.P)
    pp = case.particles
    o.set_particles(pp["x"], pp["y"], pp["z"], pp["u"], pp["v"], pp["w"], pp["T"], pp["flag"],
                    tau=case.tau, eps=case.eps)
    o.set_restart_numbers(p.P0, 0.0, 0)
    o.initialize()
    t0 = time.perf_counter()
    o.step(1)
    t_warm = time.perf_counter() - t0
    log(f"[bench] CPU warm-up step {t_warm:.2f} s on {n_all} threads")
    k = 3 if full else 1
    it0 = o.scalars()["total_iterations"]
    t0 = time.perf_counter()
    o.step(k)
    t_all = (time.perf_counter() - t0) / k
    its_all = (o.scalars()["total_iterations"] - it0) / k
    log(f"[bench] CPU {t_all:.2f} s a step on {n_all} threads ({its_all:g} Poisson iterations a step)")
    set_threads(1)
    it0 = o.scalars()["total_iterations"]
    t0 = time.perf_counter()
    if full:
        o.step(1)
        what, scale = "one whole step", 1.0
    else:
        o.params_update()
        o.set_rk4_count(0)
        for m in o.SUBSTEP_METHODS:
            getattr(o, m)()
        what, scale = "the first substep of the next step, value = 1/(4 x its time)", 4.0
    t_one = time.perf_counter() - t0
    its_one = o.scalars()["total_iterations"] - it0
    log(f"[bench] CPU {t_one:.2f} s for {what.split(',')[0]} on 1 thread ({its_one:g} Poisson iterations)")
    set_threads(n_all)
    o.close()
    return {"value": 1.0 / t_all, "unit": "steps/s", "cores": n_all, "kind": "port",
            "sample": f"oracle (C restatement of the reference step, OpenMP over {n_all} host threads) on the "
                      f"same workload from the same initial state: 1 warm-up step ({t_warm:.1f} s), then {k} timed "
                      f"whole step(s) ({t_all:.2f} s a step, {its_all:g} Poisson iterations a step)",
            "poisson_iterations_per_step": its_all, "warmup_steps": 1, "timed_steps": k,
            "single_thread": {"value": 1.0 / (scale * t_one), "unit": "steps/s", "cores": 1,
                              "sample": f"the same oracle on 1 thread after those steps: {what} "
                                        f"({t_one:.2f} s, {its_one:g} Poisson iterations)"}}


def copy_probe(mib):
    """Streaming-bandwidth probe (libtpx_probe.so, a hand-written HIP kernel):
    the best of five copies, reads and writes of `mib` MiB with 16-byte
    accesses.  The copy is the practical ceiling of a streaming kernel on this
    device and, its bytes being known, the calibration kernel of the
    FETCH_SIZE / WRITE_SIZE counters the traffic figures are scaled by."""
    import ctypes
    lib = ctypes.CDLL(os.path.join(ROOT, "triplyperiodic-modified_amd", "libtpx_probe.so"))
    lib.tpx_probe_bandwidth.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    out = (ctypes.c_double * 3)()
    nbytes = mib * (1 << 20)
    rc = lib.tpx_probe_bandwidth(nbytes, 5, 0, out)
    if rc != 0:
        raise RuntimeError(f"bandwidth probe failed ({rc})")
    print(json.dumps({"copy_probe_bytes_read": nbytes, "copy_probe_bytes_written": nbytes, "launches": 5,
                      "copy_GBps": round(out[0], 1), "read_GBps": round(out[1], 1), "write_GBps": round(out[2], 1),
                      "copy_frac_of_8TBps": round(out[0] / HBM_PEAK_GBS, 4)}))


def ensemble_measure(steps, warmup, with_streams=True):
    """Ensemble throughput: 64 independent 64^3 boxes (nu ~ U[0.01, 0.03],
    St ~ logU[0.1, 10], particles ~ logU[1e4, 2e5] per box, mass loading 0.3)
    stepped by batched kernels (one launch per kernel for all boxes), against
    the same boxes stepped one after another (each box's own tpx_step); the
    concurrent-streams form (four boxes in flight, each on its own stream) is
    reported beside it.  Returns the figures as a dict."""
    import tpx
    from tpx import ic
    t0 = time.perf_counter()
    cases = ic.ensemble_cases(64, 64, seed=0)
    solvers = []
    for c in cases:
        s = tpx.Solver(c.params, np_capacity=len(c.particles["x"]))
        s.load_case(c)
        solvers.append(s)
    ncells = sum(c.params.Nx_tot * c.params.Ny_tot * c.params.Nz_tot for c in cases)
    npart = sum(len(c.particles["x"]) for c in cases)
    log(f"[bench] ensemble of {len(cases)} boxes, {npart} particles, built in {time.perf_counter() - t0:.1f} s")
    # tile shapes tuned on the box with the median particle count, given to all
    counts = [len(c.particles["x"]) for c in cases]
    ref = int(np.argsort(counts)[len(counts) // 2])
    tiles = solvers[ref].tune_tiles(max_steps=12)
    for s in solvers:
        s.set_tiles(heat=tiles["heat"], velocity=tiles["velocity"])
    log(f"[bench] tile shapes {tiles} tuned on box {ref}")
    batch = tpx.Batch(solvers)
    t0 = time.perf_counter()
    batch.step(warmup)
    log(f"[bench] warmup {warmup} batched steps in {time.perf_counter() - t0:.2f} s")
    for s in solvers:
        s.synchronize()
    t0 = time.perf_counter()
    for b, s in enumerate(solvers):
        s.step(steps)
        s.synchronize()
    t_seq = time.perf_counter() - t0
    t0 = time.perf_counter()
    batch.step(steps)
    for s in solvers:
        s.synchronize()
    t_bat = time.perf_counter() - t0
    t_str = t_ng = None
    if with_streams:
        t0 = time.perf_counter()
        tpx.ensemble_step(solvers, steps, 4)
        t_str = time.perf_counter() - t0
        # the same batch with its segments launched kernel by kernel instead
        # of replayed from captured graphs
        os.environ["TPX_BATCH_GRAPH"] = "0"
        try:
            b2 = tpx.Batch(solvers)
        finally:
            del os.environ["TPX_BATCH_GRAPH"]
        b2.step(1)
        for s in solvers:
            s.synchronize()
        t0 = time.perf_counter()
        b2.step(steps)
        for s in solvers:
            s.synchronize()
        t_ng = time.perf_counter() - t0
        b2.close()
    batch.close()
    for s in solvers:
        s.close()
    value = ncells * steps / t_bat
    seq = ncells * steps / t_seq
    log(f"[bench] ensemble: sequential {t_seq:.3f} s, batched {t_bat:.3f} s"
        + (f", four streams {t_str:.3f} s" if t_str else "") + f" -> {value:.4g} cells/s ({value / seq:.2f}x)")
    out = {"metric": "ensemble cells/s (64 boxes of 64^3)", "value": round(value, 1), "unit": "cells/s",
           "steps": steps, "warmup": warmup, "ms_per_step": round(1000.0 * t_bat / steps, 3),
           "boxes": len(cases), "particles": npart,
           "workload": "ensemble: 64 x 64^3 boxes, particles ~ logU[1e4, 2e5], nu ~ U[0.01, 0.03], "
                       "St ~ logU[0.1, 10], mass loading 0.3, heat off; batched kernels, deposit tile shapes "
                       "from the online tuner; sequential = the same boxes stepped one after another",
           "deposit_tiles": {"heat": list(tiles["heat"]), "velocity": list(tiles["velocity"])},
           "sequential_cells_per_s": round(seq, 1), "speedup_vs_sequential": round(value / seq, 3)}
    if t_str:
        out["four_streams_cells_per_s"] = round(ncells * steps / t_str, 1)
    if t_ng:
        out["batched_without_graphs_cells_per_s"] = round(ncells * steps / t_ng, 1)
        log(f"[bench] ensemble batched without graphs {t_ng:.3f} s")
    return out


def ensemble_bench(args):
    e = ensemble_measure(args.steps, args.warmup)
    print(json.dumps({
        "metric": e["metric"], "value": e["value"], "unit": "cells/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": e["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": e["workload"], "boxes": e["boxes"], "particles": e["particles"],
                   "deposit_tiles": e["deposit_tiles"]},
        "sequential_cells_per_s": e["sequential_cells_per_s"], "speedup_vs_sequential": e["speedup_vs_sequential"],
        "four_streams_cells_per_s": e.get("four_streams_cells_per_s"),
        "roofline": None, "cpu_baseline": None}), flush=True)


def adjoint_measure():
    """Adjoint cost: 128^3, 2e5 particles, St = 2, two-way coupling at mass
    loading 0.5 and the heat exchange on, 50 steps checkpointed every 10 (gas
    and particles) and recomputed in between.  The whole reverse sweep (per
    checkpoint segment, backwards: the segment's gas replayed from its
    checkpoint, the particles' stage states recomputed; per step and substep
    the gas reversed through the projection, the Poisson iterations and the
    divergence, the heat exchange's gather and deposit, the RK bookkeeping,
    the particles' transposed stage, the momentum and density updates and
    Update_P0) is timed against the plain forward steps.  The same sweep
    with the gas tape kept for every substep, and with the gas recomputed
    from every step's start, is timed beside it.  Returns a dict."""
    import tpx
    from tpx import ic
    nsteps, every = 50, 10
    case = ic.make_case(128, 200_000, nu=0.01, seed=0, St=2.0, phi=0.5, heat=True, twc=True, nsteps=nsteps,
                        eps_range=(0.5, 1.0))
    # the same configuration without the heat exchange (uniform temperature,
    # so the Poisson solves take one or two iterations a substep)
    case_cold = ic.make_case(128, 200_000, nu=0.01, seed=0, St=2.0, phi=0.5, heat=False, twc=True, nsteps=nsteps)

    def forward(case):
        plain = tpx.Solver(case.params, np_capacity=200_000)
        plain.load_case(case)
        plain.step(1)
        plain.synchronize()
        t0 = time.perf_counter()
        plain.step(nsteps)
        plain.synchronize()
        t = time.perf_counter() - t0
        plain.close()
        return t

    t_fwd = forward(case)
    log(f"[bench] forward {nsteps} steps {t_fwd:.3f} s")

    def run(mode, particle_call=False, case=case):
        s = tpx.Solver(case.params, np_capacity=200_000)
        s.load_case(case)
        s.step(1)
        s.adjoint_set_gas_tape(mode)
        t0 = time.perf_counter()
        J = s.adjoint_record(nsteps, every)
        t_rec = time.perf_counter() - t0
        t_p = None
        if particle_call:
            t0 = time.perf_counter()
            s.adjoint_gradient()
            t_p = time.perf_counter() - t0
        t0 = time.perf_counter()
        full = s.adjoint_gradient_full()
        t_rev = time.perf_counter() - t0
        s.adjoint_release()
        s.close()
        log(f"[bench] gas tape {mode}: recorded forward {t_rec:.3f} s, full sweep {t_rev:.3f} s"
            + (f", particle-only sweep {t_p:.3f} s" if t_p is not None else "")
            + f"; J = {J:.6e}, dJ/dA = {full['A']:.6e}, |dJ/dRU| = {np.linalg.norm(full['RU']):.3e}"
            + (f", |dJ/deps| = {np.linalg.norm(full['eps']):.3e}, dJ/dP0 = {full['P0']:.6e}" if "eps" in full else ""))
        return t_rec, t_rev, t_p, full

    t_rec, t_rev, _, full = run("segments")
    t_rec_k, t_rev_k, t_rev_p, full_k = run("keep", particle_call=True)
    t_rec_r, t_rev_r, _, full_r = run("recompute")
    t_fwd_c = forward(case_cold)
    log(f"[bench] forward {nsteps} steps without the heat exchange {t_fwd_c:.3f} s")
    _, t_rev_c, _, _ = run("segments", case=case_cold)
    ref = np.max(np.abs(full_k["RU"]))
    agree = max(float(np.max(np.abs(f["RU"] - full_k["RU"])) / ref) for f in (full, full_r))
    return {"metric": "adjoint/forward cost ratio (128^3, 2e5 particles, 50 steps)", "value": round(t_rev / t_fwd, 4),
            "unit": "ratio", "steps": nsteps, "higher_is_better": False,
            "workload": "adjoint: 128^3, 2e5 particles, St=2, two-way coupling at mass loading 0.5, heat exchange "
                        "on (per-particle eps in [0.5, 1]), J = Var(C) of the final positions, gradients w.r.t. "
                        "every particle's tau, eps and initial state, the forcing coefficient A and the initial "
                        "RU, Rho, P, P0; gas and particles checkpointed every 10 steps, each segment recomputed "
                        "in the sweep (value: that sweep over the plain forward; keep_*: the gas tape kept per "
                        "substep, recompute_*: the gas recomputed from every step's start; no_heat_*: the same "
                        "segment sweep and forward without the heat exchange)",
            "forward_s": round(t_fwd, 4), "recorded_forward_s": round(t_rec, 4), "reverse_s": round(t_rev, 4),
            "record_and_reverse_over_forward": round((t_rec + t_rev) / t_fwd, 4),
            "keep_reverse_s": round(t_rev_k, 4), "keep_reverse_over_forward": round(t_rev_k / t_fwd, 4),
            "keep_recorded_forward_s": round(t_rec_k, 4),
            "recompute_reverse_s": round(t_rev_r, 4), "recompute_reverse_over_forward": round(t_rev_r / t_fwd, 4),
            "recompute_recorded_forward_s": round(t_rec_r, 4),
            "particle_reverse_s": round(t_rev_p, 4), "particle_reverse_over_forward": round(t_rev_p / t_fwd, 4),
            "recorded_over_plain": round(t_rec / t_fwd, 4), "tapes_agree_to": agree,
            "no_heat_forward_s": round(t_fwd_c, 4), "no_heat_reverse_s": round(t_rev_c, 4),
            "no_heat_reverse_over_forward": round(t_rev_c / t_fwd_c, 4)}


def adjoint_bench(args):
    a = adjoint_measure()
    print(json.dumps({
        "metric": a["metric"], "value": a["value"], "unit": "ratio", "n_gpus": 1, "steps": a["steps"], "warmup": 1,
        "higher_is_better": False, "dtype": "f64", "data": "synthetic", "config": {"workload": a["workload"]},
        **{k: v for k, v in a.items() if k not in ("metric", "value", "unit", "steps", "higher_is_better",
                                                   "workload")},
        "roofline": None, "cpu_baseline": None}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="heat256", choices=["heat256", "plain128", "plain32", "ensemble", "adjoint"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the compact ensemble and adjoint figures added to the heat256 line")
    ap.add_argument("--cpu-full", action="store_true",
                    help="CPU baseline over 3 timed steps on all threads and a whole step on 1 thread")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--timing-steps", type=int, default=2)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain copy of MIB MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU: the step is measured at N=1 only; "
                                   "multi-GPU decomposition is a design (see the sharding notes), "
                                   f"not a runnable mode (asked for --gpus {args.gpus})"}))
        sys.exit(2)

    import tpx
    from tpx import ic

    if args.workload == "adjoint":
        return adjoint_bench(args)
    if args.workload == "ensemble":
        ensemble_bench(args)
        return

    t0 = time.perf_counter()
    case = ic.config_case(args.workload, seed=0)
    p = case.params
    npart = len(case.particles["x"])
    ncell = p.Nx_tot * p.Ny_tot * p.Nz_tot
    log(f"[bench] case {args.workload}: {p.Nx_tot}^3, {npart} particles, dt={p.dt:.5g} "
        f"(built in {time.perf_counter() - t0:.1f} s)")
    # the deposit kernels' tile shapes, tuned inside steps of a second solver
    # on the same case (so the timed steps are the same steps of the run)
    t0 = time.perf_counter()
    tuner = tpx.Solver(p, np_capacity=npart)
    tuner.load_case(case)
    tiles = tuner.tune_tiles(max_steps=12)
    tuner.close()
    del tuner
    log(f"[bench] tile shapes {tiles} tuned in {time.perf_counter() - t0:.2f} s")
    s = tpx.Solver(p, np_capacity=npart)
    s.load_case(case)
    s.set_tiles(heat=tiles["heat"], velocity=tiles["velocity"])
    log("[bench] uploaded and initialised")
    s.step(args.warmup)
    s.synchronize()
    log(f"[bench] warmup {args.warmup} steps done")

    # timed region: K whole steps, device synchronised on both sides
    s.synchronize()
    it0 = s.scalars()["total_iterations"]
    t0 = time.perf_counter()
    s.step(args.steps)
    s.synchronize()
    elapsed = time.perf_counter() - t0
    mean_its = (s.scalars()["total_iterations"] - it0) / (4.0 * args.steps)
    value = args.steps / elapsed
    n_it_last = int(s.scalars()["num_iteration"])
    log(f"[bench] timed {args.steps} steps in {elapsed:.3f} s -> {value:.3f} steps/s")

    # per-kernel-class device times with HIP events on the library's stream
    s.timing(True)
    it_t0 = s.scalars()["total_iterations"]
    s.step(args.timing_steps)
    ms, launches = s.timing_read()
    s.timing(False)
    iters_timed = s.scalars()["total_iterations"] - it_t0
    per_step = {k: ms[k] / args.timing_steps for k in ms}
    log("[bench] device ms per step by kernel class: " +
        ", ".join(f"{k}={v:.3f}" for k, v in per_step.items()))
    traffic = load_traffic(args.workload)
    counted = bool(p.TWC or p.Nu > 0) and npart > 0
    per_eps, per_tau = int(case.eps is not None), int(case.tau is not None)
    kernels = {
        "advance": ("k_advance", "particle RK4 advance with its velocity gather and cell count",
                    sum(advance_bytes_per_launch(ncell, npart, rk, per_eps, per_tau, bool(p.TWC), counted)
                        for rk in range(4)) / 4.0),
        "grid_rhs": ("k_rhs_ru_kernel", "density + momentum right-hand side, fused RK4 update",
                     sum(grid_rhs_bytes_per_launch(ncell, rk, p.TWC and npart > 0) for rk in range(4)) / 4.0),
    }
    # the three passes of a Poisson iteration; their launches queued after a
    # solve has stopped return at once, so a pass's time is divided by the
    # iterations that ran
    spec = (p.Nx_tot // 2 + 1) * p.Ny_tot * p.Nz_tot * 16.0
    # pass 1 at x lengths 64, 128 and 256 is the packed real-row form unless TPX_XPACK=0
    packed = p.Nx_tot in (64, 128, 256) and os.environ.get("TPX_XPACK", "1") != "0"
    fft = {
        "fft": ("k_xy_fwd_rp" if packed else "k_xy_fwd",
                "Poisson iteration pass 1: residual, x transform, first y step", 3 * 8.0 * ncell + spec),
        "fft_yz": ("k_yz_pass", "Poisson iteration pass 2: y and z transforms, division by lambda", 2 * spec),
        "fft_inv": ("k_xy_inv", "Poisson iteration pass 3: inverse y and x transforms, P += dp", spec + 2 * 8.0 * ncell),
    }
    kernels.update(fft)
    lines = {}
    for cls, (kname, what, bytes_launch) in kernels.items():
        if launches.get(cls, 0) == 0:
            continue
        avg_ms = ms[cls] / (iters_timed if cls in fft else launches[cls])
        achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
        t = traffic.get(kname)
        lines[cls] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": round(achieved / HBM_PEAK_GBS, 4),
                      "traffic": round(t["traffic_bytes"]) if t else None,
                      "kernel": f"{kname} ({what})", "bytes_per_launch": bytes_launch,
                      "avg_launch_ms": round(avg_ms, 4), "ms_per_step": round(per_step[cls], 3)}
    dominant = max(lines, key=lambda k: lines[k]["ms_per_step"])
    roofline = lines[dominant]
    others = [v for k, v in lines.items() if k != dominant]

    # what a caller pays to take the full state back to the host (the step
    # itself never crosses PCIe; state is uploaded once and stays resident)
    t0 = time.perf_counter()
    for name in ("RU", "Rho", "P"):
        s.get(name)
    s.particles()
    download_ms = (time.perf_counter() - t0) * 1e3
    log(f"[bench] full state download {download_ms:.1f} ms")

    # whole-step bandwidth: the bytes these kernels move by their
    # algorithmic figures, and a byte model of the reference's unfused
    # whole-array passes (26 field passes a Poisson iteration), both at the
    # iteration count of the run
    whole = step_bytes(ncell, npart, mean_its) * value / 1e9
    kbytes = step_kernel_bytes(ncell, npart, 4 * mean_its, per_eps, per_tau, bool(p.TWC), p.Nu > 0, counted)
    whole_k = kbytes * value / 1e9
    s.close()
    del s

    extras = {}
    if not args.no_extras:
        # the other two headline figures, measured in the same run (compact:
        # a few batched ensemble steps; the adjoint configuration in full)
        log("[bench] ensemble (64 x 64^3) ...")
        extras["ensemble"] = ensemble_measure(steps=5, warmup=2, with_streams=False)
        log("[bench] adjoint (128^3, 2e5 particles, 50 steps) ...")
        extras["adjoint"] = adjoint_measure()

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle) ...")
        cpu = cpu_baseline(case, args.cpu_threads, full=args.cpu_full)
        log(f"[bench] CPU baseline {cpu['value']:.4f} steps/s on {cpu['cores']} threads")

    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1000.0 / value, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {p.Nx_tot}^3 FP64 staggered grid, {npart} particles, "
                               f"St=1, mass loading {case.info.get('phi')}, heat "
                               f"{'on' if p.Nu > 0 else 'off'}, FFT Poisson (epsilon {p.epsilon:g})",
                   "grid": p.Nx_tot, "particles": npart, "dt": p.dt,
                   "deposit_tiles": {"heat": list(tiles["heat"]), "velocity": list(tiles["velocity"])},
                   "poisson_iterations_last_substep": n_it_last,
                   "poisson_iterations_mean": round(mean_its, 3),
                   "poisson_iterations_per_step": round(4 * mean_its, 2)},
        "roofline": roofline,
        "roofline_next": others,
        "cpu_baseline": cpu,
        "kernel_ms_per_step": {k: round(v, 4) for k, v in per_step.items()},
        "step_kernel_bytes_GB": round(kbytes / 1e9, 2),
        "step_GBps_kernel_bytes": round(whole_k, 1),
        "step_frac_kernel_bytes": round(whole_k / HBM_PEAK_GBS, 4),
        "step_GBps_unfused_pass_model": round(whole, 1),
        "state_download_ms": round(download_ms, 1),
    }
    if cpu:
        out["speedup_vs_cpu_baseline"] = round(value / cpu["value"], 1)
    out.update(extras)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
