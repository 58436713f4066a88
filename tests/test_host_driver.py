"""The C++ host mirror of the reference's classes (params, gridsize, grid,
particle, poisson, communicator) and its unchanged driver loop, built as
tpx_sim on top of the C ABI.

CPU: the C++ input-file parser reads the same values as the Python one
(params.cpp:6-86), including the derived Tp, mp, P0 and N0.
GPU: tpx_sim run from restart files (Initial = 1) writes RU/Rho/P and the
particle files that match the oracle after two steps, and with the passive
scalar on (Initial_C = 5) its C_<n>.bin.
"""
import os
import subprocess

import numpy as np
import pytest

from helpers import oracle_from_case, particle_errors, rel_err, small_case
from tpx import ic
from tpx.params import Params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "triplyperiodic-modified_amd", "host")
SIM = os.path.join(HOST, "tpx_sim")


@pytest.fixture(scope="module")
def sim():
    if not os.path.exists(SIM):
        subprocess.run(["make", "-C", HOST], check=True, capture_output=True)
    return SIM


def test_cpp_parser_matches_python(sim, tmp_path):
    case = small_case(n=16, npart=1000, heat=True, seed=4)
    p = case.params
    p.gy1, p.gz2, p.U0, p.threshold, p.Iteration2 = 0.25, -1.5, 0.125, 0.75, 7
    path = tmp_path / "input.txt"
    p.write(str(path))
    # comments and blank lines between groups are skipped like the reference's parser
    text = path.read_text().replace("\n# Rho0", "\n\n# Rho0")
    path.write_text(text)
    out = subprocess.run([sim, "--print-params", str(path)], check=True, capture_output=True, text=True).stdout
    got = {k: float(v) for k, v in (ln.split() for ln in out.strip().splitlines())}
    q = Params.read(str(path))
    want = q.struct_fields()
    for k, v in want.items():
        assert got[k] == float(v), k
    assert got["T_final"] == q.T_final
    assert got["N0"] == q.N0
    assert got["Tp"] == q.Tp and got["mp"] == q.mp and got["P0"] == q.P0


def test_cpp_parser_rejects_a_short_file(sim, tmp_path):
    path = tmp_path / "short.txt"
    path.write_text("# too short\n32 32 32\n6.28 6.28\n")
    r = subprocess.run([sim, "--print-params", str(path)], capture_output=True, text=True)
    assert r.returncode != 0


@pytest.mark.gpu
def test_driver_from_restart_matches_oracle(gpu_available, sim, tmp_path):
    case = small_case(n=16, npart=1500, heat=True, variable_rho=True, eps_range=None, seed=23)
    p = case.params
    npart = len(case.particles["x"])
    p.np0 = (npart + 0.5) / p.Vol
    assert p.N0 == npart
    p.Initial = 1
    p.Data_dir = str(tmp_path) + "/"
    p.Stat_dir = str(tmp_path) + "/"
    p.Data_freq_fast = 1
    p.Data_freq_slow = 2
    # The loop runs while T_cur < T_final and only then asks Touch(), which
    # stores: with T_final = 2.5 dt it runs three steps and stores after the
    # first two (the reference's driver never stores its last step).
    p.T_final = 2.5 * p.dt
    p.write(str(tmp_path / "input.txt"))
    ic.write_restart(str(tmp_path), case.RU, case.Rho, case.P, p.P0, 0.0, 0, case.particles)
    print("tpx_sim 16^3 + 1500 particles, 3 steps", flush=True)
    r = subprocess.run([sim, "input.txt"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    o = oracle_from_case(case)
    o.step(2)
    n = p.Nx_tot
    RU = np.fromfile(tmp_path / "RU_2.bin").reshape(3, n, n, n)
    Rho = np.fromfile(tmp_path / "Rho_2.bin").reshape(n, n, n)
    P = np.fromfile(tmp_path / "P_2.bin").reshape(n, n, n)
    errs = {"RU": rel_err(RU, o.get("RU")), "Rho": rel_err(Rho, o.get("Rho")), "P": rel_err(P, o.get("P"))}
    parts = {q: np.fromfile(tmp_path / f"Part_{q}_2.bin") for q in "xyzuvwT"}
    parts["flag"] = np.fromfile(tmp_path / "Part_flag_2.bin", dtype=np.int32)
    errs.update(particle_errors(parts, o.particles("base"), (p.Lx, p.Ly, p.Lz)))
    with open(tmp_path / "numbers_2.dat") as fh:
        P0, T_cur, nts = fh.read().split()
    assert int(nts) == 2 and abs(float(T_cur) - 2 * p.dt) < 1e-15
    assert abs(float(P0) - o.scalars()["P0"]) <= 1e-12 * abs(o.scalars()["P0"])
    assert os.path.exists(tmp_path / "RU_1.bin") and not os.path.exists(tmp_path / "P_1.bin")
    assert not os.path.exists(tmp_path / "RU_3.bin")
    print({k: f"{v:.2e}" for k, v in errs.items()})
    assert max(errs.values()) <= 1e-10, errs


@pytest.mark.gpu
def test_driver_runs_the_passive_scalar(gpu_available, sim, tmp_path):
    """solve_for_scalar = 1 with the analytic Initial_C = 5: the driver calls
    Update_Passive_Scalar after Update_Particle and writes the scalar as
    C_<n>.bin (grid.cpp:538-544), matching the oracle."""
    case = small_case(n=16, npart=1000, heat=True, variable_rho=True, eps_range=None, seed=29)
    p = case.params
    npart = len(case.particles["x"])
    p.np0 = (npart + 0.5) / p.Vol
    p.Initial, p.Initial_C = 1, 5
    p.solve_for_scalar, p.D_M, p.s1t = 1, 0.05, 0.0
    p.Data_dir = str(tmp_path) + "/"
    p.Stat_dir = str(tmp_path) + "/"
    p.Data_freq_fast = 1
    p.Data_freq_slow = 2
    p.T_final = 2.5 * p.dt
    p.write(str(tmp_path / "input.txt"))
    ic.write_restart(str(tmp_path), case.RU, case.Rho, case.P, p.P0, 0.0, 0, case.particles)
    print("tpx_sim 16^3 + 1000 particles with the passive scalar, 3 steps", flush=True)
    r = subprocess.run([sim, "input.txt"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    n = p.Nx_tot
    d = p.Lx / n
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    X, Y = d / 2. + i * d, d / 2. + j * d
    case.PS = np.sin(2 * np.pi / p.Lx * X) * np.sin(2 * np.pi / p.Ly * Y)  # grid.cpp:385-402
    o = oracle_from_case(case)
    o.step(2)
    C = np.fromfile(tmp_path / "C_2.bin").reshape(n, n, n)
    e = rel_err(C, o.get("Passive_Scalar"))
    print("passive scalar after 2 steps", e)
    assert e <= 1e-10
