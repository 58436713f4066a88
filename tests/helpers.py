"""Shared helpers of the test suite: building matched oracle / device runs
from one synthetic case and comparing them with the parity criterion
max|a-b| / max|b| per field (particles matched by flag)."""
from __future__ import annotations

import numpy as np

from oracle.oracle import Oracle
from tpx import ic
from tpx.params import Params

FIELDS_T1 = ("RU",)
FIELDS_T0 = ("Rho", "P")


def oracle_from_case(case: ic.Case) -> Oracle:
    o = Oracle(case.params.struct_fields(), np_capacity=max(len(case.particles["x"]), 1))
    o.set("RU", case.RU)
    o.set("Rho", case.Rho)
    o.set("P", case.P)
    pp = case.particles
    o.set_particles(pp["x"], pp["y"], pp["z"], pp["u"], pp["v"], pp["w"], pp["T"], pp["flag"],
                    tau=case.tau, eps=case.eps)
    if getattr(case, "PS", None) is not None:
        o.set("Passive_Scalar", case.PS)
    if getattr(case, "RV", None) is not None:
        o.set("RV", case.RV)
        o.set("Q", case.Q if getattr(case, "Q", None) is not None else np.zeros(case.RV.shape[1:]))
    o.set_restart_numbers(case.params.P0, 0.0, 0)
    o.initialize()
    return o


def solver_from_case(case: ic.Case):
    from tpx import Solver
    s = Solver(case.params, np_capacity=max(len(case.particles["x"]), 1))
    s.load_case(case)
    return s


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    den = float(np.max(np.abs(b)))
    num = float(np.max(np.abs(np.asarray(a) - np.asarray(b))))
    if den == 0.0:
        return num
    return num / den


def by_flag(parts: dict) -> dict:
    order = np.argsort(parts["flag"], kind="stable")
    return {k: v[order] for k, v in parts.items()}


def particle_errors(pa: dict, pb: dict, L) -> dict:
    a, b = by_flag(pa), by_flag(pb)
    assert np.array_equal(a["flag"], b["flag"]), "particle id sets differ"
    out = {}
    for c, q in enumerate("xyz"):
        d = a[q] - b[q]
        d = d - L[c] * np.round(d / L[c])
        den = float(np.max(np.abs(b[q]))) or 1.0
        out[q] = float(np.max(np.abs(d))) / den if len(d) else 0.0
    for q in "uvwT":
        out[q] = rel_err(a[q], b[q]) if len(a[q]) else 0.0
    return out


def small_case(n=16, npart=2048, nu=0.02, seed=3, heat=False, twc=True, variable_rho=False,
               cooling=False, tau_spread=False, eps_range=None, gravity=False, U0=0.0):
    """A configs-style case shrunk for tests, with switches for each physics term."""
    st = 1.0
    if tau_spread:
        st = ic.loguniform(np.random.default_rng(seed + 5), 0.1, 10.0, npart)
    case = ic.make_case(n, npart, nu=nu, seed=seed, St=st, phi=0.3, heat=heat, twc=twc,
                        eps_range=eps_range, nsteps=20)
    p = case.params
    if variable_rho:
        rng = np.random.default_rng(seed + 11)
        x = (np.arange(n) + 0.5) * (p.Lx / n)
        X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
        rho = 1.0 + 0.2 * np.sin(X.transpose(2, 1, 0)) * np.cos(2 * Y.transpose(2, 1, 0)) + \
            0.05 * rng.standard_normal((n, n, n))
        case.Rho = rho
        case.RU = case.RU * 1.0
    if cooling:
        p.cooling = 1
    if gravity:
        p.gx1 = p.gx2 = 0.3
        p.gz1 = p.gz2 = -0.5
        p.ParticleGravity = 1
    if U0:
        p.U0 = U0
    return case
