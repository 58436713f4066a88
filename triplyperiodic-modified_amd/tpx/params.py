"""Input-file parameters of the solver.

The file format is the reference's positional one (params.cpp:13-59): lines
whose first character is '#' and empty lines are skipped, every other line is
appended to one whitespace-separated stream, and 37 groups are read from it in
a fixed order.  Derived quantities follow params.cpp:61-72 and the threshold
switch follows params.cpp:75-86.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Any, Dict, List, Tuple

PI = 3.141592653589793  # params.h:5

# (field name, python type) in the order params.cpp:23-59 reads them.
_ORDER: List[Tuple[str, type]] = [
    ("Nx_tot", int), ("Ny_tot", int), ("Nz_tot", int),
    ("Lx", float), ("Ly", float), ("Lz", float),
    ("elongated_box", int),
    ("solve_for_scalar", int),
    ("solve_for_vector", int),
    ("bs", int),
    ("Np_track", int),
    ("T_final", float), ("dt", float), ("threshold", float),
    ("Statistics", int), ("Stat_print", int),
    ("Data_freq_fast", int), ("Data_freq_slow", int),
    ("Data_dir", str), ("Stat_dir", str),
    ("Initial", int),
    ("Initial_C", int),
    ("Initial_V", int),
    ("PreCond", int), ("Solver", int),
    ("cooling", int),
    ("Iteration1", int), ("Iteration2", int), ("epsilon", float),
    ("TWC", int), ("ParticleGravity", int),
    ("A1", float), ("A2", float),
    ("Rho0", float),
    ("Rho_forV", float),
    ("U0", float), ("V0", float), ("W0", float),
    ("V0_1", float), ("V0_2", float), ("V0_3", float),
    ("T0", float),
    ("Mu0", float), ("k", float),
    ("gx1", float), ("gy1", float), ("gz1", float),
    ("gx2", float), ("gy2", float), ("gz2", float),
    ("Cp", float), ("Cv", float),
    ("np0", float),
    ("Dp", float), ("Rhop", float),
    ("Cvp", float),
    ("epsilonp", float),
    ("Nu", float),
    ("I01", float), ("I02", float),
    ("D_M", float),
    ("eta", float),
    ("s1t", float),
    ("s2t", float),
]

# line grouping used when writing a file (one reference group per line)
_GROUPS = [3, 3, 1, 1, 1, 1, 1, 3, 2, 2, 2, 1, 1, 1, 2, 1, 3, 2, 2, 1, 1, 3, 3,
           1, 2, 6, 2, 1, 2, 1, 1, 1, 2, 1, 1, 1, 1]


@dataclasses.dataclass
class Params:
    Nx_tot: int = 32
    Ny_tot: int = 32
    Nz_tot: int = 32
    Lx: float = 2 * PI
    Ly: float = 2 * PI
    Lz: float = 2 * PI
    elongated_box: int = 0
    solve_for_scalar: int = 0
    solve_for_vector: int = 0
    bs: int = 2
    Np_track: int = 0
    T_final: float = 1.0
    dt: float = 1e-3
    threshold: float = 1e9
    Statistics: int = 0
    Stat_print: int = 0
    Data_freq_fast: int = 1000000
    Data_freq_slow: int = 1000000
    Data_dir: str = "./"
    Stat_dir: str = "./"
    Initial: int = 1
    Initial_C: int = 3
    Initial_V: int = 2
    PreCond: int = 0
    Solver: int = 0
    cooling: int = 0
    Iteration1: int = 50
    Iteration2: int = 50
    epsilon: float = 1e-13
    TWC: int = 1
    ParticleGravity: int = 0
    A1: float = 0.0
    A2: float = 0.0
    Rho0: float = 1.0
    Rho_forV: float = 1.0
    U0: float = 0.0
    V0: float = 0.0
    W0: float = 0.0
    V0_1: float = 0.0
    V0_2: float = 0.0
    V0_3: float = 0.0
    T0: float = 1.0
    Mu0: float = 0.02
    k: float = 0.1
    gx1: float = 0.0
    gy1: float = 0.0
    gz1: float = 0.0
    gx2: float = 0.0
    gy2: float = 0.0
    gz2: float = 0.0
    Cp: float = 3.5
    Cv: float = 2.5
    np0: float = 0.0
    Dp: float = 0.01
    Rhop: float = 1000.0
    Cvp: float = 2.5
    epsilonp: float = 1.0
    Nu: float = 0.0
    I01: float = 0.0
    I02: float = 0.0
    D_M: float = 0.0
    eta: float = 0.0
    s1t: float = 1.0
    s2t: float = 1.0

    # ---- derived (params.cpp:61-72, params.h:115-127) ----
    @property
    def R(self) -> float:
        return self.Cp - self.Cv

    @property
    def P0(self) -> float:
        return self.Rho0 * self.T0 * self.R

    @property
    def Tp(self) -> float:
        return self.Rhop * self.Dp * self.Dp / (18 * self.Mu0)

    @property
    def mp(self) -> float:
        return self.Rhop * PI * self.Dp * self.Dp * self.Dp / 6.

    @property
    def Vol(self) -> float:
        return self.Lx * self.Ly * self.Lz

    @property
    def N0(self) -> int:
        return int(self.np0 * self.Vol)

    def switched(self, time: float) -> Dict[str, Any]:
        """params::update (params.cpp:75-86): the *2 values once time > threshold."""
        late = time > self.threshold
        return {
            "Iteration": self.Iteration2 if late else self.Iteration1,
            "A": self.A2 if late else self.A1,
            "g": (self.gx2, self.gy2, self.gz2) if late else (self.gx1, self.gy1, self.gz1),
            "I0": self.I02 if late else self.I01,
        }

    # ---- I/O ----
    @classmethod
    def parse(cls, text: str) -> "Params":
        kept = [ln for ln in text.split("\n") if ln != "" and not ln.startswith("#")]
        toks = " ".join(kept).split()
        if len(toks) < len(_ORDER):
            raise ValueError(f"parameter file has {len(toks)} values, {len(_ORDER)} needed")
        vals = {}
        for (name, typ), tok in zip(_ORDER, toks):
            if typ is int:
                vals[name] = int(tok)
            elif typ is float:
                vals[name] = float(tok)
            else:
                vals[name] = tok
        return cls(**vals)

    @classmethod
    def read(cls, path: str) -> "Params":
        with open(path) as fh:
            return cls.parse(fh.read())

    def text(self) -> str:
        names = [n for n, _ in _ORDER]
        out, pos = ["# triply periodic particle-laden low-Mach solver input"], 0
        for g in _GROUPS:
            vals = []
            for n in names[pos:pos + g]:
                v = getattr(self, n)
                vals.append(repr(float(v)) if isinstance(v, float) else str(v))
            out.append("# " + " ".join(names[pos:pos + g]))
            out.append(" ".join(vals))
            pos += g
        return "\n".join(out) + "\n"

    def write(self, path: str) -> None:
        with open(path, "w") as fh:
            fh.write(self.text())

    # ---- C structs (oracle and libtpx use the same field list) ----
    def struct_fields(self) -> Dict[str, Any]:
        return dict(
            nx=self.Nx_tot, ny=self.Ny_tot, nz=self.Nz_tot,
            lx=self.Lx, ly=self.Ly, lz=self.Lz, bs=self.bs,
            dt=self.dt, threshold=self.threshold,
            iteration1=self.Iteration1, iteration2=self.Iteration2, epsilon=self.epsilon,
            twc=self.TWC, particle_gravity=self.ParticleGravity, cooling=self.cooling,
            a1=self.A1, a2=self.A2, rho0=self.Rho0,
            u0=self.U0, v0=self.V0, w0=self.W0,
            t0=self.T0, mu0=self.Mu0, k=self.k,
            gx1=self.gx1, gy1=self.gy1, gz1=self.gz1,
            gx2=self.gx2, gy2=self.gy2, gz2=self.gz2,
            cp=self.Cp, cv=self.Cv, np0=self.np0, dp=self.Dp, rhop=self.Rhop,
            cvp=self.Cvp, epsilonp=self.epsilonp, nu=self.Nu, i01=self.I01, i02=self.I02,
            solve_for_scalar=int(self.solve_for_scalar), s1_type=int(self.s1t != 0), d_m=self.D_M,
            solve_for_vector=int(self.solve_for_vector), s2_type=int(self.s2t != 0), eta=self.eta,
            rho_forv=self.Rho_forV,
        )


STRUCT_FIELDS: List[Tuple[str, str]] = [
    ("nx", "i"), ("ny", "i"), ("nz", "i"),
    ("lx", "d"), ("ly", "d"), ("lz", "d"), ("bs", "i"),
    ("dt", "d"), ("threshold", "d"),
    ("iteration1", "i"), ("iteration2", "i"), ("epsilon", "d"),
    ("twc", "i"), ("particle_gravity", "i"), ("cooling", "i"),
    ("a1", "d"), ("a2", "d"), ("rho0", "d"),
    ("u0", "d"), ("v0", "d"), ("w0", "d"),
    ("t0", "d"), ("mu0", "d"), ("k", "d"),
    ("gx1", "d"), ("gy1", "d"), ("gz1", "d"), ("gx2", "d"), ("gy2", "d"), ("gz2", "d"),
    ("cp", "d"), ("cv", "d"), ("np0", "d"), ("dp", "d"), ("rhop", "d"),
    ("cvp", "d"), ("epsilonp", "d"), ("nu", "d"), ("i01", "d"), ("i02", "d"),
    ("solve_for_scalar", "i"), ("s1_type", "i"), ("d_m", "d"),
    ("solve_for_vector", "i"), ("s2_type", "i"), ("eta", "d"), ("rho_forv", "d"),
]


def is_pow2(n: int) -> bool:
    return n > 0 and (n & (n - 1)) == 0


def cfl_dt(umax_over_dx: float, cfl: float = 0.3) -> float:
    """dt such that max_c(max U_c / dx_c) * dt = cfl (tensor1::max_cfl, tensor1.cpp:365-372)."""
    if not (umax_over_dx > 0 and math.isfinite(umax_over_dx)):
        raise ValueError("velocity field has no positive maximum")
    return cfl / umax_over_dx
