"""ctypes binding of the oracle library (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module: it is the parity checker and the CPU timing baseline, never the
product.  Arrays cross the boundary in the interior layout (nz, ny, nx),
x fastest.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")

_FIELD_FMT = [
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


class OrParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int if t == "i" else ctypes.c_double) for n, t in _FIELD_FMT]


T1 = {"RU": 0, "RU_int": 1, "RU_new": 2, "RU_np1": 3, "RHS_RU": 4, "U": 5,
      "Rho_face": 6, "dummy": 7, "dummy2": 8, "Passive_Scalar_face": 9,
      "RV": 10, "RV_int": 11, "RV_new": 12, "RV_np1": 13, "RHS_RV": 14, "V": 15, "S2": 16}
T0 = {"P": 32, "RHS_Pois": 33, "C": 34, "Rho": 35, "Rho_int": 36, "Rho_new": 37,
      "Rho_np1": 38, "T": 39, "divergence": 40, "RHS_Part_Temp": 41, "RHS_Rho": 42,
      "Passive_Scalar": 43, "Passive_Scalar_int": 44, "Passive_Scalar_new": 45,
      "Passive_Scalar_np1": 46, "RHS_Passive_Scalar": 47, "S1": 48, "Q": 49, "RHS_Pois_Q": 50}
STAGE = {"base": 0, "int": 1, "new": 2, "np1": 3}
AUX = {"RHS_u": 0, "RHS_v": 1, "RHS_w": 2, "ug": 3, "vg": 4, "wg": 5, "Tg": 6,
       "RHS_T": 7, "tau": 8, "eps": 9}
SCALARS = ["P0", "P0_int", "P0_new", "P0_np1", "T_cur", "num_timestep",
           "num_iteration", "mean_energy_transferred", "A", "I0", "Iteration",
           "gx", "gy", "gz", "dP0_dt", "total_iterations"]
OPS = {"div_f2c": 0, "grad_c2f": 1, "del2": 2, "i_c2f": 3, "i_f2c": 4,
       "ix_c2f": 5, "iy_c2f": 6, "iz_c2f": 7, "ix_f2c": 8, "iy_f2c": 9, "iz_f2c": 10,
       "fold": 11, "ccp_fft": 12}

_libs: Dict[str, ctypes.CDLL] = {}


def lib(path: Optional[str] = None):
    """The oracle library at `path` (default: the in-tree build), loaded once
    per path.  Other paths serve the mutation tests, which build copies of the
    oracle with a planted mistake and check that the closed-form pins fail."""
    path = os.path.abspath(path or LIB_PATH)
    if path not in _libs:
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run make -C oracle)")
        L = ctypes.CDLL(path)
        P, D, I = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int
        L.or_create.restype = P
        L.or_create.argtypes = [ctypes.POINTER(OrParams), I]
        L.or_destroy.argtypes = [P]
        L.or_nthreads.argtypes = [I]
        L.or_nthreads.restype = I
        for fn in ("or_set_field", "or_get_field", "or_get_field_ghosted"):
            getattr(L, fn).argtypes = [P, I, I, D]
            getattr(L, fn).restype = I
        L.or_set_particles.argtypes = [P, I] + [D] * 7 + [ctypes.POINTER(ctypes.c_int), D, D]
        L.or_get_np.argtypes = [P]
        L.or_get_np.restype = I
        L.or_get_particles.argtypes = [P, I, ctypes.POINTER(D), ctypes.POINTER(ctypes.c_int)]
        L.or_get_particle_aux.argtypes = [P, I, D]
        L.or_set_particle_aux.argtypes = [P, I, D]
        L.or_get_scalars.argtypes = [P, D]
        L.or_set_restart_numbers.argtypes = [P, ctypes.c_double, ctypes.c_double, I]
        for fn in ("or_initialize", "or_params_update", "or_update_rho", "or_update_p0", "or_update_passive_scalar",
                   "or_update_rv_woq", "or_compute_rhs_pois_q", "or_solve_poisson_q", "or_update_rv_wq",
                   "or_update_particle", "or_update_ru_wop", "or_compute_div_u_new",
                   "or_compute_rhs_pois", "or_solve_poisson", "or_update_ru_wp",
                   "or_time_advance_rk4", "or_time_advance", "or_p_gather_velocity",
                   "or_p_deposit_velocity", "or_p_deposit_concentration", "or_p_send_recv"):
            getattr(L, fn).argtypes = [P]
        L.or_set_rk4_count.argtypes = [P, I]
        L.or_make_mean_u0.argtypes = [P, D]
        L.or_step.argtypes = [P, I]
        L.or_apply.argtypes = [P, I, D, D, D, D, D, D]
        L.or_apply.restype = I
        L.or_sum.argtypes = [P, D]
        L.or_sum.restype = ctypes.c_double
        L.or_max_abs.argtypes = [P, D]
        L.or_max_abs.restype = ctypes.c_double
        L.or_p_gather_temp.argtypes = [P, I]
        L.or_p_deposit_temp.argtypes = [P, I]
        L.or_fft3d.argtypes = [D, I, I, I, I]
        L.or_poisson_solve.argtypes = [P, D, D, D, D, D, ctypes.POINTER(I)]
        _libs[path] = L
    return _libs[path]


def _dp(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _c(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def fft3d(data: np.ndarray, sign: int) -> np.ndarray:
    """Complex 3-D FFT of a (nz, ny, nx) complex array; sign -1 forward, +1 inverse, unnormalised."""
    nz, ny, nx = data.shape
    buf = np.ascontiguousarray(data, dtype=np.complex128).copy()
    lib().or_fft3d(buf.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double)), nx, ny, nz, sign)
    return buf


class Oracle:
    """One single-rank instance of the reference step, restated on the CPU."""

    def __init__(self, fields: Dict, np_capacity: int = 1, lib_path: Optional[str] = None):
        self.L = lib(lib_path)
        self.par = OrParams(**fields)
        self.shape = (fields["nz"], fields["ny"], fields["nx"])
        self.bs = fields["bs"]
        self.h = self.L.or_create(ctypes.byref(self.par), int(max(np_capacity, 1)))
        if not self.h:
            raise RuntimeError("or_create failed")

    def close(self):
        if self.h:
            self.L.or_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- fields ----
    def set(self, name: str, arr: np.ndarray) -> None:
        if name in T1:
            arr = _c(arr).reshape((3,) + self.shape)
            for c in range(3):
                a = np.ascontiguousarray(arr[c])
                assert self.L.or_set_field(self.h, T1[name], c, _dp(a)) == 0
        else:
            a = _c(arr).reshape(self.shape)
            assert self.L.or_set_field(self.h, T0[name], 0, _dp(a)) == 0

    def get(self, name: str) -> np.ndarray:
        if name in T1:
            out = np.empty((3,) + self.shape)
            for c in range(3):
                a = np.empty(self.shape)
                assert self.L.or_get_field(self.h, T1[name], c, _dp(a)) == 0
                out[c] = a
            return out
        out = np.empty(self.shape)
        assert self.L.or_get_field(self.h, T0[name], 0, _dp(out)) == 0
        return out

    # ---- particles ----
    def set_particles(self, x, y, z, u, v, w, T, flag, tau=None, eps=None) -> None:
        arrs = [_c(a) for a in (x, y, z, u, v, w, T)]
        fl = np.ascontiguousarray(flag, dtype=np.int32)
        tau = None if tau is None else _c(tau)
        eps = None if eps is None else _c(eps)
        n = len(arrs[0])
        rc = self.L.or_set_particles(self.h, n, *[_dp(a) for a in arrs],
                                     fl.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dp(tau), _dp(eps))
        assert rc == 0
        self._keep = (arrs, fl, tau, eps)

    @property
    def np(self) -> int:
        return self.L.or_get_np(self.h)

    def particles(self, stage: str = "base") -> Dict[str, np.ndarray]:
        n = self.np
        qs = [np.empty(n) for _ in range(7)]
        ptrs = (ctypes.POINTER(ctypes.c_double) * 7)(*[_dp(a) for a in qs])
        fl = np.empty(n, dtype=np.int32)
        self.L.or_get_particles(self.h, STAGE[stage], ptrs, fl.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        out = dict(zip("x y z u v w T".split(), qs))
        out["flag"] = fl
        return out

    def aux(self, name: str) -> np.ndarray:
        out = np.empty(self.np)
        self.L.or_get_particle_aux(self.h, AUX[name], _dp(out))
        return out

    def set_aux(self, name: str, arr) -> None:
        a = _c(arr)
        assert len(a) == self.np
        self.L.or_set_particle_aux(self.h, AUX[name], _dp(a))

    # ---- scalars ----
    def scalars(self) -> Dict[str, float]:
        buf = np.empty(len(SCALARS))
        self.L.or_get_scalars(self.h, _dp(buf))
        return dict(zip(SCALARS, buf.tolist()))

    def set_restart_numbers(self, P0: float, T_cur: float = 0.0, num_timestep: int = 0) -> None:
        self.L.or_set_restart_numbers(self.h, P0, T_cur, num_timestep)

    # ---- methods ----
    def initialize(self): self.L.or_initialize(self.h)
    def params_update(self): self.L.or_params_update(self.h)
    def set_rk4_count(self, rk: int): self.L.or_set_rk4_count(self.h, rk)
    def update_rho(self): self.L.or_update_rho(self.h)
    def update_p0(self): self.L.or_update_p0(self.h)
    def update_particle(self): self.L.or_update_particle(self.h)
    def update_passive_scalar(self): self.L.or_update_passive_scalar(self.h)
    def update_rv_woq(self): self.L.or_update_rv_woq(self.h)
    def compute_rhs_pois_q(self): self.L.or_compute_rhs_pois_q(self.h)
    def solve_poisson_q(self): self.L.or_solve_poisson_q(self.h)
    def update_rv_wq(self): self.L.or_update_rv_wq(self.h)
    def update_ru_wop(self): self.L.or_update_ru_wop(self.h)
    def compute_div_u_new(self): self.L.or_compute_div_u_new(self.h)
    def compute_rhs_pois(self): self.L.or_compute_rhs_pois(self.h)
    def solve_poisson(self): self.L.or_solve_poisson(self.h)
    def update_ru_wp(self): self.L.or_update_ru_wp(self.h)
    def time_advance_rk4(self): self.L.or_time_advance_rk4(self.h)
    def time_advance(self): self.L.or_time_advance(self.h)

    def poisson_solve(self, C, P, RHS):
        """poisson::Solve(C, P, RHS) with the caller's face coefficient C
        (3, nz, ny, nx); returns the new P and the iteration count."""
        C = _c(C).reshape((3,) + self.shape)
        cs = [np.ascontiguousarray(C[c]) for c in range(3)]
        Pn = _c(P).reshape(self.shape).copy()
        R = _c(RHS).reshape(self.shape)
        it = ctypes.c_int(0)
        self.L.or_poisson_solve(self.h, *[_dp(a) for a in cs], _dp(Pn), _dp(R), ctypes.byref(it))
        return Pn, it.value

    def make_mean_u0(self, a: Sequence[float]):
        arr = _c(a)
        self.L.or_make_mean_u0(self.h, _dp(arr))

    def step(self, n: int = 1):
        self.L.or_step(self.h, int(n))

    SUBSTEP_METHODS = ("update_rho", "update_p0", "update_particle", "update_ru_wop",
                       "compute_div_u_new", "compute_rhs_pois", "solve_poisson",
                       "update_ru_wp", "time_advance_rk4")

    # ---- grid::Statistics (grid.cpp:902-982) ----
    def statistics(self, p) -> Dict[str, float]:
        """The reference's statistics over the oracle's state at a step
        boundary, keyed as tpx.STATS; p is the run's tpx.params.Params.  U is
        the field the last Compute_RHS_Pois left (its mean square is TKE2),
        then U = RU_np1 / Rho_face as the reference recomputes it; C and the
        particles' ug are what the last Update_Particle left.  With the
        passive vector off its entries are 0."""
        nx, ny, nz = p.Nx_tot, p.Ny_tot, p.Nz_tot
        n = nx * ny * nz
        dx, dy, dz = p.Lx / nx, p.Ly / ny, p.Lz / nz
        rho, C = self.get("Rho"), self.get("C")
        U_last = self.get("U")
        RU, Rf = self.get("RU_np1"), self.get("Rho_face")
        U = RU / Rf
        parts = self.particles("base")
        ug = self.aux("ug")
        sc = self.scalars()
        npart = self.np
        st = dict.fromkeys(["TKEV_V1", "TKEV_V2", "TKEV_V3", "VV", "Particle_CFL_max"], 0.0)
        if p.solve_for_vector:
            # VV: V as the last Compute_RHS_Pois_Q left it; then V = RV_np1 / Rho_forV,
            # V *= RV_np1 and the component means (grid.cpp:912, 918-933)
            st["VV"] = float((self.get("V") ** 2).sum()) / n
            RVn = self.get("RV_np1")
            for c, key in enumerate(("TKEV_V1", "TKEV_V2", "TKEV_V3")):
                st[key] = float(((RVn[c] / p.Rho_forV) * RVn[c]).sum()) / n
        st["Passive_Scalar_mean"] = float((self.get("Passive_Scalar") ** 2).sum()) / n  # mean_squares
        st["Rho_mean"], st["Rho_max"], st["Rho_min"] = rho.sum() / n, rho.max(), rho.min()
        st["C_max"], st["C_min"], st["C_mean"] = C.max(), C.min(), C.sum() / n
        st["TKE2"] = float((U_last ** 2).sum()) / n
        st["u_max"] = max(U[0].max(), U[1].max(), U[2].max())
        st["Gas_CFL_max"] = max(U[0].max() * p.dt / dx, U[1].max() * p.dt / dy, U[2].max() * p.dt / dz)
        UR = U * RU
        st["TKE_U"], st["TKE_V"], st["TKE_W"] = (UR[c].sum() / n for c in range(3))
        st["Vp_max"] = parts["u"].max() if npart else -1e9
        st["ug_max"] = ug.max() if npart else -1e9
        st["Tp_max"] = parts["T"].max() if npart else -1e9
        st["Tp_mean"] = parts["T"].sum() / npart if npart else 0.0
        st["Gas_diff_CFL_max"] = p.dt / (dx * dx * st["Rho_min"] / p.Mu0) * 6.
        st["balance_index"] = npart / float(p.N0) if p.N0 > 0 else 0.0
        st["HT_mean"] = sc["mean_energy_transferred"]
        st["Tg_mean"] = ((sc["P0"] / p.R) / rho).sum() / n
        st["P0"] = sc["P0"]
        st["num_iteration"] = sc["num_iteration"]
        return {k: float(v) for k, v in st.items()}

    # ---- particle primitives ----
    def p_gather_velocity(self): self.L.or_p_gather_velocity(self.h)
    def p_deposit_velocity(self): self.L.or_p_deposit_velocity(self.h)
    def p_deposit_concentration(self): self.L.or_p_deposit_concentration(self.h)
    def p_gather_temp(self, stage: str = "new"): self.L.or_p_gather_temp(self.h, STAGE[stage])
    def p_deposit_temp(self, stage: str = "new"): self.L.or_p_deposit_temp(self.h, STAGE[stage])
    def p_send_recv(self): self.L.or_p_send_recv(self.h)

    # ---- operator primitives ----
    def apply(self, op: str, a0=None, a1=None, a2=None):
        if op == "fold":
            ins = [_c(a0).ravel(), None, None]
        else:
            ins = [None if a is None else _c(a).reshape(self.shape) for a in (a0, a1, a2)]
        outs = [np.zeros(self.shape) for _ in range(3)]
        rc = self.L.or_apply(self.h, OPS[op], *[_dp(a) for a in ins], *[_dp(o) for o in outs])
        assert rc == 0
        if op in ("grad_c2f", "i_c2f"):
            return np.stack(outs)
        return outs[0]

    def sum(self, a) -> float:
        return self.L.or_sum(self.h, _dp(_c(a).reshape(self.shape)))

    def max_abs(self, a) -> float:
        return self.L.or_max_abs(self.h, _dp(_c(a).reshape(self.shape)))


def threads(n: int = 0) -> int:
    return lib().or_nthreads(int(n))
