"""Python access to libtpx, the MI355X implementation of the RK4 time step.

This is a thin ctypes layer over the C ABI of include/tpx.h (the same entry
points the C++ host classes call).  It fails loudly when the HIP library is
missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional, Sequence

import numpy as np

from .params import STRUCT_FIELDS, Params  # noqa: F401

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_DIR, "libtpx.so")

TPX_OK, TPX_E_ARG, TPX_E_HIP, TPX_E_STATE, TPX_E_UNSUPP, TPX_E_OOM = 0, -1, -2, -3, -4, -5

FIELD = {"RU": 0, "RU_int": 1, "RU_new": 2, "RU_np1": 3, "RHS_RU": 4,
         "RV": 5, "RV_int": 6, "RV_new": 7, "RV_np1": 8,
         "P": 16, "RHS_Pois": 17, "C": 18, "Rho": 19, "Rho_int": 20, "Rho_new": 21,
         "Rho_np1": 22, "divergence": 23, "RHS_Part_Temp": 24,
         "Passive_Scalar": 25, "Passive_Scalar_int": 26, "Passive_Scalar_new": 27, "Passive_Scalar_np1": 28,
         "Q": 29, "RHS_Pois_Q": 30}
VECTOR = {"RU", "RU_int", "RU_new", "RU_np1", "RHS_RU", "RV", "RV_int", "RV_new", "RV_np1"}
STAGE = {"base": 0, "int": 1, "new": 2, "np1": 3}
SCALARS = ["P0", "P0_int", "P0_new", "P0_np1", "T_cur", "num_timestep",
           "num_iteration", "mean_energy_transferred", "A", "I0", "Iteration",
           "gx", "gy", "gz", "dP0_dt", "total_iterations"]
# tpx_statistics outputs, in TPX_ST_* order (grid.cpp:902-982)
STATS = ["Tg_mean", "Tp_mean", "HT_mean", "TKE2", "TKE_U", "TKE_V", "TKE_W", "Passive_Scalar_mean",
         "TKEV_V1", "TKEV_V2", "TKEV_V3", "VV", "P0", "C_max", "C_min", "C_mean", "Rho_max", "Rho_min",
         "Rho_mean", "Particle_CFL_max", "Gas_CFL_max", "Gas_diff_CFL_max", "num_iteration",
         "balance_index", "u_max", "Vp_max", "ug_max", "Tp_max"]
KCLASS = ["grid_rhs", "div_rhs", "fft", "project", "particle", "deposit", "sort", "other", "advance",
          "fft_yz", "fft_inv"]

# C entry points of include/tpx.h, in declaration order
EXPORTS = [
    "tpx_create", "tpx_destroy", "tpx_last_error", "tpx_create_error",
    "tpx_upload_field", "tpx_download_field", "tpx_upload_particles", "tpx_num_particles",
    "tpx_download_particles", "tpx_download_particle_rhs_t", "tpx_set_restart_numbers",
    "tpx_get_scalars", "tpx_initialize", "tpx_params_update", "tpx_update_rho",
    "tpx_update_p0", "tpx_update_particle", "tpx_update_passive_scalar",
    "tpx_update_rv_woq", "tpx_compute_rhs_pois_q", "tpx_solve_poisson_q", "tpx_update_rv_wq",
    "tpx_update_ru_wop", "tpx_compute_div_u_new",
    "tpx_compute_rhs_pois", "tpx_solve_poisson", "tpx_poisson_solve", "tpx_poisson_solve_device", "tpx_update_ru_wp", "tpx_time_advance_rk4",
    "tpx_make_mean_u0", "tpx_time_advance", "tpx_step", "tpx_ensemble_step",
    "tpx_adjoint_record", "tpx_adjoint_gradient", "tpx_adjoint_gradient_full", "tpx_adjoint_gradient_heat", "tpx_adjoint_set_gas_tape", "tpx_statistics_enable", "tpx_statistics",
    "tpx_track_particles",
    "tpx_tune_tiles", "tpx_get_tiles", "tpx_set_tiles",
    "tpx_adjoint_release", "tpx_adjoint_check",
    "tpx_batch_create", "tpx_batch_step", "tpx_batch_destroy", "tpx_batch_error", "tpx_batch_last_error",
    "tpx_field_sum",
    "tpx_field_max_abs",
    "tpx_deposit_concentration", "tpx_get_stream", "tpx_timing_enable", "tpx_timing_read",
    "tpx_synchronize",
]


class TpxParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int if t == "i" else ctypes.c_double) for n, t in STRUCT_FIELDS]


class TpxError(RuntimeError):
    pass


_lib = None


def load(path: str = LIB_PATH):
    """Load libtpx.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise TpxError(f"libtpx.so not built: {path} (run __graft_entry__.build())")
    L = ctypes.CDLL(path)
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    DP = ctypes.POINTER(ctypes.c_double)
    IP = ctypes.POINTER(ctypes.c_int32)
    I64 = ctypes.c_int64
    L.tpx_create.argtypes = [ctypes.POINTER(TpxParams), I64, ctypes.POINTER(P)]
    L.tpx_destroy.argtypes = [P]
    L.tpx_last_error.argtypes = [P]
    L.tpx_last_error.restype = ctypes.c_char_p
    L.tpx_create_error.restype = ctypes.c_char_p
    L.tpx_upload_field.argtypes = [P, I, I, DP, I]
    L.tpx_download_field.argtypes = [P, I, I, DP, I]
    L.tpx_upload_particles.argtypes = [P, I64] + [DP] * 7 + [IP, DP, DP]
    L.tpx_num_particles.argtypes = [P, ctypes.POINTER(I64)]
    L.tpx_download_particles.argtypes = [P, I] + [DP] * 7 + [IP]
    L.tpx_download_particle_rhs_t.argtypes = [P, DP]
    L.tpx_set_restart_numbers.argtypes = [P, D, D, I]
    L.tpx_get_scalars.argtypes = [P, DP]
    for fn in ("tpx_initialize", "tpx_params_update", "tpx_time_advance_rk4", "tpx_time_advance",
               "tpx_deposit_concentration", "tpx_synchronize"):
        getattr(L, fn).argtypes = [P]
    for fn in ("tpx_update_rv_woq", "tpx_compute_rhs_pois_q", "tpx_update_rv_wq"):
        getattr(L, fn).argtypes = [P, I]
    L.tpx_solve_poisson_q.argtypes = [P, I, ctypes.POINTER(I)]
    for fn in ("tpx_update_rho", "tpx_update_p0", "tpx_update_particle", "tpx_update_passive_scalar", "tpx_update_ru_wop",
               "tpx_compute_div_u_new", "tpx_compute_rhs_pois", "tpx_update_ru_wp", "tpx_step"):
        getattr(L, fn).argtypes = [P, I]
    L.tpx_solve_poisson.argtypes = [P, I, ctypes.POINTER(I)]
    L.tpx_poisson_solve.argtypes = [P, DP, DP, DP, DP, DP, I, ctypes.POINTER(I)]
    L.tpx_ensemble_step.argtypes = [ctypes.POINTER(P), I, I, I]
    L.tpx_adjoint_record.argtypes = [P, I, I, DP]
    L.tpx_adjoint_gradient.argtypes = [P, DP, DP]
    L.tpx_adjoint_gradient_full.argtypes = [P, DP, DP, DP, DP, DP, DP]
    L.tpx_adjoint_gradient_heat.argtypes = [P, DP, DP, DP]
    L.tpx_adjoint_set_gas_tape.argtypes = [P, I]
    L.tpx_statistics_enable.argtypes = [P, I]
    L.tpx_tune_tiles.argtypes = [P]
    L.tpx_get_tiles.argtypes = [P, IP]
    L.tpx_set_tiles.argtypes = [P, IP]
    L.tpx_statistics.argtypes = [P, DP]
    L.tpx_track_particles.argtypes = [P, I, I, DP]
    L.tpx_adjoint_release.argtypes = [P]
    L.tpx_adjoint_check.argtypes = [P, I, DP, DP, DP, DP, DP]
    L.tpx_batch_create.argtypes = [ctypes.POINTER(P), I, ctypes.POINTER(P)]
    L.tpx_batch_step.argtypes = [P, I]
    L.tpx_batch_destroy.argtypes = [P]
    L.tpx_batch_error.restype = ctypes.c_char_p
    L.tpx_batch_last_error.argtypes = [P]
    L.tpx_batch_last_error.restype = ctypes.c_char_p
    L.tpx_make_mean_u0.argtypes = [P, DP]
    L.tpx_field_sum.argtypes = [P, I, I, DP]
    L.tpx_field_max_abs.argtypes = [P, I, I, DP]
    L.tpx_get_stream.argtypes = [P, ctypes.POINTER(P)]
    L.tpx_timing_enable.argtypes = [P, I]
    L.tpx_timing_read.argtypes = [P, DP, ctypes.POINTER(I64)]
    _lib = L
    return L


def _dp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _c(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


class Solver:
    """One device context: grid fields, particles and the step."""

    def __init__(self, params: Params, np_capacity: int = 0):
        self.L = load()
        self.params = params
        self.st = TpxParams(**params.struct_fields())
        self.shape = (params.Nz_tot, params.Ny_tot, params.Nx_tot)
        h = ctypes.c_void_p()
        rc = self.L.tpx_create(ctypes.byref(self.st), int(np_capacity), ctypes.byref(h))
        if rc != TPX_OK:
            raise TpxError(f"tpx_create failed ({rc}): {self.L.tpx_create_error().decode()}")
        self.h = h

    def _ok(self, rc: int, what: str) -> None:
        if rc != TPX_OK:
            msg = self.L.tpx_last_error(self.h).decode()
            raise TpxError(f"{what} failed ({rc}): {msg}")

    def close(self) -> None:
        if getattr(self, "h", None):
            self.L.tpx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- fields ----
    def set(self, name: str, arr: np.ndarray) -> None:
        fid = FIELD[name]
        if name in VECTOR:
            arr = _c(arr).reshape((3,) + self.shape)
            for c in range(3):
                a = np.ascontiguousarray(arr[c])
                self._ok(self.L.tpx_upload_field(self.h, fid, c, _dp(a), 0), f"upload {name}")
        else:
            a = _c(arr).reshape(self.shape)
            self._ok(self.L.tpx_upload_field(self.h, fid, 0, _dp(a), 0), f"upload {name}")

    def get(self, name: str, bs: int = 0) -> np.ndarray:
        fid = FIELD[name]
        sh = tuple(n + 2 * bs for n in self.shape)
        if name in VECTOR:
            out = np.empty((3,) + sh)
            for c in range(3):
                a = np.empty(sh)
                self._ok(self.L.tpx_download_field(self.h, fid, c, _dp(a), bs), f"download {name}")
                out[c] = a
            return out
        out = np.empty(sh)
        self._ok(self.L.tpx_download_field(self.h, fid, 0, _dp(out), bs), f"download {name}")
        return out

    # ---- particles ----
    def set_particles(self, x, y, z, u, v, w, T, flag, tau=None, eps=None) -> None:
        arrs = [_c(a) for a in (x, y, z, u, v, w, T)]
        fl = np.ascontiguousarray(flag, dtype=np.int32)
        tau = None if tau is None else _c(tau)
        eps = None if eps is None else _c(eps)
        rc = self.L.tpx_upload_particles(self.h, len(arrs[0]), *[_dp(a) for a in arrs],
                                         fl.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _dp(tau), _dp(eps))
        self._ok(rc, "upload particles")

    @property
    def np(self) -> int:
        n = ctypes.c_int64()
        self._ok(self.L.tpx_num_particles(self.h, ctypes.byref(n)), "num_particles")
        return n.value

    def particles(self, stage: str = "base") -> Dict[str, np.ndarray]:
        n = self.np
        qs = [np.empty(n) for _ in range(7)]
        fl = np.empty(n, dtype=np.int32)
        self._ok(self.L.tpx_download_particles(self.h, STAGE[stage], *[_dp(a) for a in qs],
                                               fl.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))),
                 "download particles")
        out = dict(zip("x y z u v w T".split(), qs))
        out["flag"] = fl
        return out

    def rhs_t(self) -> np.ndarray:
        out = np.empty(self.np)
        self._ok(self.L.tpx_download_particle_rhs_t(self.h, _dp(out)), "download RHS_T")
        return out

    # ---- scalars ----
    def set_restart_numbers(self, P0: float, T_cur: float = 0.0, num_timestep: int = 0) -> None:
        self._ok(self.L.tpx_set_restart_numbers(self.h, P0, T_cur, num_timestep), "restart numbers")

    def scalars(self) -> Dict[str, float]:
        buf = np.empty(len(SCALARS))
        self._ok(self.L.tpx_get_scalars(self.h, _dp(buf)), "get scalars")
        return dict(zip(SCALARS, buf.tolist()))

    # ---- methods ----
    def initialize(self): self._ok(self.L.tpx_initialize(self.h), "initialize")
    def params_update(self): self._ok(self.L.tpx_params_update(self.h), "params_update")
    def update_rho(self, rk): self._ok(self.L.tpx_update_rho(self.h, rk), "update_rho")
    def update_p0(self, rk): self._ok(self.L.tpx_update_p0(self.h, rk), "update_p0")
    def update_particle(self, rk): self._ok(self.L.tpx_update_particle(self.h, rk), "update_particle")
    def update_passive_scalar(self, rk):
        self._ok(self.L.tpx_update_passive_scalar(self.h, rk), "update_passive_scalar")
    def update_rv_woq(self, rk): self._ok(self.L.tpx_update_rv_woq(self.h, rk), "update_rv_woq")
    def compute_rhs_pois_q(self, rk): self._ok(self.L.tpx_compute_rhs_pois_q(self.h, rk), "compute_rhs_pois_q")

    def solve_poisson_q(self, rk) -> int:
        it = ctypes.c_int(0)
        self._ok(self.L.tpx_solve_poisson_q(self.h, rk, ctypes.byref(it)), "solve_poisson_q")
        return it.value

    def update_rv_wq(self, rk): self._ok(self.L.tpx_update_rv_wq(self.h, rk), "update_rv_wq")
    def update_ru_wop(self, rk): self._ok(self.L.tpx_update_ru_wop(self.h, rk), "update_ru_wop")
    def compute_div_u_new(self, rk): self._ok(self.L.tpx_compute_div_u_new(self.h, rk), "compute_div_u_new")
    def compute_rhs_pois(self, rk): self._ok(self.L.tpx_compute_rhs_pois(self.h, rk), "compute_rhs_pois")

    def solve_poisson(self, rk) -> int:
        it = ctypes.c_int(0)
        self._ok(self.L.tpx_solve_poisson(self.h, rk, ctypes.byref(it)), "solve_poisson")
        return it.value

    def poisson_solve(self, C, P, RHS):
        """poisson::Solve(C, P, RHS) with the caller's face coefficient C
        (3, nz, ny, nx): returns the new P and the iteration count."""
        C = _c(C).reshape((3,) + self.shape)
        cs = [np.ascontiguousarray(C[c]) for c in range(3)]
        Pn = _c(P).reshape(self.shape).copy()
        R = _c(RHS).reshape(self.shape)
        it = ctypes.c_int(0)
        self._ok(self.L.tpx_poisson_solve(self.h, *[_dp(a) for a in cs], _dp(Pn), _dp(R), 0, ctypes.byref(it)),
                 "poisson_solve")
        return Pn, it.value

    def update_ru_wp(self, rk): self._ok(self.L.tpx_update_ru_wp(self.h, rk), "update_ru_wp")
    def time_advance_rk4(self): self._ok(self.L.tpx_time_advance_rk4(self.h), "time_advance_rk4")
    def time_advance(self): self._ok(self.L.tpx_time_advance(self.h), "time_advance")

    def make_mean_u0(self, a: Sequence[float]) -> None:
        arr = _c(a)
        self._ok(self.L.tpx_make_mean_u0(self.h, _dp(arr)), "make_mean_u0")

    def step(self, n: int = 1) -> None:
        self._ok(self.L.tpx_step(self.h, int(n)), "step")

    # ---- adjoint of the step (no heat exchange; one- or two-way coupling) ----
    def adjoint_record(self, nsteps: int, every: int = 10) -> float:
        """Advance nsteps whole steps (as step()) while taping what the
        adjoint needs; returns J = Var(C) of the final positions."""
        out = ctypes.c_double()
        self._ok(self.L.tpx_adjoint_record(self.h, int(nsteps), int(every), ctypes.byref(out)), "adjoint_record")
        return out.value

    def adjoint_gradient(self, state: bool = False):
        """dJ/dtau_p of the recorded run, in particle order; with state=True
        also dJ/d(x, y, z, u, v, w) of the particles when recording began,
        shape (6, np)."""
        n = ctypes.c_int64()
        self._ok(self.L.tpx_num_particles(self.h, ctypes.byref(n)), "num_particles")
        g = np.zeros(n.value)
        gs = np.zeros((6, n.value)) if state else None
        self._ok(self.L.tpx_adjoint_gradient(self.h, _dp(g), _dp(gs) if state else None), "adjoint_gradient")
        return (g, gs) if state else g

    def adjoint_gradient_full(self) -> dict:
        """Every gradient of the recorded run: "tau" (np), "state" (6, np)
        in particle order, "A" (the forcing coefficient, a float), and the
        initial gas "RU" (3, nz, ny, nx), "Rho" and "P" (nz, ny, nx).  Leaves
        the gas where the recorded run ended."""
        n = ctypes.c_int64()
        self._ok(self.L.tpx_num_particles(self.h, ctypes.byref(n)), "num_particles")
        out = dict(tau=np.zeros(n.value), state=np.zeros((6, n.value)), A=np.zeros(1),
                   RU=np.zeros((3,) + self.shape), Rho=np.zeros(self.shape), P=np.zeros(self.shape))
        self._ok(self.L.tpx_adjoint_gradient_full(self.h, _dp(out["tau"]), _dp(out["state"]), _dp(out["A"]),
                                                  _dp(out["RU"]), _dp(out["Rho"]), _dp(out["P"])),
                 "adjoint_gradient_full")
        out["A"] = float(out["A"][0])
        if self.heat:
            out.update(self.adjoint_gradient_heat())
        return out

    @property
    def heat(self) -> bool:
        """The heat exchange is on (Nu pi Dp k != 0)."""
        p = self.params
        return p.Nu * p.Dp * p.k != 0.0

    def adjoint_gradient_heat(self) -> dict:
        """For a recorded run with the heat exchange: "T" and "eps" (np, in
        particle order), dJ/d of each particle's initial temperature and of
        its absorption efficiency, and "P0" (a float), dJ/d of the initial
        thermodynamic pressure; from the last sweep (one runs if none has
        since the recording)."""
        n = ctypes.c_int64()
        self._ok(self.L.tpx_num_particles(self.h, ctypes.byref(n)), "num_particles")
        out = dict(T=np.zeros(n.value), eps=np.zeros(n.value), P0=np.zeros(1))
        self._ok(self.L.tpx_adjoint_gradient_heat(self.h, _dp(out["T"]), _dp(out["eps"]), _dp(out["P0"])),
                 "adjoint_gradient_heat")
        out["P0"] = float(out["P0"][0])
        return out

    def adjoint_set_gas_tape(self, mode: str = "auto") -> None:
        """How the next recording keeps the gas for the gas sweep: "auto"
        (keep every substep when it fits in half the free device memory),
        "keep", "recompute" (each step from its start), or "segments" (the
        gas checkpointed with the particles every `every` steps, each
        segment replayed from its checkpoint in the sweep)."""
        code = {"auto": 0, "keep": 1, "recompute": 2, "segments": 3}[mode]
        self._ok(self.L.tpx_adjoint_set_gas_tape(self.h, code), "adjoint_set_gas_tape")

    def adjoint_release(self) -> None:
        self._ok(self.L.tpx_adjoint_release(self.h), "adjoint_release")

    def adjoint_check(self, op: int, x: np.ndarray, y: np.ndarray):
        """Transpose checks of the Poisson adjoint (tpx_adjoint_check): op 0
        gives (M x, M y); op 1 gives (P_1(x), the adjoint of P for y, the
        adjoint of Rho for y) for one iteration from P = x with a zero
        right-hand side at the solver's Rho."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(self.shape)
        y = np.ascontiguousarray(y, dtype=np.float64).reshape(self.shape)
        ax, aty, atr = np.zeros(self.shape), np.zeros(self.shape), np.zeros(self.shape)
        self._ok(self.L.tpx_adjoint_check(self.h, op, _dp(x), _dp(y), _dp(ax), _dp(aty), _dp(atr)), "adjoint_check")
        return (ax, aty) if op == 0 else (ax, aty, atr)

    # ---- deposit tile shapes ----
    def tune_tiles(self, max_steps: int = 12) -> dict:
        """Tune the deposit kernels' tile shapes inside real steps (they
        advance the run; the results do not depend on the shape) and return
        the chosen shapes."""
        self._ok(self.L.tpx_tune_tiles(self.h), "tune_tiles")
        for _ in range(max_steps):
            self.step(1)
            if not self.tiles()["tuning"]:
                break
        return self.tiles()

    def tiles(self) -> dict:
        out = np.zeros(5, dtype=np.int32)
        self._ok(self.L.tpx_get_tiles(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))), "get_tiles")
        return {"heat": (int(out[0]), int(out[1])), "velocity": (int(out[2]), int(out[3])), "tuning": bool(out[4])}

    def set_tiles(self, heat=(0, 0), velocity=(0, 0)) -> None:
        arr = np.array([heat[0], heat[1], velocity[0], velocity[1]], dtype=np.int32)
        self._ok(self.L.tpx_set_tiles(self.h, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))), "set_tiles")

    # ---- grid::Statistics on the device ----
    def statistics_enable(self, on: bool = True) -> None:
        """Keep what the statistics read from inside each step's last
        substep (call before stepping)."""
        self._ok(self.L.tpx_statistics_enable(self.h, int(bool(on))), "statistics_enable")

    def statistics(self) -> dict:
        """grid::Statistics at the current step boundary (tpx.STATS names)."""
        out = np.zeros(len(STATS))
        self._ok(self.L.tpx_statistics(self.h, _dp(out)), "statistics")
        return dict(zip(STATS, out.tolist()))

    def track_particles(self, stride: int, count: int) -> np.ndarray:
        """particle::trajectory's positions: row m is (x, y, z) of the particle
        with flag m*stride (NaN if none)."""
        out = np.empty((count, 3))
        if count:
            self._ok(self.L.tpx_track_particles(self.h, int(stride), int(count), _dp(out)), "track_particles")
        return out

    def deposit_concentration(self) -> None:
        self._ok(self.L.tpx_deposit_concentration(self.h), "deposit_concentration")

    def field_sum(self, name: str, comp: int = 0) -> float:
        out = ctypes.c_double()
        self._ok(self.L.tpx_field_sum(self.h, FIELD[name], comp, ctypes.byref(out)), "field_sum")
        return out.value

    def field_max_abs(self, name: str, comp: int = 0) -> float:
        out = ctypes.c_double()
        self._ok(self.L.tpx_field_max_abs(self.h, FIELD[name], comp, ctypes.byref(out)), "field_max_abs")
        return out.value

    # ---- timing ----
    def stream(self) -> int:
        s = ctypes.c_void_p()
        self._ok(self.L.tpx_get_stream(self.h, ctypes.byref(s)), "get_stream")
        return s.value or 0

    def timing(self, on: bool) -> None:
        self._ok(self.L.tpx_timing_enable(self.h, 1 if on else 0), "timing_enable")

    def timing_read(self):
        ms = np.zeros(len(KCLASS))
        cnt = np.zeros(len(KCLASS), dtype=np.int64)
        self._ok(self.L.tpx_timing_read(self.h, _dp(ms), cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))),
                 "timing_read")
        return dict(zip(KCLASS, ms.tolist())), dict(zip(KCLASS, cnt.tolist()))

    def synchronize(self) -> None:
        self._ok(self.L.tpx_synchronize(self.h), "synchronize")

    # ---- convenience ----
    def load_case(self, case) -> None:
        """Upload a synthetic case (tpx.ic.Case) and run Initialize's tail."""
        self.set("RU", case.RU)
        self.set("Rho", case.Rho)
        self.set("P", case.P)
        pp = case.particles
        self.set_particles(pp["x"], pp["y"], pp["z"], pp["u"], pp["v"], pp["w"], pp["T"], pp["flag"],
                           tau=case.tau, eps=case.eps)
        if getattr(case, "PS", None) is not None:
            self.set("Passive_Scalar", case.PS)
        if getattr(case, "RV", None) is not None:
            self.set("RV", case.RV)
            self.set("Q", getattr(case, "Q", None) if getattr(case, "Q", None) is not None else np.zeros(self.shape))
        self.set_restart_numbers(case.params.P0, 0.0, 0)
        self.initialize()

    SUBSTEP_METHODS = ("update_rho", "update_p0", "update_particle", "update_passive_scalar",
                       "update_rv_woq", "compute_rhs_pois_q", "solve_poisson_q", "update_rv_wq", "update_ru_wop",
                       "compute_div_u_new", "compute_rhs_pois", "solve_poisson",
                       "update_ru_wp", "time_advance_rk4")

    def method_step(self) -> None:
        """One step driven method by method, in simulation.cpp:57-91's order."""
        p = self.params
        self.params_update()
        for rk in range(4):
            for m in self.SUBSTEP_METHODS:
                if m == "update_passive_scalar" and not p.solve_for_scalar:
                    continue  # simulation.cpp:65
                if m in ("update_rv_woq", "compute_rhs_pois_q", "solve_poisson_q", "update_rv_wq") and \
                        not p.solve_for_vector:
                    continue  # simulation.cpp:66-71
                fn = getattr(self, m)
                if m == "time_advance_rk4":
                    fn()
                else:
                    fn(rk)
        self.make_mean_u0([p.U0 * p.Rho0, p.V0 * p.Rho0, p.W0 * p.Rho0])
        self.time_advance()


# ---------------------------------------------------------------------------
# ensembles of independent boxes
# ---------------------------------------------------------------------------
def ensemble_step(solvers, nsteps: int = 1, concurrency: int = 0) -> None:
    """Advance every solver nsteps whole steps (each exactly as Solver.step),
    with up to `concurrency` boxes in flight at once (0: the library's
    default)."""
    L = load()
    arr = (ctypes.c_void_p * len(solvers))(*[s.h for s in solvers])
    rc = L.tpx_ensemble_step(arr, len(solvers), int(nsteps), int(concurrency))
    if rc != 0:
        msgs = [m for m in ((L.tpx_last_error(s.h) or b"").decode() for s in solvers) if m]
        raise TpxError(f"ensemble_step failed ({rc}): {msgs[0] if msgs else 'invalid arguments'}")


class Batch:
    """Boxes stepped by batched kernels: every kernel of a step is launched
    once for all boxes (tpx_batch_*).  The solvers keep their own state and
    stay usable between batch steps; each box advances exactly as its own
    Solver.step would advance it."""

    def __init__(self, solvers):
        L = load()
        self.solvers = list(solvers)
        arr = (ctypes.c_void_p * len(self.solvers))(*[s.h for s in self.solvers])
        h = ctypes.c_void_p()
        rc = L.tpx_batch_create(arr, len(self.solvers), ctypes.byref(h))
        if rc != 0:
            raise TpxError(f"batch_create failed ({rc}): {(L.tpx_batch_error() or b'').decode()}")
        self.h = h

    def step(self, nsteps: int = 1) -> None:
        L = load()
        rc = L.tpx_batch_step(self.h, int(nsteps))
        if rc != 0:
            raise TpxError(f"batch_step failed ({rc}): {(L.tpx_batch_last_error(self.h) or b'').decode()}")

    def close(self) -> None:
        if getattr(self, "h", None):
            load().tpx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tune_concurrency(solvers, candidates=(2, 4, 8, 16, 32), steps: int = 1) -> int:
    """Pick the number of boxes in flight that gives the highest throughput:
    each candidate runs `steps` whole steps of the ensemble (advancing it)."""
    import time
    best, best_t = None, None
    for c in candidates:
        if c > len(solvers) and best is not None:
            break
        for s in solvers:
            s.synchronize()
        t0 = time.perf_counter()
        ensemble_step(solvers, steps, c)
        t = time.perf_counter() - t0
        if best_t is None or t < best_t:
            best, best_t = c, t
    return best
