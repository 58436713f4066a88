"""torch shims over the adjoint of the C ABI.

`var_c(tau, make_solver, nsteps, every)` is a differentiable function of the
per-particle response times: it builds a solver whose particles have response
times `tau` (make_solver(tau_numpy) -> tpx.Solver, loaded and initialised),
advances it `nsteps` whole steps (no heat exchange) and returns
J = Var(C) of the final positions; J.backward() fills tau.grad with the
hand-written adjoint's dJ/dtau (tpx_adjoint_gradient).  The solver's
particle ids must be the indices of `tau` (as tpx.ic builds them): the
gradient comes back in the solver's particle order and is put back in the
order of `tau` through the ids.

`var_c_flow(ru, make_solver, nsteps, every)` is J as a differentiable
function of the initial momentum field RU (shape (3, nz, ny, nx)):
make_solver(ru_numpy) -> tpx.Solver, loaded and initialised; J.backward()
fills ru.grad with the gas adjoint's dJ/dRU (tpx_adjoint_gradient_full)."""
from __future__ import annotations

import numpy as np


def _torch():
    import torch
    return torch


class _VarC:
    """Built lazily so that importing tpx needs no torch."""
    fn = None

    @classmethod
    def get(cls):
        if cls.fn is None:
            torch = _torch()

            class VarC(torch.autograd.Function):
                @staticmethod
                def forward(ctx, tau, make_solver, nsteps, every):
                    s = make_solver(tau.detach().cpu().numpy().astype(np.float64))
                    J = s.adjoint_record(int(nsteps), int(every))
                    ctx.solver = s
                    ctx.tau_device, ctx.tau_dtype = tau.device, tau.dtype
                    return torch.tensor(J, dtype=torch.float64)

                @staticmethod
                def backward(ctx, gJ):
                    s = ctx.solver
                    g_solver = s.adjoint_gradient()
                    g = np.empty_like(g_solver)
                    g[s.particles()["flag"]] = g_solver
                    s.adjoint_release()
                    ctx.solver = None
                    grad = torch.from_numpy(g).to(device=ctx.tau_device, dtype=ctx.tau_dtype) * gJ
                    return grad, None, None, None

            cls.fn = VarC
        return cls.fn


def var_c(tau, make_solver, nsteps: int, every: int = 10):
    return _VarC.get().apply(tau, make_solver, nsteps, every)


class _VarCFlow:
    fn = None

    @classmethod
    def get(cls):
        if cls.fn is None:
            torch = _torch()

            class VarCFlow(torch.autograd.Function):
                @staticmethod
                def forward(ctx, ru, make_solver, nsteps, every):
                    s = make_solver(ru.detach().cpu().numpy().astype(np.float64))
                    J = s.adjoint_record(int(nsteps), int(every))
                    ctx.solver = s
                    ctx.ru_device, ctx.ru_dtype = ru.device, ru.dtype
                    return torch.tensor(J, dtype=torch.float64)

                @staticmethod
                def backward(ctx, gJ):
                    s = ctx.solver
                    g = s.adjoint_gradient_full()["RU"]
                    s.adjoint_release()
                    ctx.solver = None
                    grad = torch.from_numpy(g).to(device=ctx.ru_device, dtype=ctx.ru_dtype) * gJ
                    return grad, None, None, None

            cls.fn = VarCFlow
        return cls.fn


def var_c_flow(ru, make_solver, nsteps: int, every: int = 10):
    return _VarCFlow.get().apply(ru, make_solver, nsteps, every)
