"""torch.autograd harness around the discrete adjoint of the time step: the
Nusselt number after a run of time steps as a differentiable function of
the hexagon wall temperature and the Rayleigh number. The forward runs the
HIP time step; backward returns the adjoint gradient (nusselt_gradient)."""
import numpy as np
import torch

from . import Context, nusselt_gradient
from . import host


def make_context(case, amp=1e-3, seed=20221):
    """A context of the case with its geometry, pressure coefficients and
    seeded initial state uploaded."""
    ctx = Context(*case.dims, case)
    ctx.upload_geometry(host.build_geometry(case))
    ctx.build_pressure_coeffs()
    ctx.upload_state(host.initial_state(case, amp=amp, seed=seed))
    return ctx


class NusseltRun(torch.autograd.Function):
    """J(T_hot, Ra) = weights . Nu after `nsteps` steps of `case` with those
    two values (or its mean over the steps, run["average"]), from the seeded
    initial state. run holds case, nsteps, nslots, adj_tol, passes (or tol),
    weights, average, solver, amp, seed."""

    @staticmethod
    def forward(fctx, T_hot, Ra, run):
        case = run["case"].replace(T_hot=float(T_hot), Ra=float(Ra))
        ctx = make_context(case, run.get("amp", 1e-3), run.get("seed", 20221))
        try:
            J, grad, info = nusselt_gradient(ctx, case, run["nsteps"], run["nslots"], run["adj_tol"],
                                             tol=run.get("tol"), passes=run.get("passes"),
                                             weights=run.get("weights", (0.0, 1.0, 0.0)),
                                             average=run.get("average", False), solver=run.get("solver", "gmres"))
        finally:
            ctx.close()
        fctx.grad = (grad["T_hot"], grad["Ra"])
        fctx.info = info
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(fctx, gJ):
        gT, gR = fctx.grad
        return gJ * gT, gJ * gR, None


def nusselt(T_hot, Ra, run):
    return NusseltRun.apply(T_hot, Ra, run)
