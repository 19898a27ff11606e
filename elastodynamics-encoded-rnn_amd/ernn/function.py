"""torch.autograd.Function over the HIP forward and adjoint.

Replaces the autograd tape the reference builds through every time step
(Parameter_inversion/Inversion_adam.py:268-301, backward at :447;
Damage_inversion/Inversion_adam.py:355-387, backward at :511): forward makes
one C-ABI call that runs the whole recurrence and stores checkpoints in a
workspace; backward makes one C-ABI call that recomputes segment by segment
and runs the reverse-time adjoint with FP64 gradient accumulation.
"""
from __future__ import annotations

import torch

from .engine import Engine


class ERNNFunction(torch.autograd.Function):
    """traces = ERNN(params, field; wavelets).

    params   (n_sets*6,) fp32 stiffness constants (const6, dfield) or an empty tensor
    field    flat fp32 per-cell fields (dfield: D, field6: six C fields) or an empty tensor
    wavelets flat fp32 source time functions (not differentiated)
    """

    @staticmethod
    def forward(ctx, engine: Engine, params: torch.Tensor, field: torch.Tensor, wavelets: torch.Tensor):
        p = params if engine.mode != 2 else None
        f = field if engine.mode != 0 else None
        ws = engine.acquire_workspace()
        traces = engine.forward(p, f, wavelets, ws)
        if ctx.needs_input_grad[1] and engine.fixed_c0:
            raise RuntimeError("this plan returns dJ/dD only (Survey.fixed_c0): the constants must not require grad")
        needs = ctx.needs_input_grad[1] or ctx.needs_input_grad[2]
        if needs:
            ctx.engine = engine
            ctx.ws = ws
            ctx.save_for_backward(params, field, wavelets)
        else:
            engine.release_workspace(ws)
        return traces

    @staticmethod
    def backward(ctx, grad_traces: torch.Tensor):
        engine: Engine = ctx.engine
        params, field, wavelets = ctx.saved_tensors
        p = params if engine.mode != 2 else None
        f = field if engine.mode != 0 else None
        gp, gf = engine.adjoint(p, f, wavelets, grad_traces.to(torch.float32).contiguous(), ctx.ws)
        engine.release_workspace(ctx.ws)
        ctx.ws = None
        g_params = gp.to(params.dtype) if (gp is not None and ctx.needs_input_grad[1]) else None
        g_field = gf.to(field.dtype) if (gf is not None and ctx.needs_input_grad[2]) else None
        return None, g_params, g_field, None


def ernn_traces(engine: Engine, params=None, field=None, wavelets=None) -> torch.Tensor:
    """Differentiable flat traces for the engine's survey."""
    dev = engine.device
    empty = torch.empty(0, dtype=torch.float32, device=dev)
    return ERNNFunction.apply(engine, params if params is not None else empty,
                              field if field is not None else empty, wavelets)
