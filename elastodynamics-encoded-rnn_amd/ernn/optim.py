"""Total-variation regularised Adam on the HIP path.

The damage inversion (Damage_inversion/Inversion_adam.py, DI) trains six
log-stiffness fields theta_p, C_p = clamp(exp(theta_p), bounds_p)
(DI:304-322), on  loss = data + lambda_tv * sum_p TV(C_p)  (DI:179-195,
DI:441-463) with torch.optim.Adam (DI:478-483).  FusedTVAdam takes the data
gradient left in theta.grad by the data loss's backward and, in one C-ABI
call (ernn_tv_adam_step), adds the TV term's gradient through the clamp and
exp and applies Adam with torch's arithmetic: the TV forward, its autograd
backward and the optimizer's ~10 foreach kernels become two passes over the
fields.  There is no PyTorch fallback.
"""
from __future__ import annotations

import ctypes as C
from typing import Iterable, List, Optional, Sequence, Tuple

import torch

from . import _lib


class FusedTVAdam:
    """torch.optim.Adam's interface (params, lr, betas, eps, weight_decay;
    zero_grad, step, state_dict) plus the TV term of DI:453-460.

    params: up to 8 float32 CUDA tensors of one shape (..., nx, ny), the
    trainable fields.  bounds: per field (lo, hi) of the clamp (None = none).
    log_params: C = clamp(exp(theta)) as DI:307-313, else clamp(theta).
    step() expects theta.grad = d(data loss)/d theta; afterwards theta.grad
    holds the gradient of the whole loss and `last_tv` the per-field TV(C_p)
    of the fields before the update (loss_tv = lambda_tv * sum)."""

    def __init__(self, params: Iterable[torch.Tensor], lr: float = 1e-3, betas: Tuple[float, float] = (0.9, 0.999),
                 eps: float = 1e-8, weight_decay: float = 0.0, lambda_tv: float = 0.0,
                 bounds: Optional[Sequence[Optional[Tuple[float, float]]]] = None, log_params: bool = True):
        self.params: List[torch.Tensor] = list(params)
        n = len(self.params)
        if not 1 <= n <= _lib.TV_MAX_FIELDS:
            raise ValueError(f"FusedTVAdam takes 1 .. {_lib.TV_MAX_FIELDS} fields, got {n}")
        shape = self.params[0].shape
        for p in self.params:
            if p.dtype != torch.float32 or p.device.type != "cuda" or not p.is_contiguous() or p.shape != shape:
                raise ValueError("fields must be contiguous float32 CUDA tensors of one shape")
        if len(shape) < 2 or int(torch.tensor(shape[:-2]).prod()) != 1:
            raise ValueError("fields must be (nx, ny) or (1, nx, ny)")
        self.nx, self.ny = int(shape[-2]), int(shape[-1])
        self.lr, self.betas, self.eps, self.weight_decay = float(lr), tuple(map(float, betas)), float(eps), float(weight_decay)
        self.lambda_tv = float(lambda_tv)
        self.log_params = bool(log_params)
        bounds = list(bounds) if bounds is not None else [None] * n
        if len(bounds) != n:
            raise ValueError("one (lo, hi) or None per field")
        inf = float("inf")
        self.bounds = [(-inf, inf) if b is None else (float(b[0]), float(b[1])) for b in bounds]
        self.state = {"step": 0, "exp_avg": [torch.zeros_like(p) for p in self.params],
                      "exp_avg_sq": [torch.zeros_like(p) for p in self.params]}
        self._tv = torch.zeros(n, dtype=torch.float64, device=self.params[0].device)
        d = self._desc(1)
        nbytes = C.c_int64()
        _lib.check(_lib.lib().ernn_tv_adam_workspace_bytes(C.byref(d), C.byref(nbytes)),
                   "ernn_tv_adam_workspace_bytes")
        self._ws = torch.empty(max(8, nbytes.value), dtype=torch.uint8, device=self.params[0].device)

    def _desc(self, step: int) -> _lib.TvAdamDesc:
        d = _lib.TvAdamDesc()
        d.nx, d.ny, d.n_fields, d.log_params, d.step = self.nx, self.ny, len(self.params), int(self.log_params), step
        d.lambda_tv, d.lr, d.eps, d.weight_decay = self.lambda_tv, self.lr, self.eps, self.weight_decay
        d.beta1, d.beta2 = self.betas
        for f, (lo, hi) in enumerate(self.bounds):
            d.lo[f], d.hi[f] = lo, hi
        return d

    def zero_grad(self, set_to_none: bool = True) -> None:
        for p in self.params:
            if set_to_none:
                p.grad = None
            elif p.grad is not None:
                p.grad.zero_()

    @torch.no_grad()
    def step(self) -> None:
        for p in self.params:
            if p.grad is None:
                p.grad = torch.zeros_like(p)
            elif not p.grad.is_contiguous() or p.grad.dtype != torch.float32:
                p.grad = p.grad.contiguous().float()
        self.state["step"] += 1
        d = self._desc(self.state["step"])
        c = _lib.TvAdamCall()
        for f, p in enumerate(self.params):
            c.theta[f] = p.data_ptr()
            c.grad[f] = p.grad.data_ptr()
            c.m[f] = self.state["exp_avg"][f].data_ptr()
            c.v[f] = self.state["exp_avg_sq"][f].data_ptr()
        c.tv = self._tv.data_ptr()
        c.workspace = self._ws.data_ptr()
        dev = self.params[0].device
        with torch.cuda.device(dev):
            _lib.check(_lib.lib().ernn_tv_adam_step(C.byref(d), C.byref(c), C.c_void_p(_lib.stream_handle(dev))),
                       "ernn_tv_adam_step")

    @property
    def last_tv(self) -> torch.Tensor:
        """Per-field TV(C_p) of the fields before the last step (FP64)."""
        return self._tv

    def state_dict(self) -> dict:
        return {"step": self.state["step"], "exp_avg": [t.clone() for t in self.state["exp_avg"]],
                "exp_avg_sq": [t.clone() for t in self.state["exp_avg_sq"]]}

    def load_state_dict(self, sd: dict) -> None:
        self.state["step"] = int(sd["step"])
        for dst, src in zip(self.state["exp_avg"], sd["exp_avg"]):
            dst.copy_(src)
        for dst, src in zip(self.state["exp_avg_sq"], sd["exp_avg_sq"]):
            dst.copy_(src)
