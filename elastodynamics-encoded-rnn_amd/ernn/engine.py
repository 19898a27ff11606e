"""Engine: one C-ABI plan (libernn_hip.so) bound to torch device memory.

An Engine owns the plan built from a Survey and a small pool of workspaces.
All buffers it passes to the library are torch CUDA tensors, so PyTorch's
caching allocator sees every byte; the library itself allocates only its
tables at plan creation.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Tuple

import numpy as np
import torch

from . import _lib
from .problem import MODES, Survey

# reversed adjoint shape of a reverse_u survey that names none
DEFAULT_REV_SHAPE = (1, 4, 4)


def _ptr(t: Optional[torch.Tensor]) -> Optional[int]:
    return None if t is None else t.data_ptr()


class Engine:
    """Plan + workspaces for one Survey on one CUDA device."""

    def __init__(self, survey: Survey, device: Optional[torch.device] = None, ckpt_every: Optional[int] = None):
        if not torch.cuda.is_available():
            raise RuntimeError("ernn.Engine needs a ROCm GPU: the HIP kernels are the only implementation")
        self.survey = survey
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        if self.device.type != "cuda":
            raise RuntimeError(f"ernn.Engine needs a CUDA (ROCm) device, got {self.device}")
        self.mode = MODES[survey.mode]
        self.n_coef = {0: 0, 1: 1, 2: 6}[self.mode]
        K = ckpt_every or survey.ckpt_every
        if not K:
            from .autotune import choose_ckpt
            budget = survey.ckpt_budget
            if budget is None:
                free, _ = torch.cuda.mem_get_info(self.device)
                budget = int(0.8 * free)
                K = choose_ckpt(survey, budget_bytes=budget)
            else:   # a stated budget is a cap: MemoryError if even the least-memory interval exceeds it
                K = choose_ckpt(survey, mem_cap_bytes=budget or None, budget_bytes=budget)
        # The fused backward overlaps each segment's adjoint with the next
        # segment's recompute; with one segment (every state kept) there is no
        # recompute to overlap, so the automatic choice is the unfused one.
        self._fuse = survey.fused and not (survey.fuse is None and survey.nt_max <= K)
        self._build(K)
        self._pool: List[torch.Tensor] = []

    # ------------------------------------------------------------------ plan
    def _build(self, K: int) -> None:
        sv = self.survey
        shots = sv.shots
        n = len(shots)
        inst = (_lib.Instance * n)()
        src_list, rec_list = [], []
        for b, s in enumerate(shots):
            I = inst[b]
            I.nx, I.ny, I.nt, I.set = s.nx, s.ny, s.nt, s.set
            I.h, I.dt, I.rho = s.h, s.dt, s.rho
            I.src_begin, I.src_count = len(src_list), len(s.sources)
            src_list.extend(s.sources)
            if sv.snapshot:
                I.rec_begin, I.rec_count = 0, 0
            else:
                r = np.asarray(s.receivers, dtype=np.int64).reshape(-1, 2)
                I.rec_begin, I.rec_count = len(rec_list), len(r)
                rec_list.extend(map(tuple, r))
        src = (_lib.Source * max(len(src_list), 1))()
        for k, (i, j, w, comp) in enumerate(src_list):
            src[k].i, src[k].j, src[k].w, src[k].comp = int(i), int(j), float(w), int(comp)
        rec = (_lib.Receiver * max(len(rec_list), 1))()
        for k, (i, j) in enumerate(rec_list):
            rec[k].i, rec[k].j = int(i), int(j)
        d = _lib.Desc()
        d.n_inst, d.n_sets, d.param_mode = n, sv.n_sets, self.mode
        d.rec_stride, d.snapshot, d.ckpt_every = sv.rec_stride, int(sv.snapshot), int(K)
        t = tuple(sv.tiles) + (0,) * (6 - len(sv.tiles))
        d.fwd_v, d.fwd_r, d.adj_v, d.adj_r, d.fwd_bpc, d.adj_bpc = t[:6]
        d.use_graph = int(sv.use_graph)
        d.fuse = int(self._fuse)
        d.blk_v, d.blk_rw, d.blk_ks = (int(x) for x in sv.blocking)
        ab = tuple(sv.adj_blocking)
        if sv.reverse_u and ab[2] <= 1:
            ab = DEFAULT_REV_SHAPE
        d.ablk_v, d.ablk_rw, d.ablk_ks = (int(x) for x in ab)
        d.ablk_ff_rw = int(sv.adj_ff_rw)
        d.fixed_c0 = int(bool(sv.fixed_c0))
        d.reverse_u = int(bool(sv.reverse_u))
        d.ortho = int(bool(sv.orthotropic))
        d.inst = inst
        d.src, d.n_src = src, len(src_list)
        d.rec, d.n_rec = rec, len(rec_list)
        handle = C.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().ernn_plan_create(C.byref(d), C.byref(handle)), "ernn_plan_create")
        self._plan = handle
        self.info = _lib.PlanInfo()
        _lib.check(_lib.lib().ernn_plan_info(self._plan, C.byref(self.info)), "ernn_plan_info")
        off = (C.c_int64 * (n + 1))()
        _lib.check(_lib.lib().ernn_trace_offsets(self._plan, off), "ernn_trace_offsets")
        self.trace_off = [int(x) for x in off]

    def close(self) -> None:
        if getattr(self, "_plan", None):
            torch.cuda.synchronize(self.device)
            _lib.lib().ernn_plan_destroy(self._plan)
            self._plan = None
        self._pool.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------- buffers
    @property
    def fixed_c0(self) -> bool:
        """The plan returns dJ/dD only (dfield over known constants)."""
        return bool(self.info.fixed_c0)

    @property
    def reverse_u(self) -> bool:
        """The adjoint reconstructs u backward from the checkpoints (no bank, no recompute)."""
        return bool(self.info.reverse_u)

    @property
    def ortho(self) -> bool:
        """The damage kernels take C16 = C26 = 0 as given (Survey.orthotropic)."""
        return bool(self.info.ortho)

    @property
    def workspace_bytes(self) -> int:
        return int(self.info.workspace_bytes)

    @property
    def trace_floats(self) -> int:
        return int(self.info.trace_floats)

    def acquire_workspace(self) -> torch.Tensor:
        if self._pool:
            return self._pool.pop()
        ws = torch.empty(self.workspace_bytes, dtype=torch.uint8, device=self.device)
        if ws.data_ptr() % 256:
            raise RuntimeError("workspace allocation is not 256-byte aligned")
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().ernn_workspace_init(self._plan, C.c_void_p(ws.data_ptr()),
                                                      C.c_void_p(_lib.stream_handle(self.device))),
                       "ernn_workspace_init")
        return ws

    def release_workspace(self, ws: torch.Tensor) -> None:
        if len(self._pool) < 2:
            self._pool.append(ws)

    def trim(self) -> None:
        """Drop the pooled workspaces (torch's allocator keeps the memory
        cached until torch.cuda.empty_cache())."""
        self._pool.clear()

    def _call(self, ws, params=None, field=None, wavelets=None, traces=None, grad_traces=None,
              grad_params=None, grad_field=None) -> _lib.Call:
        c = _lib.Call()
        c.workspace = _ptr(ws)
        c.params = _ptr(params)
        c.field = _ptr(field)
        c.wavelets = _ptr(wavelets)
        c.traces = _ptr(traces)
        c.grad_traces = _ptr(grad_traces)
        c.grad_params = _ptr(grad_params)
        c.grad_field = _ptr(grad_field)
        return c

    def _check_inputs(self, params, field, wavelets):
        sv = self.survey
        dev = self.device
        def chk(t, name, numel):
            if t is None:
                raise ValueError(f"{name} is required for mode {sv.mode}")
            if t.device != dev or t.dtype != torch.float32 or not t.is_contiguous():
                raise ValueError(f"{name} must be a contiguous float32 tensor on {dev}")
            if t.numel() != numel:
                raise ValueError(f"{name} has {t.numel()} elements, expected {numel}")
        if self.mode != 2:
            chk(params, "params", sv.n_sets * 6)
        if self.mode != 0:
            chk(field, "field", int(self.info.field_floats))
        chk(wavelets, "wavelets", sum(s.nt for s in sv.shots))

    # ----------------------------------------------------------------- passes
    def forward(self, params, field, wavelets, ws: torch.Tensor, traces: Optional[torch.Tensor] = None):
        """Run the forward recurrence; returns the flat trace tensor."""
        self._check_inputs(params, field, wavelets)
        if traces is None:
            traces = torch.empty(self.trace_floats, dtype=torch.float32, device=self.device)
        c = self._call(ws, params, field, wavelets, traces=traces)
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().ernn_forward(self._plan, C.byref(c),
                                               C.c_void_p(_lib.stream_handle(self.device))), "ernn_forward")
        return traces

    def adjoint(self, params, field, wavelets, grad_traces: torch.Tensor, ws: torch.Tensor
                ) -> Tuple[Optional[torch.Tensor], Optional[torch.Tensor]]:
        """Gradient of J given dJ/d traces; returns (grad_params, grad_field) in fp64
        (grad_params None for six fields and for a fixed_c0 plan)."""
        self._check_inputs(params, field, wavelets)
        if grad_traces.numel() != self.trace_floats or grad_traces.dtype != torch.float32:
            raise ValueError("grad_traces must be float32 with the trace layout")
        grad_traces = grad_traces.contiguous()
        gp = (torch.empty(self.survey.n_sets * 6, dtype=torch.float64, device=self.device)
              if self.mode != 2 and not self.fixed_c0 else None)
        gf = (torch.empty(int(self.info.grad_field_doubles), dtype=torch.float64, device=self.device)
              if self.mode != 0 else None)
        c = self._call(ws, params, field, wavelets, grad_traces=grad_traces, grad_params=gp, grad_field=gf)
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().ernn_adjoint(self._plan, C.byref(c),
                                               C.c_void_p(_lib.stream_handle(self.device))), "ernn_adjoint")
        return gp, gf

    # ----------------------------------------------------------- layouts
    def shot_traces(self, traces: torch.Tensor, b: int) -> torch.Tensor:
        """View of shot b's traces: (2, n_frames, n_rec) or (2, n_frames, nx, ny)."""
        s = self.survey.shots[b]
        nf = self.survey.n_frames(s)
        block = traces[self.trace_off[b]: self.trace_off[b + 1]]
        if self.survey.snapshot:
            return block.view(2, nf, s.nx, s.ny)
        return block.view(2, nf, -1)

    def stacked(self, traces: torch.Tensor) -> torch.Tensor:
        """(n_shots, 2, n_frames, X...) view when all shots share one block shape."""
        shapes = {tuple(self.shot_traces(traces, b).shape) for b in range(len(self.survey.shots))}
        if len(shapes) != 1:
            raise ValueError("shots have different trace shapes; use shot_traces()")
        return traces.view(len(self.survey.shots), *shapes.pop())

    # ----------------------------------------------------------- tuning
    @property
    def tiles(self) -> Tuple[int, int, int, int, int, int]:
        i = self.info
        return (int(i.fwd_v), int(i.fwd_r), int(i.adj_v), int(i.adj_r), int(i.fwd_bpc), int(i.adj_bpc))

    @property
    def blocking(self) -> Tuple[int, int, int]:
        i = self.info
        return (int(i.blk_v), int(i.blk_rw), int(i.blk_ks))

    @property
    def adj_blocking(self) -> Tuple[int, int, int]:
        i = self.info
        return (int(i.ablk_v), int(i.ablk_rw), int(i.ablk_ks))

    @property
    def ff_rw(self) -> int:
        """Rows per wavefront of the fused kernel's recompute half (0: the plan
        has no fused time-blocked backward)."""
        return int(self.info.ablk_ff_rw)

    def set_blocking(self, v: int, rw: int, ks: int, which: int = 0) -> None:
        """Time-blocked forward (which 0) or adjoint (which 1) with ks steps
        per launch (ks <= 1: one step per launch); which 2: rows per wavefront
        rw of the fused kernel's recompute half (v, ks: the adjoint's)."""
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().ernn_plan_set_blocking(self._plan, which, v, rw, ks), "ernn_plan_set_blocking")
            _lib.check(_lib.lib().ernn_plan_info(self._plan, C.byref(self.info)), "ernn_plan_info")

    def set_tiling(self, which: int, v: int, r: int, bpc: int = 0) -> None:
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().ernn_plan_set_tiling(self._plan, which, v, r, bpc), "ernn_plan_set_tiling")
            _lib.check(_lib.lib().ernn_plan_info(self._plan, C.byref(self.info)), "ernn_plan_info")

    def time_steps(self, which: int, steps: int, params, field, wavelets, ws, traces, grad_traces) -> float:
        c = self._call(ws, params, field, wavelets, traces=traces, grad_traces=grad_traces)
        ms = C.c_float()
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().ernn_time_steps(self._plan, which, steps, C.byref(c),
                                                  C.c_void_p(_lib.stream_handle(self.device)), C.byref(ms)),
                       "ernn_time_steps")
        return float(ms.value)
