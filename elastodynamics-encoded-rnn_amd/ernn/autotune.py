"""Autotuning: checkpoint interval under a memory cap, and tile shapes by timing.

Checkpoint interval K.  Each instance keeps 4 + S state slots of six padded
fields (two forward ring slots, the S = ceil(nt / K) checkpoints, two adjoint
slots; a state is u, its compensation c and its increment v per component)
and K bank slots of two fields per bank (the recomputed u of one segment; two
banks when the adjoint is fused with the next segment's recompute).
fields(K) = 6 (4 + ceil(nt/K)) + 2 banks K is minimal near K = sqrt(3 nt /
banks).  The forward pass leaves the last segment's states in the bank, so
the backward recomputes nt - (last segment) steps: a larger K costs memory
and saves recompute, and K = nt (one segment) recomputes nothing.
``choose_ckpt`` returns, among the K whose workspace fits a memory budget,
the one that recomputes the fewest steps, or raises if even the K of least
memory exceeds the cap.  A reverse_u plan has no bank and recomputes nothing:
its K is REV_SEGMENT (or the shortest longer interval that fits).

Tiles.  ``tune_tiles`` times every candidate (V columns per lane, R rows per
wavefront) of the forward and the adjoint kernel on the real problem for a
few steps (hipEvent timing inside the library) and keeps the fastest.
"""
from __future__ import annotations

import math
from typing import Iterable, Optional, Sequence, Tuple

import torch

from .problem import Survey

LEFT_PAD = 32
COL_ALIGN = 256
ROW_ALIGN = 16


def padded_field_floats(nx: int, ny: int) -> int:
    """Floats of one padded field (the library's layout, ernn_internal.h)."""
    pitch = LEFT_PAD + (ny + COL_ALIGN - 1) // COL_ALIGN * COL_ALIGN
    return ((nx + ROW_ALIGN - 1) // ROW_ALIGN * ROW_ALIGN + 3) * pitch


STATE_FIELDS = 6   # u, c, v per component (the library's kStateFields)
BANK_FIELDS = 2    # u per component


# Checkpoint interval of a reverse_u plan when none is named: each segment's
# reversed reconstruction of u starts over from an exact checkpoint, so the
# interval bounds how far rounding can carry it.  60 is a multiple of every
# step count per launch up to 6 (2, 4, 5, 6): no segment ends in a short
# reversed launch and the forward's checkpoints fall on launch ends.
REV_SEGMENT = 60


def fields(nt: int, K: int, banks: int = 1, reverse_u: bool = False) -> int:
    """Padded fields per instance for nt steps and checkpoint interval K
    (reverse_u: no bank; S + 1 checkpoints and two more ring states)."""
    if reverse_u:
        return STATE_FIELDS * (7 + math.ceil(nt / K))
    return STATE_FIELDS * (4 + math.ceil(nt / K)) + BANK_FIELDS * banks * K


def state_bytes(survey: Survey, K: int) -> int:
    nt = survey.nt_max
    b = 2 if survey.fused else 1
    return sum(fields(nt, K, b, survey.reverse_u) * padded_field_floats(s.nx, s.ny) * 4 for s in survey.shots)


def choose_ckpt(survey: Survey, mem_cap_bytes: Optional[int] = None, budget_bytes: Optional[int] = None) -> int:
    """Checkpoint interval.  Without a budget: the K of least state memory.
    With one: among the K from that one up whose workspace fits it, the K
    that recomputes the fewest steps (nt: none).  Raises if the K of least
    memory exceeds the cap."""
    nt = survey.nt_max
    if nt < 4:
        return 2
    if survey.reverse_u:   # REV_SEGMENT steps, or the shortest interval from there that fits
        fit = [K for K in range(min(REV_SEGMENT, nt), nt + 1)
               if state_bytes(survey, K) <= (budget_bytes or mem_cap_bytes or float("inf"))]
        if not fit:
            raise MemoryError(f"checkpointed state needs {state_bytes(survey, nt) / 2**30:.2f} GiB or more "
                              f"> cap {(budget_bytes or mem_cap_bytes) / 2**30:.2f} GiB")
        return fit[0]
    b = 2 if survey.fused else 1
    best = min(range(2, nt + 1), key=lambda K: (fields(nt, K, b), -K))
    if mem_cap_bytes is not None and state_bytes(survey, best) > mem_cap_bytes:
        raise MemoryError(f"checkpointed state needs {state_bytes(survey, best) / 2**30:.2f} GiB "
                          f"> cap {mem_cap_bytes / 2**30:.2f} GiB")
    if budget_bytes:
        fit = [K for K in range(best, nt + 1) if state_bytes(survey, K) <= budget_bytes]
        if fit:   # fewest recomputed steps (every segment but the last), then least memory
            best = min(fit, key=lambda K: (recomputed_steps(nt, K), state_bytes(survey, K)))
    return best


def recomputed_steps(nt: int, K: int) -> int:
    """Steps the backward recomputes: every checkpoint segment but the last,
    whose states the forward pass leaves in the bank."""
    return (math.ceil(nt / K) - 1) * K


# every compiled tile shape: V columns per lane x R <= 8 rows, V * R <= 16
TILE_SHAPES: Tuple[Tuple[int, int], ...] = (
    (1, 1), (1, 2), (1, 4), (1, 8), (2, 1), (2, 2), (2, 4), (2, 8), (4, 1), (4, 2), (4, 4))
# grid caps in blocks per CU (0: one tile per wavefront; otherwise persistent wavefronts)
GRID_CAPS: Tuple[int, ...] = (0, 2, 4, 8)
DEFAULT_CANDIDATES: Tuple[Tuple[int, int, int], ...] = tuple(
    (v, r, b) for (v, r) in TILE_SHAPES for b in GRID_CAPS)
# time-blocked forward shapes (V, RW, KS); field modes leave out (V, RW) = (2, 8)
BLOCK_SHAPES: Tuple[Tuple[int, int, int], ...] = (
    (1, 4, 2), (1, 4, 4), (1, 8, 2), (1, 8, 4), (1, 8, 5), (1, 8, 8), (2, 4, 2), (2, 4, 4), (2, 8, 2), (2, 8, 4),
    (2, 8, 8))
# time-blocked adjoint shapes (constant and damage modes; V = 2 for constant stiffness only)
ADJ_BLOCK_SHAPES: Tuple[Tuple[int, int, int], ...] = ((1, 4, 2), (1, 4, 4), (1, 6, 4), (1, 8, 2), (1, 8, 4),
                                                      (1, 8, 5), (1, 8, 8), (2, 4, 2), (2, 4, 4))
# time-blocked adjoint shapes of the six-field mode (16-row regions: 8 waves x 2 rows or 4 x 4)
ADJ_BLOCK_SHAPES_FIELD6: Tuple[Tuple[int, int, int], ...] = ((1, 2, 2), (1, 2, 4), (1, 4, 2), (1, 4, 4))
# time-blocked shapes of the fused backward (fuse plans): both halves of a launch share one shape;
# (1, 8, 4) is built for six constants only
FUSED_BLOCK_SHAPES: Tuple[Tuple[int, int, int], ...] = ((1, 4, 2), (1, 4, 4), (1, 8, 4))
# reversed-u adjoint shapes (reverse_u plans)
REV_BLOCK_SHAPES: Tuple[Tuple[int, int, int], ...] = ((1, 4, 2), (1, 4, 4), (1, 6, 2), (1, 6, 4), (1, 8, 2),
                                                      (1, 8, 4), (1, 8, 5), (1, 8, 6), (1, 10, 5), (1, 10, 6))
# largest V * R and R of the built adjoint kernels per mode (larger shapes would spill registers)
ADJ_MAX_CELLS = {"const6": 16, "dfield": 8, "field6": 4}
ADJ_MAX_ROWS = {"const6": 16, "dfield": 4, "field6": 2}


def timing_buffers(engine):
    """(traces, grad_traces) for ernn_time_steps: the adjoint runs on seeded
    Gaussian residuals, so its timed launches carry data like a real backward
    (all-zero operands draw less power and can run faster than real ones)."""
    traces = torch.zeros(engine.trace_floats, dtype=torch.float32, device=engine.device)
    gen = torch.Generator(device="cpu").manual_seed(0)
    grad = torch.randn(engine.trace_floats, generator=gen, dtype=torch.float32).to(engine.device)
    return traces, grad


def tune_tiles(engine, params, field, wavelets, steps: int = 50,
               candidates: Sequence[Tuple[int, int, int]] = DEFAULT_CANDIDATES,
               which: Iterable[int] = (0, 1), rounds: int = 5) -> dict:
    """Time candidates for the forward (0) and adjoint (1) kernels; keep the best.

    Candidates are timed in `rounds` interleaved passes (so clock and cache
    drift hit every candidate alike) and ranked by their median.
    A candidate is (V, R, blocks per CU).  which: 0 forward, 1 adjoint,
    2 fused adjoint + recompute (sets the adjoint tiling; fuse plans).
    Returns {which: (V, R, bpc, ms_per_step, {candidate: ms})}.  A reverse_u
    plan has no single-step adjoint: its which 1 is left out.
    """
    if getattr(engine, "reverse_u", False):
        which = tuple(w for w in which if w != 1)
    ws = engine.acquire_workspace()
    traces, grads = timing_buffers(engine)
    for w in which:   # bring clocks and caches to steady state before the first candidate
        engine.time_steps(w, 4 * steps, params, field, wavelets, ws, traces, grads)
    result = {}
    for w in which:
        mode = engine.survey.mode
        cands = [c for c in candidates
                 if (w == 0 and not (mode == "field6" and c[0] * c[1] > 8))
                 or (w != 0 and c[0] * c[1] <= ADJ_MAX_CELLS[mode] and c[1] <= ADJ_MAX_ROWS[mode])]
        kernel = 1 if w == 2 else w
        samples = {c: [] for c in cands}
        for _ in range(rounds):
            for c in cands:
                engine.set_tiling(kernel, *c)
                engine.time_steps(w, 2, params, field, wavelets, ws, traces, grads)   # warm-up
                samples[c].append(engine.time_steps(w, steps, params, field, wavelets, ws,
                                                    traces, grads) / steps)
        times = {c: sorted(t)[len(t) // 2] for c, t in samples.items()}
        best = min(times, key=times.get)
        engine.set_tiling(kernel, *best)
        result[w] = (best[0], best[1], best[2], times[best], times)
    engine.release_workspace(ws)
    return result


def tune_blocking(engine, params, field, wavelets, steps: int = 64, which: int = 0,
                  shapes: Optional[Sequence[Tuple[int, int, int]]] = None, rounds: int = 5) -> tuple:
    """Time the time-blocked forward (which 0) or adjoint (which 1) shapes
    against the current single-step tiling of that pass, per step, in
    interleaved rounds; keep the fastest.  On a fuse plan the adjoint shapes
    are those of the fused time-blocked backward, timed per step of adjoint
    plus recompute against the fused single-step kernel.

    Returns (V, RW, KS, ms_per_step, {shape: ms_per_step}); (0, 0, 0) is the
    single-step kernel.  A reverse_u plan times its reversed adjoint shapes
    (REV_BLOCK_SHAPES; it has no single-step adjoint)."""
    mode = engine.survey.mode
    fused = which == 1 and bool(engine.info.fuse)
    if which == 1 and engine.reverse_u:
        cands = list(REV_BLOCK_SHAPES if shapes is None else shapes)
        if not getattr(engine, "fixed_c0", True):   # the joint gradient is not built for ten-row bands
            cands = [c for c in cands if c[1] != 10]
    elif which == 0:
        shapes = BLOCK_SHAPES if shapes is None else shapes
        cands = [(0, 0, 0)] + [b for b in shapes if mode == "const6" or b[:2] != (2, 8)]
    else:
        if mode == "field6":
            cands = [(0, 0, 0)] + ([] if fused else list(ADJ_BLOCK_SHAPES_FIELD6 if shapes is None else shapes))
        else:
            shapes = (FUSED_BLOCK_SHAPES if fused else ADJ_BLOCK_SHAPES) if shapes is None else shapes
            cands = [(0, 0, 0)] + [b for b in shapes if mode == "const6" or (b[0] == 1 and b[2] != 5 and
                                                                             not (fused and b[1] == 8))]
    ws = engine.acquire_workspace()
    traces, grads = timing_buffers(engine)
    samples = {c: [] for c in cands}

    def run(c):
        if c[2] <= 1:
            return engine.time_steps(2 if fused else which, steps, params, field, wavelets, ws, traces,
                                     grads) / steps
        launches = max(2, steps // c[2])
        return engine.time_steps(3 + which, launches, params, field, wavelets, ws, traces,
                                 grads) / (launches * c[2])

    for _ in range(rounds):
        for c in cands:
            engine.set_blocking(*c, which=which)
            run(c)   # warm-up
            samples[c].append(run(c))
    times = {c: sorted(t)[len(t) // 2] for c, t in samples.items()}
    best = min(times, key=times.get)
    engine.set_blocking(*best, which=which)
    engine.release_workspace(ws)
    return (best[0], best[1], best[2], times[best], times)


def tune_fused_recompute(engine, params, field, wavelets, steps: int = 64, rounds: int = 5) -> tuple:
    """On a fuse plan whose time-blocked adjoint has 4-row bands (V = 1, KS in
    {2, 4}), time the fused kernel with its recompute half at the adjoint's
    shape and at 8 rows per wavefront, per reverse step, in interleaved
    rounds; keep the faster.  The taller recompute blocks cut the launch's
    block count, so that on small grids every block is resident at once.

    Returns (rw, {rw: ms_per_step}); ({}: not applicable, nothing timed)."""
    v, rw, ks = engine.adj_blocking
    if not engine.info.fuse or v != 1 or rw != 4 or ks not in (2, 4):
        return engine.ff_rw, {}
    ws = engine.acquire_workspace()
    traces, grads = timing_buffers(engine)
    launches = max(2, steps // ks)
    samples = {c: [] for c in (rw, 8)}
    for _ in range(rounds):
        for c in samples:
            engine.set_blocking(v, c, ks, which=2)
            engine.time_steps(4, launches, params, field, wavelets, ws, traces, grads)   # warm-up
            samples[c].append(engine.time_steps(4, launches, params, field, wavelets, ws, traces, grads)
                              / (launches * ks))
    times = {c: sorted(t)[len(t) // 2] for c, t in samples.items()}
    best = min(times, key=times.get)
    engine.set_blocking(v, best, ks, which=2)
    engine.release_workspace(ws)
    return best, times


def tune_backward(survey, params, field, wavelets, steps: int = 48, rounds: int = 3) -> tuple:
    """Choose the backward schedule of a survey by timing the candidates on
    the real problem, per reverse step (adjoint + recompute):

    - fused: each launch runs adjoint and recompute work of two consecutive
      checkpoint segments, one step each (fused single-step kernel) or KS
      steps each (fused time-blocked kernel), at the best tiles / shape;
    - unfused: the segment recompute by the time-blocked forward, then the
      time-blocked adjoint, each at its best shape.

    The fused schedule fixes the plan's slot layout (two segment banks), so
    each schedule is timed on its own temporary plan.  The survey's `fuse` is
    left as it was.

    Returns (fuse, adj_blocking, ff_rw, {"fused": ms_per_step, "blocked": ms_per_step}):
    ff_rw is the rows per wavefront of the fused kernel's recompute half the
    fused time is taken with (Survey.adj_ff_rw), 0 for the unfused schedule."""
    from .engine import Engine
    if survey.mode == "field6":
        return survey.fused, (0, 0, 0), 0, {}
    saved = survey.fuse, survey.adj_ff_rw, survey.ckpt_budget
    times = {}
    ff_rw = 0
    try:
        # the temporary plans only time steps: checkpoint interval of least memory
        survey.fuse, survey.adj_ff_rw, survey.ckpt_budget = True, 0, 0
        eng = Engine(survey, device=params.device)
        try:
            tune_tiles(eng, params, field, wavelets, steps=steps, which=(2,), rounds=rounds)
            fz = tune_blocking(eng, params, field, wavelets, steps=steps, which=1, rounds=rounds)
            times["fused"] = fz[3]
            ff_rw, ft = tune_fused_recompute(eng, params, field, wavelets, steps=steps, rounds=rounds)
            if ft:   # the fused kernel's recompute half on taller blocks
                times["fused"] = min(ft.values())
        finally:
            eng.close()
        survey.fuse = False
        eng = Engine(survey, device=params.device)
        try:
            tune_tiles(eng, params, field, wavelets, steps=steps, which=(0, 1), rounds=rounds)
            fwd = tune_blocking(eng, params, field, wavelets, steps=steps, which=0, rounds=rounds)
            adj = tune_blocking(eng, params, field, wavelets, steps=steps, which=1, rounds=rounds)
            times["blocked"] = fwd[3] + adj[3]
        finally:
            eng.close()
    finally:
        survey.fuse, survey.adj_ff_rw, survey.ckpt_budget = saved
    fuse = times["fused"] <= times["blocked"]
    return fuse, (fz[:3] if fuse else adj[:3]), (ff_rw if fuse else 0), times
